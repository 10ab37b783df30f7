"""Pins the oracle's PLANEWAVE restatement (oracle/planewave.py) without the
reference binary: known answers of the reference's grid rule, brute-force
enumeration of the G sphere, direct DFT sums, plane-wave delta functions and
round trips.  CPU only."""
import itertools

import numpy as np
import pytest

from oracle import planewave as opw
from pwgpu import model

RY = 0.5


def test_allowed_fft_lengths_closed_form():
    # LIB$FFTADJUSTGRD (src/paw_library.f90:2336-2393): n = 2^h 3^i 5^j 7^k 11^m,
    # h>=1, i<=2, j,k,m<=1; the first list entry (2) is skipped by the search
    def allowed(n):
        if n % 2:
            return False
        m = n // 2
        for p, emax in ((3, 2), (5, 1), (7, 1), (11, 1)):
            e = 0
            while m % p == 0:
                m //= p
                e += 1
            if e > emax:
                return False
        while m % 2 == 0:
            m //= 2
        return m == 1
    for nr in range(1, 400):
        want = next(n for n in range(max(nr, 4), 100000) if allowed(n))
        assert opw.fftadjustgrd(nr) == want, nr


@pytest.mark.parametrize("cells,rho_ry,grid", [
    ((1, 1, 1), 60.0, 28), ((1, 1, 1), 120.0, 36), ((2, 2, 2), 120.0, 72), ((2, 2, 2), 60.0, 56)])
def test_grid_rule_known_answers(cells, rho_ry, grid):
    rbas = np.diag([10.26 * c for c in cells])
    assert opw.divider_grid(rho_ry * RY, rbas) == (grid, grid, grid)


def test_known_sphere_counts():
    # Si8 2 945, Si64 23 847 (30 Ry), 32 H2O box 19.7 bohr at 40 Ry: 32 675, grid 80
    assert len(opw.getig(opw.gbass(np.diag([10.26] * 3))[0], 15.0)) == 2945
    assert len(opw.getig(opw.gbass(np.diag([20.52] * 3))[0], 15.0)) == 23847
    pl = opw.Plane(np.diag([19.7] * 3), 20.0, 80.0)
    assert pl.ngl == 32675 and (pl.nr1, pl.nr2, pl.nr3) == (80, 80, 80)


def test_gbass_reciprocal():
    s = model.tiny()
    g, vol = opw.gbass(s.rbas)
    assert np.allclose(g.T @ s.rbas, 2 * np.pi * np.eye(3), atol=1e-12)
    assert abs(vol - abs(np.linalg.det(s.rbas))) < 1e-10


def test_sphere_bruteforce_set_and_order():
    s = model.tiny()
    g, _ = opw.gbass(s.rbas)
    ig = opw.getig(g, s.epw_psi)
    want = set()
    for t in itertools.product(range(-12, 13), repeat=3):
        gv = g @ np.array(t, dtype=float)
        if 0.5 * gv @ gv <= s.epw_psi:
            want.add(t)
    assert set(map(tuple, ig.tolist())) == want
    gsq = np.sum((ig @ g.T) ** 2, axis=1)
    assert np.all(np.diff(gsq) > -1e-5)
    # inside a shell of equal |G|^2 the order is lexicographic in (i1,i2,i3)
    for a in range(len(ig) - 1):
        if abs(gsq[a + 1] - gsq[a]) < 1e-9:
            assert tuple(ig[a]) < tuple(ig[a + 1])


def test_sphere_order_first_shell_cubic():
    g, _ = opw.gbass(np.diag([10.0] * 3))
    ig = opw.getig(g, 2.0)
    assert ig[0].tolist() == [0, 0, 0]
    assert ig[1:7].tolist() == [[-1, 0, 0], [0, -1, 0], [0, 0, -1], [0, 0, 1], [0, 1, 0], [1, 0, 0]]


def test_minusg_involution():
    pl = opw.Plane(model.tiny().rbas, 3.0, 12.0)
    assert np.array_equal(pl.minusg[pl.minusg], np.arange(pl.ngl))
    assert np.array_equal(pl.igvec[pl.minusg], -pl.igvec)
    assert pl.ngamma == 0 and pl.igvec[0].tolist() == [0, 0, 0]


def _grid_r(pl):
    i3, i2, i1 = np.meshgrid(np.arange(pl.nr3), np.arange(pl.nr2), np.arange(pl.nr1), indexing="ij")
    return i1 / pl.nr1, i2 / pl.nr2, i3 / pl.nr3


def test_fft_direct_dft():
    s = model.tiny()
    pl = opw.Plane(s.rbas, s.epw_psi, s.epw_rho)
    rng = np.random.default_rng(0)
    c = rng.normal(size=pl.ngl) + 1j * rng.normal(size=pl.ngl)
    x1, x2, x3 = _grid_r(pl)
    ph = 2j * np.pi * (pl.igvec[:, 0, None, None, None] * x1 + pl.igvec[:, 1, None, None, None] * x2
                       + pl.igvec[:, 2, None, None, None] * x3)
    direct = np.tensordot(c, np.exp(ph), axes=(0, 0))          # sum_G c exp(+iGr)
    assert np.max(np.abs(pl.fft_gtor(c) - direct)) < 1e-10 * np.max(np.abs(direct))
    f = rng.normal(size=(pl.nr3, pl.nr2, pl.nr1)) + 1j * rng.normal(size=(pl.nr3, pl.nr2, pl.nr1))
    direct_g = np.array([np.sum(f * np.exp(-p)) for p in ph]) / pl.nrl   # 1/N sum_r f exp(-iGr)
    assert np.max(np.abs(pl.fft_rtog(f) - direct_g)) < 1e-12 * np.max(np.abs(direct_g))


def test_fft_plane_wave_and_round_trip():
    s = model.tiny()
    pl = opw.Plane(s.rbas, s.epw_psi, s.epw_rho)
    g0 = 7
    c = np.zeros(pl.ngl, dtype=complex)
    c[g0] = 1.0
    f = pl.fft_gtor(c)
    x1, x2, x3 = _grid_r(pl)
    iv = pl.igvec[g0]
    assert np.max(np.abs(f - np.exp(2j * np.pi * (iv[0] * x1 + iv[1] * x2 + iv[2] * x3)))) < 1e-12
    back = pl.fft_rtog(f)
    assert np.max(np.abs(back - c)) < 1e-13
    rng = np.random.default_rng(2)
    c = rng.normal(size=(3, pl.ngl)) + 1j * rng.normal(size=(3, pl.ngl))
    assert np.max(np.abs(pl.fft_rtog(pl.fft_gtor(c)) - c)) < 1e-13


def test_structurefactor_direct():
    s = model.tiny()
    pl = opw.Plane(s.rbas, s.epw_psi, s.epw_rho)
    r = np.array([1.3, -0.7, 2.2])
    assert np.max(np.abs(opw.structurefactor(r, pl.gbas, pl.igvec) - np.exp(-1j * (pl.gvec @ r)))) < 1e-12


def test_scalarproduct_minus_variant():
    s = model.tiny()
    pl = opw.Plane(s.rbas, s.epw_psi, s.epw_rho)
    rng = np.random.default_rng(3)
    a = rng.normal(size=(2, pl.ngl)) + 1j * rng.normal(size=(2, pl.ngl))
    b = rng.normal(size=(3, pl.ngl)) + 1j * rng.normal(size=(3, pl.ngl))
    m = pl.scalarproduct(a, b, "-")
    direct = np.array([[pl.vol * np.sum(a[i] * b[j][pl.minusg]) for j in range(3)] for i in range(2)])
    assert np.max(np.abs(m - direct)) < 1e-12 * np.max(np.abs(direct))
