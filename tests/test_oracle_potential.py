"""Pins the oracle of the density -> potential step (oracle/potential.py,
POTENTIAL$VOFRHO, src/paw_potential.f90:324-896) without the reference:

* local exchange: the closed form of the homogeneous electron gas
  e_x = -(3/4)(3/pi)^(1/3) rho^(4/3);
* Perdew-Zunger correlation: published correlation energies per electron of
  the homogeneous gas (Ceperley-Alder as fitted by Perdew and Zunger, Phys.
  Rev. B 23, 5048 (1981), Table XII), continuity of both branches at rs = 1,
  and v_c = d(rho e_c)/drho by finite differences in both branches;
* Hartree: the closed-form energy of a single cosine density;
* the returned potential, forces and VQLM are the derivatives of the returned
  energy EHARTREE + EXC with respect to the density, the positions and QLM
  (central finite differences), with two species, the pseudo core, V-bar,
  the compensation charges and the background all active.
A dropped term, a wrong sign or factor in any of these fails one of the
finite-difference checks.
"""
import math

import numpy as np
import pytest

from oracle import planewave as opw
from oracle import potential as opot
from pwgpu import model

Y0 = 1.0 / math.sqrt(4.0 * math.pi)


def tabs_dict(t):
    return {"pscorg": t[0], "vbarg": t[1], "g0": t[2], "v0": t[3]}


def test_exchange_closed_form():
    rho = np.array([1e-4, 3e-3, 0.05, 0.7, 4.0])
    ex, vx = opot.exchange_lda(rho)
    cx = -0.75 * (3.0 / math.pi) ** (1.0 / 3.0)
    assert np.allclose(ex, cx * rho ** (4.0 / 3.0), rtol=1e-14)
    h = 1e-6 * rho
    fd = (opot.exchange_lda(rho + h)[0] - opot.exchange_lda(rho - h)[0]) / (2 * h)
    assert np.allclose(vx, fd, rtol=1e-8)


@pytest.mark.parametrize("rs,ec", [(1.0, -0.0600), (2.0, -0.0448), (5.0, -0.0282), (10.0, -0.0186)])
def test_pz_correlation_published_values(rs, ec):
    rho = 3.0 / (4.0 * math.pi * rs ** 3)
    e, _ = opot.pz_unpolarised(np.array([rho]))
    assert abs(e[0] / rho - ec) < 4e-4, (rs, e[0] / rho)


def test_pz_branches_continuous_and_derivative():
    rs1 = 3.0 / (4.0 * math.pi)     # rho at rs = 1
    lo = np.array([rs1 * (1 + 1e-9)])
    hi = np.array([rs1 * (1 - 1e-9)])
    elo = opot.pz_unpolarised(lo)[0][0] / lo[0]
    ehi = opot.pz_unpolarised(hi)[0][0] / hi[0]
    assert abs(elo - ehi) < 1e-4
    for rs in (0.3, 0.8, 1.7, 6.0):
        rho = np.array([3.0 / (4.0 * math.pi * rs ** 3)])
        h = 1e-6 * rho
        fd = (opot.pz_unpolarised(rho + h)[0] - opot.pz_unpolarised(rho - h)[0]) / (2 * h)
        assert abs(opot.pz_unpolarised(rho)[1][0] - fd[0]) < 1e-8 * abs(fd[0]), rs


def test_lda_floor():
    # below RHOTMIN the energy density scales linearly and the potential is
    # e(RHOTMIN)/RHOTMIN (src/paw_dft.f90:780-965)
    exc, vxc = opot.lda(np.array([1e-7, 5e-7]))
    e0, _ = opot.lda(np.array([opot.RHOTMIN]))
    assert np.allclose(exc, e0[0] * np.array([1e-7, 5e-7]) / opot.RHOTMIN, rtol=1e-12)
    assert np.allclose(vxc, e0[0] / opot.RHOTMIN, rtol=1e-12)


def test_hartree_single_cosine():
    # rho(r) = rho0 + 2A cos(G.r): E_H = Omega 4 pi A^2 / |G|^2, RHOB = -rho0
    s = model.si8()
    pl = opw.Plane(s.rbas, s.epw_rho, s.epw_rho)
    n = np.array([1, 2, 0])
    g = pl.gbas @ n
    i3, i2, i1 = np.meshgrid(np.arange(pl.nr3), np.arange(pl.nr2), np.arange(pl.nr1), indexing="ij")
    phase = 2 * np.pi * (n[0] * i1 / pl.nr1 + n[1] * i2 / pl.nr2 + n[2] * i3 / pl.nr3)
    rho0, amp = 0.02, 0.003
    rho = rho0 + 2 * amp * np.cos(phase)
    t = np.zeros((4, 1, pl.ngl))
    r = opot.vofrho(pl, rho, np.zeros((0, 3)), np.zeros(0, dtype=int), np.zeros(0), tabs_dict(t))
    assert abs(r["ehartree"] - pl.vol * 4 * np.pi * amp ** 2 / (g @ g)) < 1e-12 * r["ehartree"]
    assert abs(r["rhob"] + rho0) < 1e-15


def system_case():
    s = model.water(nmol=2, a=9.0, seed=4, epw_psi_ry=10.0, cdual=4.0)
    pl = opw.Plane(s.rbas, s.epw_rho, s.epw_rho)
    rho = s.valence_density((pl.nr1, pl.nr2, pl.nr3))
    return s, pl, rho, tabs_dict(s.potential_tables(pl.g2))


def energy(pl, rho, pos, isp, qlm, tabs):
    r = opot.vofrho(pl, rho, pos, isp, qlm, tabs)
    return r["ehartree"] + r["exc"], r


def test_potential_is_density_derivative():
    s, pl, rho, tabs = system_case()
    e0, r = energy(pl, rho, s.pos, s.isp, s.qlm(), tabs)
    rng = np.random.default_rng(3)
    # a smooth direction inside the density sphere
    c = (rng.normal(size=pl.ngl) + 1j * rng.normal(size=pl.ngl)) * np.exp(-pl.g2)
    c = 0.5 * (c + np.conj(c[pl.minusg]))
    d = pl.fft_gtor(c).real
    d *= 1e-3 / np.max(np.abs(d))
    eps = 1e-3
    ep, _ = energy(pl, rho + eps * d, s.pos, s.isp, s.qlm(), tabs)
    em, _ = energy(pl, rho - eps * d, s.pos, s.isp, s.qlm(), tabs)
    fd = (ep - em) / (2 * eps)
    an = float(np.sum(r["v"] * d)) * pl.rweight
    assert abs(fd - an) < 1e-7 * abs(an), (fd, an)


def test_forces_are_position_derivatives():
    s, pl, rho, tabs = system_case()
    qlm = s.qlm()
    _, r = energy(pl, rho, s.pos, s.isp, qlm, tabs)
    h = 1e-4
    for a in (0, 1, 4):
        for k in range(3):
            p = s.pos.copy()
            p[a, k] += h
            ep, _ = energy(pl, rho, p, s.isp, qlm, tabs)
            p[a, k] -= 2 * h
            em, _ = energy(pl, rho, p, s.isp, qlm, tabs)
            fd = -(ep - em) / (2 * h)
            assert abs(fd - r["force"][a, k]) < 1e-6 * max(1e-3, np.max(np.abs(r["force"]))), (a, k)


def test_vqlm_is_qlm_derivative():
    s, pl, rho, tabs = system_case()
    qlm = s.qlm()
    _, r = energy(pl, rho, s.pos, s.isp, qlm, tabs)
    h = 1e-4
    for a in range(s.nat):
        q = qlm.copy()
        q[a] += h
        ep, _ = energy(pl, rho, s.pos, s.isp, q, tabs)
        q[a] -= 2 * h
        em, _ = energy(pl, rho, s.pos, s.isp, q, tabs)
        fd = (ep - em) / (2 * h)
        assert abs(fd - r["vqlm"][a]) < 1e-7 * max(1.0, abs(fd)), a


def scf_waveset(s):
    from harness import oracle_waveset
    o = oracle_waveset(s)
    dpl = opw.Plane(s.rbas, s.epw_rho, s.epw_rho)
    o.potential = (dpl, tabs_dict(s.potential_tables(dpl.g2)), s.isp, s.qlm())
    return o


def test_scf_hpsi_is_energy_gradient():
    # with the density -> potential step linked in, ETOT is no longer linear
    # in the density; HPSI must still be its gradient: for a direction dpsi in
    # fully occupied super bands, dE/deps = 2 f Omega Re sum_G conj(dpsi) Hpsi
    from oracle import waves as owv
    s = model.tiny()
    o = scf_waveset(s)
    o.etot(s.pos)                       # Gram-Schmidt start
    psi = o.psi0.copy()
    rng = np.random.default_rng(21)
    d = np.zeros_like(psi)
    d[:2] = (rng.normal(size=(2, o.plane.ngl)) + 1j * rng.normal(size=(2, o.plane.ngl))) / (1 + o.plane.g2) ** 2
    d[:2] = 0.5 * (d[:2] + np.conj(d[:2][:, o.plane.minusg]))   # Gamma: c(-G) = conj c(G)

    def e_at(p):
        o.psi0 = p
        o.proj = owv.projections(o.plane, o.setup, o.pro, o.ylm, s.pos, p)
        return o.etot(s.pos)["etot"]

    e_at(psi)
    an = 2.0 * o.occ[0] * o.plane.vol * float(np.sum(np.conj(d[:2]) * o.hpsi[:2]).real)

    def cd(eps):
        return (e_at(psi + eps * d) - e_at(psi - eps * d)) / (2 * eps)

    # the energy is not quadratic in psi here: Richardson-extrapolated
    # central differences (error O(eps^4))
    fd = (4.0 * cd(5e-4) - cd(1e-3)) / 3.0
    assert abs(fd - an) < 1e-8 * abs(an), (fd, an)


def test_compensation_gaussians_closed_form():
    """The compensation density of a moment QLM(LM) in G space, CFAC(LM)
    YLMOFG(LM,G) G0(G) exp(-iG.R) with G0 = 4 pi exp(-G^2 rcsm^2/4)/Omega
    (POTENTIAL_HARTREE, src/paw_potential.f90:1146-1173), equals the plane-wave
    coefficients of the real-space Gaussian QLM C_l exp(-r^2/rcsm^2) r^l Y_LM
    normalised by GAUSSN (src/paw_generalpurpose.f90:1088-1121) to unit
    multipole moment, for l = 0, 1, 2 and every m: pins (-i)^l/(2l+1)!!, the
    solid harmonics and their order."""
    from oracle import augmentation as oaug
    from oracle import waves as owv
    a = 9.0
    pl = opw.Plane(np.diag([a, a, a]), 60.0, 60.0)
    rcsm = 0.9
    alpha = 1.0 / rcsm ** 2
    R = np.array([3.1, 4.7, 5.3])
    i3, i2, i1 = np.meshgrid(np.arange(pl.nr3), np.arange(pl.nr2), np.arange(pl.nr1), indexing="ij")
    pts = np.stack([i1 / pl.nr1 * a, i2 / pl.nr2 * a, i3 / pl.nr3 * a], axis=-1).reshape(-1, 3)
    g0 = 4.0 * np.pi * np.exp(-pl.g2 * rcsm ** 2 / 4.0) / pl.vol
    ylmg = opot.ylmofg(pl.gvec, 9)
    cf = opot.cfac(9)
    eig = opw.structurefactor(R, pl.gbas, pl.igvec)
    for lm in range(9):
        l = int(np.sqrt(lm + 1e-5))
        c = oaug.gaussn(l, alpha)
        f = np.zeros(len(pts))
        for sh in np.array(np.meshgrid([-1, 0, 1], [-1, 0, 1], [-1, 0, 1])).reshape(3, -1).T:
            d = pts - R - sh * a
            rr = np.sqrt(np.sum(d * d, axis=1))
            y = owv.ylm(d, 9)[lm]
            f += c * np.exp(-alpha * rr ** 2) * rr ** l * y
        want = pl.fft_rtog(f.reshape(pl.nr3, pl.nr2, pl.nr1).astype(complex))
        got = cf[lm] * ylmg[lm] * g0 * eig
        assert np.max(np.abs(got - want)) < 1e-9 * np.max(np.abs(want)), lm


def test_vqlm_and_forces_with_higher_moments():
    """LMRXX = 9 (moments to l = 2) with per-species LMRX 9 and 4: VQLM is the
    QLM derivative of EHARTREE + EXC for every LM, zero beyond a species'
    LMRX, and the forces are the position derivatives."""
    s, pl, rho, tabs = system_case()
    qlm = s.qlm(9)
    lmrx = np.array([9, 4])
    qlm[s.isp == 1, 4:] = 0.0

    def en(q, pos):
        r = opot.vofrho(pl, rho, pos, s.isp, q, tabs, lmrx=lmrx)
        return r["ehartree"] + r["exc"], r

    _, r = en(qlm, s.pos)
    assert r["vqlm"].shape == (s.nat, 9)
    assert np.all(r["vqlm"][s.isp == 1, 4:] == 0.0)
    h = 1e-4
    for a in (0, 1):
        for lm in range(int(lmrx[s.isp[a]])):
            q = qlm.copy()
            q[a, lm] += h
            ep, _ = en(q, s.pos)
            q[a, lm] -= 2 * h
            em, _ = en(q, s.pos)
            fd = (ep - em) / (2 * h)
            assert abs(fd - r["vqlm"][a, lm]) < 1e-7 * max(1.0, abs(fd)), (a, lm)
    for a in (0, 2):
        for k in range(3):
            p = s.pos.copy()
            p[a, k] += h
            ep, _ = en(qlm, p)
            p[a, k] -= 2 * h
            em, _ = en(qlm, p)
            fd = -(ep - em) / (2 * h)
            assert abs(fd - r["force"][a, k]) < 1e-6 * max(1e-3, np.max(np.abs(r["force"]))), (a, k)
