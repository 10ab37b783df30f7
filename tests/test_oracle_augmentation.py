"""Pins the oracle of the one-centre terms (oracle/augmentation.py:
AUGMENTATION$MOMENTS, AUGMENTATION$SPHERE, WAVES$SPHERE/ADDCONSTANTPOT)
without the reference:

* the shifted logarithmic grid's integrals and derivatives against closed
  forms, and RADIAL$POISSON against the potential of a Gaussian (erf) and
  against direct quadrature of the Green's function for l = 1, 2;
* the Gaunt coefficients against known values and the completeness sum rule;
* the local functional's first three derivatives by finite differences;
* AUGMENTATION_XC's second-order expansion against a direct angular
  quadrature of e(rho(r, Omega)) for a weakly non-spherical density;
* the one-centre Hamiltonian DATH as the derivative of the one-centre energy
  plus the plane-wave part's sum_LM VQLM QLM with respect to the density
  matrix (central finite differences), and QLM as the multipole moments of
  the all-electron minus pseudo one-centre density.
A dropped term, a wrong sign or a wrong factor in any of these fails one of
the checks.
"""
import math

import numpy as np
import pytest
from scipy import integrate, special

from oracle import augmentation as oaug
from oracle import waves as owv
from pwgpu import model


def grid():
    return oaug.RadialGrid(1.0e-4, 0.05, 250)


def test_radial_integrals_and_derivatives():
    g = grid()
    r = g.r
    # int_0^inf r^2 exp(-r^2) dr = sqrt(pi)/4 ; int r^4 exp(-2 r^2) = 3 sqrt(pi/2)/32
    assert abs(g.integral(r ** 2 * np.exp(-r * r)) - math.sqrt(math.pi) / 4) < 1e-10
    assert abs(g.integral(r ** 4 * np.exp(-2 * r * r)) - 3 * math.sqrt(math.pi / 2) / 32) < 1e-10
    # antiderivative of 2 r exp(-r^2) is 1 - exp(-r^2)
    anti = g.integrate(2 * r * np.exp(-r * r))
    assert np.max(np.abs(anti - (1 - np.exp(-r * r)))) < 2e-6
    # derivative of exp(-r^2/2) r (the first point sits at RMIN, not on the
    # equispaced x mesh, so the one-sided formula is checked from the second on)
    f = r * np.exp(-r * r / 2)
    d = g.derive(f)
    want = (1 - r ** 2) * np.exp(-r ** 2 / 2)
    err = np.abs(d - want)[2:]
    sel = r[2:] < 4.0
    assert np.max(err[sel]) < 5e-5, (np.argmax(err[sel]), np.max(err[sel]))


def test_poisson_closed_forms():
    g = grid()
    r = g.r
    a = 1.7
    # l = 0: rho = exp(-a r^2) -> V = pi^1.5 erf(sqrt(a) r) / (a^1.5 r)
    v = g.poisson(0, np.exp(-a * r * r))
    want = math.pi ** 1.5 * special.erf(math.sqrt(a) * r) / (a ** 1.5 * r)
    sel = (r > 1e-3) & (r < 15)
    assert np.max(np.abs(v[sel] - want[sel])) < 2e-6 * np.max(np.abs(want))
    # l = 1, 2: Green's-function quadrature at a few radii
    for l in (1, 2):
        rho = r ** l * np.exp(-a * r * r)
        v = g.poisson(l, rho)
        for i in (60, 120, 170, 200):
            x = r[i]
            inner = integrate.quad(lambda t: t ** (2 * l + 2) * math.exp(-a * t * t), 0, x, epsabs=1e-14)[0]
            outer = integrate.quad(lambda t: t * math.exp(-a * t * t), x, np.inf, epsabs=1e-14)[0]
            ref = 4 * math.pi / (2 * l + 1) * (x ** (-l - 1) * inner + x ** l * outer)
            assert abs(v[i] - ref) < 5e-6 * max(1.0, abs(ref)), (l, i, v[i], ref)


def test_gaunt_values_and_sum_rule():
    cg = oaug.gaunt(25)
    y0 = 1 / math.sqrt(4 * math.pi)
    assert abs(cg[0, 0, 0] - y0) < 1e-15
    # <Y_lm Y_lm Y_00> = Y0 and symmetry in all three indices
    for lm in range(25):
        assert abs(cg[lm, lm, 0] - y0) < 1e-14
    assert np.max(np.abs(cg - np.transpose(cg, (1, 0, 2)))) < 1e-15
    assert np.max(np.abs(cg - np.transpose(cg, (2, 1, 0)))) < 1e-15
    # p x p -> d: <x|xy> type couplings, <Y_x Y_x Y_{3z2-r2}> = -1/sqrt(20 pi)
    assert abs(cg[1, 1, 6] + 1 / math.sqrt(20 * math.pi)) < 1e-14
    assert abs(cg[2, 2, 6] - 2 / math.sqrt(20 * math.pi)) < 1e-14
    # completeness: sum_lm3 CG(a,b,lm3) CG(c,d,lm3) = int Y_a Y_b Y_c Y_d for l_a + l_b <= 4
    ct, wt = np.polynomial.legendre.leggauss(16)
    ph = 2 * math.pi * np.arange(32) / 32
    c, p = np.meshgrid(ct, ph, indexing="ij")
    s = np.sqrt(1 - c * c)
    pts = np.stack([s * np.cos(p), s * np.sin(p), c], -1).reshape(-1, 3)
    w = (wt[:, None] * np.full(32, 2 * math.pi / 32)).ravel()
    y = owv.ylm(pts, 9)
    for (a, b, cc, d) in [(1, 3, 2, 1), (4, 8, 5, 7), (2, 2, 1, 3)]:
        lhs = float(np.sum(cg[a, b, :25] * cg[cc, d, :25]))
        rhs = float(np.sum(y[a] * y[b] * y[cc] * y[d] * w))
        assert abs(lhs - rhs) < 1e-13


def test_lda3_derivatives():
    rho = np.array([3e-5, 2e-3, 0.05, 0.2, 0.7, 3.0])    # both Perdew-Zunger branches
    f0, f1, f2, f3 = oaug.lda3(rho)
    from oracle import potential as opot
    e, v = opot.lda(rho)
    assert np.allclose(f0, e, rtol=1e-14) and np.allclose(f1, v, rtol=1e-13)
    h = 1e-5 * rho
    for k, fk in ((1, f1), (2, f2), (3, f3)):
        prev = oaug.lda3(rho + h)[k - 1], oaug.lda3(rho - h)[k - 1]
        fd = (prev[0] - prev[1]) / (2 * h)
        assert np.allclose(fk, fd, rtol=2e-6), k
    # below RHOTMIN: linear continuation, no curvature
    lo = oaug.lda3(np.array([4e-7]))
    top = oaug.lda3(np.array([oaug.RHOTMIN]))
    assert abs(lo[0][0] - top[0][0] * 0.4) < 1e-18 and lo[2][0] == 0.0 and lo[3][0] == 0.0


def test_xc_expansion_against_angular_quadrature():
    """E = int d^3r e(rho0 Y0 + sum eps x_LM Y_LM): the expansion is exact to
    second order, so the error of AUGMENTATION_XC falls as eps^3."""
    g = grid()
    r = g.r
    rho00 = 2.0 * np.exp(-r * r / 1.5) / 0.2820948 + 1e-3
    x = np.zeros((9, len(r)))
    x[0] = rho00
    x[2] = 0.4 * r * np.exp(-r * r)
    x[5] = 0.3 * r * r * np.exp(-r * r)
    ct, wt = np.polynomial.legendre.leggauss(20)
    ph = 2 * math.pi * np.arange(40) / 40
    c, p = np.meshgrid(ct, ph, indexing="ij")
    s = np.sqrt(1 - c * c)
    pts = np.stack([s * np.cos(p), s * np.sin(p), c], -1).reshape(-1, 3)
    w = (wt[:, None] * np.full(40, 2 * math.pi / 40)).ravel()
    y = owv.ylm(pts, 9)
    errs = []
    for eps in (0.2, 0.1):
        xe = x.copy()
        xe[1:] *= eps
        exc, _ = oaug.aug_xc(g, 9, xe)
        dens = xe.T @ y                                   # (NR, NQ)
        eden = oaug.lda3(dens)[0] @ w
        direct = g.integral(eden * r * r)
        errs.append(abs(exc - direct))
    assert errs[0] < 1e-4 and errs[1] < errs[0] / 6.0, errs


def sp_setup(sp):
    d = sp.sphere_setup()
    g = oaug.RadialGrid(d["r1"], d["dex"], d["nr"])
    return oaug.SphereSetup(g, d["aez"], d["rnuc"], d["lox"], d["aephi"], d["psphi"], d["aecore"], d["pscore"],
                            d["rcsm"], d["vadd"], d["dover"], d["dekin"], d["lmrx"])


def rand_denmat(sp, seed):
    rng = np.random.default_rng(seed)
    n = sp.lmnx
    a = rng.normal(0, 0.3, (n, n))
    return a @ a.T / n + np.eye(n) * 0.5


@pytest.mark.parametrize("species", ["si", "sid", "o"])
def test_dath_is_energy_derivative(species):
    sp = {"si": model.si_species(), "sid": model.si_species_d(), "o": model.o_species()}[species]
    st = sp_setup(sp)
    d0 = rand_denmat(sp, 5)
    rng = np.random.default_rng(9)
    vqlm = rng.normal(0, 0.3, st.lmrx)
    # a neutral cell (RHOB = 0): with a background, the derivative of the
    # compensation charge's background energy sits in the plane-wave part's VQLM
    rhob = 0.0

    def energy(d):
        out = oaug.sphere(st, d, vqlm, rhob)
        return out["energies"]["e1c"] + float(vqlm @ out["qlm"]), out

    _, out = energy(d0)
    dath = out["dath"]
    assert np.max(np.abs(dath - dath.T)) < 1e-10 * np.max(np.abs(dath))
    h = 1e-5
    n = d0.shape[0]
    for (i, j) in [(0, 0), (0, 1), (2, 2), (2, 5), (n - 1, n - 1), (1, n - 1)]:
        dp = d0.copy()
        dp[i, j] += h
        dp[j, i] += h if i != j else 0.0
        ep, _ = energy(dp)
        dm = d0.copy()
        dm[i, j] -= h
        dm[j, i] -= h if i != j else 0.0
        em, _ = energy(dm)
        fd = (ep - em) / (2 * h)
        an = dath[i, j] * (2 if i != j else 1)
        assert abs(fd - an) < 2e-6 * max(1.0, abs(an)), (species, i, j, fd, an)


def test_qlm_is_moment_of_one_centre_density():
    sp = model.si_species_d()
    st = sp_setup(sp)
    d = rand_denmat(sp, 3)
    q = oaug.moments(st, d)
    # the same moments from the densities evaluated on an angular quadrature
    g = st.grid
    r = g.r
    ae = oaug.aug_rho(st.lox, st.aephi, d, st.lmrx)
    ps = oaug.aug_rho(st.lox, st.psphi, d, st.lmrx)
    for lm in range(st.lmrx):
        l = oaug.lofm(lm)
        m = g.integral((ae[lm] - ps[lm]) * r ** (l + 2))
        if lm == 0:
            m += g.integral((st.aecore - st.pscore) * r * r) - st.aez * oaug.Y0
        assert abs(q[lm] - m) < 1e-14 * max(1.0, abs(m))
    # the valence charge of the one-centre density: trace of D times the
    # partial-wave overlap differences (completeness of the Gaunt expansion)
    lox = st.lox
    ov = np.zeros((len(lox), len(lox)))
    for a in range(len(lox)):
        for b in range(len(lox)):
            ov[a, b] = g.integral((st.aephi[a] * st.aephi[b] - st.psphi[a] * st.psphi[b]) * r * r)
    charge = float(np.sum(d * oaug.expand_da(lox, ov)))
    core = g.integral((st.aecore - st.pscore) * r * r) - st.aez * oaug.Y0
    assert abs((q[0] - core) / oaug.Y0 - charge) < 1e-10 * max(1.0, abs(charge))


def coupled_waveset(s):
    """The oracle's CP set with the density -> potential step and the
    one-centre terms linked into ETOT (model setups, DO = partial-wave
    overlap differences so that the cell is neutral)."""
    from harness import oracle_waveset
    from oracle import planewave as opw
    o = oracle_waveset(s)
    dpl = opw.Plane(s.rbas, s.epw_rho, s.epw_rho)
    t = s.potential_tables(dpl.g2)
    o.potential = (dpl, {"pscorg": t[0], "vbarg": t[1], "g0": t[2], "v0": t[3]}, s.isp, None)
    o.sphere = [sp_setup(sp) for sp in s.species]
    return o


def test_coupled_hpsi_and_forces_are_energy_derivatives():
    """With MOMENTS -> VOFRHO -> SPHERE -> ADDCONSTANTPOT in ETOT: HPSI is
    the psi-gradient and FORCE minus the position gradient of ETOT.  The
    constant shift POTB (V + POTB, DH + POTB DO) is part of the gradient: it
    carries the one-centre energy's dependence on the background RHOB, which
    follows the total charge of the trial wave functions."""
    s = model.paw_system(model.tiny())
    s.occ = np.array([2.0, 2.0, 2.0, 1.0, 1.0, 0.0, 0.0, 0.0])      # 8 electrons: a neutral cell
    o = coupled_waveset(s)
    o.etot(s.pos)                                  # Gram-Schmidt start
    assert abs(o.pot["rhob"]) < 1e-6
    psi = o.psi0.copy()

    def e_at(p, r):
        o.psi0 = p
        o.proj = owv.projections(o.plane, o.setup, o.pro, o.ylm, r, p)
        o.rlam0 = np.zeros((o.nb, o.nb))
        return o.etot(r)["etot"]

    e_at(psi, s.pos)
    hpsi, force = o.hpsi.copy(), o.force.copy()
    assert abs(o.potb) > 1e-3                      # the shift is exercised
    rng = np.random.default_rng(21)
    d = np.zeros_like(psi)
    d[:2] = (rng.normal(size=(2, o.plane.ngl)) + 1j * rng.normal(size=(2, o.plane.ngl))) / (1 + o.plane.g2) ** 2
    d[:2] = 0.5 * (d[:2] + np.conj(d[:2][:, o.plane.minusg]))
    an = 2.0 * o.occ[0] * o.plane.vol * float(np.sum(np.conj(d[:2]) * hpsi[:2]).real)

    def cd(eps):
        return (e_at(psi + eps * d, s.pos) - e_at(psi - eps * d, s.pos)) / (2 * eps)

    fd = (4.0 * cd(5e-4) - cd(1e-3)) / 3.0
    # the radial Poisson solver and integrals are not exactly adjoint on the
    # grid (4th-order rules), which bounds the agreement at a few 1e-7
    assert abs(fd - an) < 2e-6 * abs(an), (fd, an)
    h = 1e-4
    for a in range(s.nat):
        for k in range(3):
            rp = s.pos.copy()
            rp[a, k] += h
            rm = s.pos.copy()
            rm[a, k] -= h
            fdk = -(e_at(psi, rp) - e_at(psi, rm)) / (2 * h)
            assert abs(fdk - force[a, k]) < 2e-6 * max(1.0, np.max(np.abs(force))), (a, k, fdk, force[a, k])
