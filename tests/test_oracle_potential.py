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


# ---------------------------------------------------------------------------
# functional type 10 (PBE): limits, scaling and derivatives
# ---------------------------------------------------------------------------
def test_pbe_exchange_scaling_and_limits():
    rng = np.random.default_rng(11)
    rho = 10.0 ** rng.uniform(-3, 1, size=40)
    g2 = rho ** (8.0 / 3.0) * 10.0 ** rng.uniform(-3, 2, size=40)
    ex, _, _ = opot.pbe_exchange(rho, g2)
    # uniform scaling rho -> lam^3 rho(lam r): |grad rho|^2 -> lam^8, e_x -> lam^4
    for lam in (0.5, 1.7, 3.0):
        exl, _, _ = opot.pbe_exchange(lam ** 3 * rho, lam ** 8 * g2)
        assert np.allclose(exl, lam ** 4 * ex, rtol=1e-13)
    # no gradient: the local exchange; s^2 = |grad rho|^2 / (4 kF^2 rho^2), kF = (3 pi^2 rho)^(1/3)
    ex0, vx0, _ = opot.pbe_exchange(rho, 0.0 * g2)
    exl, vxl = opot.exchange_lda(rho)
    assert np.allclose(ex0, exl, rtol=1e-14) and np.allclose(vx0, vxl, rtol=1e-13)
    kf = (3.0 * math.pi ** 2 * rho) ** (1.0 / 3.0)
    s2 = g2 / (4.0 * kf ** 2 * rho ** 2)
    fx = ex / exl
    # PBE enhancement: 1 + mu s^2 for small s, 1 + kappa (Lieb-Oxford) for large s
    assert np.allclose(fx, 1.0 + opot.PBE_KAPPA - opot.PBE_KAPPA / (1.0 + opot.PBE_MU * s2 / opot.PBE_KAPPA),
                       rtol=1e-13)
    small = opot.pbe_exchange(rho, 1e-8 * rho ** (8.0 / 3.0))[0] / exl - 1.0
    assert np.allclose(small, opot.PBE_MU * 1e-8 / (4.0 * (3.0 * math.pi ** 2) ** (2.0 / 3.0)), rtol=1e-6)
    big = opot.pbe_exchange(rho, 1e12 * rho ** (8.0 / 3.0))[0] / exl
    assert np.allclose(big, 1.0 + opot.PBE_KAPPA, rtol=1e-9)
    # mu is fixed by the correlation's gradient coefficient: mu = beta pi^2 / 3
    assert abs(opot.PBE_MU - opot.PBE_BETA * math.pi ** 2 / 3.0) < 1e-15


def test_pw92_limits_and_published_relations():
    # high density: eps_c -> A ln rs + const with A = (1 - ln 2)/pi^2
    # (Gell-Mann-Brueckner), and beta1 = exp(-c1/(2A))/(2A) with c1 = 0.046644
    assert abs(opot.PW_A - (1.0 - math.log(2.0)) / math.pi ** 2) < 1e-7
    assert abs(math.exp(-0.046644 / (2 * opot.PW_A)) / (2 * opot.PW_A) - opot.PW_BETA[0]) < 2e-3
    rs = np.array([1e-4, 2e-4])
    e, _ = opot.pw92_unpolarised(3.0 / (4.0 * math.pi * rs ** 3))
    assert abs((e[1] - e[0]) / math.log(2.0) - opot.PW_A) < 2e-3 * opot.PW_A
    # the same Ceperley-Alder energies as the Perdew-Zunger fit, to 1 mHa
    for rs1 in (1.0, 2.0, 5.0, 10.0):
        rho = np.array([3.0 / (4.0 * math.pi * rs1 ** 3)])
        epw = opot.pw92_unpolarised(rho)[0][0]
        epz = opot.pz_unpolarised(rho)[0][0] / rho[0]
        assert abs(epw - epz) < 1e-3, (rs1, epw, epz)


def test_pbe_correlation_gradient_limits():
    rho = np.array([1e-3, 0.02, 0.3, 2.0])
    eps, _ = opot.pw92_unpolarised(rho)
    tn = opot.PBE_TFAC / rho ** (7.0 / 3.0)
    # t = 0: the local correlation
    ec0, _, _ = opot.pbe_correlation(rho, 0.0 * rho)
    assert np.allclose(ec0, rho * eps, rtol=1e-14)
    # small t: H = beta t^2 (second-order gradient expansion)
    t2 = 1e-7
    ec, _, _ = opot.pbe_correlation(rho, t2 / tn)
    assert np.allclose((ec / rho - eps) / t2, opot.PBE_BETA, rtol=1e-4)
    # large t: H -> -eps_c (the correlation energy vanishes)
    ec, _, _ = opot.pbe_correlation(rho, 1e8 / tn)
    assert np.all(np.abs(ec / rho) < 1e-6 * np.abs(eps))
    # t^2 = |grad rho|^2 / (4 phi^2 ks^2 rho^2), ks^2 = 4 kF / pi
    kf = (3.0 * math.pi ** 2 * rho) ** (1.0 / 3.0)
    assert np.allclose(tn, 1.0 / (4.0 * (4.0 * kf / math.pi) * rho ** 2), rtol=1e-13)


@pytest.mark.parametrize("which", ["x", "c", "xc"])
def test_pbe_derivatives(which):
    rng = np.random.default_rng(5)
    rho = 10.0 ** rng.uniform(-5.5, 1, size=30)          # includes points below RHOTMIN
    g2 = rho ** (8.0 / 3.0) * 10.0 ** rng.uniform(-4, 3, size=30)
    f = {"x": opot.pbe_exchange, "c": opot.pbe_correlation, "xc": opot.pbe}[which]
    e, v, gv = f(rho, g2)
    h = 1e-6 * rho
    fd_r = (f(rho + h, g2)[0] - f(rho - h, g2)[0]) / (2 * h)
    fd_g = (f(rho, g2 * (1 + 1e-6))[0] - f(rho, g2 * (1 - 1e-6))[0]) / (2e-6 * g2)
    # the floor at RHOTMIN makes E piecewise in rho: skip points within a step of it
    ok = np.abs(rho - opot.RHOTMIN) > 3e-6 * rho if which == "xc" else np.ones(rho.shape, bool)
    # central differences of E carry ~1e-10 |E| of rounding per unit step
    assert np.all(np.abs(v - fd_r)[ok] * rho[ok] <= 1e-9 * np.abs(e[ok]) + 2e-6 * np.abs(v * rho)[ok]), which
    assert np.all(np.abs(gv - fd_g) * g2 <= 1e-9 * np.abs(e) + 2e-6 * np.abs(gv * g2)), which


def energy_f(pl, rho, pos, isp, qlm, tabs, functional):
    r = opot.vofrho(pl, rho, pos, isp, qlm, tabs, functional=functional)
    return r["ehartree"] + r["exc"], r


def test_pbe_potential_is_density_derivative():
    # the gradient term -div(2 dE/d|grad rho|^2 grad rho) of the returned V:
    # FD of EXC + EHARTREE along a smooth density change
    s, pl, rho, tabs = system_case()
    e0, r = energy_f(pl, rho, s.pos, s.isp, s.qlm(), tabs, 10)
    e_lda, _ = energy_f(pl, rho, s.pos, s.isp, s.qlm(), tabs, 1)
    assert abs(e0 - e_lda) > 1e-4       # the gradient correction matters here
    rng = np.random.default_rng(8)
    c = (rng.normal(size=pl.ngl) + 1j * rng.normal(size=pl.ngl)) * np.exp(-pl.g2)
    c = 0.5 * (c + np.conj(c[pl.minusg]))
    d = pl.fft_gtor(c).real
    d *= 1e-3 / np.max(np.abs(d))
    eps = 1e-3
    ep, _ = energy_f(pl, rho + eps * d, s.pos, s.isp, s.qlm(), tabs, 10)
    em, _ = energy_f(pl, rho - eps * d, s.pos, s.isp, s.qlm(), tabs, 10)
    fd = (ep - em) / (2 * eps)
    an = float(np.sum(r["v"] * d)) * pl.rweight
    assert abs(fd - an) < 1e-7 * abs(an), (fd, an)


def test_pbe_forces_are_position_derivatives():
    s, pl, rho, tabs = system_case()
    _, r = energy_f(pl, rho, s.pos, s.isp, s.qlm(), tabs, 10)
    h = 1e-4
    for a, k in ((0, 0), (1, 2), (3, 1)):
        pp = np.array(s.pos, dtype=float)
        pm = pp.copy()
        pp[a, k] += h
        pm[a, k] -= h
        ep, _ = energy_f(pl, rho, pp, s.isp, s.qlm(), tabs, 10)
        em, _ = energy_f(pl, rho, pm, s.isp, s.qlm(), tabs, 10)
        fd = -(ep - em) / (2 * h)
        assert abs(fd - r["force"][a, k]) < 1e-6 * max(1.0, abs(fd)), (a, k, fd, r["force"][a, k])
