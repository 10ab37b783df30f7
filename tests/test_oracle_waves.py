"""Pins the oracle's WAVES restatement (oracle/waves.py) without the reference:
every super-wave-function formula is checked against an independent
restatement on the unpacked real bands psi = phi_1 + i phi_2
(src/paw_fft.f90:1698-1737), plus closed forms, brute-force convolution,
finite-difference forces and the constraint invariants.  CPU only."""
import numpy as np
import pytest

from oracle import planewave as opw
from oracle import waves as owv
from pwgpu import model

from harness import oracle_waveset


@pytest.fixture(scope="module")
def tiny():
    s = model.tiny()
    o = oracle_waveset(s)
    return s, o


def unsuper(pl, psi):
    """Real bands phi(2b), phi(2b+1) of super bands psi(b)."""
    pm = opw.invertg(pl.minusg, psi)
    out = np.empty((2 * psi.shape[0], psi.shape[1]), dtype=complex)
    out[0::2] = 0.5 * (psi + pm)
    out[1::2] = -0.5j * (psi - pm)
    return out


def resuper(phi):
    return phi[0::2] + 1j * phi[1::2]


def rand_super(pl, nbh, seed):
    rng = np.random.default_rng(seed)
    c = rng.normal(size=(nbh, pl.ngl)) + 1j * rng.normal(size=(nbh, pl.ngl))
    return c / (1.0 + pl.g2[None, :]) ** 2


def test_unpacked_bands_are_real(tiny):
    s, o = tiny
    phi = unsuper(o.plane, rand_super(o.plane, 4, 1))
    assert np.max(np.abs(o.plane.fft_gtor(phi).imag)) < 1e-12
    assert np.allclose(resuper(phi), rand_super(o.plane, 4, 1))


def test_overlap_unravel_matches_real_bands(tiny):
    s, o = tiny
    pl = o.plane
    a, b = rand_super(pl, 4, 2), rand_super(pl, 4, 3)
    pa, pb = unsuper(pl, a), unsuper(pl, b)
    direct = pl.vol * (np.conj(pa) @ pb.T)
    assert np.max(np.abs(direct.imag)) < 1e-12
    assert np.max(np.abs(owv.overlap(pl, 8, a, b) - direct.real)) < 1e-12 * np.max(np.abs(direct))


def test_overlap_1c_and_denmat_match_real_bands(tiny):
    s, o = tiny
    pl, st = o.plane, o.setup
    a, b = rand_super(pl, 4, 4), rand_super(pl, 4, 5)
    pa = owv.projections(pl, st, o.pro, o.ylm, s.pos, a)
    pb = owv.projections(pl, st, o.pro, o.ylm, s.pos, b)
    ra = owv.projections(pl, st, o.pro, o.ylm, s.pos, unsuper(pl, a))
    rb = owv.projections(pl, st, o.pro, o.ylm, s.pos, unsuper(pl, b))
    assert np.max(np.abs(ra.imag)) < 1e-12            # <p|phi> of real functions is real
    do = np.zeros((st.npro, st.npro))
    for iat in range(st.nat):
        i0, n = st.ipro[iat], st.lmnx[st.isp[iat]]
        do[i0:i0 + n, i0:i0 + n] = st.do_lmn(st.isp[iat])
    direct = np.conj(ra).T @ do @ rb
    assert np.max(np.abs(owv.overlap_1c(st, 8, pa, pb) - direct.real)) < 1e-12 * np.max(np.abs(direct))
    dm = owv.denmat(st, o.occ, pa)
    for iat in range(st.nat):
        i0, n = st.ipro[iat], st.lmnx[st.isp[iat]]
        p = ra[i0:i0 + n].real
        want = (p * o.occ[None, :]) @ p.T
        assert np.max(np.abs(dm[iat] - want)) < 1e-12 * np.max(np.abs(want))


def test_density_ekin_wavekinetic_match_real_bands(tiny):
    s, o = tiny
    pl = o.plane
    psi = rand_super(pl, 4, 6)
    phi = unsuper(pl, psi)
    rho = np.einsum("b,bzyx->zyx", o.occ, np.abs(pl.fft_gtor(phi)) ** 2)
    assert np.max(np.abs(owv.density(pl, o.occ, psi) - rho)) < 1e-12 * np.max(rho)
    ek = pl.vol * 0.5 * np.sum(o.occ[:, None] * pl.g2[None, :] * np.abs(phi) ** 2)
    assert abs(owv.ekin(pl, o.occ, psi) - ek) < 1e-12 * ek
    psim = rand_super(pl, 4, 7)
    d = unsuper(pl, psi - psim)
    wk = pl.vol * np.sum(o.occ[:, None] * o.mpsi[None, :] * np.abs(d) ** 2) / o.delt ** 2
    assert abs(owv.wavekinetic(pl, o.occ, o.mpsi, o.delt, psi, psim) - wk) < 1e-12 * wk


def test_addopsi_addoproj_match_real_bands(tiny):
    s, o = tiny
    pl, st = o.plane, o.setup
    psibar, chi = rand_super(pl, 4, 8), rand_super(pl, 4, 9)
    lam = np.random.default_rng(10).normal(size=(8, 8))
    want = unsuper(pl, psibar) + lam.T @ unsuper(pl, chi)      # phi'_J = phibar_J + sum_I chi_I L_IJ
    got = owv.addopsi(pl, psibar, chi, lam)
    assert np.max(np.abs(got - resuper(want))) < 1e-12 * np.max(np.abs(want))
    pj = owv.projections(pl, st, o.pro, o.ylm, s.pos, psibar)
    oj = owv.projections(pl, st, o.pro, o.ylm, s.pos, chi)
    gotp = owv.addoproj(pj, oj, lam)
    wantp = owv.projections(pl, st, o.pro, o.ylm, s.pos, got)
    assert np.max(np.abs(gotp - wantp)) < 1e-11 * np.max(np.abs(wantp))


def test_vpsi_constant_potential_and_convolution(tiny):
    s, o = tiny
    pl = o.plane
    psi = rand_super(pl, 4, 11)
    v0 = -0.731
    h = owv.vpsi(pl, np.full((pl.nr3, pl.nr2, pl.nr1), v0), psi)
    assert np.max(np.abs(h - (v0 + 0.5 * pl.g2[None, :]) * psi)) < 1e-13
    vhat = np.fft.fftn(o.v) / pl.nrl
    n = np.array([pl.nr1, pl.nr2, pl.nr3])
    conv = np.zeros_like(psi)
    for g in range(pl.ngl):
        d = np.mod(pl.igvec[g][None, :] - pl.igvec, n)
        conv[:, g] = psi @ vhat[d[:, 2], d[:, 1], d[:, 0]]
    h = owv.vpsi(pl, o.v, psi) - 0.5 * pl.g2[None, :] * psi
    assert np.max(np.abs(h - conv)) < 1e-12 * np.max(np.abs(conv))


def test_projection_of_plane_wave_closed_form(tiny):
    s, o = tiny
    pl, st = o.plane, o.setup
    g0 = 11
    psi = np.zeros((1, pl.ngl), dtype=complex)
    psi[0, g0] = 1.0
    proj = owv.projections(pl, st, o.pro, o.ylm, s.pos, psi)
    gv = pl.gvec[g0]
    gl = np.linalg.norm(gv)
    c1 = np.sqrt(3.0 / (4 * np.pi))
    ylm = [1 / np.sqrt(4 * np.pi), c1 * gv[0] / gl, c1 * gv[2] / gl, c1 * gv[1] / gl]
    tab = s.pro_table(np.array([gl * gl]))[:, 0]
    for iat in range(st.nat):
        eig = np.exp(-1j * gv @ s.pos[iat])
        lmn = 0
        for ln, l in enumerate(st.lox[0]):
            for m in range(2 * l + 1):
                p = ylm[l * l + m] * tab[ln] * (-1j) ** l * eig
                assert abs(proj[st.ipro[iat] + lmn, 0] - pl.vol * np.conj(p)) < 1e-13
                lmn += 1


def test_projection_addpro_adjoint(tiny):
    s, o = tiny
    pl, st = o.plane, o.setup
    psi = rand_super(pl, 4, 12)
    x = np.random.default_rng(13).normal(size=(st.npro, 4)) + 1j * np.random.default_rng(14).normal(size=(st.npro, 4))
    lhs = np.sum(np.conj(x) * owv.projections(pl, st, o.pro, o.ylm, s.pos, psi))
    a = owv.addpro(pl, st, o.pro, o.ylm, s.pos, np.zeros_like(psi), x)
    rhs = pl.vol * np.sum(np.conj(a) * psi)
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)


def test_model_projector_table_bessel_transform():
    from scipy import integrate, special
    s = model.tiny()
    vol = s.vol
    for gl in (0.0, 0.7, 2.3):
        tab = s.pro_table(np.array([gl * gl]))[:, 0]
        for ln, (l, sg) in enumerate(zip(s.species[0].lox, s.species[0].sigma)):
            nrm = 1.0 / np.sqrt(integrate.quad(lambda r: (r ** l * np.exp(-r * r / (2 * sg * sg))) ** 2 * r * r, 0, 40)[0])
            val = integrate.quad(lambda r: r * r * special.spherical_jn(l, gl * r) * nrm * r ** l
                                 * np.exp(-r * r / (2 * sg * sg)), 0, 40)[0]
            assert abs(tab[ln] - 4 * np.pi * val / vol) < 1e-9 * max(1.0, abs(tab[ln]))


def _lagrangian(o, s, psi, r, eps):
    """E_1c - sum_IJ eps_IJ (f_I+f_J)/2 <phi_I|(O-1)|phi_J> on real bands."""
    pl, st = o.plane, o.setup
    ra = owv.projections(pl, st, o.pro, o.ylm, r, unsuper(pl, psi))
    e = 0.0
    do = np.zeros((st.npro, st.npro), dtype=complex)
    dh = np.zeros((st.npro, st.npro), dtype=complex)
    for iat in range(st.nat):
        i0, n = st.ipro[iat], st.lmnx[st.isp[iat]]
        do[i0:i0 + n, i0:i0 + n] = st.do_lmn(st.isp[iat])
        dh[i0:i0 + n, i0:i0 + n] = o.dh[iat]
    h = np.conj(ra).T @ dh @ ra
    ov = np.conj(ra).T @ do @ ra
    f = o.occ
    e = np.sum(f * np.diag(h)).real
    e -= np.sum(eps * 0.5 * (f[:, None] + f[None, :]) * ov).real
    return e


@pytest.mark.parametrize("with_lambda", [False, True])
def test_force_finite_difference(tiny, with_lambda):
    s, o = tiny
    psi = rand_super(o.plane, 4, 15)
    rng = np.random.default_rng(16)
    eps = np.zeros((8, 8))
    if with_lambda:
        eps = rng.normal(size=(8, 8))
        eps = 0.5 * (eps + eps.T)
    proj = owv.projections(o.plane, o.setup, o.pro, o.ylm, s.pos, psi)
    f = owv.force(o.plane, o.setup, o.pro, o.ylm, s.pos, o.occ, psi, proj, o.dh, eps)
    h = 1e-4
    for iat in range(s.nat):
        for k in range(3):
            rp, rm = s.pos.copy(), s.pos.copy()
            rp[iat, k] += h
            rm[iat, k] -= h
            fd = -(_lagrangian(o, s, psi, rp, eps) - _lagrangian(o, s, psi, rm, eps)) / (2 * h)
            assert abs(f[iat, k] - fd) < 1e-6 * max(1.0, np.max(np.abs(f))), (iat, k, f[iat, k], fd)


def test_step_energy_consistency_and_invariants():
    s = model.tiny()
    o = oracle_waveset(s)
    conv = []
    for it in range(4):
        e = o.etot(s.pos)
        # sum_n f_n <psi_n|H|psi_n> equals EKIN + int V rho + sum DH D
        assert abs(np.dot(o.occ, o.eig) - e["etot"]) < 1e-11 * abs(e["etot"])
        o.propagate()
        o.orthogonalize(s.pos, s.pos)
        prj = owv.projections(o.plane, o.setup, o.pro, o.ylm, s.pos, o.psim)
        ov = owv.overlap(o.plane, s.nb, o.psim, o.psim) + owv.overlap_1c(o.setup, s.nb, prj, prj)
        assert np.max(np.abs(ov - np.eye(s.nb))) < 1e-8
        assert np.max(np.abs(prj - o.proj)) < 1e-12 * np.max(np.abs(prj))
        lam = o.rlam0
        oi = o.occ + owv.DSMALL
        assert np.max(np.abs(lam * oi[None, :] - lam.T * oi[:, None])) < 1e-12 * np.max(np.abs(lam))
        conv.append(lam.copy())
        o.switch()
        # WAVES$SWITCH with NLAMBDAX=2: Lambda(+) = 2 Lambda(0) - Lambda(-)
        want = conv[-1] if it == 0 else 2 * conv[-1] - conv[-2]
        assert np.max(np.abs(o.rlam0 - want)) < 1e-15 * np.max(np.abs(want)) + 1e-300
    assert o.ortho_iter >= 1


def test_ortho_x_against_newton_solve():
    s = model.tiny()
    o = oracle_waveset(s)
    o.etot(s.pos)
    o.propagate()
    cap = {}
    orig = owv.ortho_x

    def spy(occ, chichi, psipsi, chipsi, lam, **kw):
        cap.update(occ=occ, c=chichi, a=psipsi, b=chipsi, lam0=np.array(lam))
        return orig(occ, chichi, psipsi, chipsi, lam, **kw)

    owv.ortho_x = spy
    try:
        o.orthogonalize(s.pos, s.pos)
    finally:
        owv.ortho_x = orig
    occ, nb = cap["occ"], len(cap["occ"])
    ds = owv.DSMALL
    # free unknowns: Lambda(i,j) with f_i > f_j or (f_i == f_j and i <= j)
    free = [(i, j) for i in range(nb) for j in range(nb)
            if occ[i] > occ[j] or (occ[i] == occ[j] and i <= j)]

    def build(x):
        lam = np.zeros((nb, nb))
        for (i, j), v in zip(free, x):
            lam[i, j] = v
            if occ[i] == occ[j]:
                lam[j, i] = v
            else:
                lam[j, i] = v * (occ[j] + ds) / (occ[i] + ds)
        return lam

    def resid(x):
        lam = build(x)
        g = lam.T @ (cap["c"] @ lam + 2 * cap["b"])
        g = 0.5 * (g + g.T) + np.triu(cap["a"]) + np.triu(cap["a"], 1).T - np.eye(nb)
        return g[np.triu_indices(nb)]

    x = np.array([cap["lam0"][i, j] for i, j in free])
    for _ in range(30):
        r = resid(x)
        if np.max(np.abs(r)) < 1e-14:
            break
        jac = np.empty((len(r), len(x)))
        for k in range(len(x)):
            dx = np.zeros_like(x)
            dx[k] = 1e-7
            jac[:, k] = (resid(x + dx) - resid(x - dx)) / 2e-7
        x = x - np.linalg.lstsq(jac, r, rcond=None)[0]
    assert np.max(np.abs(resid(x))) < 1e-11
    assert np.max(np.abs(build(x) - o.rlam0)) < 1e-5 * np.max(np.abs(o.rlam0))


def test_gramschmidt_orthonormal_and_triangular(tiny):
    s, o = tiny
    pl, st = o.plane, o.setup
    psi = rand_super(pl, 4, 17)
    new = owv.gramschmidt(pl, st, o.pro, o.ylm, s.pos, 8, psi)
    prj = owv.projections(pl, st, o.pro, o.ylm, s.pos, new)
    ov = owv.overlap(pl, 8, new, new) + owv.overlap_1c(st, 8, prj, prj)
    assert np.max(np.abs(ov - np.eye(8))) < 1e-12
    a, b = unsuper(pl, psi), unsuper(pl, new)
    t = np.linalg.lstsq(a.T, b.T, rcond=None)[0]          # new_J = sum_I old_I T_IJ
    assert np.max(np.abs(np.tril(t, -1))) < 1e-10 * np.max(np.abs(t))
    assert np.all(np.diag(t).real > 0)


def test_cp_energy_conservation_small_timestep():
    s = model.tiny()
    s.dt = 2.0
    s.emasscg2 = 0.5 * (10.0 / (2 * np.pi)) ** 2 * s.dt ** 2 / s.emass
    o = oracle_waveset(s)
    cons = []
    for it in range(25):
        e = o.etot(s.pos)
        o.propagate()
        o.orthogonalize(s.pos, s.pos)
        cons.append(e["etot"] + o.fictitious_ekin())
        o.switch()
    cons = np.array(cons[1:])
    assert np.max(cons) - np.min(cons) < 2e-3 * abs(cons[0])


def _ylm_closed_forms(g):
    """The real harmonics listed in the header of SPHERICAL$YLM
    (src/paw_spherical.f90:70-89), l <= 3, as explicit polynomials."""
    x, y, z = g[:, 0], g[:, 1], g[:, 2]
    r = np.sqrt(x * x + y * y + z * z)
    x, y, z = x / r, y / r, z / r
    pi = np.pi
    return np.array([
        np.full_like(x, np.sqrt(1 / (4 * pi))),
        np.sqrt(3 / (4 * pi)) * x, np.sqrt(3 / (4 * pi)) * z, np.sqrt(3 / (4 * pi)) * y,
        np.sqrt(15 / (16 * pi)) * (x * x - y * y), np.sqrt(60 / (16 * pi)) * x * z,
        np.sqrt(5 / (16 * pi)) * (3 * z * z - 1), np.sqrt(60 / (16 * pi)) * y * z,
        np.sqrt(60 / (16 * pi)) * x * y,
        np.sqrt(35 / (32 * pi)) * (x ** 3 - 3 * x * y * y), np.sqrt(210 / (32 * pi)) * z * (x * x - y * y),
        np.sqrt(21 / (32 * pi)) * x * (5 * z * z - 1), np.sqrt(14 / (32 * pi)) * z * (5 * z * z - 3),
        np.sqrt(21 / (32 * pi)) * y * (5 * z * z - 1), np.sqrt(840 / (32 * pi)) * x * y * z,
        np.sqrt(35 / (32 * pi)) * y * (3 * x * x - y * y)])


def test_ylm_closed_forms_and_orthonormality():
    rng = np.random.default_rng(4)
    g = rng.normal(size=(200, 3))
    g[:3] = [[0.0, 0.0, 2.0], [0.0, 0.0, -1.0], [1.5, 0.0, 0.0]]    # poles and an axis
    got = owv.ylm(g, 16)
    assert np.max(np.abs(got - _ylm_closed_forms(g))) < 1e-13
    # G = 0: only the l = 0 harmonic (src/paw_spherical.f90:117-123)
    z = owv.ylm(np.zeros((1, 3)), 16)
    assert abs(z[0, 0] - 1 / np.sqrt(4 * np.pi)) < 1e-16 and np.all(z[1:, 0] == 0.0)
    # orthonormal on the sphere (Gauss-Legendre in cos(theta) x uniform phi,
    # exact for these polynomial degrees), l up to 4
    ct, wt = np.polynomial.legendre.leggauss(12)
    phi = 2 * np.pi * np.arange(24) / 24
    c, p = np.meshgrid(ct, phi, indexing="ij")
    s = np.sqrt(1 - c * c)
    pts = np.stack([s * np.cos(p), s * np.sin(p), c], axis=-1).reshape(-1, 3)
    w = (wt[:, None] * np.full(24, 2 * np.pi / 24)[None, :]).ravel()
    y = owv.ylm(pts, 25)
    assert np.max(np.abs((y * w) @ y.T - np.eye(25))) < 1e-13


def test_propagate_friction_damped_oscillator():
    """WAVES$PROPAGATE with friction (src/paw_waves1.f90:6553-6583) on a
    harmonic Hamiltonian H psi = k m(G) psi: the coefficients follow the
    recurrence (1+a) x(n+1) = (2 - dt^2 k) x(n) - (1-a) x(n-1), solved in
    closed form by the roots of its characteristic polynomial, and for small
    dt the damped oscillator x'' = -k x - (2a/dt) x' (amplitude decaying as
    exp(-a t/dt), frequency sqrt(k - (a/dt)^2))."""
    s = model.tiny()
    o = oracle_waveset(s)
    pl = o.plane
    rng = np.random.default_rng(2)
    x0 = (rng.normal(size=(2, pl.ngl)) + 1j * rng.normal(size=(2, pl.ngl))).astype(complex)
    for annee, dt, k, nstep, tol in ((0.02, 5.0, 2.0e-3, 60, 1e-12), (0.0, 5.0, 2.0e-3, 60, 1e-12),
                                     (0.004, 0.5, 1.0e-2, 1000, 3e-3)):
        o.annee, o.delt = annee, dt
        o.psim, o.psi0 = x0[0].copy()[None], x0[1].copy()[None]
        traj = [o.psim[0].copy(), o.psi0[0].copy()]
        for n in range(nstep):
            o.hpsi = k * o.mpsi[None, :] * o.psi0
            o.propagate()
            o.psi0, o.psim = o.psim, o.psi0
            traj.append(o.psi0[0].copy())
        roots = np.roots([1 + annee, -(2 - dt * dt * k), 1 - annee])
        z1, z2 = roots
        # x(n) = A z1^n + B z2^n through x(0) = psim, x(1) = psi0
        b = (x0[1] - z1 * x0[0]) / (z2 - z1)
        a = x0[0] - b
        n = nstep + 1
        want = a * z1 ** n + b * z2 ** n
        if tol < 1e-6:
            assert np.max(np.abs(traj[-1] - want)) < tol * np.max(np.abs(x0))
        # continuum limit: decay rate gamma/2 = a/dt, frequency sqrt(k - (a/dt)^2)
        gam = annee / dt
        w = np.sqrt(k - gam * gam)
        t = n * dt
        v0 = (x0[1] - x0[0]) / dt
        xc = np.exp(-gam * t) * (x0[0] * np.cos(w * t) + (v0 + gam * x0[0]) / w * np.sin(w * t))
        if tol >= 1e-6:
            assert np.max(np.abs(traj[-1] - xc)) < tol * np.max(np.abs(x0))
        # the product of the roots is the per-step damping (1-a)/(1+a)
        assert abs(z1 * z2 - (1 - annee) / (1 + annee)) < 1e-14


def indefinite_ortho_case(nb=12, nempty=0, seed=21):
    """ORTHO_X inputs whose sym(CHIPSI) has negative eigenvalues (none of the
    pair sums e_i + e_j near zero): the Cayley series cannot apply S^-1, the
    reference's eigenvector form can."""
    rng = np.random.default_rng(seed)
    q, _ = np.linalg.qr(rng.normal(size=(nb, nb)))
    e = np.concatenate([rng.uniform(-0.3, -0.2, 4), rng.uniform(0.8, 1.2, nb - 4)])
    b = (q * e) @ q.T
    anti = 0.01 * rng.normal(size=(nb, nb))
    chipsi = b + (anti - anti.T)
    m = rng.normal(size=(nb, nb))
    chichi = 0.02 * (m @ m.T) / nb
    d = 0.01 * rng.normal(size=(nb, nb))
    psipsi = np.eye(nb) + 0.5 * (d + d.T)
    occ = np.concatenate([np.full(nb - nempty, 2.0), np.zeros(nempty)])
    return occ, chichi, psipsi, chipsi


def test_ortho_x_indefinite_chipsi():
    """With an indefinite sym(CHIPSI) the iteration still reaches the
    constraint: psi(+) = psibar + chi Lambda is orthonormal, i.e.
    PSIPSI + Lambda^T CHIPSI + CHIPSI^T Lambda + Lambda^T CHICHI Lambda = 1
    (symmetrised, src/paw_waves2.f90:1905-1930), and Lambda keeps the
    occupation relation Lambda_ij f_j = Lambda_ji f_i."""
    occ, chichi, psipsi, chipsi = indefinite_ortho_case()
    assert np.min(np.linalg.eigvalsh(0.5 * (chipsi + chipsi.T))) < -0.1
    lam, it = owv.ortho_x(occ, chichi, psipsi, chipsi, np.zeros_like(chichi))
    assert 2 <= it < 50
    g = psipsi + lam.T @ chipsi + chipsi.T @ lam + lam.T @ chichi @ lam
    g = 0.5 * (g + g.T) - np.eye(len(occ))
    assert np.max(np.abs(g)) < 1e-8
    oi = occ + owv.DSMALL
    assert np.max(np.abs(lam * oi[None, :] - lam.T * oi[:, None])) < 1e-10 * np.max(np.abs(lam))
