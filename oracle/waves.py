"""Oracle restatement of the WAVES object hot path (cp-paw src/paw_waves1.f90,
src/paw_waves2.f90) for NDIM=1, NSPIN=1, Gamma point with super wave
functions (TINV=.TRUE.).

Test infrastructure only; see oracle/__init__.py for the layout conventions.
Small matrices (NB x NB, LMNX x LMNX) use the natural numpy convention
m[i, j] = M(I+1, J+1); coefficient arrays use the reversed (Fortran memory)
layout psi[nbh, ngl], proj[npro, nbh].
"""
import numpy as np

from .planewave import invertg
from .potential import vofrho

PI = 4.0 * np.arctan(1.0)
CI = 1j
DSMALL = 1.0e-12


# ----------------------------------------------------------------------------
# set-up tables
# ----------------------------------------------------------------------------
class Setup:
    """MAP_TYPE + the per-species data WAVES reads from SETUP.

    src/paw_waves1.f90:142-154 (MAP_TYPE), :1218-1247 (LMNX, NPRO).
    isp: (NAT,) 0-based species; lox: list per species of angular momenta;
    do_ln: list per species of (LNX,LNX) one-centre overlaps DO.
    """

    def __init__(self, isp, lox, do_ln):
        self.isp = np.asarray(isp, dtype=np.int64)
        self.nat = len(self.isp)
        self.lox = [list(map(int, l)) for l in lox]
        self.nsp = len(self.lox)
        self.lnx = [len(l) for l in self.lox]
        self.lmnx = [sum(2 * l + 1 for l in lo) for lo in self.lox]
        self.do_ln = [np.asarray(d, dtype=np.float64) for d in do_ln]
        self.npro = int(sum(self.lmnx[s] for s in self.isp))
        self.lmx = max((max(lo) + 1) ** 2 for lo in self.lox)
        self.ipro = np.zeros(self.nat, dtype=np.int64)     # first projector of atom
        acc = 0
        for iat in range(self.nat):
            self.ipro[iat] = acc
            acc += self.lmnx[self.isp[iat]]
        self.ibpro = np.cumsum([0] + self.lnx)[:-1]       # first bare projector
        self.nbarepro = int(sum(self.lnx))

    def do_lmn(self, isp):
        """Expand DO(LN,LN) to LMN x LMN with delta(L,L')delta(M,M').

        src/paw_waves2.f90:497-542 (WAVES_OPROJ)."""
        lox = self.lox[isp]
        n = self.lmnx[isp]
        out = np.zeros((n, n))
        o1 = 0
        for ln1, l1 in enumerate(lox):
            o2 = 0
            for ln2, l2 in enumerate(lox):
                if l1 == l2:
                    for m in range(2 * l1 + 1):
                        out[o1 + m, o2 + m] = self.do_ln[isp][ln1, ln2]
                o2 += 2 * l2 + 1
            o1 += 2 * l1 + 1
        return out


def ylm(gvec, lmx):
    """Real spherical harmonics at each G, index LM(L,M) = 1+L^2+L-M.

    src/paw_spherical.f90:66-148 (SPHERICAL$YLM) with the associated Legendre
    recursion of SPHERICAL_PLGNDR (:366-453): P_{M,M} = (-1)^M (2M-1)!!
    sin^M, (L-M) P_{L,M} = x(2L-1) P_{L-1,M} - (L+M-1) P_{L-2,M}, scaled by
    sqrt((L-M)!/(L+M)!); Y(LM0) = sqrt((2L+1)/4pi) P_{L,0} and, for M>0,
    Y(LM0-M) = sqrt(2) c (-1)^M P_{L,M} cos(M phi), Y(LM0+M) = same sin(M phi)
    with LM0 = L(L+1)+1.  Y = (1/sqrt(4pi), 0, ...) at G = 0 (:117-123).
    Returns (LMX, NG); LMX must be a full shell (L+1)^2.
    """
    lx = int(round(np.sqrt(lmx))) - 1
    if (lx + 1) ** 2 != lmx:
        raise ValueError("LMX does not correspond to a full shell")
    g = np.asarray(gvec, dtype=np.float64)
    ng = len(g)
    out = np.zeros((lmx, ng))
    out[0] = 1.0 / np.sqrt(4.0 * PI)
    if lmx == 1:
        return out
    disxy = np.sqrt(g[:, 0] ** 2 + g[:, 1] ** 2)
    dis = np.sqrt(disxy ** 2 + g[:, 2] ** 2)
    ok = np.abs(dis) >= 1.0e-12
    x = np.where(ok, g[:, 2] / np.where(ok, dis, 1.0), 1.0)        # cos(theta)
    fact = np.sqrt(np.maximum(1.0 - x * x, 0.0))
    # P_{L,M} for M >= 0 at LM(L,M) - 1 = L(L+1) - M (0-based)
    plm = np.zeros((lmx, ng))
    pmm = np.ones(ng)
    for m in range(lx + 1):
        cm = np.zeros(ng)
        c0 = pmm.copy()
        plm[m * m] = c0
        for l in range(m + 1, lx + 1):
            cp = (x * (2 * l - 1) * c0 - (l + m - 1) * cm) / (l - m)
            plm[l * (l + 1) - m] = cp
            cm, c0 = c0, cp
        pmm = -(2 * m + 1) * fact * pmm
    for l in range(1, lx + 1):
        i1 = l * (l + 1)
        f = 1.0
        sgn = 1.0
        for m in range(1, l + 1):
            f *= np.sqrt((l - m + 1) * (l + m))
            sgn = -sgn
            plm[i1 - m] = plm[i1 - m] / f
            plm[i1 + m] = sgn * plm[i1 - m]
    sinphi = np.where(disxy != 0.0, g[:, 1] / np.where(disxy != 0.0, disxy, 1.0), 0.0)
    cosphi = np.where(disxy != 0.0, g[:, 0] / np.where(disxy != 0.0, disxy, 1.0), 1.0)
    eiphi = cosphi + 1j * sinphi
    for l in range(lx + 1):
        lm0 = l * (l + 1)
        fac = np.sqrt((2 * l + 1) / (4.0 * PI))
        out[lm0] = plm[lm0] * fac
        eim = np.ones(ng, dtype=np.complex128)
        for m in range(1, l + 1):
            eim = eim * eiphi
            svar = fac * plm[lm0 + m] * np.sqrt(2.0)
            out[lm0 - m] = svar * eim.real
            out[lm0 + m] = svar * eim.imag
    out[:, ~ok] = 0.0
    out[0, ~ok] = 1.0 / np.sqrt(4.0 * PI)
    return out


def expandpro(lox, bare, ylmt, eigr):
    """Projector functions of one atom; src/paw_waves1.f90:6070-6109.

    P(lmn,G) = YLM(G,lm) * BAREPRO(G,ln) * (-i)^l * EIGR(G).
    bare: (LNX, NG); ylmt: (LMX, NG); returns (LMNX, NG).
    """
    rows = []
    for ln, l in enumerate(lox):
        for im in range(2 * l + 1):
            lm = l * l + im
            rows.append(ylmt[lm] * bare[ln] * ((-CI) ** l) * eigr)
    return np.array(rows)


# ----------------------------------------------------------------------------
# projections
# ----------------------------------------------------------------------------
def projections(plane, setup, pro, ylmt, r, psi):
    """<p|psi> = Omega sum_G conj(P) psi; src/paw_waves1.f90:5746-5905.

    pro: (NBAREPRO, NGL) bare projector tables (already / cell volume),
    r: (NAT,3) positions, psi: (NB, NGL).  Returns proj (NPRO, NB).
    """
    from .planewave import structurefactor
    out = np.zeros((setup.npro, psi.shape[0]), dtype=np.complex128)
    for iat in range(setup.nat):
        isp = setup.isp[iat]
        ib = setup.ibpro[isp]
        eigr = structurefactor(r[iat], plane.gbas, plane.igvec)
        p = expandpro(setup.lox[isp], pro[ib:ib + setup.lnx[isp]], ylmt, eigr)
        i0 = setup.ipro[iat]
        out[i0:i0 + len(p)] = plane.scalarproduct(p, psi, " ")
    return out


def addpro(plane, setup, pro, ylmt, r, psi, proj):
    """psi += sum_i |p_i> proj_i; src/paw_waves1.f90:5908-6067 (no Omega)."""
    from .planewave import structurefactor
    out = np.array(psi, dtype=np.complex128, copy=True)
    for iat in range(setup.nat):
        isp = setup.isp[iat]
        ib = setup.ibpro[isp]
        eigr = structurefactor(r[iat], plane.gbas, plane.igvec)
        p = expandpro(setup.lox[isp], pro[ib:ib + setup.lnx[isp]], ylmt, eigr)
        i0 = setup.ipro[iat]
        out += proj[i0:i0 + len(p)].T @ p
    return out


def hproj(setup, dh, proj):
    """DH <p|psi> per atom; src/paw_waves1.f90:4919-4975 (NDIM=1).

    dh: list per atom of (LMNX,LMNX) complex, dh[iat][i,j] = DH(I,J)."""
    out = np.zeros_like(proj)
    for iat in range(setup.nat):
        i0 = setup.ipro[iat]
        n = setup.lmnx[setup.isp[iat]]
        out[i0:i0 + n] = dh[iat] @ proj[i0:i0 + n]
    return out


def oproj(setup, proj):
    """DO <p|psi>; src/paw_waves2.f90:497-542."""
    out = np.zeros_like(proj)
    for iat in range(setup.nat):
        isp = setup.isp[iat]
        i0 = setup.ipro[iat]
        n = setup.lmnx[isp]
        out[i0:i0 + n] = setup.do_lmn(isp) @ proj[i0:i0 + n]
    return out


# ----------------------------------------------------------------------------
# energies and densities
# ----------------------------------------------------------------------------
def _fpm(occ, ib):
    f1 = occ[2 * ib]
    f2 = occ[2 * ib + 1] if 2 * ib + 1 < len(occ) else 0.0
    return f1, f2


def ekin(plane, occ, psi):
    """Pseudo kinetic energy; src/paw_waves1.f90:2790-2954 (no stress/bucket).

    dmat(G) = sum_b FP|c|^2 + FM Re(conj(c) c^-), E = Omega/2 sum G^2 dmat.
    """
    dmat = np.zeros(plane.ngl)
    for ib in range(psi.shape[0]):
        f1, f2 = _fpm(occ, ib)
        fp, fm = 0.5 * (f1 + f2), 0.5 * (f1 - f2)
        if fp == 0.0 and fm == 0.0:
            continue
        dmat += fp * np.real(np.conj(psi[ib]) * psi[ib])
        if abs(fm) < DSMALL:
            continue
        dmat += fm * np.real(np.conj(psi[ib]) * invertg(plane.minusg, psi[ib]))
    return plane.gweight * 0.5 * float(np.sum(plane.g2 * dmat))


def density(plane, occ, psi):
    """rho(r)=sum_b s1|psi|^2 + s2 Re(psi^2); src/paw_waves1.f90:5377-5500.

    Returns (NR3,NR2,NR1) real.
    """
    psir = plane.fft_gtor(psi)
    rho = np.zeros((plane.nr3, plane.nr2, plane.nr1))
    nb = len(occ)
    for ib in range(psi.shape[0]):
        f1 = occ[2 * ib]
        f2 = occ[2 * ib + 1] if 2 * ib + 1 < nb else 0.0
        s1, s2 = 0.5 * (f1 + f2), 0.5 * (f1 - f2)
        if s1 == 0.0 and s2 == 0.0:
            continue
        p1 = s1 * np.conj(psir[ib])
        if s2 != 0.0:
            p1 = p1 + s2 * psir[ib]          # EI2KR = 1 at Gamma
        rho += np.real(psir[ib] * p1)
    return rho


def vpsi(plane, v, psi):
    """H_loc psi = FFT^-1[V FFT psi] + G^2/2 psi; src/paw_waves1.f90:5099-5194.

    v: (NR3,NR2,NR1) real.  One transform pair per band, as the reference.
    """
    out = np.empty_like(psi, dtype=np.complex128)
    for ib in range(psi.shape[0]):
        out[ib] = plane.fft_rtog(v * plane.fft_gtor(psi[ib]))
    return out + 0.5 * plane.g2[None, :] * psi


def denmat(setup, occ, proj):
    """One-centre density matrix D_ij per atom; src/paw_waves1.f90:3331-3413,
    3492-3517 (TINV, NDIM=1; hermitised and made real)."""
    nbh = proj.shape[1]
    out = []
    for iat in range(setup.nat):
        i0 = setup.ipro[iat]
        n = setup.lmnx[setup.isp[iat]]
        p = proj[i0:i0 + n]
        d = np.zeros((n, n), dtype=np.complex128)
        for ib in range(nbh):
            f1, f2 = _fpm(occ, ib)
            s1, s2 = 0.5 * (f1 + f2), 0.5 * (f1 - f2)
            func = s1 * np.conj(p[:, ib]) + s2 * p[:, ib]
            d += np.outer(p[:, ib], func)
        d = np.real(d).astype(np.complex128)
        d = 0.5 * (d + np.conj(d.T))
        out.append(np.real(d))
    return out


def wavekinetic(plane, occ, mpsi, delt, psi0, psim):
    """Fictitious kinetic energy; src/paw_waves1.f90:6713-6847 (TINV)."""
    ek = 0.0
    for ib in range(psi0.shape[0]):
        f1, f2 = _fpm(occ, ib)
        if f1 == 0.0 and f2 == 0.0:
            continue
        d = psi0[ib] - psim[ib]
        s = float(np.sum(mpsi * (d.real ** 2 + d.imag ** 2))) / delt ** 2
        ek += s * 0.5 * (f1 + f2)
        if f1 - f2 == 0.0:
            continue
        dm = invertg(plane.minusg, d)
        s = float(np.sum(mpsi * (d.real * dm.real + d.imag * dm.imag))) / delt ** 2
        ek += s * 0.5 * (f1 - f2)
    return ek * plane.vol


# ----------------------------------------------------------------------------
# forces
# ----------------------------------------------------------------------------
def addoproj(proj, oproj_, lam):
    """PROJ += OPROJ Lambda in super-wave form; src/paw_waves2.f90:687-758."""
    l1, l2 = _fold_lambda(lam)
    return proj + oproj_ @ l1 + np.conj(oproj_) @ l2


def dedproj(setup, occ, proj, dh, rlam0):
    """dE/d<p| = DH<p|psi>f - DO<p|psi>Lambda (f_i+f_j)/2.

    src/paw_waves1.f90:5592-5691 (TINV)."""
    nb = len(occ)
    hp = hproj(setup, dh, proj)
    out = np.empty_like(hp)
    for ib in range(proj.shape[1]):
        f1, f2 = _fpm(occ, ib)
        fa, fb = 0.5 * (f1 + f2), 0.5 * (f1 - f2)
        out[:, ib] = fa * hp[:, ib] + fb * np.conj(hp[:, ib])
    occ = np.asarray(occ)
    lam = -np.asarray(rlam0) * 0.5 * (occ[:, None] + occ[None, :])
    return addoproj(out, oproj(setup, proj), lam[:nb, :nb])


def dedpro(plane, psi, dedproj_atom):
    """|dE/dp> = psi conj(dedproj), symmetrised with INVERTG.

    src/paw_waves1.f90:5694-5743.  dedproj_atom: (LMNX, NB) -> (LMNX, NGL)."""
    d = np.conj(dedproj_atom) @ psi
    return 0.5 * (d + invertg(plane.minusg, d))


def proforce(plane, lox, bare, ylmt, eigr, dedpro_atom):
    """F = 2 Omega Re sum_G [i P(G) conj(dE/dP)] G ; src/paw_waves1.f90:6112-6234
    (TSTRESS=.FALSE. branch)."""
    f = np.zeros(3)
    lmn = 0
    for ln, l in enumerate(lox):
        cwork = CI * bare[ln] * ((-CI) ** l) * eigr
        for im in range(2 * l + 1):
            lm = l * l + im
            svar = np.real(cwork * np.conj(dedpro_atom[lmn])) * ylmt[lm]
            f += svar @ plane.gvec
            lmn += 1
    return plane.gweight * 2.0 * f


def force(plane, setup, pro, ylmt, r0, occ, psi0, proj, dh, rlam0):
    """WAVES$FORCE without stress; src/paw_waves1.f90:4570-4738."""
    from .planewave import structurefactor
    ded = dedproj(setup, occ, proj, dh, rlam0)
    out = np.zeros((setup.nat, 3))
    for iat in range(setup.nat):
        isp = setup.isp[iat]
        i0 = setup.ipro[iat]
        n = setup.lmnx[isp]
        dp = dedpro(plane, psi0, ded[i0:i0 + n])
        eigr = structurefactor(r0[iat], plane.gbas, plane.igvec)
        ib = setup.ibpro[isp]
        out[iat] = proforce(plane, setup.lox[isp], pro[ib:ib + setup.lnx[isp]],
                            ylmt, eigr, dp)
    return out


# ----------------------------------------------------------------------------
# overlaps and orthogonalisation
# ----------------------------------------------------------------------------
def overlap(plane, nb, psi1, psi2):
    """<psi1|psi2> as NB x NB real matrix from super waves.

    src/paw_waves2.f90:761-853 (TINV branch)."""
    nbh = psi1.shape[0]
    t = plane.scalarproduct(psi1, psi2, " ")
    tm = plane.scalarproduct(psi1, psi2, "-")
    mat = np.zeros((nb, nb))
    for i in range(nbh):
        for j in range(nbh):
            re, im = 0.5 * t[i, j].real, 0.5 * t[i, j].imag
            rem, imm = 0.5 * tm[i, j].real, 0.5 * tm[i, j].imag
            blk = np.array([[re + rem, im + imm], [-im + imm, re - rem]])
            ia, ja = 2 * i, 2 * j
            ni, nj = min(2, nb - ia), min(2, nb - ja)
            mat[ia:ia + ni, ja:ja + nj] = blk[:ni, :nj]
    return mat


def overlap_1c(setup, nb, proj1, proj2):
    """One-centre overlap sum <psi|p>DO<p|psi>; src/paw_waves2.f90:856-1022."""
    p2a = oproj(setup, proj2)
    c1 = np.conj(proj1).T @ p2a
    c2 = proj1.T @ p2a
    nbh = proj1.shape[1]
    mat = np.zeros((nb, nb))
    for j in range(nbh):
        for i in range(nbh):
            cs1 = 0.5 * (c1[i, j] + c2[i, j])
            cs2 = 0.5 * (c1[i, j] - c2[i, j])
            mat[2 * i, 2 * j] = cs1.real
            mat[2 * i, 2 * j + 1] = cs1.imag
            mat[2 * i + 1, 2 * j] = -cs2.imag
            mat[2 * i + 1, 2 * j + 1] = cs2.real
    return mat


def _fold_lambda(lam):
    """Lambda (NB x NB) -> (LAMBDA1, LAMBDA2) (NBH x NBH), src/paw_waves2.f90:582-593."""
    lam = np.asarray(lam)
    l11 = lam[0::2, 0::2]
    l12 = lam[0::2, 1::2]
    l21 = lam[1::2, 0::2]
    l22 = lam[1::2, 1::2]
    cs1 = l11 + CI * l12
    cs2 = CI * l21 - l22
    return 0.5 * (cs1 - cs2), 0.5 * (cs1 + cs2)


def addopsi(plane, psibar, opsi, lam):
    """psi(+) = psibar + chi Lambda (super-wave form); src/paw_waves2.f90:547-615."""
    l1, l2 = _fold_lambda(lam)
    out = psibar + l1.T @ opsi
    return out + l2.T @ invertg(plane.minusg, opsi)


class OrthoNotConverged(RuntimeError):
    pass


def ortho_x(occ, chichi, psipsi, chipsi, lam, eps=1.0e-8, maxiter=200):
    """Lagrange multipliers for the orthonormality constraint.

    src/paw_waves2.f90:1844-2022 (WAVES_ORTHO_X).  Returns (lambda, iterations).
    The eigenvectors of sym(CHIPSI) enter only through U G U^T, so numpy.linalg
    .eigh stands in for DSYEV (src/paw_library.f90:1555-1585).
    """
    nb = len(occ)
    occ = np.asarray(occ, dtype=np.float64)
    lam = np.array(lam, dtype=np.float64, copy=True)
    h = 0.5 * (chipsi + chipsi.T)
    eig, u = np.linalg.eigh(h)
    iu = np.triu_indices(nb)
    for it in range(1, maxiter + 1):
        haux = chichi @ lam + 2.0 * chipsi
        gamn = lam.T @ haux
        sym = 0.5 * (gamn + gamn.T)
        up = np.triu(sym + psipsi)                 # uses PSIPSI(I,J), I<=J
        gamn = up + np.triu(up, 1).T
        gamn[np.diag_indices(nb)] -= 1.0
        digam = float(np.max(np.abs(gamn)))
        if digam < eps:
            return lam, it
        if digam > 1.0e200:
            raise OrthoNotConverged("ORTHO_X: iteration about to diverge")
        g = (u.T @ gamn) @ u
        g = g / (eig[:, None] + eig[None, :])
        g = (u @ g) @ u.T
        oi = occ + DSMALL
        lam = lam - (2.0 * oi[:, None] / (oi[:, None] + oi[None, :])) * g
        low = occ[:, None] < occ[None, :]
        ratio = oi[:, None] / oi[None, :]
        lam = np.where(low, lam.T * ratio, lam)
    raise OrthoNotConverged("ORTHO_X: loop for orthogonalization not converged")


def gramschmidt(plane, setup, pro, ylmt, r, nb, psi):
    """Initial orthonormalisation; src/paw_waves2.f90:2025-2177.

    WAVES_ORTHO_Y_C with PHI=CHI and the identity order builds the unique upper
    triangular T = 1 + X with T^T S T = 1 and positive diagonal (the root of
    the normalisation quadratic nearest zero, src/paw_waves2.f90:1083-1094), so
    T = R^-1 with S = R^T R (Cholesky).  The fold to super waves follows
    src/paw_waves2.f90:2100-2124.
    """
    proj = projections(plane, setup, pro, ylmt, r, psi)
    s = overlap(plane, nb, psi, psi) + overlap_1c(setup, nb, proj, proj)
    rmat = np.linalg.cholesky(s).T
    x = np.linalg.inv(rmat) - np.eye(nb)
    x11 = x[0::2, 0::2]
    x12 = x[0::2, 1::2]
    x21 = x[1::2, 0::2]
    x22 = x[1::2, 1::2]
    cs1 = x11 + CI * x12
    cs2 = -CI * x21 + x22
    x1, x2 = 0.5 * (cs1 + cs2), 0.5 * (cs1 - cs2)
    out = plane.addproduct(psi, psi, x1, " ")
    out = plane.addproduct(out, psi, x2, "-")
    return out


def expectation(plane, nb, psi0, hpsi):
    """<psi_n|H|psi_n> per real band; src/paw_waves1.f90:2094-2123."""
    eig = np.zeros(nb)
    for ib in range(psi0.shape[0]):
        h = overlap(plane, min(2, nb - 2 * ib), psi0[ib:ib + 1], hpsi[ib:ib + 1])
        eig[2 * ib] = h[0, 0]
        if 2 * ib + 1 < nb:
            eig[2 * ib + 1] = h[1, 1]
    return eig


# ----------------------------------------------------------------------------
# the CP step
# ----------------------------------------------------------------------------
class WaveSet:
    """WVSET_TYPE + GSET_TYPE state (src/paw_waves1.f90:158-193) for one k-point.

    model inputs: pro (NBAREPRO,NGL), v (NR3,NR2,NR1), dh (list per atom),
    occ (NB,), delt, emass, emasscg2, annee.
    """

    def __init__(self, plane, setup, pro, nb, occ, v, dh, delt, emass,
                 emasscg2, annee=0.0):
        self.plane = plane
        self.setup = setup
        self.pro = np.asarray(pro, dtype=np.float64)
        self.nb = nb
        self.nbh = (nb + 1) // 2
        self.occ = np.asarray(occ, dtype=np.float64)
        self.v = np.asarray(v, dtype=np.float64)
        self.dh = [np.asarray(d, dtype=np.complex128) for d in dh]
        self.delt, self.emass, self.emasscg2, self.annee = delt, emass, emasscg2, annee
        self.ylm = ylm(plane.gvec, setup.lmx)
        # src/paw_waves1.f90:6595-6673 (WAVES_PSIMASS, no stress, no bucket)
        self.mpsi = emass * (1.0 + emasscg2 * plane.g2)
        self.psi0 = None
        self.psim = None
        self.proj = None
        self.rlam0 = np.zeros((nb, nb))
        # optional density -> potential step: (density Plane, tables dict,
        # species (NAT,), QLM (NAT,)); None keeps the fixed potential v
        self.potential = None
        # optional one-centre terms: list per species of
        # oracle.augmentation.SphereSetup (needs the potential step above);
        # None keeps the fixed one-centre matrices dh
        self.sphere = None
        self.rlamm = None
        self.first = True
        self.wavekin1 = 0.0
        self.wavekin2 = 0.0

    # WAVES$ETOT (src/paw_waves1.f90:1610-2233), restricted to the hot path;
    # the density->potential step and AUGMENTATION are replaced by the fixed
    # model potential v and one-centre matrices dh (linear Hamiltonian).
    def etot(self, r0, rm=None):
        pl, st = self.plane, self.setup
        if self.first:
            self.psi0 = gramschmidt(pl, st, self.pro, self.ylm, r0, self.nb, self.psi0)
            self.psim = gramschmidt(pl, st, self.pro, self.ylm,
                                    r0 if rm is None else rm, self.nb, self.psim)
            self.proj = projections(pl, st, self.pro, self.ylm, r0, self.psi0)
            self.first = False
        e = {}
        e["ekin"] = ekin(pl, self.occ, self.psi0)
        dm = denmat(st, self.occ, self.proj)
        self.rho = density(pl, self.occ, self.psi0)
        if self.potential is not None and self.sphere is not None:
            # the one-centre terms linked in (src/paw_waves1.f90:1843-1915):
            # WAVES$MOMENTS -> QLM, POTENTIAL$VOFRHO(rho, QLM) -> VQLM, RHOB,
            # WAVES$SPHERE -> DH, DO, POTB, WAVES$ADDCONSTANTPOT
            from . import augmentation as oaug
            dplane, tabs = self.potential[0], self.potential[1]
            lmrx = np.array([s.lmrx for s in self.sphere])
            lmrxx = int(np.max(lmrx))
            qlm = np.zeros((st.nat, lmrxx))
            for a in range(st.nat):
                sp = self.sphere[st.isp[a]]
                qlm[a, :sp.lmrx] = oaug.moments(sp, dm[a])
            self.qlm = qlm
            self.pot = vofrho(dplane, self.rho, r0, st.isp, qlm, tabs, lmrx=lmrx)
            sph = oaug.waves_sphere(self.sphere, st.isp, dm,
                                    [self.pot["vqlm"][a, :lmrx[st.isp[a]]] for a in range(st.nat)],
                                    self.pot["rhob"], pl.vol)
            self.potb = sph["potb"]
            self.v = self.pot["v"] + self.potb
            self.dh = [np.asarray(d, dtype=np.complex128) for d in sph["dh"]]
            e["epot"] = self.pot["ehartree"] + self.pot["exc"]
            e["e1c"] = sph["e1c"]
        elif self.potential is not None:
            # POTENTIAL$VOFRHO between WAVES_DENSITY and WAVES_VPSI: the local
            # potential of this density (src/paw_potential.f90:324-896)
            dplane, tabs, isp, qlm = self.potential[:4]
            functional = self.potential[4] if len(self.potential) > 4 else 1
            self.pot = vofrho(dplane, self.rho, r0, isp, qlm, tabs, functional=functional)
            self.v = self.pot["v"]
            e["epot"] = self.pot["ehartree"] + self.pot["exc"]
        else:
            e["epot"] = pl.rweight * float(np.sum(self.v * self.rho))
        if self.sphere is None or self.potential is None:
            e["e1c"] = float(sum(np.real(np.sum(self.dh[i] * dm[i].T)) for i in range(st.nat)))
        e["etot"] = e["ekin"] + e["epot"] + e["e1c"]
        self.force = force(pl, st, self.pro, self.ylm, r0, self.occ, self.psi0,
                           self.proj, self.dh, self.rlam0)
        if self.potential is not None:
            self.force = self.force + self.pot["force"]
        hp = vpsi(pl, self.v, self.psi0)
        self.hpsi = addpro(pl, st, self.pro, self.ylm, r0, hp,
                           hproj(st, self.dh, self.proj))
        self.eig = expectation(pl, self.nb, self.psi0, self.hpsi)
        self.wavekin1 = wavekinetic(pl, self.occ, self.mpsi, self.delt, self.psi0, self.psim)
        self.denmat = dm
        return e

    # WAVES$PROPAGATE (src/paw_waves1.f90:6479-6592)
    def propagate(self):
        a1 = 2.0 / (1.0 + self.annee)
        a2 = 1.0 - a1
        a3 = -self.delt ** 2 / (1.0 + self.annee) / self.mpsi
        self.psim = a1 * self.psi0 + a2 * self.psim + a3[None, :] * self.hpsi
        self.hpsi = None

    # WAVES$ORTHOGONALIZE (src/paw_waves2.f90:11-493), TSAFEORTHO, no stress
    def orthogonalize(self, r0, rp):
        pl, st, nb = self.plane, self.setup, self.nb
        opsi = addpro(pl, st, self.pro, self.ylm, r0, self.psi0, oproj(st, self.proj))
        opsi = (self.delt ** 2 / (1.0 + self.annee) / self.mpsi)[None, :] * opsi
        self.proj = projections(pl, st, self.pro, self.ylm, rp, self.psim)
        oprj = projections(pl, st, self.pro, self.ylm, rp, opsi)
        mat = overlap_1c(st, nb, self.proj, self.proj) + overlap(pl, nb, self.psim, self.psim)
        omat = overlap_1c(st, nb, oprj, self.proj) + overlap(pl, nb, opsi, self.psim)
        oomat = overlap_1c(st, nb, oprj, oprj) + overlap(pl, nb, opsi, opsi)
        lam = np.array(self.rlam0, dtype=np.float64)
        lam = 0.5 * (np.triu(lam, 1) + np.triu(lam.T, 1))   # hermitian average
        lam = lam + lam.T + np.diag(np.diag(self.rlam0))
        occ = self.occ
        oi = np.repeat(occ[:, None], nb, axis=1)
        oj = np.repeat(occ[None, :], nb, axis=0)
        oj = np.where(oi + oj < DSMALL, DSMALL, oj)
        lam = lam * 0.5 * oi / (oi + oj)
        lam, self.ortho_iter = ortho_x(occ, oomat, mat, omat, lam)
        self.psim = addopsi(pl, self.psim, opsi, lam)
        self.proj = addoproj(self.proj, oprj, lam)
        self.rlam0 = lam
        self.omat = omat
        self.wavekin2 = wavekinetic(pl, occ, self.mpsi, self.delt, self.psi0, self.psim)

    # WAVES$SWITCH (src/paw_waves2.f90:2180-2317), NLAMBDAX=2
    def switch(self):
        self.psi0, self.psim = self.psim, self.psi0
        if self.rlamm is None:
            rlamp = np.array(self.rlam0, copy=True)
            self.rlamm = np.array(self.rlam0, copy=True)
        else:
            rlamp = 2.0 * self.rlam0 - self.rlamm
            self.rlamm = np.array(self.rlam0, copy=True)
        self.rlam0 = rlamp

    def fictitious_ekin(self):
        """0.5*(WAVEEKIN1+WAVEEKIN2); src/paw_waves1.f90:378-385."""
        if self.wavekin2 == 0.0:
            return self.wavekin1
        return 0.5 * (self.wavekin1 + self.wavekin2)
