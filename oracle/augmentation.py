"""Oracle restatement of the one-centre (augmentation) terms of the CP step:
WAVES$MOMENTS -> AUGMENTATION$MOMENTS (src/paw_waves1.f90:4147-4204,
src/paw_augmentation.f90:130-220), WAVES$SPHERE -> AUGMENTATION$SPHERE
(src/paw_waves1.f90:4418-4512, src/paw_augmentation.f90:223-600) and
WAVES$ADDCONSTANTPOT (src/paw_waves1.f90:4515-4567), for one spin (NDIMD=1),
the local-density functional of type 1 (local exchange + Perdew-Zunger), no
soft core, no external one-centre potential.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

The radial functions live on the reference's shifted logarithmic grid
(SHLOGRADIAL, src/paw_radial.f90:1571-2200):  r_i = r1 exp(dex (i-1)) - r1 with
r_1 replaced by RMIN = 1e-9, and its integration/derivative rules are those of
RADIAL_INTEGRATEEQUISPACED / RADIAL_DERIVEEQUISPACED (:397-502) in the index
variable x with dr/dx = dex (r + r1).  Densities and potentials are radial
parts of real spherical-harmonic expansions, f(r) = sum_LM f_LM(|r|) Y_LM.
"""
import numpy as np

from .waves import ylm

PI = 4.0 * np.arctan(1.0)
Y0 = 1.0 / np.sqrt(4.0 * PI)
FPI = 4.0 * PI
RMIN = 1.0e-9                      # src/paw_radial.f90:1566
RHOTMIN = 1.0e-6                   # src/paw_dft.f90:62
EXFAC = -3.0 / (4.0 * PI) * (3.0 * PI ** 2) ** (1.0 / 3.0)      # src/paw_dft.f90:1691-1693
GAMMAU, BETA1U, BETA2U = -142.3e-3, 1052.9e-3, 333.4e-3         # src/paw_dft.f90:2783-2792
AU, BU, CU, DU = 31.10e-3, -48.0e-3, 2.0e-3, -11.6e-3
RSFAC = (3.0 / (4.0 * PI)) ** (1.0 / 3.0)


def lofm(lm):
    """Main angular momentum of the 0-based real-harmonic index lm."""
    return int(np.sqrt(lm + 1e-5))


# ----------------------------------------------------------------------------
# radial grid
# ----------------------------------------------------------------------------
class RadialGrid:
    """SHLOGRADIAL grid (src/paw_radial.f90:1626-1655, 1767-1925)."""

    def __init__(self, r1, dex, nr):
        self.r1, self.dex, self.nr = float(r1), float(dex), int(nr)
        xexp = np.exp(self.dex)
        ri = np.empty(self.nr)                     # r1 exp(dex (i-1)), by repeated products
        v = self.r1 / xexp
        for i in range(self.nr):
            v = v * xexp
            ri[i] = v
        self.ri = ri
        self.r = ri - self.r1
        self.r[0] = RMIN
        self.drdx = self.dex * ri

    def integrate(self, f):
        """Antiderivative (RADIAL$INTEGRATE -> SHLOGRADIAL$INTEGRATE)."""
        return integrate_equispaced(np.asarray(f, dtype=np.float64) * self.drdx)

    def integral(self, f):
        """RADIAL$INTEGRAL: the antiderivative at the last point."""
        return float(self.integrate(f)[-1])

    def derive(self, f):
        """RADIAL$DERIVE -> SHLOGRADIAL$DERIVE."""
        return derive_equispaced(np.asarray(f, dtype=np.float64)) / self.drdx

    def poisson(self, l, rho):
        """RADIAL$POISSON (src/paw_radial.f90:1345-1386): potential of the
        angular-momentum-l component rho(r) Y_lm."""
        r = self.r
        aux2 = self.integrate(rho * r ** (l + 2))
        fac = 4.0 * PI / (2 * l + 1)
        v = fac * r ** (-l - 1) * aux2
        aux1 = rho * r ** (-l + 1)
        aux2 = self.integrate(aux1)
        return v + fac * r ** l * (aux2[-1] - aux2)


def integrate_equispaced(f):
    """RADIAL_INTEGRATEEQUISPACED (src/paw_radial.f90:397-458)."""
    f = np.asarray(f, dtype=np.float64)
    nx = len(f)
    if nx < 10:
        raise ValueError("number of mesh points smaller than 10")
    c0, c1, c2, c3 = -31.0 / 48.0, 11.0 / 48.0, -5.0 / 48.0, 1.0 / 48.0
    g = np.empty(nx)
    g[0] = c0 * f[0] + c1 * f[1] + c2 * f[2] + c3 * f[3] + f[0]
    for i in range(1, nx):
        g[i] = g[i - 1] + f[i]
    g[3:] = g[3:] + c0 * f[3:] + c1 * f[2:-1] + c2 * f[1:-2] + c3 * f[:-3]
    a = (24.0 * f[0] - 36.0 * f[1] + 24.0 * f[2] - 6.0 * f[3]) / 6.0
    b = (-26.0 * f[0] + 57.0 * f[1] - 42.0 * f[2] + 11.0 * f[3]) / 6.0
    c = (9.0 * f[0] - 24.0 * f[1] + 21.0 * f[2] - 6.0 * f[3]) / 6.0
    d = (-f[0] + 3.0 * f[1] - 3.0 * f[2] + f[3]) / 6.0
    for i in range(3):
        x = float(i + 1)
        g[i] = a * (x - 1.0) + 0.5 * b * (x * x - 1.0) + c * (x ** 3 - 1.0) / 3.0 + 0.25 * d * (x ** 4 - 1.0)
    return g


def derive_equispaced(f):
    """RADIAL_DERIVEEQUISPACED (src/paw_radial.f90:461-500)."""
    f = np.asarray(f, dtype=np.float64)
    n = len(f)
    g = np.empty(n)
    g[0] = (-11.0 * f[0] + 18.0 * f[1] - 9.0 * f[2] + 2.0 * f[3]) / 6.0
    g[1] = (-2.0 * f[0] - 14.0 * f[1] + 24.0 * f[2] - 10.0 * f[3] + 2.0 * f[4]) / 12.0
    g[2:n - 2] = f[:n - 4] / 12.0 - 2.0 / 3.0 * f[1:n - 3] + 2.0 / 3.0 * f[3:n - 1] - f[4:] / 12.0
    g[n - 2] = -(2.0 * f[n - 5] - 10.0 * f[n - 4] + 24.0 * f[n - 3] - 14.0 * f[n - 2] - 2.0 * f[n - 1]) / 12.0
    g[n - 1] = -(2.0 * f[n - 4] - 9.0 * f[n - 3] + 18.0 * f[n - 2] - 11.0 * f[n - 1]) / 6.0
    return g


def gaussn(l, alpha):
    """GAUSSN (src/paw_generalpurpose.f90:1088-1121): C with
    int d^3r C exp(-alpha r^2) r^l Y_lm r^l Y_lm = 1."""
    ifac = 1
    for i in range(2, l + 2):
        ifac *= 2 * i - 1
    rint = np.sqrt(PI / (4.0 * alpha)) * ifac / (2.0 * alpha) ** (l + 1)
    return 1.0 / rint


# ----------------------------------------------------------------------------
# Gaunt coefficients of the real harmonics
# ----------------------------------------------------------------------------
_GAUNT = {}


def gaunt(lmx):
    """CG[lm1, lm2, lm3] = int dOmega Y_lm1 Y_lm2 Y_lm3 (SPHERICAL$GAUNT,
    src/paw_spherical.f90:1052-1077) for 0-based indices < lmx, by
    Gauss-Legendre x uniform-phi quadrature, exact for these polynomials."""
    if lmx in _GAUNT:
        return _GAUNT[lmx]
    lx = int(round(np.sqrt(lmx))) - 1
    nq = 3 * lx + 2
    ct, wt = np.polynomial.legendre.leggauss(nq)
    nphi = 3 * lx + 2
    phi = 2.0 * PI * np.arange(nphi) / nphi
    c, p = np.meshgrid(ct, phi, indexing="ij")
    s = np.sqrt(1.0 - c * c)
    pts = np.stack([s * np.cos(p), s * np.sin(p), c], axis=-1).reshape(-1, 3)
    w = (wt[:, None] * np.full(nphi, 2.0 * PI / nphi)[None, :]).ravel()
    y = ylm(pts, lmx)
    cg = np.einsum("aq,bq,cq,q->abc", y, y, y, w)
    cg[np.abs(cg) < 1.0e-14] = 0.0
    _GAUNT[lmx] = cg
    return cg


# ----------------------------------------------------------------------------
# local-density functional with derivatives to third order
# ----------------------------------------------------------------------------
def lda3(rho):
    """DFT3 for the local functional of type 1 without spin
    (src/paw_dft.f90:1144-1332): exchange EXCHANGE_X3 without gradient
    (:2358-2430) plus PERDEWZUNGER_RHODEP3 (:3019-3146); returns exc and its
    first three derivatives in rho.  Densities below RHOTMIN are raised to it
    and the results continued linearly (:1290-1306); rho <= 0 gives zeros."""
    rho = np.asarray(rho, dtype=np.float64)
    out = [np.zeros_like(rho) for _ in range(4)]
    pos = rho > 0.0
    v = np.maximum(rho[pos], RHOTMIN)
    r13 = v ** (1.0 / 3.0)
    r23 = r13 * r13
    r43 = r23 * r23
    ex = EXFAC * r43
    ex1 = EXFAC * 4.0 / 3.0 * r13
    ex2 = EXFAC * 4.0 / 9.0 / r23
    ex3 = -2.0 / 3.0 * ex2 / v
    rs = RSFAC / r13
    ec = np.empty_like(v)
    ec1 = np.empty_like(v)
    ec2 = np.empty_like(v)
    ec3 = np.empty_like(v)
    hi = rs < 1.0
    r, d = rs[hi], v[hi]
    ln = np.log(r)
    lnr = r * ln
    ec[hi] = d * (BU + DU * r + (AU + CU * r) * ln)
    ec1[hi] = (1.0 / 3.0) * (-AU + 3.0 * BU - (CU - 2.0 * DU) * r + 3.0 * AU * ln + 2.0 * CU * lnr)
    ec2[hi] = -(1.0 / 9.0) / d * (3.0 * AU + (CU + 2.0 * DU) * r + 2.0 * CU * lnr)
    ec3[hi] = (1.0 / 27.0) / d ** 2 * (9.0 * AU + (6.0 * CU + 8.0 * DU) * r + 8.0 * CU * lnr)
    lo = ~hi
    r, d = rs[lo], v[lo]
    sq = np.sqrt(r)
    sv = 1.0 / (1.0 + sq * (BETA1U + BETA2U * sq))
    sv1 = sv / (6.0 * d)
    e = GAMMAU * d * sv
    ec[lo] = e
    ec1[lo] = e * sv1 * (6.0 + sq * (7.0 * BETA1U + 8.0 * BETA2U * sq))
    ec2[lo] = e * sv1 ** 2 * sq * (5.0 * BETA1U + sq * ((7.0 * BETA1U ** 2 + 8.0 * BETA2U)
                                                       + sq * BETA2U * (21.0 * BETA1U + sq * BETA2U * 16.0)))
    ec3[lo] = -e * sv1 ** 3 * sq * (35.0 * BETA1U + sq * ((76.0 * BETA1U ** 2 + 64.0 * BETA2U)
                                    + sq * (BETA1U * (35.0 * BETA1U ** 2 + 234.0 * BETA2U)
                                            + sq * BETA2U * ((140.0 * BETA1U ** 2 + 176.0 * BETA2U)
                                                             + sq * BETA2U * (175.0 * BETA1U
                                                                              + sq * BETA2U * 64.0)))))
    f0, f1, f2, f3 = ex + ec, ex1 + ec1, ex2 + ec2, ex3 + ec3
    small = rho[pos] < RHOTMIN
    a = rho[pos]
    f3 = np.where(small, 0.0, f3)
    f2 = np.where(small, 0.0, f2)
    f1 = np.where(small, f0 / RHOTMIN, f1)
    f0 = np.where(small, f0 * a / RHOTMIN, f0)
    for k, val in enumerate((f0, f1, f2, f3)):
        out[k][pos] = val
    return out


# ----------------------------------------------------------------------------
# AUGMENTATION$SPHERE pieces
# ----------------------------------------------------------------------------
def aug_rho(lox, phi, denmat, lmrx):
    """AUGMENTATION_RHO (src/paw_augmentation.f90:603-656): one-centre density
    rho_LM(r) = sum CG(LM1,LM2,LM) Re D(LMN1,LMN2) phi_LN1 phi_LN2 as (LMRX, NR)."""
    cg = gaunt(max(lmrx, (max(lox) + 1) ** 2))
    nr = phi.shape[1]
    out = np.zeros((lmrx, nr))
    lmn1 = 0
    for ln1, l1 in enumerate(lox):
        for im1 in range(2 * l1 + 1):
            lm1 = l1 * l1 + im1
            lmn2 = 0
            for ln2, l2 in enumerate(lox):
                for im2 in range(2 * l2 + 1):
                    lm2 = l2 * l2 + im2
                    for lm3 in range(lmrx):
                        c = cg[lm1, lm2, lm3]
                        if c != 0.0:
                            out[lm3] += c * float(np.real(denmat[lmn1, lmn2])) * phi[ln1] * phi[ln2]
                    lmn2 += 1
            lmn1 += 1
    return out


def aug_qlm(grid, lmrx, aez, aecore, pscore, aerho, psrho):
    """AUGMENTATION_QLM (src/paw_augmentation.f90:659-696)."""
    r = grid.r
    q = np.zeros(lmrx)
    for lm in range(lmrx):
        l = lofm(lm)
        q[lm] = grid.integral((aerho[lm] - psrho[lm]) * r ** (l + 2))
    q[0] += grid.integral((aecore - pscore) * r ** 2) - aez * Y0
    return q


def aug_xc(grid, lmrx, rho):
    """AUGMENTATION_XC (src/paw_augmentation.f90:699-977) for NDIMD=1 and a
    local functional (IMAX=1): the energy expanded about the spherical density
    to second order in the non-spherical components.  Returns (EXC, VXC(LMRX,NR))."""
    r = grid.r
    x = np.asarray(rho, dtype=np.float64)[:lmrx]
    nr = len(r)
    work1 = np.zeros(nr)
    xder = np.zeros((lmrx, nr))
    ok = x[0] > 0.0
    e, v1, v2, v3 = lda3(x[0] * Y0)
    nons = np.sum(x[1:] ** 2, axis=0) if lmrx > 1 else np.zeros(nr)
    work1 = np.where(ok, FPI * e + 0.5 * nons * v2, 0.0)
    xder[0] = np.where(ok, FPI * v1 * Y0 + 0.5 * Y0 * nons * v3, 0.0)
    for lm in range(1, lmrx):
        # XDER(IR,I,LM) and XDER(IR,J,LM) each gain 0.5 V2 XVAL(LM)
        xder[lm] = np.where(ok, 0.5 * v2 * x[lm] + 0.5 * v2 * x[lm], 0.0)
    exc = grid.integral(work1 * r ** 2)
    vxc = xder.copy()
    if r[0] < 1.0e-5:
        vxc[:, 0] = vxc[:, 1]
    return exc, vxc


def adjust_vqlm(grid, lmrx, psrho, rcsm, qlm):
    """AUGMENTATION_ADJUSTVQLM (src/paw_augmentation.f90:2054-2104); psrho
    includes the pseudo core in LM=1."""
    r = grid.r
    alpha = 1.0 / rcsm ** 2
    out = np.zeros(lmrx)
    for lm in range(lmrx):
        l = lofm(lm)
        g = gaussn(l, alpha) * np.exp(-alpha * r ** 2) * r ** l
        aux = grid.poisson(l, psrho[lm] + qlm[lm] * g)
        out[lm] = -grid.integral(aux * g * r ** 2)
    return out


def ps_hartree(grid, lmrx, psrhoc, psrho, vadd, rcsm, qlm, vqlm, rhob):
    """AUGMENTATION_PSHARTREE (src/paw_augmentation.f90:1827-1946): (PSPOT(LMRX,NR), PSEH)."""
    r = grid.r
    alpha = 1.0 / rcsm ** 2
    rhohat = np.zeros((lmrx, len(r)))
    for lm in range(lmrx):
        l = lofm(lm)
        rhohat[lm] = qlm[lm] * gaussn(l, alpha) * np.exp(-alpha * r ** 2) * r ** l
    pot_out = np.zeros((lmrx, len(r)))
    pse = vadd * (psrho[0] + psrhoc)
    pot_out[0] += vadd
    for lm in range(lmrx):
        l = lofm(lm)
        rho1 = psrho[lm] + rhohat[lm]
        if l == 0:
            rho1 = rho1 + psrhoc
        pot = grid.poisson(l, rho1)
        pse = pse + 0.5 * pot * rho1
        pot_out[lm] += pot
    for lm in range(lmrx):
        l = lofm(lm)
        pot = vqlm[lm] * r ** l
        pot_out[lm] += pot
        rho1 = psrho[lm] + rhohat[lm]
        if l == 0:
            rho1 = rho1 + psrhoc
        pse = pse + rho1 * pot
    svar = -2.0 * PI * rhob / 3.0 * np.sqrt(4.0 * PI)
    pot = svar * r ** 2
    pot_out[0] += pot
    pse = pse + (psrho[0] + rhohat[0] + psrhoc) * pot
    return pot_out, grid.integral(pse * r ** 2)


def nucpot(r, z, rnuc):
    """RADIAL$NUCPOT (src/paw_radial.f90:1464-1498): homogeneously charged
    nucleus of radius rnuc, divided by Y0."""
    r = np.asarray(r)
    pot = np.where(r > rnuc, -z / np.where(r > rnuc, r, 1.0), -z / rnuc * (1.5 - 0.5 * (r / rnuc) ** 2))
    return pot / Y0


def ae_hartree(grid, lmrx, aez, rnuc, rhoc, aerho, vqlm, rhob):
    """AUGMENTATION_AEHARTREE (src/paw_augmentation.f90:1949-2051): (AEPOT(LMRX,NR), AEEH)."""
    r = grid.r
    pot_out = np.zeros((lmrx, len(r)))
    pot = grid.poisson(0, rhoc) + nucpot(r, aez, rnuc)
    aee = pot * aerho[0]
    pot_out[0] += pot
    for lm in range(lmrx):
        l = lofm(lm)
        pot = grid.poisson(l, aerho[lm])
        aee = aee + 0.5 * pot * aerho[lm]
        pot_out[lm] += pot
    for lm in range(lmrx):
        l = lofm(lm)
        pot = vqlm[lm] * r ** l
        pot_out[lm] += pot
        aee = aee + aerho[lm] * pot
        if l == 0:
            aee = aee + rhoc * pot
    eext = -vqlm[0] * Y0 * aez
    svar = -2.0 * PI * rhob / 3.0 * np.sqrt(4.0 * PI)
    pot = svar * r ** 2
    pot_out[0] += pot
    aee = aee + (aerho[0] + rhoc) * pot
    return pot_out, grid.integral(aee * r ** 2) + eext


def expect(grid, lox, lmrx, aepot, pspot, aephi, psphi):
    """AUGMENTATION_EXPECT (src/paw_augmentation.f90:2159-2239):
    DATH(LMN1,LMN2) = <aephi|aepot|aephi> - <psphi|pspot|psphi>."""
    r2 = grid.r ** 2
    lmnx = sum(2 * l + 1 for l in lox)
    cg = gaunt(max(lmrx, (max(lox) + 1) ** 2))
    dath = np.zeros((lmnx, lmnx))
    lmn1 = 0
    for ln1, l1 in enumerate(lox):
        for im1 in range(2 * l1 + 1):
            lm1 = l1 * l1 + im1
            lmn2 = 0
            for ln2, l2 in enumerate(lox):
                for im2 in range(2 * l2 + 1):
                    lm2 = l2 * l2 + im2
                    aedmu = np.zeros(len(r2))
                    psdmu = np.zeros(len(r2))
                    for lm3 in range(lmrx):
                        c = cg[lm1, lm2, lm3]
                        if c != 0.0:
                            aedmu = aedmu + c * aepot[lm3]
                            psdmu = psdmu + c * pspot[lm3]
                    w = (aedmu * aephi[ln1] * aephi[ln2] - psdmu * psphi[ln1] * psphi[ln2]) * r2
                    dath[lmn1, lmn2] += grid.integral(w)
                    lmn2 += 1
            lmn1 += 1
    return dath


def expand_da(lox, da1):
    """AUGMENTATION_EXPANDDA (src/paw_augmentation.f90:3045-3076)."""
    lmnx = sum(2 * l + 1 for l in lox)
    out = np.zeros((lmnx, lmnx))
    o1 = 0
    for ln1, l1 in enumerate(lox):
        o2 = 0
        for ln2, l2 in enumerate(lox):
            if l1 == l2:
                for m in range(2 * l1 + 1):
                    out[o1 + m, o2 + m] = da1[ln1, ln2]
            o2 += 2 * l2 + 1
        o1 += 2 * l1 + 1
    return out


class SphereSetup:
    """The per-species setup data AUGMENTATION$SPHERE reads from SETUP
    (src/paw_augmentation.f90:265-305): grid, AEZ, LOX, AEPHI, PSPHI, AECORE,
    PSCORE, RCSM, VADD, DO, DEKIN, LMRX, and the nuclear radius RNUC of
    RADIAL$NUCPOT.  Radial arrays are (LNX, NR) / (NR,)."""

    def __init__(self, grid, aez, rnuc, lox, aephi, psphi, aecore, pscore, rcsm, vadd, dover, dekin, lmrx):
        self.grid = grid
        self.aez, self.rnuc = float(aez), float(rnuc)
        self.lox = list(lox)
        self.aephi, self.psphi = np.asarray(aephi), np.asarray(psphi)
        self.aecore, self.pscore = np.asarray(aecore), np.asarray(pscore)
        self.rcsm = float(rcsm)
        self.vadd = np.asarray(vadd)
        self.dover, self.dekin = np.asarray(dover), np.asarray(dekin)
        self.lmrx = int(lmrx)


def moments(sp, denmat):
    """AUGMENTATION$MOMENTS (src/paw_augmentation.f90:130-220): QLM(LMRX)."""
    aerho = aug_rho(sp.lox, sp.aephi, denmat, sp.lmrx)
    psrho = aug_rho(sp.lox, sp.psphi, denmat, sp.lmrx)
    return aug_qlm(sp.grid, sp.lmrx, sp.aez, sp.aecore, sp.pscore, aerho, psrho)


def sphere(sp, denmat, vqlm, rhob):
    """AUGMENTATION$SPHERE (src/paw_augmentation.f90:223-600), NDIMD=1, no
    soft core, no external potential.  denmat (LMNX,LMNX) real symmetric;
    vqlm (LMRX,) dE/dQLM from the plane-wave part.  Returns dict dath
    (LMNX,LMNX), do, potb (this atom's POTB, before / Omega), energies."""
    g = sp.grid
    r = g.r
    lmrx = sp.lmrx
    do = expand_da(sp.lox, sp.dover)
    dtkin = expand_da(sp.lox, sp.dekin)
    ekinnl = float(np.sum(denmat * dtkin))
    aerho = aug_rho(sp.lox, sp.aephi, denmat, lmrx)
    psrho = aug_rho(sp.lox, sp.psphi, denmat, lmrx)
    qlm = aug_qlm(g, lmrx, sp.aez, sp.aecore, sp.pscore, aerho, psrho)
    rho = psrho.copy()
    rho[0] = rho[0] + sp.pscore
    vqlm1 = adjust_vqlm(g, lmrx, rho, sp.rcsm, qlm) + np.asarray(vqlm)[:lmrx]
    pshpot, pseh = ps_hartree(g, lmrx, sp.pscore, psrho, sp.vadd, sp.rcsm, qlm, vqlm1, rhob)
    aehpot, aeeh = ae_hartree(g, lmrx, sp.aez, sp.rnuc, sp.aecore, aerho, vqlm1, rhob)
    corexc, _ = aug_xc(g, 1, sp.aecore[None, :])
    rho = aerho.copy()
    rho[0] = rho[0] + sp.aecore
    aeexc, aexcpot = aug_xc(g, lmrx, rho)
    aeexc -= corexc
    rho = psrho.copy()
    rho[0] = rho[0] + sp.pscore
    psexc, psxcpot = aug_xc(g, lmrx, rho)
    aetot = aexcpot + aehpot
    pstot = psxcpot + pshpot
    dwork = -(aehpot[0] - (pshpot[0] - sp.vadd)) * r ** 2 * np.sqrt(4.0 * PI)
    potb = g.integral(dwork)
    dath = dtkin + expect(g, sp.lox, lmrx, aetot, pstot, sp.aephi, sp.psphi)
    e = {"ekin": ekinnl, "aexc": aeexc, "psexc": psexc, "aeeh": aeeh, "pseh": pseh}
    e["e1c"] = ekinnl + aeexc - psexc + aeeh - pseh
    return {"dath": dath, "do": do, "potb": potb, "qlm": qlm, "energies": e}


def waves_sphere(setups, isp, denmats, vqlm, rhob, vol):
    """WAVES$SPHERE + the DH part of WAVES$ADDCONSTANTPOT
    (src/paw_waves1.f90:4418-4567): per atom DH, DO; POTB = sum POTB1 / Omega;
    DH += POTB DO; total one-centre energy."""
    out = [sphere(setups[s], d, v, rhob) for s, d, v in zip(isp, denmats, vqlm)]
    potb = sum(o["potb"] for o in out) / vol
    dh = [o["dath"] + potb * o["do"] for o in out]
    e1c = sum(o["energies"]["e1c"] for o in out)
    return {"dh": dh, "do": [o["do"] for o in out], "potb": potb, "e1c": e1c, "atoms": out}
