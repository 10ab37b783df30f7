"""Oracle restatement of the density -> potential step POTENTIAL$VOFRHO
(cp-paw src/paw_potential.f90:324-896) for one scope: no spin (NDIMD=1), the
local-density functional of type 1 (local exchange + Perdew-Zunger
correlation, src/paw_dft.f90:124-130), compensation charges of any l
(LMRXX = (LX+1)^2, LMRX per species), no confinement, no isolation (ISOLATE)
and no pair potential of
the compensation charges (POTENTIAL_PAIRP, which depends only on QLM and the
positions, not on the density).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Sphere convention.  The reference's density object stores half of the G
sphere (TSUPER) and doubles its sums, counting G = 0 half
(src/paw_potential.f90:1199-1203,1311-1318).  This restatement sums over the
full sphere of the same plane-wave object, which gives the same numbers for a
real density: the terms at G and -G are complex conjugates, so the real parts
(energies, VQLM) and G * imaginary parts (forces) of the pair are equal.
Hence the factors GWEIGHT (= Omega) here in place of 2 GWEIGHT.
"""
import numpy as np

from .planewave import PI, structurefactor

Y0 = 1.0 / np.sqrt(4.0 * PI)
FPI = 4.0 * PI
RHOTMIN = 1.0e-6                                    # src/paw_dft.f90:62

# local exchange, src/paw_dft.f90:1691-1693
KFFAC = (3.0 * PI ** 2) ** (1.0 / 3.0)
EXFAC = -3.0 / (4.0 * PI) * KFFAC
# Perdew-Zunger, non-spin-polarised parameters (src/paw_dft.f90:2783-2792) and
# RSFAC (src/paw_dft.f90:2814)
GAMMAU, BETA1U, BETA2U = -142.3e-3, 1052.9e-3, 333.4e-3
AU, BU, CU, DU = 31.10e-3, -48.0e-3, 2.0e-3, -11.6e-3
RSFAC = (3.0 / (4.0 * PI)) ** (1.0 / 3.0)


def exchange_lda(rho):
    """EXCHANGE_X1 without gradient (src/paw_dft.f90:2137-2195): EX = EXFAC rho^(4/3)
    and dEX/drho.  rho > 0 (array)."""
    rho13 = rho ** (1.0 / 3.0)
    rho43 = (rho13 * rho13) ** 2
    ex = EXFAC * rho43
    return ex, ex * (4.0 / 3.0) / rho


def pz_unpolarised(rho):
    """PERDEWZUNGER_RHODEP, non-spin part (src/paw_dft.f90:2824-2910):
    ECU = rho eps_c and VCU = d(rho eps_c)/drho.  rho > 0 (array)."""
    rs = RSFAC / rho ** (1.0 / 3.0)
    ecu = np.empty_like(rho)
    vcu = np.empty_like(rho)
    hi = rs < 1.0
    r, d = rs[hi], rho[hi]
    rsln = np.log(r)
    ecu[hi] = d * (BU + DU * r + (AU + CU * r) * rsln)
    vcu[hi] = (1.0 / 3.0) * (-AU + 3.0 * BU - (CU - 2.0 * DU) * r + 3.0 * AU * rsln + 2.0 * CU * r * rsln)
    lo = ~hi
    r, d = rs[lo], rho[lo]
    sq = np.sqrt(r)
    svar = 1.0 / (1.0 + sq * (BETA1U + BETA2U * sq))
    svar1 = svar / (6.0 * d)
    e = GAMMAU * d * svar
    ecu[lo] = e
    vcu[lo] = e * svar1 * (6.0 + sq * (7.0 * BETA1U + 8.0 * BETA2U * sq))
    return ecu, vcu


def lda(rho):
    """Energy density and potential of one grid point set, as POTENTIAL_XC
    (src/paw_potential.f90:1397-1401: rho clamped to >= 1e-8) calls DFT
    (src/paw_dft.f90:780-965: densities below RHOTMIN are raised to RHOTMIN and
    the results scaled back linearly)."""
    rhot = np.maximum(np.asarray(rho, dtype=np.float64), 1.0e-8)
    val = np.maximum(rhot, RHOTMIN)
    ex, ex_r = exchange_lda(val)
    ec, vc = pz_unpolarised(val)
    exc = ex + ec
    dexc = ex_r + vc
    small = rhot < RHOTMIN
    dexc = np.where(small, exc / RHOTMIN, dexc)
    exc = np.where(small, exc * rhot / RHOTMIN, exc)
    return exc, dexc


def ylmofg(gvec, lmx):
    """POTENTIAL_YLMOFG (src/paw_potential.f90:1548-1614): solid harmonics
    |G|^l Y_LM(G) as (LMX, NG); (Y0, 0, ...) for |G| < 1e-7."""
    from .waves import ylm
    g = np.asarray(gvec, dtype=np.float64)
    out = ylm(g, lmx)
    gl = np.sqrt(np.sum(g * g, axis=1))
    for lm in range(lmx):
        out[lm] *= gl ** int(np.sqrt(lm + 1e-5))
    small = gl < 1.0e-7
    out[:, small] = 0.0
    out[0, small] = Y0
    return out


def cfac(lmx):
    """CFAC(LM) = (-i)^L / (2L+1)!! of POTENTIAL_HARTREE
    (src/paw_potential.f90:1146-1152, 1246-1253)."""
    out = np.zeros(lmx, dtype=np.complex128)
    c = 1j
    lm = 0
    l = 0
    while lm < lmx:
        c = -c * 1j / (2 * l + 1)
        for _ in range(2 * l + 1):
            if lm < lmx:
                out[lm] = c
            lm += 1
        l += 1
    return out


def vofrho(plane, rho_r, pos, isp, qlm, tabs, lmrx=None):
    """POTENTIAL$VOFRHO -> POTENTIAL_VOFRHO (src/paw_potential.f90:324-484,
    487-896) in the scope of the module docstring.

    plane: the density plane-wave object (oracle.planewave.Plane built with the
      density cutoff for the sphere); rho_r (NR3,NR2,NR1) real pseudo density;
    pos (NAT,3); isp (NAT) species 0..NSP-1; qlm: multipole moments of the
      one-centre densities, QLM(LMRXX,NAT) as (NAT, LMRXX), or (NAT,) for
      LMRXX = 1; lmrx: LMRX per species (default LMRXX for all);
    tabs: dict of (NSP,NGL) real tables on the sphere: 'pscorg' (PS core),
      'vbarg' (V-bar), 'g0' (compensation Gaussians), 'v0' (their potential).
    Returns dict: v (NR3,NR2,NR1) total local potential, ehartree, exc, rhob,
      vqlm (shape of qlm) = dE/dQLM, force (NAT,3) (Hartree + pseudo-core terms).
    """
    om = plane.vol
    ig0 = plane.ngamma
    g2, gv = plane.g2, plane.gvec
    nonzero = np.arange(plane.ngl) != ig0
    isp = np.asarray(isp)
    flat = np.ndim(qlm) == 1
    qlm = np.asarray(qlm, dtype=np.float64)
    qlm = qlm.reshape(len(isp), -1) if qlm.size else np.zeros((len(isp), 1))
    lmrxx = qlm.shape[1]
    nsp = int(np.max(isp)) + 1 if len(isp) else 1
    lmrx = np.full(nsp, lmrxx) if lmrx is None else np.asarray(lmrx)
    psc, vbar, g0, v0 = (np.asarray(tabs[k], dtype=np.float64) for k in ("pscorg", "vbarg", "g0", "v0"))
    eigr = [structurefactor(p, plane.gbas, plane.igvec) for p in np.asarray(pos)]
    ylmg = ylmofg(gv, lmrxx)
    cf = cfac(lmrxx)
    # PLANEWAVE$SUPFFT('RTOG') of the density
    rhog = plane.fft_rtog(np.asarray(rho_r, dtype=np.complex128))
    # POTENTIAL_ADDCORE (src/paw_potential.f90:1038-1066)
    for a, e in enumerate(eigr):
        rhog = rhog + psc[isp[a]] * Y0 * e
    # POTENTIAL_HARTREE (src/paw_potential.f90:1069-1334): compensation charges
    # sum_LM CFAC(LM) QLM(LM) YLMOFG(LM,G) G0(G) exp(-iG.R)
    rho2 = rhog.copy()
    pot = np.zeros(plane.ngl, dtype=np.complex128)
    vb = 0.0
    for a, e in enumerate(eigr):
        s = isp[a]
        n = int(lmrx[s])
        cs = (cf[:n] * qlm[a, :n]) @ ylmg[:n] * e
        rho2 = rho2 + cs * g0[s]
        pot = pot + cs * v0[s] + vbar[s] * Y0 * e
        vb += qlm[a, 0] * Y0 * v0[s, ig0]
    rhob = -rho2[ig0].real
    rho2[ig0] = 0.0
    vg = np.zeros(plane.ngl, dtype=np.complex128)
    vg[nonzero] = rho2[nonzero] * FPI / g2[nonzero]
    vg[ig0] = -vb
    esum = np.sum((0.5 * np.conj(rho2[nonzero]) * vg[nonzero] + np.conj(rhog[nonzero]) * pot[nonzero]).real)
    esum += (np.conj(rhog[ig0]) * pot[ig0] + rhob * vb).real
    ehartree = om * esum
    pot = pot + vg
    rho2b = rhog.copy()
    rho2b[ig0] += rhob
    vqlm = np.zeros((len(eigr), lmrxx))
    force = np.zeros((len(eigr), 3))
    for a, e in enumerate(eigr):
        s = isp[a]
        n = int(lmrx[s])
        aval = (g0[s] * np.conj(vg) + v0[s] * np.conj(rho2b)) * e
        csv = cf[:n, None] * ylmg[:n]                       # CFAC(LM) YLMOFG(LM,G)
        vqlm[a, :n] = om * np.sum((csv * aval[None, :]).real, axis=1)
        csum = qlm[a, :n] @ csv
        cval = vbar[s] * Y0 * e * np.conj(rhog)
        sv = (cval + csum * aval).imag
        force[a] = -om * (gv * sv[:, None]).sum(axis=0)
    if flat:
        vqlm = vqlm[:, 0]
    # PLANEWAVE$SUPFFT('GTOR') of density + core, POTENTIAL_XC (src/paw_potential.f90:1339-1545)
    rho_tot = plane.fft_gtor(rhog).real
    exc_r, vxc_r = lda(rho_tot)
    exc = float(np.sum(exc_r)) * om / float(plane.nrl)
    # no gradient correction: RHOG = RTOG(vxc) + Hartree potential
    vtot = plane.fft_rtog(vxc_r.astype(np.complex128)) + pot
    # POTENTIAL_FPSCORE (src/paw_potential.f90:897-975)
    for a, e in enumerate(eigr):
        sv = -(np.conj(vtot) * psc[isp[a]] * e).imag
        force[a] += Y0 * om * (gv * sv[:, None]).sum(axis=0)
    v = plane.fft_gtor(vtot).real
    return {"v": v, "ehartree": float(ehartree), "exc": exc, "rhob": float(rhob), "vqlm": vqlm,
            "force": force}
