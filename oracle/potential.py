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


# ---------------------------------------------------------------------------
# Functional type 10 (src/paw_dft.f90:248-262): PBE exchange (EXCHANGE type 3,
# src/paw_dft.f90:1647-1651, 1706-1720) and PBE correlation with the
# Perdew-Wang 1992 local part (PBE$EVAL1, src/paw_dft.f90:6416-6519;
# PERDEWWANG91L, :4398-4830), without spin; no table interpolation
# (TSAFE, src/paw_dft.f90:56).
# ---------------------------------------------------------------------------
# PW92 parameters of the unpolarised gas (src/paw_dft.f90:4428-4433)
PW_A, PW_ALPHA1, PW_BETA = 3.10907e-2, 2.1370e-1, (7.5957, 3.5876, 1.6382, 4.9294e-1)
# PBE: kappa, beta = nu * CC0, mu = beta pi^2 / 3, gamma = (1 - ln 2)/pi^2
# (src/paw_dft.f90:5846-5896, 1706-1720); t^2 = TFAC |grad rho|^2 / (phi^2 rho^(7/3))
# with the corrected TFAC (T_BUGFIX1, src/paw_dft.f90:5870-5879); the reduced
# exchange gradient s^2 = S2FAC |grad rho|^2 / rho^(8/3) (src/paw_dft.f90:1693)
PBE_CC0 = 4.235e-3
PBE_NU = 16.0 / PI * (3.0 * PI ** 2) ** (1.0 / 3.0)
PBE_BETA = PBE_NU * PBE_CC0
PBE_MU = PBE_BETA * PI ** 2 / 3.0
PBE_KAPPA = 0.804
PBE_GAMMA = (1.0 - np.log(2.0)) / PI ** 2
PBE_TFAC = PI / (16.0 * (3.0 * PI ** 2) ** (1.0 / 3.0))
PBE_S2FAC = 0.25 / KFFAC ** 2


def pw92_unpolarised(rho):
    """PERDEWWANG91L_RHODEP1 (src/paw_dft.f90:4519-4574): eps_c(rho) of the
    unpolarised gas, G(rs) = -2A (1 + a1 rs) ln(1 + 1/(2A (b1 rs^1/2 + b2 rs +
    b3 rs^3/2 + b4 rs^2))), and d eps_c / d rho.  rho > 0 (array)."""
    rs = RSFAC / rho ** (1.0 / 3.0)
    sq = np.sqrt(rs)
    b1, b2, b3, b4 = PW_BETA
    q0 = -2.0 * PW_A * (1.0 + PW_ALPHA1 * rs)
    q1 = 2.0 * PW_A * (b1 * sq + b2 * rs + b3 * rs * sq + b4 * rs * rs)
    dq1 = PW_A * (b1 / sq + 2.0 * b2 + 3.0 * b3 * sq + 4.0 * b4 * rs)
    lg = np.log1p(1.0 / q1)
    eps = q0 * lg
    deps_drs = -2.0 * PW_A * PW_ALPHA1 * lg - q0 * dq1 / (q1 * q1 + q1)
    return eps, deps_drs * (-rs / (3.0 * rho))


def pbe_exchange(rho, g2):
    """EXCHANGE_X1 with the PBE enhancement (src/paw_dft.f90:2137-2236,
    EXCHANGE_FX2 :2577-2603): EX = EXFAC rho^(4/3) Fx(s^2), Fx = 1 + kappa -
    kappa/(1 + mu s^2/kappa); returns EX, dEX/drho, dEX/d|grad rho|^2."""
    rho43 = rho ** (4.0 / 3.0)
    exl = EXFAC * rho43
    s2g = PBE_S2FAC / rho43 ** 2            # d s^2 / d |grad rho|^2
    s2 = s2g * g2
    den = 1.0 + PBE_MU / PBE_KAPPA * s2
    fx = 1.0 + PBE_KAPPA - PBE_KAPPA / den
    dfx = PBE_MU / (den * den)              # d Fx / d s^2
    ex = exl * fx
    ex_r = exl * (4.0 / 3.0) / rho * fx + exl * dfx * (-8.0 / 3.0) * s2 / rho
    ex_g = exl * dfx * s2g
    return ex, ex_r, ex_g


def pbe_correlation(rho, g2):
    """PBE$EVAL1 without spin (phi = 1): EC = rho (eps_c + H),
    H = gamma ln(1 + (beta/gamma) t^2 (1 + A t^2)/(1 + A t^2 + A^2 t^4)),
    A = (beta/gamma)/(exp(-eps_c/gamma) - 1) (PBE_H0OFTEG1 :5919-5977,
    PBE_FOFTV1 :6174-6234); returns EC, dEC/drho, dEC/d|grad rho|^2."""
    eps, deps = pw92_unpolarised(rho)
    tn = PBE_TFAC / rho ** (7.0 / 3.0)      # d t^2 / d |grad rho|^2
    t2 = tn * g2
    c = PBE_BETA / PBE_GAMMA
    x = -eps / PBE_GAMMA
    big = x > 100.0                          # 1/A -> infinity: F = (beta/gamma) t^2
    ex = np.exp(np.where(big, 0.0, x))
    ai = (ex - 1.0) / c                      # 1/A
    dai_deps = -ex / (c * PBE_GAMMA)
    den = ai * ai + ai * t2 + t2 * t2
    f = np.where(big, c * t2, c * ai * t2 * (ai + t2) / den)
    f_t = np.where(big, c, c * ai ** 3 * (2.0 * t2 + ai) / den ** 2)
    f_a = np.where(big, 0.0, c * t2 ** 3 * (2.0 * ai + t2) / den ** 2)
    h = PBE_GAMMA * np.log1p(f)
    h_f = PBE_GAMMA / (1.0 + f)
    h_eps = h_f * f_a * dai_deps
    h_t = h_f * f_t
    h_rho = h_t * (-7.0 / 3.0) * t2 / rho + h_eps * deps
    ec = rho * (eps + h)
    ec_r = eps + h + rho * (deps + h_rho)
    ec_g = rho * h_t * tn
    return ec, ec_r, ec_g


def pbe(rho, g2):
    """Energy density, dE/drho and dE/d|grad rho|^2 of functional type 10 at
    grid points, as POTENTIAL_XC (src/paw_potential.f90:1397-1401: rho clamped
    to >= 1e-8) calls DFT (src/paw_dft.f90:780-965: densities below RHOTMIN
    are evaluated at RHOTMIN with the same gradient and the results scaled
    back linearly)."""
    rhot = np.maximum(np.asarray(rho, dtype=np.float64), 1.0e-8)
    g2 = np.asarray(g2, dtype=np.float64)
    val = np.maximum(rhot, RHOTMIN)
    ex, ex_r, ex_g = pbe_exchange(val, g2)
    ec, ec_r, ec_g = pbe_correlation(val, g2)
    exc, vt, gv = ex + ec, ex_r + ec_r, ex_g + ec_g
    small = rhot < RHOTMIN
    vt = np.where(small, exc / RHOTMIN, vt)
    gv = np.where(small, gv * rhot / RHOTMIN, gv)
    exc = np.where(small, exc * rhot / RHOTMIN, exc)
    return exc, vt, gv


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


def vofrho(plane, rho_r, pos, isp, qlm, tabs, lmrx=None, functional=1):
    """POTENTIAL$VOFRHO -> POTENTIAL_VOFRHO (src/paw_potential.f90:324-484,
    487-896) in the scope of the module docstring.

    plane: the density plane-wave object (oracle.planewave.Plane built with the
      density cutoff for the sphere); rho_r (NR3,NR2,NR1) real pseudo density;
    pos (NAT,3); isp (NAT) species 0..NSP-1; qlm: multipole moments of the
      one-centre densities, QLM(LMRXX,NAT) as (NAT, LMRXX), or (NAT,) for
      LMRXX = 1; lmrx: LMRX per species (default LMRXX for all);
    tabs: dict of (NSP,NGL) real tables on the sphere: 'pscorg' (PS core),
      'vbarg' (V-bar), 'g0' (compensation Gaussians), 'v0' (their potential).
    functional: 1 (local exchange + Perdew-Zunger) or 10 (PBE, with the
      gradient of rho + core from its Fourier coefficients and the divergence
      of 2 dE/d|grad rho|^2 grad rho added in G space, POTENTIAL_GRADIENT
      src/paw_potential.f90:977-1034 and :749-787).
    Returns dict: v (NR3,NR2,NR1) total local potential, ehartree, exc, rhob,
      vqlm (shape of qlm) = dE/dQLM, force (NAT,3) (Hartree + pseudo-core terms).
    """
    if functional not in (1, 10):
        raise ValueError("functional type must be 1 or 10")
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
    if functional == 1:
        exc_r, vxc_r = lda(rho_tot)
        # no gradient correction: RHOG = RTOG(vxc) + Hartree potential
        vtot = plane.fft_rtog(vxc_r.astype(np.complex128)) + pot
    else:
        # grad rho(r) = GTOR(i G rho(G)); after XC the vector 2 dE/d|grad rho|^2
        # grad rho goes back to G space and -i G . F(G) joins the potential
        grad = [plane.fft_gtor(1j * gv[:, k] * rhog).real for k in range(3)]
        g2r = grad[0] ** 2 + grad[1] ** 2 + grad[2] ** 2
        exc_r, vxc_r, gvr = pbe(rho_tot, g2r)
        vtot = plane.fft_rtog(vxc_r.astype(np.complex128)) + pot
        for k in range(3):
            fk = plane.fft_rtog((2.0 * gvr * grad[k]).astype(np.complex128))
            vtot = vtot - 1j * gv[:, k] * fk
    exc = float(np.sum(exc_r)) * om / float(plane.nrl)
    # POTENTIAL_FPSCORE (src/paw_potential.f90:897-975)
    for a, e in enumerate(eigr):
        sv = -(np.conj(vtot) * psc[isp[a]] * e).imag
        force[a] += Y0 * om * (gv * sv[:, None]).sum(axis=0)
    v = plane.fft_gtor(vtot).real
    return {"v": v, "ehartree": float(ehartree), "exc": exc, "rhob": float(rhob), "vqlm": vqlm,
            "force": force}
