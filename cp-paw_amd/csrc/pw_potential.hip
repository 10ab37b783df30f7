// POTENTIAL$VOFRHO (src/paw_potential.f90:324-896) on the GPU, FP64: the
// density -> local potential step that follows WAVES_DENSITY, for one spin,
// the local-density functional of type 1 (local exchange + Perdew-Zunger,
// src/paw_dft.f90:124-130) or the gradient-corrected type 10 (PBE,
// :248-262), compensation charges of the moments QLM(LMRXX,NAT)
// with l <= 4 (LMRXX <= 25), no pair potential of the compensation charges
// and no isolation.
//
// Pipeline on the density plane-wave object (sphere of the density cutoff):
//   rho(G)   = RTOG rho(r) + sum_a PSCORG Y0 e_a          (SUPFFT, ADDCORE :1038-1066)
//   Hartree  per G: compensation sum_LM CFAC(LM) QLM(LM) |G|^l Y_LM(G) G0 e_a,
//            VBAR, V0, background, EHARTREE partials (POTENTIAL_HARTREE
//            :1069-1334, CFAC = (-i)^l/(2l+1)!!, YLMOFG of POTENTIAL_YLMOFG :1548)
//   rho(r)   = GTOR rho(G);  exc(r), vxc(r) by LDA       (POTENTIAL_XC :1339-1545)
//            type 10: grad rho = GTOR(i G rho(G)), PBE, and the divergence
//            -i G . RTOG(2 de/d|grad rho|^2 grad rho) added to vxc(G)
//   vtot(G)  = RTOG vxc(r) + Hartree potential
//   per atom: VQLM(LM), Hartree/VBAR forces and the pseudo-core force
//            (POTENTIAL_FPSCORE :897-975) in one G pass
//   v(r)     = GTOR vtot(G)
// e_a(G) = exp(-i G.R_a) is formed from the integer G components and the
// per-axis phases b_k.R_a (PLANEWAVE$STRUCTUREFACTOR, src/paw_fft.f90:1597).
// The sums are block partials in a fixed order, then one ordered reduction,
// so results do not depend on scheduling.
#include <algorithm>
#include <cmath>

#include "pw_common.h"

namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr double kY0 = 0.28209479177387814347;   // 1/sqrt(4 pi)
constexpr int kPotThreads = 256;

// src/paw_dft.f90:62, 1691-1693, 2783-2792, 2814
constexpr double kRhoTMin = 1.0e-6;
constexpr double kGammaU = -142.3e-3, kBeta1U = 1052.9e-3, kBeta2U = 333.4e-3;
constexpr double kAU = 31.10e-3, kBU = -48.0e-3, kCU = 2.0e-3, kDU = -11.6e-3;

struct PotGeo {
  int ngl, nrl, nat, nsp, ngamma, lmrxx;
  const int *lmrx;       // nsp: LMRX of each species
  const int *igvec;      // 3 x ngl
  const double *gvec;    // 3 x ngl
  const double *g2;      // ngl
  const double *tabs;    // [4][nsp][ngl]: PSCORG, VBARG, G0, V0
  const int *isp;        // nat, 0-based
  const double *gr;      // nat x 3: b_k . R_a
  const double *qlm;     // nat x lmrxx (QLM(LMRXX,NAT))
  double vol;
};

constexpr int kLmrxxMax = 25;   // l <= 4

// |G|^l Y_LM(G) for LM < lmx (POTENTIAL_YLMOFG, src/paw_potential.f90:1548-1614):
// the real harmonics of SPHERICAL$YLM (src/paw_spherical.f90:66-148, recursion
// of SPHERICAL_PLGNDR :366-453) times |G|^l; (Y0, 0, ...) for |G| < 1e-7
__device__ inline void solid_ylm(int lmx, double gx, double gy, double gz, double *y) {
  const double disxy = sqrt(gx * gx + gy * gy);
  const double dis = sqrt(disxy * disxy + gz * gz);
  for (int lm = 0; lm < lmx; ++lm) y[lm] = 0.0;
  y[0] = kY0;
  if (lmx == 1 || dis < 1.0e-7) return;
  const int lx = lmx <= 1 ? 0 : (lmx <= 4 ? 1 : (lmx <= 9 ? 2 : (lmx <= 16 ? 3 : 4)));
  const double x = gz / dis;
  const double fact = sqrt(fmax(1.0 - x * x, 0.0));
  double plm[kLmrxxMax];
  double pmm = 1.0;
  for (int m = 0; m <= lx; ++m) {
    double cm = 0.0, c0 = pmm;
    plm[m * m] = c0;
    for (int l = m + 1; l <= lx; ++l) {
      const double cp = (x * (2 * l - 1) * c0 - (double)(l + m - 1) * cm) / (double)(l - m);
      plm[l * (l + 1) - m] = cp;
      cm = c0;
      c0 = cp;
    }
    pmm = -(double)(2 * m + 1) * fact * pmm;
  }
  for (int l = 1; l <= lx; ++l) {
    const int i1 = l * (l + 1);
    double f = 1.0, sg = 1.0;
    for (int m = 1; m <= l; ++m) {
      f *= sqrt((double)((l - m + 1) * (l + m)));
      sg = -sg;
      plm[i1 - m] /= f;
      plm[i1 + m] = sg * plm[i1 - m];
    }
  }
  const double sinphi = disxy != 0.0 ? gy / disxy : 0.0;
  const double cosphi = disxy != 0.0 ? gx / disxy : 1.0;
  double gl = 1.0;
  for (int l = 0; l <= lx; ++l) {
    const int lm0 = l * (l + 1);
    const double fac = sqrt((2 * l + 1) / (4.0 * kPi)) * gl;
    y[lm0] = plm[lm0] * fac;
    double er = 1.0, ei = 0.0;
    for (int m = 1; m <= l; ++m) {
      const double nr = er * cosphi - ei * sinphi, ni = er * sinphi + ei * cosphi;
      er = nr;
      ei = ni;
      const double sv = fac * plm[lm0 + m] * sqrt(2.0);
      y[lm0 - m] = sv * er;
      y[lm0 + m] = sv * ei;
    }
    gl *= dis;
  }
}

// CFAC(LM) = (-i)^l/(2l+1)!! as (re, im) (src/paw_potential.f90:1146-1152)
__device__ inline double2 cfac_l(int l) {
  double df = 1.0;
  for (int k = 3; k <= 2 * l + 1; k += 2) df *= (double)k;
  switch (l & 3) {
    case 0: return make_double2(1.0 / df, 0.0);
    case 1: return make_double2(0.0, -1.0 / df);
    case 2: return make_double2(-1.0 / df, 0.0);
    default: return make_double2(0.0, 1.0 / df);
  }
}

__device__ inline int l_of(int lm) { return lm < 1 ? 0 : (lm < 4 ? 1 : (lm < 9 ? 2 : (lm < 16 ? 3 : 4))); }

// sum_{LM < LMRX} CFAC(LM) QLM(LM) YLMOFG(LM) (complex)
__device__ inline double2 comp_factor(const PotGeo &g, int a, int n, const double *y) {
  double re = 0.0, im = 0.0;
  const double *q = g.qlm + (size_t)a * g.lmrxx;
  for (int lm = 0; lm < n; ++lm) {
    const double2 c = cfac_l(l_of(lm));
    re += c.x * q[lm] * y[lm];
    im += c.y * q[lm] * y[lm];
  }
  return make_double2(re, im);
}

__device__ inline double2 eigr(const PotGeo &g, int a, int ig) {
  const double ph = (double)g.igvec[3 * ig] * g.gr[3 * a] + (double)g.igvec[3 * ig + 1] * g.gr[3 * a + 1] +
                    (double)g.igvec[3 * ig + 2] * g.gr[3 * a + 2];
  double s, c;
  sincos(ph, &s, &c);
  return make_double2(c, -s);
}

__device__ inline double tab(const PotGeo &g, int which, int s, int ig) {
  return g.tabs[((size_t)which * g.nsp + s) * g.ngl + ig];
}

// deterministic block sum of v over kPotThreads threads; result in thread 0
__device__ inline double block_sum(double v, double *sh) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) sh[wv] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0)
    for (int w = 0; w < kPotThreads / 64; ++w) r += sh[w];
  __syncthreads();
  return r;
}

__global__ void k_real_to_c(int n, const double *__restrict__ r, double2 *__restrict__ c) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) c[i] = make_double2(r[i], 0.0);
}

__global__ void k_c_to_real(int n, const double2 *__restrict__ c, double *__restrict__ r) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) r[i] = c[i].x;
}

// Per G: core, compensation, VBAR and V0 sums over atoms; the background and
// the Hartree potential; EHARTREE partial sums (src/paw_potential.f90:
// 1038-1066, 1145-1205).  rhog gains the core; vg = VG; vh = POT + VG.
// scal[0] = RHOB (written by the thread of G = 0).
__global__ void __launch_bounds__(kPotThreads) k_pot_g(PotGeo g, double2 *__restrict__ rhog,
                                                       double2 *__restrict__ vg_out, double2 *__restrict__ vh,
                                                       double *__restrict__ epart, double *__restrict__ scal) {
  __shared__ double sh[kPotThreads / 64];
  const int ig = blockIdx.x * kPotThreads + threadIdx.x;
  double e = 0.0;
  if (ig < g.ngl) {
    double2 core = {0.0, 0.0}, comp = {0.0, 0.0}, pot = {0.0, 0.0};
    double vb = 0.0;
    double y[kLmrxxMax];
    solid_ylm(g.lmrxx, g.gvec[3 * ig], g.gvec[3 * ig + 1], g.gvec[3 * ig + 2], y);
    for (int a = 0; a < g.nat; ++a) {
      const int s = g.isp[a];
      const double2 ea = eigr(g, a, ig);
      // CSVAR = sum_LM CFAC(LM) QLM(LM) YLMOFG(LM,G) EIGR(G)
      const double2 cs = cmul(comp_factor(g, a, g.lmrx[s], y), ea);
      const double fc = tab(g, 0, s, ig) * kY0;
      const double g0 = tab(g, 2, s, ig), v0 = tab(g, 3, s, ig), vbar = tab(g, 1, s, ig) * kY0;
      core.x += fc * ea.x;
      core.y += fc * ea.y;
      comp.x += cs.x * g0;
      comp.y += cs.y * g0;
      pot.x += cs.x * v0 + vbar * ea.x;
      pot.y += cs.y * v0 + vbar * ea.y;
      if (ig == g.ngamma) vb += g.qlm[(size_t)a * g.lmrxx] * kY0 * v0;
    }
    double2 rg = rhog[ig];
    rg.x += core.x;
    rg.y += core.y;
    rhog[ig] = rg;
    double2 r2 = {rg.x + comp.x, rg.y + comp.y};
    double2 v;
    if (ig == g.ngamma) {
      const double rhob = -r2.x;
      v = make_double2(-vb, 0.0);
      e = (rg.x * pot.x + rg.y * pot.y) + rhob * vb;
      scal[0] = rhob;
    } else {
      const double f = 4.0 * kPi / g.g2[ig];
      v = make_double2(r2.x * f, r2.y * f);
      e = 0.5 * (r2.x * v.x + r2.y * v.y) + (rg.x * pot.x + rg.y * pot.y);
    }
    vg_out[ig] = v;
    vh[ig] = make_double2(pot.x + v.x, pot.y + v.y);
  }
  const double s = block_sum(e, sh);
  if (threadIdx.x == 0) epart[blockIdx.x] = s;
}

// LDA on the grid (POTENTIAL_XC, src/paw_potential.f90:1397-1401, and DFT,
// src/paw_dft.f90:780-965, 2137-2195, 2824-2910): exc(r) partial sums, and
// c(r) <- vxc(r) for the RTOG that follows
__device__ inline void lda_point(double rho, double &exc, double &vxc) {
  const double rhot = fmax(rho, 1.0e-8);
  const double val = fmax(rhot, kRhoTMin);
  const double kf = cbrt(3.0 * kPi * kPi);
  const double exfac = -3.0 / (4.0 * kPi) * kf;
  const double r13 = cbrt(val);
  const double r43 = (r13 * r13) * (r13 * r13);
  const double ex = exfac * r43;
  const double vx = ex * (4.0 / 3.0) / val;
  const double rs = cbrt(3.0 / (4.0 * kPi)) / r13;
  double ec, vc;
  if (rs < 1.0) {
    const double l = log(rs);
    ec = val * (kBU + kDU * rs + (kAU + kCU * rs) * l);
    vc = (1.0 / 3.0) * (-kAU + 3.0 * kBU - (kCU - 2.0 * kDU) * rs + 3.0 * kAU * l + 2.0 * kCU * rs * l);
  } else {
    const double sq = sqrt(rs);
    const double sv = 1.0 / (1.0 + sq * (kBeta1U + kBeta2U * sq));
    const double sv1 = sv / (6.0 * val);
    ec = kGammaU * val * sv;
    vc = ec * sv1 * (6.0 + sq * (7.0 * kBeta1U + 8.0 * kBeta2U * sq));
  }
  exc = ex + ec;
  vxc = vx + vc;
  if (rhot < kRhoTMin) {
    vxc = exc / kRhoTMin;
    exc = exc * rhot / kRhoTMin;
  }
}

__global__ void __launch_bounds__(kPotThreads) k_xc(int n, double2 *__restrict__ c, double *__restrict__ xpart) {
  __shared__ double sh[kPotThreads / 64];
  const int i = blockIdx.x * kPotThreads + threadIdx.x;
  double e = 0.0;
  if (i < n) {
    double vx;
    lda_point(c[i].x, e, vx);
    c[i] = make_double2(vx, 0.0);
  }
  const double s = block_sum(e, sh);
  if (threadIdx.x == 0) xpart[blockIdx.x] = s;
}

// Functional type 10 (src/paw_dft.f90:248-262) without spin: PBE exchange
// (EXCHANGE_X1 with EXCHANGE_FX2's PBE enhancement, src/paw_dft.f90:2137-2236,
// 2577-2603) and PBE correlation over the Perdew-Wang 1992 local part
// (PBE$EVAL1 :6416-6519, PBE_H0OFTEG1 :5919, PBE_FOFTV1 :6174,
// PERDEWWANG91L_RHODEP1 :4519), called as DFT does (:780-965: densities below
// RHOTMIN evaluated at RHOTMIN with the same gradient and scaled back).
// e: energy density; vt = de/drho; gv = de/d|grad rho|^2.
constexpr double kPwA = 3.10907e-2, kPwAlpha1 = 2.1370e-1;
constexpr double kPwB1 = 7.5957, kPwB2 = 3.5876, kPwB3 = 1.6382, kPwB4 = 4.9294e-1;
constexpr double kPbeKappa = 0.804, kPbeCC0 = 4.235e-3;

__device__ inline void pbe_point(double rho, double g2, double &e, double &vt, double &gv) {
  const double rhot = fmax(rho, 1.0e-8);
  const double val = fmax(rhot, kRhoTMin);
  const double kf3 = 3.0 * kPi * kPi, kffac = cbrt(kf3);
  const double beta = 16.0 / kPi * kffac * kPbeCC0;
  const double mu = beta * kPi * kPi / 3.0;
  const double gam = (1.0 - log(2.0)) / (kPi * kPi);
  // exchange: EXFAC rho^(4/3) (1 + kappa - kappa/(1 + mu s^2/kappa))
  const double r13 = cbrt(val), r43 = r13 * r13 * r13 * r13;
  const double exl = -3.0 / (4.0 * kPi) * kffac * r43;
  const double s2g = 0.25 / (kffac * kffac) / (r43 * r43);
  const double s2 = s2g * g2;
  const double den = 1.0 + mu / kPbeKappa * s2;
  const double fx = 1.0 + kPbeKappa - kPbeKappa / den;
  const double dfx = mu / (den * den);
  const double ex = exl * fx;
  const double ex_r = exl * (4.0 / 3.0) / val * fx - exl * dfx * (8.0 / 3.0) * s2 / val;
  const double ex_g = exl * dfx * s2g;
  // PW92 eps_c(rs) and d eps_c/d rho
  const double rs = cbrt(3.0 / (4.0 * kPi)) / r13, sq = sqrt(rs);
  const double q0 = -2.0 * kPwA * (1.0 + kPwAlpha1 * rs);
  const double q1 = 2.0 * kPwA * (kPwB1 * sq + kPwB2 * rs + kPwB3 * rs * sq + kPwB4 * rs * rs);
  const double dq1 = kPwA * (kPwB1 / sq + 2.0 * kPwB2 + 3.0 * kPwB3 * sq + 4.0 * kPwB4 * rs);
  const double lg = log1p(1.0 / q1);
  const double eps = q0 * lg;
  const double deps = (-2.0 * kPwA * kPwAlpha1 * lg - q0 * dq1 / (q1 * q1 + q1)) * (-rs / (3.0 * val));
  // H(t^2, eps) = gamma ln(1 + F), F = (beta/gamma) t^2 (1 + A t^2)/(1 + A t^2 + A^2 t^4)
  const double tn = kPi / (16.0 * kffac) / (r43 * r13 * r13 * r13);   // t^2 / |grad rho|^2 = TFAC rho^(-7/3)
  const double t2 = tn * g2;
  const double c = beta / gam;
  const double x = -eps / gam;
  double f, f_t, h_eps;
  if (x > 100.0) {
    f = c * t2;
    f_t = c;
    h_eps = 0.0;
  } else {
    const double ee = exp(x);
    const double ai = (ee - 1.0) / c;          // 1/A
    const double d2 = ai * ai + ai * t2 + t2 * t2;
    f = c * ai * t2 * (ai + t2) / d2;
    f_t = c * ai * ai * ai * (2.0 * t2 + ai) / (d2 * d2);
    const double f_a = c * t2 * t2 * t2 * (2.0 * ai + t2) / (d2 * d2);
    h_eps = gam / (1.0 + f) * f_a * (-ee / (c * gam));
  }
  const double h = gam * log1p(f);
  const double h_t = gam / (1.0 + f) * f_t;
  const double h_r = h_t * (-7.0 / 3.0) * t2 / val + h_eps * deps;
  const double ec = val * (eps + h);
  const double ec_r = eps + h + val * (deps + h_r);
  const double ec_g = val * h_t * tn;
  e = ex + ec;
  vt = ex_r + ec_r;
  gv = ex_g + ec_g;
  if (rhot < kRhoTMin) {
    vt = e / kRhoTMin;
    gv = gv * rhot / kRhoTMin;
    e = e * rhot / kRhoTMin;
  }
}

// F_k(G) = i G_k rho(G), k = 0..2 (POTENTIAL_GRADIENT 'RHO', src/paw_potential.f90:1003-1012)
__global__ void k_grad_g(int ngl, const double *__restrict__ gvec, const double2 *__restrict__ rhog,
                         double2 *__restrict__ f) {
  const int ig = blockIdx.x * blockDim.x + threadIdx.x;
  if (ig >= ngl) return;
  const double2 r = rhog[ig];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const double gk = gvec[3 * ig + k];
    f[(size_t)k * ngl + ig] = make_double2(-gk * r.y, gk * r.x);
  }
}

// PBE on the grid: c = rho + core (real part), d[k] = d rho/dx_k; c <- vt,
// d[k] <- 2 de/d|grad rho|^2 d rho/dx_k (POTENTIAL_XC :1405-1470); exc partials
__global__ void __launch_bounds__(kPotThreads) k_xc_gga(int n, double2 *__restrict__ c, double2 *__restrict__ d,
                                                        double *__restrict__ xpart) {
  __shared__ double sh[kPotThreads / 64];
  const int i = blockIdx.x * kPotThreads + threadIdx.x;
  double e = 0.0;
  if (i < n) {
    const double gx = d[i].x, gy = d[(size_t)n + i].x, gz = d[2 * (size_t)n + i].x;
    double vt, gv;
    pbe_point(c[i].x, gx * gx + gy * gy + gz * gz, e, vt, gv);
    c[i] = make_double2(vt, 0.0);
    const double s = 2.0 * gv;
    d[i] = make_double2(s * gx, 0.0);
    d[(size_t)n + i] = make_double2(s * gy, 0.0);
    d[2 * (size_t)n + i] = make_double2(s * gz, 0.0);
  }
  const double s = block_sum(e, sh);
  if (threadIdx.x == 0) xpart[blockIdx.x] = s;
}

// v(G) += sum_k -i G_k F_k(G) (POTENTIAL_GRADIENT 'POT', :1013-1027)
__global__ void k_div_g(int ngl, const double *__restrict__ gvec, const double2 *__restrict__ f,
                        double2 *__restrict__ v) {
  const int ig = blockIdx.x * blockDim.x + threadIdx.x;
  if (ig >= ngl) return;
  double2 o = v[ig];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const double gk = gvec[3 * ig + k];
    const double2 fk = f[(size_t)k * ngl + ig];
    o.x += gk * fk.y;
    o.y -= gk * fk.x;
  }
  v[ig] = o;
}

__global__ void k_zero_d(int n, double *x) {
  if ((int)threadIdx.x < n) x[threadIdx.x] = 0.0;
}

__global__ void k_copy1(const double *a, double *b) {
  if (threadIdx.x == 0) b[0] = a[0];
}

__global__ void k_add_c(int n, const double2 *__restrict__ a, double2 *__restrict__ b) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = make_double2(a[i].x + b[i].x, a[i].y + b[i].y);
}

// Per (atom, G block): VQLM(LM) and the Hartree/VBAR force terms
// (src/paw_potential.f90:1255-1290) and the pseudo-core force
// (POTENTIAL_FPSCORE, :897-975), as block partial sums
// apart[(a * nblk + block) * (lmrxx + 3) + {0..lmrxx-1: VQLM, lmrxx..+2: force}]
__global__ void __launch_bounds__(kPotThreads) k_pot_atoms(PotGeo g, const double2 *__restrict__ rhog,
                                                           const double2 *__restrict__ vg,
                                                           const double2 *__restrict__ vtot,
                                                           const double *__restrict__ scal,
                                                           double *__restrict__ apart) {
  __shared__ double sh[kLmrxxMax + 3][kPotThreads / 64];
  const int ig = blockIdx.x * kPotThreads + threadIdx.x;
  const int a = blockIdx.y;
  const int nval = g.lmrxx + 3;
  double val[kLmrxxMax + 3];
  for (int k = 0; k < nval; ++k) val[k] = 0.0;
  if (ig < g.ngl) {
    const int s = g.isp[a];
    const int n = g.lmrx[s];
    const double2 ea = eigr(g, a, ig);
    const double2 rg = rhog[ig], v = vg[ig], vt = vtot[ig];
    const double2 r2b = {rg.x + (ig == g.ngamma ? scal[0] : 0.0), rg.y};
    const double g0 = tab(g, 2, s, ig), v0 = tab(g, 3, s, ig);
    // AVAL = (G0 conj(VG) + V0 conj(RHO2)) EIGR
    const double2 b = {g0 * v.x + v0 * r2b.x, -(g0 * v.y + v0 * r2b.y)};
    const double2 aval = {b.x * ea.x - b.y * ea.y, b.x * ea.y + b.y * ea.x};
    double y[kLmrxxMax];
    solid_ylm(g.lmrxx, g.gvec[3 * ig], g.gvec[3 * ig + 1], g.gvec[3 * ig + 2], y);
    // VQLM(LM) += Re(CFAC(LM) YLMOFG(LM,G) AVAL)
    for (int lm = 0; lm < n; ++lm) {
      const double2 c = cfac_l(l_of(lm));
      val[lm] = c.x * y[lm] * aval.x - c.y * y[lm] * aval.y;
    }
    // CSUM = sum_LM QLM(LM) CFAC(LM) YLMOFG(LM,G)
    const double2 csum = comp_factor(g, a, n, y);
    // CVAL = VBARG Y0 EIGR conj(RHO)
    const double vb = tab(g, 1, s, ig) * kY0;
    const double cval_im = vb * (ea.y * rg.x - ea.x * rg.y);
    const double sv = cval_im + csum.x * aval.y + csum.y * aval.x;
    // FPSCORE: -Im(conj(VTOT) PSCORG EIGR) Y0
    const double pc = tab(g, 0, s, ig);
    const double sc = -pc * (vt.x * ea.y - vt.y * ea.x);
    const double t = -sv + kY0 * sc;
#pragma unroll
    for (int k = 0; k < 3; ++k) val[g.lmrxx + k] = g.gvec[3 * ig + k] * t;
  }
  // all values of the block at once: wave sums, then the waves in order
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (int k = 0; k < nval; ++k) {
    double v = val[k];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    if (lane == 0) sh[k][wv] = v;
  }
  __syncthreads();
  if ((int)threadIdx.x < nval) {
    double r = 0.0;
    for (int w = 0; w < kPotThreads / 64; ++w) r += sh[threadIdx.x][w];
    apart[((size_t)a * gridDim.x + blockIdx.x) * nval + threadIdx.x] = r;
  }
}

// Ordered reduction: output o sums in[o * stride * n + k * stride] over
// k < n (one workgroup per output, fixed assignment of terms to threads)
__global__ void __launch_bounds__(kPotThreads) k_sum_out(int n, int stride, const double *__restrict__ in,
                                                         double scale, double *__restrict__ out) {
  __shared__ double sh[kPotThreads / 64];
  const int o = blockIdx.x;
  const int c = o % stride, row = o / stride;
  double s = 0.0;
  for (int k = threadIdx.x; k < n; k += kPotThreads) s += in[((size_t)row * n + k) * stride + c];
  const double r = block_sum(s, sh);
  if (threadIdx.x == 0) out[o] = scale * r;
}

}  // namespace

// Issues the whole step on stream st; the results land in pinned host memory
// (p->h_pot) once the stream has passed this point (vofrho_results reads them)
// and stay on the device (vofrho_device_results).  QLM(LMRXX,NAT) comes from
// the host (qlm) or from device memory (qlm_dev).
int vofrho_enqueue(pw_plan *p, int nat, int lmrxx, const int *isp, const double *pos, const double *qlm,
                   const double *qlm_dev, const double *rho, double *vofr, hipStream_t st) {
  const int ngl = p->ngl, nrl = p->nrl, nsp = p->pot_nsp;
  if (!p->d_pot_tabs || nsp <= 0)
    return pw_fail(PW_ERR_STATE, "POTENTIAL$VOFRHO: setup tables not set (pw_potential_tables)");
  if (lmrxx < 1 || lmrxx > kLmrxxMax)
    return pw_fail(PW_ERR_UNSUPPORTED, "POTENTIAL$VOFRHO: LMRXX must be 1..25 (l <= 4)");
  const int lx = (int)std::lround(std::sqrt((double)lmrxx)) - 1;
  if ((lx + 1) * (lx + 1) != lmrxx) return pw_fail(PW_ERR_ARG, "POTENTIAL$VOFRHO: LMRXX is not a full shell");
  for (int s = 0; s < nsp; ++s)
    if (p->pot_lmrx[s] > lmrxx) return pw_fail(PW_ERR_ARG, "POTENTIAL$VOFRHO: LMRX of a species exceeds LMRXX");
  PW_TRY(plan_device(p));
  const int nbg = (ngl + kPotThreads - 1) / kPotThreads;
  const int nbr = (nrl + kPotThreads - 1) / kPotThreads;
  const int nval = lmrxx + 3;
  // work: grid (complex), rhog, vg, vh, vtot (sphere), partials, per-atom data
  const bool gga = p->pot_xc == 10;
  const size_t need = (size_t)nrl * 16 + 4 * (size_t)ngl * 16 + (gga ? 3 * ((size_t)nrl + ngl) * 16 : 0) +
                      ((size_t)nbg * (nval * (size_t)nat + 1) + nbr + 8) * 8 +
                      ((size_t)nat * nval + 3) * 8 + (size_t)nat * (8 + 3 * 8 + lmrxx * 8) + (size_t)nsp * 8 + 4096;
  if (need > p->pot_bytes) {
    if (p->d_pot) PW_HIP(hipFree(p->d_pot));
    p->d_pot = nullptr;
    p->pot_bytes = 0;
    PW_HIP(hipMalloc(&p->d_pot, need));
    p->pot_bytes = need;
  }
  // pinned: results (nat * nval + 3), then staged isp, b.R, QLM, LMRX
  const size_t hres = (size_t)nat * nval + 3;
  const size_t hneed =
      (hres + (size_t)nat + 3 * (size_t)nat + (size_t)nat * lmrxx + (size_t)nsp + 8) * sizeof(double);
  if (hneed > p->h_pot_bytes) {
    if (p->h_pot) PW_HIP(hipHostFree(p->h_pot));
    p->h_pot = nullptr;
    p->h_pot_bytes = 0;
    PW_HIP(hipHostMalloc((void **)&p->h_pot, hneed, hipHostMallocDefault));
    p->h_pot_bytes = hneed;
  }
  char *w = (char *)p->d_pot;
  double2 *grid = (double2 *)w; w += (size_t)nrl * 16;
  double2 *rhog = (double2 *)w; w += (size_t)ngl * 16;
  double2 *vg = (double2 *)w; w += (size_t)ngl * 16;
  double2 *vh = (double2 *)w; w += (size_t)ngl * 16;
  double2 *vtot = (double2 *)w; w += (size_t)ngl * 16;
  double *epart = (double *)w; w += (size_t)nbg * 8;
  double *xpart = (double *)w; w += (size_t)nbr * 8;
  double *apart = (double *)w; w += (size_t)nbg * nval * nat * 8;
  double *scal = (double *)w; w += 8 * 8;
  double *res = (double *)w; w += hres * 8;   // per atom VQLM(LMRXX), force(3); EHARTREE, EXC, RHOB
  int *d_isp = (int *)w; w += (size_t)nat * 8;
  int *d_lmrx = (int *)w; w += (size_t)nsp * 8;
  double *d_gr = (double *)w; w += (size_t)nat * 3 * 8;
  double *d_qlm = (double *)w; w += (size_t)nat * lmrxx * 8;
  w = (char *)p->d_pot + ((size_t)(w - (char *)p->d_pot) + 255) / 256 * 256;   // inside the 4096-byte slack
  // the gradient of rho + core on the grid and its Fourier coefficients (type 10)
  double2 *grid3 = gga ? (double2 *)w : nullptr;
  double2 *fg3 = gga ? grid3 + 3 * (size_t)nrl : nullptr;
  // per-atom data: 0-based species, b_k . R_a, QLM (staged in pinned memory)
  double *hbase = p->h_pot + hres;
  int *hisp = (int *)hbase;
  int *hlmrx = hisp + nat;
  double *hgr = hbase + nat + (nsp + 1) / 2 + 1;
  double *hq = hgr + 3 * (size_t)nat;
  for (int a = 0; a < nat; ++a) {
    if (isp[a] < 1 || isp[a] > nsp) return pw_fail(PW_ERR_ARG, "POTENTIAL$VOFRHO: species index out of range");
    hisp[a] = isp[a] - 1;
    for (int k = 0; k < 3; ++k)
      hgr[3 * a + k] = p->gbas[0 + 3 * k] * pos[3 * a] + p->gbas[1 + 3 * k] * pos[3 * a + 1] +
                       p->gbas[2 + 3 * k] * pos[3 * a + 2];
    if (qlm)
      for (int lm = 0; lm < lmrxx; ++lm) hq[(size_t)a * lmrxx + lm] = qlm[(size_t)a * lmrxx + lm];
  }
  for (int s = 0; s < nsp; ++s) hlmrx[s] = p->pot_lmrx[s];
  PW_HIP(hipMemcpyAsync(d_lmrx, hlmrx, nsp * sizeof(int), hipMemcpyHostToDevice, st));
  if (nat > 0) {
    PW_HIP(hipMemcpyAsync(d_isp, hisp, nat * sizeof(int), hipMemcpyHostToDevice, st));
    PW_HIP(hipMemcpyAsync(d_gr, hgr, 3 * nat * sizeof(double), hipMemcpyHostToDevice, st));
    if (qlm)
      PW_HIP(hipMemcpyAsync(d_qlm, hq, (size_t)nat * lmrxx * sizeof(double), hipMemcpyHostToDevice, st));
    else
      PW_HIP(hipMemcpyAsync(d_qlm, qlm_dev, (size_t)nat * lmrxx * sizeof(double), hipMemcpyDeviceToDevice, st));
  }
  PotGeo g{};
  g.ngl = ngl;
  g.nrl = nrl;
  g.nat = nat;
  g.nsp = nsp;
  g.ngamma = p->ngamma;
  g.lmrxx = lmrxx;
  g.lmrx = d_lmrx;
  g.igvec = p->d_igvec;
  g.gvec = p->d_gvec;
  g.g2 = p->d_g2;
  g.tabs = p->d_pot_tabs;
  g.isp = d_isp;
  g.gr = d_gr;
  g.qlm = d_qlm;
  g.vol = p->vol;
  // rho(G) of the valence density
  k_real_to_c<<<nbr, kPotThreads, 0, st>>>(nrl, rho, grid);
  PW_TRY(fft_rtog_full(p, 1, rhog, grid, PW_PREC_FP64, st));
  // core, compensation, Hartree
  k_zero_d<<<1, 64, 0, st>>>(8, scal);
  k_pot_g<<<nbg, kPotThreads, 0, st>>>(g, rhog, vg, vh, epart, scal);
  // density + core on the grid, XC, back to G
  PW_TRY(fft_gtor_full(p, 1, rhog, grid, PW_PREC_FP64, st));
  if (gga) {
    // POTENTIAL_GRADIENT('RHO'), DFT type 10, POTENTIAL_GRADIENT('POT')
    k_grad_g<<<nbg, kPotThreads, 0, st>>>(ngl, p->d_gvec, rhog, fg3);
    PW_TRY(fft_gtor_full(p, 3, fg3, grid3, PW_PREC_FP64, st));
    k_xc_gga<<<nbr, kPotThreads, 0, st>>>(nrl, grid, grid3, xpart);
    PW_TRY(fft_rtog_full(p, 1, vtot, grid, PW_PREC_FP64, st));
    PW_TRY(fft_rtog_full(p, 3, fg3, grid3, PW_PREC_FP64, st));
    k_div_g<<<nbg, kPotThreads, 0, st>>>(ngl, p->d_gvec, fg3, vtot);
  } else {
    k_xc<<<nbr, kPotThreads, 0, st>>>(nrl, grid, xpart);
    PW_TRY(fft_rtog_full(p, 1, vtot, grid, PW_PREC_FP64, st));
  }
  k_add_c<<<nbg, kPotThreads, 0, st>>>(ngl, vh, vtot);
  // per-atom sums
  if (nat > 0) {
    k_pot_atoms<<<dim3(nbg, nat), kPotThreads, 0, st>>>(g, rhog, vg, vtot, scal, apart);
    k_sum_out<<<nval * nat, kPotThreads, 0, st>>>(nbg, nval, apart, p->vol, res);
  }
  k_sum_out<<<1, kPotThreads, 0, st>>>(nbg, 1, epart, p->vol, res + (size_t)nval * nat);
  k_sum_out<<<1, kPotThreads, 0, st>>>(nbr, 1, xpart, p->vol / (double)nrl, res + (size_t)nval * nat + 1);
  k_copy1<<<1, 64, 0, st>>>(scal, res + (size_t)nval * nat + 2);
  // v(r)
  PW_TRY(fft_gtor_full(p, 1, vtot, grid, PW_PREC_FP64, st));
  k_c_to_real<<<nbr, kPotThreads, 0, st>>>(nrl, grid, vofr);
  PW_HIP(hipGetLastError());
  PW_HIP(hipMemcpyAsync(p->h_pot, res, hres * sizeof(double), hipMemcpyDeviceToHost, st));
  p->pot_nat = nat;
  p->pot_lmrxx = lmrxx;
  p->d_pot_res = res;
  return PW_OK;
}

// host results: energies {EHARTREE, EXC, RHOB}, VQLM(LMRXX,NAT), force(3,NAT)
void vofrho_results(const pw_plan *p, int nat, double *energies, double *vqlm, double *force) {
  const double *h = p->h_pot;
  const int lmrxx = p->pot_lmrxx, nval = lmrxx + 3;
  energies[0] = h[(size_t)nval * nat];
  energies[1] = h[(size_t)nval * nat + 1];
  energies[2] = p->ngamma >= 0 ? h[(size_t)nval * nat + 2] : 0.0;
  for (int a = 0; a < nat; ++a) {
    if (vqlm)
      for (int lm = 0; lm < lmrxx; ++lm) vqlm[(size_t)a * lmrxx + lm] = h[(size_t)a * nval + lm];
    if (force)
      for (int k = 0; k < 3; ++k) force[3 * a + k] = h[(size_t)a * nval + lmrxx + k];
  }
}


// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" int pw_potential_tables(pw_plan *p, int nsp, const int *lmrx, const double *tabs) {
  if (!p || nsp <= 0 || !tabs) return pw_fail(PW_ERR_ARG, "pw_potential_tables: bad argument");
  std::vector<int> l(nsp, 1);
  if (lmrx)
    for (int s = 0; s < nsp; ++s) {
      const int lx = (int)std::lround(std::sqrt((double)lmrx[s])) - 1;
      if (lmrx[s] < 1 || lmrx[s] > kLmrxxMax || (lx + 1) * (lx + 1) != lmrx[s])
        return pw_fail(PW_ERR_ARG, "pw_potential_tables: LMRX must be a full shell (L+1)^2 with L <= 4");
      l[s] = lmrx[s];
    }
  PW_TRY(plan_device(p));
  const size_t n = 4 * (size_t)nsp * p->ngl;
  if (p->d_pot_tabs) PW_HIP(hipFree(p->d_pot_tabs));
  p->d_pot_tabs = nullptr;
  p->pot_nsp = 0;
  PW_HIP(hipMalloc(&p->d_pot_tabs, n * sizeof(double)));
  PW_HIP(hipMemcpy(p->d_pot_tabs, tabs, n * sizeof(double), hipMemcpyHostToDevice));
  p->pot_nsp = nsp;
  p->pot_lmrx = l;
  return PW_OK;
}

extern "C" int pw_potential_functional(pw_plan *p, int type) {
  if (!p) return pw_fail(PW_ERR_ARG, "pw_potential_functional: null plan");
  if (type != 1 && type != 10)
    return pw_fail(PW_ERR_UNSUPPORTED, "DFT$SETI4('TYPE'): only type 1 (LDA, Perdew-Zunger) and 10 (PBE) "
                                       "are supported");
  PW_HIP(hipStreamSynchronize(p->stream));
  p->pot_xc = type;
  return PW_OK;
}

extern "C" int pw_vofrho(pw_plan *p, int nat, int lmrxx, const int *isp, const double *pos, const double *qlm,
                         const double *rho, double *vofr, double *energies, double *vqlm, double *force) {
  if (!p || nat < 0 || (nat > 0 && (!isp || !pos || !qlm)) || !rho || !vofr || !energies)
    return pw_fail(PW_ERR_ARG, "pw_vofrho: bad argument");
  PW_TRY(vofrho_enqueue(p, nat, lmrxx, isp, pos, qlm, nullptr, rho, vofr, p->stream));
  PW_HIP(hipStreamSynchronize(p->stream));
  vofrho_results(p, nat, energies, vqlm, force);
  return PW_OK;
}
