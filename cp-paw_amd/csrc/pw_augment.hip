// The one-centre terms of the CP step on the GPU, FP64: WAVES$MOMENTS ->
// AUGMENTATION$MOMENTS (src/paw_waves1.f90:4147-4204, src/paw_augmentation.f90:
// 130-220), WAVES$SPHERE -> AUGMENTATION$SPHERE (src/paw_waves1.f90:4418-4512,
// src/paw_augmentation.f90:223-600) and WAVES$ADDCONSTANTPOT
// (src/paw_waves1.f90:4515-4567) for one spin (NDIMD=1), the local-density
// functional of type 1, no soft core and no external one-centre potential.
//
// One workgroup per atom.  The radial functions live on the reference's
// shifted logarithmic grid (SHLOGRADIAL, src/paw_radial.f90:1571-2200);
// RADIAL$INTEGRAL is a weighted sum with the end corrections of
// RADIAL_INTEGRATEEQUISPACED (:397-458) folded into per-point weights, and
// RADIAL$INTEGRATE / RADIAL$POISSON (:1345-1386) run one radial function per
// wavefront: each lane holds consecutive points, the prefix sum is a lane
// scan followed by a wavefront scan of the lane totals.
//   k_sph_rho:    DENMAT (hermitised) -> AE and PS one-centre densities
//                 rho_LM(r) = sum CG(LM1,LM2,LM) D phi phi (AUGMENTATION_RHO
//                 :603-656) and the moments QLM (AUGMENTATION_QLM :659-696)
//   (POTENTIAL$VOFRHO with these QLM gives VQLM and RHOB)
//   k_sph_sphere: ADJUSTVQLM, PS/AE Hartree, XC to second order in the
//                 non-spherical parts, POTB, DATH = DEKIN + <aephi|AEPOT|aephi>
//                 - <psphi|PSPOT|psphi> (AUGMENTATION_EXPECT :2159-2239)
//   k_sph_finish: POTB = sum POTB_a / Omega, E1C, DH += POTB DO, V += POTB
// Every sum runs in a fixed order, so the results do not depend on scheduling.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "pw_common.h"

namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr double kY0 = 0.28209479177387814347;   // 1/sqrt(4 pi)
constexpr int kSphThreads = 256;                 // four wavefronts
constexpr int kSphWaves = kSphThreads / 64;
constexpr int kMaxNr = 512;
constexpr int kPerLane = kMaxNr / 64;
constexpr int kMaxLmrx = 25;
constexpr int kMaxLnx = 8;
constexpr int kMaxLmnx = 32;

// src/paw_dft.f90:62, 1691-1693, 2783-2792, 2814
constexpr double kRhoTMin = 1.0e-6;
constexpr double kGammaU = -142.3e-3, kBeta1U = 1052.9e-3, kBeta2U = 333.4e-3;
constexpr double kAU = 31.10e-3, kBU = -48.0e-3, kCU = 2.0e-3, kDU = -11.6e-3;

// per species: scalars and offsets into the packed table (doubles)
struct SphSpec {
  int nr, lnx, lmnx, lmrx;
  double aez, rnuc, rcsm, r1;
  int o_r, o_drdx, o_wint, o_aephi, o_psphi, o_aecore, o_pscore, o_vadd, o_dekin;
  // Gaunt entries (lmn1, lmn2, lm3, cg) in two orders: by (lm3, ln1, ln2) with
  // offsets per (lm3, ln1, ln2), and by (lmn1, lmn2) with offsets per pair
  int o_rho_ent, o_rho_off, o_exp_ent, o_exp_off;
  double core_q;   // int (AECORE - PSCORE) r^2 dr - AEZ Y0
};

struct SphEnt {
  short lmn1, lmn2, ln1, ln2, lm3, pad;
  double cg;
};

struct SphGeo {
  int nat, lmnxx, lmrxx, nvq;          // nvq = LMRXX + 3: stride of VQLM in the potential results
  const int *isp;                      // nat, 0-based
  const SphSpec *spec;                 // nsp
  const double *tab;                   // packed radial tables
  const SphEnt *ent;                   // packed Gaunt entries
  const int *off;                      // packed offsets
  const double *denmat;                // [nat][L][L] row-major, not yet hermitised
  double *work;                        // per atom [4][lmrxx][kMaxNr]: rho_ae, rho_ps, pot_ae, pot_ps
  double *qlm;                         // [nat][lmrxx]
  const double *vqlm;                  // VQLM (stride nvq per atom); nullptr: zero
  const double *rhob;                  // RHOB; nullptr: zero
  double2 *dh;                         // [nat][L][L] out (DATH)
  double *out;                         // per atom [8]: e1c, potb1, ekin, aexc, psexc, aeeh, pseh, corexc
};

__device__ inline int l_of(int lm) { return lm < 1 ? 0 : (lm < 4 ? 1 : (lm < 9 ? 2 : (lm < 16 ? 3 : 4))); }

__device__ inline double ipow(double r, int k) {
  double p = 1.0;
  for (int i = 0; i < k; ++i) p *= r;
  return p;
}

// GAUSSN (src/paw_generalpurpose.f90:1088-1121)
__device__ inline double gaussn(int l, double alpha) {
  double ifac = 1.0;
  for (int i = 2; i <= l + 1; ++i) ifac *= (double)(2 * i - 1);
  const double rint = sqrt(kPi / (4.0 * alpha)) * ifac / ipow(2.0 * alpha, l + 1);
  return 1.0 / rint;
}

// orders the LDS / global accesses of the lanes of one wavefront around this
// point (memory written by one lane and read by another of the same wave)
__device__ inline void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
  __builtin_amdgcn_wave_barrier();
}

__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

// RADIAL$INTEGRAL of f (nr values) with the per-point weights wint
// (RADIAL_INTEGRATEEQUISPACED's G(NR), times dr/dx); one wavefront
__device__ inline double wave_integral(const double *f, const double *wint, int nr, int lane) {
  double s = 0.0;
  for (int i = lane; i < nr; i += 64) s += wint[i] * f[i];
  return wave_sum(s);
}

// RADIAL$INTEGRATE (SHLOGRADIAL$INTEGRATE -> RADIAL_INTEGRATEEQUISPACED,
// src/paw_radial.f90:1868-1894, 397-458): g = antiderivative of fb/drdx,
// where fb = f dr/dx is given; one wavefront, lane holds kPerLane points.
__device__ void wave_integrate(const double *fb, double *g, int nr, int lane) {
  constexpr double c0 = -31.0 / 48.0, c1 = 11.0 / 48.0, c2 = -5.0 / 48.0, c3 = 1.0 / 48.0;
  wave_sync();
  const int i0 = lane * kPerLane;
  double loc[kPerLane];
  double run = 0.0;
#pragma unroll
  for (int j = 0; j < kPerLane; ++j) {
    const int i = i0 + j;
    run += (i < nr) ? fb[i] : 0.0;
    loc[j] = run;
  }
  // exclusive scan of the lane totals
  double tot = run;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const double t = __shfl_up(tot, o);
    if (lane >= o) tot += t;
  }
  const double before = tot - run;
  const double corr1 = c0 * fb[0] + c1 * fb[1] + c2 * fb[2] + c3 * fb[3];
#pragma unroll
  for (int j = 0; j < kPerLane; ++j) {
    const int i = i0 + j;
    if (i >= nr) break;
    double v = before + loc[j] + corr1;          // G(1) + sum_{k=2..i} f_k
    if (i >= 3) v += c0 * fb[i] + c1 * fb[i - 1] + c2 * fb[i - 2] + c3 * fb[i - 3];
    g[i] = v;
  }
  if (lane == 0) {
    const double f0 = fb[0], f1 = fb[1], f2 = fb[2], f3 = fb[3];
    const double a = (24.0 * f0 - 36.0 * f1 + 24.0 * f2 - 6.0 * f3) / 6.0;
    const double b = (-26.0 * f0 + 57.0 * f1 - 42.0 * f2 + 11.0 * f3) / 6.0;
    const double c = (9.0 * f0 - 24.0 * f1 + 21.0 * f2 - 6.0 * f3) / 6.0;
    const double d = (-f0 + 3.0 * f1 - 3.0 * f2 + f3) / 6.0;
    for (int i = 0; i < 3; ++i) {
      const double x = (double)(i + 1);
      g[i] = a * (x - 1.0) + 0.5 * b * (x * x - 1.0) + c * (x * x * x - 1.0) / 3.0 + 0.25 * d * (x * x * x * x - 1.0);
    }
  }
  wave_sync();
}

// RADIAL$POISSON (src/paw_radial.f90:1345-1386) of rho (LDS, nr points) into v;
// b1, b2 are wavefront scratch of kMaxNr doubles
__device__ void wave_poisson(int l, const double *rho, double *v, const double *r, const double *drdx, int nr,
                             double *b1, double *b2, int lane) {
  for (int i = lane; i < nr; i += 64) b1[i] = rho[i] * ipow(r[i], l + 2) * drdx[i];
  wave_integrate(b1, b2, nr, lane);
  const double fac = 4.0 * kPi / (double)(2 * l + 1);
  for (int i = lane; i < nr; i += 64) {
    v[i] = fac / ipow(r[i], l + 1) * b2[i];
    // rho r^(1-l)
    b1[i] = (l <= 1 ? rho[i] * ipow(r[i], 1 - l) : rho[i] / ipow(r[i], l - 1)) * drdx[i];
  }
  wave_integrate(b1, b2, nr, lane);
  const double top = b2[nr - 1];
  for (int i = lane; i < nr; i += 64) v[i] += fac * ipow(r[i], l) * (top - b2[i]);
}

// DFT3 of the local functional of type 1 without spin (src/paw_dft.f90:
// 1144-1332, EXCHANGE_X3 :2358-2430, PERDEWZUNGER_RHODEP3 :3019-3146):
// exc and its first three derivatives in rho; RHOTMIN continuation; rho <= 0: 0
__device__ void lda3(double rho, double *f) {
  f[0] = f[1] = f[2] = f[3] = 0.0;
  if (!(rho > 0.0)) return;
  const double v = fmax(rho, kRhoTMin);
  const double exfac = -3.0 / (4.0 * kPi) * cbrt(3.0 * kPi * kPi);
  const double r13 = cbrt(v), r23 = r13 * r13, r43 = r23 * r23;
  const double ex = exfac * r43, ex1 = exfac * 4.0 / 3.0 * r13, ex2 = exfac * 4.0 / 9.0 / r23;
  const double ex3 = -2.0 / 3.0 * ex2 / v;
  const double rs = cbrt(3.0 / (4.0 * kPi)) / r13;
  double ec, ec1, ec2, ec3;
  if (rs < 1.0) {
    const double ln = log(rs), lnr = rs * ln;
    ec = v * (kBU + kDU * rs + (kAU + kCU * rs) * ln);
    ec1 = (1.0 / 3.0) * (-kAU + 3.0 * kBU - (kCU - 2.0 * kDU) * rs + 3.0 * kAU * ln + 2.0 * kCU * lnr);
    ec2 = -(1.0 / 9.0) / v * (3.0 * kAU + (kCU + 2.0 * kDU) * rs + 2.0 * kCU * lnr);
    ec3 = (1.0 / 27.0) / (v * v) * (9.0 * kAU + (6.0 * kCU + 8.0 * kDU) * rs + 8.0 * kCU * lnr);
  } else {
    const double sq = sqrt(rs);
    const double sv = 1.0 / (1.0 + sq * (kBeta1U + kBeta2U * sq));
    const double sv1 = sv / (6.0 * v);
    const double b1 = kBeta1U, b2 = kBeta2U;
    ec = kGammaU * v * sv;
    ec1 = ec * sv1 * (6.0 + sq * (7.0 * b1 + 8.0 * b2 * sq));
    ec2 = ec * sv1 * sv1 * sq * (5.0 * b1 + sq * ((7.0 * b1 * b1 + 8.0 * b2) + sq * b2 * (21.0 * b1 + sq * b2 * 16.0)));
    ec3 = -ec * sv1 * sv1 * sv1 * sq *
          (35.0 * b1 +
           sq * ((76.0 * b1 * b1 + 64.0 * b2) +
                 sq * (b1 * (35.0 * b1 * b1 + 234.0 * b2) +
                       sq * b2 * ((140.0 * b1 * b1 + 176.0 * b2) + sq * b2 * (175.0 * b1 + sq * b2 * 64.0)))));
  }
  double f0 = ex + ec, f1 = ex1 + ec1, f2 = ex2 + ec2, f3 = ex3 + ec3;
  if (rho < kRhoTMin) {
    f3 = 0.0;
    f2 = 0.0;
    f1 = f0 / kRhoTMin;
    f0 = f0 * rho / kRhoTMin;
  }
  f[0] = f0;
  f[1] = f1;
  f[2] = f2;
  f[3] = f3;
}

// AUGMENTATION_XC (src/paw_augmentation.f90:699-977), NDIMD=1, local
// functional: energy density work1 (to be integrated with r^2) and the
// potential components vxc[lm] at one radial point of density x[lm]
__device__ inline double xc_point(const double *x, int lmrx, double *vxc) {
  for (int lm = 0; lm < lmrx; ++lm) vxc[lm] = 0.0;
  if (!(x[0] > 0.0)) return 0.0;
  double f[4];
  lda3(x[0] * kY0, f);
  double nons = 0.0;
  for (int lm = 1; lm < lmrx; ++lm) nons += x[lm] * x[lm];
  vxc[0] = 4.0 * kPi * f[1] * kY0 + 0.5 * kY0 * nons * f[3];
  for (int lm = 1; lm < lmrx; ++lm) vxc[lm] = 0.5 * f[2] * x[lm] + 0.5 * f[2] * x[lm];
  return 4.0 * kPi * f[0] + 0.5 * nons * f[2];
}

__device__ inline double *atom_work(const SphGeo &g, int a, int which) {
  return g.work + ((size_t)a * 4 + which) * (size_t)g.lmrxx * kMaxNr;
}

__device__ inline double dsym(const double *dm, int L, int i, int j) {
  return 0.5 * (dm[i * L + j] + dm[j * L + i]);
}

// ---------------------------------------------------------------------------
// one-centre densities and their moments
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kSphThreads) k_sph_rho(SphGeo g) {
  __shared__ double coef[kMaxLmrx * kMaxLnx * kMaxLnx];
  __shared__ double qpart[kMaxLmrx];
  const int a = blockIdx.x;
  const SphSpec sp = g.spec[g.isp[a]];
  const int L = g.lmnxx, nr = sp.nr, lnx = sp.lnx, lmrx = sp.lmrx;
  const double *dm = g.denmat + (size_t)a * L * L;
  const double *tab = g.tab;
  // C(lm3, ln1, ln2) = sum CG(LM1,LM2,LM3) D(LMN1,LMN2), entries in a fixed order
  const int ncoef = lmrx * lnx * lnx;
  for (int k = threadIdx.x; k < ncoef; k += kSphThreads) {
    const int b = g.off[sp.o_rho_off + k], e = g.off[sp.o_rho_off + k + 1];
    double s = 0.0;
    for (int t = b; t < e; ++t) {
      const SphEnt en = g.ent[sp.o_rho_ent + t];
      s += en.cg * dsym(dm, L, en.lmn1, en.lmn2);
    }
    coef[k] = s;
  }
  __syncthreads();
  double *rae = atom_work(g, a, 0), *rps = atom_work(g, a, 1);
  for (int k = threadIdx.x; k < lmrx * nr; k += kSphThreads) {
    const int lm = k / nr, i = k - lm * nr;
    double sa = 0.0, sp_ = 0.0;
    for (int l1 = 0; l1 < lnx; ++l1) {
      const double pa1 = tab[sp.o_aephi + l1 * nr + i], pp1 = tab[sp.o_psphi + l1 * nr + i];
      for (int l2 = 0; l2 < lnx; ++l2) {
        const double c = coef[(lm * lnx + l1) * lnx + l2];
        sa += c * pa1 * tab[sp.o_aephi + l2 * nr + i];
        sp_ += c * pp1 * tab[sp.o_psphi + l2 * nr + i];
      }
    }
    rae[lm * kMaxNr + i] = sa;
    rps[lm * kMaxNr + i] = sp_;
  }
  __syncthreads();
  // QLM(LM) = int (AERHO - PSRHO) r^(l+2) dr (+ core and nucleus for LM = 1)
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const double *r = tab + sp.o_r, *wint = tab + sp.o_wint;
  for (int lm = wv; lm < lmrx; lm += kSphWaves) {
    const int l = l_of(lm);
    double s = 0.0;
    for (int i = lane; i < nr; i += 64) s += wint[i] * (rae[lm * kMaxNr + i] - rps[lm * kMaxNr + i]) * ipow(r[i], l + 2);
    s = wave_sum(s);
    if (lane == 0) qpart[lm] = s;
  }
  __syncthreads();
  for (int lm = threadIdx.x; lm < g.lmrxx; lm += kSphThreads)
    g.qlm[(size_t)a * g.lmrxx + lm] = lm < lmrx ? qpart[lm] + (lm == 0 ? sp.core_q : 0.0) : 0.0;
}

// ---------------------------------------------------------------------------
// AUGMENTATION$SPHERE
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kSphThreads) k_sph_sphere(SphGeo g) {
  __shared__ double buf[kSphWaves][3][kMaxNr];
  __shared__ double vq1[kMaxLmrx], eps_[kMaxLmrx], eae[kMaxLmrx];
  __shared__ double cpot[kMaxNr];                        // core + nucleus potential (LM = 1)
  __shared__ double integ[kMaxLmrx * kMaxLnx * kMaxLnx];
  __shared__ double red[kSphWaves][4];
  __shared__ double scal[8];
  const int a = blockIdx.x;
  const SphSpec sp = g.spec[g.isp[a]];
  const int L = g.lmnxx, nr = sp.nr, lnx = sp.lnx, lmnx = sp.lmnx, lmrx = sp.lmrx;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const double *tab = g.tab;
  const double *r = tab + sp.o_r, *drdx = tab + sp.o_drdx, *wint = tab + sp.o_wint;
  const double *aecore = tab + sp.o_aecore, *pscore = tab + sp.o_pscore, *vadd = tab + sp.o_vadd;
  const double *dm = g.denmat + (size_t)a * L * L;
  const double *rae = atom_work(g, a, 0), *rps = atom_work(g, a, 1);
  double *pae = atom_work(g, a, 2), *pps = atom_work(g, a, 3);
  const double *q = g.qlm + (size_t)a * g.lmrxx;
  const double rhob = g.rhob ? *g.rhob : 0.0;
  const double alpha = 1.0 / (sp.rcsm * sp.rcsm);
  const double svar_b = -2.0 * kPi * rhob / 3.0 * sqrt(4.0 * kPi);
  double *b0 = buf[wv][0], *b1 = buf[wv][1], *b2 = buf[wv][2];

  // ADJUSTVQLM (src/paw_augmentation.f90:2054-2104) plus the plane-wave VQLM:
  // VQLM1(LM) = -int V[PSRHO (+ PS core) + QLM g_l] g_l r^2 + VQLM(LM)
  for (int lm = wv; lm < lmrx; lm += kSphWaves) {
    const int l = l_of(lm);
    const double cl = gaussn(l, alpha);
    for (int i = lane; i < nr; i += 64)
      b0[i] = rps[lm * kMaxNr + i] + (lm == 0 ? pscore[i] : 0.0) + q[lm] * cl * exp(-alpha * r[i] * r[i]) * ipow(r[i], l);
    wave_sync();
    // the potential goes to b0 after its last read as the density
    for (int i = lane; i < nr; i += 64) b1[i] = b0[i] * ipow(r[i], l + 2) * drdx[i];
    wave_sync();
    wave_integrate(b1, b2, nr, lane);
    wave_sync();
    const double fac = 4.0 * kPi / (double)(2 * l + 1);
    for (int i = lane; i < nr; i += 64) {
      b1[i] = (l <= 1 ? b0[i] * ipow(r[i], 1 - l) : b0[i] / ipow(r[i], l - 1)) * drdx[i];
      b0[i] = fac / ipow(r[i], l + 1) * b2[i];
    }
    wave_sync();
    wave_integrate(b1, b2, nr, lane);
    wave_sync();
    const double top = b2[nr - 1];
    double s = 0.0;
    for (int i = lane; i < nr; i += 64) {
      const double v = b0[i] + fac * ipow(r[i], l) * (top - b2[i]);
      s += wint[i] * v * cl * exp(-alpha * r[i] * r[i]) * ipow(r[i], l) * r[i] * r[i];
    }
    s = wave_sum(s);
    if (lane == 0) vq1[lm] = -s + (g.vqlm ? g.vqlm[(size_t)a * g.nvq + lm] : 0.0);
    wave_sync();
  }
  // potential of the core density and the nucleus (AUGMENTATION_AEHARTREE
  // :1983-1987, RADIAL$NUCPOT :1464-1498) by wavefront 0
  if (wv == 0) {
    wave_poisson(0, aecore, b0, r, drdx, nr, b1, b2, lane);
    for (int i = lane; i < nr; i += 64) {
      const double ri = r[i];
      const double nuc = ri > sp.rnuc ? -sp.aez / ri : -sp.aez / sp.rnuc * (1.5 - 0.5 * (ri / sp.rnuc) * (ri / sp.rnuc));
      cpot[i] = b0[i] + nuc / kY0;
    }
  }
  __syncthreads();
  // PS Hartree (AUGMENTATION_PSHARTREE :1827-1946) and AE Hartree
  // (AUGMENTATION_AEHARTREE :1949-2051), one LM per wavefront
  for (int lm = wv; lm < lmrx; lm += kSphWaves) {
    const int l = l_of(lm);
    const double cl = gaussn(l, alpha);
    // PS: rho1 = PSRHO + RHOHAT (+ PS core for l = 0)
    for (int i = lane; i < nr; i += 64)
      b0[i] = rps[lm * kMaxNr + i] + q[lm] * cl * exp(-alpha * r[i] * r[i]) * ipow(r[i], l) +
              (l == 0 ? pscore[i] : 0.0);
    wave_poisson(l, b0, pps + lm * kMaxNr, r, drdx, nr, b1, b2, lane);
    double s = 0.0;
    for (int i = lane; i < nr; i += 64) {
      const double ri = r[i], rho1 = b0[i];
      const double pext = vq1[lm] * ipow(ri, l);
      double pot = pps[lm * kMaxNr + i];
      double e = 0.5 * pot * rho1 + rho1 * pext;
      pot += pext;
      if (lm == 0) {
        const double pb = svar_b * ri * ri;
        e += vadd[i] * (rps[i] + pscore[i]) + rho1 * pb;   // rho1 of LM = 1 holds PSRHO + RHOHAT + core
        pot += vadd[i] + pb;
      }
      pps[lm * kMaxNr + i] = pot;
      s += wint[i] * e * ri * ri;
    }
    s = wave_sum(s);
    if (lane == 0) eps_[lm] = s;
    // AE
    wave_poisson(l, rae + lm * kMaxNr, pae + lm * kMaxNr, r, drdx, nr, b1, b2, lane);
    s = 0.0;
    for (int i = lane; i < nr; i += 64) {
      const double ri = r[i], rho = rae[lm * kMaxNr + i];
      const double pext = vq1[lm] * ipow(ri, l);
      double pot = pae[lm * kMaxNr + i];
      double e = 0.5 * pot * rho + rho * pext;
      pot += pext;
      if (l == 0) e += aecore[i] * pext;
      if (lm == 0) {
        const double pb = svar_b * ri * ri;
        e += cpot[i] * rho + (rho + aecore[i]) * pb;
        pot += cpot[i] + pb;
      }
      pae[lm * kMaxNr + i] = pot;
      s += wint[i] * e * ri * ri;
    }
    s = wave_sum(s);
    if (lane == 0) eae[lm] = s;
  }
  __syncthreads();
  // POTB of this atom (src/paw_augmentation.f90:463-470) from the Hartree
  // potentials, and the one-centre kinetic energy (:372-376)
  if (wv == 0) {
    double s = 0.0;
    for (int i = lane; i < nr; i += 64)
      s += wint[i] * (-(pae[i] - (pps[i] - vadd[i])) * r[i] * r[i] * sqrt(4.0 * kPi));
    s = wave_sum(s);
    if (lane == 0) scal[0] = s;
  } else if (wv == 1) {
    double s = 0.0;
    for (int k = lane; k < lmnx * lmnx; k += 64) {
      const int i = k / lmnx, j = k - i * lmnx;
      // DEKIN expanded with delta(L,L') delta(M,M') (AUGMENTATION_EXPANDDA)
      const int li = g.off[sp.o_exp_off + (lmnx * lmnx + 1) + i], lj = g.off[sp.o_exp_off + (lmnx * lmnx + 1) + j];
      const int lni = li & 0xffff, lnj = lj & 0xffff, mi = li >> 16, mj = lj >> 16;
      if (mi == mj) s += dsym(dm, L, i, j) * tab[sp.o_dekin + lni * lnx + lnj];
    }
    s = wave_sum(s);
    if (lane == 0) scal[1] = s;
  }
  __syncthreads();
  // XC, one radial point per thread: AE (valence + core), PS (valence + core),
  // and the core alone; potentials added to AEPOT / PSPOT
  double eae_x = 0.0, eps_x = 0.0, ecore = 0.0;
  for (int i = threadIdx.x; i < nr; i += kSphThreads) {
    // the first point sits at RMIN: its potential is taken from the second
    // (src/paw_augmentation.f90:957-959)
    const int iv = (i == 0 && r[0] < 1.0e-5) ? 1 : i;
    double x[kMaxLmrx], v[kMaxLmrx], vdum[kMaxLmrx];
    const double r2 = r[i] * r[i];
    for (int lm = 0; lm < lmrx; ++lm) x[lm] = rae[lm * kMaxNr + i] + (lm == 0 ? aecore[i] : 0.0);
    eae_x += wint[i] * xc_point(x, lmrx, vdum) * r2;
    for (int lm = 0; lm < lmrx; ++lm) x[lm] = rps[lm * kMaxNr + i] + (lm == 0 ? pscore[i] : 0.0);
    eps_x += wint[i] * xc_point(x, lmrx, vdum) * r2;
    x[0] = aecore[i];
    ecore += wint[i] * xc_point(x, 1, vdum) * r2;
    for (int lm = 0; lm < lmrx; ++lm) x[lm] = rae[lm * kMaxNr + iv] + (lm == 0 ? aecore[iv] : 0.0);
    xc_point(x, lmrx, v);
    for (int lm = 0; lm < lmrx; ++lm) pae[lm * kMaxNr + i] += v[lm];
    for (int lm = 0; lm < lmrx; ++lm) x[lm] = rps[lm * kMaxNr + iv] + (lm == 0 ? pscore[iv] : 0.0);
    xc_point(x, lmrx, v);
    for (int lm = 0; lm < lmrx; ++lm) pps[lm * kMaxNr + i] += v[lm];
  }
  eae_x = wave_sum(eae_x);
  eps_x = wave_sum(eps_x);
  ecore = wave_sum(ecore);
  if (lane == 0) {
    red[wv][0] = eae_x;
    red[wv][1] = eps_x;
    red[wv][2] = ecore;
  }
  __syncthreads();
  // EXPECT: I(lm3, ln1, ln2) = int (AEPOT phi_ae phi_ae - PSPOT phi_ps phi_ps) r^2
  const int nint = lmrx * lnx * lnx;
  for (int k = wv; k < nint; k += kSphWaves) {
    const int lm = k / (lnx * lnx), rem = k - lm * lnx * lnx, l1 = rem / lnx, l2 = rem - l1 * lnx;
    if (l2 < l1) continue;
    double s = 0.0;
    for (int i = lane; i < nr; i += 64)
      s += wint[i] * (pae[lm * kMaxNr + i] * tab[sp.o_aephi + l1 * nr + i] * tab[sp.o_aephi + l2 * nr + i] -
                      pps[lm * kMaxNr + i] * tab[sp.o_psphi + l1 * nr + i] * tab[sp.o_psphi + l2 * nr + i]) *
           r[i] * r[i];
    s = wave_sum(s);
    if (lane == 0) {
      integ[k] = s;
      integ[(lm * lnx + l2) * lnx + l1] = s;
    }
  }
  __syncthreads();
  // DATH = DEKIN + sum_LM3 CG(LM1,LM2,LM3) I(LM3,LN1,LN2)
  double2 *dh = g.dh + (size_t)a * L * L;
  for (int k = threadIdx.x; k < L * L; k += kSphThreads) {
    const int i = k / L, j = k - i * L;
    double v = 0.0;
    if (i < lmnx && j < lmnx) {
      const int p = i * lmnx + j;
      const int b = g.off[sp.o_exp_off + p], e = g.off[sp.o_exp_off + p + 1];
      for (int t = b; t < e; ++t) {
        const SphEnt en = g.ent[sp.o_exp_ent + t];
        v += en.cg * integ[(en.lm3 * lnx + en.ln1) * lnx + en.ln2];
      }
      const int li = g.off[sp.o_exp_off + (lmnx * lmnx + 1) + i], lj = g.off[sp.o_exp_off + (lmnx * lmnx + 1) + j];
      if ((li >> 16) == (lj >> 16)) v += tab[sp.o_dekin + (li & 0xffff) * lnx + (lj & 0xffff)];
    }
    dh[k] = make_double2(v, 0.0);
  }
  if (threadIdx.x == 0) {
    double aeeh = 0.0, pseh = 0.0;
    for (int lm = 0; lm < lmrx; ++lm) {
      aeeh += eae[lm];
      pseh += eps_[lm];
    }
    aeeh += -vq1[0] * kY0 * sp.aez;   // EEXT
    double aexc = 0.0, psexc = 0.0, corexc = 0.0;
    for (int w = 0; w < kSphWaves; ++w) {
      aexc += red[w][0];
      psexc += red[w][1];
      corexc += red[w][2];
    }
    aexc -= corexc;
    double *o = g.out + (size_t)a * 8;
    o[0] = scal[1] + aexc - psexc + aeeh - pseh;
    o[1] = scal[0];
    o[2] = scal[1];
    o[3] = aexc;
    o[4] = psexc;
    o[5] = aeeh;
    o[6] = pseh;
    o[7] = corexc;
  }
}

// POTB = sum_a POTB_a / Omega and E1C = sum_a E1C_a (atoms in order);
// res = {E1C, POTB}
__global__ void k_sph_finish(int nat, const double *out, double vol, double *res) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double e = 0.0, p = 0.0;
  for (int a = 0; a < nat; ++a) {
    e += out[(size_t)a * 8];
    p += out[(size_t)a * 8 + 1];
  }
  res[0] = e;
  res[1] = p / vol;
}

// WAVES$ADDCONSTANTPOT: DH += POTB DO (per atom L x L), V += POTB
__global__ void k_sph_addconst(int nat, int L, const double *res, const double2 *domat, double2 *dh, int nrl,
                               double *v) {
  const double potb = res[1];
  const size_t n = (size_t)nat * L * L;
  for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n + nrl; k += (size_t)gridDim.x * blockDim.x) {
    if (k < n) dh[k].x += potb * domat[k].x;
    else v[k - n] += potb;
  }
}

// ---------------------------------------------------------------------------
// host set-up
// ---------------------------------------------------------------------------
// CG(lm1, lm2, lm3) = int dOmega Y Y Y (SPHERICAL$GAUNT, src/paw_spherical.f90:
// 1052-1077) by Gauss-Legendre x uniform-phi quadrature, exact here
std::vector<double> gaunt_table(int lx) {
  const int lmx = (lx + 1) * (lx + 1);
  const int nq = 3 * lx + 2, nphi = 3 * lx + 2;
  // Gauss-Legendre nodes by Newton iteration on P_nq
  std::vector<double> ct(nq), wt(nq);
  for (int i = 0; i < nq; ++i) {
    double x = std::cos(kPi * (i + 0.75) / (nq + 0.5));
    for (int it = 0; it < 100; ++it) {
      double p0 = 1.0, p1 = x;
      for (int k = 2; k <= nq; ++k) {
        const double p2 = ((2 * k - 1) * x * p1 - (k - 1) * p0) / k;
        p0 = p1;
        p1 = p2;
      }
      const double dp = nq * (x * p1 - p0) / (x * x - 1.0);
      const double dx = p1 / dp;
      x -= dx;
      if (std::fabs(dx) < 1e-16) {
        wt[i] = 2.0 / ((1.0 - x * x) * dp * dp);
        break;
      }
      wt[i] = 2.0 / ((1.0 - x * x) * dp * dp);
    }
    ct[i] = x;
  }
  std::vector<double> cg((size_t)lmx * lmx * lmx, 0.0);
  std::vector<double> y(lmx);
  for (int i = 0; i < nq; ++i)
    for (int k = 0; k < nphi; ++k) {
      const double phi = 2.0 * kPi * k / nphi, s = std::sqrt(1.0 - ct[i] * ct[i]);
      ylm_dir(lx, s * std::cos(phi), s * std::sin(phi), ct[i], y.data());
      const double w = wt[i] * 2.0 * kPi / nphi;
      for (int a1 = 0; a1 < lmx; ++a1)
        for (int a2 = 0; a2 < lmx; ++a2) {
          const double ya = w * y[a1] * y[a2];
          for (int a3 = 0; a3 < lmx; ++a3) cg[((size_t)a1 * lmx + a2) * lmx + a3] += ya * y[a3];
        }
    }
  for (auto &c : cg)
    if (std::fabs(c) < 1.0e-14) c = 0.0;
  return cg;
}

}  // namespace

struct SphHost {
  bool set = false;
  int nr = 0, lmrx = 0;
  double r1 = 0, dex = 0, aez = 0, rnuc = 0, rcsm = 0;
  std::vector<double> aephi, psphi, aecore, pscore, vadd, dekin;   // LNX*NR, NR, LNX*LNX
};

struct SphState {
  std::vector<SphHost> spec;
  bool ready = false, uploaded = false;
  int lmrxx = 1, nrmax = 0;
  SphSpec *d_spec = nullptr;
  double *d_tab = nullptr;
  SphEnt *d_ent = nullptr;
  int *d_off = nullptr;
  double *d_work = nullptr, *d_qlm = nullptr, *d_out = nullptr, *d_res = nullptr;
  double *h_res = nullptr;   // pinned {E1C, POTB}
};

static SphState *sph_of(pw_waveset *ws) { return (SphState *)ws->sph; }

void sphere_release(pw_waveset *ws) {
  SphState *s = sph_of(ws);
  if (!s) return;
  void *ptrs[] = {s->d_spec, s->d_tab, s->d_ent, s->d_off, s->d_work, s->d_qlm, s->d_out, s->d_res};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  if (s->h_res) (void)hipHostFree(s->h_res);
  delete s;
  ws->sph = nullptr;
}

bool sphere_ready(const pw_waveset *ws) { return ws->sph && ((const SphState *)ws->sph)->ready; }
int sphere_lmrxx(const pw_waveset *ws) {
  if (!ws->sph) return 1;
  int m = 1;
  for (const SphHost &h : ((const SphState *)ws->sph)->spec) m = std::max(m, h.lmrx);
  return m;
}
int sphere_lmrx(const pw_waveset *ws, int isp) {
  return ws->sph ? ((const SphState *)ws->sph)->spec[isp].lmrx : 1;
}

// packs the tables of every species and uploads them (first ETOT after set-up)
static int sphere_upload(pw_waveset *ws) {
  SphState *s = sph_of(ws);
  std::vector<double> tab;
  std::vector<SphEnt> ent;
  std::vector<int> off;
  std::vector<SphSpec> spec(ws->nsp);
  int lxmax = 0;
  for (int k = 0; k < ws->nsp; ++k) {
    for (int ln = 0; ln < ws->lnx[k]; ++ln) lxmax = std::max(lxmax, ws->lox[ln + ws->lnxx * k]);
    const int lr = (int)std::lround(std::sqrt((double)s->spec[k].lmrx)) - 1;
    lxmax = std::max(lxmax, lr);
  }
  const std::vector<double> cg = gaunt_table(lxmax);
  const int lmxg = (lxmax + 1) * (lxmax + 1);
  s->lmrxx = 1;
  s->nrmax = 0;
  for (int k = 0; k < ws->nsp; ++k) {
    const SphHost &h = s->spec[k];
    SphSpec &o = spec[k];
    const int nr = h.nr, lnx = ws->lnx[k];
    o.nr = nr;
    o.lnx = lnx;
    o.lmnx = ws->lmnx[k];
    o.lmrx = h.lmrx;
    o.aez = h.aez;
    o.rnuc = h.rnuc;
    o.rcsm = h.rcsm;
    o.r1 = h.r1;
    s->lmrxx = std::max(s->lmrxx, h.lmrx);
    s->nrmax = std::max(s->nrmax, nr);
    // the shifted logarithmic grid as SHLOGRADIAL_RESOLVE builds it
    // (src/paw_radial.f90:1640-1652): r_i = r1 x^(i-1) - r1 by repeated
    // products, r_1 = RMIN; dr/dx = dex r1 x^(i-1) (:1884-1888)
    std::vector<double> r(nr), drdx(nr), wint(nr);
    const double xexp = std::exp(h.dex);
    double ri = h.r1 / xexp;
    for (int i = 0; i < nr; ++i) {
      ri *= xexp;
      r[i] = ri - h.r1;
      drdx[i] = ri * h.dex;
    }
    r[0] = 1.0e-9;
    // G(NR) of RADIAL_INTEGRATEEQUISPACED as weights (src/paw_radial.f90:418-429)
    const double cc[4] = {-31.0 / 48.0, 11.0 / 48.0, -5.0 / 48.0, 1.0 / 48.0};
    for (int i = 0; i < nr; ++i) wint[i] = 1.0;
    for (int j = 0; j < 4; ++j) {
      wint[j] += cc[j];
      wint[nr - 1 - j] += cc[j];
    }
    for (int i = 0; i < nr; ++i) wint[i] *= drdx[i];
    auto put = [&](const double *v, size_t n) {
      const int at = (int)tab.size();
      tab.insert(tab.end(), v, v + n);
      return at;
    };
    o.o_r = put(r.data(), nr);
    o.o_drdx = put(drdx.data(), nr);
    o.o_wint = put(wint.data(), nr);
    o.o_aephi = put(h.aephi.data(), (size_t)lnx * nr);
    o.o_psphi = put(h.psphi.data(), (size_t)lnx * nr);
    o.o_aecore = put(h.aecore.data(), nr);
    o.o_pscore = put(h.pscore.data(), nr);
    o.o_vadd = put(h.vadd.data(), nr);
    o.o_dekin = put(h.dekin.data(), (size_t)lnx * lnx);
    // core moment: int (AECORE - PSCORE) r^2 - AEZ Y0
    double cq = 0.0;
    for (int i = 0; i < nr; ++i) cq += wint[i] * (h.aecore[i] - h.pscore[i]) * r[i] * r[i];
    o.core_q = cq - h.aez * kY0;
    // Gaunt entries of this species
    struct E { int lmn1, lmn2, ln1, ln2, lm3; double cg; };
    std::vector<E> list;
    std::vector<int> lmn_ln, lmn_lm;
    for (int ln = 0; ln < lnx; ++ln) {
      const int l = ws->lox[ln + ws->lnxx * k];
      for (int m = 0; m < 2 * l + 1; ++m) {
        lmn_ln.push_back(ln);
        lmn_lm.push_back(l * l + m);
      }
    }
    const int lmnx = (int)lmn_ln.size();
    for (int i = 0; i < lmnx; ++i)
      for (int j = 0; j < lmnx; ++j)
        for (int lm3 = 0; lm3 < h.lmrx; ++lm3) {
          const double c = cg[((size_t)lmn_lm[i] * lmxg + lmn_lm[j]) * lmxg + lm3];
          if (c != 0.0) list.push_back({i, j, lmn_ln[i], lmn_ln[j], lm3, c});
        }
    // order 1: by (lm3, ln1, ln2), then (lmn1, lmn2) -- offsets per (lm3, ln1, ln2)
    std::vector<E> byc = list;
    std::stable_sort(byc.begin(), byc.end(), [&](const E &x, const E &y) {
      const int kx = (x.lm3 * lnx + x.ln1) * lnx + x.ln2, ky = (y.lm3 * lnx + y.ln1) * lnx + y.ln2;
      return kx < ky;
    });
    o.o_rho_ent = (int)ent.size();
    o.o_rho_off = (int)off.size();
    {
      const int nc = h.lmrx * lnx * lnx;
      std::vector<int> cnt(nc + 1, 0);
      for (const E &e : byc) cnt[(e.lm3 * lnx + e.ln1) * lnx + e.ln2 + 1]++;
      for (int c = 0; c < nc; ++c) cnt[c + 1] += cnt[c];
      off.insert(off.end(), cnt.begin(), cnt.end());
      for (const E &e : byc)
        ent.push_back({(short)e.lmn1, (short)e.lmn2, (short)e.ln1, (short)e.ln2, (short)e.lm3, 0, e.cg});
    }
    // order 2: by (lmn1, lmn2), then lm3 -- offsets per pair, then per lmn
    // (ln | m-slot << 16) for the delta(L,L') delta(M,M') expansions
    o.o_exp_ent = (int)ent.size();
    o.o_exp_off = (int)off.size();
    {
      const int np = lmnx * lmnx;
      std::vector<int> cnt(np + 1, 0);
      for (const E &e : list) cnt[e.lmn1 * lmnx + e.lmn2 + 1]++;
      for (int c = 0; c < np; ++c) cnt[c + 1] += cnt[c];
      off.insert(off.end(), cnt.begin(), cnt.end());
      for (int i = 0; i < lmnx; ++i) {
        // equal (l, m) slot <=> same harmonic: compare lm and the l of ln
        off.push_back(lmn_ln[i] | (lmn_lm[i] << 16));
      }
      for (const E &e : list)
        ent.push_back({(short)e.lmn1, (short)e.lmn2, (short)e.ln1, (short)e.ln2, (short)e.lm3, 0, e.cg});
    }
    if (nr > kMaxNr || nr < 10) return pw_fail(PW_ERR_UNSUPPORTED, "WAVES$SPHERE: NR must be 10..512");
    if (lnx > kMaxLnx || lmnx > kMaxLmnx) return pw_fail(PW_ERR_UNSUPPORTED, "WAVES$SPHERE: LNX > 8 or LMNX > 32");
  }
  if (s->lmrxx > kMaxLmrx) return pw_fail(PW_ERR_UNSUPPORTED, "WAVES$SPHERE: LMRX > 25");
  if (ws->lmnxx > kMaxLmnx) return pw_fail(PW_ERR_UNSUPPORTED, "WAVES$SPHERE: LMNXX > 32");
  auto up = [&](void **dst, const void *src, size_t bytes) -> int {
    if (*dst) PW_HIP(hipFree(*dst));
    *dst = nullptr;
    PW_HIP(hipMalloc(dst, std::max<size_t>(bytes, 8)));
    if (bytes && src) PW_HIP(hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice));
    return PW_OK;
  };
  PW_TRY(up((void **)&s->d_spec, spec.data(), spec.size() * sizeof(SphSpec)));
  PW_TRY(up((void **)&s->d_tab, tab.data(), tab.size() * sizeof(double)));
  PW_TRY(up((void **)&s->d_ent, ent.data(), ent.size() * sizeof(SphEnt)));
  PW_TRY(up((void **)&s->d_off, off.data(), off.size() * sizeof(int)));
  const size_t nw = (size_t)ws->nat * 4 * s->lmrxx * kMaxNr;
  PW_TRY(up((void **)&s->d_work, nullptr, nw * sizeof(double)));
  PW_TRY(up((void **)&s->d_qlm, nullptr, (size_t)ws->nat * s->lmrxx * sizeof(double)));
  PW_TRY(up((void **)&s->d_out, nullptr, (size_t)ws->nat * 8 * sizeof(double)));
  PW_TRY(up((void **)&s->d_res, nullptr, 8 * sizeof(double)));
  PW_HIP(hipMemset(s->d_work, 0, nw * sizeof(double)));
  if (!s->h_res) PW_HIP(hipHostMalloc((void **)&s->h_res, 8 * sizeof(double), hipHostMallocDefault));
  s->uploaded = true;
  return PW_OK;
}

static SphGeo sphere_geo(pw_waveset *ws) {
  SphState *s = sph_of(ws);
  SphGeo g{};
  g.nat = ws->nat;
  g.lmnxx = ws->lmnxx;
  g.lmrxx = s->lmrxx;
  g.nvq = s->lmrxx + 3;
  g.isp = ws->d_atom_isp;
  g.spec = s->d_spec;
  g.tab = s->d_tab;
  g.ent = s->d_ent;
  g.off = s->d_off;
  g.denmat = ws->d_res + ws->r_denmat;
  g.work = s->d_work;
  g.qlm = s->d_qlm;
  g.dh = ws->d_dhmat;
  g.out = s->d_out;
  return g;
}

// WAVES$MOMENTS on the device: QLM(LMRXX,NAT) of the current DENMAT (the
// partial sums that k_denmat left in the step's result block)
int sphere_moments(pw_waveset *ws, const double **qlm_dev) {
  SphState *s = sph_of(ws);
  if (!s->uploaded) PW_TRY(sphere_upload(ws));
  SphGeo g = sphere_geo(ws);
  ScopeTimer tm(ws, "W:MOMENTS");
  k_sph_rho<<<ws->nat, kSphThreads, 0, ws->stream>>>(g);
  PW_HIP(hipGetLastError());
  *qlm_dev = s->d_qlm;
  return PW_OK;
}

// WAVES$SPHERE and WAVES$ADDCONSTANTPOT on the device after POTENTIAL$VOFRHO:
// DH (ws->d_dhmat), POTB, E1C; V += POTB, DH += POTB DO.  The scalars reach
// the pinned buffer when the stream passes (sphere_results).
int sphere_terms(pw_waveset *ws, const double *vqlm_dev, const double *rhob_dev) {
  SphState *s = sph_of(ws);
  SphGeo g = sphere_geo(ws);
  g.vqlm = vqlm_dev;
  g.rhob = rhob_dev;
  {
    ScopeTimer tm(ws, "W:SPHERE");
    k_sph_sphere<<<ws->nat, kSphThreads, 0, ws->stream>>>(g);
    k_sph_finish<<<1, 64, 0, ws->stream>>>(ws->nat, s->d_out, ws->plan->vol, s->d_res);
    PW_HIP(hipGetLastError());
  }
  {
    ScopeTimer tm(ws, "W:ADDCONSTANTPOT");
    const size_t n = (size_t)ws->nat * ws->lmnxx * ws->lmnxx + ws->plan->nrl;
    const int blocks = (int)std::min<size_t>(4096, (n + 255) / 256);
    k_sph_addconst<<<blocks, 256, 0, ws->stream>>>(ws->nat, ws->lmnxx, s->d_res, ws->d_domat, ws->d_dhmat,
                                                    ws->plan->nrl, ws->d_v);
    PW_HIP(hipGetLastError());
  }
  PW_HIP(hipMemcpyAsync(s->h_res, s->d_res, 2 * sizeof(double), hipMemcpyDeviceToHost, ws->stream));
  return PW_OK;
}

void sphere_results(const pw_waveset *ws, double *e1c, double *potb) {
  const SphState *s = (const SphState *)ws->sph;
  *e1c = s->h_res[0];
  *potb = s->h_res[1];
}

const double *sphere_qlm_dev(const pw_waveset *ws) { return ws->sph ? ((const SphState *)ws->sph)->d_qlm : nullptr; }

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" int pw_waves_sphere_setup(pw_waveset *ws, int isp, int nr, double r1, double dex, double aez, double rnuc,
                                     double rcsm, int lmrx, const double *aephi, const double *psphi,
                                     const double *aecore, const double *pscore, const double *vadd,
                                     const double *dekin) {
  if (!ws || !aephi || !psphi || !aecore || !pscore || !vadd || !dekin)
    return pw_fail(PW_ERR_ARG, "pw_waves_sphere_setup: null argument");
  PW_TRY(ws_finish(ws));
  if (isp < 1 || isp > ws->nsp) return pw_fail(PW_ERR_ARG, "pw_waves_sphere_setup: species index out of range");
  if (nr < 10 || nr > kMaxNr) return pw_fail(PW_ERR_UNSUPPORTED, "pw_waves_sphere_setup: NR must be 10..512");
  if (!(r1 > 0.0) || !(dex > 0.0) || !(rcsm > 0.0) || !(rnuc > 0.0))
    return pw_fail(PW_ERR_ARG, "pw_waves_sphere_setup: R1, DEX, RCSM and RNUC must be positive");
  const int lr = (int)std::lround(std::sqrt((double)lmrx)) - 1;
  if (lmrx < 1 || lmrx > kMaxLmrx || (lr + 1) * (lr + 1) != lmrx)
    return pw_fail(PW_ERR_ARG, "pw_waves_sphere_setup: LMRX must be (L+1)^2 with L <= 4");
  const int k = isp - 1, lnx = ws->lnx[k];
  if (lnx > kMaxLnx || ws->lmnx[k] > kMaxLmnx)
    return pw_fail(PW_ERR_UNSUPPORTED, "pw_waves_sphere_setup: LNX > 8 or LMNX > 32");
  PW_HIP(hipStreamSynchronize(ws->stream));
  if (!ws->sph) {
    SphState *s = new SphState();
    s->spec.resize(ws->nsp);
    ws->sph = s;
  }
  SphState *s = sph_of(ws);
  SphHost &h = s->spec[k];
  h.set = true;
  h.nr = nr;
  h.lmrx = lmrx;
  h.r1 = r1;
  h.dex = dex;
  h.aez = aez;
  h.rnuc = rnuc;
  h.rcsm = rcsm;
  h.aephi.assign(aephi, aephi + (size_t)lnx * nr);
  h.psphi.assign(psphi, psphi + (size_t)lnx * nr);
  h.aecore.assign(aecore, aecore + nr);
  h.pscore.assign(pscore, pscore + nr);
  h.vadd.assign(vadd, vadd + nr);
  h.dekin.resize((size_t)lnx * lnx);
  // DEKIN(LNX,LNX) column-major -> [ln1][ln2]
  for (int i = 0; i < lnx; ++i)
    for (int j = 0; j < lnx; ++j) h.dekin[(size_t)i * lnx + j] = dekin[(size_t)j * lnx + i];
  s->uploaded = false;
  s->ready = std::all_of(s->spec.begin(), s->spec.end(), [](const SphHost &x) { return x.set; });
  return PW_OK;
}
