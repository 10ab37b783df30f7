// WAVES kernels for gfx950 (FP64): projector products, overlaps, the
// constraint update of the wave functions and the reductions of the step.
//
// Projector functions are never stored: each kernel rebuilds
//   P(atom,lmn,G) = YLM(G,lm) PRO(G,ln) (-i)^l exp(-iG.R)
// (WAVES_EXPANDPRO, src/paw_waves1.f90:6070-6109) from per-axis phase tables,
// so the NPRO x NGL projector matrix costs no HBM traffic.
#include "pw_common.h"

namespace {

constexpr int kB = 256;

struct ProjGeo {
  int ngl, nbh, nat, lmnxx, nphase, nmaxx;
  const int *igvec;        // 3*ngl
  const double *pro;       // [nbarepro][ngl]
  const double *ylm;       // [lmx][ngl]
  const double2 *phase;    // [nat][3][nphase]
  const int *atom_isp, *ipro0, *ibpro, *lmnx;
  const int *lmn_ln, *lmn_lm, *lmn_l;   // [nsp][lmnxx]
};

__device__ inline double2 czero() { return make_double2(0.0, 0.0); }

// XCD-aware tile order.  Workgroups are dealt to the eight XCDs in turn
// (linear id mod 8), and each XCD has its own L2.  The tiles that read the
// same operand block (one G range) get linear ids of one residue mod 8 that
// follow each other in that XCD's order, so the block is fetched from HBM once
// instead of once per tile.  Grid: xcd_blocks(ninner, nshared) in x.
constexpr int kXcd = 8;
struct XcdTile {
  int inner, shared;
  bool valid;
};
__device__ inline XcdTile xcd_tile(int ninner, int nshared) {
  const int L = blockIdx.x, x = L % kXcd, slot = L / kXcd;
  XcdTile t;
  t.inner = slot % ninner;
  t.shared = (slot / ninner) * kXcd + x;
  t.valid = t.shared < nshared;
  return t;
}
inline unsigned xcd_blocks(int ninner, int nshared) {
  return (unsigned)((nshared + kXcd - 1) / kXcd) * kXcd * ninner;
}

__device__ inline double2 eigr_of(const ProjGeo &q, int atom, int ig) {
  const int *iv = q.igvec + 3 * ig;
  const double2 *ph = q.phase + (size_t)atom * 3 * q.nphase;
  double2 e = cmul(ph[iv[0] + q.nmaxx], ph[q.nphase + iv[1] + q.nmaxx]);
  return cmul(e, ph[2 * q.nphase + iv[2] + q.nmaxx]);
}

// sum of v over the 16 lanes of a DPP row (lanes 16k .. 16k+15), in every lane
// of the row: four rotate-and-add steps on the cross-lane (DPP) path of the
// VALU (row_ror 8, 4, 2, 1), no LDS traffic
template <int CTRL>
__device__ inline double dpp_ror(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}
__device__ inline double row16_sum(double v) {
  v += dpp_ror<0x128>(v);
  v += dpp_ror<0x124>(v);
  v += dpp_ror<0x122>(v);
  v += dpp_ror<0x121>(v);
  return v;
}

// z * (-i)^l
__device__ inline double2 rot_mil(double2 z, int l) {
  switch (l & 3) {
    case 0: return z;
    case 1: return make_double2(z.y, -z.x);
    case 2: return make_double2(-z.x, -z.y);
    default: return make_double2(-z.y, z.x);
  }
}

// P(atom, lmn, G) given eigr; returns zero for an empty slot
__device__ inline double2 pro_val(const ProjGeo &q, int atom, int lmn, int ig, double2 eigr) {
  const int isp = q.atom_isp[atom];
  if (lmn >= q.lmnx[isp]) return czero();
  const int k = isp * q.lmnxx + lmn;
  const int ln = q.lmn_ln[k], lm = q.lmn_lm[k], l = q.lmn_l[k];
  const double s = q.ylm[(size_t)lm * q.ngl + ig] * q.pro[(size_t)(q.ibpro[isp] + ln) * q.ngl + ig];
  return rot_mil(make_double2(s * eigr.x, s * eigr.y), l);
}

// ---------------------------------------------------------------------------
// phase tables exp(-i n b_k.R), n in [-nmax, nmax]
// ---------------------------------------------------------------------------
// GBAS(:,k).R of up to kPhaseAtoms atoms, passed by value with the launch
constexpr int kPhaseAtoms = 128;
struct PhaseArgs {
  double gr[3 * kPhaseAtoms];
};

__global__ void k_phase(int nat, int nphase, int nmaxx, PhaseArgs a, double2 *phase) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int total = nat * 3 * nphase;
  if (i >= total) return;
  const int n = i % nphase - nmaxx;
  const int ak = i / nphase;                 // atom*3 + axis
  double s, c;
  sincos((double)n * a.gr[ak], &s, &c);
  phase[i] = make_double2(c, -s);
}

// ---------------------------------------------------------------------------
// FP64 MFMA building blocks.  v_mfma_f64_16x16x4f64 lane layout: A(i,k) in
// lane i+16k, B(k,j) in lane j+16k, C(i,j) in lane j+16(i%4), element i/4.
// A complex product is four real MFMAs on split re/im operands.  LDS rows are
// padded to 16 (mod 32) doubles so that the 16 x 4 operand reads of a wave
// are free of bank conflicts.
// ---------------------------------------------------------------------------
typedef double dbl4 __attribute__((ext_vector_type(4)));

constexpr int KC = 16;   // K depth staged in LDS per step (four MFMA K slices)

constexpr int padw(int w) { return w + ((16 - w % 32) + 32) % 32; }

__device__ inline dbl4 mfma4(double a, double b, dbl4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// (cr + i ci) += (ar + i ai)(br + i bi)
__device__ inline void cmfma(dbl4 &cr, dbl4 &ci, double ar, double ai, double br, double bi) {
  cr = mfma4(ar, br, cr);
  cr = mfma4(-ai, bi, cr);
  ci = mfma4(ar, bi, ci);
  ci = mfma4(ai, br, ci);
}

// conj(P(ipro, G)) for the projections
struct GenProj {
  ProjGeo q;
  const int2 *iptab;
  __device__ double2 operator()(int r, int ig) const {
    const int2 al = iptab[r];
    return cconj(pro_val(q, al.x, al.y, ig, eigr_of(q, al.x, ig)));
  }
};

// stacked rows [conj(psi1_i(G)) ; psi1_i(-G)] for the band overlaps; rows
// [0,nbhp) give T, rows [nbhp, 2 nbhp) give TM (nbhp = nbh rounded up to 16)
struct GenOvl {
  const double2 *psi1;
  const int *minusg;
  int ngl, nbh, nbhp;
  __device__ double2 operator()(int r, int ig) const {
    if (r < nbhp) return r < nbh ? cconj(psi1[(size_t)r * ngl + ig]) : czero();
    r -= nbhp;
    return r < nbh ? psi1[(size_t)r * ngl + minusg[ig]] : czero();
  }
};

// part[split][r][b] = sum_{G in split} A(r,G) psi[b][G], A from a generator.
// Workgroup: NW waves, one 16-row block each, all NCB 16-band blocks; the
// psi chunk is shared through LDS.
template <class Gen, int NCB, int NW>
__global__ void __launch_bounds__(64 * NW) k_rowgemm(Gen gen, int nrow, int ngl, int nbh,
                                                     const double2 *__restrict__ psi, int gsplit, int tr,
                                                     int tb, int ns, double2 *__restrict__ part) {
  constexpr int WR = 16 * NW, SA = padw(WR), NB16 = 16 * NCB, SB = padw(NB16);
  __shared__ double sAr[KC][SA], sAi[KC][SA], sBr[KC][SB], sBi[KC][SB];
  const XcdTile xt = xcd_tile(tr * tb, ns);
  if (!xt.valid) return;
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6, li = lane & 15, lk = lane >> 4;
  const int r0 = (xt.inner % tr) * WR, b0 = (xt.inner / tr) * NB16;
  const int g0 = xt.shared * gsplit, g1 = min(ngl, g0 + gsplit);
  dbl4 accr[NCB], acci[NCB];
#pragma unroll
  for (int c = 0; c < NCB; ++c) {
    accr[c] = dbl4{0.0, 0.0, 0.0, 0.0};
    acci[c] = dbl4{0.0, 0.0, 0.0, 0.0};
  }
  for (int gc = g0; gc < g1; gc += KC) {
    for (int e = t; e < WR * KC; e += 64 * NW) {
      const int k = e % KC, r = e / KC, ig = gc + k;
      double2 v = czero();
      if (r0 + r < nrow && ig < g1) v = gen(r0 + r, ig);
      sAr[k][r] = v.x;
      sAi[k][r] = v.y;
    }
    for (int e = t; e < NB16 * KC; e += 64 * NW) {
      const int k = e % KC, b = e / KC, ig = gc + k;
      double2 v = czero();
      if (b0 + b < nbh && ig < g1) v = psi[(size_t)(b0 + b) * ngl + ig];
      sBr[k][b] = v.x;
      sBi[k][b] = v.y;
    }
    __syncthreads();
#pragma unroll
    for (int ks = 0; ks < KC / 4; ++ks) {
      const int k = 4 * ks + lk;
      const double ar = sAr[k][16 * wv + li], ai = sAi[k][16 * wv + li];
#pragma unroll
      for (int c = 0; c < NCB; ++c) cmfma(accr[c], acci[c], ar, ai, sBr[k][16 * c + li], sBi[k][16 * c + li]);
    }
    __syncthreads();
  }
  double2 *dst = part + (size_t)xt.shared * nrow * nbh;
#pragma unroll
  for (int c = 0; c < NCB; ++c) {
    const int j = b0 + 16 * c + li;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int i = r0 + 16 * wv + lk + 4 * r;
      if (i < nrow && j < nbh) dst[(size_t)i * nbh + j] = make_double2(accr[c][r], acci[c][r]);
    }
  }
}

// sum of split partials: out[i] = scale * sum_s part[s][i] in a fixed order.
// A block takes 32 outputs; its 8 lanes per output stride over the splits
// with four accumulators each and are combined through LDS.
constexpr int kRedC = 32, kRedS = 8;

__device__ inline double rz(double) { return 0.0; }
__device__ inline double2 rz(double2) { return make_double2(0.0, 0.0); }
__device__ inline double radd(double a, double b) { return a + b; }
__device__ inline double2 radd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ inline double rscale(double a, double s) { return a * s; }
__device__ inline double2 rscale(double2 a, double s) { return make_double2(a.x * s, a.y * s); }

// out[i] = scale * sum_k part[k][i]: kRedS lanes per column take the splits
// k = sl, sl + kRedS, ...; each lane issues its loads 16 at a time and adds
// them in k order, then the lanes are added in order (deterministic)
template <typename T>
__global__ void __launch_bounds__(kRedC * kRedS) k_reduce_split(const T *__restrict__ part, int ns, int n,
                                                                double scale, T *__restrict__ out) {
  __shared__ T red[kRedS][kRedC];
  const int c = threadIdx.x % kRedC, sl = threadIdx.x / kRedC;
  const int i = blockIdx.x * kRedC + c;
  T acc = rz(T());
  if (i < n) {
    constexpr int V = 16;
    for (int k0 = sl; k0 < ns; k0 += V * kRedS) {
      T v[V];
#pragma unroll
      for (int u = 0; u < V; ++u) {
        const int k = k0 + u * kRedS;
        v[u] = k < ns ? part[(size_t)k * n + i] : rz(T());
      }
#pragma unroll
      for (int u = 0; u < V; ++u) acc = radd(acc, v[u]);
    }
  }
  red[sl][c] = acc;
  __syncthreads();
  if (sl != 0 || i >= n) return;
  T r = red[0][c];
#pragma unroll
  for (int q = 1; q < kRedS; ++q) r = radd(r, red[q][c]);
  out[i] = rscale(r, scale);
}


// ---------------------------------------------------------------------------
// psi[b][G] += sum_k A(b,k) B(k,G): the band-major output is C^T, so a lane's
// MFMA results are 16 consecutive G of one band (coalesced read-modify-write).
// A^T(k,b) and B(k,G) come from generators; NW waves of 16 G each, NRB
// 16-band blocks per wave.
// ---------------------------------------------------------------------------
// ADDPRO (src/paw_waves1.f90:5930-5980 WAVES_ADDPRO): A^T(ipro,b) = X[ipro][b], B = P
struct AddproA {
  const double2 *x;
  int nbh;
  __device__ double2 operator()(int k, int b) const { return x[(size_t)k * nbh + b]; }
};
struct AddproB {
  ProjGeo q;
  const int2 *iptab;
  __device__ double2 operator()(int k, int ig) const {
    const int2 al = iptab[k];
    return pro_val(q, al.x, al.y, ig, eigr_of(q, al.x, ig));
  }
};

// Destination of the column products: psi[b][G] = (src or psi)[b][G] + C,
// then times num / m(G) when mpsi is set (OPSI: the copy of PSI0, ADDPRO and
// the mass scaling of WAVES_OPSI in one pass, the same operations in order)
struct ColOut {
  double2 *psi;
  const double2 *src;
  const double *mpsi;
  double num;
  __device__ void store(size_t i, int ig, double cr, double ci) const {
    const double2 s = src ? src[i] : psi[i];
    double2 v = make_double2(s.x + cr, s.y + ci);
    if (mpsi) {
      const double f = num / mpsi[ig];
      v = make_double2(f * v.x, f * v.y);
    }
    psi[i] = v;
  }
};

template <class GA, class GB, int NRB, int NW>
__global__ void __launch_bounds__(64 * NW) k_colgemm(GA ga, GB gb, int nk, int ngl, int nbh, ColOut out) {
  constexpr int WG = 16 * NW, SP = padw(WG), NB16 = 16 * NRB, SX = padw(NB16);
  __shared__ double sBr[KC][SP], sBi[KC][SP], sAr[KC][SX], sAi[KC][SX];
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6, li = lane & 15, lk = lane >> 4;
  const int g0 = blockIdx.x * WG, b0 = blockIdx.y * NB16;
  dbl4 accr[NRB], acci[NRB];
#pragma unroll
  for (int c = 0; c < NRB; ++c) {
    accr[c] = dbl4{0.0, 0.0, 0.0, 0.0};
    acci[c] = dbl4{0.0, 0.0, 0.0, 0.0};
  }
  for (int k0 = 0; k0 < nk; k0 += KC) {
    for (int e = t; e < WG * KC; e += 64 * NW) {
      const int g = e % WG, k = e / WG, ig = g0 + g;
      double2 v = czero();
      if (k0 + k < nk && ig < ngl) v = gb(k0 + k, ig);
      sBr[k][g] = v.x;
      sBi[k][g] = v.y;
    }
    for (int e = t; e < NB16 * KC; e += 64 * NW) {
      const int b = e % NB16, k = e / NB16;
      double2 v = czero();
      if (k0 + k < nk && b0 + b < nbh) v = ga(k0 + k, b0 + b);
      sAr[k][b] = v.x;
      sAi[k][b] = v.y;
    }
    __syncthreads();
#pragma unroll
    for (int ks = 0; ks < KC / 4; ++ks) {
      const int k = 4 * ks + lk;
      const double br = sBr[k][16 * wv + li], bi = sBi[k][16 * wv + li];
#pragma unroll
      for (int c = 0; c < NRB; ++c) cmfma(accr[c], acci[c], sAr[k][16 * c + li], sAi[k][16 * c + li], br, bi);
    }
    __syncthreads();
  }
  const int ig = g0 + 16 * wv + li;
  if (ig >= ngl) return;
#pragma unroll
  for (int c = 0; c < NRB; ++c) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int b = b0 + 16 * c + lk + 4 * r;
      if (b < nbh) out.store((size_t)b * ngl + ig, ig, accr[c][r], acci[c][r]);
    }
  }
}

// ---------------------------------------------------------------------------
// forces: F(atom) = 2 Omega sum_G G Re[i P conj(D)], D = psi conj(dE/dproj)
// (WAVES_DEDPRO + WAVES_PROFORCE fused; the INVERTG symmetrisation of DEDPRO
//  leaves this sum unchanged because P(-G) = conj(P(G))).
// D(ipro,G) = sum_b conj(ded[ipro][b]) psi[b][G] on MFMA (NW waves x 16 ipro
// rows, NGB 16-G blocks); the epilogue contracts D with P and G and sums the
// workgroup's G range: part[gtile][ipro][3].
// ---------------------------------------------------------------------------
template <int NGB, int NW>
__global__ void __launch_bounds__(64 * NW) k_force_m(ProjGeo q, const int2 *iptab, int npro,
                                                     const double2 *__restrict__ psi,
                                                     const double2 *__restrict__ ded,
                                                     const double *__restrict__ gvec, int nrt, int ngt,
                                                     double *part) {
  constexpr int WR = 16 * NW, SA = padw(WR), NG16 = 16 * NGB, SB = padw(NG16);
  __shared__ double sAr[KC][SA], sAi[KC][SA], sBr[KC][SB], sBi[KC][SB];
  const XcdTile xt = xcd_tile(nrt, ngt);
  if (!xt.valid) return;
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6, li = lane & 15, lk = lane >> 4;
  const int g0 = xt.shared * NG16, r0 = xt.inner * WR;
  const int nbh = q.nbh, ngl = q.ngl;
  dbl4 accr[NGB], acci[NGB];
#pragma unroll
  for (int c = 0; c < NGB; ++c) {
    accr[c] = dbl4{0.0, 0.0, 0.0, 0.0};
    acci[c] = dbl4{0.0, 0.0, 0.0, 0.0};
  }
  for (int k0 = 0; k0 < nbh; k0 += KC) {
    for (int e = t; e < WR * KC; e += 64 * NW) {
      const int k = e % KC, r = e / KC, b = k0 + k, ip = r0 + r;
      double2 v = czero();
      if (ip < npro && b < nbh) v = ded[(size_t)ip * nbh + b];
      sAr[k][r] = v.x;
      sAi[k][r] = -v.y;
    }
    for (int e = t; e < NG16 * KC; e += 64 * NW) {
      const int g = e % NG16, k = e / NG16, b = k0 + k, ig = g0 + g;
      double2 v = czero();
      if (b < nbh && ig < ngl) v = psi[(size_t)b * ngl + ig];
      sBr[k][g] = v.x;
      sBi[k][g] = v.y;
    }
    __syncthreads();
#pragma unroll
    for (int ks = 0; ks < KC / 4; ++ks) {
      const int k = 4 * ks + lk;
      const double ar = sAr[k][16 * wv + li], ai = sAi[k][16 * wv + li];
#pragma unroll
      for (int c = 0; c < NGB; ++c) cmfma(accr[c], acci[c], ar, ai, sBr[k][16 * c + li], sBi[k][16 * c + li]);
    }
    __syncthreads();
  }
  double f[4][3];
#pragma unroll
  for (int r = 0; r < 4; ++r) f[r][0] = f[r][1] = f[r][2] = 0.0;
#pragma unroll
  for (int c = 0; c < NGB; ++c) {
    const int ig = g0 + 16 * c + li;
    if (ig >= ngl) continue;
    const double gx = gvec[3 * ig], gy = gvec[3 * ig + 1], gz = gvec[3 * ig + 2];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int ip = r0 + 16 * wv + lk + 4 * r;
      if (ip >= npro) continue;
      const int2 al = iptab[ip];
      const double2 pv = pro_val(q, al.x, al.y, ig, eigr_of(q, al.x, ig));
      // Re(i p conj(d)) = p.x d.y - p.y d.x
      const double cc = pv.x * acci[c][r] - pv.y * accr[c][r];
      f[r][0] += cc * gx;
      f[r][1] += cc * gy;
      f[r][2] += cc * gz;
    }
  }
  // sum over the 16 lanes of a row
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int d = 0; d < 3; ++d) f[r][d] = row16_sum(f[r][d]);
  if (li != 0) return;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int ip = r0 + 16 * wv + lk + 4 * r;
    if (ip < npro) {
      double *dst = part + ((size_t)xt.shared * npro + ip) * 3;
      dst[0] = f[r][0];
      dst[1] = f[r][1];
      dst[2] = f[r][2];
    }
  }
}


// unravel super-wave products into the NB x NB real matrix (row-major),
// WAVES_OVERLAP (src/paw_waves2.f90:805-848); t, tm already times Omega
__global__ void k_unravel_pw(int nbh, int nb, const double2 *t, const double2 *tm, double *mat) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= nbh * nbh) return;
  const int i = idx / nbh, j = idx % nbh;
  const double re = 0.5 * t[idx].x, im = 0.5 * t[idx].y;
  const double rem = 0.5 * tm[idx].x, imm = 0.5 * tm[idx].y;
  const int ia = 2 * i, ja = 2 * j;
  mat[(size_t)ia * nb + ja] = re + rem;
  mat[(size_t)ia * nb + ja + 1] = im + imm;
  mat[(size_t)(ia + 1) * nb + ja] = -im + imm;
  mat[(size_t)(ia + 1) * nb + ja + 1] = re - rem;
}

// ---------------------------------------------------------------------------
// one-centre overlap (WAVES_1COVERLAP, src/paw_waves2.f90:856-1022), added to mat
// c1 = sum_k conj(p1_ki) p2a_kj, c2 = sum_k p1_ki p2a_kj.  The projector index
// k is split over blockIdx.z in runs of k1c rows; part[z][i][j] = (c1, c2)
// of the run, summed in z order by k_overlap_finish.
// ---------------------------------------------------------------------------
constexpr int k1c = 64;
__global__ void __launch_bounds__(256) k_overlap1c(int npro, int nbh, const double2 *p1, const double2 *p2a,
                                                   double4 *part) {
  __shared__ double2 sa[k1c][17];
  __shared__ double2 sb[k1c][17];
  const int ti = threadIdx.x / 16, tj = threadIdx.x % 16;
  const int i0 = blockIdx.x * 16, j0 = blockIdx.y * 16, k0 = blockIdx.z * k1c;
  const int i = i0 + ti, j = j0 + tj;
  for (int w = threadIdx.x; w < k1c * 16; w += 256) {
    const int k = w / 16, c = w % 16;
    const bool ok = k0 + k < npro;
    sa[k][c] = (ok && i0 + c < nbh) ? p1[(size_t)(k0 + k) * nbh + i0 + c] : czero();
    sb[k][c] = (ok && j0 + c < nbh) ? p2a[(size_t)(k0 + k) * nbh + j0 + c] : czero();
  }
  __syncthreads();
  double2 c1 = czero(), c2 = czero();
#pragma unroll 8
  for (int k = 0; k < k1c; ++k) {
    const double2 a = sa[k][ti], b = sb[k][tj];
    c1.x += a.x * b.x + a.y * b.y;
    c1.y += a.x * b.y - a.y * b.x;
    c2.x += a.x * b.x - a.y * b.y;
    c2.y += a.x * b.y + a.y * b.x;
  }
  if (i >= nbh || j >= nbh) return;
  part[((size_t)blockIdx.z * nbh + i) * nbh + j] = make_double4(c1.x, c1.y, c2.x, c2.y);
}

// WAVES_OVERLAP's matrix in one pass from the G-split partials of the
// stacked product (rows 0..nbhp-1: T, rows nbhp..: TM): the sums of
// k_reduce_split, the unravel of k_unravel_pw and, with part1c, the
// one-centre sum over the runs of k_overlap1c in z order.
__global__ void __launch_bounds__(kRedC * kRedS) k_overlap_finish(const double2 *__restrict__ part, int ns, size_t n,
                                                                  double scale, int nbh, int nbhp, int nb,
                                                                  const double4 *__restrict__ part1c, int nz,
                                                                  double *mat) {
  __shared__ double2 red[2][kRedS][kRedC];
  const int c = threadIdx.x % kRedC, sl = threadIdx.x / kRedC;
  const int idx = blockIdx.x * kRedC + c;
  const bool in = idx < nbh * nbh;
  const int i = in ? idx / nbh : 0, j = in ? idx - i * nbh : 0;
  const size_t it = (size_t)i * nbh + j, itm = (size_t)(nbhp + i) * nbh + j;
  double2 at = rz(double2()), am = rz(double2());
  if (in) {
    constexpr int V = 16;
    for (int k0 = sl; k0 < ns; k0 += V * kRedS) {
      double2 vt[V], vm[V];
#pragma unroll
      for (int u = 0; u < V; ++u) {
        const int k = k0 + u * kRedS;
        vt[u] = k < ns ? part[(size_t)k * n + it] : rz(double2());
        vm[u] = k < ns ? part[(size_t)k * n + itm] : rz(double2());
      }
#pragma unroll
      for (int u = 0; u < V; ++u) {
        at = radd(at, vt[u]);
        am = radd(am, vm[u]);
      }
    }
  }
  red[0][sl][c] = at;
  red[1][sl][c] = am;
  __syncthreads();
  if (sl != 0 || !in) return;
  double2 t = red[0][0][c], tm = red[1][0][c];
#pragma unroll
  for (int q = 1; q < kRedS; ++q) {
    t = radd(t, red[0][q][c]);
    tm = radd(tm, red[1][q][c]);
  }
  t = rscale(t, scale);
  tm = rscale(tm, scale);
  const double re = 0.5 * t.x, im = 0.5 * t.y, rem = 0.5 * tm.x, imm = 0.5 * tm.y;
  double m00 = re + rem, m01 = im + imm, m10 = -im + imm, m11 = re - rem;
  if (part1c) {
    double4 cc = make_double4(0.0, 0.0, 0.0, 0.0);
    for (int z = 0; z < nz; ++z) {
      const double4 v = part1c[(size_t)z * nbh * nbh + idx];
      cc.x += v.x;
      cc.y += v.y;
      cc.z += v.z;
      cc.w += v.w;
    }
    const double2 s1 = make_double2(0.5 * (cc.x + cc.z), 0.5 * (cc.y + cc.w));
    const double2 s2 = make_double2(0.5 * (cc.x - cc.z), 0.5 * (cc.y - cc.w));
    m00 += s1.x;
    m01 += s1.y;
    m10 += -s2.y;
    m11 += s2.x;
  }
  const int ia = 2 * i, ja = 2 * j;
  mat[(size_t)ia * nb + ja] = m00;
  mat[(size_t)ia * nb + ja + 1] = m01;
  mat[(size_t)(ia + 1) * nb + ja] = m10;
  mat[(size_t)(ia + 1) * nb + ja + 1] = m11;
}

// per-atom matrix on projections: out[ipro0+i][b] = sum_j M_atom[i][j] proj[ipro0+j][b]
// grid (NAT, LMNXX*NBH/64), 64 threads: one output element per thread
__global__ void __launch_bounds__(64) k_atom_mat(ProjGeo q, const double2 *proj, const double2 *mats,
                                                 double2 *out) {
  const int atom = blockIdx.x;
  const int lmnx = q.lmnx[q.atom_isp[atom]];
  const int w = blockIdx.y * 64 + threadIdx.x;
  if (w >= lmnx * q.nbh) return;
  const int base = q.ipro0[atom];
  const double2 *m = mats + (size_t)atom * q.lmnxx * q.lmnxx;
  const int i = w / q.nbh, b = w % q.nbh;
  double2 s = czero();
#pragma unroll 8
  for (int j = 0; j < lmnx; ++j) s = cadd(s, cmul(m[i * q.lmnxx + j], proj[(size_t)(base + j) * q.nbh + b]));
  out[(size_t)(base + i) * q.nbh + b] = s;
}

// lambda (real NB x NB, row-major) -> LAMBDA1/LAMBDA2 element (i,j) of src/paw_waves2.f90:582-593
__device__ inline void fold_lambda(const double *lam, int nb, int i, int j, double2 &l1, double2 &l2) {
  const double a11 = lam[(size_t)(2 * i) * nb + 2 * j];
  const double a12 = lam[(size_t)(2 * i) * nb + 2 * j + 1];
  const double a21 = lam[(size_t)(2 * i + 1) * nb + 2 * j];
  const double a22 = lam[(size_t)(2 * i + 1) * nb + 2 * j + 1];
  // cs1 = a11 + i a12 ; cs2 = i a21 - a22
  l1 = make_double2(0.5 * (a11 + a22), 0.5 * (a12 - a21));
  l2 = make_double2(0.5 * (a11 - a22), 0.5 * (a12 + a21));
}

// ADDOPSI: psibar[j][G] += sum_i L1[i][j] chi[i][G] + L2[i][j] conj(chi[i][-G])
// as one column GEMM over the stacked K = 2 NBH
struct AddopsiA {
  const double *lam;
  int nb, nbh;
  __device__ double2 operator()(int k, int j) const {
    double2 l1, l2;
    fold_lambda(lam, nb, k < nbh ? k : k - nbh, j, l1, l2);
    return k < nbh ? l1 : l2;
  }
};
struct AddopsiB {
  const double2 *chi;
  const int *minusg;
  int ngl, nbh;
  __device__ double2 operator()(int k, int ig) const {
    return k < nbh ? chi[(size_t)k * ngl + ig] : cconj(chi[(size_t)(k - nbh) * ngl + minusg[ig]]);
  }
};

// ---------------------------------------------------------------------------
// Split-BF16x3 MFMA building blocks (mixed form).  A value x is carried as
// three bf16 pieces x = h + m + l (h = bf16(x), m = bf16(x - h),
// l = bf16(x - h - m): 24 significant bits, the precision of FP32), and a
// product a*b as the six piece products of order >= 2^-16,
//   ah bh + ah bm + am bh + ah bl + am bm + al bh,
// on v_mfma_f32_16x16x32_bf16 with FP32 accumulation.  Lane layout (checked
// on the device): A(i,k) in lane i + 16(k/8), element k%8; B(k,j) in lane
// j + 16(k/8), element k%8; C(i,j) in lane j + 16(i/4), element i%4.
// Operands live in LDS as six planes (re h,m,l, im h,m,l), K contiguous per
// row/column, rows padded to 80 B so that the 16-byte operand reads of 16
// lanes cover all 64 banks once.
// ---------------------------------------------------------------------------
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float flt4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

constexpr int KB = 32;        // K of one MFMA = K of one staged chunk
constexpr int KBP = KB + 8;   // LDS row stride in bf16
constexpr int NPL = 6;        // planes per complex operand

__device__ inline flt4 mfma_bf(const bf16x8 &a, const bf16x8 &b, flt4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

__device__ inline bf16x8 bf_neg(const bf16x8 &a) {
  return __builtin_bit_cast(bf16x8, __builtin_bit_cast(u32x4, a) ^ 0x80008000u);
}

// split planes of 8 consecutive complex values: p[0..2] re h,m,l, p[3..5] im h,m,l
struct Planes8 {
  bf16x8 p[NPL];
};

struct Split3 {
  __bf16 h, m, l;
};

// f rounded to bf16 (nearest, ties to even) on the integer path; returns the
// bf16 value as a float, its 16 bits in b
__device__ inline float bf16_rne(float f, __bf16 &b) {
  unsigned u = __float_as_uint(f);
  u = (u + 0x7FFFu + ((u >> 16) & 1u)) & 0xFFFF0000u;
  b = __builtin_bit_cast(__bf16, (unsigned short)(u >> 16));
  return __uint_as_float(u);
}

// x -> FP32 -> h + m + l, each a bf16 and each remainder exact in FP32
__device__ inline Split3 split3(double x) {
  Split3 s;
  const float f = (float)x;
  const float r = f - bf16_rne(f, s.h);
  const float q = r - bf16_rne(r, s.m);
  bf16_rne(q, s.l);
  return s;
}

__device__ inline void put(Planes8 &o, int e, double2 v) {
  const Split3 a = split3(v.x), b = split3(v.y);
  o.p[0][e] = a.h;
  o.p[1][e] = a.m;
  o.p[2][e] = a.l;
  o.p[3][e] = b.h;
  o.p[4][e] = b.m;
  o.p[5][e] = b.l;
}

__device__ inline void zero8(Planes8 &o) {
  const bf16x8 z = __builtin_bit_cast(bf16x8, u32x4{0u, 0u, 0u, 0u});
#pragma unroll
  for (int q = 0; q < NPL; ++q) o.p[q] = z;
}

// real product x*y of split operands (three pieces each) accumulated in c
__device__ inline flt4 rmfma3(const bf16x8 *x, const bf16x8 *y, flt4 c) {
  c = mfma_bf(x[2], y[0], c);
  c = mfma_bf(x[1], y[1], c);
  c = mfma_bf(x[0], y[2], c);
  c = mfma_bf(x[1], y[0], c);
  c = mfma_bf(x[0], y[1], c);
  c = mfma_bf(x[0], y[0], c);
  return c;
}

// (cr + i ci) += a b, a and b split; an = the negated im pieces of a
__device__ inline void cmfma_bf(flt4 &cr, flt4 &ci, const bf16x8 *a, const bf16x8 *an, const bf16x8 *b) {
  cr = rmfma3(a, b, cr);
  cr = rmfma3(an, b + 3, cr);
  ci = rmfma3(a, b + 3, ci);
  ci = rmfma3(a + 3, b, ci);
}
// the same with the sign on B's imaginary planes (bn = -Im b): for products
// whose B operand is shared by the A tiles of a wave, negated once per chunk
__device__ inline void cmfma_bf_nb(flt4 &cr, flt4 &ci, const bf16x8 *a, const bf16x8 *b, const bf16x8 *bn) {
  cr = rmfma3(a, b, cr);
  cr = rmfma3(a + 3, bn, cr);
  ci = rmfma3(a, b + 3, ci);
  ci = rmfma3(a + 3, b, ci);
}

__device__ inline void store_planes(__bf16 (*pl)[KBP], int plane_rows, int row, int k8, const Planes8 &v) {
#pragma unroll
  for (int q = 0; q < NPL; ++q) *reinterpret_cast<bf16x8 *>(&pl[q * plane_rows + row][8 * k8]) = v.p[q];
}

__device__ inline void load_planes(const __bf16 (*pl)[KBP], int plane_rows, int row, int k8, bf16x8 *v) {
#pragma unroll
  for (int q = 0; q < NPL; ++q) v[q] = *reinterpret_cast<const bf16x8 *>(&pl[q * plane_rows + row][8 * k8]);
}

// Projector functions of the mixed form from two factor tables,
//   P(atom,lmn,G) = (-i)^l F(isp,lmn,G) E(atom,G),
//   F = YLM(G,lm) PRO(G,ln) [NSP*LMNXX][NGLP],  E = exp(-iG.R_atom) [NAT][NGLP]
// (WAVES_EXPANDPRO, src/paw_waves1.f90:6070-6109).  Both tables are FP64 and
// padded with zeros to NGLP (a multiple of 64) G; they hold (NSP*LMNXX*8 +
// NAT*16) bytes per G instead of the NPRO*16 of a stored projector matrix, so
// the products read them from cache, and each product forms P in FP64 while
// it stages its operands.
__global__ void k_pfac_f(ProjGeo q, int nsp, int nglp, double *f, float *f32) {
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (size_t)nsp * q.lmnxx * nglp) return;
  const int k = (int)(idx / nglp), ig = (int)(idx % nglp);
  const int isp = k / q.lmnxx, lmn = k % q.lmnxx;
  double v = 0.0;
  if (ig < q.ngl && lmn < q.lmnx[isp])
    v = q.ylm[(size_t)q.lmn_lm[k] * q.ngl + ig] * q.pro[(size_t)(q.ibpro[isp] + q.lmn_ln[k]) * q.ngl + ig];
  f[idx] = v;
  f32[idx] = (float)v;
}

__global__ void k_pfac_e(ProjGeo q, int nglp, double2 *e, float2 *e32) {
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (size_t)q.nat * nglp) return;
  const int atom = (int)(idx / nglp), ig = (int)(idx % nglp);
  const double2 v = ig < q.ngl ? eigr_of(q, atom, ig) : czero();
  e[idx] = v;
  e32[idx] = make_float2((float)v.x, (float)v.y);
}

// (-i)^l = cr + i ci
__device__ inline void mil(int l, float &cr, float &ci) {
  const int m = l & 3;
  cr = (m & 1) ? 0.f : (m ? -1.f : 1.f);
  ci = (m & 1) ? (m == 1 ? -1.f : 1.f) : 0.f;
}

// split of an FP32 value: h + m + l, each a bf16 and each remainder exact
__device__ inline Split3 split3f(float f) {
  Split3 s;
  const float r = f - bf16_rne(f, s.h);
  const float q = r - bf16_rne(r, s.m);
  // q holds at most 8 significant bits: its bf16 is exact
  s.l = __builtin_bit_cast(__bf16, (unsigned short)(__float_as_uint(q) >> 16));
  return s;
}

__device__ inline void put_f(Planes8 &o, int e, float re, float im) {
  const Split3 a = split3f(re), b = split3f(im);
  o.p[0][e] = a.h;
  o.p[1][e] = a.m;
  o.p[2][e] = a.l;
  o.p[3][e] = b.h;
  o.p[4][e] = b.m;
  o.p[5][e] = b.l;
}

struct PFac {
  const double *f;
  const double2 *e;
  const float *f32;    // F and E rounded to FP32: the products stage P from these
  const float2 *e32;
  const int2 *iptab;
  const int *atom_isp, *lmn_l;
  int lmnxx, nglp;
  // row of projector ip: F row, E row and l
  __device__ void row(int ip, const double *&fr, const double2 *&er, int &l) const {
    const int2 al = iptab[ip];
    const int k = atom_isp[al.x] * lmnxx + al.y;
    fr = f + (size_t)k * nglp;
    er = e + (size_t)al.x * nglp;
    l = lmn_l[k];
  }
  __device__ double2 at(int ip, int ig) const {
    const double *fr;
    const double2 *er;
    int l;
    row(ip, fr, er, l);
    const double s = fr[ig];
    const double2 ee = er[ig];
    return rot_mil(make_double2(s * ee.x, s * ee.y), l);
  }
};

// ---- row operands (8 consecutive G of one row) -------------------------------
// conj(P(row, G)) from the projector factors
struct RowPf {
  PFac pf;
  // a thread's row is the same in every K chunk: its FP32 table rows and
  // (-i)^l are looked up once
  struct Row {
    const float *fr;
    const float2 *er;
    float cr, ci;
  };
  __device__ Row row(int r) const {
    const int2 al = pf.iptab[r];
    const int k = pf.atom_isp[al.x] * pf.lmnxx + al.y;
    Row h;
    h.fr = pf.f32 + (size_t)k * pf.nglp;
    h.er = pf.e32 + (size_t)al.x * pf.nglp;
    mil(pf.lmn_l[k], h.cr, h.ci);
    return h;
  }
  // k_rowgemm_bf fetches the next G chunk's table values before the current chunk's MFMAs
  struct Pre {
    float sv[8];
    float2 ee[8];
  };
  __device__ void fetch(const Row &h, int g0, Pre &p) const {
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      p.sv[e] = h.fr[g0 + e];
      p.ee[e] = h.er[g0 + e];
    }
  }
  // conj(P) = conj(F (-i)^l E)
  __device__ void make(const Row &h, const Pre &p, Planes8 &o) const {
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const float2 ee = p.ee[e];
      const float tr = h.cr * ee.x - h.ci * ee.y, ti = h.cr * ee.y + h.ci * ee.x;
      put_f(o, e, p.sv[e] * tr, -(p.sv[e] * ti));
    }
  }
};

// ---- operands split once into bf16 runs --------------------------------------
// SP[row][run] holds the six planes (bf16x8 each, 96 B) of the complex values
// (row, 8 run .. 8 run + 7).  nrun covers K rounded up to whole 32-chunks and
// the tail is zero.  An operand that several tiles of a product read (the
// waves of the row GEMMs, the coefficients of the column GEMMs) is split once
// by k_split_rows / k_split_cols; the product then stages a run with six
// 16-byte copies.
// Layouts: row-major SP[row][run] (a thread's run is 8 consecutive K, the
// threads of a wave take consecutive runs) or run-major SP[run][row] (the
// threads of a wave take consecutive rows: column operands, whose K runs
// gather across bands).
// Row-major runs: one lane per complex value.  The 8 lanes of a run read 8
// consecutive values (a wave reads 1 KB contiguous), write their bf16 pieces
// into an LDS image of the run (plane q, element e at 16 q + 2 e bytes) and
// store the 96-byte run back as 12 contiguous bytes per lane.
template <class Src>
__global__ void __launch_bounds__(256) k_split_rows(Src src, int nrow, int nrun, unsigned *sp) {
  __shared__ __bf16 img[32][8 * NPL];
  const int t = threadIdx.x, e = t & 7, lr = t >> 3;
  const size_t id = (size_t)blockIdx.x * 32 + lr;
  const bool ok = id < (size_t)nrow * nrun;
  double2 v = make_double2(0.0, 0.0);
  if (ok) {
    // nrow * nrun < 2^32 (an operand of 2^32 runs would take 384 GiB): 32-bit quotients
    const unsigned u = (unsigned)id;
    const int r = (int)(u / (unsigned)nrun), run = (int)(u % (unsigned)nrun);
    v = src(r, 8 * run + e);
  }
  const Split3 a = split3(v.x), b = split3(v.y);
  img[lr][e] = a.h;
  img[lr][8 + e] = a.m;
  img[lr][16 + e] = a.l;
  img[lr][24 + e] = b.h;
  img[lr][32 + e] = b.m;
  img[lr][40 + e] = b.l;
  __syncthreads();
  if (!ok) return;
  const unsigned *w = reinterpret_cast<const unsigned *>(&img[lr][0]) + 3 * e;
  unsigned *d = sp + id * (NPL * 4) + 3 * e;
  d[0] = w[0];
  d[1] = w[1];
  d[2] = w[2];
}

// Run-major runs: one thread per run (8 K of one row; the lanes of a wave
// take consecutive rows, so each of its 8 loads is coalesced).  The wave's
// 64 runs are contiguous in the output: they go through LDS and out as six
// stores of 16 contiguous bytes per lane.
template <class Src>
__global__ void __launch_bounds__(256) k_split_cols(Src src, int nrow, int nrun, u32x4 *sp) {
  __shared__ u32x4 img[256 * NPL];
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t n = (size_t)nrow * nrun;
  Planes8 o;
  if (idx < n) {
    const unsigned u = (unsigned)idx;   // < nrow * nrun < 2^32
    const int r = (int)(u % (unsigned)nrow), run = (int)(u / (unsigned)nrow);
#pragma unroll
    for (int e = 0; e < 8; ++e) put(o, e, src(r, 8 * run + e));
  } else {
    zero8(o);
  }
#pragma unroll
  for (int q = 0; q < NPL; ++q) img[threadIdx.x * NPL + q] = __builtin_bit_cast(u32x4, o.p[q]);
  __syncthreads();
  const size_t base = (size_t)blockIdx.x * blockDim.x * NPL;
  const size_t lim = n * NPL;
#pragma unroll
  for (int q = 0; q < NPL; ++q) {
    const size_t k = (size_t)q * blockDim.x + threadIdx.x;
    if (base + k < lim) sp[base + k] = img[k];
  }
}

struct SplitRuns {
  const u32x4 *sp;
  int nrow, nrun;
  int srow, srun;   // strides in runs: row-major (nrun, 1), run-major (1, nrow)
  __device__ void load(int r, int k0, Planes8 &o) const {
    const int run = k0 >> 3;
    if (r >= nrow || run >= nrun) {
      zero8(o);
      return;
    }
    const u32x4 *src = sp + ((size_t)r * srow + (size_t)run * srun) * NPL;
#pragma unroll
    for (int q = 0; q < NPL; ++q) o.p[q] = __builtin_bit_cast(bf16x8, src[q]);
  }
};
// rows [conj(psi1_i(G)); psi1_i(-G)] of the band overlaps, i < nbhp each, from
// the runs of psi1 (conjugated by flipping the signs of its imaginary planes)
// and of psi1(-G)
struct OvlRuns {
  SplitRuns rows, minus;
  int nbhp;
  typedef int Row;
  __device__ Row row(int r) const { return r; }
  struct Pre {
    Planes8 v;
  };
  __device__ void fetch(int r, int g0, Pre &p) const {
    if (r < nbhp) rows.load(r, g0, p.v);
    else minus.load(r - nbhp, g0, p.v);
  }
  __device__ void make(int r, const Pre &p, Planes8 &o) const {
    o = p.v;
    if (r < nbhp) {
#pragma unroll
      for (int q = 3; q < NPL; ++q) o.p[q] = bf_neg(o.p[q]);
    }
  }
};
// band operand of k_rowgemm_bf: band b, G run at g0
struct BandRuns {
  SplitRuns s;
  __device__ void operator()(int b, int g0, Planes8 &o) const { s.load(b, g0, o); }
};
// A operand of k_colgemm_bf and k_force_bf: row b, K run at k0
struct ARuns {
  SplitRuns s;
  __device__ void operator()(int b, int k0, Planes8 &o) const { s.load(b, k0, o); }
};
// B operand of k_colgemm_bf and k_force_bf: K run at k0 of column G
struct ColRuns {
  SplitRuns s;
  __device__ void operator()(int k0, int ig, Planes8 &o) const { s.load(ig, k0, o); }
  // k_colgemm_bf fetches the next K chunk's operand before the current chunk's MFMAs
  struct Pre {
    Planes8 v;
  };
  __device__ void fetch(int k0, int ig, Pre &p) const { s.load(ig, k0, p.v); }
  __device__ void make(const Pre &p, int, int, Planes8 &o) const { o = p.v; }
};

// sources of k_split_rows / k_split_cols, (row, k) -> complex value, zero outside
struct SrcPsiRows {          // psi[b][G], rows b
  const double2 *psi;
  int ngl, nbh;
  __device__ double2 operator()(int b, int g) const {
    return (b < nbh && g < ngl) ? psi[(size_t)b * ngl + g] : czero();
  }
};
struct SrcPsiCols {          // psi[b][G], rows G, K = b
  const double2 *psi;
  int ngl, nbh;
  __device__ double2 operator()(int g, int b) const {
    return (b < nbh && g < ngl) ? psi[(size_t)b * ngl + g] : czero();
  }
};
struct SrcPsiMinusRows {     // psi[b][-G], rows b
  const double2 *psi;
  const int *minusg;
  int ngl, nbh;
  __device__ double2 operator()(int b, int g) const {
    return (b < nbh && g < ngl) ? psi[(size_t)b * ngl + minusg[g]] : czero();
  }
};
struct SrcXRows {            // x[k][b] (projector-major), rows b, K = k
  const double2 *x;
  int nbh, nk;
  __device__ double2 operator()(int b, int k) const {
    return (b < nbh && k < nk) ? x[(size_t)k * nbh + b] : czero();
  }
};
struct SrcDedRows {          // conj(ded[ip][b]), rows ip, K = b
  const double2 *ded;
  int npro, nbh;
  __device__ double2 operator()(int ip, int b) const {
    return (ip < npro && b < nbh) ? cconj(ded[(size_t)ip * nbh + b]) : czero();
  }
};

// part[split][r][b] = sum_{G in split} A(r,G) psi[b][G] (split-BF16x3).
// Each 32-G chunk is summed by the MFMAs from zero in FP32 and then added to
// FP64 accumulators: the FP32 roundings are those of one chunk's sum, never of
// the running total (the products of localised projectors and extended waves
// cancel strongly over G, and an FP32 running total loses ~1e-6 there).
template <class Gen, class GB, int NCB, int NW>
__global__ void __launch_bounds__(64 * NW) k_rowgemm_bf(Gen gen, GB gb, int nrow, int ngl, int nbh, int gsplit,
                                                        int tr, int tb, int ns, double2 *__restrict__ part) {
  constexpr int WR = 16 * NW, NB16 = 16 * NCB;
  __shared__ __bf16 sA[NPL * WR][KBP];
  __shared__ __bf16 sB[NPL * NB16][KBP];
  const XcdTile xt = xcd_tile(tr * tb, ns);
  if (!xt.valid) return;
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6, li = lane & 15, lk = lane >> 4;
  const int r0 = (xt.inner % tr) * WR, b0 = (xt.inner / tr) * NB16;
  const int g0 = xt.shared * gsplit, g1 = min(ngl, g0 + gsplit);
  dbl4 accr[NCB], acci[NCB];
#pragma unroll
  for (int c = 0; c < NCB; ++c) {
    accr[c] = dbl4{0.0, 0.0, 0.0, 0.0};
    acci[c] = dbl4{0.0, 0.0, 0.0, 0.0};
  }
  const int ra = t >> 2, ka = t & 3;          // WR rows x 4 runs = 64 NW runs, one per thread
  const bool rok = r0 + ra < nrow;
  const typename Gen::Row rh = gen.row(rok ? r0 + ra : 0);
  typename Gen::Pre pa;
  if (rok && g0 < g1) gen.fetch(rh, g0 + 8 * ka, pa);
  // the band runs of a chunk (NB16 x 4 runs, NSH per thread) are loaded into
  // registers one chunk ahead, so their latency overlaps the MFMAs, when at
  // most two runs fall to a thread; with more (2-wave workgroups) the
  // registers would halve the resident waves, and they are loaded in place
  constexpr int NSH = (NB16 * 4 + 64 * NW - 1) / (64 * NW);
  constexpr bool PRE = NSH <= 2;
  Planes8 shb[PRE ? NSH : 1];
  auto fetch_b = [&](int gc) {
#pragma unroll
    for (int i = 0; i < NSH; ++i) {
      const int e = t + i * 64 * NW;
      if (e < NB16 * 4) gb(b0 + (e >> 2), gc + 8 * (e & 3), shb[i]);
    }
  };
  if constexpr (PRE) {
    if (g0 < g1) fetch_b(g0);
  }
  for (int gc = g0; gc < g1; gc += KB) {
    {
      Planes8 v;
      if (rok) gen.make(rh, pa, v);
      else zero8(v);
      store_planes(sA, WR, ra, ka, v);
    }
    if constexpr (PRE) {
#pragma unroll
      for (int i = 0; i < NSH; ++i) {
        const int e = t + i * 64 * NW;
        if (e < NB16 * 4) store_planes(sB, NB16, e >> 2, e & 3, shb[i]);
      }
    } else {
      for (int e = t; e < NB16 * 4; e += 64 * NW) {
        Planes8 v;
        gb(b0 + (e >> 2), gc + 8 * (e & 3), v);
        store_planes(sB, NB16, e >> 2, e & 3, v);
      }
    }
    __syncthreads();
    if (rok && gc + KB < g1) gen.fetch(rh, gc + KB + 8 * ka, pa);
    if constexpr (PRE) {
      if (gc + KB < g1) fetch_b(gc + KB);
    }
    bf16x8 a[NPL], an[3];
    load_planes(sA, WR, 16 * wv + li, lk, a);
#pragma unroll
    for (int q = 0; q < 3; ++q) an[q] = bf_neg(a[3 + q]);
#pragma unroll
    for (int c = 0; c < NCB; ++c) {
      bf16x8 b[NPL];
      load_planes(sB, NB16, 16 * c + li, lk, b);
      flt4 cr = flt4{0.f, 0.f, 0.f, 0.f}, ci = flt4{0.f, 0.f, 0.f, 0.f};
      cmfma_bf(cr, ci, a, an, b);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        accr[c][r] += (double)cr[r];
        acci[c][r] += (double)ci[r];
      }
    }
    __syncthreads();
  }
  double2 *dst = part + (size_t)xt.shared * nrow * nbh;
#pragma unroll
  for (int c = 0; c < NCB; ++c) {
    const int j = b0 + 16 * c + li;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int i = r0 + 16 * wv + 4 * lk + r;
      if (i < nrow && j < nbh) dst[(size_t)i * nbh + j] = make_double2(accr[c][r], acci[c][r]);
    }
  }
}

// ---- column GEMM: psi[b][G] += sum_k A(b,k) B(k,G) (split-BF16x3) -----------
// B operand of ADDPRO: 8 consecutive projectors k at one G, from the factors
struct AddproBf {
  PFac pf;
  int nk;
  // P(k, G) from the FP32 tables, zero beyond NPRO
  __device__ float2 at(int k, int ig) const {
    if (k >= nk) return make_float2(0.f, 0.f);
    const int2 al = pf.iptab[k];
    const int kr = pf.atom_isp[al.x] * pf.lmnxx + al.y;
    const float sv = pf.f32[(size_t)kr * pf.nglp + ig];
    const float2 ee = pf.e32[(size_t)al.x * pf.nglp + ig];
    float cr, ci;
    mil(pf.lmn_l[kr], cr, ci);
    return make_float2(sv * (cr * ee.x - ci * ee.y), sv * (cr * ee.y + ci * ee.x));
  }
  // the table values of the next K chunk, fetched before the current chunk's
  // MFMAs: eight projectors of one atom share one structure factor; other
  // runs (atom boundaries, the tail) are formed element by element in make
  struct Pre {
    float2 ee;
    float sv[8];
    bool one;
  };
  __device__ void fetch(int k0, int ig, Pre &p) const {
    p.one = ig < pf.nglp && k0 + 7 < nk && pf.iptab[k0].x == pf.iptab[k0 + 7].x;
    if (!p.one) return;
    const int2 al = pf.iptab[k0];
    p.ee = pf.e32[(size_t)al.x * pf.nglp + ig];
    const int k = pf.atom_isp[al.x] * pf.lmnxx + al.y;
#pragma unroll
    for (int e = 0; e < 8; ++e) p.sv[e] = pf.f32[(size_t)(k + e) * pf.nglp + ig];
  }
  __device__ void make(const Pre &p, int k0, int ig, Planes8 &o) const {
    if (ig >= pf.nglp) {
      zero8(o);
      return;
    }
    if (!p.one) {
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const float2 v = at(k0 + e, ig);
        put_f(o, e, v.x, v.y);
      }
      return;
    }
    const int2 al = pf.iptab[k0];
    const int k = pf.atom_isp[al.x] * pf.lmnxx + al.y;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      float cr, ci;
      mil(pf.lmn_l[k + e], cr, ci);
      const float tr = cr * p.ee.x - ci * p.ee.y, ti = cr * p.ee.y + ci * p.ee.x;
      put_f(o, e, p.sv[e] * tr, p.sv[e] * ti);
    }
  }
};
struct SrcLamRows {          // LAMBDA1/LAMBDA2 stacked along K = 2 NBH, rows b
  const double *lam;
  int nb, nbh;
  __device__ double2 operator()(int b, int k) const {
    if (k >= 2 * nbh || b >= nbh) return czero();
    double2 l1, l2;
    fold_lambda(lam, nb, k < nbh ? k : k - nbh, b, l1, l2);
    return k < nbh ? l1 : l2;
  }
};
struct SrcChiCols {          // [chi_k(G); conj(chi_k(-G))] along K = 2 NBH, rows G
  const double2 *chi;
  const int *minusg;
  int ngl, nbh;
  __device__ double2 operator()(int g, int k) const {
    if (g >= ngl || k >= 2 * nbh) return czero();
    return k < nbh ? chi[(size_t)k * ngl + g] : cconj(chi[(size_t)(k - nbh) * ngl + minusg[g]]);
  }
};

template <class GA, class GB, int NRB, int NW>
__global__ void __launch_bounds__(64 * NW) k_colgemm_bf(GA ga, GB gb, int nk, int ngl, int nbh, ColOut out) {
  constexpr int WG = 16 * NW, NB16 = 16 * NRB;
  __shared__ __bf16 sB[NPL * WG][KBP];
  __shared__ __bf16 sA[NPL * NB16][KBP];
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6, li = lane & 15, lk = lane >> 4;
  const int g0 = blockIdx.x * WG, b0 = blockIdx.y * NB16;
  dbl4 accr[NRB], acci[NRB];   // FP64 sums of per-chunk FP32 MFMA sums
#pragma unroll
  for (int c = 0; c < NRB; ++c) {
    accr[c] = dbl4{0.0, 0.0, 0.0, 0.0};
    acci[c] = dbl4{0.0, 0.0, 0.0, 0.0};
  }
  const int gs = t % WG, ks = t / WG;          // WG columns x 4 runs, one per thread
  // with WG = 64 a wave's K run is uniform: its index lookups become scalar loads
  auto krun = [&](int k0) { return WG == 64 ? __builtin_amdgcn_readfirstlane(k0 + 8 * ks) : k0 + 8 * ks; };
  typename GB::Pre pre;
  gb.fetch(krun(0), g0 + gs, pre);
  // the coefficient runs of a chunk (NB16 x 4 runs, up to NSH per thread) are
  // loaded into registers one chunk ahead, so their latency overlaps the MFMAs
  constexpr int NSH = (NB16 * 4 + 64 * NW - 1) / (64 * NW);
  Planes8 sha[NSH];
  auto fetch_a = [&](int k0) {
#pragma unroll
    for (int i = 0; i < NSH; ++i) {
      const int e = t + i * 64 * NW;
      if (e < NB16 * 4) ga(b0 + (e >> 2), k0 + 8 * (e & 3), sha[i]);
    }
  };
  fetch_a(0);
  for (int k0 = 0; k0 < nk; k0 += KB) {
    {
      Planes8 v;
      gb.make(pre, krun(k0), g0 + gs, v);
      store_planes(sB, WG, gs, ks, v);
    }
#pragma unroll
    for (int i = 0; i < NSH; ++i) {
      const int e = t + i * 64 * NW;
      if (e < NB16 * 4) store_planes(sA, NB16, e >> 2, e & 3, sha[i]);
    }
    __syncthreads();
    if (k0 + KB < nk) {
      gb.fetch(krun(k0 + KB), g0 + gs, pre);
      fetch_a(k0 + KB);
    }
    bf16x8 bb[NPL], bn[3];
    load_planes(sB, WG, 16 * wv + li, lk, bb);
#pragma unroll
    for (int q = 0; q < 3; ++q) bn[q] = bf_neg(bb[3 + q]);
#pragma unroll
    for (int c = 0; c < NRB; ++c) {
      bf16x8 a[NPL];
      load_planes(sA, NB16, 16 * c + li, lk, a);
      flt4 cr = flt4{0.f, 0.f, 0.f, 0.f}, ci = flt4{0.f, 0.f, 0.f, 0.f};
      cmfma_bf_nb(cr, ci, a, bb, bn);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        accr[c][r] += (double)cr[r];
        acci[c][r] += (double)ci[r];
      }
    }
    __syncthreads();
  }
  const int ig = g0 + 16 * wv + li;
  if (ig >= ngl) return;
#pragma unroll
  for (int c = 0; c < NRB; ++c) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int b = b0 + 16 * c + 4 * lk + r;
      if (b < nbh) out.store((size_t)b * ngl + ig, ig, accr[c][r], acci[c][r]);
    }
  }
}

// ---- forces in the mixed form ------------------------------------------------
// D(ipro,G) = sum_b conj(ded[ipro][b]) psi[b][G] as a split-BF16x3 column
// product (rows ipro, columns G, FP32 chunk sums).  The epilogue forms
// P(ipro,G) in FP64 from the projector factors and sums Re(i P conj(D)) G
// over the block's G range in FP64: part[gtile][ipro][3] (WAVES_DEDPRO +
// WAVES_PROFORCE, src/paw_waves1.f90:5694-5745, 6112-6234).
template <class GA, class GB, int NRB, int NW>
__global__ void __launch_bounds__(64 * NW) k_force_bf(GA ga, GB gb, int npro, int nk, int ngl, PFac pf,
                                                      const double *__restrict__ gvec, int nrt, int ngt,
                                                      double *__restrict__ part) {
  constexpr int WG = 16 * NW, NB16 = 16 * NRB;
  __shared__ __bf16 sB[NPL * WG][KBP];
  __shared__ __align__(16) __bf16 sA[NPL * NB16][KBP];
  // per-wave row sums of the epilogue reuse the A staging area (free after the K loop)
  static_assert(sizeof(double) * NW * NB16 * 3 <= sizeof(__bf16) * NPL * NB16 * KBP, "sF fits in sA");
  double (*sF)[NB16][3] = reinterpret_cast<double (*)[NB16][3]>(&sA[0][0]);
  // F row, atom and l of the block's projector rows
  __shared__ int sRow[NB16], sAtom[NB16], sL[NB16];
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6, li = lane & 15, lk = lane >> 4;
  const XcdTile xt = xcd_tile(nrt, ngt);
  if (!xt.valid) return;
  const int g0 = xt.shared * WG, p0 = xt.inner * NB16;
  if (t < NB16) {
    int row = 0, atom = 0, l = 0;
    if (p0 + t < npro) {
      const int2 al = pf.iptab[p0 + t];
      row = pf.atom_isp[al.x] * pf.lmnxx + al.y;
      atom = al.x;
      l = pf.lmn_l[row];
    }
    sRow[t] = row;
    sAtom[t] = atom;
    sL[t] = l;
  }
  flt4 accr[NRB], acci[NRB];
#pragma unroll
  for (int c = 0; c < NRB; ++c) {
    accr[c] = flt4{0.f, 0.f, 0.f, 0.f};
    acci[c] = flt4{0.f, 0.f, 0.f, 0.f};
  }
  const int gs = t % WG, ks = t / WG;
  typename GB::Pre pre;
  gb.fetch(8 * ks, g0 + gs, pre);
  // the projector-row runs of a chunk, loaded into registers one chunk ahead
  constexpr int NSH = (NB16 * 4 + 64 * NW - 1) / (64 * NW);
  Planes8 sha[NSH];
  auto fetch_a = [&](int k0) {
#pragma unroll
    for (int i = 0; i < NSH; ++i) {
      const int e = t + i * 64 * NW;
      if (e < NB16 * 4) ga(p0 + (e >> 2), k0 + 8 * (e & 3), sha[i]);
    }
  };
  fetch_a(0);
  for (int k0 = 0; k0 < nk; k0 += KB) {
    {
      Planes8 v;
      gb.make(pre, k0 + 8 * ks, g0 + gs, v);
      store_planes(sB, WG, gs, ks, v);
    }
#pragma unroll
    for (int i = 0; i < NSH; ++i) {
      const int e = t + i * 64 * NW;
      if (e < NB16 * 4) store_planes(sA, NB16, e >> 2, e & 3, sha[i]);
    }
    __syncthreads();
    if (k0 + KB < nk) {
      gb.fetch(k0 + KB + 8 * ks, g0 + gs, pre);
      fetch_a(k0 + KB);
    }
    bf16x8 bb[NPL], bn[3];
    load_planes(sB, WG, 16 * wv + li, lk, bb);
#pragma unroll
    for (int q = 0; q < 3; ++q) bn[q] = bf_neg(bb[3 + q]);
#pragma unroll
    for (int c = 0; c < NRB; ++c) {
      bf16x8 a[NPL];
      load_planes(sA, NB16, 16 * c + li, lk, a);
      cmfma_bf_nb(accr[c], acci[c], a, bb, bn);
    }
    __syncthreads();
  }
  // C(i = ipro row, j = G column) sits in lane j + 16 (i/4), element i%4
  const int ig = g0 + 16 * wv + li;
  double gx = 0.0, gy = 0.0, gz = 0.0;
  if (ig < ngl) {
    gx = gvec[3 * ig];
    gy = gvec[3 * ig + 1];
    gz = gvec[3 * ig + 2];
  }
#pragma unroll
  for (int c = 0; c < NRB; ++c) {
    double f[4][3];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int ip = p0 + 16 * c + 4 * lk + r;
      double cc = 0.0;
      if (ig < ngl && ip < npro) {
        const int lr = 16 * c + 4 * lk + r;
        const double sv = pf.f[(size_t)sRow[lr] * pf.nglp + ig];
        const double2 ee = pf.e[(size_t)sAtom[lr] * pf.nglp + ig];
        const double2 pv = rot_mil(make_double2(sv * ee.x, sv * ee.y), sL[lr]);
        // Re(i p conj(d)) = p.x d.y - p.y d.x
        cc = pv.x * (double)acci[c][r] - pv.y * (double)accr[c][r];
      }
      f[r][0] = cc * gx;
      f[r][1] = cc * gy;
      f[r][2] = cc * gz;
    }
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
      for (int d = 0; d < 3; ++d) f[r][d] = row16_sum(f[r][d]);
    if (li == 0) {
#pragma unroll
      for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int d = 0; d < 3; ++d) sF[wv][16 * c + 4 * lk + r][d] = f[r][d];
    }
  }
  __syncthreads();
  for (int e = t; e < NB16 * 3; e += 64 * NW) {
    const int r = e / 3, d = e % 3, ip = p0 + r;
    if (ip >= npro) continue;
    double s = 0.0;
#pragma unroll
    for (int w = 0; w < NW; ++w) s += sF[w][r][d];
    part[((size_t)xt.shared * npro + ip) * 3 + d] = s;
  }
}

// ADDOPROJ: proj[k][j] += sum_i oproj[k][i] L1[i][j] + conj(oproj[k][i]) L2[i][j]
// (L1, L2) pairs of the whole NBH x NBH fold, l12[2 (i nbh + j) + 0/1]
__global__ void k_fold_lam(int nb, int nbh, const double *lam, double2 *l12) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= nbh * nbh) return;
  double2 l1, l2;
  fold_lambda(lam, nb, idx / nbh, idx % nbh, l1, l2);
  l12[2 * idx] = l1;
  l12[2 * idx + 1] = l2;
}

// A workgroup takes 64 consecutive elements; its four waves sum quarters of i,
// which are then added in the order of the quarters (deterministic).
__global__ void __launch_bounds__(256) k_addoproj(int npro, int nbh, double2 *proj, const double2 *oproj,
                                                  const double2 *l12) {
  __shared__ double2 part[4][64];
  const int lane = threadIdx.x & 63, q = threadIdx.x >> 6;
  const int idx = blockIdx.x * 64 + lane;
  const bool ok = idx < npro * nbh;
  double2 s = czero();
  if (ok) {
    const int k = idx / nbh, j = idx % nbh;
    const double2 *o = oproj + (size_t)k * nbh;
    const int i1 = (q + 1) * nbh / 4;
#pragma unroll 4
    for (int i = q * nbh / 4; i < i1; ++i) {
      const double2 l1 = l12[2 * ((size_t)i * nbh + j)], l2 = l12[2 * ((size_t)i * nbh + j) + 1];
      const double2 oi = o[i];
      s = cadd(s, cadd(cmul(oi, l1), cmul(cconj(oi), l2)));
    }
  }
  part[q][lane] = s;
  __syncthreads();
  if (q == 0 && ok) {
    double2 t = cadd(cadd(part[0][lane], part[1][lane]), cadd(part[2][lane], part[3][lane]));
    proj[idx] = cadd(proj[idx], t);
  }
}

// dE/dproj fold (WAVES_DEDPROJ, src/paw_waves1.f90:5651-5666): d = F1 x + F2 conj(x)
__global__ void k_fold_f(int npro, int nbh, const double *occ2, double2 *x) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= npro * nbh) return;
  const int b = idx % nbh;
  const double f1 = occ2[2 * b], f2 = occ2[2 * b + 1];
  const double2 v = x[idx];
  x[idx] = make_double2((f1 + f2) * v.x, (f1 - f2) * v.y);
}

// ---------------------------------------------------------------------------
// reductions over (band, G): EKIN, WAVEKINETIC, <psi|H|psi>
// ---------------------------------------------------------------------------
template <int NT>
__device__ inline double block_sum(double v, double *red) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  const int w = threadIdx.x / 64, l = threadIdx.x % 64;
  if (l == 0) red[w] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int k = 0; k < NT / 64; ++k) s += red[k];
  __syncthreads();
  return s;
}

// The (G, band) reductions run on a grid of G blocks (x) by groups of
// kBandGrp bands (y), one partial sum per workgroup.
constexpr int kBandGrp = 8;

// wave kinetic partial: sum_G m(G) sum_b [fa |d|^2 + fb (dr dr(-G) - di di(-G))]
__global__ void __launch_bounds__(kB) k_wavekin(int ngl, int nbh, const double *mpsi,
                                                const int *minusg, const double *fab,
                                                const double2 *p0, const double2 *pm,
                                                double *part) {
  __shared__ double red[kB / 64];
  double s = 0.0;
  const int b0 = blockIdx.y * kBandGrp, b1 = min(nbh, b0 + kBandGrp);
  for (int ig = blockIdx.x * kB + threadIdx.x; ig < ngl; ig += gridDim.x * kB) {
    double d = 0.0;
    const int mg = minusg[ig];
    for (int b = b0; b < b1; ++b) {
      const double fa = fab[2 * b], fb = fab[2 * b + 1];
      if (fa == 0.0 && fb == 0.0) continue;
      const double2 x = csub(p0[(size_t)b * ngl + ig], pm[(size_t)b * ngl + ig]);
      d += fa * (x.x * x.x + x.y * x.y);
      if (fb != 0.0) {
        const double2 y = csub(p0[(size_t)b * ngl + mg], pm[(size_t)b * ngl + mg]);
        // TPSIM = conj(TPSIP(-G)): Re*Re + Im*Im = x.x*y.x - x.y*y.y
        d += fb * (x.x * y.x - x.y * y.y);
      }
    }
    s += mpsi[ig] * d;
  }
  s = block_sum<kB>(s, red);
  if (threadIdx.x == 0) part[blockIdx.y * gridDim.x + blockIdx.x] = s;
}

// The three reductions of WAVES$ETOT over psi0 in one pass, so that psi0 is
// read once: grid (G blocks, band groups of kBandGrp) as k_wavekin, whose
// per-block partial layout it keeps for the kinetic energy
//   ekin: sum_G g2 * sum_b [fp |c|^2 + fm Re(c(G) c(-G))]
// and the wave kinetic energy (the first half, as k_wavekin), plus the
// <psi|H psi> pair of each band of the group
//   expect[2*(blockIdx.x*nbh + b)] = (sum conj(psi) hpsi, sum psi hpsi(-G)).
// Per G the band sums of ekin and the wave kinetic energy run in band order.
__global__ void __launch_bounds__(kB) k_etot_sums(int ngl, int nbh, const double *g2, const double *mpsi,
                                                  const int *minusg, const double *fpm, const double *fab,
                                                  const double2 *p0, const double2 *pm, const double2 *hp,
                                                  double *ekin, double *wk, double *expect) {
  constexpr int kV = 2 + 2 * kBandGrp;
  __shared__ double red[kV][kB / 64];
  const int b0 = blockIdx.y * kBandGrp, nbg = min(kBandGrp, nbh - b0);
  double se = 0.0, sw = 0.0, tr[kBandGrp], tmr[kBandGrp];
#pragma unroll
  for (int k = 0; k < kBandGrp; ++k) tr[k] = tmr[k] = 0.0;
  for (int ig = blockIdx.x * kB + threadIdx.x; ig < ngl; ig += gridDim.x * kB) {
    double d = 0.0, e = 0.0;
    const int mg = minusg[ig];
#pragma unroll
    for (int k = 0; k < kBandGrp; ++k) {
      if (k >= nbg) break;
      const int b = b0 + k;
      const size_t o = (size_t)b * ngl;
      const double2 c = p0[o + ig], h = hp[o + ig], hm = hp[o + mg];
      tr[k] += c.x * h.x + c.y * h.y;
      tmr[k] += c.x * hm.x - c.y * hm.y;
      const double fp = fpm[2 * b], fm = fpm[2 * b + 1];
      if (fp != 0.0) d += fp * (c.x * c.x + c.y * c.y);
      if (fm != 0.0) {
        const double2 cm = p0[o + mg];
        d += fm * (c.x * cm.x - c.y * cm.y);
      }
      const double fa = fab[2 * b], fb = fab[2 * b + 1];
      if (fa == 0.0 && fb == 0.0) continue;
      const double2 x = csub(c, pm[o + ig]);
      e += fa * (x.x * x.x + x.y * x.y);
      if (fb != 0.0) {
        const double2 y = csub(p0[o + mg], pm[o + mg]);
        e += fb * (x.x * y.x - x.y * y.y);
      }
    }
    se += g2[ig] * d;
    sw += mpsi[ig] * e;
  }
  double v[kV];
  v[0] = se;
  v[1] = sw;
#pragma unroll
  for (int k = 0; k < kBandGrp; ++k) {
    v[2 + 2 * k] = tr[k];
    v[3 + 2 * k] = tmr[k];
  }
  const int w = threadIdx.x / 64, l = threadIdx.x % 64;
#pragma unroll
  for (int k = 0; k < kV; ++k) {
    double t = v[k];
    for (int off = 32; off > 0; off >>= 1) t += __shfl_down(t, off, 64);
    if (l == 0) red[k][w] = t;
  }
  __syncthreads();
  if (threadIdx.x < kV) {
    const int k = threadIdx.x;
    double t = 0.0;
    for (int q = 0; q < kB / 64; ++q) t += red[k][q];
    if (k == 0) ekin[blockIdx.y * gridDim.x + blockIdx.x] = t;
    else if (k == 1) wk[blockIdx.y * gridDim.x + blockIdx.x] = t;
    else if ((k - 2) / 2 < nbg) expect[2 * ((size_t)blockIdx.x * nbh + b0 + (k - 2) / 2) + (k & 1)] = t;
  }
}

// one-centre density matrix (real part, before hermitisation) per atom.
// Grid (NAT, entry groups of blockDim.x): entry w = i*LMNXX + j of the atom's
// LMNXX x LMNXX block.  The atom's projections PROJ(B, BASE+I) are staged in
// LDS by chunks of BC bands (row pitch BC+1 double2, so the rows j of one
// wave fall in different banks) with the occupations beside them; each entry
// sums over the bands in order 0..NBH-1.  Dynamic LDS: (LMNXX*(BC+1)+BC)*16 B.
__global__ void __launch_bounds__(kB) k_denmat(ProjGeo q, int bc, const double *occ2, const double2 *proj,
                                               double *dm) {
  extern __shared__ double2 sp[];
  const int atom = blockIdx.x;
  const int lmnx = q.lmnx[q.atom_isp[atom]];
  const int base = q.ipro0[atom];
  const int L = q.lmnxx, pitch = bc + 1;
  double2 *so = sp + (size_t)L * pitch;
  const int w = blockIdx.y * blockDim.x + threadIdx.x;
  const int i = w / L, j = w % L;
  const bool on = w < L * L && i < lmnx && j < lmnx;
  double s = 0.0;
  for (int b0 = 0; b0 < q.nbh; b0 += bc) {
    const int nbc = min(bc, q.nbh - b0);
    __syncthreads();
    for (int t = threadIdx.x; t < lmnx * nbc; t += blockDim.x) {
      const int r = t / nbc, c = t - r * nbc;
      sp[r * pitch + c] = proj[(size_t)(base + r) * q.nbh + b0 + c];
    }
    for (int t = threadIdx.x; t < nbc; t += blockDim.x)
      so[t] = make_double2(occ2[2 * (b0 + t)], occ2[2 * (b0 + t) + 1]);
    __syncthreads();
    if (on) {
      const double2 *ri = sp + i * pitch, *rj = sp + j * pitch;
      for (int c = 0; c < nbc; ++c) {
        const double2 o = so[c], pi = ri[c], pj = rj[c];
        // Re(pi (s1 conj(pj) + s2 pj))
        s += o.x * (pi.x * pj.x + pi.y * pj.y) + o.y * (pi.x * pj.x - pi.y * pj.y);
      }
    }
  }
  if (w < L * L) dm[(size_t)atom * L * L + w] = s;
}

__global__ void __launch_bounds__(kB) k_dot(size_t n, const double *a, const double *b, double *part) {
  __shared__ double red[kB / 64];
  double s = 0.0;
  for (size_t i = (size_t)blockIdx.x * kB + threadIdx.x; i < n; i += (size_t)gridDim.x * kB) s += a[i] * b[i];
  s = block_sum<kB>(s, red);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// grid (NGL/256, NBH): G along x, the band along y
// Residual of a step: the constrained force on the wave functions,
//   F = HPSI - (1+a) m(G)/dt^2 (psi(+) - psibar) = HPSI - O psi(0) Lambda,
// from the Verlet update psibar (src/paw_waves2.f90:129-252) and the
// constraint term that ORTHOGONALIZE adds to it; F vanishes at the fixed
// point of the damped dynamics, psi(+) = psi(0) = psi(-).  part[b][blk] =
// sum over the block's G of |F(b,G)|^2 for super band b (the squares of its
// two real bands' residuals added).
__global__ void __launch_bounds__(kB) k_resid_part(int ngl, double cs, const double *mpsi, const double2 *hpsi,
                                                   const double2 *psip, const double2 *psibar, double *part) {
  __shared__ double red[kB / 64];
  const int b = blockIdx.y;
  double s = 0.0;
  for (int ig = blockIdx.x * kB + threadIdx.x; ig < ngl; ig += gridDim.x * kB) {
    const size_t i = (size_t)b * ngl + ig;
    const double c = cs * mpsi[ig];
    const double2 h = hpsi[i], pp = psip[i], pb = psibar[i];
    const double fx = h.x - c * (pp.x - pb.x), fy = h.y - c * (pp.y - pb.y);
    s += fx * fx + fy * fy;
  }
  s = block_sum<kB>(s, red);
  if (threadIdx.x == 0) part[(size_t)b * gridDim.x + blockIdx.x] = s;
}

// out = max over super bands b of sqrt(scale sum_blk part[b][blk]) (one workgroup)
__global__ void __launch_bounds__(kB) k_resid_max(int nbh, int nblk, double scale, const double *part, double *out) {
  __shared__ double red[kB / 64];
  double m = 0.0;
  for (int b = threadIdx.x; b < nbh; b += kB) {
    double s = 0.0;
    for (int k = 0; k < nblk; ++k) s += part[(size_t)b * nblk + k];
    m = fmax(m, sqrt(scale * s));
  }
  for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_down(m, off, 64));
  if (threadIdx.x % 64 == 0) red[threadIdx.x / 64] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double r = 0.0;
    for (int k = 0; k < kB / 64; ++k) r = fmax(r, red[k]);
    *out = r;
  }
}

__global__ void k_propagate_g(int ngl, int nbh, double a1, double a2, double a3n,
                              const double *mpsi, const double2 *psi0, double2 *psim,
                              const double2 *hpsi) {
  const int ig = blockIdx.x * blockDim.x + threadIdx.x;
  if (ig >= ngl) return;
  const size_t idx = (size_t)blockIdx.y * ngl + ig;
  const double a3 = a3n / mpsi[ig];
  const double2 p0 = psi0[idx], pm = psim[idx], h = hpsi[idx];
  psim[idx] = make_double2(a1 * p0.x + a2 * pm.x + a3 * h.x, a1 * p0.y + a2 * pm.y + a3 * h.y);
}

ProjGeo make_pgeo(pw_waveset *ws) {
  ProjGeo q;
  pw_plan *p = ws->plan;
  q.ngl = p->ngl; q.nbh = ws->nbh; q.nat = ws->nat; q.lmnxx = ws->lmnxx;
  q.nphase = ws->nphase; q.nmaxx = (ws->nphase - 1) / 2;
  q.igvec = p->d_igvec; q.pro = ws->d_pro; q.ylm = ws->d_ylm; q.phase = ws->d_phase;
  q.atom_isp = ws->d_atom_isp; q.ipro0 = ws->d_ipro0; q.ibpro = ws->d_ibpro; q.lmnx = ws->d_lmnx;
  q.lmn_ln = ws->d_lmn_ln; q.lmn_lm = ws->d_lmn_lm; q.lmn_l = ws->d_lmn_l;
  return q;
}

// blocks of 16 bands per workgroup (at most 8) and the number of band tiles
void band_tiles(int nbh, int *nblk, int *ntiles) {
  const int total = (nbh + 15) / 16;
  *ntiles = (total + 7) / 8;
  *nblk = (total + *ntiles - 1) / *ntiles;
}

#define PW_NB_SWITCH(nblk, LAUNCH)                                      \
  switch (nblk) {                                                       \
    case 1: LAUNCH(1); break;                                           \
    case 2: LAUNCH(2); break;                                           \
    case 3: LAUNCH(3); break;                                           \
    case 4: LAUNCH(4); break;                                           \
    case 5: LAUNCH(5); break;                                           \
    case 6: LAUNCH(6); break;                                           \
    case 7: LAUNCH(7); break;                                           \
    default: LAUNCH(8); break;                                          \
  }

}  // namespace

int ws_part(pw_waveset *ws, size_t bytes);
int ws_split(pw_waveset *ws, size_t bytes);

namespace {

// bytes of the split runs of an nrow x K operand
size_t runs_bytes(int nrow, int K) { return (size_t)nrow * ((K + KB - 1) / KB * (KB / 8)) * NPL * sizeof(u32x4); }

// split an nrow x K operand into runs at dst (nrow x K runs, runs_bytes)
template <bool RUN_MAJOR, class Src>
SplitRuns split_runs_to(pw_waveset *ws, const Src &src, int nrow, int K, void *dst) {
  const int nrun = (K + KB - 1) / KB * (KB / 8);
  const size_t n = (size_t)nrow * nrun;
  u32x4 *sp = static_cast<u32x4 *>(dst);
  if constexpr (RUN_MAJOR)
    k_split_cols<Src><<<(unsigned)((n + 255) / 256), 256, 0, ws->stream>>>(src, nrow, nrun, sp);
  else
    k_split_rows<Src><<<(unsigned)((n + 31) / 32), 256, 0, ws->stream>>>(src, nrow, nrun,
                                                                          reinterpret_cast<unsigned *>(sp));
  return RUN_MAJOR ? SplitRuns{sp, nrow, nrun, 1, nrow} : SplitRuns{sp, nrow, nrun, nrun, 1};
}
// ... at byte offset off of the scratch split buffer (sized with ws_split)
template <bool RUN_MAJOR, class Src>
SplitRuns split_runs(pw_waveset *ws, const Src &src, int nrow, int K, size_t off) {
  return split_runs_to<RUN_MAJOR>(ws, src, nrow, K, static_cast<char *>(ws->d_split) + off);
}

// runs of the rows of a wave-function array (minus: of psi(-G)).  Inside a
// runs scope they are taken from, or split into, one of four kept slots;
// otherwise they are split into the scratch buffer at byte offset off.
int wave_runs(pw_waveset *ws, const double2 *psi, bool minus, size_t off, SplitRuns *out) {
  pw_plan *p = ws->plan;
  const int nbh = ws->nbh, ngl = p->ngl;
  const size_t bytes = runs_bytes(nbh, ngl);
  auto split_to = [&](void *dst) {
    if (minus) *out = split_runs_to<false>(ws, SrcPsiMinusRows{psi, p->d_minusg, ngl, nbh}, nbh, ngl, dst);
    else *out = split_runs_to<false>(ws, SrcPsiRows{psi, ngl, nbh}, nbh, ngl, dst);
  };
  if (ws->runs_scope) {
    int free_slot = -1;
    for (int i = 0; i < 4; ++i) {
      const pw_waveset::WaveRuns &w = ws->wruns[i];
      if (w.valid && w.psi == psi && w.minus == minus) {
        const int nrun = (ngl + KB - 1) / KB * (KB / 8);
        *out = SplitRuns{reinterpret_cast<const u32x4 *>(static_cast<char *>(ws->d_wruns) + i * bytes), nbh, nrun,
                         nrun, 1};
        return PW_OK;
      }
      if (!w.valid && free_slot < 0) free_slot = i;
    }
    if (free_slot >= 0) {
      if (4 * bytes > ws->wruns_bytes) {   // only at the scope's first split (all slots free)
        if (ws->d_wruns) PW_HIP(hipFree(ws->d_wruns));
        PW_HIP(hipMalloc(&ws->d_wruns, 4 * bytes));
        ws->wruns_bytes = 4 * bytes;
      }
      ScopeTimer tm(ws, "split");
      split_to(static_cast<char *>(ws->d_wruns) + free_slot * bytes);
      ws->wruns[free_slot] = pw_waveset::WaveRuns{psi, minus, true};
      return PW_OK;
    }
  }
  PW_TRY(ws_split(ws, off + bytes));
  ScopeTimer tm(ws, "split");
  split_to(static_cast<char *>(ws->d_split) + off);
  return PW_OK;
}

// out[r][b] = scale sum_G A(r,G) psi[b][G]: split over G (about two waves per
// SIMD), deterministic reduction of the splits.  out == nullptr puts the
// result behind the partials and returns it in *res.
// G-split partials left for the caller's own reduction (defer != nullptr)
struct RowPart {
  const double2 *part = nullptr;
  int ns = 0;
  size_t n = 0;
  double scale = 1.0;
};

template <int NW, class Gen>
int run_rowgemm(pw_waveset *ws, const Gen &gen, int nrow, const double2 *psi, double scale,
                double2 *out, double2 **res, const char *name, RowPart *defer = nullptr) {
  const int ngl = ws->plan->ngl, nbh = ws->nbh;
  int nblk, tb;
  band_tiles(nbh, &nblk, &tb);
  const int tr = (nrow + 16 * NW - 1) / (16 * NW);
  const int want = std::max(1, ws->tune_split_waves / NW / std::max(1, tr * tb));
  int ns = std::min(want, (ngl + KC - 1) / KC);
  const int gsplit = ((ngl + ns - 1) / ns + KC - 1) / KC * KC;
  ns = (ngl + gsplit - 1) / gsplit;
  const size_t n = (size_t)nrow * nbh;
  PW_TRY(ws_part(ws, (size_t)(ns + (out ? 0 : 1)) * n * sizeof(double2)));
  double2 *part = (double2 *)ws->d_part;
  if (!out) out = part + (size_t)ns * n;
  {
    ScopeTimer tm(ws, name);
    const unsigned grid = xcd_blocks(tr * tb, ns);
#define PW_LAUNCH(NB) \
  k_rowgemm<Gen, NB, NW><<<grid, 64 * NW, 0, ws->stream>>>(gen, nrow, ngl, nbh, psi, gsplit, tr, tb, ns, part)
    PW_NB_SWITCH(nblk, PW_LAUNCH)
#undef PW_LAUNCH
    PW_HIP(hipGetLastError());
  }
  if (defer) {
    *defer = RowPart{part, ns, n, scale};
    return PW_OK;
  }
  {
    ScopeTimer tm(ws, "reduce");
    k_reduce_split<double2><<<(n + kRedC - 1) / kRedC, kRedC * kRedS, 0, ws->stream>>>(part, ns, (int)n, scale,
                                                                                         out);
    PW_HIP(hipGetLastError());
  }
  if (res) *res = out;
  return PW_OK;
}

// psi[b][G] += sum_k A(b,k) B(k,G)
template <int NW, class GA, class GB>
int run_colgemm(pw_waveset *ws, const GA &ga, const GB &gb, int nk, const ColOut &out, const char *name) {
  const int ngl = ws->plan->ngl, nbh = ws->nbh;
  int nblk, tb;
  band_tiles(nbh, &nblk, &tb);
  dim3 grid((ngl + 16 * NW - 1) / (16 * NW), tb);
  ScopeTimer tm(ws, name);
#define PW_LAUNCH(NB) k_colgemm<GA, GB, NB, NW><<<grid, 64 * NW, 0, ws->stream>>>(ga, gb, nk, ngl, nbh, out)
  PW_NB_SWITCH(nblk, PW_LAUNCH)
#undef PW_LAUNCH
  PW_HIP(hipGetLastError());
  return PW_OK;
}

// split-BF16x3 counterpart of run_rowgemm (G splits are whole 32-chunks)
template <int NW, class Gen>
int run_rowgemm_bf(pw_waveset *ws, const Gen &gen, int nrow, const BandRuns &gb, double scale,
                   double2 *out, double2 **res, const char *name, RowPart *defer = nullptr) {
  const int ngl = ws->plan->ngl, nbh = ws->nbh;
  int nblk, tb;
  band_tiles(nbh, &nblk, &tb);
  const int tr = (nrow + 16 * NW - 1) / (16 * NW);
  const int want = std::max(1, ws->tune_split_waves / NW / std::max(1, tr * tb));
  int ns = std::min(want, (ngl + KB - 1) / KB);
  const int gsplit = ((ngl + ns - 1) / ns + KB - 1) / KB * KB;
  ns = (ngl + gsplit - 1) / gsplit;
  const size_t n = (size_t)nrow * nbh;
  PW_TRY(ws_part(ws, (size_t)(ns + (out ? 0 : 1)) * n * sizeof(double2)));
  double2 *part = (double2 *)ws->d_part;
  if (!out) out = part + (size_t)ns * n;
  {
    ScopeTimer tm(ws, name);
    const unsigned grid = xcd_blocks(tr * tb, ns);
#define PW_LAUNCH(NB) \
  k_rowgemm_bf<Gen, BandRuns, NB, NW><<<grid, 64 * NW, 0, ws->stream>>>(gen, gb, nrow, ngl, nbh, gsplit, tr, tb, \
                                                                         ns, part)
    PW_NB_SWITCH(nblk, PW_LAUNCH)
#undef PW_LAUNCH
    PW_HIP(hipGetLastError());
  }
  if (defer) {
    *defer = RowPart{part, ns, n, scale};
    return PW_OK;
  }
  {
    ScopeTimer tm(ws, "reduce");
    k_reduce_split<double2><<<(n + kRedC - 1) / kRedC, kRedC * kRedS, 0, ws->stream>>>(part, ns, (int)n, scale,
                                                                                         out);
    PW_HIP(hipGetLastError());
  }
  if (res) *res = out;
  return PW_OK;
}

template <int NW, class GA, class GB>
int run_colgemm_bf(pw_waveset *ws, const GA &ga, const GB &gb, int nk, const ColOut &out, const char *name) {
  const int ngl = ws->plan->ngl, nbh = ws->nbh;
  int nblk, tb;
  band_tiles(nbh, &nblk, &tb);
  dim3 grid((ngl + 16 * NW - 1) / (16 * NW), tb);
  ScopeTimer tm(ws, name);
#define PW_LAUNCH(NB) k_colgemm_bf<GA, GB, NB, NW><<<grid, 64 * NW, 0, ws->stream>>>(ga, gb, nk, ngl, nbh, out)
  PW_NB_SWITCH(nblk, PW_LAUNCH)
#undef PW_LAUNCH
  PW_HIP(hipGetLastError());
  return PW_OK;
}

int nglp_of(const pw_waveset *ws) { return (ws->plan->ngl + 63) / 64 * 64; }

}  // namespace

// ---------------------------------------------------------------------------
// host launchers
// ---------------------------------------------------------------------------
int ws_part(pw_waveset *ws, size_t bytes) {
  if (bytes > ws->part_bytes) {
    if (ws->d_part) PW_HIP(hipFree(ws->d_part));
    PW_HIP(hipMalloc(&ws->d_part, bytes));
    ws->part_bytes = bytes;
  }
  return PW_OK;
}

int k_runs_scope(pw_waveset *ws, bool on) {
  ws->runs_scope = on;
  for (auto &w : ws->wruns) w.valid = false;
  return PW_OK;
}

int ws_split(pw_waveset *ws, size_t bytes) {
  if (bytes > ws->split_bytes) {
    if (ws->d_split) PW_HIP(hipFree(ws->d_split));
    PW_HIP(hipMalloc(&ws->d_split, bytes));
    ws->split_bytes = bytes;
  }
  return PW_OK;
}

int ws_small(pw_waveset *ws, size_t bytes) {
  if (bytes > ws->small_bytes) {
    if (ws->d_small) PW_HIP(hipFree(ws->d_small));
    PW_HIP(hipMalloc(&ws->d_small, bytes));
    ws->small_bytes = bytes;
  }
  return PW_OK;
}

int k_phase_tables(pw_waveset *ws, const double *r, double2 *phase) {
  // r: host positions (3*nat); gr[atom*3+k] = GBAS(:,k).R, handed to the
  // kernel by value in chunks of kPhaseAtoms atoms (no staging buffer)
  pw_plan *p = ws->plan;
  ScopeTimer tm(ws, "phase");
  for (int a0 = 0; a0 < ws->nat; a0 += kPhaseAtoms) {
    const int na = std::min(kPhaseAtoms, ws->nat - a0);
    PhaseArgs args;
    for (int a = 0; a < na; ++a)
      for (int k = 0; k < 3; ++k) {
        const double *ra = r + 3 * (a0 + a);
        args.gr[3 * a + k] = p->gbas[0 + 3 * k] * ra[0] + p->gbas[1 + 3 * k] * ra[1] + p->gbas[2 + 3 * k] * ra[2];
      }
    const int total = na * 3 * ws->nphase;
    k_phase<<<(total + 255) / 256, 256, 0, ws->stream>>>(na, ws->nphase, (ws->nphase - 1) / 2, args,
                                                          phase + (size_t)a0 * 3 * ws->nphase);
    PW_HIP(hipGetLastError());
  }
  return PW_OK;
}

// projector factors F and E for the positions of the current phase tables
int k_pfac(pw_waveset *ws) {
  const int nglp = nglp_of(ws);
  const size_t nf = (size_t)ws->nsp * ws->lmnxx * nglp, ne = (size_t)ws->nat * nglp;
  const size_t bytes = nf * (sizeof(double) + sizeof(float)) + ne * (sizeof(double2) + sizeof(float2));
  if (bytes > ws->pfac_bytes) {
    if (ws->d_pfac) PW_HIP(hipFree(ws->d_pfac));
    PW_HIP(hipMalloc(&ws->d_pfac, bytes));
    ws->pfac_bytes = bytes;
  }
  ProjGeo q = make_pgeo(ws);
  q.phase = ws->d_phase;
  double *f = (double *)ws->d_pfac;
  ScopeTimer tm(ws, "pgen");
  // [F FP64][E FP64][E FP32][F FP32]: every part starts 8-byte aligned
  double2 *e = (double2 *)(f + nf);
  float2 *e32 = (float2 *)(e + ne);
  float *f32 = (float *)(e32 + ne);
  k_pfac_f<<<(nf + 255) / 256, 256, 0, ws->stream>>>(q, ws->nsp, nglp, f, f32);
  k_pfac_e<<<(ne + 255) / 256, 256, 0, ws->stream>>>(q, nglp, e, e32);
  PW_HIP(hipGetLastError());
  return PW_OK;
}

PFac pfac_of(const pw_waveset *ws) {
  const int nglp = nglp_of(ws);
  const double *f = (const double *)ws->d_pfac;
  const size_t nf = (size_t)ws->nsp * ws->lmnxx * nglp, ne = (size_t)ws->nat * nglp;
  const double2 *e = (const double2 *)(f + nf);
  const float2 *e32 = (const float2 *)(e + ne);
  const float *f32 = (const float *)(e32 + ne);
  return PFac{f, e, f32, e32, ws->d_ipro_tab, ws->d_atom_isp, ws->d_lmn_l, ws->lmnxx, nglp};
}

int k_projections(pw_waveset *ws, const double2 *psi, const double2 *phase, double2 *proj) {
  if (ws->prec == PW_PREC_MIXED) {
    BandRuns gb;
    PW_TRY(wave_runs(ws, psi, false, 0, &gb.s));
    RowPf gen{pfac_of(ws)};
    return run_rowgemm_bf<4>(ws, gen, ws->npro, gb, ws->plan->vol, proj, nullptr, "projections");
  }
  ProjGeo q = make_pgeo(ws);
  q.phase = phase;
  GenProj gen{q, ws->d_ipro_tab};
  return run_rowgemm<4>(ws, gen, ws->npro, psi, ws->plan->vol, proj, nullptr, "projections");
}

// psi = (src or psi) + sum_i |p_i> x_i, then times num / m(G) when scaled
static int addpro_to(pw_waveset *ws, const ColOut &out, const double2 *phase, const double2 *x) {
  if (ws->prec == PW_PREC_MIXED) {
    PW_TRY(ws_split(ws, runs_bytes(ws->nbh, ws->npro)));
    ARuns ga;
    {
      ScopeTimer tm(ws, "split");
      ga = ARuns{split_runs<false>(ws, SrcXRows{x, ws->nbh, ws->npro}, ws->nbh, ws->npro, 0)};
    }
    AddproBf gb{pfac_of(ws), ws->npro};
    if (ws->tune_col_waves == 2) return run_colgemm_bf<2>(ws, ga, gb, ws->npro, out, "addpro");
    return run_colgemm_bf<4>(ws, ga, gb, ws->npro, out, "addpro");
  }
  ProjGeo q = make_pgeo(ws);
  q.phase = phase;
  AddproA ga{x, ws->nbh};
  AddproB gb{q, ws->d_ipro_tab};
  return run_colgemm<2>(ws, ga, gb, ws->npro, out, "addpro");
}

int k_addpro(pw_waveset *ws, double2 *psi, const double2 *phase, const double2 *x) {
  return addpro_to(ws, ColOut{psi, nullptr, nullptr, 0.0}, phase, x);
}

// WAVES_OPSI: opsi = dt^2/((1+a) m(G)) (psi0 + sum_i |p_i> x_i), x = DO <p|psi0>,
// in the ADDPRO pass (no copy of PSI0, no separate scaling pass)
int k_addpro_opsi(pw_waveset *ws, double2 *opsi, const double2 *psi0, const double2 *phase, const double2 *x) {
  return addpro_to(ws, ColOut{opsi, psi0, ws->d_mpsi, ws->delt * ws->delt / (1.0 + ws->annee)}, phase, x);
}

int k_force(pw_waveset *ws, const double2 *psi, const double2 *phase, const double2 *ded, double *dst) {
  constexpr int NW = 4, NGB = 4;
  ProjGeo q = make_pgeo(ws);
  q.phase = phase;
  const size_t n = (size_t)ws->npro * 3;
  if (ws->prec == PW_PREC_MIXED) {
    // the projector factors hold P at the positions of the last phase_at
    constexpr int NRB = 4;
    const int ngt = (q.ngl + 16 * NW - 1) / (16 * NW);
    PW_TRY(ws_part(ws, (size_t)ngt * n * sizeof(double)));
    const size_t ba = runs_bytes(ws->npro, ws->nbh), bb = runs_bytes(q.ngl, ws->nbh);
    PW_TRY(ws_split(ws, ba + bb));
    ARuns ga;
    ColRuns gb;
    {
      ScopeTimer tm(ws, "split");
      ga = ARuns{split_runs<false>(ws, SrcDedRows{ded, ws->npro, ws->nbh}, ws->npro, ws->nbh, 0)};
      gb = ColRuns{split_runs<true>(ws, SrcPsiCols{psi, q.ngl, ws->nbh}, q.ngl, ws->nbh, ba)};
    }
    {
      ScopeTimer tm(ws, "force");
      const int nrt = (ws->npro + 16 * NRB - 1) / (16 * NRB);
      k_force_bf<ARuns, ColRuns, NRB, NW><<<xcd_blocks(nrt, ngt), 64 * NW, 0, ws->stream>>>(
          ga, gb, ws->npro, ws->nbh, q.ngl, pfac_of(ws), ws->plan->d_gvec, nrt, ngt, ws->d_part);
      PW_HIP(hipGetLastError());
    }
    k_reduce_split<double><<<(n + kRedC - 1) / kRedC, kRedC * kRedS, 0, ws->stream>>>(ws->d_part, ngt, (int)n,
                                                                                        2.0 * ws->plan->vol, dst);
    PW_HIP(hipGetLastError());
    return PW_OK;
  }
  const int ngt = (q.ngl + 16 * NGB - 1) / (16 * NGB);
  PW_TRY(ws_part(ws, (size_t)ngt * n * sizeof(double)));
  {
    ScopeTimer tm(ws, "force");
    const int nrt = (ws->npro + 16 * NW - 1) / (16 * NW);
    k_force_m<NGB, NW><<<xcd_blocks(nrt, ngt), 64 * NW, 0, ws->stream>>>(q, ws->d_ipro_tab, ws->npro, psi, ded,
                                                                         ws->plan->d_gvec, nrt, ngt, ws->d_part);
    PW_HIP(hipGetLastError());
  }
  k_reduce_split<double><<<(n + kRedC - 1) / kRedC, kRedC * kRedS, 0, ws->stream>>>(ws->d_part, ngt, (int)n,
                                                                                      2.0 * ws->plan->vol, dst);
  PW_HIP(hipGetLastError());
  return PW_OK;
}

// mat_dev (row-major NB x NB) = WAVES_OVERLAP(psi1, psi2)
// T = Omega sum_G conj(psi1_i) psi2_j and TM = Omega sum_G psi1_i psi2_j(-G)
//   = Omega sum_G psi1_i(-G) psi2_j(G), one GEMM over stacked rows
int k_overlap(pw_waveset *ws, const double2 *psi1, const double2 *psi2, double *mat, bool, bool) {
  pw_plan *p = ws->plan;
  const int nbh = ws->nbh, nbhp = (nbh + 15) / 16 * 16;
  double2 *red = nullptr;
  if (ws->prec == PW_PREC_MIXED) {
    const size_t bytes = runs_bytes(nbh, p->ngl);
    PW_TRY(ws_split(ws, 3 * bytes));   // outside a runs scope all three go to the scratch
    OvlRuns gen;
    gen.nbhp = nbhp;
    BandRuns gb;
    PW_TRY(wave_runs(ws, psi1, false, 0, &gen.rows));
    PW_TRY(wave_runs(ws, psi1, true, bytes, &gen.minus));
    if (psi2 == psi1) gb.s = gen.rows;
    else PW_TRY(wave_runs(ws, psi2, false, 2 * bytes, &gb.s));
    PW_TRY(run_rowgemm_bf<2>(ws, gen, 2 * nbhp, gb, p->vol, nullptr, &red, "overlap"));
  } else {
    GenOvl gen{psi1, p->d_minusg, p->ngl, nbh, nbhp};
    PW_TRY(run_rowgemm<2>(ws, gen, 2 * nbhp, psi2, p->vol, nullptr, &red, "overlap"));
  }
  k_unravel_pw<<<(nbh * nbh + 255) / 256, 256, 0, ws->stream>>>(nbh, ws->nb, red, red + (size_t)nbhp * nbh,
                                                                  mat);
  PW_HIP(hipGetLastError());
  return PW_OK;
}

// mat_dev = WAVES_OVERLAP(psi1, psi2) + WAVES_1COVERLAP(proj1, proj2): the
// one-centre partials first (DO*proj2 in d_xproj, the partials in the small
// scratch), then the G-split product with its reduction deferred, then one
// k_overlap_finish pass (four launches)
int k_overlap_full(pw_waveset *ws, const double2 *psi1, const double2 *psi2, const double2 *proj1,
                   const double2 *proj2, double *mat) {
  pw_plan *p = ws->plan;
  ProjGeo q = make_pgeo(ws);
  const int nbh = ws->nbh, nbhp = (nbh + 15) / 16 * 16, nz = (ws->npro + k1c - 1) / k1c;
  PW_TRY(ws_small(ws, (size_t)nz * nbh * nbh * sizeof(double4)));
  double4 *p1c = (double4 *)ws->d_small;
  {
    ScopeTimer tm(ws, "overlap1c");
    k_atom_mat<<<dim3(ws->nat, (ws->lmnxx * nbh + 63) / 64), 64, 0, ws->stream>>>(q, proj2, ws->d_domat,
                                                                                   ws->d_xproj);
    dim3 grid((nbh + 15) / 16, (nbh + 15) / 16, nz);
    k_overlap1c<<<grid, 256, 0, ws->stream>>>(ws->npro, nbh, proj1, ws->d_xproj, p1c);
    PW_HIP(hipGetLastError());
  }
  RowPart rp;
  if (ws->prec == PW_PREC_MIXED) {
    const size_t bytes = runs_bytes(nbh, p->ngl);
    PW_TRY(ws_split(ws, 3 * bytes));   // outside a runs scope all three go to the scratch
    OvlRuns gen;
    gen.nbhp = nbhp;
    BandRuns gb;
    PW_TRY(wave_runs(ws, psi1, false, 0, &gen.rows));
    PW_TRY(wave_runs(ws, psi1, true, bytes, &gen.minus));
    if (psi2 == psi1) gb.s = gen.rows;
    else PW_TRY(wave_runs(ws, psi2, false, 2 * bytes, &gb.s));
    PW_TRY(run_rowgemm_bf<2>(ws, gen, 2 * nbhp, gb, p->vol, nullptr, nullptr, "overlap", &rp));
  } else {
    GenOvl gen{psi1, p->d_minusg, p->ngl, nbh, nbhp};
    PW_TRY(run_rowgemm<2>(ws, gen, 2 * nbhp, psi2, p->vol, nullptr, nullptr, "overlap", &rp));
  }
  {
    ScopeTimer tm(ws, "reduce");
    k_overlap_finish<<<(nbh * nbh + kRedC - 1) / kRedC, kRedC * kRedS, 0, ws->stream>>>(
        rp.part, rp.ns, rp.n, rp.scale, nbh, nbhp, ws->nb, p1c, nz, mat);
    PW_HIP(hipGetLastError());
  }
  return PW_OK;
}

int k_atom_matrix(pw_waveset *ws, const double2 *proj, double2 *out, const double2 *mats) {
  ProjGeo q = make_pgeo(ws);
  ScopeTimer tm(ws, "atom_mat");
  k_atom_mat<<<dim3(ws->nat, (ws->lmnxx * ws->nbh + 63) / 64), 64, 0, ws->stream>>>(q, proj, mats, out);
  PW_HIP(hipGetLastError());
  return PW_OK;
}

int k_addopsi(pw_waveset *ws, double2 *psibar, const double2 *chi, const double *lam) {
  pw_plan *p = ws->plan;
  if (ws->prec == PW_PREC_MIXED) {
    const int nk = 2 * ws->nbh;
    const size_t ba = runs_bytes(ws->nbh, nk), bb = runs_bytes(p->ngl, nk);
    PW_TRY(ws_split(ws, ba + bb));
    ARuns ga;
    ColRuns gb;
    {
      ScopeTimer tm(ws, "split");
      ga = ARuns{split_runs<false>(ws, SrcLamRows{lam, ws->nb, ws->nbh}, ws->nbh, nk, 0)};
      gb = ColRuns{split_runs<true>(ws, SrcChiCols{chi, p->d_minusg, p->ngl, ws->nbh}, p->ngl, nk, ba)};
    }
    if (ws->tune_col_waves == 2) return run_colgemm_bf<2>(ws, ga, gb, nk, ColOut{psibar}, "addopsi");
    return run_colgemm_bf<4>(ws, ga, gb, nk, ColOut{psibar}, "addopsi");
  }
  AddopsiA ga{lam, ws->nb, ws->nbh};
  AddopsiB gb{chi, p->d_minusg, p->ngl, ws->nbh};
  return run_colgemm<2>(ws, ga, gb, 2 * ws->nbh, ColOut{psibar}, "addopsi");
}

int k_addoproj(pw_waveset *ws, double2 *proj, const double2 *oproj, const double *lam) {
  const int n = ws->npro * ws->nbh;
  const int nbh = ws->nbh;
  PW_TRY(ws_small(ws, (size_t)2 * nbh * nbh * sizeof(double2)));
  ScopeTimer tm(ws, "addoproj");
  k_fold_lam<<<(nbh * nbh + 255) / 256, 256, 0, ws->stream>>>(ws->nb, nbh, lam, (double2 *)ws->d_small);
  k_addoproj<<<(n + 63) / 64, 256, 0, ws->stream>>>(ws->npro, nbh, proj, oproj, (const double2 *)ws->d_small);
  PW_HIP(hipGetLastError());
  return PW_OK;
}

int k_fold_occ(pw_waveset *ws, double2 *x) {
  const int n = ws->npro * ws->nbh;
  k_fold_f<<<(n + 255) / 256, 256, 0, ws->stream>>>(ws->npro, ws->nbh, ws->d_occ2, x);
  PW_HIP(hipGetLastError());
  return PW_OK;
}

static int reduce_gblocks(pw_plan *p) { return std::min(1024, (p->ngl + kB - 1) / kB); }
static int band_groups(int nbh) { return (nbh + kBandGrp - 1) / kBandGrp; }

// Reductions of the step leave their partial sums on the device (dst); the
// host reads them back with one copy per step section (see pw_waves.cpp).
int red_blocks(const pw_waveset *ws) { return reduce_gblocks(ws->plan) * band_groups(ws->nbh); }
// G chunks of the <psi|H|psi> partials: the G blocks of k_etot_sums
int expect_chunks(const pw_waveset *ws) { return reduce_gblocks(ws->plan); }
int vrho_blocks(const pw_waveset *ws) { return std::min(1024, (ws->plan->nrl + kB - 1) / kB); }

int k_wavekinetic(pw_waveset *ws, const double2 *psi0, const double2 *psim, const double *fab_dev,
                  double *dst) {
  pw_plan *p = ws->plan;
  ScopeTimer tm(ws, "wavekinetic");
  const dim3 grid(reduce_gblocks(p), band_groups(ws->nbh));
  k_wavekin<<<grid, kB, 0, ws->stream>>>(p->ngl, ws->nbh, ws->d_mpsi, p->d_minusg, fab_dev, psi0, psim, dst);
  PW_HIP(hipGetLastError());
  return PW_OK;
}

int k_etot_sums(pw_waveset *ws, double *ekin, double *wk, double *expect) {
  pw_plan *p = ws->plan;
  ScopeTimer tm(ws, "etot_sums");
  const dim3 grid(reduce_gblocks(p), band_groups(ws->nbh));
  k_etot_sums<<<grid, kB, 0, ws->stream>>>(p->ngl, ws->nbh, p->d_g2, ws->d_mpsi, p->d_minusg, ws->d_fpm, ws->d_fab,
                                           ws->d_psi0, ws->d_psim, ws->d_hpsi, ekin, wk, expect);
  PW_HIP(hipGetLastError());
  return PW_OK;
}

int k_denmat(pw_waveset *ws, const double2 *proj, double *dst) {
  ProjGeo q = make_pgeo(ws);
  ScopeTimer tm(ws, "denmat");
  // band chunk: the staged rows and occupations within 64 KB of LDS
  int bc = 64;
  while (bc > 1 && ((size_t)q.lmnxx * (bc + 1) + bc) * sizeof(double2) > 65536) bc /= 2;
  const size_t lds = ((size_t)q.lmnxx * (bc + 1) + bc) * sizeof(double2);
  if (lds > 65536) return pw_fail(PW_ERR_ARG, "WAVES$DENMAT: LMNXX too large");
  const dim3 grid(ws->nat, (q.lmnxx * q.lmnxx + kB - 1) / kB);
  k_denmat<<<grid, kB, lds, ws->stream>>>(q, bc, ws->d_occ2, proj, dst);
  PW_HIP(hipGetLastError());
  return PW_OK;
}

int k_dot_vrho(pw_waveset *ws, double *dst) {
  pw_plan *p = ws->plan;
  ScopeTimer tm(ws, "vrho");
  k_dot<<<vrho_blocks(ws), kB, 0, ws->stream>>>((size_t)p->nrl, ws->d_v, ws->d_rho, dst);
  PW_HIP(hipGetLastError());
  return PW_OK;
}

int k_propagate(pw_waveset *ws) {
  pw_plan *p = ws->plan;
  const double a1 = 2.0 / (1.0 + ws->annee), a2 = 1.0 - a1;
  const double a3n = -ws->delt * ws->delt / (1.0 + ws->annee);
  ScopeTimer tm(ws, "propagate");
  k_propagate_g<<<dim3((p->ngl + 255) / 256, ws->nbh), 256, 0, ws->stream>>>(p->ngl, ws->nbh, a1, a2, a3n, ws->d_mpsi,
                                                         ws->d_psi0, ws->d_psim, ws->d_hpsi);
  PW_HIP(hipGetLastError());
  return PW_OK;
}

int k_residual(pw_waveset *ws, const double2 *psibar, double *dst) {
  pw_plan *p = ws->plan;
  const int nblk = reduce_gblocks(p);
  PW_TRY(ws_part(ws, (size_t)ws->nbh * nblk * sizeof(double)));
  ScopeTimer tm(ws, "residual");
  const double cs = (1.0 + ws->annee) / (ws->delt * ws->delt);
  k_resid_part<<<dim3(nblk, ws->nbh), kB, 0, ws->stream>>>(p->ngl, cs, ws->d_mpsi, ws->d_hpsi, ws->d_psim, psibar,
                                                           ws->d_part);
  k_resid_max<<<1, kB, 0, ws->stream>>>(ws->nbh, nblk, p->vol, ws->d_part, dst);
  PW_HIP(hipGetLastError());
  return PW_OK;
}

int k_copy(pw_waveset *ws, double2 *dst, const double2 *src, size_t n) {
  PW_HIP(hipMemcpyAsync(dst, src, n * sizeof(double2), hipMemcpyDeviceToDevice, ws->stream));
  return PW_OK;
}
