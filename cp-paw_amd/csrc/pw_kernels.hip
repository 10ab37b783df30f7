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

__device__ inline double2 eigr_of(const ProjGeo &q, int atom, int ig) {
  const int *iv = q.igvec + 3 * ig;
  const double2 *ph = q.phase + (size_t)atom * 3 * q.nphase;
  double2 e = cmul(ph[iv[0] + q.nmaxx], ph[q.nphase + iv[1] + q.nmaxx]);
  return cmul(e, ph[2 * q.nphase + iv[2] + q.nmaxx]);
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
__global__ void k_phase(int nat, int nphase, int nmaxx, const double *gr, double2 *phase) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int total = nat * 3 * nphase;
  if (i >= total) return;
  const int n = i % nphase - nmaxx;
  const int ak = i / nphase;                 // atom*3 + axis
  double s, c;
  sincos((double)n * gr[ak], &s, &c);
  phase[i] = make_double2(c, -s);
}

// ---------------------------------------------------------------------------
// projections: part[split][ipro][b] = sum_G conj(P) psi over the split's G range
// tile: 64 projector rows x 32 bands, K chunk 32
// ---------------------------------------------------------------------------
constexpr int PR = 64, PBT = 32, PKC = 32;

__global__ void __launch_bounds__(kB) k_proj(ProjGeo q, const double2 *psi, int gsplit, int npro,
                                             double2 *part) {
  __shared__ double2 sP[PKC][PR];
  __shared__ double2 sS[PKC][PBT];
  __shared__ double2 sE[8][PKC];
  const int apt = min(8, PR / q.lmnxx);             // atoms per tile
  const int atom0 = blockIdx.x * apt;
  const int b0 = blockIdx.y * PBT;
  const int split = blockIdx.z;
  const int g0 = split * gsplit, g1 = min(q.ngl, g0 + gsplit);
  const int t = threadIdx.x;
  const int tr = t / 16, tb = t % 16;
  double2 acc[4][2];
#pragma unroll
  for (int i = 0; i < 4; ++i) { acc[i][0] = czero(); acc[i][1] = czero(); }
  for (int gc = g0; gc < g1; gc += PKC) {
    // structure factors of the tile's atoms
    for (int w = t; w < apt * PKC; w += kB) {
      const int a = w / PKC, k = w % PKC, ig = gc + k, atom = atom0 + a;
      sE[a][k] = (atom < q.nat && ig < g1) ? eigr_of(q, atom, ig) : czero();
    }
    for (int w = t; w < PKC * PBT; w += kB) {
      const int b = w / PKC, k = w % PKC, ig = gc + k;
      sS[k][b] = (b0 + b < q.nbh && ig < g1) ? psi[(size_t)(b0 + b) * q.ngl + ig] : czero();
    }
    __syncthreads();
    for (int w = t; w < PKC * PR; w += kB) {
      const int k = w % PKC, r = w / PKC, ig = gc + k;
      const int a = r / q.lmnxx, lmn = r % q.lmnxx, atom = atom0 + a;
      double2 v = czero();
      if (a < apt && atom < q.nat && ig < g1) v = cconj(pro_val(q, atom, lmn, ig, sE[a][k]));
      sP[k][r] = v;
    }
    __syncthreads();
#pragma unroll 4
    for (int k = 0; k < PKC; ++k) {
      double2 a[4], b[2];
#pragma unroll
      for (int i = 0; i < 4; ++i) a[i] = sP[k][tr * 4 + i];
      b[0] = sS[k][tb * 2];
      b[1] = sS[k][tb * 2 + 1];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          acc[i][j].x += a[i].x * b[j].x - a[i].y * b[j].y;
          acc[i][j].y += a[i].x * b[j].y + a[i].y * b[j].x;
        }
    }
    __syncthreads();
  }
  double2 *dst = part + (size_t)split * npro * q.nbh;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int r = tr * 4 + i;
    const int a = r / q.lmnxx, lmn = r % q.lmnxx, atom = atom0 + a;
    if (a >= apt || atom >= q.nat) continue;
    if (lmn >= q.lmnx[q.atom_isp[atom]]) continue;
    const int ipro = q.ipro0[atom] + lmn;
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int b = b0 + tb * 2 + j;
      if (b < q.nbh) dst[(size_t)ipro * q.nbh + b] = acc[i][j];
    }
  }
}

// sum the split partials: out[i] = scale * sum_s part[s][i]
__global__ void k_reduce_c(const double2 *part, int nsplit, int n, double scale, double2 *out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double2 s = czero();
  for (int k = 0; k < nsplit; ++k) s = cadd(s, part[(size_t)k * n + i]);
  out[i] = make_double2(s.x * scale, s.y * scale);
}

// ---------------------------------------------------------------------------
// ADDPRO: psi[b][G] += sum_ipro P[ipro][G] X[ipro][b]; tile 64 G x 32 bands
// ---------------------------------------------------------------------------
constexpr int AG = 64, ABT = 32, AAT = 4;

__global__ void __launch_bounds__(kB) k_addpro(ProjGeo q, double2 *psi, const double2 *x) {
  __shared__ double2 sP[AAT * 16][AG];
  __shared__ double2 sX[AAT * 16][ABT];
  __shared__ double2 sE[AAT][AG];
  const int g0 = blockIdx.x * AG;
  const int b0 = blockIdx.y * ABT;
  const int t = threadIdx.x;
  const int tb = t / 16, tg = t % 16;
  const int rows = AAT * q.lmnxx;
  double2 acc[2][4];
#pragma unroll
  for (int j = 0; j < 2; ++j)
#pragma unroll
    for (int i = 0; i < 4; ++i) acc[j][i] = czero();
  for (int atom0 = 0; atom0 < q.nat; atom0 += AAT) {
    for (int w = t; w < AAT * AG; w += kB) {
      const int a = w / AG, k = w % AG, ig = g0 + k, atom = atom0 + a;
      sE[a][k] = (atom < q.nat && ig < q.ngl) ? eigr_of(q, atom, ig) : czero();
    }
    for (int w = t; w < rows * ABT; w += kB) {
      const int r = w / ABT, b = w % ABT;
      const int a = r / q.lmnxx, lmn = r % q.lmnxx, atom = atom0 + a;
      double2 v = czero();
      if (atom < q.nat && b0 + b < q.nbh && lmn < q.lmnx[q.atom_isp[atom]])
        v = x[(size_t)(q.ipro0[atom] + lmn) * q.nbh + b0 + b];
      sX[r][b] = v;
    }
    __syncthreads();
    for (int w = t; w < rows * AG; w += kB) {
      const int r = w / AG, k = w % AG, ig = g0 + k;
      const int a = r / q.lmnxx, lmn = r % q.lmnxx, atom = atom0 + a;
      sP[r][k] = (atom < q.nat && ig < q.ngl) ? pro_val(q, atom, lmn, ig, sE[a][k]) : czero();
    }
    __syncthreads();
    for (int r = 0; r < rows; ++r) {
      const double2 x0 = sX[r][tb * 2], x1 = sX[r][tb * 2 + 1];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const double2 p = sP[r][tg + 16 * i];
        acc[0][i].x += p.x * x0.x - p.y * x0.y;
        acc[0][i].y += p.x * x0.y + p.y * x0.x;
        acc[1][i].x += p.x * x1.x - p.y * x1.y;
        acc[1][i].y += p.x * x1.y + p.y * x1.x;
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    const int b = b0 + tb * 2 + j;
    if (b >= q.nbh) continue;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int ig = g0 + tg + 16 * i;
      if (ig < q.ngl) {
        double2 &o = psi[(size_t)b * q.ngl + ig];
        o = cadd(o, acc[j][i]);
      }
    }
  }
}

// ---------------------------------------------------------------------------
// forces: F(atom) = 2 Omega sum_G G Re[i P conj(D)], D = psi conj(dE/dproj)
// (WAVES_DEDPRO + WAVES_PROFORCE fused; the INVERTG symmetrisation of DEDPRO
//  leaves this sum unchanged because P(-G) = conj(P(G)))
// one wavefront per atom, one lane per G; part[gtile][atom][3]
// ---------------------------------------------------------------------------
constexpr int FG = 64, FAT = 4, FBK = 32, FMAXL = 16;

__global__ void __launch_bounds__(kB) k_force(ProjGeo q, const double2 *psi, const double2 *ded,
                                              const double *gvec, double *part) {
  __shared__ double2 sS[FBK][FG];
  __shared__ double2 sD[FAT][FMAXL][FBK];
  const int g0 = blockIdx.x * FG;
  const int atom0 = blockIdx.y * FAT;
  const int t = threadIdx.x;
  const int a = t / 64, lane = t % 64;
  const int atom = atom0 + a;
  const int ig = g0 + lane;
  const bool act = atom < q.nat;
  const int lmnx = act ? q.lmnx[q.atom_isp[atom]] : 0;
  double2 acc[FMAXL];
#pragma unroll
  for (int i = 0; i < FMAXL; ++i) acc[i] = czero();
  for (int bc = 0; bc < q.nbh; bc += FBK) {
    for (int w = t; w < FBK * FG; w += kB) {
      const int b = w / FG, k = w % FG, gg = g0 + k;
      sS[b][k] = (bc + b < q.nbh && gg < q.ngl) ? psi[(size_t)(bc + b) * q.ngl + gg] : czero();
    }
    for (int w = t; w < FAT * FMAXL * FBK; w += kB) {
      const int aa = w / (FMAXL * FBK), rem = w % (FMAXL * FBK);
      const int lmn = rem / FBK, b = rem % FBK, at = atom0 + aa;
      double2 v = czero();
      if (at < q.nat && lmn < q.lmnx[q.atom_isp[at]] && bc + b < q.nbh)
        v = cconj(ded[(size_t)(q.ipro0[at] + lmn) * q.nbh + bc + b]);
      sD[aa][lmn][b] = v;
    }
    __syncthreads();
    const int nbk = min(FBK, q.nbh - bc);
    for (int b = 0; b < nbk; ++b) {
      const double2 s = sS[b][lane];
#pragma unroll
      for (int i = 0; i < FMAXL; ++i) {
        if (i < lmnx) {
          const double2 d = sD[a][i][b];
          acc[i].x += d.x * s.x - d.y * s.y;
          acc[i].y += d.x * s.y + d.y * s.x;
        }
      }
    }
    __syncthreads();
  }
  double f0 = 0.0, f1 = 0.0, f2 = 0.0;
  if (act && ig < q.ngl) {
    const double2 e = eigr_of(q, atom, ig);
    double c = 0.0;
#pragma unroll
    for (int i = 0; i < FMAXL; ++i) {
      if (i < lmnx) {
        const double2 p = pro_val(q, atom, i, ig, e);
        // Re(i p conj(d)) = -Im(p conj(d)) = -(p.y d.x - p.x d.y)
        c += p.x * acc[i].y - p.y * acc[i].x;
      }
    }
    f0 = c * gvec[3 * ig];
    f1 = c * gvec[3 * ig + 1];
    f2 = c * gvec[3 * ig + 2];
  }
  // wavefront reduction (fixed tree)
  for (int off = 32; off > 0; off >>= 1) {
    f0 += __shfl_down(f0, off, 64);
    f1 += __shfl_down(f1, off, 64);
    f2 += __shfl_down(f2, off, 64);
  }
  if (lane == 0 && act) {
    double *dst = part + ((size_t)blockIdx.x * q.nat + atom) * 3;
    dst[0] = f0; dst[1] = f1; dst[2] = f2;
  }
}

__global__ void k_reduce_r(const double *part, int nsplit, int n, double scale, double *out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s = 0.0;
  for (int k = 0; k < nsplit; ++k) s += part[(size_t)k * n + i];
  out[i] = s * scale;
}

// ---------------------------------------------------------------------------
// overlaps: T = sum_G conj(psi1_i) psi2_j, TM = sum_G psi1_i psi2_j(-G)
// tile 32 x 32 bands, split over G; part[split][2][nbh][nbh]
// ---------------------------------------------------------------------------
constexpr int OT = 32, OKC = 32;

__global__ void __launch_bounds__(kB) k_overlap(int ngl, int nbh, const int *minusg,
                                                const double2 *psi1, const double2 *psi2,
                                                int gsplit, double2 *part) {
  __shared__ double2 s1[OKC][OT];
  __shared__ double2 s2[OKC][OT];
  __shared__ double2 s2m[OKC][OT];
  const int i0 = blockIdx.x * OT, j0 = blockIdx.y * OT, split = blockIdx.z;
  const int g0 = split * gsplit, g1 = min(ngl, g0 + gsplit);
  const int t = threadIdx.x, ti = t / 16, tj = t % 16;
  double2 at[2][2], am[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) { at[a][b] = czero(); am[a][b] = czero(); }
  for (int gc = g0; gc < g1; gc += OKC) {
    for (int w = t; w < OKC * OT; w += kB) {
      const int b = w / OKC, k = w % OKC, ig = gc + k;
      const bool ok = ig < g1;
      s1[k][b] = (ok && i0 + b < nbh) ? psi1[(size_t)(i0 + b) * ngl + ig] : czero();
      s2[k][b] = (ok && j0 + b < nbh) ? psi2[(size_t)(j0 + b) * ngl + ig] : czero();
      s2m[k][b] = (ok && j0 + b < nbh) ? psi2[(size_t)(j0 + b) * ngl + minusg[ig]] : czero();
    }
    __syncthreads();
#pragma unroll 4
    for (int k = 0; k < OKC; ++k) {
      double2 a[2] = {s1[k][ti * 2], s1[k][ti * 2 + 1]};
      double2 b[2] = {s2[k][tj * 2], s2[k][tj * 2 + 1]};
      double2 bm[2] = {s2m[k][tj * 2], s2m[k][tj * 2 + 1]};
#pragma unroll
      for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y) {
          // conj(a) b
          at[x][y].x += a[x].x * b[y].x + a[x].y * b[y].y;
          at[x][y].y += a[x].x * b[y].y - a[x].y * b[y].x;
          // a bm
          am[x][y].x += a[x].x * bm[y].x - a[x].y * bm[y].y;
          am[x][y].y += a[x].x * bm[y].y + a[x].y * bm[y].x;
        }
    }
    __syncthreads();
  }
  double2 *dst = part + (size_t)split * 2 * nbh * nbh;
#pragma unroll
  for (int x = 0; x < 2; ++x)
#pragma unroll
    for (int y = 0; y < 2; ++y) {
      const int i = i0 + ti * 2 + x, j = j0 + tj * 2 + y;
      if (i < nbh && j < nbh) {
        dst[(size_t)i * nbh + j] = at[x][y];
        dst[(size_t)nbh * nbh + (size_t)i * nbh + j] = am[x][y];
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
// c1 = sum_k conj(p1_ki) p2a_kj, c2 = sum_k p1_ki p2a_kj
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_overlap1c(int npro, int nbh, int nb, const double2 *p1,
                                                   const double2 *p2a, double *mat) {
  __shared__ double2 sa[64][17];
  __shared__ double2 sb[64][17];
  const int ti = threadIdx.x / 16, tj = threadIdx.x % 16;
  const int i0 = blockIdx.x * 16, j0 = blockIdx.y * 16;
  const int i = i0 + ti, j = j0 + tj;
  double2 c1 = czero(), c2 = czero();
  for (int k0 = 0; k0 < npro; k0 += 64) {
    for (int w = threadIdx.x; w < 64 * 16; w += 256) {
      const int k = w / 16, c = w % 16;
      const bool ok = k0 + k < npro;
      sa[k][c] = (ok && i0 + c < nbh) ? p1[(size_t)(k0 + k) * nbh + i0 + c] : czero();
      sb[k][c] = (ok && j0 + c < nbh) ? p2a[(size_t)(k0 + k) * nbh + j0 + c] : czero();
    }
    __syncthreads();
#pragma unroll 8
    for (int k = 0; k < 64; ++k) {
      const double2 a = sa[k][ti], b = sb[k][tj];
      c1.x += a.x * b.x + a.y * b.y;
      c1.y += a.x * b.y - a.y * b.x;
      c2.x += a.x * b.x - a.y * b.y;
      c2.y += a.x * b.y + a.y * b.x;
    }
    __syncthreads();
  }
  if (i >= nbh || j >= nbh) return;
  const double2 s1 = make_double2(0.5 * (c1.x + c2.x), 0.5 * (c1.y + c2.y));
  const double2 s2 = make_double2(0.5 * (c1.x - c2.x), 0.5 * (c1.y - c2.y));
  const int ia = 2 * i, ja = 2 * j;
  mat[(size_t)ia * nb + ja] += s1.x;
  mat[(size_t)ia * nb + ja + 1] += s1.y;
  mat[(size_t)(ia + 1) * nb + ja] += -s2.y;
  mat[(size_t)(ia + 1) * nb + ja + 1] += s2.x;
}

// per-atom matrix on projections: out[ipro0+i][b] = sum_j M_atom[i][j] proj[ipro0+j][b]
__global__ void k_atom_mat(ProjGeo q, const double2 *proj, const double2 *mats, double2 *out) {
  const int atom = blockIdx.x;
  const int lmnx = q.lmnx[q.atom_isp[atom]];
  const int base = q.ipro0[atom];
  const double2 *m = mats + (size_t)atom * q.lmnxx * q.lmnxx;
  for (int w = threadIdx.x; w < lmnx * q.nbh; w += blockDim.x) {
    const int i = w / q.nbh, b = w % q.nbh;
    double2 s = czero();
    for (int j = 0; j < lmnx; ++j) s = cadd(s, cmul(m[i * q.lmnxx + j], proj[(size_t)(base + j) * q.nbh + b]));
    out[(size_t)(base + i) * q.nbh + b] = s;
  }
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
constexpr int XG = 64, XJ = 32, XK = 16;

__global__ void __launch_bounds__(kB) k_addopsi(int ngl, int nbh, int nb, const int *minusg,
                                                double2 *psibar, const double2 *chi,
                                                const double *lam) {
  __shared__ double2 sC[XK][XG];
  __shared__ double2 sM[XK][XG];
  __shared__ double2 sL1[XK][XJ];
  __shared__ double2 sL2[XK][XJ];
  const int g0 = blockIdx.x * XG, j0 = blockIdx.y * XJ;
  const int t = threadIdx.x, tj = t / 16, tg = t % 16;
  double2 acc[2][4];
#pragma unroll
  for (int j = 0; j < 2; ++j)
#pragma unroll
    for (int i = 0; i < 4; ++i) acc[j][i] = czero();
  for (int i0 = 0; i0 < nbh; i0 += XK) {
    for (int w = t; w < XK * XG; w += kB) {
      const int i = w / XG, k = w % XG, ig = g0 + k;
      const bool ok = i0 + i < nbh && ig < ngl;
      sC[i][k] = ok ? chi[(size_t)(i0 + i) * ngl + ig] : czero();
      sM[i][k] = ok ? cconj(chi[(size_t)(i0 + i) * ngl + minusg[ig]]) : czero();
    }
    for (int w = t; w < XK * XJ; w += kB) {
      const int i = w / XJ, j = w % XJ;
      double2 l1 = czero(), l2 = czero();
      if (i0 + i < nbh && j0 + j < nbh) fold_lambda(lam, nb, i0 + i, j0 + j, l1, l2);
      sL1[i][j] = l1;
      sL2[i][j] = l2;
    }
    __syncthreads();
    for (int i = 0; i < XK; ++i) {
      const double2 a0 = sL1[i][tj * 2], a1 = sL1[i][tj * 2 + 1];
      const double2 c0 = sL2[i][tj * 2], c1 = sL2[i][tj * 2 + 1];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const double2 x = sC[i][tg + 16 * q], xm = sM[i][tg + 16 * q];
        acc[0][q] = cadd(acc[0][q], cadd(cmul(a0, x), cmul(c0, xm)));
        acc[1][q] = cadd(acc[1][q], cadd(cmul(a1, x), cmul(c1, xm)));
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    const int jj = j0 + tj * 2 + j;
    if (jj >= nbh) continue;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int ig = g0 + tg + 16 * q;
      if (ig < ngl) {
        double2 &o = psibar[(size_t)jj * ngl + ig];
        o = cadd(o, acc[j][q]);
      }
    }
  }
}

// ADDOPROJ: proj[k][j] += sum_i oproj[k][i] L1[i][j] + conj(oproj[k][i]) L2[i][j]
__global__ void k_addoproj(int npro, int nbh, int nb, double2 *proj, const double2 *oproj,
                           const double *lam) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= npro * nbh) return;
  const int k = idx / nbh, j = idx % nbh;
  double2 s = czero();
  for (int i = 0; i < nbh; ++i) {
    double2 l1, l2;
    fold_lambda(lam, nb, i, j, l1, l2);
    const double2 o = oproj[(size_t)k * nbh + i];
    s = cadd(s, cadd(cmul(o, l1), cmul(cconj(o), l2)));
  }
  proj[idx] = cadd(proj[idx], s);
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

// ekin partial per block: sum_G g2 * sum_b [fp |c|^2 + fm Re(c(G) c(-G))]
__global__ void __launch_bounds__(kB) k_ekin(int ngl, int nbh, const double *g2, const int *minusg,
                                             const double *fpm, const double2 *psi, double *part) {
  __shared__ double red[kB / 64];
  double s = 0.0;
  for (int ig = blockIdx.x * kB + threadIdx.x; ig < ngl; ig += gridDim.x * kB) {
    double d = 0.0;
    const int mg = minusg[ig];
    for (int b = 0; b < nbh; ++b) {
      const double fp = fpm[2 * b], fm = fpm[2 * b + 1];
      const double2 c = psi[(size_t)b * ngl + ig];
      if (fp != 0.0) d += fp * (c.x * c.x + c.y * c.y);
      if (fm != 0.0) {
        const double2 cm = psi[(size_t)b * ngl + mg];
        d += fm * (c.x * cm.x - c.y * cm.y);
      }
    }
    s += g2[ig] * d;
  }
  s = block_sum<kB>(s, red);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// wave kinetic partial: sum_G m(G) sum_b [fa |d|^2 + fb (dr dr(-G) - di di(-G))]
__global__ void __launch_bounds__(kB) k_wavekin(int ngl, int nbh, const double *mpsi,
                                                const int *minusg, const double *fab,
                                                const double2 *p0, const double2 *pm,
                                                double *part) {
  __shared__ double red[kB / 64];
  double s = 0.0;
  for (int ig = blockIdx.x * kB + threadIdx.x; ig < ngl; ig += gridDim.x * kB) {
    double d = 0.0;
    const int mg = minusg[ig];
    for (int b = 0; b < nbh; ++b) {
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
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// <psi|H psi> per super band: out[b] = (sum conj(psi) hpsi, sum psi hpsi(-G))
__global__ void __launch_bounds__(kB) k_expect(int ngl, const int *minusg, const double2 *psi,
                                               const double2 *hpsi, double *out) {
  __shared__ double red[kB / 64];
  const int b = blockIdx.x;
  const double2 *p = psi + (size_t)b * ngl;
  const double2 *h = hpsi + (size_t)b * ngl;
  double tr = 0.0, tmr = 0.0;
  for (int ig = threadIdx.x; ig < ngl; ig += kB) {
    const double2 a = p[ig], c = h[ig], cm = h[minusg[ig]];
    tr += a.x * c.x + a.y * c.y;
    tmr += a.x * cm.x - a.y * cm.y;
  }
  tr = block_sum<kB>(tr, red);
  tmr = block_sum<kB>(tmr, red);
  if (threadIdx.x == 0) {
    out[2 * b] = tr;
    out[2 * b + 1] = tmr;
  }
}

// one-centre density matrix (real part, before hermitisation) per atom
__global__ void k_denmat(ProjGeo q, const double *occ2, const double2 *proj, double *dm) {
  const int atom = blockIdx.x;
  const int lmnx = q.lmnx[q.atom_isp[atom]];
  const int base = q.ipro0[atom];
  for (int w = threadIdx.x; w < lmnx * lmnx; w += blockDim.x) {
    const int i = w / lmnx, j = w % lmnx;
    double s = 0.0;
    for (int b = 0; b < q.nbh; ++b) {
      const double s1 = occ2[2 * b], s2 = occ2[2 * b + 1];
      const double2 pi = proj[(size_t)(base + i) * q.nbh + b];
      const double2 pj = proj[(size_t)(base + j) * q.nbh + b];
      // Re(pi (s1 conj(pj) + s2 pj))
      s += s1 * (pi.x * pj.x + pi.y * pj.y) + s2 * (pi.x * pj.x - pi.y * pj.y);
    }
    dm[((size_t)atom * q.lmnxx + i) * q.lmnxx + j] = s;
  }
}

__global__ void __launch_bounds__(kB) k_dot(size_t n, const double *a, const double *b, double *part) {
  __shared__ double red[kB / 64];
  double s = 0.0;
  for (size_t i = (size_t)blockIdx.x * kB + threadIdx.x; i < n; i += (size_t)gridDim.x * kB) s += a[i] * b[i];
  s = block_sum<kB>(s, red);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ void k_propagate_g(int ngl, int nbh, double a1, double a2, double a3n,
                              const double *mpsi, const double2 *psi0, double2 *psim,
                              const double2 *hpsi) {
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (size_t)ngl * nbh) return;
  const int ig = idx % ngl;
  const double a3 = a3n / mpsi[ig];
  const double2 p0 = psi0[idx], pm = psim[idx], h = hpsi[idx];
  psim[idx] = make_double2(a1 * p0.x + a2 * pm.x + a3 * h.x, a1 * p0.y + a2 * pm.y + a3 * h.y);
}

__global__ void k_scale_mass(int ngl, int nbh, double num, const double *mpsi, double2 *x) {
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (size_t)ngl * nbh) return;
  const double f = num / mpsi[idx % ngl];
  x[idx] = make_double2(f * x[idx].x, f * x[idx].y);
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

int nsplit_for(int ngl, int tiles, int chunk) {
  int want = std::max(1, 1024 / std::max(tiles, 1));
  int maxs = std::max(1, (ngl + chunk - 1) / chunk);
  return std::min(want, maxs);
}

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

int ws_small(pw_waveset *ws, size_t bytes) {
  if (bytes > ws->small_bytes) {
    if (ws->d_small) PW_HIP(hipFree(ws->d_small));
    PW_HIP(hipMalloc(&ws->d_small, bytes));
    ws->small_bytes = bytes;
  }
  return PW_OK;
}

int k_phase_tables(pw_waveset *ws, const double *r, double2 *phase) {
  // r: host positions (3*nat); gr[atom*3+k] = GBAS(:,k).R
  pw_plan *p = ws->plan;
  std::vector<double> gr(3 * ws->nat);
  for (int a = 0; a < ws->nat; ++a)
    for (int k = 0; k < 3; ++k)
      gr[3 * a + k] = p->gbas[0 + 3 * k] * r[3 * a] + p->gbas[1 + 3 * k] * r[3 * a + 1] +
                      p->gbas[2 + 3 * k] * r[3 * a + 2];
  PW_TRY(ws_small(ws, 3 * ws->nat * sizeof(double)));
  PW_HIP(hipMemcpyAsync(ws->d_small, gr.data(), gr.size() * sizeof(double), hipMemcpyHostToDevice,
                        ws->stream));
  const int total = ws->nat * 3 * ws->nphase;
  ScopeTimer tm(ws, "phase");
  k_phase<<<(total + 255) / 256, 256, 0, ws->stream>>>(ws->nat, ws->nphase, (ws->nphase - 1) / 2,
                                                        ws->d_small, phase);
  PW_HIP(hipGetLastError());
  // the scratch is reused by the next call; keep the copy ordered before it
  PW_HIP(hipStreamSynchronize(ws->stream));
  return PW_OK;
}

int k_projections(pw_waveset *ws, const double2 *psi, const double2 *phase, double2 *proj) {
  ProjGeo q = make_pgeo(ws);
  q.phase = phase;
  const int apt = std::min(8, PR / ws->lmnxx);
  const int ta = (ws->nat + apt - 1) / apt;
  const int tb = (ws->nbh + PBT - 1) / PBT;
  const int ns = nsplit_for(q.ngl, ta * tb, PKC);
  const int gsplit = ((q.ngl + ns - 1) / ns + PKC - 1) / PKC * PKC;
  const int nsr = (q.ngl + gsplit - 1) / gsplit;
  const size_t n = (size_t)ws->npro * ws->nbh;
  PW_TRY(ws_part(ws, (size_t)nsr * n * sizeof(double2)));
  {
    ScopeTimer tm(ws, "projections");
    k_proj<<<dim3(ta, tb, nsr), kB, 0, ws->stream>>>(q, psi, gsplit, ws->npro, (double2 *)ws->d_part);
    PW_HIP(hipGetLastError());
  }
  {
    ScopeTimer tm(ws, "reduce");
    k_reduce_c<<<(n + 255) / 256, 256, 0, ws->stream>>>((const double2 *)ws->d_part, nsr, (int)n,
                                                        ws->plan->vol, proj);
    PW_HIP(hipGetLastError());
  }
  return PW_OK;
}

int k_addpro(pw_waveset *ws, double2 *psi, const double2 *phase, const double2 *x) {
  if (ws->lmnxx > 16) return pw_fail(PW_ERR_UNSUPPORTED, "WAVES_ADDPRO: LMNX > 16");
  ProjGeo q = make_pgeo(ws);
  q.phase = phase;
  dim3 grid((q.ngl + AG - 1) / AG, (ws->nbh + ABT - 1) / ABT);
  ScopeTimer tm(ws, "addpro");
  k_addpro<<<grid, kB, 0, ws->stream>>>(q, psi, x);
  PW_HIP(hipGetLastError());
  return PW_OK;
}

int k_force(pw_waveset *ws, const double2 *psi, const double2 *phase, const double2 *ded,
            double *force_host) {
  if (ws->lmnxx > FMAXL) return pw_fail(PW_ERR_UNSUPPORTED, "WAVES$FORCE: LMNX > 16");
  ProjGeo q = make_pgeo(ws);
  q.phase = phase;
  const int ngt = (q.ngl + FG - 1) / FG;
  const size_t n = (size_t)ws->nat * 3;
  PW_TRY(ws_part(ws, (size_t)ngt * n * sizeof(double)));
  PW_TRY(ws_small(ws, n * sizeof(double)));
  {
    ScopeTimer tm(ws, "force");
    k_force<<<dim3(ngt, (ws->nat + FAT - 1) / FAT), kB, 0, ws->stream>>>(q, psi, ded, ws->plan->d_gvec,
                                                                        ws->d_part);
    PW_HIP(hipGetLastError());
  }
  k_reduce_r<<<(n + 255) / 256, 256, 0, ws->stream>>>(ws->d_part, ngt, (int)n, 2.0 * ws->plan->vol,
                                                      ws->d_small);
  PW_HIP(hipGetLastError());
  PW_HIP(hipMemcpyAsync(force_host, ws->d_small, n * sizeof(double), hipMemcpyDeviceToHost, ws->stream));
  PW_HIP(hipStreamSynchronize(ws->stream));
  return PW_OK;
}

// mat_dev (row-major NB x NB) = WAVES_OVERLAP(psi1, psi2)
int k_overlap(pw_waveset *ws, const double2 *psi1, const double2 *psi2, double *mat, bool, bool) {
  pw_plan *p = ws->plan;
  const int nbh = ws->nbh;
  const int tiles = ((nbh + OT - 1) / OT) * ((nbh + OT - 1) / OT);
  const int ns = nsplit_for(p->ngl, tiles, OKC);
  const int gsplit = ((p->ngl + ns - 1) / ns + OKC - 1) / OKC * OKC;
  const int nsr = (p->ngl + gsplit - 1) / gsplit;
  const size_t n = (size_t)2 * nbh * nbh;
  PW_TRY(ws_part(ws, (size_t)(nsr + 1) * n * sizeof(double2)));
  double2 *red = (double2 *)ws->d_part + (size_t)nsr * n;
  {
    ScopeTimer tm(ws, "overlap");
    k_overlap<<<dim3((nbh + OT - 1) / OT, (nbh + OT - 1) / OT, nsr), kB, 0, ws->stream>>>(
        p->ngl, nbh, p->d_minusg, psi1, psi2, gsplit, (double2 *)ws->d_part);
    PW_HIP(hipGetLastError());
  }
  k_reduce_c<<<(n + 255) / 256, 256, 0, ws->stream>>>((const double2 *)ws->d_part, nsr, (int)n, p->vol,
                                                      red);
  k_unravel_pw<<<(nbh * nbh + 255) / 256, 256, 0, ws->stream>>>(nbh, ws->nb, red, red + (size_t)nbh * nbh,
                                                                  mat);
  PW_HIP(hipGetLastError());
  return PW_OK;
}

// mat_dev += WAVES_1COVERLAP(proj1, proj2); uses d_xproj as DO*proj2 scratch
int k_overlap_1c(pw_waveset *ws, const double2 *proj1, const double2 *proj2, double *mat) {
  ProjGeo q = make_pgeo(ws);
  ScopeTimer tm(ws, "overlap1c");
  k_atom_mat<<<ws->nat, 256, 0, ws->stream>>>(q, proj2, ws->d_domat, ws->d_xproj);
  dim3 grid((ws->nbh + 15) / 16, (ws->nbh + 15) / 16);
  k_overlap1c<<<grid, 256, 0, ws->stream>>>(ws->npro, ws->nbh, ws->nb, proj1, ws->d_xproj, mat);
  PW_HIP(hipGetLastError());
  return PW_OK;
}

int k_atom_matrix(pw_waveset *ws, const double2 *proj, double2 *out, const double2 *mats) {
  ProjGeo q = make_pgeo(ws);
  ScopeTimer tm(ws, "atom_mat");
  k_atom_mat<<<ws->nat, 256, 0, ws->stream>>>(q, proj, mats, out);
  PW_HIP(hipGetLastError());
  return PW_OK;
}

int k_addopsi(pw_waveset *ws, double2 *psibar, const double2 *chi, const double *lam) {
  pw_plan *p = ws->plan;
  dim3 grid((p->ngl + XG - 1) / XG, (ws->nbh + XJ - 1) / XJ);
  ScopeTimer tm(ws, "addopsi");
  k_addopsi<<<grid, kB, 0, ws->stream>>>(p->ngl, ws->nbh, ws->nb, p->d_minusg, psibar, chi, lam);
  PW_HIP(hipGetLastError());
  return PW_OK;
}

int k_addoproj(pw_waveset *ws, double2 *proj, const double2 *oproj, const double *lam) {
  const int n = ws->npro * ws->nbh;
  ScopeTimer tm(ws, "addoproj");
  k_addoproj<<<(n + 255) / 256, 256, 0, ws->stream>>>(ws->npro, ws->nbh, ws->nb, proj, oproj, lam);
  PW_HIP(hipGetLastError());
  return PW_OK;
}

int k_fold_occ(pw_waveset *ws, double2 *x) {
  const int n = ws->npro * ws->nbh;
  k_fold_f<<<(n + 255) / 256, 256, 0, ws->stream>>>(ws->npro, ws->nbh, ws->d_occ2, x);
  PW_HIP(hipGetLastError());
  return PW_OK;
}

static int reduce_blocks(pw_plan *p) { return std::min(1024, (p->ngl + kB - 1) / kB); }

int k_ekin(pw_waveset *ws, const double2 *psi, const double *fpm_dev, double *out_host) {
  pw_plan *p = ws->plan;
  const int nbk = reduce_blocks(p);
  PW_TRY(ws_part(ws, nbk * sizeof(double)));
  {
    ScopeTimer tm(ws, "ekin");
    k_ekin<<<nbk, kB, 0, ws->stream>>>(p->ngl, ws->nbh, p->d_g2, p->d_minusg, fpm_dev, psi, ws->d_part);
    PW_HIP(hipGetLastError());
  }
  std::vector<double> h(nbk);
  PW_HIP(hipMemcpyAsync(h.data(), ws->d_part, nbk * sizeof(double), hipMemcpyDeviceToHost, ws->stream));
  PW_HIP(hipStreamSynchronize(ws->stream));
  double s = 0.0;
  for (double v : h) s += v;
  *out_host = 0.5 * s * p->vol;
  return PW_OK;
}

int k_wavekinetic(pw_waveset *ws, const double2 *psi0, const double2 *psim, const double *fab_dev,
                  double *out_host) {
  pw_plan *p = ws->plan;
  const int nbk = reduce_blocks(p);
  PW_TRY(ws_part(ws, nbk * sizeof(double)));
  {
    ScopeTimer tm(ws, "wavekinetic");
    k_wavekin<<<nbk, kB, 0, ws->stream>>>(p->ngl, ws->nbh, ws->d_mpsi, p->d_minusg, fab_dev, psi0, psim,
                                          ws->d_part);
    PW_HIP(hipGetLastError());
  }
  std::vector<double> h(nbk);
  PW_HIP(hipMemcpyAsync(h.data(), ws->d_part, nbk * sizeof(double), hipMemcpyDeviceToHost, ws->stream));
  PW_HIP(hipStreamSynchronize(ws->stream));
  double s = 0.0;
  for (double v : h) s += v;
  *out_host = s / (ws->delt * ws->delt) * p->vol;
  return PW_OK;
}

int k_expect(pw_waveset *ws, const double2 *psi, const double2 *hpsi, double *eig_host) {
  pw_plan *p = ws->plan;
  PW_TRY(ws_part(ws, 2 * ws->nbh * sizeof(double)));
  {
    ScopeTimer tm(ws, "expect");
    k_expect<<<ws->nbh, kB, 0, ws->stream>>>(p->ngl, p->d_minusg, psi, hpsi, ws->d_part);
    PW_HIP(hipGetLastError());
  }
  std::vector<double> h(2 * ws->nbh);
  PW_HIP(hipMemcpyAsync(h.data(), ws->d_part, h.size() * sizeof(double), hipMemcpyDeviceToHost, ws->stream));
  PW_HIP(hipStreamSynchronize(ws->stream));
  for (int b = 0; b < ws->nbh; ++b) {
    const double re = 0.5 * h[2 * b] * p->vol, rem = 0.5 * h[2 * b + 1] * p->vol;
    eig_host[2 * b] = re + rem;
    if (2 * b + 1 < ws->nb) eig_host[2 * b + 1] = re - rem;
  }
  return PW_OK;
}

int k_denmat(pw_waveset *ws, const double2 *proj, double *dm_host) {
  ProjGeo q = make_pgeo(ws);
  const size_t n = (size_t)ws->nat * ws->lmnxx * ws->lmnxx;
  PW_TRY(ws_part(ws, n * sizeof(double)));
  PW_HIP(hipMemsetAsync(ws->d_part, 0, n * sizeof(double), ws->stream));
  {
    ScopeTimer tm(ws, "denmat");
    k_denmat<<<ws->nat, 256, 0, ws->stream>>>(q, ws->d_occ2, proj, ws->d_part);
    PW_HIP(hipGetLastError());
  }
  PW_HIP(hipMemcpyAsync(dm_host, ws->d_part, n * sizeof(double), hipMemcpyDeviceToHost, ws->stream));
  PW_HIP(hipStreamSynchronize(ws->stream));
  return PW_OK;
}

int k_dot_vrho(pw_waveset *ws, double *out_host) {
  pw_plan *p = ws->plan;
  const int nbk = std::min(1024, (p->nrl + kB - 1) / kB);
  PW_TRY(ws_part(ws, nbk * sizeof(double)));
  {
    ScopeTimer tm(ws, "vrho");
    k_dot<<<nbk, kB, 0, ws->stream>>>((size_t)p->nrl, ws->d_v, ws->d_rho, ws->d_part);
    PW_HIP(hipGetLastError());
  }
  std::vector<double> h(nbk);
  PW_HIP(hipMemcpyAsync(h.data(), ws->d_part, nbk * sizeof(double), hipMemcpyDeviceToHost, ws->stream));
  PW_HIP(hipStreamSynchronize(ws->stream));
  double s = 0.0;
  for (double v : h) s += v;
  *out_host = s * p->vol / (double)p->nrl;
  return PW_OK;
}

int k_propagate(pw_waveset *ws) {
  pw_plan *p = ws->plan;
  const double a1 = 2.0 / (1.0 + ws->annee), a2 = 1.0 - a1;
  const double a3n = -ws->delt * ws->delt / (1.0 + ws->annee);
  const size_t n = (size_t)p->ngl * ws->nbh;
  ScopeTimer tm(ws, "propagate");
  k_propagate_g<<<(n + 255) / 256, 256, 0, ws->stream>>>(p->ngl, ws->nbh, a1, a2, a3n, ws->d_mpsi,
                                                         ws->d_psi0, ws->d_psim, ws->d_hpsi);
  PW_HIP(hipGetLastError());
  return PW_OK;
}

int k_opsi_scale(pw_waveset *ws, double2 *opsi) {
  pw_plan *p = ws->plan;
  const size_t n = (size_t)p->ngl * ws->nbh;
  ScopeTimer tm(ws, "opsi_scale");
  k_scale_mass<<<(n + 255) / 256, 256, 0, ws->stream>>>(p->ngl, ws->nbh,
                                                        ws->delt * ws->delt / (1.0 + ws->annee),
                                                        ws->d_mpsi, opsi);
  PW_HIP(hipGetLastError());
  return PW_OK;
}

int k_copy(pw_waveset *ws, double2 *dst, const double2 *src, size_t n) {
  PW_HIP(hipMemcpyAsync(dst, src, n * sizeof(double2), hipMemcpyDeviceToDevice, ws->stream));
  return PW_OK;
}
