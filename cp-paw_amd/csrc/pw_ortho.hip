// WAVES_ORTHO_X (src/paw_waves2.f90:1844-2022) in FP64 on the GPU.
//
// The reference iterates
//   GAMN   = sym[LAMBDA^T (CHICHI LAMBDA + 2 CHIPSI)] + PSIPSI - 1
//   LAMBDA -= 2 f_i/(f_i+f_j) * S^-1(GAMN),   S(X) = B X + X B,  B = sym(CHIPSI)
// until max|GAMN| < 1e-8, applying S^-1 through the eigenvectors of B (DSYEV).
// The iterates depend on S^-1 only, not on an eigen-basis, so this
// implementation keeps the reference's outer iteration and stopping rule and
// applies S^-1 exactly (to rounding) with GEMMs only, by the Cayley transform
// and the squared Smith iteration:
//   M = (B + qI)^-1 (Newton-Schulz), A = I - 2qM = (B - qI)M,
//   S^-1(G) = sum_k A^k (2q M G M) A^k, summed by X <- X + A_j X A_j with
//   A_j = A^(2^j); the error after J levels is rho^(2^J), rho =
//   (sqrt(k)-1)/(sqrt(k)+1) for the spectral bounds of B (host Lanczos on
//   sym(B), widened).  B is positive definite in the CP step (B ~ dt^2/m).
// Every product is an NB^3 GEMM on the device; nothing is diagonalised.
#include <algorithm>
#include <cmath>

#include "pw_common.h"

namespace {

constexpr int T16 = 16;

// C = alpha op(A) op(B) + beta D   (row-major n x n).  One 16x16 tile of C
// per workgroup; the A row panel and B column panel (16 x n each) are loaded
// into LDS in one pass, so the kernel pays one memory latency, not n/16.
__global__ void __launch_bounds__(256) k_gemm(int n, const double *A, int ta, const double *B, int tb,
                                              double *C, double alpha, const double *D, double beta) {
  extern __shared__ double smem[];
  double *sa = smem;                 // [16][n+1]
  double *sb = smem + T16 * (n + 1); // [n][17]
  const int ty = threadIdx.x / T16, tx = threadIdx.x % T16;
  const int r0 = blockIdx.y * T16, c0 = blockIdx.x * T16;
  for (int w = threadIdx.x; w < T16 * n; w += 256) {
    int r, k;
    if (ta) { r = w % T16; k = w / T16; } else { k = w % n; r = w / n; }
    double v = 0.0;
    if (r0 + r < n) v = ta ? A[(size_t)k * n + r0 + r] : A[(size_t)(r0 + r) * n + k];
    sa[r * (n + 1) + k] = v;
  }
  for (int w = threadIdx.x; w < T16 * n; w += 256) {
    int k, c;
    if (tb) { k = w % n; c = w / n; } else { c = w % T16; k = w / T16; }
    double v = 0.0;
    if (c0 + c < n) v = tb ? B[(size_t)(c0 + c) * n + k] : B[(size_t)k * n + c0 + c];
    sb[k * (T16 + 1) + c] = v;
  }
  __syncthreads();
  double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
  const double *ra = sa + ty * (n + 1);
  int k = 0;
  for (; k + 3 < n; k += 4) {
    a0 += ra[k] * sb[k * (T16 + 1) + tx];
    a1 += ra[k + 1] * sb[(k + 1) * (T16 + 1) + tx];
    a2 += ra[k + 2] * sb[(k + 2) * (T16 + 1) + tx];
    a3 += ra[k + 3] * sb[(k + 3) * (T16 + 1) + tx];
  }
  for (; k < n; ++k) a0 += ra[k] * sb[k * (T16 + 1) + tx];
  const int row = r0 + ty, col = c0 + tx;
  if (row < n && col < n) {
    double v = alpha * ((a0 + a1) + (a2 + a3));
    if (D) v += beta * D[(size_t)row * n + col];
    C[(size_t)row * n + col] = v;
  }
}

__global__ void k_symm(int n, const double *B, double *Bs) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n * n) return;
  const int i = idx / n, j = idx % n;
  Bs[idx] = 0.5 * (B[(size_t)i * n + j] + B[(size_t)j * n + i]);
}

// GAMN(i,j) = 0.5(T(i,j)+T(j,i)) + A(min,max) - delta; block maxima of |GAMN|
__global__ void k_gamma(int n, const double *T, const double *A, double *G, double *bmax) {
  __shared__ double red[256];
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  double m = 0.0;
  if (idx < n * n) {
    const int i = idx / n, j = idx % n;
    const int lo = min(i, j), hi = max(i, j);
    double v = 0.5 * (T[(size_t)lo * n + hi] + T[(size_t)hi * n + lo]) + A[(size_t)lo * n + hi];
    if (i == j) v -= 1.0;
    G[idx] = v;
    m = fabs(v);
  }
  red[threadIdx.x] = m;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) bmax[blockIdx.x] = red[0];
}

// LAMBDA(i,j) -= 2 (f_i+ds)/(f_i+f_j+2ds) X(i,j)   (src/paw_waves2.f90:1979-1987)
__global__ void k_lam_update(int n, const double *occ, const double *X, double *lam) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n * n) return;
  const int i = idx / n, j = idx % n;
  const double oi = occ[i] + 1.0e-12, oj = occ[j] + 1.0e-12;
  lam[idx] -= 2.0 * oi / (oi + oj) * X[idx];
}

// LAMBDA(i,j) = LAMBDA(j,i) (f_i+ds)/(f_j+ds) where f_i < f_j (src/paw_waves2.f90:1999-2005)
__global__ void k_lam_sym(int n, const double *occ, const double *lin, double *lout) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n * n) return;
  const int i = idx / n, j = idx % n;
  if (occ[i] < occ[j])
    lout[idx] = lin[(size_t)j * n + i] * (occ[i] + 1.0e-12) / (occ[j] + 1.0e-12);
  else
    lout[idx] = lin[idx];
}

}  // namespace

// cyclic Jacobi: a (n x n row-major symmetric) -> eigenvalues e, eigenvectors
// as columns of u (row-major u[i*n+k] = component i of vector k)
void host_syev(int n, std::vector<double> &a, std::vector<double> &e, std::vector<double> &u) {
  u.assign((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i) u[(size_t)i * n + i] = 1.0;
  for (int sweep = 0; sweep < 100; ++sweep) {
    bool rotated = false;
    for (int p = 0; p < n - 1; ++p) {
      for (int q = p + 1; q < n; ++q) {
        const double apq = a[(size_t)p * n + q];
        const double app = a[(size_t)p * n + p], aqq = a[(size_t)q * n + q];
        if (std::fabs(apq) <= 1.0e-18 * std::sqrt(std::fabs(app * aqq)) || apq == 0.0) continue;
        rotated = true;
        const double theta = (aqq - app) / (2.0 * apq);
        const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < n; ++k) {
          const double akp = a[(size_t)k * n + p], akq = a[(size_t)k * n + q];
          a[(size_t)k * n + p] = c * akp - s * akq;
          a[(size_t)k * n + q] = s * akp + c * akq;
        }
        for (int k = 0; k < n; ++k) {
          const double apk = a[(size_t)p * n + k], aqk = a[(size_t)q * n + k];
          a[(size_t)p * n + k] = c * apk - s * aqk;
          a[(size_t)q * n + k] = s * apk + c * aqk;
        }
        for (int k = 0; k < n; ++k) {
          const double ukp = u[(size_t)k * n + p], ukq = u[(size_t)k * n + q];
          u[(size_t)k * n + p] = c * ukp - s * ukq;
          u[(size_t)k * n + q] = s * ukp + c * ukq;
        }
      }
    }
    if (!rotated) break;
  }
  e.resize(n);
  for (int i = 0; i < n; ++i) e[i] = a[(size_t)i * n + i];
}

namespace {

__global__ void k_scaled_identity(int n, double a, double *X) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n * n) return;
  X[idx] = (idx / n == idx % n) ? a : 0.0;
}

// A = I - 2 q M
__global__ void k_cayley(int n, double q, const double *M, double *A) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n * n) return;
  A[idx] = ((idx / n == idx % n) ? 1.0 : 0.0) - 2.0 * q * M[idx];
}

constexpr int kMaxLevels = 12;

struct OrthoWork {
  double *bs, *m0, *m1, *gam, *x0, *x1, *t1, *t2, *lam2, *occ, *red;
  double *apow[kMaxLevels];
};

int ortho_work(pw_waveset *ws, OrthoWork &w) {
  const size_t n2 = (size_t)ws->nb * ws->nb;
  const size_t nblk = (n2 + 255) / 256;
  const size_t need = (10 + kMaxLevels) * n2 + ws->nb + 2 * nblk + 64;
  if (need > ws->ortho_cap) {
    if (ws->d_ortho) PW_HIP(hipFree(ws->d_ortho));
    PW_HIP(hipMalloc(&ws->d_ortho, need * sizeof(double)));
    ws->ortho_cap = need;
  }
  double *p = ws->d_ortho;
  w.bs = p; p += n2; w.m0 = p; p += n2; w.m1 = p; p += n2; w.gam = p; p += n2;
  w.x0 = p; p += n2; w.x1 = p; p += n2; w.t1 = p; p += n2; w.t2 = p; p += n2;
  w.lam2 = p; p += n2; p += n2;
  for (int j = 0; j < kMaxLevels; ++j) { w.apow[j] = p; p += n2; }
  w.occ = p; p += ws->nb; w.red = p;
  return PW_OK;
}

int gemm(pw_waveset *ws, const double *A, int ta, const double *B, int tb, double *C,
         double alpha = 1.0, const double *D = nullptr, double beta = 0.0) {
  const int n = ws->nb;
  dim3 grid((n + T16 - 1) / T16, (n + T16 - 1) / T16);
  const size_t lds = (size_t)(T16 * (n + 1) + n * (T16 + 1)) * sizeof(double);
  if (lds > 65536) {
    static bool raised = false;
    if (!raised) {
      PW_HIP(hipFuncSetAttribute((const void *)k_gemm, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
      raised = true;
    }
  }
  k_gemm<<<grid, 256, lds, ws->stream>>>(n, A, ta, B, tb, C, alpha, D, beta);
  PW_HIP(hipGetLastError());
  return PW_OK;
}

// Extreme eigenvalues of a symmetric matrix (row-major) by Lanczos with full
// re-orthogonalisation; returns inner (Ritz) bounds.
void host_lanczos_bounds(int n, const std::vector<double> &a, double *lo, double *hi) {
  const int m = std::min(n, 40);
  std::vector<double> V((size_t)(m + 1) * n, 0.0), alpha(m, 0.0), beta(m + 1, 0.0), w(n);
  // start vector with all components (deterministic)
  double nrm = 0.0;
  for (int i = 0; i < n; ++i) { V[i] = 1.0 + 0.37 * std::sin(1.7 * i + 0.3); nrm += V[i] * V[i]; }
  nrm = std::sqrt(nrm);
  for (int i = 0; i < n; ++i) V[i] /= nrm;
  int k = 0;
  for (; k < m; ++k) {
    const double *v = &V[(size_t)k * n];
    for (int i = 0; i < n; ++i) {
      double s = 0.0;
      const double *ar = &a[(size_t)i * n];
      for (int j = 0; j < n; ++j) s += ar[j] * v[j];
      w[i] = s;
    }
    double al = 0.0;
    for (int i = 0; i < n; ++i) al += w[i] * v[i];
    alpha[k] = al;
    for (int rep = 0; rep < 2; ++rep)
      for (int j = 0; j <= k; ++j) {
        const double *vj = &V[(size_t)j * n];
        double c = 0.0;
        for (int i = 0; i < n; ++i) c += w[i] * vj[i];
        for (int i = 0; i < n; ++i) w[i] -= c * vj[i];
      }
    double b = 0.0;
    for (int i = 0; i < n; ++i) b += w[i] * w[i];
    b = std::sqrt(b);
    beta[k + 1] = b;
    if (b < 1.0e-14 * std::fabs(al) || k + 1 == m) { ++k; break; }
    double *vn = &V[(size_t)(k + 1) * n];
    for (int i = 0; i < n; ++i) vn[i] = w[i] / b;
  }
  std::vector<double> t((size_t)k * k, 0.0), e, u;
  for (int i = 0; i < k; ++i) {
    t[(size_t)i * k + i] = alpha[i];
    if (i + 1 < k) t[(size_t)i * k + i + 1] = t[(size_t)(i + 1) * k + i] = beta[i + 1];
  }
  host_syev(k, t, e, u);
  *lo = *std::min_element(e.begin(), e.end());
  *hi = *std::max_element(e.begin(), e.end());
}

}  // namespace

int ortho_x_device(pw_waveset *ws, double *d_lam, const double *d_chichi, const double *d_psipsi,
                   const double *d_chipsi, int *iter) {
  ScopeTimer tm(ws, "ortho_x");
  const int n = ws->nb;
  const size_t n2 = (size_t)n * n;
  const int nblk = (int)((n2 + 255) / 256);
  OrthoWork w;
  PW_TRY(ortho_work(ws, w));
  PW_HIP(hipMemcpyAsync(w.occ, ws->occ.data(), n * sizeof(double), hipMemcpyHostToDevice, ws->stream));
  k_symm<<<nblk, 256, 0, ws->stream>>>(n, d_chipsi, w.bs);
  PW_HIP(hipGetLastError());
  // spectral bounds of B on the host (inner Ritz bounds, widened by 3%)
  std::vector<double> hb(n2);
  PW_HIP(hipMemcpyAsync(hb.data(), w.bs, n2 * sizeof(double), hipMemcpyDeviceToHost, ws->stream));
  PW_HIP(hipStreamSynchronize(ws->stream));
  double lo, hi;
  host_lanczos_bounds(n, hb, &lo, &hi);
  if (!(lo > 0.0) || !(hi >= lo))
    return pw_fail(PW_ERR_NOCONV, "WAVES_ORTHO_X: sym(CHIPSI) is not positive definite");
  lo *= 0.97;
  hi *= 1.03;
  const double q = std::sqrt(lo * hi);
  const double kap = hi / lo;
  const double rho = (std::sqrt(kap) - 1.0) / (std::sqrt(kap) + 1.0);
  int nlev = 1;
  if (rho > 1.0e-300) {
    const double need = std::log(1.0e-17) / std::log(rho);
    nlev = std::max(1, (int)std::ceil(std::log2(std::max(need, 1.0)))) + 1;
  }
  if (nlev > kMaxLevels) return pw_fail(PW_ERR_NOCONV, "WAVES_ORTHO_X: CHIPSI too ill-conditioned");
  // M = (B + qI)^-1 by Newton-Schulz, X0 = 2/(lo+hi+2q) I
  const double kp = (hi + q) / (lo + q);
  const double e0 = (kp - 1.0) / (kp + 1.0);
  int nns = 1;
  if (e0 > 1.0e-300) nns = (int)std::ceil(std::log2(std::max(1.0, std::log(1.0e-18) / std::log(e0)))) + 1;
  ws->ortho_levels = nlev;
  ws->ortho_ns = nns;
  k_scaled_identity<<<nblk, 256, 0, ws->stream>>>(n, 2.0 / (lo + hi + 2.0 * q), w.m0);
  double *mc = w.m0, *mn = w.m1;
  for (int k = 0; k < nns; ++k) {
    PW_TRY(gemm(ws, w.bs, 0, mc, 0, w.t1, 1.0, mc, q));       // T = (B+qI) X
    PW_TRY(gemm(ws, mc, 0, w.t1, 0, mn, -1.0, mc, 2.0));      // X' = 2X - X T
    std::swap(mc, mn);
  }
  // A^(2^j), j = 0..nlev-1
  k_cayley<<<nblk, 256, 0, ws->stream>>>(n, q, mc, w.apow[0]);
  for (int j = 1; j < nlev; ++j) PW_TRY(gemm(ws, w.apow[j - 1], 0, w.apow[j - 1], 0, w.apow[j]));
  PW_HIP(hipGetLastError());
  const double eps = 1.0e-8;
  std::vector<double> hm(nblk);
  for (int it = 1; it <= 200; ++it) {
    // HAUX = CHICHI LAMBDA + 2 CHIPSI ; T = LAMBDA^T HAUX
    PW_TRY(gemm(ws, d_chichi, 0, d_lam, 0, w.t1, 1.0, d_chipsi, 2.0));
    PW_TRY(gemm(ws, d_lam, 1, w.t1, 0, w.t2));
    k_gamma<<<nblk, 256, 0, ws->stream>>>(n, w.t2, d_psipsi, w.gam, w.red);
    PW_HIP(hipGetLastError());
    PW_HIP(hipMemcpyAsync(hm.data(), w.red, nblk * sizeof(double), hipMemcpyDeviceToHost, ws->stream));
    PW_HIP(hipStreamSynchronize(ws->stream));
    double digam = 0.0;
    for (double v : hm) digam = std::max(digam, v);
    if (!(digam == digam)) return pw_fail(PW_ERR_NOCONV, "WAVES_ORTHO_X: NaN in overlap");
    if (digam < eps) {
      *iter = it;
      return PW_OK;
    }
    if (digam > 1.0e200)
      return pw_fail(PW_ERR_NOCONV, "WAVES_ORTHO_X: iteration loop for orthogonalization is about to diverge");
    // X = S^-1(GAMN): X0 = 2q M G M ; X <- X + A_j X A_j
    PW_TRY(gemm(ws, w.gam, 0, mc, 0, w.t1));
    PW_TRY(gemm(ws, mc, 0, w.t1, 0, w.x0, 2.0 * q));
    double *xc = w.x0, *xn = w.x1;
    for (int j = 0; j < nlev; ++j) {
      PW_TRY(gemm(ws, xc, 0, w.apow[j], 0, w.t1));
      PW_TRY(gemm(ws, w.apow[j], 0, w.t1, 0, xn, 1.0, xc, 1.0));
      std::swap(xc, xn);
    }
    k_lam_update<<<nblk, 256, 0, ws->stream>>>(n, w.occ, xc, d_lam);
    k_lam_sym<<<nblk, 256, 0, ws->stream>>>(n, w.occ, d_lam, w.lam2);
    PW_HIP(hipMemcpyAsync(d_lam, w.lam2, n2 * sizeof(double), hipMemcpyDeviceToDevice, ws->stream));
    PW_HIP(hipGetLastError());
  }
  return pw_fail(PW_ERR_NOCONV,
                 "WAVES_ORTHO_X: loop for orthogonalization is not converged (specify SAFEORTHO=F, "
                 "reduce the cutoff or the time step)");
}

// ---------------------------------------------------------------------------
// C ABI: WAVES_ORTHO_X on host matrices (column-major NB x NB)
// ---------------------------------------------------------------------------
extern "C" int pw_ortho_x(int nb, const double *occ, const double *chichi, const double *psipsi,
                          const double *chipsi, double *lambda, int *iter) {
  if (nb <= 0 || !occ || !chichi || !psipsi || !chipsi || !lambda || !iter)
    return pw_fail(PW_ERR_ARG, "pw_ortho_x: bad argument");
  pw_waveset ws;
  ws.nb = nb;
  ws.occ.assign(occ, occ + nb);
  const size_t n2 = (size_t)nb * nb;
  std::vector<double> h(4 * n2);
  const double *src[4] = {chichi, psipsi, chipsi, lambda};
  for (int m = 0; m < 4; ++m)
    for (int i = 0; i < nb; ++i)
      for (int j = 0; j < nb; ++j) h[m * n2 + (size_t)i * nb + j] = src[m][(size_t)j * nb + i];
  double *d = nullptr;
  PW_HIP(hipMalloc(&d, 4 * n2 * sizeof(double)));
  int rc = PW_OK;
  if (hipMemcpy(d, h.data(), 4 * n2 * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
    rc = pw_fail(PW_ERR_HIP, "pw_ortho_x: upload failed");
  if (rc == PW_OK) rc = ortho_x_device(&ws, d + 3 * n2, d, d + n2, d + 2 * n2, iter);
  if (rc == PW_OK && hipMemcpy(h.data() + 3 * n2, d + 3 * n2, n2 * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
    rc = pw_fail(PW_ERR_HIP, "pw_ortho_x: download failed");
  if (rc == PW_OK)
    for (int i = 0; i < nb; ++i)
      for (int j = 0; j < nb; ++j) lambda[(size_t)j * nb + i] = h[3 * n2 + (size_t)i * nb + j];
  (void)hipFree(d);
  if (ws.d_ortho) (void)hipFree(ws.d_ortho);
  return rc;
}
