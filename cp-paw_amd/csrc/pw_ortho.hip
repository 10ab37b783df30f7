// WAVES_ORTHO_X (src/paw_waves2.f90:1844-2022) in FP64 on the GPU.
//
// The reference iterates
//   GAMN   = sym[LAMBDA^T (CHICHI LAMBDA + 2 CHIPSI)] + PSIPSI - 1
//   LAMBDA -= 2 f_i/(f_i+f_j) * S^-1(GAMN),   S(X) = B X + X B,  B = sym(CHIPSI)
// until max|GAMN| < 1e-8, with S^-1 applied through the eigenvectors of B
// (DSYEV).  The iterates depend on S^-1 only, not on the eigen-basis, so this
// implementation keeps the reference's outer iteration and stopping rule and
// applies S^-1 exactly (to rounding) in a slowly varying basis U:
//   U^T B U = D + E (E off-diagonal, small once U is an eigen-basis of a
//   nearby B), Y solves (D+E) Y + Y (D+E) = U^T GAMN U by the contraction
//   Y <- (Gt - E Y - (E Y)^T) / (d_i + d_j), and S^-1(GAMN) = U Y U^T.
// U is refreshed with a host Jacobi eigen-decomposition of B when the
// contraction bound ||E||_F / min(d) exceeds 0.05 (first call, or when B has
// drifted).  Every matrix product is an NB^3 GEMM on the device.
#include <algorithm>
#include <cmath>

#include "pw_common.h"

namespace {

constexpr int T16 = 16;

// C = alpha op(A) op(B) + beta D   (row-major n x n; D may alias nothing)
__global__ void __launch_bounds__(256) k_gemm(int n, const double *A, int ta, const double *B, int tb,
                                              double *C, double alpha, const double *D, double beta) {
  __shared__ double sa[T16][T16 + 1];
  __shared__ double sb[T16][T16 + 1];
  const int ty = threadIdx.x / T16, tx = threadIdx.x % T16;
  const int row = blockIdx.y * T16 + ty, col = blockIdx.x * T16 + tx;
  double acc = 0.0;
  for (int k0 = 0; k0 < n; k0 += T16) {
    {
      const int r = blockIdx.y * T16 + ty, k = k0 + tx;   // A(r,k)
      double v = 0.0;
      if (r < n && k < n) v = ta ? A[(size_t)k * n + r] : A[(size_t)r * n + k];
      sa[ty][tx] = v;
    }
    {
      const int k = k0 + ty, c = blockIdx.x * T16 + tx;   // B(k,c)
      double v = 0.0;
      if (k < n && c < n) v = tb ? B[(size_t)c * n + k] : B[(size_t)k * n + c];
      sb[ty][tx] = v;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < T16; ++k) acc += sa[ty][k] * sb[k][tx];
    __syncthreads();
  }
  if (row < n && col < n) {
    double v = alpha * acc;
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

// d = diag(Bt), E = Bt - diag(d); per-block sum of E^2 and min d
__global__ void k_split(int n, const double *Bt, double *d, double *E, double *bsum, double *bmin) {
  __shared__ double rs[256], rm[256];
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  double s = 0.0, mn = 1.0e300;
  if (idx < n * n) {
    const int i = idx / n, j = idx % n;
    const double v = Bt[idx];
    if (i == j) { d[i] = v; E[idx] = 0.0; mn = v; }
    else { E[idx] = v; s = v * v; }
  }
  rs[threadIdx.x] = s;
  rm[threadIdx.x] = mn;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) {
      rs[threadIdx.x] += rs[threadIdx.x + k];
      rm[threadIdx.x] = fmin(rm[threadIdx.x], rm[threadIdx.x + k]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) { bsum[blockIdx.x] = rs[0]; bmin[blockIdx.x] = rm[0]; }
}

__global__ void k_sylv(int n, const double *Gt, const double *EY, const double *d, double *Y) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n * n) return;
  const int i = idx / n, j = idx % n;
  double r = Gt[idx];
  if (EY) r -= EY[idx] + EY[(size_t)j * n + i];
  Y[idx] = r / (d[i] + d[j]);
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

struct OrthoWork {
  double *bs, *bt, *E, *d, *gam, *gt, *y, *ey, *x, *t1, *t2, *lam2, *occ, *red;
};

int ortho_work(pw_waveset *ws, OrthoWork &w) {
  const size_t n2 = (size_t)ws->nb * ws->nb;
  const size_t nblk = (n2 + 255) / 256;
  const size_t need = 12 * n2 + 2 * ws->nb + 2 * nblk + 64;
  static thread_local double *buf = nullptr;
  static thread_local size_t cap = 0;
  if (need > cap) {
    if (buf) PW_HIP(hipFree(buf));
    PW_HIP(hipMalloc(&buf, need * sizeof(double)));
    cap = need;
  }
  double *p = buf;
  w.bs = p; p += n2; w.bt = p; p += n2; w.E = p; p += n2; w.gam = p; p += n2;
  w.gt = p; p += n2; w.y = p; p += n2; w.ey = p; p += n2; w.x = p; p += n2;
  w.t1 = p; p += n2; w.t2 = p; p += n2; w.lam2 = p; p += n2; p += n2;
  w.d = p; p += ws->nb; w.occ = p; p += ws->nb; w.red = p;
  return PW_OK;
}

int gemm(pw_waveset *ws, const double *A, int ta, const double *B, int tb, double *C,
         double alpha = 1.0, const double *D = nullptr, double beta = 0.0) {
  const int n = ws->nb;
  dim3 grid((n + T16 - 1) / T16, (n + T16 - 1) / T16);
  k_gemm<<<grid, 256, 0, ws->stream>>>(n, A, ta, B, tb, C, alpha, D, beta);
  PW_HIP(hipGetLastError());
  return PW_OK;
}

int refresh_basis(pw_waveset *ws, const double *d_bs) {
  const int n = ws->nb;
  std::vector<double> a((size_t)n * n), e, u;
  PW_HIP(hipMemcpyAsync(a.data(), d_bs, a.size() * sizeof(double), hipMemcpyDeviceToHost, ws->stream));
  PW_HIP(hipStreamSynchronize(ws->stream));
  host_syev(n, a, e, u);
  if (!ws->d_ux) PW_HIP(hipMalloc(&ws->d_ux, (size_t)n * n * sizeof(double)));
  PW_HIP(hipMemcpyAsync(ws->d_ux, u.data(), u.size() * sizeof(double), hipMemcpyHostToDevice, ws->stream));
  PW_HIP(hipStreamSynchronize(ws->stream));
  ws->have_basis = true;
  ws->eig_refresh++;
  return PW_OK;
}

// rotated split of Bs; returns the contraction bound
int split_basis(pw_waveset *ws, OrthoWork &w, double *rho) {
  const int n = ws->nb;
  const int nblk = (n * n + 255) / 256;
  PW_TRY(gemm(ws, ws->d_ux, 1, w.bs, 0, w.t1));
  PW_TRY(gemm(ws, w.t1, 0, ws->d_ux, 0, w.bt));
  k_split<<<nblk, 256, 0, ws->stream>>>(n, w.bt, w.d, w.E, w.red, w.red + nblk);
  PW_HIP(hipGetLastError());
  std::vector<double> h(2 * nblk);
  PW_HIP(hipMemcpyAsync(h.data(), w.red, h.size() * sizeof(double), hipMemcpyDeviceToHost, ws->stream));
  PW_HIP(hipStreamSynchronize(ws->stream));
  double s = 0.0, mn = 1.0e300;
  for (int k = 0; k < nblk; ++k) { s += h[k]; mn = std::min(mn, h[nblk + k]); }
  *rho = mn > 0.0 ? std::sqrt(s) / mn : 1.0e300;
  return PW_OK;
}

}  // namespace

int ortho_x_device(pw_waveset *ws, double *d_lam, const double *d_chichi, const double *d_psipsi,
                   const double *d_chipsi, int *iter) {
  ScopeTimer tm(ws, "ortho_x");
  const int n = ws->nb;
  const int nblk = (n * n + 255) / 256;
  OrthoWork w;
  PW_TRY(ortho_work(ws, w));
  PW_HIP(hipMemcpyAsync(w.occ, ws->occ.data(), n * sizeof(double), hipMemcpyHostToDevice, ws->stream));
  k_symm<<<nblk, 256, 0, ws->stream>>>(n, d_chipsi, w.bs);
  PW_HIP(hipGetLastError());
  double rho = 1.0e300;
  if (ws->have_basis) PW_TRY(split_basis(ws, w, &rho));
  if (!(rho <= 0.05)) {
    PW_TRY(refresh_basis(ws, w.bs));
    PW_TRY(split_basis(ws, w, &rho));
    if (!(rho < 0.5))
      return pw_fail(PW_ERR_NOCONV, "WAVES_ORTHO_X: CHIPSI is not positive definite");
  }
  // inner iterations needed for a 1e-17 relative fixed point
  int ninner = 1;
  if (rho > 1.0e-17) ninner = std::min(80, (int)std::ceil(std::log(1.0e-17) / std::log(rho)) + 1);
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
    // Gt = U^T GAMN U ; Y = S^-1 in the rotated basis ; X = U Y U^T
    PW_TRY(gemm(ws, ws->d_ux, 1, w.gam, 0, w.t1));
    PW_TRY(gemm(ws, w.t1, 0, ws->d_ux, 0, w.gt));
    k_sylv<<<nblk, 256, 0, ws->stream>>>(n, w.gt, nullptr, w.d, w.y);
    for (int k = 0; k < ninner; ++k) {
      PW_TRY(gemm(ws, w.E, 0, w.y, 0, w.ey));
      k_sylv<<<nblk, 256, 0, ws->stream>>>(n, w.gt, w.ey, w.d, w.y);
    }
    PW_TRY(gemm(ws, ws->d_ux, 0, w.y, 0, w.t1));
    PW_TRY(gemm(ws, w.t1, 0, ws->d_ux, 1, w.x));
    k_lam_update<<<nblk, 256, 0, ws->stream>>>(n, w.occ, w.x, d_lam);
    k_lam_sym<<<nblk, 256, 0, ws->stream>>>(n, w.occ, d_lam, w.lam2);
    PW_HIP(hipMemcpyAsync(d_lam, w.lam2, (size_t)n * n * sizeof(double), hipMemcpyDeviceToDevice,
                          ws->stream));
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
  if (ws.d_ux) (void)hipFree(ws.d_ux);
  return rc;
}
