// PLANEWAVE instance: real-space grid, G sphere, index maps, FFT set-up.
//
// Restates src/paw_fft.f90:823-1110 (DIVIDERGRIDONTASKS, INITIALIZE,
// COUNTG/GETIG, MINUSG) and src/paw_library.f90:2336-2393 (FFTADJUSTGRD),
// src/paw_generalpurpose.f90:72-201 (GBASS, BOXSPH) for one task.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <numeric>
#include <unordered_map>

#include "pw_common.h"

static thread_local std::string g_last_error;

void pw_set_error(const std::string &msg) { g_last_error = msg; }
int pw_fail(int code, const std::string &msg) {
  g_last_error = msg;
  return code;
}

extern "C" int pw_version(void) { return 1; }

extern "C" int pw_last_error(char *buf, int len) {
  if (!buf || len <= 0) return PW_ERR_ARG;
  std::strncpy(buf, g_last_error.c_str(), (size_t)len - 1);
  buf[len - 1] = 0;
  return PW_OK;
}

extern "C" int pw_malloc(void **ptr, long long nbytes) {
  if (!ptr || nbytes < 0) return pw_fail(PW_ERR_ARG, "pw_malloc: bad argument");
  PW_HIP(hipMalloc(ptr, (size_t)std::max<long long>(nbytes, 1)));
  return PW_OK;
}
extern "C" int pw_free(void *ptr) {
  if (ptr) PW_HIP(hipFree(ptr));
  return PW_OK;
}
extern "C" int pw_memcpy_htod(void *dst, const void *src, long long n) {
  PW_HIP(hipMemcpy(dst, src, (size_t)n, hipMemcpyHostToDevice));
  return PW_OK;
}
extern "C" int pw_memcpy_dtoh(void *dst, const void *src, long long n) {
  PW_HIP(hipMemcpy(dst, src, (size_t)n, hipMemcpyDeviceToHost));
  return PW_OK;
}
extern "C" int pw_device_synchronize(void) {
  PW_HIP(hipDeviceSynchronize());
  return PW_OK;
}

// ---------------------------------------------------------------------------
// grid rule
// ---------------------------------------------------------------------------
static void gbass(const double *r, double *g, double &vol) {
  // r[i+3j] = RBAS(i+1,j+1); same cofactor formulas as GBASS
  auto R = [&](int i, int j) { return r[(i - 1) + 3 * (j - 1)]; };
  double G[3][3];
  G[0][0] = R(2, 2) * R(3, 3) - R(3, 2) * R(2, 3);
  G[1][0] = R(3, 2) * R(1, 3) - R(1, 2) * R(3, 3);
  G[2][0] = R(1, 2) * R(2, 3) - R(2, 2) * R(1, 3);
  G[0][1] = R(2, 3) * R(3, 1) - R(3, 3) * R(2, 1);
  G[1][1] = R(3, 3) * R(1, 1) - R(1, 3) * R(3, 1);
  G[2][1] = R(1, 3) * R(2, 1) - R(2, 3) * R(1, 1);
  G[0][2] = R(2, 1) * R(3, 2) - R(3, 1) * R(2, 2);
  G[1][2] = R(3, 1) * R(1, 2) - R(1, 1) * R(3, 2);
  G[2][2] = R(1, 1) * R(2, 2) - R(2, 1) * R(1, 2);
  double det = R(1, 1) * G[0][0] + R(2, 1) * G[1][0] + R(3, 1) * G[2][0];
  double fac = 2.0 * M_PI / det;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) g[i + 3 * j] = G[i][j] * fac;
  vol = std::fabs(det);
}

static void boxsph(const double *t, double x0, double y0, double z0, double rmax,
                   int mn[3], int mx[3]) {
  // t[i+3j]: column j is a displacement vector
  const double *t1 = t, *t2 = t + 3, *t3 = t + 6;
  double g1[3] = {t2[1] * t3[2] - t2[2] * t3[1], t2[2] * t3[0] - t2[0] * t3[2],
                  t2[0] * t3[1] - t2[1] * t3[0]};
  double g2[3] = {t3[1] * t1[2] - t3[2] * t1[1], t3[2] * t1[0] - t3[0] * t1[2],
                  t3[0] * t1[1] - t3[1] * t1[0]};
  double g3[3] = {t1[1] * t2[2] - t1[2] * t2[1], t1[2] * t2[0] - t1[0] * t2[2],
                  t1[0] * t2[1] - t1[1] * t2[0]};
  double det = t1[0] * g1[0] + t1[1] * g1[1] + t1[2] * g1[2];
  for (int i = 0; i < 3; ++i) { g1[i] /= det; g2[i] /= det; g3[i] /= det; }
  double xn[3] = {std::sqrt(g1[0] * g1[0] + g1[1] * g1[1] + g1[2] * g1[2]) * rmax,
                  std::sqrt(g2[0] * g2[0] + g2[1] * g2[1] + g2[2] * g2[2]) * rmax,
                  std::sqrt(g3[0] * g3[0] + g3[1] * g3[1] + g3[2] * g3[2]) * rmax};
  double p0[3] = {g1[0] * x0 + g1[1] * y0 + g1[2] * z0, g2[0] * x0 + g2[1] * y0 + g2[2] * z0,
                  g3[0] * x0 + g3[1] * y0 + g3[2] * z0};
  const double ibig = 1000000.0;
  for (int k = 0; k < 3; ++k) {
    mn[k] = (int)(long long)(p0[k] - xn[k] + 1.0 + ibig) - 1000000;
    mx[k] = (int)(long long)(p0[k] + xn[k] + ibig) - 1000000;
  }
}

static std::vector<int> allowed_lengths() {
  std::vector<int> out;
  for (int h = 1; h <= 25; ++h)
    for (int i = 0; i <= 2; ++i)
      for (int j = 0; j <= 1; ++j)
        for (int k = 0; k <= 1; ++k)
          for (int m = 0; m <= 1; ++m) {
            if ((int)out.size() >= 300) goto done;
            double v = std::pow(2.0, h) * std::pow(3.0, i) * std::pow(5.0, j) *
                       std::pow(7.0, k) * std::pow(11.0, m);
            if (v > 37748736.0) continue;
            out.push_back((int)std::lround(v));
          }
done:
  std::sort(out.begin(), out.end());
  return out;
}

static int fftadjustgrd(int nr) {
  static const std::vector<int> ifr = allowed_lengths();
  for (size_t i = 1; i < ifr.size(); ++i)
    if (ifr[i] >= nr) return ifr[i];
  return -1;
}

// Position of element e in the digit-reversed order of a mixed-radix
// decimation-in-time transform with the stage radices f.radix[0..nstage):
// the element index of position p reads p's digits (most significant digit
// for the last stage) from the least significant end.
std::vector<int> digit_reversal_positions(const Fft1D &f) {
  std::vector<int> pos(f.n);
  for (int p = 0; p < f.n; ++p) {
    int rem = p, size = f.n, idx = 0, mult = 1;
    for (int st = f.nstage - 1; st >= 0; --st) {
      size /= f.radix[st];
      const int d = rem / size;
      rem %= size;
      idx += d * mult;
      mult *= f.radix[st];
    }
    pos[idx] = p;
  }
  return pos;
}

bool fft_factor(int n, Fft1D &f, bool inplace) {
  f = Fft1D();
  f.n = n;
  int m = n;
  // Odd radices first, then the largest powers of two (16 only for the
  // Stockham passes; its butterfly doubles the registers of the in-place ones)
  const int order_s[] = {9, 3, 5, 7, 11, 16, 8, 4, 2};
  const int order_i[] = {9, 3, 5, 7, 11, 8, 4, 2};
  const int *order = inplace ? order_i : order_s;
  const int norder = inplace ? 8 : 9;
  for (int io = 0; io < norder; ++io) {
    const int r = order[io];
    while (m % r == 0) {
      f.radix[f.nstage++] = r;
      m /= r;
      if (f.nstage >= 12) return false;
    }
  }
  int L = 1;
  for (int st = 0; st < f.nstage; ++st) {
    f.mnbf[st] = fdiv_magic((unsigned)(n / f.radix[st]));
    f.mlp[st] = fdiv_magic((unsigned)L);
    L *= f.radix[st];
  }
  return m == 1;
}

// ---------------------------------------------------------------------------
// plan creation
// ---------------------------------------------------------------------------
static int build_plan(pw_plan *p, const double *rbas, double epw_psi, double epw_rho) {
  std::memcpy(p->rbas, rbas, 9 * sizeof(double));
  gbass(rbas, p->gbas, p->vol);
  if (!(p->vol > 0.0)) return pw_fail(PW_ERR_ARG, "GBASS: lattice vectors linearly dependent");
  // DIVIDERGRIDONTASKS
  int mn[3], mx[3];
  boxsph(p->gbas, 0.0, 0.0, 0.0, std::sqrt(2.0 * epw_rho) + 1.0e-6, mn, mx);
  p->nr1 = fftadjustgrd(mx[0] - mn[0] + 1);
  p->nr2 = fftadjustgrd(mx[1] - mn[1] + 1);
  p->nr3 = fftadjustgrd(mx[2] - mn[2] + 1);
  if (p->nr1 <= 0 || p->nr2 <= 0 || p->nr3 <= 0)
    return pw_fail(PW_ERR_ARG, "PLANEWAVE$ADJUSTFFTGRD: requested gridpoint out of range");
  p->nrl = p->nr1 * p->nr2 * p->nr3;
  for (int k = 0; k < 3; ++k)
    if (!fft_factor(k == 0 ? p->nr1 : (k == 1 ? p->nr2 : p->nr3), p->f[k]))
      return pw_fail(PW_ERR_UNSUPPORTED, "FFT length not factorisable");
  if (!fft_factor(p->nr2, p->fyz[0], true) || !fft_factor(p->nr3, p->fyz[1], true))
    return pw_fail(PW_ERR_UNSUPPORTED, "FFT length not factorisable");
  // GETIG
  boxsph(p->gbas, 0.0, 0.0, 0.0, std::sqrt(2.0 * epw_psi) + 1.0e-6, mn, mx);
  const double *G = p->gbas;
  std::vector<int> ig;
  std::vector<double> g2a;
  for (int i = mn[0]; i <= mx[0]; ++i) {
    double t1 = i;
    for (int j = mn[1]; j <= mx[1]; ++j) {
      double t2 = j;
      for (int k = mn[2]; k <= mx[2]; ++k) {
        double t3 = k;
        double g1 = 0.0 + G[0] * t1 + G[3] * t2 + G[6] * t3;
        double gy = 0.0 + G[1] * t1 + G[4] * t2 + G[7] * t3;
        double g3 = 0.0 + G[2] * t1 + G[5] * t2 + G[8] * t3;
        double gs = g1 * g1 + gy * gy + g3 * g3;
        if (0.5 * gs > epw_psi) continue;
        ig.push_back(i); ig.push_back(j); ig.push_back(k);
        g2a.push_back(gs);
      }
    }
  }
  int ng = (int)g2a.size();
  std::vector<int> ord(ng);
  std::iota(ord.begin(), ord.end(), 0);
  std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return g2a[a] < g2a[b]; });
  std::vector<int> sig(3 * ng);
  std::vector<double> sg2(ng);
  for (int n = 0; n < ng; ++n) {
    for (int c = 0; c < 3; ++c) sig[3 * n + c] = ig[3 * ord[n] + c];
    sg2[n] = g2a[ord[n]];
  }
  // runs of |G|^2 within 1e-5 of the run's first element, sorted
  // lexicographically; the last element always joins the current run
  // (src/paw_fft.f90:1284-1324)
  int istart = 0;
  while (istart < ng) {
    double svar = sg2[istart];
    int stop = istart + 1;
    while (stop < ng && std::fabs(sg2[stop] - svar) < 1.0e-5) ++stop;
    if (stop == ng - 1) stop = ng;
    std::vector<std::array<int, 3>> run;
    for (int n = istart; n < stop; ++n) run.push_back({sig[3 * n], sig[3 * n + 1], sig[3 * n + 2]});
    std::sort(run.begin(), run.end());
    for (int n = istart; n < stop; ++n)
      for (int c = 0; c < 3; ++c) sig[3 * n + c] = run[n - istart][c];
    istart = stop;
  }
  p->ngl = ng;
  p->igvec = sig;
  // MINUSG (Gamma point: -G)
  std::unordered_map<long long, int> key;
  auto pack = [](int a, int b, int c) {
    return ((long long)(a + 4096) << 26) | ((long long)(b + 4096) << 13) | (long long)(c + 4096);
  };
  for (int n = 0; n < ng; ++n) key[pack(sig[3 * n], sig[3 * n + 1], sig[3 * n + 2])] = n;
  p->minusg.assign(ng, -1);
  for (int n = 0; n < ng; ++n) {
    auto it = key.find(pack(-sig[3 * n], -sig[3 * n + 1], -sig[3 * n + 2]));
    if (it == key.end()) return pw_fail(PW_ERR_ARG, "MINUSG: inversion symmetric partner not found");
    p->minusg[n] = it->second;
    if (it->second == n) p->ngamma = n;
  }
  // G vectors and |G|^2 (PLANEWAVE_GVECTORS, GETR8A('G2'))
  p->gvec.resize(3 * ng);
  p->g2.resize(ng);
  for (int n = 0; n < ng; ++n) {
    double t1 = sig[3 * n], t2 = sig[3 * n + 1], t3 = sig[3 * n + 2];
    double gx = 0.0 + G[0] * t1 + G[3] * t2 + G[6] * t3;
    double gy = 0.0 + G[1] * t1 + G[4] * t2 + G[7] * t3;
    double gz = 0.0 + G[2] * t1 + G[5] * t2 + G[8] * t3;
    p->gvec[3 * n] = gx; p->gvec[3 * n + 1] = gy; p->gvec[3 * n + 2] = gz;
    p->g2[n] = gx * gx + gy * gy + gz * gz;
    for (int c = 0; c < 3; ++c) p->nmax[c] = std::max(p->nmax[c], std::abs(sig[3 * n + c]));
  }
  // stripes (src/paw_fft.f90:1063-1073 with one task): lines (y,z) along x
  p->yz2stripe.assign((size_t)p->nr2 * p->nr3, -1);
  std::vector<int> cnt;
  for (int n = 0; n < ng; ++n) {
    int y = ((sig[3 * n + 1] % p->nr2) + p->nr2) % p->nr2;
    int z = ((sig[3 * n + 2] % p->nr3) + p->nr3) % p->nr3;
    int &s = p->yz2stripe[y + p->nr2 * z];
    if (s < 0) { s = (int)cnt.size(); cnt.push_back(0); }
    cnt[s]++;
  }
  p->nstripe = (int)cnt.size();
  p->stripe_off.assign(p->nstripe + 1, 0);
  for (int s = 0; s < p->nstripe; ++s) p->stripe_off[s + 1] = p->stripe_off[s] + cnt[s];
  p->stripe_g.assign(ng, 0);
  p->stripe_x.assign(ng, 0);
  p->stripe_s.assign(ng, 0);
  std::vector<int> fill(p->stripe_off.begin(), p->stripe_off.end() - 1);
  for (int n = 0; n < ng; ++n) {
    int x = ((sig[3 * n] % p->nr1) + p->nr1) % p->nr1;
    int y = ((sig[3 * n + 1] % p->nr2) + p->nr2) % p->nr2;
    int z = ((sig[3 * n + 2] % p->nr3) + p->nr3) % p->nr3;
    int s = p->yz2stripe[y + p->nr2 * z];
    p->stripe_g[fill[s]] = n;
    p->stripe_x[fill[s]] = x;
    p->stripe_s[fill[s]] = s;
    fill[s]++;
  }
  p->z2iz.assign(p->nr3, -1);
  p->nzp = 0;
  for (int z = 0; z < p->nr3; ++z) {
    bool any = false;
    for (int y = 0; y < p->nr2 && !any; ++y) any = p->yz2stripe[y + p->nr2 * z] >= 0;
    if (any) {
      p->z2iz[z] = p->nzp++;
      p->zlist.push_back(z);
    }
  }
  // positions of the (y, z) plane in digit-reversed order, the input order
  // of the in-place decimation-in-time transforms of the yz pass
  // rows of the plane are yzpitch apart: an odd pitch puts the rows of the
  // y transforms, which neighbouring lanes work on, into distinct LDS banks
  const std::vector<int> py = digit_reversal_positions(p->fyz[0]), pz = digit_reversal_positions(p->fyz[1]);
  p->yzpitch = p->nr2 | 1;
  p->stripe_pos.assign(p->nstripe, 0);
  for (int z = 0; z < p->nr3; ++z)
    for (int y = 0; y < p->nr2; ++y) {
      const int s = p->yz2stripe[y + p->nr2 * z];
      if (s >= 0) p->stripe_pos[s] = pz[z] * p->yzpitch + py[y];
    }
  p->zpos.assign(p->nzp, 0);
  for (int i = 0; i < p->nzp; ++i) p->zpos[i] = pz[p->zlist[i]] * p->yzpitch;
  // Lane slots of the y transforms (one stripe row per lane).  A 64-bit LDS
  // access is served 16 lanes at a time over 32 four-byte banks, so 16 lanes
  // that read the same element of 16 rows are conflict-free when the row
  // offsets differ mod 16.  The k-th row whose offset is r mod 16 takes slot
  // 16 k + r; when some residue holds more rows than ceil(nzp / 16) windows
  // the rows keep their order, one slot each.
  {
    std::vector<std::vector<int>> byres(16);
    for (int i = 0; i < p->nzp; ++i) byres[p->zpos[i] % 16].push_back(p->zpos[i]);
    size_t nwin = 0;
    for (const auto &v : byres) nwin = std::max(nwin, v.size());
    if (nwin == (size_t)(p->nzp + 15) / 16) {
      p->ypen.assign(16 * nwin, -1);
      for (int r = 0; r < 16; ++r)
        for (size_t k = 0; k < byres[r].size(); ++k) p->ypen[16 * k + r] = byres[r][k];
    } else {
      p->ypen = p->zpos;
    }
  }
  return PW_OK;
}

template <typename T>
static int upload(T **dst, const std::vector<T> &src) {
  PW_HIP(hipMalloc((void **)dst, std::max<size_t>(1, src.size()) * sizeof(T)));
  if (!src.empty())
    PW_HIP(hipMemcpy(*dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice));
  return PW_OK;
}

static int upload_plan(pw_plan *p) {
  if (p->uploaded) return PW_OK;
  p->uploaded = true;
  PW_TRY(upload(&p->d_minusg, p->minusg));
  PW_TRY(upload(&p->d_igvec, p->igvec));
  PW_TRY(upload(&p->d_stripe_off, p->stripe_off));
  PW_TRY(upload(&p->d_stripe_g, p->stripe_g));
  PW_TRY(upload(&p->d_stripe_x, p->stripe_x));
  PW_TRY(upload(&p->d_stripe_s, p->stripe_s));
  PW_TRY(upload(&p->d_zlist, p->zlist));
  PW_TRY(upload(&p->d_yz2stripe, p->yz2stripe));
  PW_TRY(upload(&p->d_stripe_pos, p->stripe_pos));
  PW_TRY(upload(&p->d_zpos, p->zpos));
  PW_TRY(upload(&p->d_ypen, p->ypen));
  PW_TRY(upload(&p->d_z2iz, p->z2iz));
  PW_TRY(upload(&p->d_g2, p->g2));
  PW_TRY(upload(&p->d_gvec, p->gvec));
  for (int k = 0; k < 3; ++k) {
    int n = p->f[k].n;
    std::vector<double2> tw(n);
    std::vector<float2> twf(n);
    for (int m = 0; m < n; ++m) {
      // exp(+2 pi i m / n), evaluated in long double
      long double a = 2.0L * 3.141592653589793238462643383279502884L * (long double)m / (long double)n;
      double c = (double)std::cos(a), s = (double)std::sin(a);
      tw[m] = make_double2(c, s);
      twf[m] = make_float2((float)c, (float)s);
    }
    PW_TRY(upload(&p->d_tw[k], tw));
    PW_TRY(upload(&p->d_twf[k], twf));
  }
  return PW_OK;
}

int plan_device(pw_plan *p) { return upload_plan(p); }

int plan_work(pw_plan *p, int nbatch, int prec, bool planes) {
  PW_TRY(upload_plan(p));
  size_t es = prec == PW_PREC_FP64 ? sizeof(double2) : sizeof(float2);
  size_t sb = (size_t)nbatch * p->nstripe * p->nr1 * es;
  size_t pb = planes ? (size_t)nbatch * p->nzp * p->nr2 * p->nr1 * es : 0;
  if (sb > p->stripe_bytes) {
    if (p->d_stripe_buf) PW_HIP(hipFree(p->d_stripe_buf));
    PW_HIP(hipMalloc(&p->d_stripe_buf, sb));
    p->stripe_bytes = sb;
  }
  if (pb > p->plane_bytes) {
    if (p->d_plane_buf) PW_HIP(hipFree(p->d_plane_buf));
    PW_HIP(hipMalloc(&p->d_plane_buf, pb));
    p->plane_bytes = pb;
  }
  return PW_OK;
}

// ---------------------------------------------------------------------------
// C ABI: plan
// ---------------------------------------------------------------------------
extern "C" int pw_plan_create(const double rbas[9], const double kvec[3], int tinv,
                              double epw_psi, double epw_rho, pw_plan **out) {
  if (!rbas || !out) return pw_fail(PW_ERR_ARG, "pw_plan_create: null argument");
  if (kvec && (kvec[0] != 0.0 || kvec[1] != 0.0 || kvec[2] != 0.0))
    return pw_fail(PW_ERR_UNSUPPORTED, "pw_plan_create: only the Gamma point is supported");
  if (!tinv)
    return pw_fail(PW_ERR_UNSUPPORTED, "pw_plan_create: only super wave functions (TINV) are supported");
  if (!(epw_psi > 0.0) || !(epw_rho >= epw_psi))
    return pw_fail(PW_ERR_ARG, "pw_plan_create: need 0 < epw_psi <= epw_rho");
  {
    std::lock_guard<std::mutex> lk(capture_gate().mu);
    capture_gate().plans++;
  }
  pw_plan *p = new pw_plan();
  int rc = build_plan(p, rbas, epw_psi, epw_rho);
  if (rc != PW_OK) {
    pw_plan_destroy(p);
    return rc;
  }
  *out = p;
  return PW_OK;
}

CaptureGate &capture_gate() {
  static CaptureGate g;
  return g;
}

extern "C" int pw_plan_destroy(pw_plan *p) {
  if (!p) return PW_OK;
  void *ptrs[] = {p->d_minusg, p->d_igvec, p->d_stripe_off, p->d_stripe_g, p->d_stripe_x,
                  p->d_stripe_s, p->d_zlist, p->d_yz2stripe, p->d_z2iz, p->d_g2, p->d_gvec, p->d_stripe_buf,
                  p->d_plane_buf, p->d_dens_part, p->d_vt, p->d_stripe_pos, p->d_zpos, p->d_ypen, p->d_tw[0], p->d_tw[1], p->d_tw[2], p->d_twf[0],
                  p->d_twf[1], p->d_twf[2], p->d_pot_tabs, p->d_pot};
  for (void *q : ptrs)
    if (q) (void)hipFree(q);
  if (p->h_pot) (void)hipHostFree(p->h_pot);
  delete p;
  std::lock_guard<std::mutex> lk(capture_gate().mu);
  capture_gate().plans--;
  return PW_OK;
}

extern "C" int pw_plan_geti4(const pw_plan *p, const char *id, int *val) {
  if (!p || !id || !val) return pw_fail(PW_ERR_ARG, "pw_plan_geti4: null argument");
  std::string s(id);
  if (s == "NGL") *val = p->ngl;
  else if (s == "NRL") *val = p->nrl;
  else if (s == "NR1") *val = p->nr1;
  else if (s == "NR2") *val = p->nr2;
  else if (s == "NR3") *val = p->nr3;
  else if (s == "NGAMMA") *val = p->ngamma + 1;
  else if (s == "NSTRIPE") *val = p->nstripe;
  else if (s == "NZPLANE") *val = p->nzp;
  else return pw_fail(PW_ERR_ARG, "pw_plan_geti4: ID not recognised: " + s);
  return PW_OK;
}

extern "C" int pw_plan_getr8(const pw_plan *p, const char *id, double *val) {
  if (!p || !id || !val) return pw_fail(PW_ERR_ARG, "pw_plan_getr8: null argument");
  std::string s(id);
  if (s == "GWEIGHT") *val = p->vol;
  else if (s == "RWEIGHT") *val = p->vol / (double)p->nrl;
  else return pw_fail(PW_ERR_ARG, "pw_plan_getr8: ID not recognised: " + s);
  return PW_OK;
}

extern "C" int pw_plan_getr8a(const pw_plan *p, const char *id, int len, double *val) {
  if (!p || !id || !val) return pw_fail(PW_ERR_ARG, "pw_plan_getr8a: null argument");
  std::string s(id);
  const std::vector<double> *src = nullptr;
  if (s == "GVEC") src = &p->gvec;
  else if (s == "G2") src = &p->g2;
  else return pw_fail(PW_ERR_ARG, "pw_plan_getr8a: ID not recognised: " + s);
  if ((size_t)len != src->size()) return pw_fail(PW_ERR_ARG, "pw_plan_getr8a: size inconsistent");
  std::memcpy(val, src->data(), src->size() * sizeof(double));
  return PW_OK;
}

extern "C" int pw_plan_geti4a(const pw_plan *p, const char *id, int len, int *val) {
  if (!p || !id || !val) return pw_fail(PW_ERR_ARG, "pw_plan_geti4a: null argument");
  std::string s(id);
  if (s == "IGVEC") {
    if (len != 3 * p->ngl) return pw_fail(PW_ERR_ARG, "pw_plan_geti4a: size inconsistent");
    std::memcpy(val, p->igvec.data(), (size_t)len * sizeof(int));
  } else if (s == "MINUSG") {
    if (len != p->ngl) return pw_fail(PW_ERR_ARG, "pw_plan_geti4a: size inconsistent");
    for (int n = 0; n < p->ngl; ++n) val[n] = p->minusg[n] + 1;
  } else {
    return pw_fail(PW_ERR_ARG, "pw_plan_geti4a: ID not recognised: " + s);
  }
  return PW_OK;
}

extern "C" int pw_plan_set_stream(pw_plan *p, void *stream) {
  if (!p) return pw_fail(PW_ERR_ARG, "pw_plan_set_stream: null plan");
  p->stream = (hipStream_t)stream;
  return PW_OK;
}

extern "C" int pw_fft(pw_plan *p, const char *id, int nfft, int ngl, void *fofg, int nrl,
                      void *fofr, int prec) {
  if (!p || !id || !fofg || !fofr) return pw_fail(PW_ERR_ARG, "pw_fft: null argument");
  if (ngl != p->ngl) return pw_fail(PW_ERR_ARG, "PLANEWAVE$FFT: size of F(G) inconsistent");
  if (nrl != p->nrl) return pw_fail(PW_ERR_ARG, "PLANEWAVE$FFT: size of F(R) inconsistent");
  if (nfft <= 0) return pw_fail(PW_ERR_ARG, "PLANEWAVE$FFT: NFFT must be positive");
  if (prec != PW_PREC_FP64 && prec != PW_PREC_MIXED)
    return pw_fail(PW_ERR_ARG, "pw_fft: unknown precision");
  std::string s(id);
  if (s != "GTOR" && s != "RTOG") return pw_fail(PW_ERR_ARG, "PLANEWAVE$FFT: ID must be either GTOR or RTOG");
  PW_TRY(upload_plan(p));
  if (s == "GTOR") return fft_gtor_full(p, nfft, (const double2 *)fofg, fofr, prec, p->stream);
  return fft_rtog_full(p, nfft, (double2 *)fofg, fofr, prec, p->stream);
}
