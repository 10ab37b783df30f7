// Batched 3D plane-wave FFTs for gfx950.
//
// Semantics follow PLANEWAVE$FFT (src/paw_fft.f90:1948-2341) with the
// conventions of LIB$FFTC8 (src/paw_library.f90:2396-2426):
//   GTOR: f(r) = sum_G f(G) exp(+iGr)          (unnormalised)
//   RTOG: f(G) = 1/N sum_r f(r) exp(-iGr)      (on the G sphere)
// The reference's pruning (x only on stripes, y only on z planes holding
// stripes) is kept, so the result equals the full 3D DFT of the scattered
// sphere.  Unlike the reference, which runs one band and one FFTW line at a
// time, each pass here is one batched kernel over all bands:
//   x pass : sphere -> stripe rows (scatter fused), radix-R Stockham in LDS
//   yz pass: one (x, band) plane of NR2 x NR3 resident in LDS: y transforms
//            on the z planes that hold stripes, full z transforms, then
//            V(r) and the inverse transforms (VPSI) or the occupation-
//            weighted |psi|^2 sum (density); no grid round trip through HBM
//   y / z passes over HBM z planes serve the full-grid PLANEWAVE$FFT calls
// Work buffers: stripes [band][x][stripe], planes [band][iz][y][x].
#include <type_traits>

#include "pw_common.h"

namespace {

constexpr int kBlock = 256;
constexpr int kTX = 16;       // x columns per y/z tile

template <typename T2> struct Prec;
template <> struct Prec<double2> {
  typedef double T;
  static __device__ inline double2 from(double2 v) { return v; }
};
template <> struct Prec<float2> {
  typedef float T;
  static __device__ inline float2 from(double2 v) { return make_float2((float)v.x, (float)v.y); }
};

template <typename T2>
__device__ inline T2 czero() { T2 r; r.x = 0; r.y = 0; return r; }

// a * (sign * i)
template <typename T2>
__device__ inline T2 mul_si(T2 a, int sign) {
  T2 r;
  r.x = sign > 0 ? -a.y : a.y;
  r.y = sign > 0 ? a.x : -a.x;
  return r;
}

// cos/sin(2 pi m / R)
template <int R> struct Roots;
template <> struct Roots<3> {
  static constexpr double c[3] = {1.0, -0.5, -0.5};
  static constexpr double s[3] = {0.0, 0.86602540378443864676, -0.86602540378443864676};
};
template <> struct Roots<5> {
  static constexpr double c[5] = {1.0, 0.30901699437494742410, -0.80901699437494742410,
                                  -0.80901699437494742410, 0.30901699437494742410};
  static constexpr double s[5] = {0.0, 0.95105651629515357212, 0.58778525229247312917,
                                  -0.58778525229247312917, -0.95105651629515357212};
};
template <> struct Roots<7> {
  static constexpr double c[7] = {1.0, 0.62348980185873353053, -0.22252093395631440429,
                                  -0.90096886790241912624, -0.90096886790241912624,
                                  -0.22252093395631440429, 0.62348980185873353053};
  static constexpr double s[7] = {0.0, 0.78183148246802980871, 0.97492791218182360702,
                                  0.43388373911755812048, -0.43388373911755812048,
                                  -0.97492791218182360702, -0.78183148246802980871};
};
template <> struct Roots<11> {
  static constexpr double c[11] = {1.0, 0.84125353283118116886, 0.41541501300188642553,
                                   -0.14231483827328514044, -0.65486073394528506406,
                                   -0.95949297361449738989, -0.95949297361449738989,
                                   -0.65486073394528506406, -0.14231483827328514044,
                                   0.41541501300188642553, 0.84125353283118116886};
  static constexpr double s[11] = {0.0, 0.54064081745559758211, 0.90963199535451837141,
                                   0.98982144188093273238, 0.75574957435425828377,
                                   0.28173255684142969771, -0.28173255684142969771,
                                   -0.75574957435425828377, -0.98982144188093273238,
                                   -0.90963199535451837141, -0.54064081745559758211};
};

template <> struct Roots<9> {
  static constexpr double c[9] = {1.0, 0.766044443118978035202, 0.173648177666930348852, -0.5,
                                  -0.939692620785908384054, -0.939692620785908384054, -0.5,
                                  0.173648177666930348852, 0.766044443118978035202};
  static constexpr double s[9] = {0.0, 0.642787609686539326323, 0.984807753012208059367,
                                  0.866025403784438646764, 0.342020143325668733044,
                                  -0.342020143325668733044, -0.866025403784438646764,
                                  -0.984807753012208059367, -0.642787609686539326323};
};
template <> struct Roots<16> {
  static constexpr double c[16] = {1.0, 0.923879532511286756128, 0.707106781186547524401,
                                   0.382683432365089771728, 0.0, -0.382683432365089771728,
                                   -0.707106781186547524401, -0.923879532511286756128, -1.0,
                                   -0.923879532511286756128, -0.707106781186547524401,
                                   -0.382683432365089771728, 0.0, 0.382683432365089771728,
                                   0.707106781186547524401, 0.923879532511286756128};
  static constexpr double s[16] = {0.0, 0.382683432365089771728, 0.707106781186547524401,
                                   0.923879532511286756128, 1.0, 0.923879532511286756128,
                                   0.707106781186547524401, 0.382683432365089771728, 0.0,
                                   -0.382683432365089771728, -0.707106781186547524401,
                                   -0.923879532511286756128, -1.0, -0.923879532511286756128,
                                   -0.707106781186547524401, -0.382683432365089771728};
};

template <typename T2, int R>
__device__ inline void dft(T2 *v, int sign);

// DFT of size R = P Q from P-point and Q-point DFTs: input r = Q a + b,
// output q = q1 + P q2; Y_b = DFT_P over a, times exp(sign 2 pi i b q1 / R),
// then DFT_Q over b.
template <typename T2, int P, int Q>
__device__ inline void dft_pq(T2 *v, int sign) {
  typedef typename Prec<T2>::T T;
  constexpr int R = P * Q;
  T2 y[Q][P];
#pragma unroll
  for (int b = 0; b < Q; ++b) {
    T2 t[P];
#pragma unroll
    for (int a = 0; a < P; ++a) t[a] = v[Q * a + b];
    dft<T2, P>(t, sign);
#pragma unroll
    for (int q1 = 0; q1 < P; ++q1) y[b][q1] = t[q1];
  }
#pragma unroll
  for (int b = 1; b < Q; ++b)
#pragma unroll
    for (int q1 = 1; q1 < P; ++q1) {
      const int m = (b * q1) % R;
      T2 w;
      w.x = (T)Roots<R>::c[m];
      w.y = sign > 0 ? (T)Roots<R>::s[m] : -(T)Roots<R>::s[m];
      y[b][q1] = cmul(y[b][q1], w);
    }
#pragma unroll
  for (int q1 = 0; q1 < P; ++q1) {
    T2 t[Q];
#pragma unroll
    for (int b = 0; b < Q; ++b) t[b] = y[b][q1];
    dft<T2, Q>(t, sign);
#pragma unroll
    for (int q2 = 0; q2 < Q; ++q2) v[q1 + P * q2] = t[q2];
  }
}

// in-register DFT of size R: v[q] <- sum_r v[r] exp(sign 2 pi i r q / R)
template <typename T2, int R>
__device__ inline void dft(T2 *v, int sign) {
  typedef typename Prec<T2>::T T;
  if constexpr (R == 2) {
    T2 a = v[0], b = v[1];
    v[0] = cadd(a, b); v[1] = csub(a, b);
  } else if constexpr (R == 3) {
    // m = v0 - t/2, out(1,2) = m +- sign sin(2 pi/3) i d: 12 operations
    const T c = (T)-0.5, s = (T)0.86602540378443864676;
    const T2 t = cadd(v[1], v[2]), d = csub(v[1], v[2]);
    T2 m;
    m.x = v[0].x + c * t.x;
    m.y = v[0].y + c * t.y;
    v[0] = cadd(v[0], t);
    const T sx = sign > 0 ? s : -s;
    T2 o1, o2;
    o1.x = m.x - sx * d.y;
    o1.y = m.y + sx * d.x;
    o2.x = m.x + sx * d.y;
    o2.y = m.y - sx * d.x;
    v[1] = o1;
    v[2] = o2;
  } else if constexpr (R == 4) {
    T2 t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
    T2 t2 = cadd(v[1], v[3]), t3 = mul_si(csub(v[1], v[3]), sign);
    v[0] = cadd(t0, t2); v[2] = csub(t0, t2);
    v[1] = cadd(t1, t3); v[3] = csub(t1, t3);
  } else if constexpr (R == 9) {
    dft_pq<T2, 3, 3>(v, sign);
  } else if constexpr (R == 16) {
    dft_pq<T2, 4, 4>(v, sign);
  } else if constexpr (R == 8) {
    T2 e[4] = {v[0], v[2], v[4], v[6]};
    T2 o[4] = {v[1], v[3], v[5], v[7]};
    dft<T2, 4>(e, sign);
    dft<T2, 4>(o, sign);
    const T r2 = (T)0.70710678118654752440;
    // w^q, w = exp(sign 2 pi i/8)
    T2 w1; w1.x = r2; w1.y = sign > 0 ? r2 : -r2;
    T2 w3; w3.x = -r2; w3.y = sign > 0 ? r2 : -r2;
    T2 o1 = cmul(o[1], w1), o2 = mul_si(o[2], sign), o3 = cmul(o[3], w3);
    v[0] = cadd(e[0], o[0]); v[4] = csub(e[0], o[0]);
    v[1] = cadd(e[1], o1);   v[5] = csub(e[1], o1);
    v[2] = cadd(e[2], o2);   v[6] = csub(e[2], o2);
    v[3] = cadd(e[3], o3);   v[7] = csub(e[3], o3);
  } else {
    constexpr int H = (R - 1) / 2;
    T2 sp[H], sm[H];
#pragma unroll
    for (int r = 1; r <= H; ++r) {
      sp[r - 1] = cadd(v[r], v[R - r]);
      sm[r - 1] = mul_si(csub(v[r], v[R - r]), sign);
    }
    // out[q], out[R-q] = a_q +- b_q with a_q = v0 + sum_r cos(2 pi r q/R) sp_r
    // and b_q = sum_r sin(2 pi r q/R) sm_r (cos even, sin odd in q)
    T2 out[R];
    out[0] = v[0];
#pragma unroll
    for (int r = 0; r < H; ++r) out[0] = cadd(out[0], sp[r]);
#pragma unroll
    for (int q = 1; q <= H; ++q) {
      T2 a = v[0], b;
#pragma unroll
      for (int r = 1; r <= H; ++r) {
        const int m = (r * q) % R;
        const T c = (T)Roots<R>::c[m];
        a.x += sp[r - 1].x * c;
        a.y += sp[r - 1].y * c;
      }
      b.x = sm[0].x * (T)Roots<R>::s[q % R];
      b.y = sm[0].y * (T)Roots<R>::s[q % R];
#pragma unroll
      for (int r = 2; r <= H; ++r) {
        const T s = (T)Roots<R>::s[(r * q) % R];
        b.x += sm[r - 1].x * s;
        b.y += sm[r - 1].y * s;
      }
      out[q] = cadd(a, b);
      out[R - q] = csub(a, b);
    }
#pragma unroll
    for (int q = 0; q < R; ++q) v[q] = out[q];
  }
}

// One Stockham stage over npen pencils of length n in LDS.
// Element e of pencil p lives at base[p*ps + e*es].
template <typename T2, int R>
__device__ inline void stage(const T2 *__restrict__ in, T2 *__restrict__ out, int n, int ns,
                             int npen, int es, int ps, const T2 *__restrict__ tw, int sign,
                             bool pen_fast) {
  const int nbf = n / R;
  const int total = nbf * npen;
  const int twm = n / (ns * R);
  for (int w = threadIdx.x; w < total; w += blockDim.x) {
    int pen, j;
    if (pen_fast) { pen = w % npen; j = w / npen; }
    else { j = w % nbf; pen = w / nbf; }
    const int k = j % ns;
    const T2 *src = in + pen * ps;
    T2 v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) v[r] = src[(j + r * nbf) * es];
    if (ns > 1) {
#pragma unroll
      for (int r = 1; r < R; ++r) {
        T2 t = tw[twm * k * r];
        if (sign < 0) t.y = -t.y;
        v[r] = cmul(v[r], t);
      }
    }
    dft<T2, R>(v, sign);
    const int idxd = (j / ns) * ns * R + k;
    T2 *dst = out + pen * ps;
#pragma unroll
    for (int r = 0; r < R; ++r) dst[(idxd + r * ns) * es] = v[r];
  }
}

// Full 1D FFT of npen pencils; returns the buffer holding the result.
template <typename T2>
__device__ T2 *lds_fft(T2 *b0, T2 *b1, const Fft1D &f, int npen, int es, int ps,
                       const T2 *tw, int sign, bool pen_fast) {
  int ns = 1;
  T2 *in = b0, *out = b1;
  for (int st = 0; st < f.nstage; ++st) {
    const int r = f.radix[st];
    switch (r) {
      case 2: stage<T2, 2>(in, out, f.n, ns, npen, es, ps, tw, sign, pen_fast); break;
      case 3: stage<T2, 3>(in, out, f.n, ns, npen, es, ps, tw, sign, pen_fast); break;
      case 4: stage<T2, 4>(in, out, f.n, ns, npen, es, ps, tw, sign, pen_fast); break;
      case 5: stage<T2, 5>(in, out, f.n, ns, npen, es, ps, tw, sign, pen_fast); break;
      case 7: stage<T2, 7>(in, out, f.n, ns, npen, es, ps, tw, sign, pen_fast); break;
      case 8: stage<T2, 8>(in, out, f.n, ns, npen, es, ps, tw, sign, pen_fast); break;
      case 11: stage<T2, 11>(in, out, f.n, ns, npen, es, ps, tw, sign, pen_fast); break;
      case 9: stage<T2, 9>(in, out, f.n, ns, npen, es, ps, tw, sign, pen_fast); break;
      case 16: stage<T2, 16>(in, out, f.n, ns, npen, es, ps, tw, sign, pen_fast); break;
      default: break;
    }
    ns *= r;
    __syncthreads();
    T2 *t = in; in = out; out = t;
  }
  return in;
}

// The Stockham chain of a length N fixed at compile time (fft_factor's rule
// for these passes: odd radices, then 16, 8, 4, 2); STAGE ST of NST with the
// span NS = product of the earlier radices.  Returns the result buffer.
__host__ __device__ constexpr Fft1D fft_stockham_ct(int n) {
  Fft1D f{};
  f.n = n;
  int m = n;
  const int order[9] = {9, 3, 5, 7, 11, 16, 8, 4, 2};
  for (int io = 0; io < 9; ++io)
    while (m % order[io] == 0 && f.nstage < 12) {
      f.radix[f.nstage++] = order[io];
      m /= order[io];
    }
  return f;
}
template <typename T2, int N, int ST, int NS>
__device__ inline T2 *lds_fft_ct(T2 *in, T2 *out, int npen, int es, int ps, const T2 *tw, int sign, bool pen_fast) {
  constexpr Fft1D f = fft_stockham_ct(N);
  if constexpr (ST < f.nstage) {
    constexpr int R = f.radix[ST];
    stage<T2, R>(in, out, N, NS, npen, es, ps, tw, sign, pen_fast);
    __syncthreads();
    return lds_fft_ct<T2, N, ST + 1, NS * R>(out, in, npen, es, ps, tw, sign, pen_fast);
  } else {
    return in;
  }
}

struct Geo {
  int nr1, nr2, nr3, ngl, nstripe, nzp;
  int pitch;                      // row pitch of the LDS-resident (y, z) plane (odd)
  int nypen;                      // lane slots of the y transforms (ypen)
  unsigned long long mnzp, mnr2;  // fdiv magics of nzp and nr2
  const int *stripe_off, *stripe_g, *stripe_x, *stripe_s, *yz2stripe, *z2iz, *zlist;
  const int *stripe_pos, *zpos;   // digit-reversed plane offsets of stripes / stripe rows
  const int *ypen;                // stripe-row offset of each lane slot of the y transforms, -1: none
  const double *g2;
};

// stripes per x-pass workgroup (LDS holds two kXT x NR1 buffers)
template <typename T2> struct XTile;
template <> struct XTile<double2> { static constexpr int n = 16; };
template <> struct XTile<float2> { static constexpr int n = 32; };

// Workgroup -> (stripe group, band) of the x passes, a 1D grid of
// 8 ceil(nb/8) ngroups workgroups: consecutive workgroups go to the eight XCDs
// in turn, and each XCD takes every stripe group of its bands (band = xcd +
// 8 k), so the coefficients of a band, gathered and scattered in G order by
// all its stripe groups, pass through one XCD's L2 instead of eight.
__device__ inline bool x_block(int ngroups, int nb, int &group, int &band) {
  const int id = blockIdx.x, x = id % 8, j = id / 8;
  band = x + 8 * (j / ngroups);
  group = j % ngroups;
  return band < nb;
}
inline dim3 x_grid(int ngroups, int nb) { return dim3((unsigned)(8 * ((nb + 7) / 8) * ngroups)); }
// LDS pitch of the x passes' stripe tiles: odd, so that the tile's transposes
// (neighbouring lanes on neighbouring stripes, one pitch apart) fall into
// distinct banks
__host__ __device__ inline int x_pitch(int nr1) { return nr1 | 1; }

// ---- x pass, forward: sphere -> stripes [band][x][stripe] --------------------
template <typename T2, int NX = 0>
__global__ void __launch_bounds__(kBlock) k_xfwd(Geo g, Fft1D f, const T2 *tw, const double2 *psi,
                                                 T2 *stripes, int nb) {
  constexpr int PX = XTile<T2>::n;
  if constexpr (NX > 0) g.nr1 = NX;
  int grp, band;
  if (!x_block((g.nstripe + PX - 1) / PX, nb, grp, band)) return;
  extern __shared__ unsigned char smem_raw[];
  const int xp = x_pitch(g.nr1);
  T2 *b0 = reinterpret_cast<T2 *>(smem_raw);
  T2 *b1 = b0 + PX * xp;
  const int s0 = grp * PX;
  const int ns = min(PX, g.nstripe - s0);
  for (int i = threadIdx.x; i < PX * xp; i += blockDim.x) b0[i] = czero<T2>();
  __syncthreads();
  const double2 *src = psi + (size_t)band * g.ngl;
  const int e0 = g.stripe_off[s0], e1 = g.stripe_off[s0 + ns];
  for (int e = e0 + threadIdx.x; e < e1; e += blockDim.x)
    b0[(g.stripe_s[e] - s0) * xp + g.stripe_x[e]] = Prec<T2>::from(src[g.stripe_g[e]]);
  __syncthreads();
  T2 *r;
  if constexpr (NX > 0) r = lds_fft_ct<T2, NX, 0, 1>(b0, b1, PX, 1, xp, tw, +1, false);
  else r = lds_fft<T2>(b0, b1, f, PX, 1, xp, tw, +1, false);
  T2 *dst = stripes + (size_t)band * g.nr1 * g.nstripe + s0;
  for (int i = threadIdx.x; i < ns * g.nr1; i += blockDim.x) {
    const int x = i / ns, sl = i % ns;
    dst[(size_t)x * g.nstripe + sl] = r[sl * xp + x];
  }
}

// ---- x pass, inverse: stripes -> sphere (gather, scale, optional G^2/2 psi) ---
template <typename T2, int NX = 0>
__global__ void __launch_bounds__(kBlock) k_xinv(Geo g, Fft1D f, const T2 *tw, const T2 *stripes,
                                                 double2 *out, const double2 *psi, double scale,
                                                 int add_kin, int nb) {
  constexpr int PX = XTile<T2>::n;
  if constexpr (NX > 0) g.nr1 = NX;
  int grp, band;
  if (!x_block((g.nstripe + PX - 1) / PX, nb, grp, band)) return;
  extern __shared__ unsigned char smem_raw[];
  const int xp = x_pitch(g.nr1);
  T2 *b0 = reinterpret_cast<T2 *>(smem_raw);
  T2 *b1 = b0 + PX * xp;
  const int s0 = grp * PX;
  const int ns = min(PX, g.nstripe - s0);
  const T2 *src = stripes + (size_t)band * g.nr1 * g.nstripe + s0;
  for (int i = threadIdx.x; i < PX * g.nr1; i += blockDim.x) {
    const int x = i / PX, sl = i % PX;
    b0[sl * xp + x] = sl < ns ? src[(size_t)x * g.nstripe + sl] : czero<T2>();
  }
  __syncthreads();
  T2 *r;
  if constexpr (NX > 0) r = lds_fft_ct<T2, NX, 0, 1>(b0, b1, PX, 1, xp, tw, -1, false);
  else r = lds_fft<T2>(b0, b1, f, PX, 1, xp, tw, -1, false);
  const int e0 = g.stripe_off[s0], e1 = g.stripe_off[s0 + ns];
  double2 *dst = out + (size_t)band * g.ngl;
  const double2 *ps = psi + (size_t)band * g.ngl;
  for (int e = e0 + threadIdx.x; e < e1; e += blockDim.x) {
    const int ig = g.stripe_g[e];
    T2 v = r[(g.stripe_s[e] - s0) * xp + g.stripe_x[e]];
    double2 o = make_double2((double)v.x * scale, (double)v.y * scale);
    if (add_kin) {
      const double h = 0.5 * g.g2[ig];
      const double2 p = ps[ig];
      o.x += h * p.x;
      o.y += h * p.y;
    }
    dst[ig] = o;
  }
}

// ---- y pass forward: stripes -> planes -----------------------------------------
template <typename T2>
__global__ void __launch_bounds__(kBlock) k_yfwd(Geo g, Fft1D f, const T2 *tw, const T2 *stripes,
                                                 T2 *planes) {
  extern __shared__ unsigned char smem_raw[];
  T2 *b0 = reinterpret_cast<T2 *>(smem_raw);
  T2 *b1 = b0 + kTX * g.nr2;
  const int ntx = (g.nr1 + kTX - 1) / kTX;
  const int iz = blockIdx.x / ntx;
  const int x0 = (blockIdx.x % ntx) * kTX;
  const int band = blockIdx.y;
  const int z = g.zlist[iz];
  const T2 *src = stripes + (size_t)band * g.nstripe * g.nr1;
  for (int w = threadIdx.x; w < g.nr2 * kTX; w += blockDim.x) {
    const int y = w / kTX, x = w % kTX, gx = x0 + x;
    const int s = g.yz2stripe[y + g.nr2 * z];
    b0[w] = (s >= 0 && gx < g.nr1) ? src[(size_t)gx * g.nstripe + s] : czero<T2>();
  }
  __syncthreads();
  T2 *r = lds_fft<T2>(b0, b1, f, kTX, kTX, 1, tw, +1, true);
  T2 *dst = planes + (((size_t)band * g.nzp + iz) * g.nr2) * g.nr1;
  for (int w = threadIdx.x; w < g.nr2 * kTX; w += blockDim.x) {
    const int y = w / kTX, x = w % kTX, gx = x0 + x;
    if (gx < g.nr1) dst[(size_t)y * g.nr1 + gx] = r[w];
  }
}

// ---- y pass inverse: planes -> stripes ------------------------------------------
template <typename T2>
__global__ void __launch_bounds__(kBlock) k_yinv(Geo g, Fft1D f, const T2 *tw, const T2 *planes,
                                                 T2 *stripes) {
  extern __shared__ unsigned char smem_raw[];
  T2 *b0 = reinterpret_cast<T2 *>(smem_raw);
  T2 *b1 = b0 + kTX * g.nr2;
  const int ntx = (g.nr1 + kTX - 1) / kTX;
  const int iz = blockIdx.x / ntx;
  const int x0 = (blockIdx.x % ntx) * kTX;
  const int band = blockIdx.y;
  const int z = g.zlist[iz];
  const T2 *src = planes + (((size_t)band * g.nzp + iz) * g.nr2) * g.nr1;
  for (int w = threadIdx.x; w < g.nr2 * kTX; w += blockDim.x) {
    const int y = w / kTX, x = w % kTX, gx = x0 + x;
    b0[w] = gx < g.nr1 ? src[(size_t)y * g.nr1 + gx] : czero<T2>();
  }
  __syncthreads();
  T2 *r = lds_fft<T2>(b0, b1, f, kTX, kTX, 1, tw, -1, true);
  T2 *dst = stripes + (size_t)band * g.nstripe * g.nr1;
  for (int w = threadIdx.x; w < g.nr2 * kTX; w += blockDim.x) {
    const int y = w / kTX, x = w % kTX, gx = x0 + x;
    const int s = g.yz2stripe[y + g.nr2 * z];
    if (s >= 0 && gx < g.nr1) dst[(size_t)gx * g.nstripe + s] = r[w];
  }
}

// z pencils of one (y, x tile) from the plane buffer (zero on empty planes)
template <typename T2>
__device__ inline void load_zpencils(const Geo &g, const T2 *planes_band, int y, int x0, T2 *b0) {
  for (int w = threadIdx.x; w < g.nr3 * kTX; w += blockDim.x) {
    const int z = w / kTX, x = w % kTX, gx = x0 + x;
    const int iz = g.z2iz[z];
    b0[w] = (iz >= 0 && gx < g.nr1) ? planes_band[((size_t)iz * g.nr2 + y) * g.nr1 + gx]
                                    : czero<T2>();
  }
}

template <typename T2>
__device__ inline void store_zpencils(const Geo &g, T2 *planes_band, int y, int x0, const T2 *r) {
  for (int w = threadIdx.x; w < g.nr3 * kTX; w += blockDim.x) {
    const int z = w / kTX, x = w % kTX, gx = x0 + x;
    const int iz = g.z2iz[z];
    if (iz >= 0 && gx < g.nr1) planes_band[((size_t)iz * g.nr2 + y) * g.nr1 + gx] = r[w];
  }
}

// ---- in-place LDS transforms for the resident yz plane -----------------------
// Mixed-radix Cooley-Tukey in place: every butterfly reads and writes the
// same R positions, so a thread handles one butterfly at a time and only the
// stages are separated by barriers.  DIT takes digit-reversed input and gives
// natural order; DIF takes natural order and gives digit-reversed output
// (the reversal of a length follows its radix sequence, Fft1D::radix).
constexpr int kYZ = 512;

// quotient of w < 2^16 by 0 < d < 2^16 from m = ceil(2^32/d) = fdiv_magic(d)
// (exact in that range, d = 1 included); the magics come from the host
__device__ inline unsigned fdiv(unsigned w, unsigned long long m) { return (unsigned)(((unsigned long long)w * m) >> 32); }

// one in-place butterfly stage; dif=false: twiddle then DFT (DIT), dif=true:
// DFT then twiddle (DIF).  tw: LDS copy of exp(+2 pi i m/n), m < n.
// mpen, mnbf, mlp: fdiv magics of npen, n/R and L/R.
template <typename T2, int R, bool DIF>
__device__ inline void ip_stage(T2 *buf, int n, int L, int npen, int es, int ps, const int *pmap,
                                const T2 *tw, int sign, bool pen_fast, unsigned long long mpen,
                                unsigned long long mnbf, unsigned long long mlp) {
  const int Lp = L / R, nbf = n / R, total = nbf * npen, tws = n / L;
  for (int w = threadIdx.x; w < total; w += kYZ) {
    int pen, b;
    if (pen_fast) {
      b = (int)fdiv(w, mpen);
      pen = w - b * npen;
    } else {
      pen = (int)fdiv(w, mnbf);
      b = w - pen * nbf;
    }
    const int g = (int)fdiv(b, mlp), j = b - g * Lp;
    T2 *base = buf + (pmap ? pmap[pen] : pen * ps) + (g * L + j) * es;
    T2 v[R];
#pragma unroll
    for (int q = 0; q < R; ++q) v[q] = base[q * Lp * es];
    if (DIF) dft<T2, R>(v, sign);
    if (Lp > 1) {
#pragma unroll
      for (int q = 1; q < R; ++q) {
        T2 t = tw[j * q * tws];
        if (sign < 0) t.y = -t.y;
        v[q] = cmul(v[q], t);
      }
    }
    if (!DIF) dft<T2, R>(v, sign);
#pragma unroll
    for (int q = 0; q < R; ++q) base[q * Lp * es] = v[q];
  }
  __syncthreads();
}

// radices of the in-place transforms (fft_factor with inplace: no 16); 11
// only in the BIG instantiations, whose butterfly needs twice the registers
#define PW_RADIX_SWITCH(r, CALL)               \
  switch (r) {                                 \
    case 2: CALL(2); break;                    \
    case 3: CALL(3); break;                    \
    case 4: CALL(4); break;                    \
    case 5: CALL(5); break;                    \
    case 7: CALL(7); break;                    \
    case 8: CALL(8); break;                    \
    case 9: CALL(9); break;                    \
    case 11: if constexpr (BIG) CALL(11); break; \
    default: break;                            \
  }

// stages 0..nst-1 of the DIT transform (nst = f.nstage: all of it)
template <typename T2, bool BIG>
__device__ void fft_dit(T2 *buf, const Fft1D &f, int nst, int npen, unsigned long long mpen, int es, int ps,
                        const int *pmap, const T2 *tw, int sign, bool pen_fast) {
  int L = 1;
  for (int st = 0; st < nst; ++st) {
    const int r = f.radix[st];
    L *= r;
#define PW_CALL(R) \
  ip_stage<T2, R, false>(buf, f.n, L, npen, es, ps, pmap, tw, sign, pen_fast, mpen, f.mnbf[st], f.mlp[st])
    PW_RADIX_SWITCH(r, PW_CALL)
#undef PW_CALL
  }
}

// stages nst-1..0 of the DIF transform (nst = f.nstage: all of it)
template <typename T2, bool BIG>
__device__ void fft_dif(T2 *buf, const Fft1D &f, int nst, int npen, unsigned long long mpen, int es, int ps,
                        const int *pmap, const T2 *tw, int sign, bool pen_fast) {
  int L = 1;
  for (int st = 0; st < nst; ++st) L *= f.radix[st];
  for (int st = nst - 1; st >= 0; --st) {
    const int r = f.radix[st];
#define PW_CALL(R) \
  ip_stage<T2, R, true>(buf, f.n, L, npen, es, ps, pmap, tw, sign, pen_fast, mpen, f.mnbf[st], f.mlp[st])
    PW_RADIX_SWITCH(r, PW_CALL)
#undef PW_CALL
    L /= r;
  }
}

// ---- plane lengths fixed at compile time ---------------------------------------
// A yz kernel instantiated for plane lengths NY x NZ (0: taken from the plan)
// runs its stages from a compile-time factorisation: the stage sequence is
// unrolled, radices, spans and strides are constants, each thread keeps one
// pencil per transform and the twiddles sit in per-stage tables (ct_stage).
// The factorisation follows fft_factor's in-place rule; the launcher checks it
// against the plan before it picks such a kernel.
__host__ __device__ constexpr unsigned long long fdiv_magic_ct(unsigned d) {
  return (0x100000000ull + d - 1) / d;
}
__host__ __device__ constexpr Fft1D fft_inplace_ct(int n) {
  Fft1D f{};
  f.n = n;
  int m = n;
  const int order[8] = {9, 3, 5, 7, 11, 8, 4, 2};
  for (int io = 0; io < 8; ++io)
    while (m % order[io] == 0 && f.nstage < 12) {
      f.radix[f.nstage++] = order[io];
      m /= order[io];
    }
  int L = 1;
  for (int st = 0; st < f.nstage; ++st) {
    f.mnbf[st] = fdiv_magic_ct((unsigned)(n / f.radix[st]));
    f.mlp[st] = fdiv_magic_ct((unsigned)L);
    L *= f.radix[st];
  }
  return f;
}
__host__ __device__ constexpr int fft_lprod(const Fft1D &f, int st) {
  int L = 1;
  for (int k = 0; k <= st; ++k) L *= f.radix[k];
  return L;
}

template <int N> constexpr int nstage_ct() { return fft_inplace_ct(N).nstage; }

// the plane's scalar lengths as constants
template <int NY, int NZ>
__device__ inline void yz_fix(Geo &g) {
  if constexpr (NY > 0) {
    g.nr2 = NY;
    g.pitch = NY | 1;
    g.mnr2 = fdiv_magic_ct(NY);
  }
  if constexpr (NZ > 0) g.nr3 = NZ;
}

// ---- stages of the compile-time plane lengths ------------------------------------
// A thread keeps one pencil through every stage of a transform: PenSet holds
// the pencil's LDS offset and the butterflies b0, b0 + bstep, ... it takes in
// each stage, so a butterfly's position needs no division.  The twiddles of
// stage ST sit in LDS in butterfly order,
//   tab[tw_off_ct(N, ST) + j (R - 1) + q - 1] = exp(+2 pi i j q N / L),
// j < L/R, 1 <= q < R; the stages together hold N - 1 of them.
struct PenSet {
  int off, b0, bstep;   // b0 >= every stage's butterfly count: the thread has none
};
__host__ __device__ constexpr int tw_off_ct(int n, int st) {
  const Fft1D f = fft_inplace_ct(n);
  int off = 0, L = 1;
  for (int s = 0; s < st; ++s) {
    L *= f.radix[s];
    off += (L / f.radix[s]) * (f.radix[s] - 1);
  }
  return off;
}

// pencils p = tid mod npen with offsets p * ps (pmap == nullptr) or pmap[p]
// (a negative entry: no pencil); every thread below (kYZ / npen) npen takes
// part (npen <= kYZ)
__device__ inline PenSet pen_set(int npen, int ps, const int *pmap) {
  PenSet s;
  const int t = threadIdx.x, step = npen > 0 ? kYZ / npen : 0;
  const int q = npen > 0 ? t / npen : 0, p = t - q * npen;
  const int off = npen > 0 && t < step * npen ? (pmap ? pmap[p] : p * ps) : -1;
  if (off >= 0) {
    s.off = off;
    s.b0 = q;
    s.bstep = step;
  } else {
    s.off = 0;
    s.b0 = 1 << 30;
    s.bstep = 1;
  }
  return s;
}

// the stage twiddle table of length N from root[m] = exp(+2 pi i m / N)
template <typename T2, int N>
__device__ inline void tw_table_ct(const T2 *root, T2 *tab) {
  constexpr Fft1D f = fft_inplace_ct(N);
  int off = 0, L = 1;
  for (int st = 0; st < f.nstage; ++st) {
    const int R = f.radix[st];
    L *= R;
    const int cnt = (L / R) * (R - 1);
    for (int e = threadIdx.x; e < cnt; e += kYZ) {
      const int j = e / (R - 1), q = e - j * (R - 1) + 1;
      tab[off + e] = root[j * q * (N / L)];
    }
    off += cnt;
  }
}

// stage ST of length N over the pencils of ps, elements ES apart; DIF: DFT
// then twiddle, else twiddle then DFT; SIGN: exp(SIGN 2 pi i ...)
template <typename T2, int N, int ST, bool DIF, int SIGN, int ES>
__device__ inline void ct_stage(T2 *buf, const PenSet &ps, const T2 *tab) {
  constexpr Fft1D f = fft_inplace_ct(N);
  constexpr int R = f.radix[ST], L = fft_lprod(f, ST), Lp = L / R, NBF = N / R, TO = tw_off_ct(N, ST);
  for (int b = ps.b0; b < NBF; b += ps.bstep) {
    int gg, j;
    if constexpr (Lp == 1) {
      gg = b;
      j = 0;
    } else if constexpr (Lp == NBF) {
      gg = 0;
      j = b;
    } else {
      gg = b / Lp;
      j = b - gg * Lp;
    }
    T2 *base = buf + ps.off + __mul24(__mul24(gg, L) + j, ES);
    T2 v[R];
#pragma unroll
    for (int q = 0; q < R; ++q) v[q] = base[q * Lp * ES];
    if constexpr (DIF) dft<T2, R>(v, SIGN);
    if constexpr (Lp > 1) {
      const T2 *t = tab + TO + __mul24(j, R - 1);
#pragma unroll
      for (int q = 1; q < R; ++q) {
        T2 w = t[q - 1];
        if (SIGN < 0) w.y = -w.y;
        v[q] = cmul(v[q], w);
      }
    }
    if constexpr (!DIF) dft<T2, R>(v, SIGN);
#pragma unroll
    for (int q = 0; q < R; ++q) base[q * Lp * ES] = v[q];
  }
  __syncthreads();
}

// stages ST..NST-1 of the DIT transform, stages ST..0 of the DIF transform
template <typename T2, int N, int ST, int NST, int SIGN, int ES>
__device__ inline void ct_dit(T2 *buf, const PenSet &ps, const T2 *tab) {
  if constexpr (ST < NST) {
    ct_stage<T2, N, ST, false, SIGN, ES>(buf, ps, tab);
    ct_dit<T2, N, ST + 1, NST, SIGN, ES>(buf, ps, tab);
  }
}
template <typename T2, int N, int ST, int SIGN, int ES>
__device__ inline void ct_dif(T2 *buf, const PenSet &ps, const T2 *tab) {
  if constexpr (ST >= 0) {
    ct_stage<T2, N, ST, true, SIGN, ES>(buf, ps, tab);
    ct_dif<T2, N, ST - 1, SIGN, ES>(buf, ps, tab);
  }
}

// zero the plane, scatter the x column of one band's stripes into the
// digit-reversed (z, y) positions, transform y (DIT, on the rows of the z
// planes that hold stripes) and z (DIT) up to, not including, its last
// stage, which the caller fuses with its use of psi(x, y, z) (z_last_*).
// Rows of the plane are g.pitch apart (forward, +i).  zpos: offset of each
// stripe-holding row.  Both transforms run pencil-fast: neighbouring lanes
// take neighbouring pencils, whose rows (y) or columns (z) fall into distinct
// LDS banks.
template <typename T2, bool BIG>
__device__ inline void yz_forward(const Geo &g, const Fft1D &fy, const Fft1D &fz, const T2 *twy,
                                  const T2 *twz, const T2 *col, T2 *pl, const int *zpos) {
  const int npl = g.pitch * g.nr3;
  for (int i = threadIdx.x; i < npl; i += kYZ) pl[i] = czero<T2>();
  __syncthreads();
  for (int s = threadIdx.x; s < g.nstripe; s += kYZ) pl[g.stripe_pos[s]] = col[s];
  __syncthreads();
  fft_dit<T2, BIG>(pl, fy, fy.nstage, g.nzp, g.mnzp, 1, 0, zpos, twy, +1, true);
  fft_dit<T2, BIG>(pl, fz, fz.nstage - 1, g.nr2, g.mnr2, g.pitch, 1, nullptr, twz, +1, true);
}

// The last DIT stage of the z transforms, one butterfly per (column y, j):
// it yields psi(z = j + q Lp) for q < R (natural order), Lp = NR3/R.
template <typename T2, int R>
__device__ inline void z_last_load(const T2 *base, int j, int Lp, int es, const T2 *tw, T2 *v) {
#pragma unroll
  for (int q = 0; q < R; ++q) v[q] = base[q * Lp * es];
  if (Lp > 1) {
#pragma unroll
    for (int q = 1; q < R; ++q) v[q] = cmul(v[q], tw[j * q]);
  }
  dft<T2, R>(v, +1);
}

// VPSI: the last forward z stage, psi V(r), and the first inverse (DIF) z
// stage, which works on the same R positions, in one pass over the plane
template <typename T2, int R>
__device__ inline void z_last_vpsi(const Geo &g, T2 *pl, int nr3, const T2 *tw, const typename Prec<T2>::T *vx) {
  typedef typename Prec<T2>::T T;
  const int Lp = nr3 / R, total = Lp * g.nr2;
  for (int w = threadIdx.x; w < total; w += kYZ) {
    const int j = (int)fdiv(w, g.mnr2), y = w - j * g.nr2;
    T2 *base = pl + y + j * g.pitch;
    T2 v[R];
    z_last_load<T2, R>(base, j, Lp, g.pitch, tw, v);
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const T vv = vx[(j + q * Lp) * g.nr2 + y];
      v[q].x *= vv;
      v[q].y *= vv;
    }
    dft<T2, R>(v, -1);
    if (Lp > 1) {
#pragma unroll
      for (int q = 1; q < R; ++q) {
        T2 t = tw[j * q];
        t.y = -t.y;
        v[q] = cmul(v[q], t);
      }
    }
#pragma unroll
    for (int q = 0; q < R; ++q) base[q * Lp * g.pitch] = v[q];
  }
  __syncthreads();
}

// DENSITY: the last forward z stage and the band's term of the density sum
// acc[z][y] += s1 |psi|^2 + s2 Re(psi^2) = s1 (re^2 + im^2) + s2 (re^2 - im^2)
template <typename T2, int R>
__device__ inline void z_last_dens(const Geo &g, const T2 *pl, int nr3, const T2 *tw, double s1, double s2,
                                   double *acc) {
  const int Lp = nr3 / R, total = Lp * g.nr2;
  for (int w = threadIdx.x; w < total; w += kYZ) {
    const int j = (int)fdiv(w, g.mnr2), y = w - j * g.nr2;
    T2 v[R];
    z_last_load<T2, R>(pl + y + j * g.pitch, j, Lp, g.pitch, tw, v);
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const double re = v[q].x, im = v[q].y;
      acc[(j + q * Lp) * g.nr2 + y] += s1 * (re * re + im * im) + s2 * (re * re - im * im);
    }
  }
}

// The same with the band sum in registers: thread t owns the positions of the
// butterflies t and t + kYZ of that stage (R <= kRegR, at most 2 kYZ butterflies;
// the launcher checks both),
// the same in every band; flushed to part by dens_reg_store.
constexpr int kRegR = 8;
template <typename T2, int R>
__device__ inline void z_last_dens_reg(const Geo &g, const T2 *pl, int nr3, const T2 *tw, double s1, double s2,
                                       double (&accr)[2][kRegR]) {
  const int Lp = nr3 / R, total = Lp * g.nr2;
#pragma unroll
  for (int it = 0; it < 2; ++it) {
    const int w = threadIdx.x + it * kYZ;
    if (w < total) {
      const int j = (int)fdiv(w, g.mnr2), y = w - j * g.nr2;
      T2 v[R];
      z_last_load<T2, R>(pl + y + j * g.pitch, j, Lp, g.pitch, tw, v);
#pragma unroll
      for (int q = 0; q < R; ++q) {
        const double re = v[q].x, im = v[q].y;
        accr[it][q] += s1 * (re * re + im * im) + s2 * (re * re - im * im);
      }
    }
  }
}

template <int R>
__device__ inline void dens_reg_store(const Geo &g, int nr3, const double (&accr)[2][kRegR], double *dst) {
  const int Lp = nr3 / R, total = Lp * g.nr2;
#pragma unroll
  for (int it = 0; it < 2; ++it) {
    const int w = threadIdx.x + it * kYZ;
    if (w < total) {
      const int j = (int)fdiv(w, g.mnr2), y = w - j * g.nr2;
#pragma unroll
      for (int q = 0; q < R; ++q) dst[(j + q * Lp) * g.nr2 + y] = accr[it][q];
    }
  }
}

// ---- the same for compile-time plane lengths NY x NZ (row pitch P = NY | 1) ----------
// py: the stripe rows (pencils of the y transforms), pz: the columns y (pencils
// of the z transforms, off = y); tby, tbz: the stage twiddle tables.
template <typename T2, int NY, int NZ>
__device__ inline void yz_forward_ct(const Geo &g, const T2 *col, T2 *pl, const PenSet &py, const PenSet &pz,
                                     const T2 *tby, const T2 *tbz) {
  constexpr int P = NY | 1;
  for (int i = threadIdx.x; i < P * NZ; i += kYZ) pl[i] = czero<T2>();
  __syncthreads();
  for (int s = threadIdx.x; s < g.nstripe; s += kYZ) pl[g.stripe_pos[s]] = col[s];
  __syncthreads();
  ct_dit<T2, NY, 0, nstage_ct<NY>(), +1, 1>(pl, py, tby);
  ct_dit<T2, NZ, 0, nstage_ct<NZ>() - 1, +1, P>(pl, pz, tbz);
}

// the last forward z stage of butterfly j of column pz.off: psi(z = j + q Lp)
template <typename T2, int NY, int NZ>
__device__ inline void z_last_load_ct(const T2 *pl, int y, int j, const T2 *tbz, T2 *v) {
  constexpr Fft1D f = fft_inplace_ct(NZ);
  constexpr int ST = f.nstage - 1, R = f.radix[ST], Lp = NZ / R, P = NY | 1, TO = tw_off_ct(NZ, ST);
  const T2 *base = pl + y + __mul24(j, P);
#pragma unroll
  for (int q = 0; q < R; ++q) v[q] = base[q * Lp * P];
  if constexpr (Lp > 1) {
    const T2 *t = tbz + TO + __mul24(j, R - 1);
#pragma unroll
    for (int q = 1; q < R; ++q) v[q] = cmul(v[q], t[q - 1]);
  }
  dft<T2, R>(v, +1);
}

// VPSI: last forward z stage, psi V(r), first inverse (DIF) z stage
template <typename T2, int NY, int NZ>
__device__ inline void z_last_vpsi_ct(T2 *pl, const PenSet &pz, const T2 *tbz, const typename Prec<T2>::T *vx) {
  typedef typename Prec<T2>::T T;
  constexpr Fft1D f = fft_inplace_ct(NZ);
  constexpr int ST = f.nstage - 1, R = f.radix[ST], Lp = NZ / R, P = NY | 1, TO = tw_off_ct(NZ, ST);
  const int y = pz.off;
  for (int j = pz.b0; j < Lp; j += pz.bstep) {
    T2 v[R];
    z_last_load_ct<T2, NY, NZ>(pl, y, j, tbz, v);
    const T *vj = vx + (unsigned)(__mul24(j, NY) + y);
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const T vv = vj[q * Lp * NY];
      v[q].x *= vv;
      v[q].y *= vv;
    }
    dft<T2, R>(v, -1);
    if constexpr (Lp > 1) {
      const T2 *t = tbz + TO + __mul24(j, R - 1);
#pragma unroll
      for (int q = 1; q < R; ++q) {
        T2 w = t[q - 1];
        w.y = -w.y;
        v[q] = cmul(v[q], w);
      }
    }
    T2 *base = pl + y + __mul24(j, P);
#pragma unroll
    for (int q = 0; q < R; ++q) base[q * Lp * P] = v[q];
  }
  __syncthreads();
}

// DENSITY: the band's term of the sum at the butterfly positions of the last
// z stage, into LDS (acc[z][y]) or into registers (accr[it], butterfly
// b0 + it bstep; needs L/R <= 2 bstep, which the launcher checks)
template <typename T2, int NY, int NZ>
__device__ inline void z_last_dens_ct(const T2 *pl, const PenSet &pz, const T2 *tbz, double s1, double s2,
                                      double *acc) {
  constexpr Fft1D f = fft_inplace_ct(NZ);
  constexpr int R = f.radix[f.nstage - 1], Lp = NZ / R;
  const int y = pz.off;
  for (int j = pz.b0; j < Lp; j += pz.bstep) {
    T2 v[R];
    z_last_load_ct<T2, NY, NZ>(pl, y, j, tbz, v);
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const double re = v[q].x, im = v[q].y;
      acc[(j + q * Lp) * NY + y] += s1 * (re * re + im * im) + s2 * (re * re - im * im);
    }
  }
}
template <typename T2, int NY, int NZ>
__device__ inline void z_last_dens_reg_ct(const T2 *pl, const PenSet &pz, const T2 *tbz, double s1, double s2,
                                          double (&accr)[2][kRegR]) {
  constexpr Fft1D f = fft_inplace_ct(NZ);
  constexpr int R = f.radix[f.nstage - 1], Lp = NZ / R;
  static_assert(R <= kRegR, "register band sum: radix above kRegR");
#pragma unroll
  for (int it = 0; it < 2; ++it) {
    const int j = pz.b0 + it * pz.bstep;
    if (j < Lp) {
      T2 v[R];
      z_last_load_ct<T2, NY, NZ>(pl, pz.off, j, tbz, v);
#pragma unroll
      for (int q = 0; q < R; ++q) {
        const double re = v[q].x, im = v[q].y;
        accr[it][q] += s1 * (re * re + im * im) + s2 * (re * re - im * im);
      }
    }
  }
}
template <int NY, int NZ>
__device__ inline void dens_reg_store_ct(const PenSet &pz, const double (&accr)[2][kRegR], double *dst) {
  constexpr Fft1D f = fft_inplace_ct(NZ);
  constexpr int R = f.radix[f.nstage - 1], Lp = NZ / R;
#pragma unroll
  for (int it = 0; it < 2; ++it) {
    const int j = pz.b0 + it * pz.bstep;
    if (j < Lp) {
#pragma unroll
      for (int q = 0; q < R; ++q) dst[(j + q * Lp) * NY + pz.off] = accr[it][q];
    }
  }
}
// the register band sum fits: every thread has at most two last-stage butterflies
template <int NY, int NZ>
constexpr bool dens_reg_ct() {
  constexpr Fft1D f = fft_inplace_ct(NZ);
  return f.radix[f.nstage - 1] <= kRegR && NZ / f.radix[f.nstage - 1] <= 2 * (kYZ / NY);
}

// ---- yz pass of WAVES_VPSI: psi(x,y,z) * V(r), back to the stripes ----------------
// One workgroup per (x, band); vt = V as [x][z][y] in the work precision.
template <typename T2, bool BIG, int NY = 0, int NZ = 0>
__global__ void __launch_bounds__(kYZ) k_yz_vpsi(Geo g, Fft1D fy, Fft1D fz, const T2 *twy, const T2 *twz,
                                                 T2 *stripes, const typename Prec<T2>::T *vt) {
  typedef typename Prec<T2>::T T;
  yz_fix<NY, NZ>(g);
  extern __shared__ unsigned char smem_raw[];
  const int n23 = g.nr2 * g.nr3;
  T2 *pl = reinterpret_cast<T2 *>(smem_raw);
  T2 *ty = pl + g.pitch * g.nr3, *tz = ty + g.nr2;
  const int x = blockIdx.x, band = blockIdx.y;
  T2 *col = stripes + ((size_t)band * g.nr1 + x) * g.nstripe;
  const T *vx = vt + (size_t)x * n23;
  if constexpr (NY > 0 && NZ > 0) {
    tw_table_ct<T2, NY>(twy, ty);
    tw_table_ct<T2, NZ>(twz, tz);
    const PenSet py = pen_set(g.nypen, 0, g.ypen), pz = pen_set(NY, 1, nullptr);
    yz_forward_ct<T2, NY, NZ>(g, col, pl, py, pz, ty, tz);
    z_last_vpsi_ct<T2, NY, NZ>(pl, pz, tz, vx);
    ct_dif<T2, NZ, nstage_ct<NZ>() - 2, -1, NY | 1>(pl, pz, tz);
    ct_dif<T2, NY, nstage_ct<NY>() - 1, -1, 1>(pl, py, ty);
  } else {
    int *zpos = reinterpret_cast<int *>(tz + g.nr3);
    for (int i = threadIdx.x; i < g.nzp; i += kYZ) zpos[i] = g.zpos[i];
    for (int i = threadIdx.x; i < g.nr2; i += kYZ) ty[i] = twy[i];
    for (int i = threadIdx.x; i < g.nr3; i += kYZ) tz[i] = twz[i];
    yz_forward<T2, BIG>(g, fy, fz, ty, tz, col, pl, zpos);
#define PW_CALL(R) z_last_vpsi<T2, R>(g, pl, g.nr3, tz, vx)
    PW_RADIX_SWITCH(fz.radix[fz.nstage - 1], PW_CALL)
#undef PW_CALL
    fft_dif<T2, BIG>(pl, fz, fz.nstage - 1, g.nr2, g.mnr2, g.pitch, 1, nullptr, tz, -1, true);
    fft_dif<T2, BIG>(pl, fy, fy.nstage, g.nzp, g.mnzp, 1, 0, zpos, ty, -1, true);
  }
  for (int s = threadIdx.x; s < g.nstripe; s += kYZ) col[s] = pl[g.stripe_pos[s]];
}

// ---- yz pass of WAVES_DENSITY: sum over a group of bands ----------------------
// part[group][x][z][y] = sum_b s1 |psi|^2 + s2 Re(psi^2) over the group's bands
// REG: the band sum in registers (z_last_dens_reg), LDS holds only the plane;
// the FP32 form is held to 128 registers so that two workgroups share a SIMD set
template <typename T2, bool REG> struct DensWaves { static constexpr int v = 1; };
template <> struct DensWaves<float2, true> { static constexpr int v = 4; };
template <typename T2, bool BIG, bool REG, int NY = 0, int NZ = 0>
__global__ void __launch_bounds__(kYZ) __attribute__((amdgpu_waves_per_eu(DensWaves<T2, REG>::v))) k_yz_dens(Geo g, Fft1D fy, Fft1D fz, const T2 *twy, const T2 *twz,
                                                 const T2 *stripes, int nb, int bpg, const double *occ2,
                                                 double *part) {
  static_assert(!(REG && BIG), "the register band sum takes radices up to kRegR");
  yz_fix<NY, NZ>(g);
  extern __shared__ unsigned char smem_raw[];
  const int n23 = g.nr2 * g.nr3;
  double *acc = reinterpret_cast<double *>(smem_raw);
  T2 *pl = reinterpret_cast<T2 *>(REG ? acc : acc + n23);
  double accr[2][kRegR];
#pragma unroll
  for (int it = 0; it < 2; ++it)
#pragma unroll
    for (int q = 0; q < kRegR; ++q) accr[it][q] = 0.0;
  T2 *ty = pl + g.pitch * g.nr3, *tz = ty + g.nr2;
  int *zpos = reinterpret_cast<int *>(tz + g.nr3);
  const int x = blockIdx.x, grp = blockIdx.y;
  const int bstart = grp * bpg, bend = min(nb, bstart + bpg);
  double *dst = part + ((size_t)grp * g.nr1 + x) * n23;
  if (!REG)
    for (int i = threadIdx.x; i < n23; i += kYZ) acc[i] = 0.0;
  if constexpr (NY > 0 && NZ > 0) {
    if constexpr (!REG || dens_reg_ct<NY, NZ>()) {
      tw_table_ct<T2, NY>(twy, ty);
      tw_table_ct<T2, NZ>(twz, tz);
      const PenSet py = pen_set(g.nypen, 0, g.ypen), pz = pen_set(NY, 1, nullptr);
      for (int band = bstart; band < bend; ++band) {
        const double s1 = occ2[2 * band], s2 = occ2[2 * band + 1];
        if (s1 == 0.0 && s2 == 0.0) continue;   // uniform across the workgroup
        __syncthreads();
        yz_forward_ct<T2, NY, NZ>(g, stripes + ((size_t)band * g.nr1 + x) * g.nstripe, pl, py, pz, ty, tz);
        // each acc position belongs to one butterfly of a fixed thread in every band
        if constexpr (REG) z_last_dens_reg_ct<T2, NY, NZ>(pl, pz, tz, s1, s2, accr);
        else z_last_dens_ct<T2, NY, NZ>(pl, pz, tz, s1, s2, acc);
      }
      if constexpr (REG) dens_reg_store_ct<NY, NZ>(pz, accr, dst);
    }
  } else {
    for (int i = threadIdx.x; i < g.nzp; i += kYZ) zpos[i] = g.zpos[i];
    for (int i = threadIdx.x; i < g.nr2; i += kYZ) ty[i] = twy[i];
    for (int i = threadIdx.x; i < g.nr3; i += kYZ) tz[i] = twz[i];
    for (int band = bstart; band < bend; ++band) {
      const double s1 = occ2[2 * band], s2 = occ2[2 * band + 1];
      if (s1 == 0.0 && s2 == 0.0) continue;   // uniform across the workgroup
      __syncthreads();
      yz_forward<T2, BIG>(g, fy, fz, ty, tz, stripes + ((size_t)band * g.nr1 + x) * g.nstripe, pl, zpos);
      // each acc position belongs to one butterfly of a fixed thread in every band
      if constexpr (REG) {
#define PW_CALL(R) if constexpr (R <= kRegR) z_last_dens_reg<T2, R>(g, pl, g.nr3, tz, s1, s2, accr)
        PW_RADIX_SWITCH(fz.radix[fz.nstage - 1], PW_CALL)
#undef PW_CALL
      } else {
#define PW_CALL(R) z_last_dens<T2, R>(g, pl, g.nr3, tz, s1, s2, acc)
        PW_RADIX_SWITCH(fz.radix[fz.nstage - 1], PW_CALL)
#undef PW_CALL
      }
    }
    if constexpr (REG) {
#define PW_CALL(R) if constexpr (R <= kRegR) dens_reg_store<R>(g, g.nr3, accr, dst)
      PW_RADIX_SWITCH(fz.radix[fz.nstage - 1], PW_CALL)
#undef PW_CALL
    }
  }
  if constexpr (!REG) {
    __syncthreads();
    for (int i = threadIdx.x; i < n23; i += kYZ) dst[i] = acc[i];
  }
}

// out[b][a] = (sum_{k<nsum} in[k][a][b]) for in as nsum x [A][B]; 32x32 LDS tiles
template <typename TI, typename TO>
__global__ void k_transpose_sum(int A, int B, int nsum, const TI *in, TO *out) {
  __shared__ double tile[32][33];
  const int a0 = blockIdx.y * 32, b0 = blockIdx.x * 32;
  const int tx = threadIdx.x % 32, ty = threadIdx.x / 32;   // 32 x 8 threads
  for (int r = ty; r < 32; r += 8) {
    const int a = a0 + r, b = b0 + tx;
    double v = 0.0;
    if (a < A && b < B)
      for (int k = 0; k < nsum; ++k) v += (double)in[((size_t)k * A + a) * B + b];
    tile[r][tx] = v;
  }
  __syncthreads();
  for (int r = ty; r < 32; r += 8) {
    const int b = b0 + r, a = a0 + tx;
    if (a < A && b < B) out[(size_t)b * A + a] = (TO)tile[tx][r];
  }
}

// ---- z pass with full-grid output (PLANEWAVE$FFT 'GTOR') ----------------------------
template <typename T2>
__global__ void __launch_bounds__(kBlock) k_zout(Geo g, Fft1D f, const T2 *tw, const T2 *planes,
                                                 T2 *fofr) {
  extern __shared__ unsigned char smem_raw[];
  T2 *b0 = reinterpret_cast<T2 *>(smem_raw);
  T2 *b1 = b0 + kTX * g.nr3;
  const int ntx = (g.nr1 + kTX - 1) / kTX;
  const int y = blockIdx.x / ntx;
  const int x0 = (blockIdx.x % ntx) * kTX;
  const int band = blockIdx.y;
  load_zpencils<T2>(g, planes + (size_t)band * g.nzp * g.nr2 * g.nr1, y, x0, b0);
  __syncthreads();
  T2 *r = lds_fft<T2>(b0, b1, f, kTX, kTX, 1, tw, +1, true);
  T2 *dst = fofr + (size_t)band * g.nr1 * g.nr2 * g.nr3;
  for (int w = threadIdx.x; w < g.nr3 * kTX; w += blockDim.x) {
    const int z = w / kTX, x = w % kTX, gx = x0 + x;
    if (gx < g.nr1) dst[((size_t)z * g.nr2 + y) * g.nr1 + gx] = r[w];
  }
}

// ---- z pass with full-grid input (PLANEWAVE$FFT 'RTOG') -----------------------------
template <typename T2>
__global__ void __launch_bounds__(kBlock) k_zin(Geo g, Fft1D f, const T2 *tw, const T2 *fofr,
                                                T2 *planes) {
  extern __shared__ unsigned char smem_raw[];
  T2 *b0 = reinterpret_cast<T2 *>(smem_raw);
  T2 *b1 = b0 + kTX * g.nr3;
  const int ntx = (g.nr1 + kTX - 1) / kTX;
  const int y = blockIdx.x / ntx;
  const int x0 = (blockIdx.x % ntx) * kTX;
  const int band = blockIdx.y;
  const T2 *src = fofr + (size_t)band * g.nr1 * g.nr2 * g.nr3;
  for (int w = threadIdx.x; w < g.nr3 * kTX; w += blockDim.x) {
    const int z = w / kTX, x = w % kTX, gx = x0 + x;
    b0[w] = gx < g.nr1 ? src[((size_t)z * g.nr2 + y) * g.nr1 + gx] : czero<T2>();
  }
  __syncthreads();
  T2 *r = lds_fft<T2>(b0, b1, f, kTX, kTX, 1, tw, -1, true);
  store_zpencils<T2>(g, planes + (size_t)band * g.nzp * g.nr2 * g.nr1, y, x0, r);
}

Geo make_geo(const pw_plan *p) {
  Geo g;
  g.nr1 = p->nr1; g.nr2 = p->nr2; g.nr3 = p->nr3; g.ngl = p->ngl;
  g.nstripe = p->nstripe; g.nzp = p->nzp; g.pitch = p->yzpitch;
  g.mnzp = fdiv_magic((unsigned)std::max(1, p->nzp)); g.mnr2 = fdiv_magic((unsigned)p->nr2);
  g.stripe_off = p->d_stripe_off; g.stripe_g = p->d_stripe_g; g.stripe_x = p->d_stripe_x;
  g.stripe_s = p->d_stripe_s; g.yz2stripe = p->d_yz2stripe; g.z2iz = p->d_z2iz;
  g.zlist = p->d_zlist; g.stripe_pos = p->d_stripe_pos; g.zpos = p->d_zpos; g.g2 = p->d_g2;
  g.ypen = p->d_ypen; g.nypen = (int)p->ypen.size();
  return g;
}

template <typename T2> const T2 *twid(const pw_plan *p, int k);
template <> const double2 *twid<double2>(const pw_plan *p, int k) { return p->d_tw[k]; }
template <> const float2 *twid<float2>(const pw_plan *p, int k) { return p->d_twf[k]; }

// the plan's x factorisation equals the compile-time one of length n
bool x_is(const Fft1D &f, int n) {
  const Fft1D c = fft_stockham_ct(n);
  if (f.n != n || f.nstage != c.nstage) return false;
  for (int st = 0; st < c.nstage; ++st)
    if (f.radix[st] != c.radix[st]) return false;
  return true;
}

// calls f(IC<NX>) with the x length when a compile-time instantiation exists
template <int N> using ICX = std::integral_constant<int, N>;
template <class F>
int with_x(const pw_plan *p, F &&f) {
  if (x_is(p->f[0], 72)) return f(ICX<72>{});
  if (x_is(p->f[0], 36)) return f(ICX<36>{});
  if (x_is(p->f[0], 64)) return f(ICX<64>{});
  if (x_is(p->f[0], 80)) return f(ICX<80>{});
  return f(ICX<0>{});
}

template <typename T2>
int run_forward_x(pw_plan *p, const Geo &g, int nb, const double2 *psi, hipStream_t s) {
  T2 *stripes = (T2 *)p->d_stripe_buf;
  constexpr int PX = XTile<T2>::n;
  const dim3 gx = x_grid((p->nstripe + PX - 1) / PX, nb);
  const size_t lds = 2 * PX * x_pitch(p->nr1) * sizeof(T2);
  PW_TRY(with_x(p, [&](auto nx) -> int {
    k_xfwd<T2, decltype(nx)::value><<<gx, kBlock, lds, s>>>(g, p->f[0], twid<T2>(p, 0), psi, stripes, nb);
    return PW_OK;
  }));
  PW_HIP(hipGetLastError());
  return PW_OK;
}

template <typename T2>
int run_inverse_x(pw_plan *p, const Geo &g, int nb, double2 *out, const double2 *psi, double scale,
                  int add_kin, hipStream_t s) {
  T2 *stripes = (T2 *)p->d_stripe_buf;
  constexpr int PX = XTile<T2>::n;
  const dim3 gx = x_grid((p->nstripe + PX - 1) / PX, nb);
  const size_t lds = 2 * PX * x_pitch(p->nr1) * sizeof(T2);
  PW_TRY(with_x(p, [&](auto nx) -> int {
    k_xinv<T2, decltype(nx)::value><<<gx, kBlock, lds, s>>>(g, p->f[0], twid<T2>(p, 0), stripes, out, psi, scale,
                                                            add_kin, nb);
    return PW_OK;
  }));
  PW_HIP(hipGetLastError());
  return PW_OK;
}

template <typename T2>
int run_forward_xy(pw_plan *p, const Geo &g, int nb, const double2 *psi, T2 *stripes, T2 *planes,
                   hipStream_t s) {
  const int ntx = (p->nr1 + kTX - 1) / kTX;
  constexpr int PX = XTile<T2>::n;
  const dim3 gx = x_grid((p->nstripe + PX - 1) / PX, nb);
  k_xfwd<T2><<<gx, kBlock, 2 * PX * x_pitch(p->nr1) * sizeof(T2), s>>>(g, p->f[0], twid<T2>(p, 0), psi,
                                                              stripes, nb);
  dim3 gy(p->nzp * ntx, nb);
  k_yfwd<T2><<<gy, kBlock, 2 * kTX * p->nr2 * sizeof(T2), s>>>(g, p->f[1], twid<T2>(p, 1),
                                                               stripes, planes);
  PW_HIP(hipGetLastError());
  return PW_OK;
}

template <typename T2>
int run_inverse_yx(pw_plan *p, const Geo &g, int nb, double2 *out, T2 *stripes, T2 *planes,
                   hipStream_t s) {
  const int ntx = (p->nr1 + kTX - 1) / kTX;
  dim3 gy(p->nzp * ntx, nb);
  k_yinv<T2><<<gy, kBlock, 2 * kTX * p->nr2 * sizeof(T2), s>>>(g, p->f[1], twid<T2>(p, 1),
                                                               planes, stripes);
  constexpr int PX = XTile<T2>::n;
  const dim3 gx = x_grid((p->nstripe + PX - 1) / PX, nb);
  k_xinv<T2><<<gx, kBlock, 2 * PX * x_pitch(p->nr1) * sizeof(T2), s>>>(g, p->f[0], twid<T2>(p, 0),
                                                              stripes, out, nullptr, 1.0 / (double)p->nrl, 0, nb);
  PW_HIP(hipGetLastError());
  return PW_OK;
}

// dynamic LDS of a yz workgroup; beyond 64 KiB the kernel's limit is raised
template <typename K>
int yz_lds(K kernel, size_t lds) {
  if (lds > 65536)
    PW_HIP(hipFuncSetAttribute((const void *)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  return PW_OK;
}

// a radix-11 stage in the in-place y or z transforms needs the BIG kernels
bool yz_big(const pw_plan *p) {
  for (int k = 0; k < 2; ++k)
    for (int st = 0; st < p->fyz[k].nstage; ++st)
      if (p->fyz[k].radix[st] == 11) return true;
  return false;
}

int yz_fits(const pw_plan *p, size_t lds) {
  // one (x, band) plane must stay resident in LDS
  if (lds > 160 * 1024 - 1024)
    return pw_fail(PW_ERR_UNSUPPORTED, "WAVES_VPSI/DENSITY: the NR2 x NR3 plane does not fit in LDS");
  return PW_OK;
}

template <typename T2>
int vt_of(pw_plan *p, const double *v, hipStream_t s) {
  typedef typename Prec<T2>::T T;
  const size_t need = (size_t)p->nrl * sizeof(T);
  if (need > p->vt_bytes) {
    if (p->d_vt) PW_HIP(hipFree(p->d_vt));
    PW_HIP(hipMalloc(&p->d_vt, need));
    p->vt_bytes = need;
  }
  const int A = p->nr2 * p->nr3, B = p->nr1;   // v is [zy][x]; vt is [x][zy]
  dim3 grid((B + 31) / 32, (A + 31) / 32);
  k_transpose_sum<double, T><<<grid, 256, 0, s>>>(A, B, 1, v, (T *)p->d_vt);
  PW_HIP(hipGetLastError());
  return PW_OK;
}

// the plan's in-place factorisation of length n equals the compile-time one
bool yz_is(const Fft1D &f, int n) {
  const Fft1D c = fft_inplace_ct(n);
  if (f.n != n || f.nstage != c.nstage) return false;
  for (int st = 0; st < c.nstage; ++st)
    if (f.radix[st] != c.radix[st] || f.mnbf[st] != c.mnbf[st] || f.mlp[st] != c.mlp[st]) return false;
  return true;
}
// the plan's yz plane is NY x NZ with the compile-time factorisations
bool yz_match(const pw_plan *p, int ny, int nz) {
  return p->nr2 == ny && p->nr3 == nz && p->yzpitch == (ny | 1) && yz_is(p->fyz[0], ny) && yz_is(p->fyz[1], nz);
}
template <int N> using IC = std::integral_constant<int, N>;
// calls f(IC<NY>, IC<NZ>) with the plane's lengths when a compile-time
// instantiation exists (the cells of the benchmarks and of the request mix:
// 72^3 Si64, 36^3 Si8, 36 x 72 planes of Si16/Si32, 64^3, 80^3), else with 0, 0
template <class F>
int with_yz(const pw_plan *p, F &&f) {
  if (yz_match(p, 72, 72)) return f(IC<72>{}, IC<72>{});
  if (yz_match(p, 36, 36)) return f(IC<36>{}, IC<36>{});
  if (yz_match(p, 36, 72)) return f(IC<36>{}, IC<72>{});
  if (yz_match(p, 72, 36)) return f(IC<72>{}, IC<36>{});
  if (yz_match(p, 64, 64)) return f(IC<64>{}, IC<64>{});
  if (yz_match(p, 80, 80)) return f(IC<80>{}, IC<80>{});
  return f(IC<0>{}, IC<0>{});
}

template <typename T2>
int vpsi_t(pw_plan *p, int nb, const double2 *psi, const double *v, double2 *hpsi, int add_kin,
           hipStream_t s) {
  typedef typename Prec<T2>::T T;
  const size_t lds = ((size_t)p->yzpitch * p->nr3 + p->nr2 + p->nr3) * sizeof(T2) + (size_t)p->nzp * sizeof(int);
  PW_TRY(yz_fits(p, lds));
  Geo g = make_geo(p);
  PW_TRY(vt_of<T2>(p, v, s));
  PW_TRY(run_forward_x<T2>(p, g, nb, psi, s));
  if (yz_big(p)) {
    PW_TRY(yz_lds(k_yz_vpsi<T2, true>, lds));
    k_yz_vpsi<T2, true><<<dim3(p->nr1, nb), kYZ, lds, s>>>(g, p->fyz[0], p->fyz[1], twid<T2>(p, 1), twid<T2>(p, 2),
                                                           (T2 *)p->d_stripe_buf, (const T *)p->d_vt);
  } else {
    PW_TRY(with_yz(p, [&](auto ny, auto nz) -> int {
      constexpr int NY = decltype(ny)::value, NZ = decltype(nz)::value;
      PW_TRY(yz_lds(k_yz_vpsi<T2, false, NY, NZ>, lds));
      k_yz_vpsi<T2, false, NY, NZ><<<dim3(p->nr1, nb), kYZ, lds, s>>>(
          g, p->fyz[0], p->fyz[1], twid<T2>(p, 1), twid<T2>(p, 2), (T2 *)p->d_stripe_buf, (const T *)p->d_vt);
      return PW_OK;
    }));
  }
  PW_HIP(hipGetLastError());
  return run_inverse_x<T2>(p, g, nb, hpsi, psi, 1.0 / (double)p->nrl, add_kin, s);
}

template <typename T2>
int density_t(pw_plan *p, int nb, const double2 *psi, const double *occ2, double *rho, int target_blocks,
              hipStream_t s) {
  const size_t n23 = (size_t)p->nr2 * p->nr3;
  // the band sum in registers when no thread has more than two butterflies of
  // the last z stage (pencil y = tid mod NR2, butterflies tid / NR2 + k (kYZ / NR2))
  const Fft1D &fzl = p->fyz[1];
  const int rl = fzl.radix[fzl.nstage - 1];
  const bool reg = !yz_big(p) && rl <= kRegR && p->nr3 / rl <= 2 * (kYZ / p->nr2);
  const size_t lds = (reg ? 0 : n23 * sizeof(double)) +
                     ((size_t)p->yzpitch * p->nr3 + p->nr2 + p->nr3) * sizeof(T2) + (size_t)p->nzp * sizeof(int);
  PW_TRY(yz_fits(p, lds));
  Geo g = make_geo(p);
  PW_TRY(run_forward_x<T2>(p, g, nb, psi, s));
  // band groups so that the grid holds about target_blocks workgroups
  int ngrp = std::max(1, std::min(nb, (target_blocks + p->nr1 - 1) / p->nr1));
  const int bpg = (nb + ngrp - 1) / ngrp;
  ngrp = (nb + bpg - 1) / bpg;
  const size_t need = (size_t)ngrp * p->nrl * sizeof(double);
  if (need > p->dens_bytes) {
    if (p->d_dens_part) PW_HIP(hipFree(p->d_dens_part));
    PW_HIP(hipMalloc(&p->d_dens_part, need));
    p->dens_bytes = need;
  }
#define PW_DENS(BIG, REG, ...)                                                                             \
  do {                                                                                                     \
    PW_TRY(yz_lds(k_yz_dens<T2, BIG, REG, ##__VA_ARGS__>, lds));                                           \
    k_yz_dens<T2, BIG, REG, ##__VA_ARGS__><<<dim3(p->nr1, ngrp), kYZ, lds, s>>>(g, p->fyz[0], p->fyz[1], twid<T2>(p, 1),  \
                                                                 twid<T2>(p, 2), (const T2 *)p->d_stripe_buf, \
                                                                 nb, bpg, occ2, (double *)p->d_dens_part);  \
  } while (0)
  if (yz_big(p)) {
    PW_DENS(true, false);
  } else {
    PW_TRY(with_yz(p, [&](auto ny, auto nz) -> int {
      constexpr int NY = decltype(ny)::value, NZ = decltype(nz)::value;
      if (reg) PW_DENS(false, true, NY, NZ);
      else PW_DENS(false, false, NY, NZ);
      return PW_OK;
    }));
  }
#undef PW_DENS
  // rho[zy][x] = sum_group part[group][x][zy]
  const int A = p->nr1, B = (int)n23;
  dim3 grid((B + 31) / 32, (A + 31) / 32);
  k_transpose_sum<double, double><<<grid, 256, 0, s>>>(A, B, ngrp, (const double *)p->d_dens_part, rho);
  PW_HIP(hipGetLastError());
  return PW_OK;
}

// full-grid transforms of nb bands through the work buffers stripes
// (nb x NSTRIPE x NR1) and planes (nb x NZPLANE x NR2 x NR1); every element of
// both is written before it is read, so neither needs clearing
template <typename T2>
int gtor_t(pw_plan *p, int nb, const double2 *fofg, T2 *fofr, T2 *stripes, T2 *planes, hipStream_t s) {
  Geo g = make_geo(p);
  PW_TRY(run_forward_xy<T2>(p, g, nb, fofg, stripes, planes, s));
  const int ntx = (p->nr1 + kTX - 1) / kTX;
  dim3 gz(p->nr2 * ntx, nb);
  k_zout<T2><<<gz, kBlock, 2 * kTX * p->nr3 * sizeof(T2), s>>>(g, p->f[2], twid<T2>(p, 2), planes, fofr);
  PW_HIP(hipGetLastError());
  return PW_OK;
}

template <typename T2>
int rtog_t(pw_plan *p, int nb, double2 *fofg, const T2 *fofr, T2 *stripes, T2 *planes, hipStream_t s) {
  Geo g = make_geo(p);
  const int ntx = (p->nr1 + kTX - 1) / kTX;
  dim3 gz(p->nr2 * ntx, nb);
  k_zin<T2><<<gz, kBlock, 2 * kTX * p->nr3 * sizeof(T2), s>>>(g, p->f[2], twid<T2>(p, 2), fofr, planes);
  PW_HIP(hipGetLastError());
  return run_inverse_yx<T2>(p, g, nb, fofg, stripes, planes, s);
}

// full-grid transforms in batches of at most c bands through the given buffers
int full_fft(pw_plan *p, bool gtor, int nfft, int c, double2 *fofg, void *fofr, int prec, void *stripes,
             void *planes, hipStream_t s) {
  for (int b0 = 0; b0 < nfft; b0 += c) {
    const int n = std::min(c, nfft - b0);
    double2 *fg = fofg + (size_t)b0 * p->ngl;
    if (prec == PW_PREC_FP64) {
      double2 *fr = (double2 *)fofr + (size_t)b0 * p->nrl;
      if (gtor) PW_TRY(gtor_t<double2>(p, n, fg, fr, (double2 *)stripes, (double2 *)planes, s));
      else PW_TRY(rtog_t<double2>(p, n, fg, fr, (double2 *)stripes, (double2 *)planes, s));
    } else {
      float2 *fr = (float2 *)fofr + (size_t)b0 * p->nrl;
      if (gtor) PW_TRY(gtor_t<float2>(p, n, fg, fr, (float2 *)stripes, (float2 *)planes, s));
      else PW_TRY(rtog_t<float2>(p, n, fg, fr, (float2 *)stripes, (float2 *)planes, s));
    }
  }
  return PW_OK;
}

}  // namespace

// Largest batch per launch sequence; keeps work buffers bounded.
static int chunk_of(const pw_plan *p, int prec) {
  size_t es = prec == PW_PREC_FP64 ? sizeof(double2) : sizeof(float2);
  size_t per = es * ((size_t)p->nstripe * p->nr1 + (size_t)p->nzp * p->nr2 * p->nr1);
  size_t budget = (size_t)1 << 30;   // 1 GiB of work buffers
  size_t c = budget / std::max<size_t>(per, 1);
  return (int)std::max<size_t>(1, std::min<size_t>(c, 4096));
}

int fft_vpsi(pw_plan *p, int nb, const double2 *psi, const double *v, double2 *hpsi, int add_kin,
             int prec, hipStream_t s) {
  const int c = chunk_of(p, prec);
  PW_TRY(plan_work(p, std::min(nb, c), prec, false));
  for (int b0 = 0; b0 < nb; b0 += c) {
    const int n = std::min(c, nb - b0);
    const double2 *ps = psi + (size_t)b0 * p->ngl;
    double2 *hs = hpsi + (size_t)b0 * p->ngl;
    if (prec == PW_PREC_FP64) PW_TRY(vpsi_t<double2>(p, n, ps, v, hs, add_kin, s));
    else PW_TRY(vpsi_t<float2>(p, n, ps, v, hs, add_kin, s));
  }
  return PW_OK;
}

int fft_density(pw_plan *p, int nb, const double2 *psi, const double *occ2, double *rho, int prec,
                hipStream_t s, int target_blocks) {
  if (nb > chunk_of(p, prec))
    return pw_fail(PW_ERR_UNSUPPORTED, "WAVES_DENSITY: band set exceeds the work-buffer budget");
  PW_TRY(plan_work(p, nb, prec, false));
  if (prec == PW_PREC_FP64) return density_t<double2>(p, nb, psi, occ2, rho, target_blocks, s);
  return density_t<float2>(p, nb, psi, occ2, rho, target_blocks, s);
}

int fft_gtor_full(pw_plan *p, int nfft, const double2 *fofg, void *fofr, int prec, hipStream_t s) {
  const int c = chunk_of(p, prec);
  PW_TRY(plan_work(p, std::min(nfft, c), prec, true));
  return full_fft(p, true, nfft, c, const_cast<double2 *>(fofg), fofr, prec, p->d_stripe_buf, p->d_plane_buf, s);
}

int fft_rtog_full(pw_plan *p, int nfft, double2 *fofg, const void *fofr, int prec, hipStream_t s) {
  const int c = chunk_of(p, prec);
  PW_TRY(plan_work(p, std::min(nfft, c), prec, true));
  return full_fft(p, false, nfft, c, fofg, const_cast<void *>(fofr), prec, p->d_stripe_buf, p->d_plane_buf, s);
}

// ---------------------------------------------------------------------------
// PLANEWAVE$FFT without allocation (pw_fft_noalloc).
//
// The stripe and plane work buffers of the full-grid transforms live in the
// file-scope device array `scratch` (kNoAllocScratch bytes per GPU, held
// until the process ends) instead of the plan's lazily allocated buffers, so
// the call makes no hipMalloc/hipFree.  In host code `scratch` names a host
// stand-in; the device address comes from hipGetSymbolAddress, taken once per
// device with that device current on the first call there, and only that
// address is handed to kernels.  Every call on one GPU shares the array:
// calls that overlap on one GPU (two streams, two threads) race.  The
// kernels write every element of the work buffers before reading it, so the
// contents left by an earlier call need no clearing.
// ---------------------------------------------------------------------------
constexpr size_t kNoAllocScratch = (size_t)64 << 20;
constexpr int kNoAllocMaxDev = 64;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratch];

static std::mutex g_scratch_mu;
static void *g_scratch_dev[kNoAllocMaxDev];

// device address of `scratch` on the current device
static int scratch_ptr(void **out) {
  int dev = 0;
  PW_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kNoAllocMaxDev) return pw_fail(PW_ERR_UNSUPPORTED, "pw_fft_noalloc: device index too large");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      return pw_fail(PW_ERR_HIP, std::string("pw_fft_noalloc: hipGetSymbolAddress(scratch): ") +
                                     hipGetErrorString(e));
    g_scratch_dev[dev] = ptr;
  }
  *out = g_scratch_dev[dev];
  return PW_OK;
}

extern "C" int pw_fft_noalloc(pw_plan *p, const char *id, int nfft, int ngl, void *fofg, int nrl, void *fofr,
                              int prec) {
  if (!p || !id || !fofg || !fofr) return pw_fail(PW_ERR_ARG, "pw_fft_noalloc: null argument");
  if (ngl != p->ngl) return pw_fail(PW_ERR_ARG, "PLANEWAVE$FFT: size of F(G) inconsistent");
  if (nrl != p->nrl) return pw_fail(PW_ERR_ARG, "PLANEWAVE$FFT: size of F(R) inconsistent");
  if (nfft <= 0) return pw_fail(PW_ERR_ARG, "PLANEWAVE$FFT: NFFT must be positive");
  if (prec != PW_PREC_FP64 && prec != PW_PREC_MIXED) return pw_fail(PW_ERR_ARG, "pw_fft_noalloc: unknown precision");
  const std::string s(id);
  if (s != "GTOR" && s != "RTOG") return pw_fail(PW_ERR_ARG, "PLANEWAVE$FFT: ID must be either GTOR or RTOG");
  // work bytes of one transform: stripes, then planes at a 256-byte boundary
  const size_t es = prec == PW_PREC_FP64 ? sizeof(double2) : sizeof(float2);
  const size_t sb = es * (size_t)p->nstripe * p->nr1, pb = es * (size_t)p->nzp * p->nr2 * p->nr1;
  const size_t per = sb + pb + 256;
  if (per > kNoAllocScratch)
    return pw_fail(PW_ERR_ARG, "pw_fft_noalloc: one transform needs " + std::to_string(per) +
                                   " bytes of work space, more than the " + std::to_string(kNoAllocScratch) +
                                   " bytes of the static scratch; use pw_fft");
  if (!p->uploaded)
    return pw_fail(PW_ERR_STATE, "pw_fft_noalloc: the plan's tables are not on the device yet "
                                 "(any pw_fft or waveset call on the plan uploads them)");
  void *base = nullptr;
  PW_TRY(scratch_ptr(&base));
  const int c = (int)std::min<size_t>((size_t)nfft, kNoAllocScratch / per);
  const size_t soff = ((size_t)c * sb + 255) / 256 * 256;
  unsigned char *b = (unsigned char *)base;
  return full_fft(p, s == "GTOR", nfft, c, (double2 *)fofg, fofr, prec, b, b + soff, p->stream);
}
