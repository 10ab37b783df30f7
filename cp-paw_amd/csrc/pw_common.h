// Internal declarations of the pwgpu library (not part of the C ABI).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/pwgpu.h"

// ---------------------------------------------------------------------------
// error plumbing
// ---------------------------------------------------------------------------
void pw_set_error(const std::string &msg);
int pw_fail(int code, const std::string &msg);

// Stream capture (the hipGraph of ORTHO_X) runs only while a single plan and a
// single waveset exist in the process: HIP calls that other threads make
// during a capture (another request's frees among them) invalidate it.  Plans
// and wavesets are counted under the same lock that a capture holds, before
// their first allocation and after their last free.
struct CaptureGate {
  std::mutex mu;
  int plans = 0, wavesets = 0;
};
CaptureGate &capture_gate();

#define PW_HIP(call)                                                         \
  do {                                                                       \
    hipError_t e_ = (call);                                                  \
    if (e_ != hipSuccess)                                                    \
      return pw_fail(PW_ERR_HIP, std::string(#call) + ": " +                 \
                                     hipGetErrorString(e_));                 \
  } while (0)

#define PW_TRY(call)                                                         \
  do {                                                                       \
    int rc_ = (call);                                                        \
    if (rc_ != PW_OK) return rc_;                                            \
  } while (0)

// ---------------------------------------------------------------------------
// complex helpers (interleaved re,im as double2 / float2)
// ---------------------------------------------------------------------------
__host__ __device__ inline double2 cmake(double r, double i) { return make_double2(r, i); }
__host__ __device__ inline float2 cmake(float r, float i) { return make_float2(r, i); }
template <typename T2> struct Real;
template <> struct Real<double2> { typedef double type; };
template <> struct Real<float2> { typedef float type; };

template <typename T2>
__host__ __device__ inline T2 cadd(T2 a, T2 b) { T2 r; r.x = a.x + b.x; r.y = a.y + b.y; return r; }
template <typename T2>
__host__ __device__ inline T2 csub(T2 a, T2 b) { T2 r; r.x = a.x - b.x; r.y = a.y - b.y; return r; }
template <typename T2>
__host__ __device__ inline T2 cmul(T2 a, T2 b) {
  T2 r; r.x = a.x * b.x - a.y * b.y; r.y = a.x * b.y + a.y * b.x; return r;
}
template <typename T2>
__host__ __device__ inline T2 cconj(T2 a) { a.y = -a.y; return a; }
template <typename T2, typename S>
__host__ __device__ inline T2 cscale(T2 a, S s) { a.x *= s; a.y *= s; return a; }

// ---------------------------------------------------------------------------
// 1-D FFT factorisation
// ---------------------------------------------------------------------------
struct Fft1D {
  int n = 0;
  int nstage = 0;
  int radix[12] = {0};
  // ceil(2^32/d) for the in-place stages' quotients by n/radix[st] and by
  // L/radix[st], L = radix[0]...radix[st] (see fdiv in pw_fft.hip)
  unsigned long long mnbf[12] = {0}, mlp[12] = {0};
};
inline unsigned long long fdiv_magic(unsigned d) { return (0x100000000ull + d - 1) / d; }
// inplace: factors for the in-place LDS transforms of the yz pass, which
// take no radix 16 (8 x 2 instead) so that their kernels stay small in registers
bool fft_factor(int n, Fft1D &f, bool inplace = false);

// ---------------------------------------------------------------------------
// PLANEWAVE instance
// ---------------------------------------------------------------------------
struct pw_plan {
  double rbas[9];
  double gbas[9];     // GBAS(3,3) column-major: gbas[i + 3*j] = GBAS(i+1,j+1)
  double vol = 0.0;
  int nr1 = 0, nr2 = 0, nr3 = 0, nrl = 0;
  int ngl = 0;
  int ngamma = -1;    // 0-based
  std::vector<int> igvec;   // 3*ngl
  std::vector<int> minusg;  // 0-based
  std::vector<double> gvec; // 3*ngl
  std::vector<double> g2;   // ngl
  // stripes: (y,z) lines along x that hold sphere vectors
  int nstripe = 0;
  std::vector<int> stripe_off; // nstripe+1, CSR into stripe_g / stripe_x
  std::vector<int> stripe_g;   // ngl sphere indices grouped by stripe
  std::vector<int> stripe_x;   // ngl x-index of each entry
  std::vector<int> stripe_s;   // ngl stripe of each entry
  std::vector<int> yz2stripe;  // nr2*nr3, -1 if empty
  std::vector<int> stripe_pos; // nstripe: digit-reversed z*yzpitch + y of each stripe
  std::vector<int> zpos;       // nzp: digit-reversed z*yzpitch of each stripe-holding z plane
  std::vector<int> ypen;       // the zpos rows in windows of 16 lane slots (-1: empty slot)
  int yzpitch = 0;             // row pitch of the LDS-resident (y, z) plane: nr2 made odd
  int nzp = 0;                 // z planes holding stripes
  std::vector<int> z2iz;       // nr3, -1 if empty
  std::vector<int> zlist;      // nzp: z of each used plane
  int nmax[3] = {0, 0, 0};     // max |IGVEC(k)|
  Fft1D f[3];
  Fft1D fyz[2];                // y and z factors of the in-place yz pass (no radix 16)
  // device copies
  int *d_minusg = nullptr, *d_igvec = nullptr;
  int *d_stripe_off = nullptr, *d_stripe_g = nullptr, *d_stripe_x = nullptr;
  int *d_stripe_s = nullptr;
  int *d_yz2stripe = nullptr, *d_z2iz = nullptr, *d_zlist = nullptr;
  int *d_stripe_pos = nullptr, *d_zpos = nullptr, *d_ypen = nullptr;
  double *d_g2 = nullptr, *d_gvec = nullptr;
  double2 *d_tw[3] = {nullptr, nullptr, nullptr};
  float2 *d_twf[3] = {nullptr, nullptr, nullptr};
  hipStream_t stream = 0;
  bool uploaded = false;
  // growable work buffers
  void *d_stripe_buf = nullptr; size_t stripe_bytes = 0;
  void *d_plane_buf = nullptr; size_t plane_bytes = 0;
  void *d_dens_part = nullptr; size_t dens_bytes = 0;
  void *d_vt = nullptr; size_t vt_bytes = 0;   // V(r) as [x][z][y] in the work precision
  // POTENTIAL$VOFRHO on this (density) plane-wave object: setup tables and work
  double *d_pot_tabs = nullptr; int pot_nsp = 0;
  std::vector<int> pot_lmrx;   // LMRX per species (POTENTIAL_HARTREE's LMRX(NSP))
  void *d_pot = nullptr; size_t pot_bytes = 0;
  double *h_pot = nullptr; size_t h_pot_bytes = 0; int pot_nat = 0;   // pinned: results and staged atom data
  int pot_lmrxx = 1;           // LMRXX of the last call
  int pot_xc = 1;              // functional type (DFT$SETI4('TYPE')): 1 LDA (PZ), 10 PBE
  double *d_pot_res = nullptr; // device results of the last call: per atom VQLM(LMRXX), force(3);
                               // then EHARTREE, EXC, RHOB
};

// POTENTIAL$VOFRHO on a density plan (pw_potential.hip): enqueue on stream s
// (isp 1-based, pos R(3,NAT), QLM(LMRXX,NAT) from the host or the device);
// results are valid after the stream passes.
int vofrho_enqueue(pw_plan *dp, int nat, int lmrxx, const int *isp, const double *pos, const double *qlm,
                   const double *qlm_dev, const double *rho_dev, double *vofr_dev, hipStream_t s);
void vofrho_results(const pw_plan *dp, int nat, double *energies, double *vqlm, double *force);

// stripe work buffer for nbatch bands (and the z-plane buffer of the full-grid
// transforms when planes is set)
int plan_work(pw_plan *p, int nbatch, int prec, bool planes);
int plan_device(pw_plan *p);   // uploads maps and twiddles on first use

// FFT pipelines (pw_fft.hip).  T selects the arithmetic (FP64 / FP32).
// psi/fofg are complex128 sphere arrays (NGL per band).
struct FftCtx;
int fft_gtor_full(pw_plan *p, int nfft, const double2 *fofg, void *fofr, int prec, hipStream_t s);
int fft_rtog_full(pw_plan *p, int nfft, double2 *fofg, const void *fofr, int prec, hipStream_t s);
// hpsi = FFT^-1[v FFT psi] + (add_kin ? G^2/2 psi : 0)
int fft_vpsi(pw_plan *p, int nb, const double2 *psi, const double *v, double2 *hpsi,
             int add_kin, int prec, hipStream_t s);
// rho(r) = sum_b s1[b] |psi_b(r)|^2 + s2[b] Re psi_b(r)^2; occ2 = (s1,s2) per band
int fft_density(pw_plan *p, int nb, const double2 *psi, const double *occ2_dev,
                double *rho, int prec, hipStream_t s,
                int target_blocks = 512);

// ---------------------------------------------------------------------------
// WAVES state
// ---------------------------------------------------------------------------
struct KernelTimer;

struct OrthoGraphKey {
  double lo = 0.0, hi = 0.0;
  int nlev = 0, nns = 0, batch = 0;
  const void *lam = nullptr, *chichi = nullptr, *psipsi = nullptr, *chipsi = nullptr, *work = nullptr;
  bool operator==(const OrthoGraphKey &o) const {
    return lo == o.lo && hi == o.hi && nlev == o.nlev && nns == o.nns && batch == o.batch && lam == o.lam &&
           chichi == o.chichi && psipsi == o.psipsi && chipsi == o.chipsi && work == o.work;
  }
};

struct pw_waveset {
  pw_plan *plan = nullptr;
  int prec = PW_PREC_FP64;
  int nb = 0, nbh = 0, nat = 0, nsp = 0, lnxx = 0, lmnxx = 0, lmx = 0;
  int npro = 0, nbarepro = 0;
  std::vector<int> isp;        // 0-based
  std::vector<int> lnx, lmnx;  // per species
  std::vector<int> lox;        // lnxx*nsp
  std::vector<int> ipro0;      // per atom first projector
  std::vector<int> ibpro;      // per species first bare projector
  // per species lmn -> (ln, lm, l), lmnxx slots each
  std::vector<int> lmn_ln, lmn_lm, lmn_l;
  // parameters
  double delt = 5.0, emass = 1000.0, emasscg2 = 0.0, annee = 0.0;
  bool have_mass = false;
  std::vector<double> occ;
  std::vector<double> do_ln;   // lnxx*lnxx*nsp
  std::vector<double> do_lmn;  // per species lmnxx*lmnxx (row-major [i*lmnxx+j])
  std::vector<double> dh;      // per atom lmnxx*lmnxx complex interleaved
  std::vector<double> r0, rp, rm;
  // density -> potential step linked into ETOT (pw_waves_set_potential)
  pw_plan *dplan = nullptr;
  std::vector<int> isp1;             // NAT, 1-based
  std::vector<double> qlm;           // QLM(LMRXX,NAT)
  int qlm_lmrxx = 1;
  void *sph = nullptr;               // one-centre set-up and work (pw_augment.hip)
  double potb = 0.0;                 // POTB of the last ETOT with the one-centre terms
  double ehartree = 0.0, exc = 0.0, rhob = 0.0;
  std::vector<double> vqlm, fpot;    // dE/dQLM (NAT), potential forces (3 NAT)
  // lambda (real, Gamma point)
  std::vector<double> rlam0, rlamm;  // host copies, filled by the getters
  bool have_rlam0 = false, have_rlamm = false;   // history depth of the multipliers (NLAMBDA)
  double *d_rlam0 = nullptr, *d_rlamm = nullptr, *d_occ = nullptr;   // RLAM0, RLAMM (NB x NB), OCC (NB)
  bool have_mats = false;            // MAT/OMAT/OOMAT of the last ORTHOGONALIZE on the device
  bool first = true;
  // results
  double ekin = 0, epot = 0, e1c = 0, etot = 0, wavekin1 = 0, wavekin2 = 0;
  int ortho_iter = 0;
  int ortho_levels = 0, ortho_ns = 0, ortho_series = 0;
  // spectral bounds of sym(CHIPSI) from the last Lanczos run, reused while
  // the device checks of ORTHO_X accept them
  double lz_lo = 0.0, lz_hi = 0.0;
  bool lz_valid = false;
  int ortho_lanczos = 0;       // Lanczos runs in the last ORTHO_X call
  int ortho_eigen = 0;         // ORTHO_X solves through the eigenvectors of sym(CHIPSI) (B not positive definite)
  int ortho_graph_misses = 0;  // ORTHO_X solves issued without a graph (other objects alive)
  // launch shapes chosen by pw_waves_autotune
  int tune_split_waves = 2048;  // waves of a G-split GEMM launch
  int tune_col_waves = 4;       // waves (16 G each) of a column-product workgroup (mixed form: 2 or 4)
  // device results of a step section and their pinned host copy; the host
  // block has NB*NB more doubles that stage the DEDPROJ constraint matrix
  double *d_res = nullptr, *h_res = nullptr;
  size_t res_n = 0;
  size_t r_ekin = 0, r_vrho = 0, r_wk = 0, r_expect = 0, r_denmat = 0, r_force = 0, r_resid = 0;
  // ETOT's and ORTHOGONALIZE's sums are copied into h_res without a wait:
  // res_ev follows the copy, and ws_finish (called by every entry point that
  // reads or changes the host state) waits for it and completes them
  hipEvent_t res_ev = nullptr;
  int res_pending = 0;         // 1: ETOT's sums, 2: ORTHOGONALIZE's
  double pend_wk1 = 0.0;       // WAVEKIN1 when ORTHOGONALIZE's sums were enqueued
  bool pend_switched = false;  // SWITCH ran after them: WAVEKIN2 stays zero
  // residual of the last ORTHOGONALIZE (max over super bands of the norm of
  // the constrained force), evaluated when residual_on
  bool residual_on = false;
  double residual = 0.0;
  int tune_dens_blocks = 512;   // workgroups of the density yz pass
  std::vector<double> force;   // 3*nat
  std::vector<double> eig;     // nb
  std::vector<double> mat, omat, oomat;
  std::vector<double> denmat;  // lmnxx*lmnxx*nat
  bool have_hpsi = false;
  bool have_v = false, have_psi0 = false, have_psim = false, have_dh = false;
  double efict = 0.0;
  // device arrays
  double2 *d_psi0 = nullptr, *d_psim = nullptr, *d_hpsi = nullptr, *d_opsi = nullptr;
  double2 *d_tmp = nullptr;    // NBH*NGL scratch (Gram-Schmidt copy)
  double2 *d_proj = nullptr, *d_oproj = nullptr, *d_xproj = nullptr;
  double *d_pro = nullptr, *d_ylm = nullptr, *d_mpsi = nullptr;
  double *d_v = nullptr, *d_rho = nullptr;
  double *d_occ2 = nullptr;    // (s1,s2) per super band
  double *d_fpm = nullptr;     // (FP,FM) of WAVES_EKIN per super band
  double *d_fab = nullptr;     // ((f1+f2)/2,(f1-f2)/2) of WAVES_WAVEKINETIC
  double2 *d_domat = nullptr;  // per atom DO (LMNXX x LMNXX) as complex
  double2 *d_dhmat = nullptr;  // per atom DH (LMNXX x LMNXX)
  int *d_atom_isp = nullptr, *d_ipro0 = nullptr, *d_ibpro = nullptr, *d_lmnx = nullptr;
  int *d_lmn_ln = nullptr, *d_lmn_lm = nullptr, *d_lmn_l = nullptr;
  int2 *d_ipro_tab = nullptr;  // (atom, lmn) of each projector
  // mixed form: factors of the projector functions for the positions
  // pfac_pos, P(atom,lmn,G) = (-i)^l F(isp,lmn,G) E(atom,G):
  // F = YLM*PRO [NSP*LMNXX][NGLP] doubles, then E = exp(-iG.R) [NAT][NGLP]
  // double2, NGLP = NGL padded to 64 with zeros
  void *d_pfac = nullptr;
  size_t pfac_bytes = 0;
  std::vector<double> pfac_pos;
  bool pfac_valid = false;
  double2 *d_phase = nullptr;  // nat*3*(2*nmaxx+1)
  int nphase = 0;              // 2*nmaxx+1
  std::vector<double> phase_pos;   // positions of the tables in d_phase
  bool phase_valid = false;
  double *d_part = nullptr; size_t part_bytes = 0;   // split-K partials
  void *d_split = nullptr; size_t split_bytes = 0;   // mixed form: operands split into bf16 runs
  // mixed form: bf16 runs of wave-function arrays (rows, or rows at -G) kept
  // for the products of one step section while runs_scope is set (the arrays
  // do not change inside such a section)
  struct WaveRuns {
    const void *psi = nullptr;
    bool minus = false, valid = false;
  };
  void *d_wruns = nullptr; size_t wruns_bytes = 0;
  WaveRuns wruns[4];
  bool runs_scope = false;
  double *d_small = nullptr;   // small matrices / scalars scratch
  double *d_small_mats = nullptr;  // MAT, OMAT, OOMAT, LAMBDA, GS work (NB*NB each)
  size_t small_bytes = 0;
  // orthogonalisation work (device, FP64)
  double *d_ortho = nullptr;
  void *h_ortho_state = nullptr;   // pinned: ORTHO_X loop state read back per batch
  size_t ortho_cap = 0;
  // captured graph of the ORTHO_X preparation and first iteration batch
  void *ortho_exec = nullptr;
  OrthoGraphKey ortho_key;
  hipStream_t stream = 0;
  bool own_stream = false;
  // timing
  bool profile = false;
  std::map<std::string, std::pair<double, int>> ktime;
  std::vector<std::pair<std::string, std::pair<hipEvent_t, hipEvent_t>>> pending;
};

// kernel launch bracket for optional per-kernel timing
struct ScopeTimer {
  pw_waveset *ws; const char *name; hipEvent_t a = nullptr, b = nullptr;
  ScopeTimer(pw_waveset *w, const char *n);
  ~ScopeTimer();
};
void ws_flush_timers(pw_waveset *ws);
// a waveset as created, keeping its buffers, stream, graph and launch shapes
int ws_finish(pw_waveset *ws);
int ws_reset_run(pw_waveset *ws);

// real spherical harmonics at one direction, LM(L,M) = 1+L^2+L-M, L <= lx <= 6
// (SPHERICAL$YLM; pw_waves.cpp)
void ylm_dir(int lx, double gx, double gy, double gz, double *y);

// one-centre terms (pw_augment.hip): set-up state lives in ws->sph
void sphere_release(pw_waveset *ws);
bool sphere_ready(const pw_waveset *ws);
int sphere_lmrxx(const pw_waveset *ws);
int sphere_lmrx(const pw_waveset *ws, int isp);
int sphere_moments(pw_waveset *ws, const double **qlm_dev);
int sphere_terms(pw_waveset *ws, const double *vqlm_dev, const double *rhob_dev);
void sphere_results(const pw_waveset *ws, double *e1c, double *potb);
const double *sphere_qlm_dev(const pw_waveset *ws);

// kernels in pw_kernels.hip
int k_phase_tables(pw_waveset *ws, const double *r_dev_host, double2 *phase);
int k_pfac(pw_waveset *ws);
int k_runs_scope(pw_waveset *ws, bool on);
int k_projections(pw_waveset *ws, const double2 *psi, const double2 *phase, double2 *proj);
int k_addpro(pw_waveset *ws, double2 *psi, const double2 *phase, const double2 *proj);
// opsi = dt^2/((1+a) m(G)) (psi0 + |p> proj) in one pass
int k_addpro_opsi(pw_waveset *ws, double2 *opsi, const double2 *psi0, const double2 *phase, const double2 *proj);
// reductions: partial sums into device memory dst (read back once per section)
int k_force(pw_waveset *ws, const double2 *psi, const double2 *phase, const double2 *dedproj, double *dst);
// max over super bands of |HPSI - (1+a) m/dt^2 (psi(+) - psibar)| into dst (device)
int k_residual(pw_waveset *ws, const double2 *psibar, double *dst);
int k_overlap(pw_waveset *ws, const double2 *psi1, const double2 *psi2, double *mat_host_or_dev, bool to_host, bool tid);
// mat_dev = WAVES_OVERLAP(psi1, psi2) + WAVES_1COVERLAP(proj1, proj2) in one sequence
int k_overlap_full(pw_waveset *ws, const double2 *psi1, const double2 *psi2, const double2 *proj1,
                   const double2 *proj2, double *mat_dev);
int k_addopsi(pw_waveset *ws, double2 *psibar, const double2 *chi, const double *lam_dev);
int k_addoproj(pw_waveset *ws, double2 *proj, const double2 *oproj, const double *lam_dev);
int k_atom_matrix(pw_waveset *ws, const double2 *proj, double2 *out, const double2 *mats);
int k_fold_occ(pw_waveset *ws, double2 *x);
int red_blocks(const pw_waveset *ws);
int expect_chunks(const pw_waveset *ws);
int vrho_blocks(const pw_waveset *ws);
int k_wavekinetic(pw_waveset *ws, const double2 *psi0, const double2 *psim, const double *fab_dev, double *dst);
// EKIN, WAVEKINETIC (first half) and <psi|H|psi> of WAVES$ETOT in one pass
int k_etot_sums(pw_waveset *ws, double *ekin, double *wk, double *expect);
int k_propagate(pw_waveset *ws);
int k_copy(pw_waveset *ws, double2 *dst, const double2 *src, size_t n);
int k_denmat(pw_waveset *ws, const double2 *proj, double *dst);
int k_dot_vrho(pw_waveset *ws, double *dst);
int ws_part(pw_waveset *ws, size_t bytes);
int ws_small(pw_waveset *ws, size_t bytes);

// orthogonalisation (pw_ortho.hip)
void ortho_graph_release(pw_waveset *ws);
int ortho_x_device(pw_waveset *ws, double *d_lam, const double *d_chichi,
                   const double *d_psipsi, const double *d_chipsi, int *iter);
