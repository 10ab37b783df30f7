// WAVES object on the device: set-up, the CP step and the C ABI.
//
// Step order follows WAVES$ETOT (src/paw_waves1.f90:1610-2233),
// WAVES$PROPAGATE (:6479), WAVES$ORTHOGONALIZE (src/paw_waves2.f90:11) and
// WAVES$SWITCH (src/paw_waves2.f90:2180) for one k-point, one spin, NDIM=1
// and super wave functions.  The density->potential step and AUGMENTATION are
// outside this library: the caller provides V(r) ("VOFR") and the one-centre
// matrices DH and DO; the energies reported are the ones the wave-function
// step computes from them (EKIN, int V rho, sum DH D).
#include <algorithm>
#include <cmath>
#include <cstring>

#include "pw_common.h"

int k_fold_occ(pw_waveset *ws, double2 *x);

// ---------------------------------------------------------------------------
// timers
// ---------------------------------------------------------------------------
ScopeTimer::ScopeTimer(pw_waveset *w, const char *n) : ws(w), name(n) {
  if (!ws || !ws->profile) return;
  if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) { a = b = nullptr; return; }
  (void)hipEventRecord(a, ws->stream);
}
ScopeTimer::~ScopeTimer() {
  if (!a) return;
  (void)hipEventRecord(b, ws->stream);
  ws->pending.push_back({name, {a, b}});
}
void ws_flush_timers(pw_waveset *ws) {
  if (ws->pending.empty()) return;
  (void)hipStreamSynchronize(ws->stream);
  for (auto &p : ws->pending) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, p.second.first, p.second.second) == hipSuccess) {
      auto &t = ws->ktime[p.first];
      t.first += ms;
      t.second += 1;
    }
    (void)hipEventDestroy(p.second.first);
    (void)hipEventDestroy(p.second.second);
  }
  ws->pending.clear();
}

// ---------------------------------------------------------------------------
// helpers
// ---------------------------------------------------------------------------
static int dalloc(void **p, size_t bytes) {
  PW_HIP(hipMalloc(p, std::max<size_t>(bytes, 8)));
  PW_HIP(hipMemset(*p, 0, std::max<size_t>(bytes, 8)));
  return PW_OK;
}

static int need(bool ok, const char *what) {
  if (!ok) return pw_fail(PW_ERR_STATE, std::string("WAVES: ") + what + " has not been set");
  return PW_OK;
}

static double *d_mat(pw_waveset *ws, int k) {   // 0 MAT, 1 OMAT, 2 OOMAT, 3 LAMBDA, 4 GS
  return ws->d_small_mats + (size_t)k * ws->nb * ws->nb;
}

// LAMBDA of DEDPROJ: -RLAM0 (f_i + f_j)/2 (src/paw_waves1.f90:5592-5691)
__global__ void k_lam_dedproj(int nb, const double *occ, const double *rlam0, double *lam) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= nb * nb) return;
  const int i = idx / nb, j = idx - i * nb;
  lam[idx] = -rlam0[idx] * 0.5 * (occ[i] + occ[j]);
}

// initial LAMBDA of ORTHO_X: hermitised RLAM0 times 0.5 f_i/(f_i+f_j)
// (src/paw_waves2.f90:295-359)
__global__ void k_lam_guess(int nb, const double *occ, const double *rlam0, double *lam) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= nb * nb) return;
  const int i = idx / nb, j = idx - i * nb;
  double v = rlam0[idx];
  if (i != j) v = 0.5 * (rlam0[idx] + rlam0[(size_t)j * nb + i]);
  const double oi = occ[i];
  double oj = occ[j];
  if (oi + oj < 1.0e-12) oj = 1.0e-12;
  lam[idx] = v * 0.5 * oi / (oi + oj);
}

// WAVES$SWITCH with NLAMBDAX=2 (src/paw_waves2.f90:2236-2297):
// first: RLAMM = RLAM0; then RLAM(+) = 2 RLAM0 - RLAMM, RLAMM = RLAM0
__global__ void k_lam_switch(int n, int first, double *rlam0, double *rlamm) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n) return;
  const double l0 = rlam0[idx];
  if (!first) rlam0[idx] = 2.0 * l0 - rlamm[idx];
  rlamm[idx] = l0;
}

__global__ void k_copy_mat(int n, const double *src, double *dst) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx < n) dst[idx] = src[idx];
}

static int upload_occ(pw_waveset *ws) {
  const int nbh = ws->nbh;
  std::vector<double> occ2(2 * nbh), fpm(2 * nbh), fab(2 * nbh);
  for (int b = 0; b < nbh; ++b) {
    const double f1 = ws->occ[2 * b];
    const double f2 = (2 * b + 1 < ws->nb) ? ws->occ[2 * b + 1] : 0.0;
    occ2[2 * b] = 0.5 * (f1 + f2);
    occ2[2 * b + 1] = 0.5 * (f1 - f2);
    // WAVES_EKIN (src/paw_waves1.f90:2848-2884)
    double fp = 0.5 * (f1 + f2), fm = 0.5 * (f1 - f2);
    if (std::fabs(fm) < 1.0e-12) fm = 0.0;
    fpm[2 * b] = fp;
    fpm[2 * b + 1] = fm;
    // WAVES_WAVEKINETIC (src/paw_waves1.f90:6778-6822)
    if (f1 == 0.0 && f2 == 0.0) { fab[2 * b] = 0.0; fab[2 * b + 1] = 0.0; }
    else { fab[2 * b] = 0.5 * (f1 + f2); fab[2 * b + 1] = (f1 - f2 == 0.0) ? 0.0 : 0.5 * (f1 - f2); }
  }
  PW_HIP(hipMemcpy(ws->d_occ2, occ2.data(), occ2.size() * sizeof(double), hipMemcpyHostToDevice));
  PW_HIP(hipMemcpy(ws->d_fpm, fpm.data(), fpm.size() * sizeof(double), hipMemcpyHostToDevice));
  PW_HIP(hipMemcpy(ws->d_fab, fab.data(), fab.size() * sizeof(double), hipMemcpyHostToDevice));
  PW_HIP(hipMemcpy(ws->d_occ, ws->occ.data(), ws->nb * sizeof(double), hipMemcpyHostToDevice));
  return PW_OK;
}

static int upload_do(pw_waveset *ws) {
  // DO(LN,LN) -> per atom LMN x LMN (WAVES_OPROJ, src/paw_waves2.f90:497-542)
  const int L = ws->lmnxx;
  std::vector<double2> m((size_t)ws->nat * L * L, make_double2(0.0, 0.0));
  for (int a = 0; a < ws->nat; ++a) {
    const int isp = ws->isp[a];
    int o1 = 0;
    for (int ln1 = 0; ln1 < ws->lnx[isp]; ++ln1) {
      const int l1 = ws->lox[ln1 + ws->lnxx * isp];
      int o2 = 0;
      for (int ln2 = 0; ln2 < ws->lnx[isp]; ++ln2) {
        const int l2 = ws->lox[ln2 + ws->lnxx * isp];
        if (l1 == l2) {
          const double v = ws->do_ln[ln1 + ws->lnxx * (ln2 + ws->lnxx * isp)];
          for (int mm = 0; mm < 2 * l1 + 1; ++mm)
            m[((size_t)a * L + o1 + mm) * L + o2 + mm] = make_double2(v, 0.0);
        }
        o2 += 2 * l2 + 1;
      }
      o1 += 2 * l1 + 1;
    }
  }
  PW_HIP(hipMemcpy(ws->d_domat, m.data(), m.size() * sizeof(double2), hipMemcpyHostToDevice));
  return PW_OK;
}

// Real spherical harmonics at one direction, LM(L,M) = 1+L^2+L-M
// (SPHERICAL$YLM, src/paw_spherical.f90:66-148): associated Legendre
// functions by the recursion of SPHERICAL_PLGNDR (:366-453) scaled by
// sqrt((L-M)!/(L+M)!), then Y(LM0) = c P_{L,0}, Y(LM0-M) = sqrt2 c
// (-1)^M P_{L,M} cos(M phi), Y(LM0+M) = ... sin(M phi), c = sqrt((2L+1)/4pi).
void ylm_dir(int lx, double gx, double gy, double gz, double *y) {
  const int lmx = (lx + 1) * (lx + 1);
  const double disxy = std::sqrt(gx * gx + gy * gy);
  const double dis = std::sqrt(disxy * disxy + gz * gz);
  for (int lm = 0; lm < lmx; ++lm) y[lm] = 0.0;
  y[0] = 1.0 / std::sqrt(4.0 * M_PI);
  if (lx == 0 || std::fabs(dis) < 1.0e-12) return;
  const double x = gz / dis;
  const double fact = std::sqrt(std::max(1.0 - x * x, 0.0));
  double plm[64];
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
      f *= std::sqrt((double)((l - m + 1) * (l + m)));
      sg = -sg;
      plm[i1 - m] /= f;
      plm[i1 + m] = sg * plm[i1 - m];
    }
  }
  const double sinphi = disxy != 0.0 ? gy / disxy : 0.0;
  const double cosphi = disxy != 0.0 ? gx / disxy : 1.0;
  for (int l = 0; l <= lx; ++l) {
    const int lm0 = l * (l + 1);
    const double fac = std::sqrt((2 * l + 1) / (4.0 * M_PI));
    y[lm0] = plm[lm0] * fac;
    double er = 1.0, ei = 0.0;
    for (int m = 1; m <= l; ++m) {
      const double nr = er * cosphi - ei * sinphi, ni = er * sinphi + ei * cosphi;
      er = nr;
      ei = ni;
      const double s = fac * plm[lm0 + m] * std::sqrt(2.0);
      y[lm0 - m] = s * er;
      y[lm0 + m] = s * ei;
    }
  }
}

static int ylm_table(pw_waveset *ws) {
  pw_plan *p = ws->plan;
  const int lx = (int)std::lround(std::sqrt((double)ws->lmx)) - 1;
  if (lx > 6) return pw_fail(PW_ERR_UNSUPPORTED, "WAVES: projectors with l>6 are not supported");
  std::vector<double> y((size_t)ws->lmx * p->ngl, 0.0);
  double yl[64];
  for (int g = 0; g < p->ngl; ++g) {
    ylm_dir(lx, p->gvec[3 * g], p->gvec[3 * g + 1], p->gvec[3 * g + 2], yl);
    for (int lm = 0; lm < ws->lmx; ++lm) y[(size_t)lm * p->ngl + g] = yl[lm];
  }
  PW_HIP(hipMemcpy(ws->d_ylm, y.data(), y.size() * sizeof(double), hipMemcpyHostToDevice));
  return PW_OK;
}

static int upload_mass(pw_waveset *ws) {
  // WAVES_PSIMASS (src/paw_waves1.f90:6641-6666), no bucket
  pw_plan *p = ws->plan;
  std::vector<double> m(p->ngl);
  for (int g = 0; g < p->ngl; ++g) m[g] = ws->emass * (1.0 + ws->emasscg2 * p->g2[g]);
  PW_HIP(hipMemcpy(ws->d_mpsi, m.data(), m.size() * sizeof(double), hipMemcpyHostToDevice));
  ws->have_mass = true;
  return PW_OK;
}

static int phase_at(pw_waveset *ws, const std::vector<double> &r) {
  // the tables of these positions, unless they are already there (ETOT
  // recomputes them at least once per step)
  if (!ws->phase_valid || ws->phase_pos != r) {
    PW_TRY(k_phase_tables(ws, r.data(), ws->d_phase));
    ws->phase_pos = r;
    ws->phase_valid = true;
  }
  // the mixed form's products read the projector factors of these positions
  if (ws->prec == PW_PREC_MIXED && (!ws->pfac_valid || ws->pfac_pos != r)) {
    PW_TRY(k_pfac(ws));
    ws->pfac_pos = r;
    ws->pfac_valid = true;
  }
  return PW_OK;
}

// DEDPROJ (src/paw_waves1.f90:5592-5691) into d_xproj
static int dedproj(pw_waveset *ws) {
  const int nb = ws->nb;
  PW_TRY(k_atom_matrix(ws, ws->d_proj, ws->d_xproj, ws->d_dhmat));
  PW_TRY(k_fold_occ(ws, ws->d_xproj));
  // LAMBDA = -RLAM0 (f_i+f_j)/2 from the device-resident RLAM0
  k_lam_dedproj<<<(nb * nb + 255) / 256, 256, 0, ws->stream>>>(nb, ws->d_occ, ws->d_rlam0, d_mat(ws, 3));
  PW_HIP(hipGetLastError());
  PW_TRY(k_atom_matrix(ws, ws->d_proj, ws->d_oproj, ws->d_domat));
  PW_TRY(k_addoproj(ws, ws->d_xproj, ws->d_oproj, d_mat(ws, 3)));
  return PW_OK;
}

// read the device results of a step section back (one copy, one wait)
// the copy of a section's sums into h_res is in flight: mark it (what: 1
// ETOT, 2 ORTHOGONALIZE) behind an event for ws_finish
static int defer_results(pw_waveset *ws, int what) {
  if (!ws->res_ev) PW_HIP(hipEventCreateWithFlags(&ws->res_ev, hipEventDisableTiming));
  PW_HIP(hipEventRecord(ws->res_ev, ws->stream));
  ws->res_pending |= what;
  return PW_OK;
}

static double sum_of(const double *v, int n) {
  double s = 0.0;
  for (int i = 0; i < n; ++i) s += v[i];
  return s;
}

// ---------------------------------------------------------------------------
// Gram-Schmidt (WAVES_GRAMSCHMIDT, src/paw_waves2.f90:2025-2177)
// ---------------------------------------------------------------------------
static int gramschmidt_one(pw_waveset *ws, double2 *psi, const std::vector<double> &r) {
  const int nb = ws->nb;
  pw_plan *p = ws->plan;
  PW_TRY(phase_at(ws, r));
  PW_TRY(k_runs_scope(ws, true));
  PW_TRY(k_projections(ws, psi, ws->d_phase, ws->d_oproj));
  PW_TRY(k_overlap_full(ws, psi, psi, ws->d_oproj, ws->d_oproj, d_mat(ws, 4)));
  PW_TRY(k_runs_scope(ws, false));
  std::vector<double> s((size_t)nb * nb);
  PW_HIP(hipMemcpyAsync(s.data(), d_mat(ws, 4), s.size() * sizeof(double), hipMemcpyDeviceToHost, ws->stream));
  PW_HIP(hipStreamSynchronize(ws->stream));
  // Cholesky S = R^T R (R upper, positive diagonal); T = R^-1; X = T - 1.
  // WAVES_ORTHO_Y_C with PHI=CHI builds this unique triangular transform.
  std::vector<double> R((size_t)nb * nb, 0.0);
  for (int j = 0; j < nb; ++j) {
    for (int i = 0; i <= j; ++i) {
      double v = s[(size_t)i * nb + j];
      for (int k = 0; k < i; ++k) v -= R[(size_t)k * nb + i] * R[(size_t)k * nb + j];
      if (i == j) {
        if (!(v > 0.0)) return pw_fail(PW_ERR_NOCONV, "WAVES_GRAMSCHMIDT: overlap not positive definite");
        R[(size_t)i * nb + j] = std::sqrt(v);
      } else {
        R[(size_t)i * nb + j] = v / R[(size_t)i * nb + i];
      }
    }
  }
  std::vector<double> T((size_t)nb * nb, 0.0);
  for (int j = 0; j < nb; ++j) {          // column j of R^-1 (upper triangular)
    T[(size_t)j * nb + j] = 1.0 / R[(size_t)j * nb + j];
    for (int i = j - 1; i >= 0; --i) {
      double v = 0.0;
      for (int k = i + 1; k <= j; ++k) v += R[(size_t)i * nb + k] * T[(size_t)k * nb + j];
      T[(size_t)i * nb + j] = -v / R[(size_t)i * nb + i];
    }
  }
  for (int i = 0; i < nb; ++i) T[(size_t)i * nb + i] -= 1.0;
  PW_HIP(hipMemcpyAsync(d_mat(ws, 3), T.data(), T.size() * sizeof(double), hipMemcpyHostToDevice, ws->stream));
  PW_TRY(k_copy(ws, ws->d_tmp, psi, (size_t)ws->nbh * p->ngl));
  PW_TRY(k_addopsi(ws, psi, ws->d_tmp, d_mat(ws, 3)));
  PW_HIP(hipStreamSynchronize(ws->stream));
  return PW_OK;
}

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" int pw_waves_create(pw_plan *plan, int nb, int nat, int nsp, const int *isp,
                               const int *lnx, int lnxx, const int *lox, int prec, pw_waveset **out) {
  if (!plan || !isp || !lnx || !lox || !out) return pw_fail(PW_ERR_ARG, "pw_waves_create: null argument");
  // the reference's input layer admits only an even number of states
  // (NB=2*INT((NB+1)/2), src/paw_ioroutines.f90:4813), so the last super
  // band is never a lone real band
  if (nb <= 0 || nb % 2 != 0)
    return pw_fail(PW_ERR_UNSUPPORTED, "pw_waves_create: NB must be positive and even (the reference rounds NB up "
                                       "to even, src/paw_ioroutines.f90:4813)");
  if (nat <= 0 || nsp <= 0 || lnxx <= 0) return pw_fail(PW_ERR_ARG, "pw_waves_create: bad sizes");
  if (nb > 1024) return pw_fail(PW_ERR_UNSUPPORTED, "pw_waves_create: NB > 1024 is not supported");
  if (prec != PW_PREC_FP64 && prec != PW_PREC_MIXED) return pw_fail(PW_ERR_ARG, "pw_waves_create: bad precision");
  PW_TRY(plan_device(plan));
  {
    std::lock_guard<std::mutex> lk(capture_gate().mu);
    capture_gate().wavesets++;
  }
  pw_waveset *ws = new pw_waveset();
  ws->plan = plan;
  ws->prec = prec;
  ws->nb = nb; ws->nbh = (nb + 1) / 2; ws->nat = nat; ws->nsp = nsp; ws->lnxx = lnxx;
  ws->isp.resize(nat);
  for (int a = 0; a < nat; ++a) {
    if (isp[a] < 1 || isp[a] > nsp) {
      pw_waves_destroy(ws);
      return pw_fail(PW_ERR_ARG, "pw_waves_create: ISP out of range");
    }
    ws->isp[a] = isp[a] - 1;
  }
  ws->lnx.assign(lnx, lnx + nsp);
  ws->lox.assign(lox, lox + (size_t)lnxx * nsp);
  ws->lmnx.assign(nsp, 0);
  ws->ibpro.assign(nsp, 0);
  int lmax = 0;
  for (int s = 0; s < nsp; ++s) {
    if (lnx[s] < 0 || lnx[s] > lnxx) {
      pw_waves_destroy(ws);
      return pw_fail(PW_ERR_ARG, "pw_waves_create: LNX > LNXX");
    }
    ws->ibpro[s] = ws->nbarepro;
    ws->nbarepro += lnx[s];
    for (int ln = 0; ln < lnx[s]; ++ln) {
      const int l = lox[ln + lnxx * s];
      ws->lmnx[s] += 2 * l + 1;
      lmax = std::max(lmax, l);
    }
    ws->lmnxx = std::max(ws->lmnxx, ws->lmnx[s]);
  }
  ws->lmx = (lmax + 1) * (lmax + 1);
  ws->lmn_ln.assign((size_t)nsp * ws->lmnxx, 0);
  ws->lmn_lm.assign((size_t)nsp * ws->lmnxx, 0);
  ws->lmn_l.assign((size_t)nsp * ws->lmnxx, 0);
  for (int s = 0; s < nsp; ++s) {
    int lmn = 0;
    for (int ln = 0; ln < lnx[s]; ++ln) {
      const int l = lox[ln + lnxx * s];
      for (int m = 0; m < 2 * l + 1; ++m, ++lmn) {
        ws->lmn_ln[s * ws->lmnxx + lmn] = ln;
        ws->lmn_lm[s * ws->lmnxx + lmn] = l * l + m;
        ws->lmn_l[s * ws->lmnxx + lmn] = l;
      }
    }
  }
  ws->ipro0.assign(nat, 0);
  for (int a = 0; a < nat; ++a) {
    ws->ipro0[a] = ws->npro;
    ws->npro += ws->lmnx[ws->isp[a]];
  }
  ws->emasscg2 = 0.0;
  ws->rlam0.assign((size_t)nb * nb, 0.0);
  ws->force.assign(3 * nat, 0.0);
  ws->eig.assign(nb, 0.0);
  ws->denmat.assign((size_t)ws->lmnxx * ws->lmnxx * nat, 0.0);
  const int nmaxx = std::max(plan->nmax[0], std::max(plan->nmax[1], plan->nmax[2]));
  ws->nphase = 2 * nmaxx + 1;
  auto fail = [&](int rc) { pw_waves_destroy(ws); return rc; };
  const size_t psz = (size_t)ws->nbh * plan->ngl * sizeof(double2);
  const size_t jsz = (size_t)ws->npro * ws->nbh * sizeof(double2);
  int rc = PW_OK;
  double2 **cbufs[] = {&ws->d_psi0, &ws->d_psim, &ws->d_hpsi, &ws->d_opsi, &ws->d_tmp};
  for (auto b : cbufs) if (rc == PW_OK) rc = dalloc((void **)b, psz);
  double2 **pbufs[] = {&ws->d_proj, &ws->d_oproj, &ws->d_xproj};
  for (auto b : pbufs) if (rc == PW_OK) rc = dalloc((void **)b, jsz);
  if (rc == PW_OK) rc = dalloc((void **)&ws->d_pro, (size_t)ws->nbarepro * plan->ngl * sizeof(double));
  if (rc == PW_OK) rc = dalloc((void **)&ws->d_ylm, (size_t)ws->lmx * plan->ngl * sizeof(double));
  if (rc == PW_OK) rc = dalloc((void **)&ws->d_mpsi, plan->ngl * sizeof(double));
  if (rc == PW_OK) rc = dalloc((void **)&ws->d_v, plan->nrl * sizeof(double));
  if (rc == PW_OK) rc = dalloc((void **)&ws->d_rho, plan->nrl * sizeof(double));
  if (rc == PW_OK) rc = dalloc((void **)&ws->d_occ2, 2 * ws->nbh * sizeof(double));
  if (rc == PW_OK) rc = dalloc((void **)&ws->d_occ, ws->nb * sizeof(double));
  if (rc == PW_OK) rc = dalloc((void **)&ws->d_rlam0, (size_t)nb * nb * sizeof(double));
  if (rc == PW_OK) rc = dalloc((void **)&ws->d_rlamm, (size_t)nb * nb * sizeof(double));
  if (rc == PW_OK) rc = dalloc((void **)&ws->d_fpm, 2 * ws->nbh * sizeof(double));
  if (rc == PW_OK) rc = dalloc((void **)&ws->d_fab, 2 * ws->nbh * sizeof(double));
  if (rc == PW_OK) rc = dalloc((void **)&ws->d_domat, (size_t)nat * ws->lmnxx * ws->lmnxx * sizeof(double2));
  if (rc == PW_OK) rc = dalloc((void **)&ws->d_dhmat, (size_t)nat * ws->lmnxx * ws->lmnxx * sizeof(double2));
  if (rc == PW_OK) rc = dalloc((void **)&ws->d_phase, (size_t)nat * 3 * ws->nphase * sizeof(double2));
  if (rc == PW_OK) rc = dalloc((void **)&ws->d_small_mats, (size_t)6 * nb * nb * sizeof(double));
  if (rc == PW_OK) {
    // result layout of a step section (partial sums, see fetch_results)
    const size_t rb = red_blocks(ws), vb = vrho_blocks(ws);
    ws->r_ekin = 0;
    ws->r_vrho = ws->r_ekin + rb;
    ws->r_wk = ws->r_vrho + vb;
    ws->r_expect = ws->r_wk + rb;
    ws->r_denmat = ws->r_expect + 2 * (size_t)ws->nbh * expect_chunks(ws);
    ws->r_force = ws->r_denmat + (size_t)nat * ws->lmnxx * ws->lmnxx;
    ws->r_resid = ws->r_force + 3 * (size_t)ws->npro;
    ws->res_n = ws->r_resid + 1;
    rc = dalloc((void **)&ws->d_res, ws->res_n * sizeof(double));
    if (rc == PW_OK && hipHostMalloc((void **)&ws->h_res, (ws->res_n + (size_t)nb * nb) * sizeof(double)) != hipSuccess)
      rc = pw_fail(PW_ERR_HIP, "pw_waves_create: pinned result buffer");
  }
  std::vector<int> iv;
  auto up = [&](int **dst, const std::vector<int> &src) {
    if (rc != PW_OK) return;
    rc = dalloc((void **)dst, src.size() * sizeof(int));
    if (rc == PW_OK && !src.empty() &&
        hipMemcpy(*dst, src.data(), src.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess)
      rc = pw_fail(PW_ERR_HIP, "pw_waves_create: upload failed");
  };
  up(&ws->d_atom_isp, ws->isp);
  up(&ws->d_ipro0, ws->ipro0);
  up(&ws->d_ibpro, ws->ibpro);
  up(&ws->d_lmnx, ws->lmnx);
  up(&ws->d_lmn_ln, ws->lmn_ln);
  up(&ws->d_lmn_lm, ws->lmn_lm);
  up(&ws->d_lmn_l, ws->lmn_l);
  {
    std::vector<int> tab;
    for (int a = 0; a < nat; ++a)
      for (int l = 0; l < ws->lmnx[ws->isp[a]]; ++l) {
        tab.push_back(a);
        tab.push_back(l);
      }
    up((int **)&ws->d_ipro_tab, tab);
  }
  if (rc == PW_OK) rc = ylm_table(ws);
  if (rc != PW_OK) return fail(rc);
  if (hipStreamCreateWithFlags(&ws->stream, hipStreamNonBlocking) != hipSuccess) return fail(pw_fail(PW_ERR_HIP, "hipStreamCreate failed"));
  ws->own_stream = true;
  *out = ws;
  return PW_OK;
}

// A waveset as pw_waves_create left it, keeping its device buffers, stream,
// captured graph and tuned launch shapes: the request loop hands a finished
// set to the next request of the same shape.  The state of a run (settings,
// positions, multipliers, results, flags, cached factors and runs, the linked
// potential and one-centre set-up) comes from a default-constructed set, and
// the device arrays a new set starts from zero are cleared again: the
// multipliers RLAM0/RLAMM (the first constraint solve starts from RLAM0 and
// the first DEDPROJ reads it), the NB x NB matrices and the projections.
int ws_reset_run(pw_waveset *ws) {
  PW_TRY(ws_finish(ws));
  ws_flush_timers(ws);
  sphere_release(ws);
  const size_t n2 = (size_t)ws->nb * ws->nb, jsz = (size_t)ws->npro * ws->nbh * sizeof(double2);
  PW_HIP(hipMemsetAsync(ws->d_rlam0, 0, n2 * sizeof(double), ws->stream));
  PW_HIP(hipMemsetAsync(ws->d_rlamm, 0, n2 * sizeof(double), ws->stream));
  PW_HIP(hipMemsetAsync(ws->d_small_mats, 0, 6 * n2 * sizeof(double), ws->stream));
  PW_HIP(hipMemsetAsync(ws->d_proj, 0, jsz, ws->stream));
  PW_HIP(hipMemsetAsync(ws->d_oproj, 0, jsz, ws->stream));
  PW_HIP(hipMemsetAsync(ws->d_xproj, 0, jsz, ws->stream));
  PW_HIP(hipStreamSynchronize(ws->stream));
  pw_waveset f;
  // structure (pw_waves_create)
  f.plan = ws->plan;
  f.prec = ws->prec;
  f.nb = ws->nb; f.nbh = ws->nbh; f.nat = ws->nat; f.nsp = ws->nsp; f.lnxx = ws->lnxx;
  f.lmnxx = ws->lmnxx; f.lmx = ws->lmx; f.npro = ws->npro; f.nbarepro = ws->nbarepro;
  f.isp = std::move(ws->isp); f.lnx = std::move(ws->lnx); f.lmnx = std::move(ws->lmnx); f.lox = std::move(ws->lox);
  f.ipro0 = std::move(ws->ipro0); f.ibpro = std::move(ws->ibpro);
  f.lmn_ln = std::move(ws->lmn_ln); f.lmn_lm = std::move(ws->lmn_lm); f.lmn_l = std::move(ws->lmn_l);
  f.nphase = ws->nphase;
  f.rlam0.assign((size_t)f.nb * f.nb, 0.0);
  f.force.assign(3 * (size_t)f.nat, 0.0);
  f.eig.assign(f.nb, 0.0);
  f.denmat.assign((size_t)f.lmnxx * f.lmnxx * f.nat, 0.0);
  f.res_n = ws->res_n;
  f.r_ekin = ws->r_ekin; f.r_vrho = ws->r_vrho; f.r_wk = ws->r_wk; f.r_expect = ws->r_expect;
  f.r_denmat = ws->r_denmat; f.r_force = ws->r_force; f.r_resid = ws->r_resid;
  // device and pinned buffers (fixed and grown on demand)
  f.d_psi0 = ws->d_psi0; f.d_psim = ws->d_psim; f.d_hpsi = ws->d_hpsi; f.d_opsi = ws->d_opsi; f.d_tmp = ws->d_tmp;
  f.d_proj = ws->d_proj; f.d_oproj = ws->d_oproj; f.d_xproj = ws->d_xproj;
  f.d_pro = ws->d_pro; f.d_ylm = ws->d_ylm; f.d_mpsi = ws->d_mpsi; f.d_v = ws->d_v; f.d_rho = ws->d_rho;
  f.d_occ2 = ws->d_occ2; f.d_occ = ws->d_occ; f.d_rlam0 = ws->d_rlam0; f.d_rlamm = ws->d_rlamm;
  f.d_fpm = ws->d_fpm; f.d_fab = ws->d_fab; f.d_domat = ws->d_domat; f.d_dhmat = ws->d_dhmat;
  f.d_atom_isp = ws->d_atom_isp; f.d_ipro0 = ws->d_ipro0; f.d_ibpro = ws->d_ibpro; f.d_lmnx = ws->d_lmnx;
  f.d_lmn_ln = ws->d_lmn_ln; f.d_lmn_lm = ws->d_lmn_lm; f.d_lmn_l = ws->d_lmn_l; f.d_ipro_tab = ws->d_ipro_tab;
  f.d_phase = ws->d_phase; f.d_small_mats = ws->d_small_mats;
  f.d_res = ws->d_res; f.h_res = ws->h_res;
  f.d_pfac = ws->d_pfac; f.pfac_bytes = ws->pfac_bytes;
  f.d_part = ws->d_part; f.part_bytes = ws->part_bytes;
  f.d_split = ws->d_split; f.split_bytes = ws->split_bytes;
  f.d_wruns = ws->d_wruns; f.wruns_bytes = ws->wruns_bytes;
  f.d_small = ws->d_small; f.small_bytes = ws->small_bytes;
  f.d_ortho = ws->d_ortho; f.h_ortho_state = ws->h_ortho_state; f.ortho_cap = ws->ortho_cap;
  f.ortho_exec = ws->ortho_exec; f.ortho_key = ws->ortho_key;
  f.stream = ws->stream; f.own_stream = ws->own_stream; f.res_ev = ws->res_ev;
  // launch shapes
  f.tune_split_waves = ws->tune_split_waves; f.tune_col_waves = ws->tune_col_waves;
  f.tune_dens_blocks = ws->tune_dens_blocks;
  *ws = std::move(f);
  return PW_OK;
}

extern "C" int pw_waves_destroy(pw_waveset *ws) {
  if (!ws) return PW_OK;
  PW_TRY(ws_finish(ws));
  ws_flush_timers(ws);
  ortho_graph_release(ws);
  sphere_release(ws);
  void *ptrs[] = {ws->d_psi0, ws->d_psim, ws->d_hpsi, ws->d_opsi, ws->d_tmp, ws->d_proj, ws->d_oproj,
                  ws->d_xproj, ws->d_pro, ws->d_ylm, ws->d_mpsi, ws->d_v, ws->d_rho, ws->d_occ2,
                  ws->d_occ, ws->d_rlam0, ws->d_rlamm,
                  ws->d_fpm, ws->d_fab, ws->d_domat, ws->d_dhmat, ws->d_atom_isp, ws->d_ipro0,
                  ws->d_ibpro, ws->d_lmnx, ws->d_lmn_ln, ws->d_lmn_lm, ws->d_lmn_l, ws->d_phase,
                  ws->d_part, ws->d_small, ws->d_ortho, ws->d_small_mats, ws->d_ipro_tab, ws->d_pfac,
                  ws->d_split, ws->d_wruns};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  if (ws->d_res) (void)hipFree(ws->d_res);
  if (ws->h_res) (void)hipHostFree(ws->h_res);
  if (ws->h_ortho_state) (void)hipHostFree(ws->h_ortho_state);
  if (ws->res_ev) (void)hipEventDestroy(ws->res_ev);
  if (ws->own_stream && ws->stream) (void)hipStreamDestroy(ws->stream);
  delete ws;
  std::lock_guard<std::mutex> lk(capture_gate().mu);
  capture_gate().wavesets--;
  return PW_OK;
}

extern "C" int pw_waves_set_stream(pw_waveset *ws, void *stream) {
  if (!ws) return pw_fail(PW_ERR_ARG, "pw_waves_set_stream: null waveset");
  PW_TRY(ws_finish(ws));
  if (ws->own_stream && ws->stream) (void)hipStreamDestroy(ws->stream);
  ws->stream = (hipStream_t)stream;
  ws->own_stream = false;
  return PW_OK;
}

extern "C" int pw_waves_setr8(pw_waveset *ws, const char *id, double val) {
  if (!ws || !id) return pw_fail(PW_ERR_ARG, "pw_waves_setr8: null argument");
  PW_TRY(ws_finish(ws));
  // inputs are copied synchronously; work still queued on the waveset's
  // (non-blocking) stream must not see them change
  PW_HIP(hipStreamSynchronize(ws->stream));
  std::string s(id);
  if (s == "TIMESTEP") ws->delt = val;
  else if (s == "EMASS") ws->emass = val;
  else if (s == "EMASSCG2") ws->emasscg2 = val;
  else if (s == "FRICTION") ws->annee = val;
  else if (s == "RESIDUALCHECK") ws->residual_on = val != 0.0;
  else return pw_fail(PW_ERR_ARG, "WAVES$SETR8: ID not recognised: " + s);
  return upload_mass(ws);
}

extern "C" int pw_waves_setr8a(pw_waveset *ws, const char *id, int len, const double *val) {
  if (!ws || !id || !val) return pw_fail(PW_ERR_ARG, "pw_waves_setr8a: null argument");
  PW_TRY(ws_finish(ws));
  // inputs are copied synchronously; work still queued on the waveset's
  // (non-blocking) stream must not see them change
  PW_HIP(hipStreamSynchronize(ws->stream));
  std::string s(id);
  pw_plan *p = ws->plan;
  auto chk = [&](size_t n) {
    return (size_t)len == n ? PW_OK : pw_fail(PW_ERR_ARG, "WAVES$SETR8A: size inconsistent for " + s);
  };
  if (s == "PRO") {
    PW_TRY(chk((size_t)p->ngl * ws->nbarepro));
    // caller layout PRO(NGL,NBAREPRO) column-major == [ibare][ngl]
    PW_HIP(hipMemcpy(ws->d_pro, val, (size_t)len * sizeof(double), hipMemcpyHostToDevice));
    ws->pfac_valid = false;
  } else if (s == "OCC") {
    PW_TRY(chk(ws->nb));
    ws->occ.assign(val, val + ws->nb);
    PW_TRY(upload_occ(ws));
  } else if (s == "DO") {
    PW_TRY(chk((size_t)ws->lnxx * ws->lnxx * ws->nsp));
    ws->do_ln.assign(val, val + len);
    PW_TRY(upload_do(ws));
  } else if (s == "VOFR") {
    PW_TRY(chk(p->nrl));
    PW_HIP(hipMemcpy(ws->d_v, val, (size_t)len * sizeof(double), hipMemcpyHostToDevice));
    ws->have_v = true;
  } else if (s == "R0" || s == "RP" || s == "RM") {
    PW_TRY(chk((size_t)3 * ws->nat));
    std::vector<double> r(val, val + len);
    if (s == "R0") ws->r0 = r; else if (s == "RP") ws->rp = r; else ws->rm = r;
  } else if (s == "QLM") {
    // QLM(LMRXX,NAT): LMRXX = LEN/NAT must be a full shell (L+1)^2, L <= 4
    const int lmrxx = ws->nat > 0 ? len / ws->nat : 0;
    const int lx = (int)std::lround(std::sqrt((double)std::max(lmrxx, 1))) - 1;
    if (lmrxx < 1 || (size_t)lmrxx * ws->nat != (size_t)len || (lx + 1) * (lx + 1) != lmrxx || lx > 4)
      return pw_fail(PW_ERR_ARG, "WAVES$SETR8A: QLM must be QLM(LMRXX,NAT) with LMRXX = (L+1)^2, L <= 4");
    ws->qlm.assign(val, val + len);
    ws->qlm_lmrxx = lmrxx;
  } else if (s == "ORTHOBOUNDS") {
    // spectral bounds of sym(CHIPSI) for the next ORTHO_X call; ORTHO_X checks
    // them on the device and recomputes them when they are not accepted
    PW_TRY(chk(2));
    ws->lz_lo = val[0];
    ws->lz_hi = val[1];
    ws->lz_valid = true;
  } else {
    return pw_fail(PW_ERR_ARG, "WAVES$SETR8A: ID not recognised: " + s);
  }
  return PW_OK;
}

extern "C" int pw_waves_setc8a(pw_waveset *ws, const char *id, int len, const double *val) {
  if (!ws || !id || !val) return pw_fail(PW_ERR_ARG, "pw_waves_setc8a: null argument");
  PW_TRY(ws_finish(ws));
  // inputs are copied synchronously; work still queued on the waveset's
  // (non-blocking) stream must not see them change
  PW_HIP(hipStreamSynchronize(ws->stream));
  std::string s(id);
  pw_plan *p = ws->plan;
  const size_t npsi = (size_t)ws->nbh * p->ngl;
  if (s == "PSI0" || s == "PSIM") {
    if ((size_t)len != npsi) return pw_fail(PW_ERR_ARG, "WAVES$SETC8A: size inconsistent for " + s);
    PW_HIP(hipMemcpy(s == "PSI0" ? ws->d_psi0 : ws->d_psim, val, npsi * sizeof(double2), hipMemcpyHostToDevice));
    if (s == "PSI0") ws->have_psi0 = true; else ws->have_psim = true;
  } else if (s == "DH") {
    const size_t n = (size_t)ws->lmnxx * ws->lmnxx * ws->nat;
    if ((size_t)len != n) return pw_fail(PW_ERR_ARG, "WAVES$SETC8A: size inconsistent for DH");
    // DH(LMNXX,LMNXX,NAT) column-major -> per atom row-major [i][j]
    std::vector<double2> m(n);
    const int L = ws->lmnxx;
    for (int a = 0; a < ws->nat; ++a)
      for (int i = 0; i < L; ++i)
        for (int j = 0; j < L; ++j) {
          const size_t src = ((size_t)a * L + j) * L + i;   // DH(i,j,a)
          m[((size_t)a * L + i) * L + j] = make_double2(val[2 * src], val[2 * src + 1]);
        }
    PW_HIP(hipMemcpy(ws->d_dhmat, m.data(), n * sizeof(double2), hipMemcpyHostToDevice));
    ws->dh.assign(val, val + 2 * n);
    ws->have_dh = true;
  } else if (s == "RLAM0" || s == "RLAMM") {
    const size_t n = (size_t)ws->nb * ws->nb;
    if ((size_t)len != n) return pw_fail(PW_ERR_ARG, "WAVES$SETC8A: size inconsistent for " + s);
    std::vector<double> h(n, 0.0);
    for (int i = 0; i < ws->nb; ++i)
      for (int j = 0; j < ws->nb; ++j) h[(size_t)i * ws->nb + j] = val[2 * ((size_t)j * ws->nb + i)];
    PW_HIP(hipMemcpy(s == "RLAM0" ? ws->d_rlam0 : ws->d_rlamm, h.data(), n * sizeof(double),
                     hipMemcpyHostToDevice));
    if (s == "RLAMM") ws->have_rlamm = true; else ws->have_rlam0 = true;
  } else {
    return pw_fail(PW_ERR_ARG, "WAVES$SETC8A: ID not recognised: " + s);
  }
  return PW_OK;
}

static int get_dev_c(pw_waveset *ws, const double2 *src, size_t n, int len, double *val) {
  if ((size_t)len != n) return pw_fail(PW_ERR_ARG, "WAVES$GETC8A: size inconsistent");
  PW_HIP(hipStreamSynchronize(ws->stream));
  PW_HIP(hipMemcpy(val, src, n * sizeof(double2), hipMemcpyDeviceToHost));
  return PW_OK;
}

extern "C" int pw_waves_getc8a(pw_waveset *ws, const char *id, int len, double *val) {
  if (!ws || !id || !val) return pw_fail(PW_ERR_ARG, "pw_waves_getc8a: null argument");
  PW_TRY(ws_finish(ws));
  std::string s(id);
  const size_t npsi = (size_t)ws->nbh * ws->plan->ngl;
  const size_t nproj = (size_t)ws->nbh * ws->npro;
  if (s == "PSI0") return get_dev_c(ws, ws->d_psi0, npsi, len, val);
  if (s == "DH") {
    // DH(LMNXX,LMNXX,NAT) of the last ETOT (or as set), column-major
    const int L = ws->lmnxx;
    const size_t n = (size_t)L * L * ws->nat;
    if ((size_t)len != n) return pw_fail(PW_ERR_ARG, "WAVES$GETC8A: size inconsistent");
    std::vector<double2> m(n);
    PW_HIP(hipStreamSynchronize(ws->stream));
    PW_HIP(hipMemcpy(m.data(), ws->d_dhmat, n * sizeof(double2), hipMemcpyDeviceToHost));
    for (int a = 0; a < ws->nat; ++a)
      for (int i = 0; i < L; ++i)
        for (int j = 0; j < L; ++j) {
          const size_t dst = ((size_t)a * L + j) * L + i;
          val[2 * dst] = m[((size_t)a * L + i) * L + j].x;
          val[2 * dst + 1] = m[((size_t)a * L + i) * L + j].y;
        }
    return PW_OK;
  }
  if (s == "PSIM") return get_dev_c(ws, ws->d_psim, npsi, len, val);
  if (s == "HPSI") return get_dev_c(ws, ws->d_hpsi, npsi, len, val);
  if (s == "OPSI") return get_dev_c(ws, ws->d_opsi, npsi, len, val);
  if (s == "PROJ") return get_dev_c(ws, ws->d_proj, nproj, len, val);
  if (s == "OPROJ") return get_dev_c(ws, ws->d_oproj, nproj, len, val);
  if (s == "RLAM0" || s == "RLAMM") {
    const size_t n = (size_t)ws->nb * ws->nb;
    if ((size_t)len != n) return pw_fail(PW_ERR_ARG, "WAVES$GETC8A: size inconsistent");
    const bool zero = s == "RLAM0" ? !ws->have_rlam0 : !ws->have_rlamm;
    if (s == "RLAMM" && zero) return pw_fail(PW_ERR_STATE, "WAVES$GETC8A: RLAMM not set (NLAMBDA < 2)");
    std::vector<double> &h = s == "RLAM0" ? ws->rlam0 : ws->rlamm;
    h.resize(n);
    PW_HIP(hipStreamSynchronize(ws->stream));
    PW_HIP(hipMemcpy(h.data(), s == "RLAM0" ? ws->d_rlam0 : ws->d_rlamm, n * sizeof(double),
                     hipMemcpyDeviceToHost));
    for (int i = 0; i < ws->nb; ++i)
      for (int j = 0; j < ws->nb; ++j) {
        val[2 * ((size_t)j * ws->nb + i)] = h[(size_t)i * ws->nb + j];
        val[2 * ((size_t)j * ws->nb + i) + 1] = 0.0;
      }
    return PW_OK;
  }
  return pw_fail(PW_ERR_ARG, "WAVES$GETC8A: ID not recognised: " + s);
}

static void to_colmajor(const std::vector<double> &m, int n, double *out) {
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) out[(size_t)j * n + i] = m[(size_t)i * n + j];
}

extern "C" int pw_waves_getr8a(pw_waveset *ws, const char *id, int len, double *val) {
  if (!ws || !id || !val) return pw_fail(PW_ERR_ARG, "pw_waves_getr8a: null argument");
  PW_TRY(ws_finish(ws));
  std::string s(id);
  auto chk = [&](size_t n) {
    return (size_t)len == n ? PW_OK : pw_fail(PW_ERR_ARG, "WAVES$GETR8A: size inconsistent for " + s);
  };
  if (s == "RHO") {
    PW_TRY(chk(ws->plan->nrl));
    PW_HIP(hipStreamSynchronize(ws->stream));
    PW_HIP(hipMemcpy(val, ws->d_rho, (size_t)len * sizeof(double), hipMemcpyDeviceToHost));
  } else if (s == "FORCE") {
    PW_TRY(chk((size_t)3 * ws->nat));
    std::memcpy(val, ws->force.data(), (size_t)len * sizeof(double));
  } else if (s == "EIG") {
    PW_TRY(chk(ws->nb));
    std::memcpy(val, ws->eig.data(), (size_t)len * sizeof(double));
  } else if (s == "MAT" || s == "OMAT" || s == "OOMAT") {
    PW_TRY(chk((size_t)ws->nb * ws->nb));
    if (!ws->have_mats) return pw_fail(PW_ERR_STATE, "WAVES: no orthogonalisation yet");
    std::vector<double> &m = s == "MAT" ? ws->mat : (s == "OMAT" ? ws->omat : ws->oomat);
    m.resize((size_t)len);
    PW_HIP(hipStreamSynchronize(ws->stream));
    PW_HIP(hipMemcpy(m.data(), d_mat(ws, s == "MAT" ? 0 : (s == "OMAT" ? 1 : 2)), (size_t)len * sizeof(double),
                     hipMemcpyDeviceToHost));
    to_colmajor(m, ws->nb, val);
  } else if (s == "VQLM" || s == "QLM") {
    if (s == "QLM" && ws->dplan && sphere_ready(ws)) {
      // the moments WAVES$MOMENTS made on the device in the last ETOT
      ws->qlm_lmrxx = sphere_lmrxx(ws);
      ws->qlm.resize((size_t)ws->nat * ws->qlm_lmrxx);
      PW_HIP(hipStreamSynchronize(ws->stream));
      PW_HIP(hipMemcpy(ws->qlm.data(), sphere_qlm_dev(ws), ws->qlm.size() * sizeof(double), hipMemcpyDeviceToHost));
    }
    // VQLM(LMRXX,NAT) = dE/dQLM of the last ETOT, and the moments it used
    if (!ws->dplan) return pw_fail(PW_ERR_STATE, "WAVES$GETR8A: " + s + " needs the density -> potential step");
    const std::vector<double> &src = s == "VQLM" ? ws->vqlm : ws->qlm;
    PW_TRY(chk(src.size()));
    std::memcpy(val, src.data(), (size_t)len * sizeof(double));
  } else if (s == "DENMAT") {
    PW_TRY(chk(ws->denmat.size()));
    // row-major per atom -> DENMAT(LMNXX,LMNXX,NAT) column-major
    const int L = ws->lmnxx;
    for (int a = 0; a < ws->nat; ++a)
      for (int i = 0; i < L; ++i)
        for (int j = 0; j < L; ++j) val[((size_t)a * L + j) * L + i] = ws->denmat[((size_t)a * L + i) * L + j];
  } else {
    return pw_fail(PW_ERR_ARG, "WAVES$GETR8A: ID not recognised: " + s);
  }
  return PW_OK;
}

// The orthonormality test of WAVES$ORTHOGONALIZE's TTEST branch
// (src/paw_waves2.f90:425-464): <psi|S|psi> of PSI0 at R(0) from the
// projections, WAVES_1COVERLAP and WAVES_OVERLAP, returned as
// max_ij |<psi_i|S|psi_j> - delta_ij|.  Always evaluated with the FP64
// kernels, also for a mixed-form set, so that it measures the mixed form's
// wave functions in the FP64 metric.  Uses OPROJ and the GS matrix as
// scratch (both are recomputed before their next use in the step).
static int ortho_error(pw_waveset *ws, double *out) {
  PW_TRY(need(ws->have_psi0, "PSI0"));
  PW_TRY(need(!ws->r0.empty(), "R0"));
  PW_TRY(need(!ws->do_ln.empty(), "DO"));
  struct Restore {
    pw_waveset *w;
    int prec;
    ~Restore() { w->prec = prec; }
  } restore{ws, ws->prec};
  ws->prec = PW_PREC_FP64;
  const int nb = ws->nb;
  PW_TRY(phase_at(ws, ws->r0));
  PW_TRY(k_projections(ws, ws->d_psi0, ws->d_phase, ws->d_oproj));
  PW_TRY(k_overlap_full(ws, ws->d_psi0, ws->d_psi0, ws->d_oproj, ws->d_oproj, d_mat(ws, 4)));
  std::vector<double> s((size_t)nb * nb);
  PW_HIP(hipMemcpyAsync(s.data(), d_mat(ws, 4), s.size() * sizeof(double), hipMemcpyDeviceToHost, ws->stream));
  PW_HIP(hipStreamSynchronize(ws->stream));
  double m = 0.0;
  for (int i = 0; i < nb; ++i)
    for (int j = 0; j < nb; ++j) m = std::max(m, std::fabs(s[(size_t)i * nb + j] - (i == j ? 1.0 : 0.0)));
  // the phase tables are those of R(0) again; the mixed form's projector
  // factors were not touched
  *out = m;
  return PW_OK;
}

extern "C" int pw_waves_getr8(pw_waveset *ws, const char *id, double *val) {
  if (!ws || !id || !val) return pw_fail(PW_ERR_ARG, "pw_waves_getr8: null argument");
  PW_TRY(ws_finish(ws));
  std::string s(id);
  if (s == "ORTHOERR") return ortho_error(ws, val);
  if (s == "EKIN") *val = ws->ekin;
  else if (s == "EPOT") *val = ws->epot;
  else if (s == "E1C") *val = ws->e1c;
  else if (s == "ETOT") *val = ws->etot;
  else if (s == "WAVEEKIN1") *val = ws->wavekin1;
  else if (s == "WAVEEKIN2") *val = ws->wavekin2;
  else if (s == "EFICT") *val = ws->efict;
  else if (s == "RESIDUAL") *val = ws->residual;
  else if (s == "EHARTREE") *val = ws->ehartree;
  else if (s == "EXC") *val = ws->exc;
  else if (s == "RHOB") *val = ws->rhob;
  else if (s == "POTB") *val = ws->potb;
  else return pw_fail(PW_ERR_ARG, "WAVES$GETR8: ID not recognised: " + s);
  return PW_OK;
}

extern "C" int pw_waves_geti4(pw_waveset *ws, const char *id, int *val) {
  if (!ws || !id || !val) return pw_fail(PW_ERR_ARG, "pw_waves_geti4: null argument");
  PW_TRY(ws_finish(ws));
  std::string s(id);
  if (s == "ORTHOITER") *val = ws->ortho_iter;
  else if (s == "NPRO") *val = ws->npro;
  else if (s == "NBH") *val = ws->nbh;
  else if (s == "LMNXX") *val = ws->lmnxx;
  else if (s == "ORTHOLEVELS") *val = ws->ortho_levels;
  else if (s == "ORTHONS") *val = ws->ortho_ns;
  else if (s == "ORTHOSERIES") *val = ws->ortho_series;
  else if (s == "SPLITWAVES") *val = ws->tune_split_waves;
  else if (s == "COLWAVES") *val = ws->tune_col_waves;
  else if (s == "DENSBLOCKS") *val = ws->tune_dens_blocks;
  else if (s == "ORTHOLANCZOS") *val = ws->ortho_lanczos;
  else if (s == "ORTHOEIGEN") *val = ws->ortho_eigen;
  else if (s == "ORTHOGRAPHMISSES") *val = ws->ortho_graph_misses;
  // depth of the multiplier history as WAVES_WRITELAMBDA counts it
  // (src/paw_waves2.f90:3948-3954): 0 none, 1 RLAM0, 2 RLAM0 and RLAMM
  else if (s == "NLAMBDA") *val = ws->have_rlamm ? 2 : (ws->have_rlam0 ? 1 : 0);
  else return pw_fail(PW_ERR_ARG, "WAVES$GETI4: ID not recognised: " + s);
  return PW_OK;
}

extern "C" int pw_waves_devptr(pw_waveset *ws, const char *id, void **ptr) {
  if (!ws || !id || !ptr) return pw_fail(PW_ERR_ARG, "pw_waves_devptr: null argument");
  PW_TRY(ws_finish(ws));
  std::string s(id);
  if (s == "PSI0") *ptr = ws->d_psi0;
  else if (s == "PSIM") *ptr = ws->d_psim;
  else if (s == "HPSI") *ptr = ws->d_hpsi;
  else if (s == "PROJ") *ptr = ws->d_proj;
  else if (s == "VOFR") *ptr = ws->d_v;
  else if (s == "RHO") *ptr = ws->d_rho;
  else return pw_fail(PW_ERR_ARG, "pw_waves_devptr: ID not recognised: " + s);
  return PW_OK;
}

extern "C" int pw_waves_profile(pw_waveset *ws, int enable) {
  if (!ws) return pw_fail(PW_ERR_ARG, "pw_waves_profile: null waveset");
  PW_TRY(ws_finish(ws));
  ws_flush_timers(ws);
  ws->profile = enable != 0;
  ws->ktime.clear();
  return PW_OK;
}

extern "C" int pw_waves_kernel_time(pw_waveset *ws, const char *name, double *ms, int *count) {
  if (!ws || !name || !ms || !count) return pw_fail(PW_ERR_ARG, "pw_waves_kernel_time: null argument");
  PW_TRY(ws_finish(ws));
  ws_flush_timers(ws);
  auto it = ws->ktime.find(name);
  *ms = it == ws->ktime.end() ? 0.0 : it->second.first;
  *count = it == ws->ktime.end() ? 0 : it->second.second;
  return PW_OK;
}

static int ready(pw_waveset *ws) {
  PW_TRY(need(ws->have_psi0, "PSI0"));
  PW_TRY(need(ws->have_psim, "PSIM"));
  PW_TRY(need(!ws->occ.empty(), "OCC"));
  PW_TRY(need(!ws->do_ln.empty(), "DO"));
  // DH is an input unless the one-centre terms produce it
  PW_TRY(need(ws->have_dh || (ws->dplan && sphere_ready(ws)), "DH"));
  PW_TRY(need(ws->have_v, "VOFR"));
  PW_TRY(need(!ws->r0.empty(), "R0"));
  PW_TRY(need(ws->have_mass, "TIMESTEP/EMASS"));
  return PW_OK;
}

extern "C" int pw_waves_gramschmidt(pw_waveset *ws) {
  if (!ws) return pw_fail(PW_ERR_ARG, "pw_waves_gramschmidt: null waveset");
  PW_TRY(ws_finish(ws));
  PW_TRY(ready(ws));
  PW_TRY(gramschmidt_one(ws, ws->d_psi0, ws->r0));
  PW_TRY(gramschmidt_one(ws, ws->d_psim, ws->rm.empty() ? ws->r0 : ws->rm));
  return PW_OK;
}

extern "C" int pw_waves_projections(pw_waveset *ws) {
  if (!ws) return pw_fail(PW_ERR_ARG, "pw_waves_projections: null waveset");
  PW_TRY(ws_finish(ws));
  PW_TRY(need(ws->have_psi0, "PSI0"));
  PW_TRY(need(!ws->r0.empty(), "R0"));
  PW_TRY(phase_at(ws, ws->r0));
  PW_TRY(k_projections(ws, ws->d_psi0, ws->d_phase, ws->d_proj));
  PW_HIP(hipStreamSynchronize(ws->stream));
  return PW_OK;
}

// ETOT's results from its sums in h_res (energies, DENMAT, forces, band
// energies)
static int etot_host(pw_waveset *ws) {
  const pw_plan *p = ws->plan;
  const double *hr = ws->h_res;
  ws->ekin = 0.5 * sum_of(hr + ws->r_ekin, red_blocks(ws)) * p->vol;
  if (ws->dplan) {
    double en[3];
    ws->vqlm.resize((size_t)ws->nat * ws->dplan->pot_lmrxx);
    vofrho_results(ws->dplan, ws->nat, en, ws->vqlm.data(), ws->fpot.data());
    ws->ehartree = en[0];
    ws->exc = en[1];
    ws->rhob = en[2];
    ws->epot = en[0] + en[1];
  } else {
    ws->epot = sum_of(hr + ws->r_vrho, vrho_blocks(ws)) * p->vol / (double)p->nrl;
  }
  ws->wavekin1 = sum_of(hr + ws->r_wk, red_blocks(ws)) / (ws->delt * ws->delt) * p->vol;
  // density matrix (hermitised, real) and the one-centre energy sum_ij DH_ij D_ji
  const int L = ws->lmnxx;
  const double *dm = hr + ws->r_denmat;
  for (int a = 0; a < ws->nat; ++a)
    for (int i = 0; i < L; ++i)
      for (int j = 0; j < L; ++j) {
        const size_t ij = ((size_t)a * L + i) * L + j, ji = ((size_t)a * L + j) * L + i;
        ws->denmat[ij] = 0.5 * (dm[ij] + dm[ji]);
      }
  if (ws->dplan && sphere_ready(ws)) {
    // AUGMENTATION$SPHERE's one-centre energy and POTB, from the device
    sphere_results(ws, &ws->e1c, &ws->potb);
  } else {
    double e1 = 0.0;
    for (int a = 0; a < ws->nat; ++a)
      for (int i = 0; i < L; ++i)
        for (int j = 0; j < L; ++j) {
          const size_t src = ((size_t)a * L + j) * L + i;   // DH(i,j,a) column-major
          e1 += ws->dh[2 * src] * ws->denmat[((size_t)a * L + j) * L + i];
        }
    ws->e1c = e1;
  }
  ws->etot = ws->ekin + ws->epot + ws->e1c;
  // forces: sum the projector contributions of each atom (fixed order)
  const double *fp = hr + ws->r_force;
  for (int a = 0; a < ws->nat; ++a) {
    const int i0 = ws->ipro0[a], nl = ws->lmnx[ws->isp[a]];
    for (int d = 0; d < 3; ++d) {
      double f = 0.0;
      for (int l = 0; l < nl; ++l) f += fp[(size_t)(i0 + l) * 3 + d];
      if (ws->dplan) f += ws->fpot[3 * a + d];
      ws->force[3 * a + d] = f;
    }
  }
  // band energies from the super-band expectation values
  const double *ex = hr + ws->r_expect;
  const int nc = expect_chunks(ws);
  for (int b = 0; b < ws->nbh; ++b) {
    double t = 0.0, tm = 0.0;
    for (int c = 0; c < nc; ++c) {
      t += ex[2 * ((size_t)c * ws->nbh + b)];
      tm += ex[2 * ((size_t)c * ws->nbh + b) + 1];
    }
    const double re = 0.5 * t * p->vol, rem = 0.5 * tm * p->vol;
    ws->eig[2 * b] = re + rem;
    if (2 * b + 1 < ws->nb) ws->eig[2 * b + 1] = re - rem;
  }
  return PW_OK;
}

// ORTHOGONALIZE's results from its sums in h_res
static void ortho_host(pw_waveset *ws) {
  const double wk2 = sum_of(ws->h_res + ws->r_wk, red_blocks(ws)) / (ws->delt * ws->delt) * ws->plan->vol;
  if (ws->residual_on) ws->residual = ws->h_res[ws->r_resid];
  ws->efict = 0.5 * (ws->pend_wk1 + wk2);
  if (!ws->pend_switched) ws->wavekin2 = wk2;
}

int ws_finish(pw_waveset *ws) {
  if (!ws || !ws->res_pending) return PW_OK;
  const int what = ws->res_pending;
  ws->res_pending = 0;
  PW_HIP(hipEventSynchronize(ws->res_ev));
  if (what & 1) PW_TRY(etot_host(ws));
  if (what & 2) ortho_host(ws);
  return PW_OK;
}

extern "C" int pw_waves_etot(pw_waveset *ws) {
  if (!ws) return pw_fail(PW_ERR_ARG, "pw_waves_etot: null waveset");
  PW_TRY(ready(ws));
  pw_plan *p = ws->plan;
  // the phase tables and projector factors are redone at least once per step
  ws->phase_valid = false;
  ws->pfac_valid = false;
  ScopeTimer whole(ws, "WAVES$ETOT");
  if (ws->first) {
    PW_TRY(pw_waves_gramschmidt(ws));
    PW_TRY(pw_waves_projections(ws));
    ws->first = false;
  }
  // WAVES$DENMAT
  PW_TRY(k_denmat(ws, ws->d_proj, ws->d_res + ws->r_denmat));
  // WAVES_DENSITY
  {
    ScopeTimer tm(ws, "density");
    PW_TRY(fft_density(p, ws->nbh, ws->d_psi0, ws->d_occ2, ws->d_rho, ws->prec, ws->stream, ws->tune_dens_blocks));
  }
  if (ws->dplan) {
    // POTENTIAL$VOFRHO: the local potential of this density, used by VPSI
    // below; its energy EHARTREE + EXC and forces are read with the results
    if (sphere_ready(ws)) {
      if (ws->dplan->pot_xc != 1)
        return pw_fail(PW_ERR_UNSUPPORTED, "WAVES$ETOT: the one-centre terms take the local functional (type 1) "
                                           "only; gradient corrections inside the spheres are not implemented");
      // the one-centre terms linked in (src/paw_waves1.f90:1843-1915):
      // WAVES$MOMENTS -> QLM, POTENTIAL$VOFRHO(rho, QLM) -> VQLM and RHOB,
      // WAVES$SPHERE -> DH, POTB, E1C, then WAVES$ADDCONSTANTPOT
      const int lmrxx = sphere_lmrxx(ws);
      for (int s = 0; s < ws->nsp; ++s)
        if (ws->dplan->pot_lmrx[s] != sphere_lmrx(ws, s))
          return pw_fail(PW_ERR_STATE, "WAVES$ETOT: LMRX of the density object's tables differs from the "
                                       "one-centre set-up");
      const double *qlm_dev = nullptr;
      PW_TRY(sphere_moments(ws, &qlm_dev));
      {
        ScopeTimer tm(ws, "vofrho");
        PW_TRY(vofrho_enqueue(ws->dplan, ws->nat, lmrxx, ws->isp1.data(), ws->r0.data(), nullptr, qlm_dev,
                              ws->d_rho, ws->d_v, ws->stream));
      }
      const double *res = ws->dplan->d_pot_res;
      PW_TRY(sphere_terms(ws, res, res + (size_t)(lmrxx + 3) * ws->nat + 2));
    } else {
      if (ws->qlm.size() != (size_t)ws->nat * ws->qlm_lmrxx || ws->qlm.empty())
        return pw_fail(PW_ERR_STATE, "WAVES$ETOT: QLM not set");
      ScopeTimer tm(ws, "vofrho");
      PW_TRY(vofrho_enqueue(ws->dplan, ws->nat, ws->qlm_lmrxx, ws->isp1.data(), ws->r0.data(), ws->qlm.data(),
                            nullptr, ws->d_rho, ws->d_v, ws->stream));
    }
  } else {
    // energy of the fixed local potential: RWEIGHT sum_r V rho
    PW_TRY(k_dot_vrho(ws, ws->d_res + ws->r_vrho));
  }
  // WAVES$FORCE
  PW_TRY(phase_at(ws, ws->r0));
  {
    ScopeTimer sec(ws, "W:FORCE");
    PW_TRY(dedproj(ws));
    PW_TRY(k_force(ws, ws->d_psi0, ws->d_phase, ws->d_xproj, ws->d_res + ws->r_force));
  }
  // WAVES$HPSI: VPSI + HPROJ + ADDPRO
  {
    ScopeTimer sec(ws, "W:HPSI");
    {
      ScopeTimer tm(ws, "vpsi");
      PW_TRY(fft_vpsi(p, ws->nbh, ws->d_psi0, ws->d_v, ws->d_hpsi, 1, ws->prec, ws->stream));
    }
    PW_TRY(k_atom_matrix(ws, ws->d_proj, ws->d_xproj, ws->d_dhmat));
    PW_TRY(k_addpro(ws, ws->d_hpsi, ws->d_phase, ws->d_xproj));
  }
  // WAVES_EKIN, <psi|H|psi>, WAVES_WAVEKINETIC (first half): one pass over psi0
  PW_TRY(k_etot_sums(ws, ws->d_res + ws->r_ekin, ws->d_res + ws->r_wk, ws->d_res + ws->r_expect));
  // the sums are copied without a wait (ws_finish completes them) unless the
  // density -> potential step needs its results now
  PW_TRY(ws_finish(ws));
  PW_HIP(hipMemcpyAsync(ws->h_res, ws->d_res, ws->res_n * sizeof(double), hipMemcpyDeviceToHost, ws->stream));
  if (ws->dplan) {
    PW_HIP(hipStreamSynchronize(ws->stream));
    PW_TRY(etot_host(ws));
  } else {
    PW_TRY(defer_results(ws, 1));
  }
  ws->have_hpsi = true;
  return PW_OK;
}

extern "C" int pw_waves_set_potential(pw_waveset *ws, pw_plan *dplan) {
  if (!ws) return pw_fail(PW_ERR_ARG, "pw_waves_set_potential: null waveset");
  PW_TRY(ws_finish(ws));
  if (dplan) {
    const pw_plan *p = ws->plan;
    if (dplan->nr1 != p->nr1 || dplan->nr2 != p->nr2 || dplan->nr3 != p->nr3)
      return pw_fail(PW_ERR_ARG, "pw_waves_set_potential: the density object has another r-space grid");
    for (int k = 0; k < 9; ++k)
      if (std::fabs(dplan->rbas[k] - p->rbas[k]) > 1e-12 * (1.0 + std::fabs(p->rbas[k])))
        return pw_fail(PW_ERR_ARG, "pw_waves_set_potential: the density object has another cell");
    if (!dplan->d_pot_tabs || dplan->pot_nsp < ws->nsp)
      return pw_fail(PW_ERR_STATE, "pw_waves_set_potential: setup tables of the density object not set");
  }
  PW_HIP(hipStreamSynchronize(ws->stream));
  ws->dplan = dplan;
  ws->isp1.resize(ws->nat);
  for (int a = 0; a < ws->nat; ++a) ws->isp1[a] = ws->isp[a] + 1;
  ws->vqlm.assign((size_t)ws->nat * std::max(ws->qlm_lmrxx, 1), 0.0);
  ws->fpot.assign(3 * (size_t)ws->nat, 0.0);
  if (dplan) ws->have_v = true;   // VOFR is produced by ETOT
  return PW_OK;
}

extern "C" int pw_waves_propagate(pw_waveset *ws) {
  if (!ws) return pw_fail(PW_ERR_ARG, "pw_waves_propagate: null waveset");
  if (!ws->have_hpsi)
    return pw_fail(PW_ERR_STATE, "WAVES$PROPAGATE: HPSI not available (call WAVES$ETOT first)");
  {
    ScopeTimer sec(ws, "WAVES$PROPAGATE");
    PW_TRY(k_propagate(ws));
  }
  ws->have_hpsi = false;
  return PW_OK;
}

extern "C" int pw_waves_orthogonalize(pw_waveset *ws) {
  if (!ws) return pw_fail(PW_ERR_ARG, "pw_waves_orthogonalize: null waveset");
  PW_TRY(ready(ws));
  pw_plan *p = ws->plan;
  const int nb = ws->nb;
  const std::vector<double> &rp = ws->rp.empty() ? ws->r0 : ws->rp;
  ScopeTimer whole(ws, "WAVES$ORTHOGONALIZE");
  // OPSI = dt^2/((1+a) m(G)) (psi0 + |p>DO<p|psi0>) at R(0)
  {
    ScopeTimer sec(ws, "WAVES_OPSI");
    PW_TRY(phase_at(ws, ws->r0));
    PW_TRY(k_atom_matrix(ws, ws->d_proj, ws->d_xproj, ws->d_domat));
    PW_TRY(k_addpro_opsi(ws, ws->d_opsi, ws->d_psi0, ws->d_phase, ws->d_xproj));
  }
  // projections at R(+); psibar and chi stay fixed until the overlaps are done
  {
    ScopeTimer sec(ws, "WAVES_OVERLAP");
    PW_TRY(phase_at(ws, rp));
    PW_TRY(k_runs_scope(ws, true));
    PW_TRY(k_projections(ws, ws->d_psim, ws->d_phase, ws->d_proj));
    PW_TRY(k_projections(ws, ws->d_opsi, ws->d_phase, ws->d_oproj));
    // MAT, OMAT, OOMAT
    PW_TRY(k_overlap_full(ws, ws->d_psim, ws->d_psim, ws->d_proj, ws->d_proj, d_mat(ws, 0)));
    PW_TRY(k_overlap_full(ws, ws->d_opsi, ws->d_psim, ws->d_oproj, ws->d_proj, d_mat(ws, 1)));
    PW_TRY(k_overlap_full(ws, ws->d_opsi, ws->d_opsi, ws->d_oproj, ws->d_oproj, d_mat(ws, 2)));
    PW_TRY(k_runs_scope(ws, false));
  }
  // initial LAMBDA: hermitised RLAM0 times 0.5 f_i/(f_i+f_j) (src/paw_waves2.f90:295-359)
  k_lam_guess<<<(nb * nb + 255) / 256, 256, 0, ws->stream>>>(nb, ws->d_occ, ws->d_rlam0, d_mat(ws, 3));
  PW_HIP(hipGetLastError());
  PW_TRY(ortho_x_device(ws, d_mat(ws, 3), d_mat(ws, 2), d_mat(ws, 0), d_mat(ws, 1), &ws->ortho_iter));
  // psi(+) = psibar + chi Lambda ; proj(+) = proj + oproj Lambda
  {
    ScopeTimer sec(ws, "WAVES_ADDOPSI");
    if (ws->residual_on) PW_TRY(k_copy(ws, ws->d_tmp, ws->d_psim, (size_t)ws->nbh * p->ngl));   // psibar
    PW_TRY(k_addopsi(ws, ws->d_psim, ws->d_opsi, d_mat(ws, 3)));
    PW_TRY(k_addoproj(ws, ws->d_proj, ws->d_oproj, d_mat(ws, 3)));
  }
  PW_TRY(ws_finish(ws));   // ETOT's sums: their copy ended before ORTHO_X's readback
  if (ws->residual_on) {
    PW_TRY(k_residual(ws, ws->d_tmp, ws->d_res + ws->r_resid));
    PW_HIP(hipMemcpyAsync(ws->h_res + ws->r_resid, ws->d_res + ws->r_resid, sizeof(double), hipMemcpyDeviceToHost,
                          ws->stream));
  }
  // RLAM0 stays on the device; MAT, OMAT, OOMAT go to the host only when asked for
  k_copy_mat<<<(nb * nb + 255) / 256, 256, 0, ws->stream>>>(nb * nb, d_mat(ws, 3), ws->d_rlam0);
  PW_HIP(hipGetLastError());
  ws->have_mats = true;
  ws->have_rlam0 = true;
  PW_TRY(k_wavekinetic(ws, ws->d_psi0, ws->d_psim, ws->d_fab, ws->d_res + ws->r_wk));
  PW_HIP(hipMemcpyAsync(ws->h_res + ws->r_wk, ws->d_res + ws->r_wk, red_blocks(ws) * sizeof(double),
                        hipMemcpyDeviceToHost, ws->stream));
  ws->pend_wk1 = ws->wavekin1;
  ws->pend_switched = false;
  return defer_results(ws, 2);
}

extern "C" int pw_waves_switch(pw_waveset *ws) {
  if (!ws) return pw_fail(PW_ERR_ARG, "pw_waves_switch: null waveset");
  std::swap(ws->d_psi0, ws->d_psim);
  // NLAMBDAX=2 extrapolation (src/paw_waves2.f90:2236-2297), on the device
  const int n2 = ws->nb * ws->nb;
  k_lam_switch<<<(n2 + 255) / 256, 256, 0, ws->stream>>>(n2, ws->have_rlamm ? 0 : 1, ws->d_rlam0, ws->d_rlamm);
  PW_HIP(hipGetLastError());
  // NLAMBDA=0: RLAM0 becomes the zero guess; NLAMBDA>=1: the old RLAM0 becomes RLAMM
  ws->have_rlamm = ws->have_rlam0;
  ws->have_rlam0 = true;
  if (ws->res_pending & 2) ws->pend_switched = true;
  ws->wavekin1 = 0.0;
  ws->wavekin2 = 0.0;
  return PW_OK;
}

extern "C" int pw_waves_step(pw_waveset *ws) {
  PW_TRY(pw_waves_etot(ws));
  PW_TRY(pw_waves_propagate(ws));
  PW_TRY(pw_waves_orthogonalize(ws));
  PW_TRY(pw_waves_switch(ws));
  return PW_OK;
}

extern "C" int pw_waves_vpsi(pw_waveset *ws, const void *psi, void *hpsi) {
  if (!ws || !psi || !hpsi) return pw_fail(PW_ERR_ARG, "pw_waves_vpsi: null argument");
  PW_TRY(ws_finish(ws));
  PW_TRY(need(ws->have_v, "VOFR"));
  ScopeTimer tm(ws, "vpsi");
  return fft_vpsi(ws->plan, ws->nbh, (const double2 *)psi, ws->d_v, (double2 *)hpsi, 1, ws->prec, ws->stream);
}

extern "C" int pw_waves_density(pw_waveset *ws, const void *psi, double *rho) {
  if (!ws || !psi || !rho) return pw_fail(PW_ERR_ARG, "pw_waves_density: null argument");
  PW_TRY(ws_finish(ws));
  PW_TRY(need(!ws->occ.empty(), "OCC"));
  ScopeTimer tm(ws, "density");
  return fft_density(ws->plan, ws->nbh, (const double2 *)psi, ws->d_occ2, rho, ws->prec, ws->stream);
}
