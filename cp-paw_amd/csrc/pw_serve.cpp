// Request loop for independent wave-function relaxations on one GPU.
//
// The reference relaxes one system per run (damped CP dynamics with the
// autopilot's energy window, src/paw_assist.f90:297-400).  The server here
// keeps many such jobs in flight: requests are queued as they arrive, admitted
// first-fit while their device footprint fits the memory budget and a slot is
// free, and each admitted job runs its CP loop on its own HIP stream, so the
// kernels of jobs of different sizes overlap on the device.  A job's state
// (coefficients, projections, potential, captured ORTHO_X graph) stays
// resident in HBM from admission to completion.  Results do not depend on
// what else runs: every kernel sums in a fixed order and no state is shared.
// Launch shapes come from pw_waves_autotune, once per system shape.
// Finished plans and wavesets stay allocated in a pool, per system shape, and
// serve the next request of that shape (reset to their created state): once
// each shape has been seen, a request makes no plan set-up and no device
// allocation or free.  Pooled sets count against the memory budget and the
// oldest are freed when a request does not fit otherwise.
#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <tuple>
#include <vector>

#include "pw_common.h"

namespace {

struct Job {
  long long id = 0;
  std::string key;                 // system shape (pool key)
  pw_plan *plan = nullptr;         // a pooled plan and waveset of this shape, if one was free
  pw_waveset *ws = nullptr;
  double rbas[9];
  double epw_psi = 0, epw_rho = 0;
  int nb = 0, nat = 0, nsp = 0, lnxx = 0, prec = 0, max_steps = 0, nconv = 0;
  std::vector<int> isp, lnx, lox;
  std::vector<double> pro, occ, dov, dh, vofr, r, psi;
  double dt = 0, emass = 0, emasscg2 = 0, friction = 0, etol = 0, rtol = 0;
  long long footprint = 0;
  bool tune = true;
  // results
  bool done = false;
  int status = PW_OK;
  std::string err;
  int steps = 0, converged = 0;
  double etot = 0, ekin = 0, epot = 0, e1c = 0, residual = 0;
  double t_submit = 0, t_start = 0, t_end = 0;
};

using Clock = std::chrono::steady_clock;

// launch shapes per system shape, shared by all jobs of the process
std::mutex g_tune_mu;
std::map<std::tuple<int, int, int, int, int>, std::array<int, 3>> g_tuned;

}  // namespace

struct pw_server {
  long long budget = 0;
  int max_active = 1;
  struct Idle {
    std::string key;
    pw_plan *plan;
    pw_waveset *ws;
    long long footprint;
  };
  std::deque<Idle> idle;           // finished sets, oldest first
  long long idle_bytes = 0;
  long long reused = 0, created = 0;
  // per shape: footprint, NGL, NRL (computed once from a host-only plan)
  std::map<std::string, std::array<long long, 3>> shapes;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<std::shared_ptr<Job>> pending;
  std::map<long long, std::shared_ptr<Job>> jobs;
  long long used = 0, next_id = 1, peak_used = 0;
  int active = 0, peak_active = 0;
  bool stop = false;
  std::vector<std::thread> workers;
  Clock::time_point t0 = Clock::now();
  double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }
};

namespace {

// the job's CP loop on its plan and waveset (pooled ones reset, or new);
// on success they are left in j->plan / j->ws for the pool, on failure freed
int run_job(Job *j) {
  pw_plan *plan = j->plan;
  pw_waveset *ws = j->ws;
  int rc = PW_OK;
  if (ws) {
    rc = ws_reset_run(ws);
  } else {
    const double k0[3] = {0.0, 0.0, 0.0};
    rc = pw_plan_create(j->rbas, k0, 1, j->epw_psi, j->epw_rho, &plan);
    if (rc == PW_OK)
      rc = pw_waves_create(plan, j->nb, j->nat, j->nsp, j->isp.data(), j->lnx.data(), j->lnxx, j->lox.data(),
                           j->prec, &ws);
  }
  auto r8a = [&](const char *id, const std::vector<double> &v) {
    if (rc == PW_OK) rc = pw_waves_setr8a(ws, id, (int)v.size(), v.data());
  };
  auto c8a = [&](const char *id, const std::vector<double> &v) {
    if (rc == PW_OK) rc = pw_waves_setc8a(ws, id, (int)(v.size() / 2), v.data());
  };
  auto r8 = [&](const char *id, double v) {
    if (rc == PW_OK) rc = pw_waves_setr8(ws, id, v);
  };
  r8a("PRO", j->pro);
  r8a("DO", j->dov);
  c8a("DH", j->dh);
  r8a("OCC", j->occ);
  r8a("VOFR", j->vofr);
  r8("TIMESTEP", j->dt);
  r8("EMASS", j->emass);
  r8("EMASSCG2", j->emasscg2);
  r8("FRICTION", j->friction);
  r8a("R0", j->r);
  r8a("RM", j->r);
  r8a("RP", j->r);
  c8a("PSI0", j->psi);
  c8a("PSIM", j->psi);
  if (rc == PW_OK && j->tune) rc = pw_waves_autotune(ws);
  if (rc == PW_OK && j->rtol > 0.0) rc = pw_waves_setr8(ws, "RESIDUALCHECK", 1.0);
  // damped CP steps until |dE| < etol (or, with rtol, the residual of the
  // previous step's ORTHOGONALIZE < rtol) on nconv consecutive steps
  double eprev = 0.0;
  int calm = 0;
  for (int it = 0; rc == PW_OK && it < j->max_steps; ++it) {
    rc = pw_waves_etot(ws);
    if (rc == PW_OK) rc = ws_finish(ws);   // ETOT's energies and the last residual on the host
    if (rc != PW_OK) break;
    j->etot = ws->etot;
    j->ekin = ws->ekin;
    j->epot = ws->epot;
    j->e1c = ws->e1c;
    j->residual = ws->residual;
    j->steps = it + 1;
    const bool calm_now = j->rtol > 0.0 ? ws->residual < j->rtol : std::fabs(ws->etot - eprev) < j->etol;
    calm = (it > 0 && calm_now) ? calm + 1 : 0;
    eprev = ws->etot;
    if (calm >= j->nconv) {
      j->converged = 1;
      break;
    }
    rc = pw_waves_propagate(ws);
    if (rc == PW_OK) rc = pw_waves_orthogonalize(ws);
    if (rc == PW_OK) rc = pw_waves_switch(ws);
  }
  if (rc != PW_OK) {
    char buf[512];
    pw_last_error(buf, sizeof(buf));
    j->err = buf;
    if (ws) pw_waves_destroy(ws);
    if (plan) pw_plan_destroy(plan);
    ws = nullptr;
    plan = nullptr;
  }
  j->plan = plan;
  j->ws = ws;
  return rc;
}

void destroy_idle(std::vector<pw_server::Idle> &v) {
  for (auto &e : v) {
    pw_waves_destroy(e.ws);
    pw_plan_destroy(e.plan);
  }
  v.clear();
}

// first pending job that fits the free budget (FIFO among those that fit)
std::deque<std::shared_ptr<Job>>::iterator admissible(pw_server *s) {
  if (s->active >= s->max_active) return s->pending.end();
  for (auto it = s->pending.begin(); it != s->pending.end(); ++it)
    if (s->used + (*it)->footprint <= s->budget) return it;
  return s->pending.end();
}

void worker(pw_server *s) {
  for (;;) {
    std::shared_ptr<Job> job;
    std::vector<pw_server::Idle> evict;
    {
      std::unique_lock<std::mutex> lk(s->mu);
      s->cv_work.wait(lk, [&] { return (s->stop && s->pending.empty()) || admissible(s) != s->pending.end(); });
      auto it = admissible(s);
      if (it == s->pending.end()) return;   // stopping and drained
      job = *it;
      s->pending.erase(it);
      // a pooled set of the same shape, or room for a new one (oldest pooled
      // sets freed first)
      auto same = std::find_if(s->idle.begin(), s->idle.end(),
                               [&](const pw_server::Idle &e) { return e.key == job->key; });
      if (same != s->idle.end()) {
        job->plan = same->plan;
        job->ws = same->ws;
        s->idle_bytes -= same->footprint;
        s->idle.erase(same);
        s->reused++;
      } else {
        while (!s->idle.empty() && s->used + s->idle_bytes + job->footprint > s->budget) {
          evict.push_back(s->idle.front());
          s->idle_bytes -= s->idle.front().footprint;
          s->idle.pop_front();
        }
        s->created++;
      }
      s->used += job->footprint;
      s->active++;
      s->peak_used = std::max(s->peak_used, s->used + s->idle_bytes);
      s->peak_active = std::max(s->peak_active, s->active);
      job->t_start = s->now();
    }
    destroy_idle(evict);
    const int rc = run_job(job.get());
    {
      std::lock_guard<std::mutex> lk(s->mu);
      job->status = rc;
      job->t_end = s->now();
      job->done = true;
      s->used -= job->footprint;
      s->active--;
      if (job->ws) {
        s->idle.push_back(pw_server::Idle{job->key, job->plan, job->ws, job->footprint});
        s->idle_bytes += job->footprint;
        job->plan = nullptr;
        job->ws = nullptr;
      }
    }
    s->cv_done.notify_all();
    s->cv_work.notify_all();
  }
}

}  // namespace

// ---------------------------------------------------------------------------
// autotuner: G-split width of the projection/overlap GEMMs and the band
// grouping of the density pass, timed with HIP events on the waveset's
// stream; the choice is cached per (NGL, NBH, NPRO, NR1, prec)
// ---------------------------------------------------------------------------
extern "C" int pw_waves_autotune(pw_waveset *ws) {
  if (!ws) return pw_fail(PW_ERR_ARG, "pw_waves_autotune: null waveset");
  PW_TRY(ws_finish(ws));
  if (!ws->have_psi0 || ws->r0.empty())
    return pw_fail(PW_ERR_STATE, "pw_waves_autotune: PSI0 and R0 must be set");
  pw_plan *p = ws->plan;
  const auto key = std::make_tuple(p->ngl, ws->nbh, ws->npro, p->nr1, ws->prec);
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tuned.find(key);
    if (it != g_tuned.end()) {
      ws->tune_split_waves = it->second[0];
      ws->tune_dens_blocks = it->second[1];
      ws->tune_col_waves = it->second[2];
      return PW_OK;
    }
  }
  hipEvent_t a, b;
  PW_HIP(hipEventCreate(&a));
  PW_HIP(hipEventCreate(&b));
  auto timed = [&](auto &&fn) -> double {
    double best = 1e30;
    for (int rep = 0; rep < 3; ++rep) {
      if (hipEventRecord(a, ws->stream) != hipSuccess) return 1e30;
      if (fn() != PW_OK) return 1e30;
      if (hipEventRecord(b, ws->stream) != hipSuccess || hipEventSynchronize(b) != hipSuccess) return 1e30;
      float ms = 0.f;
      hipEventElapsedTime(&ms, a, b);
      if (rep > 0) best = std::min(best, (double)ms);   // first run warms up
    }
    return best;
  };
  int rc = pw_waves_projections(ws);   // phase tables (and split projectors) for R0
  int best_split = ws->tune_split_waves, best_dens = ws->tune_dens_blocks, best_col = ws->tune_col_waves;
  if (rc == PW_OK) {
    // G-split width of the row products: more splits fill the GPU, fewer
    // shrink the partial sums their ordered reduction reads back; timed on
    // the projections and one band overlap (the two row-product shapes),
    // reductions included, into the Gram-Schmidt scratch matrix
    double tbest = 1e30;
    double *mat = ws->d_small_mats + (size_t)4 * ws->nb * ws->nb;
    for (int cand : {256, 512, 1024, 2048, 4096}) {
      ws->tune_split_waves = cand;
      const double t = timed([&] {
        int r = k_projections(ws, ws->d_psi0, ws->d_phase, ws->d_oproj);
        if (r == PW_OK) r = k_overlap(ws, ws->d_psi0, ws->d_psi0, mat, false, true);
        return r;
      });
      if (t < tbest) { tbest = t; best_split = cand; }
    }
    tbest = 1e30;
    for (int cand : {192, 256, 384, 448, 512, 1024}) {
      const double t = timed([&] {
        return fft_density(p, ws->nbh, ws->d_psi0, ws->d_occ2, ws->d_rho, ws->prec, ws->stream, cand);
      });
      if (t < tbest) { tbest = t; best_dens = cand; }
    }
    if (ws->prec == PW_PREC_MIXED) {
      // workgroup width of the column products (ADDPRO, ADDOPSI): 4 waves
      // share the coefficient tile over 64 G, 2 waves give twice the
      // workgroups over 32 G; timed on ADDPRO into the OPSI buffer, which
      // every ORTHOGONALIZE rebuilds from PSI0
      tbest = 1e30;
      for (int cand : {2, 4}) {
        ws->tune_col_waves = cand;
        const double t = timed([&] { return k_addpro(ws, ws->d_opsi, ws->d_phase, ws->d_proj); });
        if (t < tbest) { tbest = t; best_col = cand; }
      }
    }
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  if (rc != PW_OK) return rc;
  ws->tune_split_waves = best_split;
  ws->tune_dens_blocks = best_dens;
  ws->tune_col_waves = best_col;
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_tuned[key] = {best_split, best_dens, best_col};
  return PW_OK;
}

extern "C" int pw_waves_footprint(pw_plan *plan, int nb, int nat, int nsp, const int *isp, const int *lnx,
                                  int lnxx, const int *lox, int prec, long long *bytes) {
  if (!plan || nb <= 0 || nat <= 0 || nsp <= 0 || !isp || !lnx || !lox || !bytes)
    return pw_fail(PW_ERR_ARG, "pw_waves_footprint: bad argument");
  long long npro = 0, lmx = 1, nbare = 0, lmnxx = 1;
  for (int s = 0; s < nsp; ++s) {
    long long lmn = 0;
    for (int ln = 0; ln < lnx[s]; ++ln) {
      const int l = lox[ln + lnxx * s];
      lmn += 2 * l + 1;
      lmx = std::max<long long>(lmx, (l + 1) * (l + 1));
    }
    nbare += lnx[s];
    lmnxx = std::max(lmnxx, lmn);
  }
  for (int a = 0; a < nat; ++a) {
    const int s = isp[a] - 1;
    if (s < 0 || s >= nsp) return pw_fail(PW_ERR_ARG, "pw_waves_footprint: species out of range");
    long long lmn = 0;
    for (int ln = 0; ln < lnx[s]; ++ln) lmn += 2 * lox[ln + lnxx * s] + 1;
    npro += lmn;
  }
  const long long ngl = plan->ngl, nrl = plan->nrl, nbh = (nb + 1) / 2;
  const long long es = prec == PW_PREC_FP64 ? 16 : 8;
  const long long nglp = (ngl + 63) / 64 * 64;
  long long b = 0;
  b += 5 * nbh * ngl * 16;                          // PSI0, PSIM, HPSI, OPSI, scratch
  b += 3 * npro * nbh * 16;                         // PROJ, OPROJ, XPROJ
  b += (nbare + lmx + 1) * ngl * 8 + 3 * nrl * 8;   // tables, V, rho
  b += 64 * npro * nbh * 16;                        // G-split partials
  b += nbh * (long long)plan->nstripe * plan->nr1 * es + nrl * es;   // stripes, V transposed
  b += 16 * nrl * 8;                                // density partials
  b += (long long)(21 + 6) * nb * nb * 8;           // NB x NB matrices, ORTHO_X work
  b += ngl * 48 + (long long)nat * lmnxx * lmnxx * 32;
  if (prec == PW_PREC_MIXED) b += (nsp * lmnxx * 12 + nat * 24) * nglp;  // projector factors, FP64 and FP32
  *bytes = b + b / 10 + (8ll << 20);
  return PW_OK;
}

// ---------------------------------------------------------------------------
// C ABI: server
// ---------------------------------------------------------------------------
extern "C" int pw_server_create(long long mem_budget, int max_active, pw_server **srv) {
  if (!srv || mem_budget <= 0 || max_active <= 0 || max_active > 64)
    return pw_fail(PW_ERR_ARG, "pw_server_create: bad argument");
  pw_server *s = new pw_server();
  s->budget = mem_budget;
  s->max_active = max_active;
  for (int i = 0; i < max_active; ++i) s->workers.emplace_back(worker, s);
  *srv = s;
  return PW_OK;
}

extern "C" int pw_server_destroy(pw_server *s) {
  if (!s) return PW_OK;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv_work.notify_all();
  for (auto &t : s->workers) t.join();
  std::vector<pw_server::Idle> all(s->idle.begin(), s->idle.end());
  destroy_idle(all);
  delete s;
  return PW_OK;
}

extern "C" int pw_server_submit(pw_server *s, const pw_request *q, long long *id) {
  if (!s || !q || !id) return pw_fail(PW_ERR_ARG, "pw_server_submit: null argument");
  if (q->nb <= 0 || q->nat <= 0 || q->nsp <= 0 || q->lnxx <= 0 || !q->isp || !q->lnx || !q->lox || !q->pro ||
      !q->occ || !q->dov || !q->dh || !q->vofr || !q->r || !q->psi || q->nbarepro <= 0 || q->max_steps <= 0)
    return pw_fail(PW_ERR_ARG, "pw_server_submit: incomplete request");
  // the system shape: cell, cutoffs, bands, atoms, species, projectors, form
  std::string key;
  auto put = [&](const void *p, size_t n) { key.append(static_cast<const char *>(p), n); };
  put(q->rbas, 9 * sizeof(double));
  put(&q->epw_psi, sizeof(double));
  put(&q->epw_rho, sizeof(double));
  const int dims[5] = {q->nb, q->nat, q->nsp, q->lnxx, q->prec};
  put(dims, sizeof(dims));
  put(q->isp, (size_t)q->nat * sizeof(int));
  put(q->lnx, (size_t)q->nsp * sizeof(int));
  put(q->lox, (size_t)q->lnxx * q->nsp * sizeof(int));
  // footprint and sizes, once per shape from a host-only plan (no device memory)
  std::array<long long, 3> shape{};
  bool known = false;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->shapes.find(key);
    if (it != s->shapes.end()) {
      shape = it->second;
      known = true;
    }
  }
  if (!known) {
    pw_plan *plan = nullptr;
    const double k0[3] = {0.0, 0.0, 0.0};
    PW_TRY(pw_plan_create(q->rbas, k0, 1, q->epw_psi, q->epw_rho, &plan));
    long long f = 0;
    const int rc = pw_waves_footprint(plan, q->nb, q->nat, q->nsp, q->isp, q->lnx, q->lnxx, q->lox, q->prec, &f);
    shape = {f, plan->ngl, plan->nrl};
    pw_plan_destroy(plan);
    if (rc != PW_OK) return rc;
    std::lock_guard<std::mutex> lk(s->mu);
    s->shapes[key] = shape;
  }
  auto job = std::make_shared<Job>();
  const long long fp = shape[0];
  const int ngl = (int)shape[1], nrl = (int)shape[2];
  if (fp > s->budget) return pw_fail(PW_ERR_UNSUPPORTED, "pw_server_submit: request larger than the memory budget");
  int lmnxx = 1;
  for (int sp = 0; sp < q->nsp; ++sp) {
    int lmn = 0;
    for (int ln = 0; ln < q->lnx[sp]; ++ln) lmn += 2 * q->lox[ln + q->lnxx * sp] + 1;
    lmnxx = std::max(lmnxx, lmn);
  }
  const int nbh = (q->nb + 1) / 2;
  Job &j = *job;
  j.key = std::move(key);
  std::copy(q->rbas, q->rbas + 9, j.rbas);
  j.epw_psi = q->epw_psi;
  j.epw_rho = q->epw_rho;
  j.nb = q->nb;
  j.nat = q->nat;
  j.nsp = q->nsp;
  j.lnxx = q->lnxx;
  j.prec = q->prec;
  j.max_steps = q->max_steps;
  j.nconv = std::max(1, q->nconv);
  j.isp.assign(q->isp, q->isp + q->nat);
  j.lnx.assign(q->lnx, q->lnx + q->nsp);
  j.lox.assign(q->lox, q->lox + (size_t)q->lnxx * q->nsp);
  j.pro.assign(q->pro, q->pro + (size_t)ngl * q->nbarepro);
  j.occ.assign(q->occ, q->occ + q->nb);
  j.dov.assign(q->dov, q->dov + (size_t)q->lnxx * q->lnxx * q->nsp);
  j.dh.assign(q->dh, q->dh + (size_t)2 * lmnxx * lmnxx * q->nat);
  j.vofr.assign(q->vofr, q->vofr + nrl);
  j.r.assign(q->r, q->r + 3 * q->nat);
  j.psi.assign(q->psi, q->psi + (size_t)2 * ngl * nbh);
  j.dt = q->dt;
  j.emass = q->emass;
  j.emasscg2 = q->emasscg2;
  j.friction = q->friction;
  j.etol = q->etol;
  j.rtol = q->rtol;
  j.tune = q->autotune != 0;
  j.footprint = fp;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stop) return pw_fail(PW_ERR_STATE, "pw_server_submit: server is stopping");
    j.id = s->next_id++;
    j.t_submit = s->now();
    s->jobs[j.id] = job;
    s->pending.push_back(job);
    *id = j.id;
  }
  s->cv_work.notify_all();
  return PW_OK;
}

extern "C" int pw_server_wait(pw_server *s, long long id, pw_result *res) {
  if (!s || !res) return pw_fail(PW_ERR_ARG, "pw_server_wait: null argument");
  std::shared_ptr<Job> j;
  {
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->jobs.find(id);
    if (it == s->jobs.end()) return pw_fail(PW_ERR_ARG, "pw_server_wait: unknown request id");
    j = it->second;
    s->cv_done.wait(lk, [&] { return j->done; });
    s->jobs.erase(it);
  }
  res->status = j->status;
  res->steps = j->steps;
  res->converged = j->converged;
  res->etot = j->etot;
  res->ekin = j->ekin;
  res->epot = j->epot;
  res->e1c = j->e1c;
  res->t_submit = j->t_submit;
  res->t_start = j->t_start;
  res->t_end = j->t_end;
  res->footprint = j->footprint;
  res->residual = j->residual;
  if (j->status != PW_OK) return pw_fail(j->status, "request " + std::to_string(id) + ": " + j->err);
  return PW_OK;
}

extern "C" int pw_server_stats(pw_server *s, long long *peak_bytes, int *peak_active, int *pending) {
  if (!s || !peak_bytes || !peak_active || !pending) return pw_fail(PW_ERR_ARG, "pw_server_stats: null argument");
  std::lock_guard<std::mutex> lk(s->mu);
  *peak_bytes = s->peak_used;
  *peak_active = s->peak_active;
  *pending = (int)s->pending.size();
  return PW_OK;
}

extern "C" int pw_server_pool_stats(pw_server *s, long long *reused, long long *created, long long *idle_bytes) {
  if (!s || !reused || !created || !idle_bytes) return pw_fail(PW_ERR_ARG, "pw_server_pool_stats: null argument");
  std::lock_guard<std::mutex> lk(s->mu);
  *reused = s->reused;
  *created = s->created;
  *idle_bytes = s->idle_bytes;
  return PW_OK;
}
