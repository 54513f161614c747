// Serving loop: CD3D / TCD3D requests admitted by an HBM budget and run
// concurrently, each on its own hierarchy and stream, over the public C ABI.
//
// One scheduler thread admits queued requests in arrival order (a request
// that does not fit may be overtaken at most max_bypass times) and hands
// them to worker threads.  Device memory is accounted per hierarchy: an
// admitted request first reserves the estimate (cd3d_estimate_bytes), and the
// reservation is replaced by the hierarchy's real size once it has run.  Idle
// hierarchies stay resident for later requests of the same mesh and are
// evicted least recently used first when a new one needs room.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <list>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

#include "../../include/cd3d_server.h"

// capi.cpp: the handle's coarse inverse stays on its main stream
void cd3d_internal_serving_handle(cd3d_problem *p);

namespace {

double now_ms() {
  using namespace std::chrono;
  return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

bool same_mesh(const cd3d_mesh_desc &a, const cd3d_mesh_desc &b) {
  return a.order == b.order && a.m == b.m && a.k0 == b.k0 && a.levels == b.levels && a.box[0] == b.box[0] &&
         a.box[1] == b.box[1] && a.box[2] == b.box[2];
}

int64_t finest_dofs(const cd3d_mesh_desc &m) {
  const int64_t nn = (int64_t)m.order * ((int64_t)m.m << (m.k0 + m.levels - 1)) + 1;
  return nn * nn * nn;
}

struct Resident {
  cd3d_mesh_desc mesh{};
  cd3d_problem *h = nullptr;
  size_t bytes = 0;  // accounted bytes (reservation until the first run ends)
  bool busy = false;
  double last_used = 0;
};

struct Pending {
  uint64_t id = 0;
  cd3d_request req{};
  double t_submit = 0;
  size_t est = 0;
  int bypassed = 0;
};

struct Work {
  Pending p;
  Resident *slot = nullptr;
  bool reused = false;
  double t_admit = 0;
};

struct Done {
  cd3d_result res{};
  std::vector<double> sol;
};

}  // namespace

struct cd3d_server {
  cd3d_server_params prm{};
  int device = 0;
  std::mutex mu;
  std::condition_variable cv_sched, cv_work, cv_done;
  std::deque<Pending> queue;
  std::deque<Work> ready;
  std::list<Resident> resident;  // stable addresses
  std::map<uint64_t, Done> done;
  std::map<uint64_t, int> outstanding;  // submitted, not yet waited for
  uint64_t next_id = 1;
  size_t bytes = 0, peak = 0;
  int inflight = 0;
  bool stop = false;
  uint64_t submitted = 0, completed = 0, failed = 0, rejected = 0;
  double t_first = -1, t_last = 0;
  std::vector<double> latencies;
  std::thread scheduler;
  std::vector<std::thread> workers;

  size_t estimate(const cd3d_request &r) const {
    const bool cellwise = r.coeffs.conv_kind != 0 || r.coeffs.f_kind != 0;
    const int restart = std::max(1, std::min(r.solver.restart, r.solver.maxit));
    const int tcd = r.kind == CD3D_REQ_TCD ? (r.coeffs.supg ? 2 : 1) : 0;
    return cd3d_estimate_bytes(&r.mesh, restart, cellwise ? 1 : 0, tcd);
  }

  // evict idle hierarchies (least recently used first, never `keep`) until
  // `need` more bytes fit; returns whether they do
  bool make_room(size_t need, const Resident *keep, std::vector<cd3d_problem *> &to_destroy) {
    while (bytes + need > prm.hbm_budget) {
      auto victim = resident.end();
      for (auto it = resident.begin(); it != resident.end(); ++it)
        if (!it->busy && &*it != keep && (victim == resident.end() || it->last_used < victim->last_used)) victim = it;
      if (victim == resident.end()) return false;
      bytes -= victim->bytes;
      if (victim->h) to_destroy.push_back(victim->h);
      resident.erase(victim);
    }
    return true;
  }

  void schedule_loop() {
    std::unique_lock<std::mutex> lk(mu);
    while (true) {
      // woken by a submission or a completion
      cv_sched.wait(lk, [&] { return stop || (wake && !queue.empty() && inflight < prm.max_inflight); });
      if (stop) return;
      wake = false;
      std::vector<cd3d_problem *> to_destroy;
      bool admitted_any = false;
      for (auto it = queue.begin(); it != queue.end() && inflight < prm.max_inflight;) {
        Pending &p = *it;
        Resident *slot = nullptr;
        for (auto &r : resident)
          if (!r.busy && same_mesh(r.mesh, p.req.mesh)) {
            slot = &r;
            break;
          }
        const size_t need = slot ? (p.est > slot->bytes ? p.est - slot->bytes : 0) : p.est;
        if (!make_room(need, slot, to_destroy)) {
          if (it == queue.begin() && p.bypassed >= prm.max_bypass) break;  // no overtaking past the limit
          ++it;
          continue;
        }
        if (it != queue.begin()) queue.front().bypassed++;
        Work w;
        w.p = p;
        w.t_admit = now_ms();
        if (slot) {
          slot->bytes += need;
          w.reused = true;
        } else {
          resident.emplace_back();
          slot = &resident.back();
          slot->mesh = p.req.mesh;
          slot->bytes = need;
        }
        slot->busy = true;
        w.slot = slot;
        bytes += need;
        peak = std::max(peak, bytes);
        inflight++;
        ready.push_back(w);
        admitted_any = true;
        it = queue.erase(it);
      }
      if (!to_destroy.empty()) {
        lk.unlock();
        for (auto *h : to_destroy) cd3d_destroy(h);
        lk.lock();
      }
      if (admitted_any) cv_work.notify_all();
    }
  }
  bool wake = false;  // something changed since the last admission pass

  // the continuation of a Jacobi-smoothed solve that has not converged
  // after `done` iterations: the multicolour SSOR smoother with 4 + 4 sweeps
  // (with 2 + 2, a Q2 request of 80^3 cells on 5^3 macro bricks with eps
  // 3.2e-6 still stands at 0.8 of its initial residual after 400 iterations;
  // with 4 + 4 it converges in 57), within the request's iteration budget
  // and to its tolerance relative to the request's initial residual
  static cd3d_solver_params escalated(const cd3d_solver_params &req, int done, double r0, double rfinal) {
    cd3d_solver_params sp = req;
    sp.smoother = 13;
    sp.omega = 1.0;
    sp.pre = sp.post = 4;
    sp.maxit = req.maxit - done;
    if (req.rel_tol > 0) sp.rel_tol = req.rel_tol * r0 / rfinal;
    return sp;
  }

  void run_request(Work &w, Done &d) {
    const cd3d_request &r = w.p.req;
    cd3d_result &res = d.res;
    res.n_dof = (int)finest_dofs(r.mesh);
    if (prm.simulate) {
      std::this_thread::sleep_for(std::chrono::microseconds((int64_t)(prm.sim_us_per_dof * res.n_dof)));
      res.status = CD3D_OK;
      if (r.want_solution) d.sol.assign(res.n_dof, 0.0);
      return;
    }
    cd3d_problem *&h = w.slot->h;
    cd3d_status st = CD3D_OK;
    if (!h) {
      st = cd3d_create(&r.mesh, &h);
      if (st == CD3D_OK) cd3d_internal_serving_handle(h);
    }
    if (st != CD3D_OK) {
      h = nullptr;
      res.status = st;
      return;
    }
    cd3d_stats s{};
    if (r.kind == CD3D_REQ_CD) {
      // Jacobi-smoothed solves of convection-dominated problems can stall
      // far from the tolerance (the reference's algorithm does too): after
      // robust_after iterations such a solve continues from its iterate with
      // the multicolour SSOR smoother, to the same tolerance relative to the
      // request's initial residual and within its iteration budget
      cd3d_solver_params sp = r.solver;
      const bool robust = prm.robust_after > 0 && sp.smoother == 1 && sp.maxit > prm.robust_after;
      if (robust) sp.maxit = prm.robust_after;
      st = cd3d_assemble(h, &r.coeffs);
      if (st == CD3D_OK) st = cd3d_solve(h, &sp, nullptr, nullptr, 0, &s);
      res.iterations = s.iterations;
      res.rel_residual = s.r0 > 0 ? s.rfinal / s.r0 : 0.0;
      if (robust && st == CD3D_NOT_CONVERGED && s.rfinal > 0) {
        std::vector<double> x(res.n_dof);
        st = cd3d_get_solution(h, x.data());
        const cd3d_solver_params sp2 = escalated(r.solver, s.iterations, s.r0, s.rfinal);
        const double r0 = s.r0;
        cd3d_stats s2{};
        if (st == CD3D_OK) st = cd3d_solve(h, &sp2, nullptr, x.data(), 1, &s2);
        res.iterations += s2.iterations;
        res.rel_residual = r0 > 0 ? s2.rfinal / r0 : 0.0;
        res.escalated = 1;
      }
      if ((st == CD3D_OK || st == CD3D_NOT_CONVERGED) && r.want_solution) {
        d.sol.resize(res.n_dof);
        const cd3d_status g = cd3d_get_solution(h, d.sol.data());
        if (g != CD3D_OK) st = g;
      }
    } else {
      std::vector<double> u(res.n_dof);
      st = cd3d_interp_pulse(h, r.mesh.levels - 1, &r.coeffs, u.data());
      if (st == CD3D_OK) st = cd3d_tcd_init(h, &r.coeffs, r.t0, u.data());
      const bool initialised = st == CD3D_OK;
      // a step that does not converge ends the run: later steps would start
      // from an unconverged state
      for (int k = 0; k < r.nsteps && st == CD3D_OK; k++) {
        st = cd3d_tcd_step(h, &r.coeffs, &r.solver, r.tau, r.theta, nullptr, &s);
        res.iterations += s.iterations;
        if (s.r0 > 0) res.rel_residual = std::max(res.rel_residual, s.rfinal / s.r0);
        if (st == CD3D_OK) res.steps_done = k + 1;
      }
      if ((st == CD3D_OK || st == CD3D_NOT_CONVERGED) && initialised && r.want_solution) {
        d.sol.resize(res.n_dof);
        const cd3d_status g = cd3d_tcd_get_state(h, d.sol.data());
        if (g != CD3D_OK) st = g;
      }
    }
    res.status = st;
  }

  // CD requests that cd3d_solve_batch runs in lockstep: FP64, damped Jacobi,
  // V-cycle, exact coarse solve, at least one pre-smoothing sweep
  bool batchable(const cd3d_request &r) const {
    const cd3d_solver_params &s = r.solver;
    return !prm.simulate && prm.max_batch > 1 && r.kind == CD3D_REQ_CD &&
           (s.precision == CD3D_FP64 || s.precision == CD3D_MIXED) &&
           s.smoother == 1 && s.coarse_smoother == 17 && s.cycle == 1 && s.pre >= 1 &&
           std::min(s.restart, s.maxit) <= 200;
  }
  static bool same_batch(const cd3d_request &a, const cd3d_request &b) {
    return a.mesh.levels == b.mesh.levels && memcmp(&a.solver, &b.solver, sizeof(a.solver)) == 0;
  }

  // admitted CD requests of one batch: each assembles on its own hierarchy,
  // all solve together; a member that has not converged after robust_after
  // iterations continues alone, as in run_request
  void run_batch(std::vector<Work> &ws, std::vector<Done> &ds) {
    const int n = (int)ws.size();
    std::vector<cd3d_problem *> hs;
    std::vector<int> idx;
    cd3d_solver_params sp = ws[0].p.req.solver;
    const bool robust = prm.robust_after > 0 && sp.maxit > prm.robust_after;
    if (robust) sp.maxit = prm.robust_after;
    // hierarchies and assembly of the members, on up to 8 host threads (each
    // member has its own stream; assembly ends with a synchronisation)
    auto prepare = [&](int k) {
      const cd3d_request &r = ws[k].p.req;
      ds[k].res.n_dof = (int)finest_dofs(r.mesh);
      cd3d_problem *&h = ws[k].slot->h;
      cd3d_status st = CD3D_OK;
      if (!h) {
        st = cd3d_create(&r.mesh, &h);
        if (st == CD3D_OK) cd3d_internal_serving_handle(h);
        else h = nullptr;
      }
      if (st == CD3D_OK) st = cd3d_assemble(h, &r.coeffs);
      ds[k].res.status = st;
    };
    {
      std::atomic<int> next{0};
      std::vector<std::thread> th;
      const int nt = std::min(n, 8) - 1;  // the worker itself takes part
      auto drain = [&] {
        for (int k; (k = next++) < n;) prepare(k);
      };
      for (int t = 0; t < nt; t++)
        th.emplace_back([&] {
          (void)hipSetDevice(device);
          drain();
        });
      drain();
      for (auto &x : th) x.join();
    }
    for (int k = 0; k < n; k++)
      if (ds[k].res.status == CD3D_OK) {
        hs.push_back(ws[k].slot->h);
        idx.push_back(k);
      }
    if (hs.empty()) return;
    std::vector<cd3d_stats> st(hs.size());
    const cd3d_status bst = cd3d_solve_batch(hs.data(), (int)hs.size(), &sp, st.data());
    for (size_t q = 0; q < hs.size(); q++) {
      const int k = idx[q];
      const cd3d_request &r = ws[k].p.req;
      cd3d_result &res = ds[k].res;
      cd3d_problem *h = hs[q];
      cd3d_status s1 = (bst == CD3D_OK || bst == CD3D_NOT_CONVERGED || bst == CD3D_DIVERGED ||
                        bst == CD3D_NONFINITE) ? (cd3d_status)st[q].status : bst;
      res.iterations = st[q].iterations;
      res.rel_residual = st[q].r0 > 0 ? st[q].rfinal / st[q].r0 : 0.0;
      if (robust && s1 == CD3D_NOT_CONVERGED && st[q].rfinal > 0) {
        std::vector<double> x(res.n_dof);
        s1 = cd3d_get_solution(h, x.data());
        const cd3d_solver_params sp2 = escalated(r.solver, st[q].iterations, st[q].r0, st[q].rfinal);
        cd3d_stats s2{};
        if (s1 == CD3D_OK) s1 = cd3d_solve(h, &sp2, nullptr, x.data(), 1, &s2);
        res.iterations += s2.iterations;
        res.rel_residual = st[q].r0 > 0 ? s2.rfinal / st[q].r0 : 0.0;
        res.escalated = 1;
      }
      if ((s1 == CD3D_OK || s1 == CD3D_NOT_CONVERGED) && r.want_solution) {
        ds[k].sol.resize(res.n_dof);
        const cd3d_status g = cd3d_get_solution(h, ds[k].sol.data());
        if (g != CD3D_OK) s1 = g;
      }
      res.status = s1;
    }
    for (auto &d : ds) d.res.batch_size = n;
  }

  void work_loop() {
    if (!prm.simulate) (void)hipSetDevice(device);
    std::unique_lock<std::mutex> lk(mu);
    while (true) {
      cv_work.wait(lk, [&] { return stop || !ready.empty(); });
      if (ready.empty()) return;  // stop, and nothing admitted is left
      std::vector<Work> ws{ready.front()};
      ready.pop_front();
      if (batchable(ws[0].p.req))
        for (auto it = ready.begin(); it != ready.end() && (int)ws.size() < prm.max_batch;) {
          if (batchable(it->p.req) && same_batch(ws[0].p.req, it->p.req)) {
            ws.push_back(*it);
            it = ready.erase(it);
          } else {
            ++it;
          }
        }
      lk.unlock();
      std::vector<Done> ds(ws.size());
      if (ws.size() == 1) {
        run_request(ws[0], ds[0]);
        ds[0].res.batch_size = 1;
      }
      else run_batch(ws, ds);
      const double t_end = now_ms();
      lk.lock();
      for (size_t k = 0; k < ws.size(); k++) finish(ws[k], ds[k], t_end, lk);
    }
  }

  // bookkeeping of one finished request (called with mu held; may drop it
  // to destroy a hierarchy)
  void finish(Work &w, Done &d, double t_end, std::unique_lock<std::mutex> &lk) {
    {
      cd3d_problem *destroy = nullptr;
      Resident *slot = w.slot;
      size_t actual = slot->h ? cd3d_device_bytes(slot->h) : (prm.simulate ? slot->bytes : 0);
      bytes = bytes - slot->bytes + actual;
      peak = std::max(peak, bytes);
      slot->bytes = actual;
      slot->busy = false;
      slot->last_used = t_end;
      const bool bad = d.res.status != CD3D_OK && d.res.status != CD3D_NOT_CONVERGED;
      if (!prm.keep_resident || (!slot->h && !prm.simulate) || bad) {  // a failed run leaves no hierarchy behind
        bytes -= slot->bytes;
        destroy = slot->h;
        for (auto it = resident.begin(); it != resident.end(); ++it)
          if (&*it == slot) {
            resident.erase(it);
            break;
          }
      }
      d.res.id = w.p.id;
      d.res.queue_ms = w.t_admit - w.p.t_submit;
      d.res.run_ms = t_end - w.t_admit;
      d.res.latency_ms = t_end - w.p.t_submit;
      d.res.device_bytes = actual;
      d.res.reused_hierarchy = w.reused ? 1 : 0;
      if (d.res.status == CD3D_OK) {
        completed++;
        latencies.push_back(d.res.latency_ms);
        t_last = std::max(t_last, t_end);
      } else {
        failed++;
      }
      inflight--;
      done[w.p.id] = std::move(d);
      wake = true;
      cv_sched.notify_all();
      cv_done.notify_all();
      if (destroy) {
        lk.unlock();
        cd3d_destroy(destroy);
        lk.lock();
      }
    }
  }
};

extern "C" {

void cd3d_server_default_params(cd3d_server_params *p) {
  if (!p) return;
  memset(p, 0, sizeof(*p));
  p->hbm_budget = 0;
  p->max_inflight = 8;
  p->workers = 0;
  p->keep_resident = 1;
  p->max_bypass = 16;
  p->simulate = 0;
  p->sim_us_per_dof = 0.01;
  p->robust_after = 200;
  p->max_batch = 8;
}

cd3d_status cd3d_server_create(const cd3d_server_params *prm, cd3d_server **out) {
  if (!prm || !out || prm->max_inflight < 1 || prm->workers < 0 || prm->max_bypass < 0 || prm->max_batch < 1)
    return CD3D_ERR_ARG;
  *out = nullptr;
  cd3d_server *s = new (std::nothrow) cd3d_server();
  if (!s) return CD3D_ERR_NOMEM;
  s->prm = *prm;
  if (s->prm.workers == 0) s->prm.workers = s->prm.max_inflight;
  if (!s->prm.simulate) {
    if (hipGetDevice(&s->device) != hipSuccess) {
      delete s;
      return CD3D_ERR_HIP;
    }
    if (s->prm.hbm_budget == 0) {
      size_t free_b = 0, total_b = 0;
      if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) {
        delete s;
        return CD3D_ERR_HIP;
      }
      s->prm.hbm_budget = (size_t)(0.85 * (double)free_b);
    }
  } else if (s->prm.hbm_budget == 0) {
    delete s;
    return CD3D_ERR_ARG;  // a simulated server needs an explicit budget
  }
  s->scheduler = std::thread([s] { s->schedule_loop(); });
  for (int i = 0; i < s->prm.workers; i++) s->workers.emplace_back([s] { s->work_loop(); });
  *out = s;
  return CD3D_OK;
}

cd3d_status cd3d_server_submit(cd3d_server *s, const cd3d_request *req, uint64_t *id) {
  if (!s || !req || !id) return CD3D_ERR_ARG;
  if (req->kind != CD3D_REQ_CD && req->kind != CD3D_REQ_TCD) return CD3D_ERR_ARG;
  if (req->kind == CD3D_REQ_TCD && (req->nsteps < 0 || !(req->tau > 0) || !(req->theta >= 0.5 && req->theta <= 1)))
    return CD3D_ERR_ARG;
  Pending p;
  p.req = *req;
  p.est = s->estimate(*req);
  if (p.est == 0) return CD3D_ERR_ARG;
  std::lock_guard<std::mutex> g(s->mu);
  if (p.est > s->prm.hbm_budget) {
    s->rejected++;
    return CD3D_ERR_NOMEM;
  }
  p.id = s->next_id++;
  p.t_submit = now_ms();
  if (s->t_first < 0) s->t_first = p.t_submit;
  s->queue.push_back(p);
  s->outstanding[p.id] = 1;
  s->submitted++;
  *id = p.id;
  s->wake = true;
  s->cv_sched.notify_all();
  return CD3D_OK;
}

cd3d_status cd3d_server_wait(cd3d_server *s, uint64_t id, double timeout_ms, cd3d_result *res, double *sol,
                             int64_t sol_len) {
  if (!s) return CD3D_ERR_ARG;
  std::unique_lock<std::mutex> lk(s->mu);
  if (!s->outstanding.count(id)) return CD3D_ERR_ARG;
  auto ready = [&] { return s->done.count(id) > 0; };
  if (timeout_ms < 0) s->cv_done.wait(lk, ready);
  else if (!s->cv_done.wait_for(lk, std::chrono::duration<double, std::milli>(timeout_ms), ready))
    return CD3D_ERR_STATE;
  Done &d = s->done[id];
  if (res) *res = d.res;
  if (sol && !d.sol.empty()) {
    if (sol_len < (int64_t)d.sol.size()) return CD3D_ERR_ARG;  // the result stays for a second call
    memcpy(sol, d.sol.data(), sizeof(double) * d.sol.size());
  }
  const cd3d_status st = (cd3d_status)d.res.status;
  s->done.erase(id);
  s->outstanding.erase(id);
  return st;
}

cd3d_status cd3d_server_drain(cd3d_server *s) {
  if (!s) return CD3D_ERR_ARG;
  std::unique_lock<std::mutex> lk(s->mu);
  s->cv_done.wait(lk, [&] { return s->queue.empty() && s->ready.empty() && s->inflight == 0; });
  return CD3D_OK;
}

cd3d_status cd3d_server_get_stats(cd3d_server *s, cd3d_server_stats *st) {
  if (!s || !st) return CD3D_ERR_ARG;
  std::lock_guard<std::mutex> g(s->mu);
  memset(st, 0, sizeof(*st));
  st->submitted = s->submitted;
  st->completed = s->completed;
  st->failed = s->failed;
  st->rejected = s->rejected;
  if (!s->latencies.empty()) {
    std::vector<double> v = s->latencies;
    std::sort(v.begin(), v.end());
    auto pct = [&](double q) {
      const double pos = q * (double)(v.size() - 1);
      const size_t i = (size_t)pos;
      const double f = pos - (double)i;
      return i + 1 < v.size() ? v[i] * (1 - f) + v[i + 1] * f : v[i];
    };
    st->p50_ms = pct(0.50);
    st->p99_ms = pct(0.99);
    double sum = 0;
    for (double x : v) sum += x;
    st->mean_ms = sum / (double)v.size();
    const double span = s->t_last - s->t_first;
    st->solves_per_s = span > 0 ? 1e3 * (double)s->completed / span : 0.0;
  }
  st->bytes_in_use = s->bytes;
  st->peak_bytes = s->peak;
  st->budget = s->prm.hbm_budget;
  st->inflight = s->inflight;
  st->queued = (int)s->queue.size();
  st->resident = (int)s->resident.size();
  return CD3D_OK;
}

cd3d_status cd3d_server_destroy(cd3d_server *s) {
  if (!s) return CD3D_ERR_ARG;
  {
    std::unique_lock<std::mutex> lk(s->mu);
    s->queue.clear();  // queued requests are dropped; admitted ones finish
    s->cv_done.wait(lk, [&] { return s->ready.empty() && s->inflight == 0; });
    s->stop = true;
    s->cv_sched.notify_all();
    s->cv_work.notify_all();
  }
  s->scheduler.join();
  for (auto &t : s->workers) t.join();
  for (auto &r : s->resident)
    if (r.h) cd3d_destroy(r.h);
  delete s;
  return CD3D_OK;
}

}  // extern "C"
