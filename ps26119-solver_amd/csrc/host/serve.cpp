// serve.cpp — continuous batching of LP solves on one GPU (see serve.h).
#include "serve.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>

#include "engine.h"
#include "la.h"
#include "ps26119_hip.h"
#include "scaling.h"
#include "solve.h"
#include "verify.h"

namespace ps26119::amd {

struct Server::Request {
  std::int64_t id = 0;
  Model model;
  Options options;
  EngineOptions eo;
  ServedResult result;
  bool done = false;
  // prepared on a host worker
  std::unique_ptr<ScaledProblem> sp;
  std::vector<double> cmin, x_init;
  psh_problem prob{};
  double eta = 1.0;
  std::size_t bytes = 0;
  bool mixed = false;  // increment form (precision = mixed, hip_refine = 1)
  RefineGuard guard;  // stall guard of the increment form
  // live on the device
  int slot = -1;
  PidState pid;
  std::int64_t it = 0, inner = 0, fast_it = -1;
  double fast_s = -1, r0 = -1, r_last = kInf;
  // finishing
  Status status = Status::NotSolved;
  std::string message;
  std::vector<double> xs, ys;
  std::vector<double> dual_ray, primal_ray;  // certificate of an Infeasible / Unbounded verdict
};

namespace {

double clock_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

KktStats kkt_from(const psh_kkt& d) {
  KktStats k;
  k.primal_residual = d.primal_residual;
  k.dual_residual = d.dual_residual;
  k.primal_obj = d.primal_obj;
  k.dual_obj = d.dual_obj;
  k.b_norm = d.b_norm;
  k.c_norm = d.c_norm;
  k.primal_max_rel = d.primal_max_rel;
  k.dual_max_rel = d.dual_max_rel;
  return k;
}

}  // namespace

Server::Server(const ServerOptions& o) : opt_(o), t0_(clock_s()) {
  if (opt_.check_every < 1) opt_.check_every = 64;
  const int rc = psh_server_create(opt_.device, opt_.budget_bytes, std::max(1, opt_.max_instances), &dev_);
  if (rc != PSH_OK) {
    error_ = rc == PSH_E_NO_DEVICE ? "HIP backend unavailable: no usable MI355X/gfx950 device is visible"
             : rc == PSH_E_OOM    ? "out of memory: the HBM budget could not be allocated"
                                  : "serving arena could not be created";
    dev_ = nullptr;
    return;
  }
  psh_server_graphs(dev_, opt_.graphs ? 1 : 0);
  psh_server_timing(dev_, opt_.timing ? 1 : 0);
  capacity_ = psh_server_free_bytes(dev_);
  for (int k = 0; k < std::max(1, opt_.setup_threads); ++k) workers_.emplace_back([this] { setup_worker(); });
  loop_ = std::thread([this] { serve_loop(); });
}

Server::~Server() {
  {
    std::lock_guard<std::mutex> g(mu_);
    stop_ = true;
  }
  cv_work_.notify_all();
  cv_serve_.notify_all();
  if (loop_.joinable()) loop_.join();
  for (auto& t : workers_) t.join();
  if (dev_) psh_server_destroy(dev_);
}

double Server::now() const { return clock_s() - t0_; }

bool Server::kernel_time(long long& iterations, double& total_ms, double& bytes) {
  if (!dev_ || !opt_.timing) return false;
  std::lock_guard<std::mutex> g(dev_mu_);  // between two rounds of the serving thread
  return psh_server_timing_get(dev_, &iterations, &total_ms, &bytes) == PSH_OK;
}

int Server::in_flight() const {
  std::lock_guard<std::mutex> g(mu_);
  int n = 0;
  for (const auto& kv : all_) n += kv.second->done ? 0 : 1;
  return n;
}

std::int64_t Server::submit(Model model, const Options& options) {
  auto r = std::make_shared<Request>();
  r->model = std::move(model);
  r->options = options;
  r->result.submit_time = now();
  std::lock_guard<std::mutex> g(mu_);
  r->id = next_id_++;
  all_[r->id] = r;
  to_prepare_.push_back(r);
  cv_work_.notify_one();
  return r->id;
}

bool Server::wait(std::int64_t id, ServedResult& out, double timeout_s) {
  std::unique_lock<std::mutex> lk(mu_);
  auto it = all_.find(id);
  if (it == all_.end()) return false;
  auto r = it->second;
  auto pred = [&] { return r->done; };
  if (timeout_s < 0) cv_done_.wait(lk, pred);
  else if (!cv_done_.wait_for(lk, std::chrono::duration<double>(timeout_s), pred)) return false;
  out = r->result;
  all_.erase(id);
  return true;
}

void Server::complete(const std::shared_ptr<Request>& r) {
  {
    std::lock_guard<std::mutex> g(mu_);
    r->result.done_time = now();
    r->result.solution.seconds = r->result.done_time - r->result.submit_time;
    r->done = true;
    r->sp.reset();
    r->xs.clear();
    r->ys.clear();
  }
  cv_done_.notify_all();
}

// Host side of admission: validation, scaling and ‖Ã‖₂ (reference EngineContext ctor,
// engine.cpp:195-235), without presolve.
void Server::prepare(Request& r) {
  Solution& s = r.result.solution;
  s.engine = "r2hpdhg-served";
  s.precision = r.options.precision == Precision::Mixed ? "mixed" : "fp64";
  s.model_fingerprint = r.model.fingerprint_hex();
  auto reject = [&](Status st, const std::string& why) {
    s.status = st;
    s.message = why;
    r.status = st;
  };
  if (auto err = r.model.validate(); !err.empty()) return reject(Status::NotSolved, "invalid model: " + err);
  if (!r.model.is_integer.empty() &&
      std::any_of(r.model.is_integer.begin(), r.model.is_integer.end(), [](std::uint8_t v) { return v != 0; }) &&
      !r.options.relax_integrality)
    return reject(Status::NotSolved, "integer columns: the server solves LPs");
  if (auto why = r.model.crossed_bounds(); !why.empty()) {
    reject(Status::Infeasible, "bounds cross: " + why);
    s.check = "PASS";
    return;
  }
  r.eo = engine_options_from(r.options);
  if (auto bad = apply_engine_params(r.options, r.eo); !bad.empty())
    return reject(Status::NotSolved, "unknown engine parameter '" + bad + "'");
  if (r.eo.max_epoch_gap_only == 2)
    return reject(Status::NotSolved, "hip_max_epoch_gap_only = 2 is implemented for single solves only");
  r.sp = std::make_unique<ScaledProblem>(make_scaled_problem(r.model, r.eo.scaling));
  const ScaledProblem& sp = *r.sp;
  if (sp.A.nnz() >= (1LL << 31)) return reject(Status::NotSolved, "matrix beyond the 32-bit device index range");
  const bool empty = sp.n == 0 || sp.A.nnz() == 0;
  const double norm = empty ? 0.0 : estimate_norm2_host(sp);
  r.eta = norm > 0 ? 0.998 / norm : 1.0;
  r.prob = problem_of(sp, r.cmin);
  r.mixed = r.options.precision == Precision::Mixed && r.eo.refine != 0;
  if (r.mixed) r.message = "fp32 increments on fp64 anchors";
  r.bytes = psh_server_bytes(sp.m, sp.n, sp.A.nnz(), r.mixed ? PSH_SERVE_INCREMENT : 0);
  if (r.bytes > capacity_) return reject(Status::NotSolved, "out of memory: the model exceeds the HBM budget");
  r.x_init.resize(sp.n);  // x = proj_X(0), y = 0 (reference r2hpdhg.cpp:31-34)
  for (int j = 0; j < sp.n; ++j) r.x_init[j] = std::min(std::max(0.0, sp.col_lower[j]), sp.col_upper[j]);
  r.pid.omega = r.pid.best_omega = 1.0;
  s.setup_seconds = now() - r.result.submit_time;
  r.status = Status::NotSolved;
}

// Host side of retirement: unscale, KKT on the original model, certified bound and the gate
// (reference EngineContext::finish, engine.cpp:268-307; solve.cpp:119-185).
void Server::finish(Request& r) {
  Solution& sol = r.result.solution;
  const ScaledProblem& sp = *r.sp;
  const Model& M = r.model;
  std::vector<double> y_min;
  sp.unscale_primal(r.xs, sol.x);
  sp.unscale_dual(r.ys, y_min);
  const KktStats k = kkt_on_original(sp, sol.x, y_min);
  Status status = r.status;
  if (!k.finite()) status = Status::NumericalError;
  if (status == Status::Optimal && !k.converged(r.eo.tolerance)) status = Status::NumericalError;
  sol.y.resize(sp.m);
  for (int i = 0; i < sp.m; ++i) sol.y[i] = M.sense * y_min[i];
  sol.z = M.obj;
  for (int j = 0; j < sp.n; ++j)
    for (int p = M.col_start[j]; p < M.col_start[j + 1]; ++p) sol.z[j] -= M.value[p] * sol.y[M.row_index[p]];
  sol.row_activity = activity(M, sol.x);
  sol.objective = M.objective_value(sol.x);
  sol.dual_objective = M.sense * k.dual_obj + M.obj_offset;
  sol.primal_residual = k.rel_primal();
  sol.dual_residual = k.rel_dual();
  sol.gap = k.rel_gap();
  sol.status = status;
  sol.message = r.message;
  if (status == Status::Infeasible) sol.dual_ray = r.dual_ray;
  if (status == Status::Unbounded) sol.primal_ray = r.primal_ray;
  sol.iterations = r.it;
  sol.iterations_to_fast = r.fast_it;
  sol.seconds_to_fast = r.fast_s;
  sol.primal_weight = r.pid.omega;
  if (M.num_rows + M.num_cols > 0) sol.certified_bound = certified_dual_bound(M, sol.y, true).bound;
  gate(M, r.options, sol);
}

void Server::setup_worker() {
  for (;;) {
    std::shared_ptr<Request> r;
    bool is_finish = false;
    {
      std::unique_lock<std::mutex> lk(mu_);
      cv_work_.wait(lk, [&] { return stop_ || !to_finish_.empty() || !to_prepare_.empty(); });
      if (stop_) return;  // destruction abandons the requests still in flight
      if (!to_finish_.empty()) {  // finishing first: it frees a caller
        r = to_finish_.front();
        to_finish_.pop_front();
        is_finish = true;
      } else if (!to_prepare_.empty()) {
        r = to_prepare_.front();
        to_prepare_.pop_front();
      } else {
        return;  // stop_ with nothing left
      }
    }
    try {
      if (is_finish) {
        finish(*r);
        complete(r);
        continue;
      }
      prepare(*r);
    } catch (const std::bad_alloc&) {
      r->status = r->result.solution.status = Status::NotSolved;
      r->result.solution.message = "out of memory";
    } catch (const std::exception& e) {
      r->status = r->result.solution.status = Status::NumericalError;
      r->result.solution.message = e.what();
    }
    if (is_finish) {
      complete(r);
      continue;
    }
    if (!r->sp || !r->result.solution.message.empty()) {  // rejected during preparation
      complete(r);
      continue;
    }
    {
      std::lock_guard<std::mutex> g(mu_);
      ready_.push_back(r);
    }
    cv_serve_.notify_one();
  }
}

void Server::serve_loop() {
  std::vector<std::shared_ptr<Request>> live;
  std::vector<psh_serve_param> params;
  std::vector<int> ids, restart_ids;
  std::vector<psh_stats> stats;
  const int K = opt_.check_every;
  // A ray verdict stands only when the ray passes the certificate check on the request's own
  // model (the finishing gate checks it again).
  auto certify = [&](Request& r, Status v, double robj, double rviol) {
    std::vector<double> dr, pr;
    if (v == Status::Infeasible) dr.assign(r.sp->m, 0.0);
    else pr.assign(r.sp->n, 0.0);
    if (psh_server_ray_vectors(dev_, r.slot, dr.empty() ? nullptr : dr.data(), pr.empty() ? nullptr : pr.data()) != PSH_OK)
      return false;
    char buf[200];
    bool ok;
    if (v == Status::Infeasible) {
      ok = check_infeasibility_certificate(r.model, dr).passed;
      std::snprintf(buf, sizeof buf, "primal infeasible: dual ray certificate (objective %f, violation %f)", robj, rviol);
    } else {
      std::vector<double> xs(r.sp->n), x;
      if (psh_server_download(dev_, r.slot, xs.data(), nullptr) != PSH_OK) return false;
      r.sp->unscale_primal(xs, x);
      ok = check_unboundedness_certificate(r.model, x, pr).passed;
      std::snprintf(buf, sizeof buf, "unbounded: primal ray certificate (c'd %f, violation %f)", robj, rviol);
    }
    if (!ok) return false;
    r.message = r.message.empty() ? std::string(buf) : std::string(buf) + "; " + r.message;
    r.dual_ray = std::move(dr);
    r.primal_ray = std::move(pr);
    return true;
  };
  // A device error ends every live request; their slots and arena space are released so that
  // later requests can still be admitted.
  auto fail_live = [&](const std::string& why) {
    for (auto& r : live) {
      (void)psh_server_remove(dev_, r->slot);
      r->status = Status::NumericalError;
      r->result.solution.status = Status::NumericalError;
      r->result.solution.message = why;
      {
        std::lock_guard<std::mutex> g(mu_);
        --live_;
      }
      complete(r);
    }
    live.clear();
  };
  for (;;) {
    std::unique_lock<std::mutex> round;  // this round's device calls (released at the end of the round)
    // admission: prepared requests in arrival order while the budget has room
    std::vector<std::shared_ptr<Request>> admit;
    {
      std::unique_lock<std::mutex> lk(mu_);
      if (live.empty()) cv_serve_.wait(lk, [&] { return stop_ || !ready_.empty(); });
      if (stop_) break;  // destruction abandons the requests still in flight
      round = std::unique_lock<std::mutex>(dev_mu_);
      std::size_t free_bytes = psh_server_free_bytes(dev_);
      while (!ready_.empty() && static_cast<int>(live.size() + admit.size()) < opt_.max_instances) {
        auto& r = ready_.front();
        if (r->bytes > free_bytes) break;
        free_bytes -= r->bytes;
        admit.push_back(r);
        ready_.pop_front();
      }
    }
    for (auto& r : admit) {
      int id = -1;
      int rc = psh_server_add(dev_, &r->prob, r->mixed ? PSH_SERVE_INCREMENT : 0, &id);
      if (rc == PSH_OK) {
        std::vector<double> y0(r->sp->m, 0.0);
        rc = psh_server_set_iterate(dev_, id, r->x_init.data(), y0.data());
      }
      if (rc != PSH_OK) {
        if (rc == PSH_E_OOM && !live.empty()) {  // fragmentation: back to the queue head until a retirement
          std::lock_guard<std::mutex> g(mu_);
          ready_.push_front(r);
          continue;
        }
        r->status = r->result.solution.status = Status::NumericalError;
        r->result.solution.message = psh_server_error(dev_);
        complete(r);
        continue;
      }
      r->slot = id;
      r->result.admit_time = now();
      live.push_back(r);
      std::lock_guard<std::mutex> g(mu_);
      ++live_;
    }
    if (live.empty()) continue;

    // one chunk for every live instance
    params.clear();
    ids.clear();
    for (auto& r : live) {
      params.push_back(psh_serve_param{r->slot, r->eta / r->pid.omega, r->eta * r->pid.omega, r->inner, r->eo.reflection});
      ids.push_back(r->slot);
    }
    int rc = psh_server_chunk(dev_, K, params.data(), static_cast<int>(params.size()));
    bool want_ray = false;  // ray tests from an instance's fourth check on (reference r2hpdhg.cpp:93-111)
    for (auto& r : live) want_ray = want_ray || r->it + K >= 4 * K;
    if (rc == PSH_OK) {
      stats.assign(live.size(), psh_stats{});
      rc = psh_server_check(dev_, ids.data(), static_cast<int>(ids.size()), want_ray ? 1 : 0, stats.data());
    }
    if (rc != PSH_OK) {
      fail_live(std::string("serving device error: ") + psh_server_error(dev_));
      continue;
    }
    // per-instance decisions (the single-LP engine's rules)
    std::vector<std::shared_ptr<Request>> keep, retire;
    std::vector<int> to_fp64;
    restart_ids.clear();
    const double t_now = now();
    for (std::size_t q = 0; q < live.size(); ++q) {
      auto& r = live[q];
      const psh_stats& st = stats[q];
      const EngineOptions& o = r->eo;
      r->it += K;
      r->inner += K;
      const double om = r->pid.omega, eta = r->eta;
      auto resid = [&](double px, double py, double cr) {
        return std::sqrt(std::max(om * px + py / om + 2.0 * eta * cr, 0.0));
      };
      if (st.have_r0) r->r0 = resid(st.r0_px, st.r0_py, st.r0_cross);
      const double rr = resid(st.px, st.py, st.cross);
      const KktStats k = kkt_from(st.kkt);
      Status verdict = Status::NotSolved;
      if (!k.finite() || !std::isfinite(rr) || !st.finite) {
        verdict = Status::NumericalError;
        r->message = "non-finite iterate";
      } else {
        if (r->fast_it < 0 && k.rel_kkt() <= tol::kFirstOrderFast) r->fast_it = r->it, r->fast_s = t_now - r->result.submit_time;
        if (k.converged(o.tolerance)) verdict = Status::Optimal;
        else if (r->it >= o.iteration_limit) verdict = Status::IterationLimit;
        else if (t_now - r->result.submit_time > o.time_limit) verdict = Status::TimeLimit;
      }
      if (o.verbosity >= 2)
        std::fprintf(stderr, "served %lld it %8lld  rp %.2e  rd %.2e  gap %.2e\n", static_cast<long long>(r->id),
                     static_cast<long long>(r->it), k.rel_primal(), k.rel_dual(), k.rel_gap());
      if (verdict == Status::NotSolved && want_ray && r->it >= 4 * K) {
        double robj = 0, rviol = 0;
        const Status v = ray_decision(st.ray, robj, rviol);
        if (v != Status::NotSolved && certify(*r, v, robj, rviol)) verdict = v;
      }
      if (verdict != Status::NotSolved) {
        r->status = verdict;
        retire.push_back(r);
        continue;
      }
      keep.push_back(r);
      bool promoted = false;  // increment form stalled: fp64 from the next chunk, after a restart
      if (r->mixed) {
        const double rel = k.rel_kkt();
        if (r->guard.stalled(rel, r->it, o)) {
          promoted = true;
          r->mixed = false;
          r->message += "; fp32 -> fp64 at iteration " + std::to_string(r->it);
          to_fp64.push_back(r->slot);
        }
      }
      const bool restart =
          promoted || restart_due(o, rr, r->r0, r->r_last, r->inner, r->it) || increment_epoch_full(o, r->mixed, r->inner) ||
          epoch_cap_reached(o, r->inner, k);
      r->r_last = rr;
      if (!restart) continue;
      pid_update(r->pid, o, st.dx_norm, st.dy_norm, k.rel_primal(), k.rel_dual());
      restart_ids.push_back(r->slot);
      r->inner = 0;
      r->r_last = kInf;
    }
    rc = psh_server_restart(dev_, restart_ids.data(), static_cast<int>(restart_ids.size()));
    for (int id : to_fp64)
      if (rc == PSH_OK) rc = psh_server_set_precision(dev_, id, 0);
    for (auto& r : retire) {
      if (rc == PSH_OK) {
        r->xs.resize(r->sp->n);
        r->ys.resize(r->sp->m);
        rc = psh_server_download(dev_, r->slot, r->xs.data(), r->ys.data());
      }
      if (rc == PSH_OK) rc = psh_server_remove(dev_, r->slot);
      {
        std::lock_guard<std::mutex> g(mu_);
        --live_;
        to_finish_.push_back(r);
      }
      cv_work_.notify_one();
    }
    live.swap(keep);
    if (rc != PSH_OK) fail_live(std::string("serving device error: ") + psh_server_error(dev_));
  }
}

}  // namespace ps26119::amd
