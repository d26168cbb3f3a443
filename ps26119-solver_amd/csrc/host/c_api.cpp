// c_api.cpp — the C API (include/ps26119/ps26119.h, ps26119_amd.h). Same validation, option
// translation and exception firewall as the reference's C API (reference src/core/c_api.cpp:29-202).
#include <cmath>
#include <cstdio>
#include <cstring>
#include <exception>
#include <map>
#include <memory>
#include <mutex>

#include "batch.h"
#include "engine.h"
#include "presolve.h"
#include "ps26119/ps26119.h"
#include "ps26119/ps26119_amd.h"
#include "ps26119_hip.h"
#include "scaling.h"
#include "serve.h"
#include "solve.h"
#include "verify.h"

using namespace ps26119;

namespace {

constexpr const char* kVersion = "0.1.0-mi355x";

void set_text(char* dst, std::size_t cap, const std::string& s) { std::snprintf(dst, cap, "%s", s.c_str()); }

void clear(ps26119_result* r) {
  std::memset(r, 0, sizeof *r);
  r->status = PS26119_NOT_SOLVED;
  r->check = -1;
}

const char* to_options(const ps26119_options* opt, int num_rows, int num_cols, Options& o) {
  if (!opt) return nullptr;
  switch (opt->algorithm) {
    case PS26119_ALG_ORACLE: o.algorithm = Algorithm::Oracle; break;
    case PS26119_ALG_PDLP: o.algorithm = Algorithm::Pdlp; break;
    case PS26119_ALG_R2HPDHG: o.algorithm = Algorithm::R2hpdhg; break;
    case PS26119_ALG_SIMPLEX: o.algorithm = Algorithm::Simplex; break;
    case PS26119_ALG_AUTO: o.algorithm = Algorithm::Auto; break;
    default: return "unknown algorithm";
  }
  if (opt->precision != PS26119_PREC_FP64 && opt->precision != PS26119_PREC_MIXED) return "unknown precision";
  o.precision = opt->precision == PS26119_PREC_MIXED ? Precision::Mixed : Precision::Fp64;
  o.use_gpu = opt->use_gpu != 0;
  if (!(opt->tolerance >= 0) || !std::isfinite(opt->tolerance)) return "tolerance must be a finite number >= 0";
  if (!(opt->time_limit >= 0)) return "time_limit must be >= 0 (0 = default)";
  if (opt->iteration_limit < 0) return "iteration_limit must be >= 0 (0 = default)";
  if (opt->threads < 0) return "threads must be >= 0 (0 = all cores)";
  if (opt->tolerance > 0) o.tolerance = opt->tolerance;
  if (opt->time_limit > 0) o.time_limit = opt->time_limit;
  if (opt->iteration_limit > 0) o.iteration_limit = opt->iteration_limit;
  o.verbosity = opt->verbosity;
  o.threads = opt->threads;
  if (opt->warm_x) o.warm_x.assign(opt->warm_x, opt->warm_x + num_cols);
  if (opt->warm_y) o.warm_y.assign(opt->warm_y, opt->warm_y + num_rows);
  return nullptr;
}

void fill(const Solution& s, ps26119_result* r) {
  r->status = static_cast<int>(s.status);
  r->objective = s.objective;
  r->dual_objective = s.dual_objective;
  r->primal_residual = s.primal_residual;
  r->dual_residual = s.dual_residual;
  r->gap = s.gap;
  r->certified_bound = s.certified_bound;
  r->iterations = s.iterations;
  r->seconds = s.seconds;
  r->check = s.check == "PASS" ? 1 : s.check == "FAIL" ? 0 : -1;
  set_text(r->engine, sizeof r->engine, s.engine);
  set_text(r->message, sizeof r->message, s.message);
}

// Builds the Model from C arrays; returns an error text or nullptr.
const char* make_model(int m, int n, int sense, double offset, const double* c, const double* cl, const double* cu,
                       const double* rl, const double* ru, const int* cs, const int* ri, const double* v, Model& M,
                       std::string& err) {
  if (m < 0 || n < 0) return "negative dimension";
  if (!cs) return "col_start is NULL";
  if (n > 0 && (!c || !cl || !cu)) return "column data is NULL";
  if (m > 0 && (!rl || !ru)) return "row bounds are NULL";
  const int nnz = cs[n];
  if (nnz < 0) return "col_start[n] < 0";
  if (nnz > 0 && (!ri || !v)) return "matrix arrays are NULL";
  M.num_rows = m;
  M.num_cols = n;
  M.sense = sense;
  M.obj_offset = offset;
  M.obj.assign(c, c + n);
  M.col_lower.assign(cl, cl + n);
  M.col_upper.assign(cu, cu + n);
  M.row_lower.assign(rl, rl + m);
  M.row_upper.assign(ru, ru + m);
  M.col_start.assign(cs, cs + n + 1);
  M.row_index.assign(ri, ri + nnz);
  M.value.assign(v, v + nnz);
  err = M.validate();
  if (!err.empty()) {
    err = "invalid model: " + err;
    return err.c_str();
  }
  return nullptr;
}

// The extension's named parameters into Options: "warm_primal_weight" is the reference's
// Options::warm_primal_weight (the first primal weight of a warm-started solve or batch; the C
// struct has no field for it), every other name goes to the engine, which checks it.
const char* take_params(const ps26119_amd_options* ext, Options& o) {
  for (int k = 0; k < ext->n_params; ++k) {
    if (!ext->param_names || !ext->param_values || !ext->param_names[k]) return "engine parameter list is NULL";
    const double v = ext->param_values[k];
    if (std::strcmp(ext->param_names[k], "warm_primal_weight") == 0) {
      if (!(v >= 0 && std::isfinite(v))) return "warm_primal_weight must be finite and non-negative";
      o.warm_primal_weight = v;
    } else {
      o.engine_params.emplace_back(ext->param_names[k], v);
    }
  }
  return nullptr;
}

template <class F>
int guarded(ps26119_result* result, F&& f) {
  try {
    return f();
  } catch (const std::exception& e) {
    result->status = PS26119_NUMERICAL_ERROR;
    set_text(result->message, sizeof result->message, e.what());
    return result->status;
  } catch (...) {
    result->status = PS26119_NUMERICAL_ERROR;
    set_text(result->message, sizeof result->message, "unknown error");
    return result->status;
  }
}

int solve_common(int num_rows, int num_cols, int sense, double obj_offset, const double* c, const double* col_lower,
                 const double* col_upper, const double* row_lower, const double* row_upper, const int* col_start,
                 const int* row_index, const double* value, const ps26119_options* opt, const ps26119_amd_options* ext,
                 ps26119_result* result, ps26119_amd_stats* stats, double* x, double* y, double* z, double* dual_ray,
                 double* primal_ray) {
  if (!result) return PS26119_INVALID_ARGUMENT;
  clear(result);
  if (stats) std::memset(stats, 0, sizeof *stats);
  auto invalid = [&](const char* why) {
    result->status = PS26119_INVALID_ARGUMENT;
    set_text(result->message, sizeof result->message, why);
    return result->status;
  };
  return guarded(result, [&]() -> int {
    Model m;
    std::string err;
    if (const char* bad = make_model(num_rows, num_cols, sense, obj_offset, c, col_lower, col_upper, row_lower, row_upper,
                                     col_start, row_index, value, m, err))
      return invalid(bad);
    Options o;
    if (const char* bad = to_options(opt, num_rows, num_cols, o)) return invalid(bad);
    if (ext) {
      o.presolve = ext->presolve != 0;
      if (ext->check_every > 0) o.termination_check_every = ext->check_every;
      if (const char* bad = take_params(ext, o)) return invalid(bad);
    }
    const Solution s = amd::solve(m, o);
    fill(s, result);
    auto put = [](double* dst, const std::vector<double>& v, int len) {
      if (dst && static_cast<int>(v.size()) == len && len > 0) std::memcpy(dst, v.data(), sizeof(double) * len);
    };
    put(x, s.x, num_cols);
    put(y, s.y, num_rows);
    put(z, s.z, num_cols);
    put(dual_ray, s.dual_ray, num_rows);
    put(primal_ray, s.primal_ray, num_cols);
    if (stats) {
      stats->setup_seconds = s.setup_seconds;
      stats->iterations_to_fast = s.iterations_to_fast;
      stats->seconds_to_fast = s.seconds_to_fast;
      stats->primal_weight = s.primal_weight;
      stats->check_primal = s.check_primal;
      stats->check_dual = s.check_dual;
      stats->check_gap = s.check_gap;
      for (int k = 0; k < 3; ++k) stats->kernel_launches[k] = s.kernel_launches[k], stats->kernel_ms[k] = s.kernel_ms[k];
      stats->device_bytes = s.device_bytes;
      stats->iterate_seconds = s.iterate_seconds;
    }
    return result->status;
  });
}

}  // namespace

extern "C" {

const char* ps26119_version(void) { return kVersion; }

void ps26119_default_options(ps26119_options* opt) {
  if (!opt) return;
  const Options d;
  opt->algorithm = PS26119_ALG_AUTO;
  opt->precision = PS26119_PREC_FP64;
  opt->use_gpu = 0;
  opt->tolerance = d.tolerance;
  opt->time_limit = d.time_limit;
  opt->iteration_limit = d.iteration_limit;
  opt->verbosity = 0;
  opt->threads = 1;
  opt->warm_x = nullptr;
  opt->warm_y = nullptr;
}

int ps26119_solve_lp(int num_rows, int num_cols, int sense, double obj_offset, const double* c, const double* col_lower,
                     const double* col_upper, const double* row_lower, const double* row_upper, const int* col_start,
                     const int* row_index, const double* value, const ps26119_options* opt, ps26119_result* result,
                     double* x, double* y, double* z) {
  return solve_common(num_rows, num_cols, sense, obj_offset, c, col_lower, col_upper, row_lower, row_upper, col_start,
                      row_index, value, opt, nullptr, result, nullptr, x, y, z, nullptr, nullptr);
}

int ps26119_solve_lp_ex(int num_rows, int num_cols, int sense, double obj_offset, const double* c,
                        const double* col_lower, const double* col_upper, const double* row_lower,
                        const double* row_upper, const int* col_start, const int* row_index, const double* value,
                        const ps26119_options* opt, ps26119_result* result, double* x, double* y, double* z,
                        double* dual_ray, double* primal_ray) {
  return solve_common(num_rows, num_cols, sense, obj_offset, c, col_lower, col_upper, row_lower, row_upper, col_start,
                      row_index, value, opt, nullptr, result, nullptr, x, y, z, dual_ray, primal_ray);
}

int ps26119_solve_mps(const char* path, const ps26119_options* opt, ps26119_result* result, const char* solution_path) {
  (void)opt;
  (void)solution_path;
  if (!result) return PS26119_INVALID_ARGUMENT;
  clear(result);
  result->status = PS26119_INVALID_ARGUMENT;
  set_text(result->message, sizeof result->message,
           path ? "read error: MPS files are not read by the MI355X backend library; pass the model to "
                  "ps26119_solve_lp"
                : "path is NULL");
  return result->status;
}

void ps26119_amd_default_options(ps26119_amd_options* o) {
  if (!o) return;
  std::memset(o, 0, sizeof *o);
  o->presolve = 1;
}

int ps26119_amd_solve_lp(int num_rows, int num_cols, int sense, double obj_offset, const double* c,
                         const double* col_lower, const double* col_upper, const double* row_lower,
                         const double* row_upper, const int* col_start, const int* row_index, const double* value,
                         const ps26119_options* opt, const ps26119_amd_options* ext, ps26119_result* result,
                         ps26119_amd_stats* stats, double* x, double* y, double* z, double* dual_ray, double* primal_ray) {
  return solve_common(num_rows, num_cols, sense, obj_offset, c, col_lower, col_upper, row_lower, row_upper, col_start,
                      row_index, value, opt, ext, result, stats, x, y, z, dual_ray, primal_ray);
}

int ps26119_amd_solve_batch(int num_rows, int num_cols, int sense, double obj_offset, const double* c,
                            const double* col_lower, const double* col_upper, const double* row_lower,
                            const double* row_upper, const int* col_start, const int* row_index, const double* value,
                            int num_scenarios, const ps26119_amd_scenario* scenarios, const ps26119_options* opt,
                            const ps26119_amd_options* ext, ps26119_result* results, double* xs, double* ys) {
  if (!results || num_scenarios < 0 || (num_scenarios > 0 && !scenarios)) return PS26119_INVALID_ARGUMENT;
  for (int k = 0; k < num_scenarios; ++k) clear(&results[k]);
  auto invalid_all = [&](const std::string& why) {
    for (int k = 0; k < num_scenarios; ++k) {
      results[k].status = PS26119_INVALID_ARGUMENT;
      set_text(results[k].message, sizeof results[k].message, why);
    }
    return static_cast<int>(PS26119_INVALID_ARGUMENT);
  };
  ps26119_result scratch;
  clear(&scratch);
  const int rc = guarded(&scratch, [&]() -> int {
    Model m;
    std::string err;
    if (const char* bad = make_model(num_rows, num_cols, sense, obj_offset, c, col_lower, col_upper, row_lower, row_upper,
                                     col_start, row_index, value, m, err))
      return invalid_all(bad);
    Options o;
    if (const char* bad = to_options(opt, num_rows, num_cols, o)) return invalid_all(bad);
    if (ext) {
      if (ext->check_every > 0) o.termination_check_every = ext->check_every;
      if (const char* bad = take_params(ext, o)) return invalid_all(bad);
    }
    std::vector<Scenario> S(num_scenarios);
    for (int k = 0; k < num_scenarios; ++k) {
      const ps26119_amd_scenario& q = scenarios[k];
      if (q.obj) S[k].obj.assign(q.obj, q.obj + num_cols);
      if (q.col_lower) S[k].col_lower.assign(q.col_lower, q.col_lower + num_cols);
      if (q.col_upper) S[k].col_upper.assign(q.col_upper, q.col_upper + num_cols);
      if (q.row_lower) S[k].row_lower.assign(q.row_lower, q.row_lower + num_rows);
      if (q.row_upper) S[k].row_upper.assign(q.row_upper, q.row_upper + num_rows);
    }
    const std::vector<Solution> sol = amd::solve_batch(m, S, o);
    int first = PS26119_OPTIMAL;
    for (int k = 0; k < num_scenarios; ++k) {
      fill(sol[k], &results[k]);
      if (first == PS26119_OPTIMAL && results[k].status != PS26119_OPTIMAL) first = results[k].status;
      const bool have = sol[k].x.size() == static_cast<std::size_t>(num_cols) && sol[k].y.size() == static_cast<std::size_t>(num_rows);
      if (xs && have && num_cols > 0)
        std::memcpy(xs + static_cast<std::size_t>(k) * num_cols, sol[k].x.data(), sizeof(double) * num_cols);
      if (ys && have && num_rows > 0)
        std::memcpy(ys + static_cast<std::size_t>(k) * num_rows, sol[k].y.data(), sizeof(double) * num_rows);
    }
    return first;
  });
  if (scratch.status == PS26119_NUMERICAL_ERROR && scratch.message[0]) {  // an exception escaped the solver
    for (int k = 0; k < num_scenarios; ++k) results[k] = scratch;
    return PS26119_NUMERICAL_ERROR;
  }
  return rc;
}

struct ps26119_amd_server {
  std::unique_ptr<amd::Server> srv;
  std::mutex mu;
  std::map<long long, std::pair<int, int>> dims;  // id → (m, n)
};

ps26119_amd_server* ps26119_amd_server_create(const ps26119_amd_server_options* o, char* error, int error_len) {
  try {
    amd::ServerOptions so;
    if (o) {
      so.device = o->device;
      if (o->budget_bytes > 0) so.budget_bytes = static_cast<std::size_t>(o->budget_bytes);
      if (o->max_instances > 0) so.max_instances = o->max_instances;
      if (o->check_every > 0) so.check_every = o->check_every;
      if (o->setup_threads > 0) so.setup_threads = o->setup_threads;
      so.graphs = o->eager_launches == 0;
      so.timing = o->timing != 0;
    }
    auto s = std::make_unique<ps26119_amd_server>();
    s->srv = std::make_unique<amd::Server>(so);
    if (!s->srv->ok()) {
      if (error && error_len > 0) set_text(error, error_len, s->srv->error());
      return nullptr;
    }
    return s.release();
  } catch (const std::exception& e) {
    if (error && error_len > 0) set_text(error, error_len, e.what());
    return nullptr;
  }
}

void ps26119_amd_server_destroy(ps26119_amd_server* s) { delete s; }

long long ps26119_amd_server_submit(ps26119_amd_server* s, int num_rows, int num_cols, int sense, double obj_offset,
                                    const double* c, const double* col_lower, const double* col_upper,
                                    const double* row_lower, const double* row_upper, const int* col_start,
                                    const int* row_index, const double* value, const ps26119_options* opt,
                                    const ps26119_amd_options* ext) {
  if (!s) return -1;
  try {
    Model m;
    std::string err;
    if (num_rows < 0 || num_cols < 0 || !col_start) return -1;
    // an invalid model is still accepted: it finishes at once with its status and message
    (void)make_model(num_rows, num_cols, sense, obj_offset, c, col_lower, col_upper, row_lower, row_upper, col_start,
                     row_index, value, m, err);
    if (m.num_rows != num_rows || m.num_cols != num_cols) return -1;
    Options o;
    if (to_options(opt, num_rows, num_cols, o)) return -1;
    o.use_gpu = true;
    if (ext) {
      if (ext->check_every > 0) o.termination_check_every = ext->check_every;
      if (take_params(ext, o)) return -1;
    }
    const long long id = s->srv->submit(std::move(m), o);
    std::lock_guard<std::mutex> g(s->mu);
    s->dims[id] = {num_rows, num_cols};
    return id;
  } catch (...) {
    return -1;
  }
}

int ps26119_amd_server_wait(ps26119_amd_server* s, long long id, double timeout_s, ps26119_result* result, double* x,
                            double* y, double* z, double* times) {
  if (!s) return -1;
  int m = 0, n = 0;
  {
    std::lock_guard<std::mutex> g(s->mu);
    auto it = s->dims.find(id);
    if (it == s->dims.end()) return -1;
    m = it->second.first;
    n = it->second.second;
  }
  amd::ServedResult r;
  if (!s->srv->wait(id, r, timeout_s)) return 1;
  {
    std::lock_guard<std::mutex> g(s->mu);
    s->dims.erase(id);
  }
  if (result) {
    clear(result);
    fill(r.solution, result);
  }
  auto put = [](double* dst, const std::vector<double>& v, int len) {
    if (dst && static_cast<int>(v.size()) == len && len > 0) std::memcpy(dst, v.data(), sizeof(double) * len);
  };
  put(x, r.solution.x, n);
  put(y, r.solution.y, m);
  put(z, r.solution.z, n);
  if (times) {
    times[0] = r.submit_time;
    times[1] = r.admit_time;
    times[2] = r.done_time;
    times[3] = r.latency();
  }
  return 0;
}

int ps26119_amd_server_in_flight(ps26119_amd_server* s) { return s ? s->srv->in_flight() : 0; }

int ps26119_amd_server_kernel_time(ps26119_amd_server* s, long long* launches, double* total_ms, double* bytes) {
  long long l = 0;
  double ms = 0.0, b = 0.0;
  if (!s || !s->srv->kernel_time(l, ms, b)) return -1;
  if (launches) *launches = l;
  if (total_ms) *total_ms = ms;
  if (bytes) *bytes = b;
  return 0;
}

int ps26119_amd_device_count(void) { return psh_device_count(); }

int ps26119_amd_scale(int num_rows, int num_cols, int sense, const double* c, const double* col_lower,
                      const double* col_upper, const double* row_lower, const double* row_upper, const int* col_start,
                      const int* row_index, const double* value, double* row_scale, double* col_scale, double* scalars,
                      double* a_vals, double* at_vals) {
  try {
    Model m;
    std::string err;
    if (make_model(num_rows, num_cols, sense, 0.0, c, col_lower, col_upper, row_lower, row_upper, col_start, row_index,
                   value, m, err))
      return PS26119_INVALID_ARGUMENT;
    const amd::ScaledProblem sp = amd::make_scaled_problem(m, amd::ScalingOptions{});
    if (row_scale) std::memcpy(row_scale, sp.row_scale.data(), sizeof(double) * sp.m);
    if (col_scale) std::memcpy(col_scale, sp.col_scale.data(), sizeof(double) * sp.n);
    if (scalars) {
      scalars[0] = sp.bound_scale;
      scalars[1] = sp.obj_scale;
      scalars[2] = sp.log10_range;
      scalars[3] = sp.geometric_mean_applied ? 1.0 : 0.0;
    }
    if (a_vals && !sp.A.val.empty()) std::memcpy(a_vals, sp.A.val.data(), sizeof(double) * sp.A.val.size());
    if (at_vals && !sp.At.val.empty()) std::memcpy(at_vals, sp.At.val.data(), sizeof(double) * sp.At.val.size());
    return PS26119_OPTIMAL;
  } catch (...) {
    return PS26119_NUMERICAL_ERROR;
  }
}

double ps26119_amd_certified_bound(int num_rows, int num_cols, int sense, double obj_offset, const double* c,
                                   const double* col_lower, const double* col_upper, const double* row_lower,
                                   const double* row_upper, const int* col_start, const int* row_index,
                                   const double* value, const double* y, int use_implied_bounds) {
  try {
    Model m;
    std::string err;
    if (!y || make_model(num_rows, num_cols, sense, obj_offset, c, col_lower, col_upper, row_lower, row_upper, col_start,
                         row_index, value, m, err))
      return kNaN;
    return amd::certified_dual_bound(m, std::vector<double>(y, y + num_rows), use_implied_bounds != 0).bound;
  } catch (...) {
    return kNaN;
  }
}

int ps26119_amd_check_solution(int num_rows, int num_cols, int sense, double obj_offset, const double* c,
                               const double* col_lower, const double* col_upper, const double* row_lower,
                               const double* row_upper, const int* col_start, const int* row_index, const double* value,
                               const double* x, const double* y, double* out) {
  try {
    Model m;
    std::string err;
    if (!x || !y || !out ||
        make_model(num_rows, num_cols, sense, obj_offset, c, col_lower, col_upper, row_lower, row_upper, col_start,
                   row_index, value, m, err))
      return PS26119_INVALID_ARGUMENT;
    const amd::CheckReport r =
        amd::check_solution(m, std::vector<double>(x, x + num_cols), std::vector<double>(y, y + num_rows));
    const double v[6] = {r.max_primal_violation, r.max_dual_violation, r.relative_gap, r.primal_objective,
                         r.dual_objective, r.passed() ? 1.0 : 0.0};
    std::memcpy(out, v, sizeof v);
    return PS26119_OPTIMAL;
  } catch (...) {
    return PS26119_NUMERICAL_ERROR;
  }
}

int ps26119_amd_presolve_info(int num_rows, int num_cols, int sense, double obj_offset, const double* c,
                              const double* col_lower, const double* col_upper, const double* row_lower,
                              const double* row_upper, const int* col_start, const int* row_index, const double* value,
                              double* out) {
  try {
    Model m;
    std::string err;
    if (!out || make_model(num_rows, num_cols, sense, obj_offset, c, col_lower, col_upper, row_lower, row_upper, col_start,
                           row_index, value, m, err))
      return PS26119_INVALID_ARGUMENT;
    const amd::Presolved r = amd::presolve(m);
    using O = amd::Presolved::Outcome;
    out[0] = r.outcome == O::Reduced ? 0 : (r.outcome == O::ProvedInfeasible ? 1 : 2);
    out[1] = r.rows_dropped;
    out[2] = r.cols_dropped;
    out[3] = r.reduced.num_rows;
    out[4] = r.reduced.num_cols;
    out[5] = static_cast<double>(r.reduced.nnz());
    out[6] = r.reduced.obj_offset;
    return PS26119_OPTIMAL;
  } catch (...) {
    return PS26119_NUMERICAL_ERROR;
  }
}

int ps26119_amd_presolve_model(int num_rows, int num_cols, int sense, double obj_offset, const double* c,
                               const double* col_lower, const double* col_upper, const double* row_lower,
                               const double* row_upper, const int* col_start, const int* row_index, const double* value,
                               double* info, int* kept_rows, int* kept_cols, double* rl, double* ru, double* cl,
                               double* cu, int* r_col_start, int* r_row_index, double* r_value) {
  try {
    Model m;
    std::string err;
    if (!info || make_model(num_rows, num_cols, sense, obj_offset, c, col_lower, col_upper, row_lower, row_upper, col_start,
                            row_index, value, m, err))
      return PS26119_INVALID_ARGUMENT;
    const amd::Presolved r = amd::presolve(m);
    using O = amd::Presolved::Outcome;
    info[0] = r.outcome == O::Reduced ? 0 : (r.outcome == O::ProvedInfeasible ? 1 : 2);
    info[1] = r.rows_dropped;
    info[2] = r.cols_dropped;
    info[3] = r.reduced.num_rows;
    info[4] = r.reduced.num_cols;
    info[5] = static_cast<double>(r.reduced.nnz());
    info[6] = r.reduced.obj_offset;
    if (r.outcome != O::Reduced) return PS26119_OPTIMAL;
    const Model& R = r.reduced;
    auto put = [](auto* dst, const auto& src) {
      if (dst && !src.empty()) std::copy(src.begin(), src.end(), dst);
    };
    put(kept_rows, r.kept_rows);
    put(kept_cols, r.kept_cols);
    put(rl, R.row_lower);
    put(ru, R.row_upper);
    put(cl, R.col_lower);
    put(cu, R.col_upper);
    put(r_col_start, R.col_start);
    put(r_row_index, R.row_index);
    put(r_value, R.value);
    return PS26119_OPTIMAL;
  } catch (...) {
    return PS26119_NUMERICAL_ERROR;
  }
}

}  // extern "C"
