// solve.cpp — solve(): the path from a caller's model to a verified answer on the MI355X.
// A Pipeline runs the stages in order — admission, presolve, the engine on the reduced model,
// lifting the answer back and re-checking it on the original model, the certified bound, the
// gate — with one budget: every follow-up solve gets only what is left of the caller's time and
// iterations. The decisions and their order are the reference's (src/core/solve.cpp:119-185,
// 266-467, LP part); requests for a CPU engine or another algorithm are answered NotSolved with
// a reason instead of being run here.
#include "solve.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <memory>
#include <new>

#include "engine.h"
#include "la.h"
#include "presolve.h"
#include "verify.h"

namespace ps26119::amd {

namespace {

using Clock = std::chrono::steady_clock;

double seconds_since(Clock::time_point t0) { return std::chrono::duration<double>(Clock::now() - t0).count(); }

void append(Solution& s, const std::string& t) { s.message += std::string(s.message.empty() ? "" : "; ") + t; }

bool all_finite(const std::vector<double>& v) {
  return std::all_of(v.begin(), v.end(), [](double a) { return std::isfinite(a); });
}

bool warm_start_sized(const Model& model, const Options& o) {
  return (o.warm_x.empty() || static_cast<int>(o.warm_x.size()) == model.num_cols) &&
         (o.warm_y.empty() || static_cast<int>(o.warm_y.size()) == model.num_rows);
}

// The certified bound may also use implied bounds while the call is still within its time.
bool may_use_implied_bounds(double elapsed, double time_limit) { return elapsed < 1.1 * time_limit + 1.0; }

// Relative KKT error of a point on the original model (the reduced solve's yardstick too).
KktStats kkt_of_point(const Model& M, const Solution& s) {
  ScaledProblem sp;  // kkt_on_original reads only the original model
  sp.m = M.num_rows;
  sp.n = M.num_cols;
  sp.original = &M;
  std::vector<double> y_min(s.y);
  for (double& v : y_min) v *= M.sense;
  return kkt_on_original(sp, s.x, y_min);
}

Solution infeasible_by_bounds(const Model& model, const Options& options, const std::string& which) {
  Solution s;
  s.status = Status::Infeasible;
  s.engine = to_string(options.algorithm);
  s.precision = to_string(options.precision);
  s.model_fingerprint = model.fingerprint_hex();
  s.message = "bounds cross: " + which;
  return s;
}

Solution refuse(const Model& model, const char* engine, const std::string& why, Status st = Status::NotSolved) {
  Solution s;
  s.status = st;
  s.engine = engine;
  s.message = why;
  s.model_fingerprint = model.fingerprint_hex();
  return s;
}

// The engine on one model, exactly as given (no presolve): argument checks, r²HPDHG or PDLP on
// the device, and for an Unbounded verdict whose point is not primal-feasible a zero-objective
// solve for a point the ray can start from. `certify`: also the certified dual bound (with the
// implied bounds of `implied` when it was started for this model).
Solution engine_solve(const Model& model, const Options& options, bool certify, ImpliedBoundsJob* implied = nullptr) {
  const auto t0 = Clock::now();
  Solution sol;
  sol.model_fingerprint = model.fingerprint_hex();
  sol.precision = to_string(options.precision);
  sol.engine = to_string(options.algorithm);
  if (auto err = model.validate(); !err.empty()) {
    sol.status = Status::NotSolved;
    sol.message = "invalid model: " + err;
    return sol;
  }
  if (!warm_start_sized(model, options)) {
    sol.status = Status::NotSolved;
    sol.message = "a warm start vector does not match the model's size";
    return sol;
  }
  if (!all_finite(options.warm_x) || !all_finite(options.warm_y) ||
      !(options.warm_primal_weight >= 0 && std::isfinite(options.warm_primal_weight))) {
    sol.status = Status::NotSolved;
    sol.message = "a warm start holds a non-finite value";
    return sol;
  }
  if (auto which = model.crossed_bounds(); !which.empty()) return infeasible_by_bounds(model, options, which);
  try {
    EngineOptions eo = engine_options_from(options);
    if (auto bad = apply_engine_params(options, eo); !bad.empty()) {
      sol.status = Status::NotSolved;
      sol.message = "unknown engine parameter '" + bad + "'";
      return sol;
    }
    sol = options.algorithm == Algorithm::Pdlp ? solve_pdlp(model, eo) : solve_r2hpdhg(model, eo);
    if (sol.status == Status::Unbounded && !check_unboundedness_certificate(model, sol.x, sol.primal_ray).passed) {
      // The ray is only a certificate from a feasible point: look for one with c = 0 in what is
      // left of the budget (reference solve.cpp:552-582).
      Model feasibility = model;
      std::fill(feasibility.obj.begin(), feasibility.obj.end(), 0.0);
      Options fo = options;
      fo.presolve = false;
      fo.warm_x.clear();
      fo.warm_y.clear();
      fo.time_limit = std::max(0.0, options.time_limit - seconds_since(t0));
      fo.iteration_limit = std::max<std::int64_t>(0, options.iteration_limit - sol.iterations);
      Solution f = engine_solve(feasibility, fo, true);
      const std::int64_t both = sol.iterations + f.iterations;
      if (f.status == Status::Optimal) {
        sol.x = f.x;
        sol.row_activity = activity(model, sol.x);
        append(sol, "a zero-objective solve supplied the feasible point (" + std::to_string(f.iterations) +
                        " iterations)");
      } else if (f.status == Status::Infeasible) {
        f.message = "a direction of unboundedness exists but no feasible point does: " + f.message;
        f.primal_ray.clear();
        sol = f;
      } else {
        sol.status = f.status;
        append(sol, std::string("the search for a feasible point ended ") + to_string(f.status));
        sol.primal_ray.clear();
      }
      sol.iterations = both;
    }
  } catch (const std::bad_alloc&) {
    sol = Solution{};
    sol.status = Status::NotSolved;
    sol.message = "out of memory";
  } catch (const std::exception& e) {
    sol = Solution{};
    sol.status = Status::NumericalError;
    sol.message = e.what();
  }
  sol.model_fingerprint = model.fingerprint_hex();
  if (certify && static_cast<int>(sol.y.size()) == model.num_rows && model.num_rows + model.num_cols > 0)
    sol.certified_bound =
        certified_dual_bound(model, sol.y, may_use_implied_bounds(seconds_since(t0), options.time_limit), implied).bound;
  sol.seconds = seconds_since(t0);
  return sol;
}

class Pipeline {
 public:
  Pipeline(const Model& model, const Options& options, ImpliedBoundsJob* implied)
      : M_(model), opt_(options), t0_(Clock::now()), implied_(implied) {}

  Solution run() {
    const bool usable = options_usable();
    if (!opt_.presolve || !usable || !M_.validate().empty()) return engine_solve(M_, opt_, true, implied_);
    if (auto which = M_.crossed_bounds(); !which.empty()) return infeasible_by_bounds(M_, opt_, which);
    const auto tp = Clock::now();
    P_ = presolve(M_);
    if (opt_.verbosity >= 1) std::fprintf(stderr, "presolve: %.3fs\n", seconds_since(tp));
    switch (P_.outcome) {
      case Presolved::Outcome::NothingToDo:
        return engine_solve(M_, opt_, true, implied_);
      case Presolved::Outcome::ProvedInfeasible:
        return solve_original(0, "it proved the model infeasible (" + P_.reason + ")");
      case Presolved::Outcome::Reduced:
        break;
    }
    summary_ = "presolve: " + std::to_string(P_.rows_dropped) + " rows and " + std::to_string(P_.cols_dropped) +
               " columns eliminated";
    const Solution red = solve_reduced();
    if (Solution lifted; lift_certificate(red, lifted)) return lifted;
    if (red.status == Status::Infeasible || red.status == Status::Unbounded)
      return solve_original(red.iterations, std::string("the reduced model came back ") + to_string(red.status));
    return finish_optimal_path(red);
  }

 private:
  bool options_usable() const {
    return warm_start_sized(M_, opt_) && all_finite(opt_.warm_x) && all_finite(opt_.warm_y);
  }

  // The caller's options with what is left of its time and iterations.
  Options remaining(Options o, std::int64_t spent_iterations) const {
    o.time_limit = std::max(0.0, opt_.time_limit - seconds_since(t0_));
    o.iteration_limit = std::max<std::int64_t>(0, opt_.iteration_limit - spent_iterations);
    return o;
  }

  // Options for solves of the reduced model: no presolve, and the original model's ‖b‖, ‖c‖ and
  // objective shift in the relative KKT terms, so that a tolerance means the same as on the
  // original.
  Options reduced_options() const {
    Options o = opt_;
    o.presolve = false;
    double b2 = 0.0, c2 = 0.0;
    for (int i = 0; i < M_.num_rows; ++i) {
      const double lo = std::isfinite(M_.row_lower[i]) ? std::fabs(M_.row_lower[i]) : 0.0;
      const double up = std::isfinite(M_.row_upper[i]) ? std::fabs(M_.row_upper[i]) : 0.0;
      const double big = std::max(lo, up);
      b2 += big * big;
    }
    for (double c : M_.obj) c2 += c * c;
    o.kkt_b_norm = std::sqrt(b2);
    o.kkt_c_norm = std::sqrt(c2);
    o.kkt_obj_shift = M_.sense * (P_.reduced.obj_offset - M_.obj_offset);
    return o;
  }

  Solution solve_reduced() {
    Options o = reduced_options();
    if (!opt_.warm_x.empty()) {
      o.warm_x.clear();
      for (int j : P_.kept_cols) o.warm_x.push_back(opt_.warm_x[j]);
    }
    if (!opt_.warm_y.empty()) {
      o.warm_y.clear();
      for (int i : P_.kept_rows) o.warm_y.push_back(opt_.warm_y[i]);
    }
    if (P_.reduced.num_cols == 0 && P_.reduced.num_rows == 0) {  // presolve decided everything
      Solution s;
      s.status = Status::Optimal;
      s.engine = "presolve";
      return s;
    }
    return engine_solve(P_.reduced, remaining(o, 0), false);
  }

  // An Infeasible / Unbounded verdict of the reduced model whose certificate, lifted to the
  // original model, holds there: true with the answer in `out`.
  bool lift_certificate(const Solution& red, Solution& out) const {
    if (red.status == Status::Infeasible && static_cast<int>(red.dual_ray.size()) == P_.reduced.num_rows) {
      std::vector<double> ray = lift_dual_ray(M_, P_, red.dual_ray);
      if (ray.empty() || !check_infeasibility_certificate(M_, ray).passed) return false;
      out = red;
      out.x.clear();
      out.y.clear();
      out.z.clear();
      out.row_activity.clear();
      out.dual_ray = std::move(ray);
      append(out, summary_ + "; the reduced model's dual ray, lifted to the original rows, holds there");
    } else if (red.status == Status::Unbounded && static_cast<int>(red.primal_ray.size()) == P_.reduced.num_cols &&
               static_cast<int>(red.x.size()) == P_.reduced.num_cols) {
      // removed columns move by 0 along the ray; a bound a removed row created is a reduced
      // column bound, which the ray already respects
      std::vector<double> dir(M_.num_cols, 0.0);
      for (int k = 0; k < P_.reduced.num_cols; ++k) dir[P_.kept_cols[k]] = red.primal_ray[k];
      Solution s = postsolve(M_, P_, red);
      if (!check_unboundedness_certificate(M_, s.x, dir).passed) return false;
      out = std::move(s);
      out.status = Status::Unbounded;
      out.primal_ray = std::move(dir);
      out.dual_ray.clear();
      out.y.clear();
      out.z.clear();
      out.objective = out.dual_objective = kNaN;
      append(out, summary_ + "; the reduced model's ray and point, lifted, hold on the original model");
    } else {
      return false;
    }
    out.model_fingerprint = M_.fingerprint_hex();
    out.seconds = seconds_since(t0_);
    return true;
  }

  // The original model solved directly, for a verdict with a certificate in its own space.
  Solution solve_original(std::int64_t spent, const std::string& why) const {
    Options o = remaining(opt_, spent);
    o.presolve = false;
    Solution s = engine_solve(M_, o, true, implied_);
    s.iterations += spent;
    s.seconds = seconds_since(t0_);
    append(s, "presolve: " + why + "; the original model was solved directly");
    return s;
  }

  bool meets_tolerance(const Solution& s) const {
    return static_cast<int>(s.x.size()) == M_.num_cols && static_cast<int>(s.y.size()) == M_.num_rows &&
           kkt_of_point(M_, s).converged(opt_.tolerance);
  }

  Solution finish_optimal_path(const Solution& red) const {
    const auto tf = Clock::now();
    Solution post = postsolve(M_, P_, red);
    append(post, summary_);
    post.model_fingerprint = M_.fingerprint_hex();
    const double t_post = seconds_since(tf);
    const bool met = post.status != Status::Optimal || meets_tolerance(post);
    const double t_met = seconds_since(tf);
    if (post.status == Status::Optimal && !met) {
      // The reduced model's KKT terms differ from the original's: first the reduced model again,
      // warm-started from its answer at 100× the tolerance, then the original from scratch.
      Options tighter = reduced_options();
      tighter.tolerance = std::max(opt_.tolerance * 1e-2, 1e-13);
      tighter.warm_x = red.x;
      tighter.warm_y = red.y;
      const Solution red2 = engine_solve(P_.reduced, remaining(tighter, red.iterations), true);
      Solution post2 = postsolve(M_, P_, red2);
      post2.iterations += red.iterations;
      if (post2.status == Status::Optimal && meets_tolerance(post2)) {
        char tol[32];
        std::snprintf(tol, sizeof tol, "%.0e", tighter.tolerance);
        post2.message = summary_ + "; the reduced model was solved again to " + tol + " to hold on the original";
        post = std::move(post2);
      } else {
        const std::int64_t spent = red.iterations + red2.iterations;
        Solution direct = engine_solve(M_, remaining(reduced_options(), spent), true, implied_);
        direct.iterations += spent;
        append(direct, summary_ + "; the lifted point missed the tolerance on the original model, which was solved "
                                   "directly");
        post = std::move(direct);
      }
    } else if (post.status == Status::Optimal && static_cast<int>(post.x.size()) != M_.num_cols) {
      post.status = Status::NumericalError;
    }
    const double t_before_bound = seconds_since(tf);
    if (static_cast<int>(post.y.size()) == M_.num_rows && M_.num_rows + M_.num_cols > 0)
      post.certified_bound =
          certified_dual_bound(M_, post.y, may_use_implied_bounds(seconds_since(t0_), opt_.time_limit), implied_).bound;
    if (opt_.verbosity >= 1)
      std::fprintf(stderr, "postsolve %.3fs, KKT of the lifted point %.3fs, certified bound %.3fs\n", t_post,
                   t_met - t_post, seconds_since(tf) - t_before_bound);
    post.seconds = seconds_since(t0_);
    return post;
  }

  const Model& M_;
  const Options& opt_;
  const Clock::time_point t0_;
  ImpliedBoundsJob* implied_;  // the original model's implied bounds, computed while the engine runs
  Presolved P_;
  std::string summary_;
};

}  // namespace

void gate(const Model& model, const Options& options, Solution& s) {
  if (s.status == Status::Infeasible || s.status == Status::Unbounded) {
    s.objective = s.dual_objective = kNaN;
    if (s.status == Status::Infeasible && !model.crossed_bounds().empty()) {
      s.check = "PASS";  // crossed bounds certify themselves
      return;
    }
    const CertificateCheck c = s.status == Status::Infeasible ? check_infeasibility_certificate(model, s.dual_ray)
                                                              : check_unboundedness_certificate(model, s.x, s.primal_ray);
    if (!c.present) {
      s.check.clear();
      append(s, "no certificate came back from the engine");
    } else if (c.passed) {
      s.check = "PASS";
      append(s, c.detail);
    } else {
      s.check = "FAIL";
      append(s, std::string(to_string(s.status)) + " revoked: " + c.detail);
      s.status = Status::NumericalError;
    }
    return;
  }
  if (s.status != Status::Optimal) return;
  if (static_cast<int>(s.x.size()) != model.num_cols || static_cast<int>(s.y.size()) != model.num_rows) {
    s.status = Status::NumericalError;
    append(s, "Optimal came back without its primal and dual vectors");
    return;
  }
  const CheckReport rep = check_solution(model, s.x, s.y);
  s.check = rep.passed() ? "PASS" : "FAIL";
  s.check_primal = rep.max_primal_violation;
  s.check_dual = rep.max_dual_violation;
  s.check_gap = rep.relative_gap;
  if (std::isnan(s.dual_objective)) s.dual_objective = rep.dual_objective;
  if (std::isnan(s.primal_residual)) s.primal_residual = rep.max_primal_violation;
  if (std::isnan(s.dual_residual)) s.dual_residual = rep.max_dual_violation;
  if (std::isnan(s.gap)) s.gap = rep.relative_gap;
  if (rep.passed()) return;
  const std::string finding = "the check on the original model found " + rep.summary();
  if (options.tolerance < tol::kVerifyPrimal) {  // a tolerance that asked for verifier-grade accuracy
    s.status = Status::NumericalError;
    append(s, "Optimal revoked: " + finding);
  } else {
    append(s, finding + " (the requested tolerance is looser than the checker's)");
  }
}

Solution solve(const Model& model, const Options& options) {
  try {
    if (options.algorithm != Algorithm::Auto && options.algorithm != Algorithm::R2hpdhg &&
        options.algorithm != Algorithm::Pdlp)
      return refuse(model, to_string(options.algorithm),
                    std::string("the MI355X backend runs the first-order engines (r2hpdhg, pdlp); algorithm '") +
                        to_string(options.algorithm) + "' is served by the reference CPU library");
    if (!options.use_gpu)
      return refuse(model, "r2hpdhg",
                    "the MI355X backend serves GPU solves: set use_gpu = 1 (CPU engines are in the reference library)");
    const bool integers = std::any_of(model.is_integer.begin(), model.is_integer.end(), [](auto v) { return v != 0; });
    if (integers && !options.relax_integrality)
      return refuse(model, "r2hpdhg",
                    "integer columns: branch-and-bound is not part of the MI355X backend (solve the LP relaxation "
                    "with relax_integrality)");
    Options resolved = options;
    const bool chose = options.algorithm == Algorithm::Auto;
    if (chose) resolved.algorithm = Algorithm::R2hpdhg;  // a GPU request defaults to r2hpdhg
    set_host_threads(options.threads);
    // The certified bound needs the implied column bounds only when a column has an infinite
    // bound; they depend on the model alone, so they are computed while the engine runs.
    std::unique_ptr<ImpliedBoundsJob> implied;
    auto open = [](const std::vector<double>& b) {
      return std::any_of(b.begin(), b.end(), [](double v) { return !std::isfinite(v); });
    };
    if ((open(model.col_lower) || open(model.col_upper)) && model.num_rows > 0 && model.validate().empty())
      implied = std::make_unique<ImpliedBoundsJob>(model);
    Solution s = Pipeline(model, resolved, implied.get()).run();
    gate(model, resolved, s);
    if (chose) {
      char head[96];
      const double work = static_cast<double>(std::max(model.num_rows, 1)) * static_cast<double>(model.nnz());
      std::snprintf(head, sizeof head, "auto: r2hpdhg (rows*nnz = %.2g)", work);
      s.message = s.message.empty() ? head : std::string(head) + "; " + s.message;
    }
    if (s.engine.empty()) s.engine = resolved.algorithm == Algorithm::Pdlp ? "pdlp" : "r2hpdhg";
    return s;
  } catch (const std::bad_alloc&) {
    return refuse(model, "", "out of memory");
  } catch (const std::exception& e) {
    return refuse(model, "", std::string("internal error: ") + e.what(), Status::NumericalError);
  }
}

}  // namespace ps26119::amd
