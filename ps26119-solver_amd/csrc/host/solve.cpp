// solve.cpp — solve(): validation, presolve wrapper, dispatch to the MI355X r²HPDHG engine,
// the verification gate and the certified bound, with the reference's semantics
// (reference src/core/solve.cpp:119-609, LP part). What this library serves is the GPU
// first-order path: a request for a CPU engine, or for another algorithm, is answered with
// NotSolved and a reason instead of being run on the CPU.
#include "solve.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <new>

#include "engine.h"
#include "la.h"
#include "presolve.h"
#include "verify.h"

namespace ps26119::amd {

namespace {

bool within_bound_budget(double elapsed, double time_limit) { return elapsed < 1.1 * time_limit + 1.0; }

bool has_integers(const Model& m) {
  for (auto v : m.is_integer)
    if (v) return true;
  return false;
}

KktStats original_kkt(const Model& M, const Solution& s) {
  ScaledProblem sp;  // only the original matrix is used by kkt_on_original
  sp.m = M.num_rows;
  sp.n = M.num_cols;
  sp.original = &M;
  sp.A = transpose(csc_as_transpose(M));
  sp.a_orig = sp.A.val;
  std::vector<double> ymin(s.y);
  for (double& v : ymin) v *= M.sense;
  return kkt_on_original(sp, s.x, ymin);
}

Solution trivially_infeasible(const Model& model, const Options& options, const std::string& why) {
  Solution s;
  s.status = Status::Infeasible;
  s.engine = to_string(options.algorithm);
  s.precision = to_string(options.precision);
  s.model_fingerprint = model.fingerprint_hex();
  s.message = "bounds cross: " + why;
  return s;
}

void note(Solution& s, const std::string& t) { s.message += std::string(s.message.empty() ? "" : "; ") + t; }

Solution solve_direct(const Model& model, const Options& options, bool certify) {
  const auto t0 = std::chrono::steady_clock::now();
  Solution sol;
  sol.model_fingerprint = model.fingerprint_hex();
  sol.precision = to_string(options.precision);
  sol.engine = to_string(options.algorithm);
  if (auto err = model.validate(); !err.empty()) {
    sol.status = Status::NotSolved;
    sol.message = "invalid model: " + err;
    return sol;
  }
  if ((!options.warm_x.empty() && static_cast<int>(options.warm_x.size()) != model.num_cols) ||
      (!options.warm_y.empty() && static_cast<int>(options.warm_y.size()) != model.num_rows)) {
    sol.status = Status::NotSolved;
    sol.message = "warm start vectors have the wrong size";
    return sol;
  }
  auto finite = [](const std::vector<double>& v) { return std::all_of(v.begin(), v.end(), [](double a) { return std::isfinite(a); }); };
  if (!finite(options.warm_x) || !finite(options.warm_y) ||
      !(options.warm_primal_weight >= 0 && std::isfinite(options.warm_primal_weight))) {
    sol.status = Status::NotSolved;
    sol.message = "warm start contains non-finite values";
    return sol;
  }
  if (auto why = model.crossed_bounds(); !why.empty()) return trivially_infeasible(model, options, why);
  try {
    EngineOptions eo = engine_options_from(options);
    if (auto bad = apply_engine_params(options, eo); !bad.empty()) {
      sol.status = Status::NotSolved;
      sol.message = "unknown engine parameter '" + bad + "'";
      return sol;
    }
    sol = options.algorithm == Algorithm::Pdlp ? solve_pdlp(model, eo) : solve_r2hpdhg(model, eo);
    if (sol.status == Status::Unbounded && !check_unboundedness_certificate(model, sol.x, sol.primal_ray).passed) {
      // a valid ray from a point that is not primal-feasible: find a feasible point with a
      // zero-objective solve within the remaining budget (reference solve.cpp:552-582)
      Model feas = model;
      std::fill(feas.obj.begin(), feas.obj.end(), 0.0);
      Options fo = options;
      fo.presolve = false;
      fo.warm_x.clear();
      fo.warm_y.clear();
      fo.time_limit = std::max(0.0, options.time_limit - std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
      fo.iteration_limit = std::max<std::int64_t>(0, options.iteration_limit - sol.iterations);
      Solution f = solve_direct(feas, fo, true);
      const std::int64_t total = sol.iterations + f.iterations;
      if (f.status == Status::Optimal) {
        sol.x = f.x;
        sol.row_activity = activity(model, sol.x);
        note(sol, "feasible point from a zero-objective solve (" + std::to_string(f.iterations) + " iterations)");
      } else if (f.status == Status::Infeasible) {
        f.message = "unbounded direction found, but the model is infeasible: " + f.message;
        f.primal_ray.clear();
        sol = f;
      } else {
        sol.status = f.status;
        note(sol, std::string("no feasible point found for the ray (") + to_string(f.status) + ")");
        sol.primal_ray.clear();
      }
      sol.iterations = total;
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
        certified_dual_bound(model, sol.y,
                             within_bound_budget(std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(),
                                                 options.time_limit))
            .bound;
  sol.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return sol;
}

Solution solve_impl(const Model& model, const Options& options) {
  const bool bad_warm = (!options.warm_x.empty() && static_cast<int>(options.warm_x.size()) != model.num_cols) ||
                        (!options.warm_y.empty() && static_cast<int>(options.warm_y.size()) != model.num_rows);
  auto finite = [](const std::vector<double>& v) { return std::all_of(v.begin(), v.end(), [](double a) { return std::isfinite(a); }); };
  const bool bad_values = !finite(options.warm_x) || !finite(options.warm_y);
  if (!options.presolve || bad_warm || bad_values || !model.validate().empty()) return solve_direct(model, options, true);
  if (auto why = model.crossed_bounds(); !why.empty()) return trivially_infeasible(model, options, why);
  const auto t0 = std::chrono::steady_clock::now();
  PresolveResult pr = presolve(model);
  if (options.verbosity >= 1)
    std::fprintf(stderr, "presolve: %.3fs\n", std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
  if (pr.outcome == PresolveResult::Outcome::Unchanged) return solve_direct(model, options, true);
  Options inner = options;
  inner.presolve = false;
  {  // assess the reduced solve with the ORIGINAL model's norms (reference solve.cpp:282-294)
    double b2 = 0, c2 = 0;
    for (int i = 0; i < model.num_rows; ++i) {
      const double lo = std::isfinite(model.row_lower[i]) ? std::fabs(model.row_lower[i]) : 0.0;
      const double up = std::isfinite(model.row_upper[i]) ? std::fabs(model.row_upper[i]) : 0.0;
      b2 += std::max(lo, up) * std::max(lo, up);
    }
    for (double c : model.obj) c2 += c * c;
    inner.kkt_b_norm = std::sqrt(b2);
    inner.kkt_c_norm = std::sqrt(c2);
    inner.kkt_obj_shift = model.sense * (pr.reduced.obj_offset - model.obj_offset);
  }
  auto elapsed = [&] { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); };
  auto budget = [&](Options o, std::int64_t used) {
    o.time_limit = std::max(0.0, options.time_limit - elapsed());
    o.iteration_limit = std::max<std::int64_t>(0, options.iteration_limit - used);
    return o;
  };
  auto on_original = [&](const Solution& why, const std::string& what) {
    Options direct = budget(options, why.iterations);
    direct.presolve = false;
    Solution d = solve_direct(model, direct, true);
    d.iterations += why.iterations;
    d.seconds = elapsed();
    note(d, "presolve: " + what + "; re-solved the original model for an original-space certificate");
    return d;
  };
  if (pr.outcome == PresolveResult::Outcome::Infeasible) return on_original(Solution{}, "reduction found infeasibility (" + pr.message + ")");
  if (!options.warm_x.empty()) {
    inner.warm_x.clear();
    for (int j : pr.col_map) inner.warm_x.push_back(options.warm_x[j]);
  }
  if (!options.warm_y.empty()) {
    inner.warm_y.clear();
    for (int i : pr.row_map) inner.warm_y.push_back(options.warm_y[i]);
  }
  Solution red;
  if (pr.reduced.num_cols == 0 && pr.reduced.num_rows == 0) {
    red.status = Status::Optimal;
    red.engine = "presolve";
  } else {
    red = solve_direct(pr.reduced, budget(inner, 0), false);
  }
  const std::string removed =
      "presolve removed " + std::to_string(pr.removed_rows) + " rows, " + std::to_string(pr.removed_cols) + " cols";
  if (red.status == Status::Infeasible && static_cast<int>(red.dual_ray.size()) == pr.reduced.num_rows) {
    std::vector<double> r = postsolve_farkas(model, pr, red.dual_ray);
    if (!r.empty() && check_infeasibility_certificate(model, r).passed) {
      Solution s = red;
      s.x.clear();
      s.y.clear();
      s.z.clear();
      s.row_activity.clear();
      s.dual_ray = std::move(r);
      note(s, removed + "; reduced-model Farkas certificate postsolved to the original rows");
      s.model_fingerprint = model.fingerprint_hex();
      s.seconds = elapsed();
      return s;
    }
  }
  if (red.status == Status::Unbounded && static_cast<int>(red.primal_ray.size()) == pr.reduced.num_cols &&
      static_cast<int>(red.x.size()) == pr.reduced.num_cols) {
    std::vector<double> d(model.num_cols, 0.0);
    for (int j = 0; j < pr.reduced.num_cols; ++j) d[pr.col_map[j]] = red.primal_ray[j];
    Solution s = postsolve(model, pr, red);
    if (check_unboundedness_certificate(model, s.x, d).passed) {
      s.status = Status::Unbounded;
      s.primal_ray = std::move(d);
      s.dual_ray.clear();
      s.y.clear();
      s.z.clear();
      s.objective = s.dual_objective = kNaN;
      note(s, removed + "; reduced-model ray and point postsolved to the original model");
      s.model_fingerprint = model.fingerprint_hex();
      s.seconds = elapsed();
      return s;
    }
  }
  if (red.status == Status::Infeasible || red.status == Status::Unbounded)
    return on_original(red, std::string("reduced model ") + to_string(red.status));
  Solution post = postsolve(model, pr, red);
  post.message = post.message.empty() ? removed : post.message + "; " + removed;
  post.model_fingerprint = model.fingerprint_hex();
  auto passes = [&](const Solution& s) {
    return static_cast<int>(s.x.size()) == model.num_cols && static_cast<int>(s.y.size()) == model.num_rows &&
           original_kkt(model, s).converged(options.tolerance);
  };
  if (post.status == Status::Optimal && !passes(post)) {
    // tighten the reduced solve 100× warm-started, else solve the original cold
    Options tight = inner;
    tight.tolerance = std::max(options.tolerance * 1e-2, 1e-13);
    tight.warm_x = red.x;
    tight.warm_y = red.y;
    const Solution red2 = solve_direct(pr.reduced, budget(tight, red.iterations), true);
    Solution post2 = postsolve(model, pr, red2);
    post2.iterations += red.iterations;
    if (post2.status == Status::Optimal && passes(post2)) {
      char buf[64];
      std::snprintf(buf, sizeof buf, "%.0e", tight.tolerance);
      post2.message = removed + "; reduced model re-solved to " + buf + " to pass on the original";
      post = post2;
    } else {
      Solution cold = solve_direct(model, budget(inner, red.iterations + red2.iterations), true);
      cold.iterations += red.iterations + red2.iterations;
      note(cold, removed + "; postsolved point missed the tolerance, fell back to solving the original");
      post = cold;
    }
  } else if (post.status == Status::Optimal && static_cast<int>(post.x.size()) != model.num_cols) {
    post.status = Status::NumericalError;
  }
  if (static_cast<int>(post.y.size()) == model.num_rows && model.num_rows + model.num_cols > 0)
    post.certified_bound = certified_dual_bound(model, post.y, within_bound_budget(elapsed(), options.time_limit)).bound;
  post.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return post;
}

}  // namespace

void gate(const Model& model, const Options& options, Solution& s) {  // reference solve.cpp:119-185 (LP)
  if (s.status == Status::Infeasible || s.status == Status::Unbounded) {
    s.objective = s.dual_objective = kNaN;
    if (s.status == Status::Infeasible && !model.crossed_bounds().empty()) {
      s.check = "PASS";
      return;
    }
    const CertificateCheck c = s.status == Status::Infeasible ? check_infeasibility_certificate(model, s.dual_ray)
                                                              : check_unboundedness_certificate(model, s.x, s.primal_ray);
    if (!c.present) {
      s.check.clear();
      note(s, "not certified (the engine returned no certificate)");
    } else if (c.passed) {
      s.check = "PASS";
      note(s, c.detail);
    } else {
      s.check = "FAIL";
      note(s, std::string(to_string(s.status)) + " withdrawn, " + c.detail);
      s.status = Status::NumericalError;
    }
    return;
  }
  if (s.status != Status::Optimal) return;
  if (static_cast<int>(s.x.size()) != model.num_cols || static_cast<int>(s.y.size()) != model.num_rows) {
    s.status = Status::NumericalError;
    note(s, "engine reported Optimal without primal and dual vectors");
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
  const std::string what = "in-process check on the original model: " + rep.summary();
  if (options.tolerance < tol::kVerifyPrimal) {  // first-order engine that promised verifier grade
    s.status = Status::NumericalError;
    note(s, "Optimal withdrawn, " + what);
  } else {
    note(s, what + " (requested tolerance is looser than the verifier's)");
  }
}

Solution solve(const Model& model, const Options& options) {
  try {
    Solution s;
    if (options.algorithm != Algorithm::Auto && options.algorithm != Algorithm::R2hpdhg &&
        options.algorithm != Algorithm::Pdlp) {
      s.status = Status::NotSolved;
      s.engine = to_string(options.algorithm);
      s.message = std::string("the MI355X backend runs the first-order engines (r2hpdhg, pdlp); algorithm '") +
                  to_string(options.algorithm) + "' is served by the reference CPU library";
      s.model_fingerprint = model.fingerprint_hex();
      return s;
    }
    if (!options.use_gpu) {
      s.status = Status::NotSolved;
      s.engine = "r2hpdhg";
      s.message = "the MI355X backend serves GPU solves: set use_gpu = 1 (CPU engines are in the reference library)";
      s.model_fingerprint = model.fingerprint_hex();
      return s;
    }
    if (has_integers(model) && !options.relax_integrality) {
      s.status = Status::NotSolved;
      s.engine = "r2hpdhg";
      s.message = "integer columns: branch-and-bound is not part of the MI355X backend (solve the LP relaxation with "
                  "relax_integrality)";
      s.model_fingerprint = model.fingerprint_hex();
      return s;
    }
    Options resolved = options;
    const bool was_auto = options.algorithm == Algorithm::Auto;
    if (was_auto) resolved.algorithm = Algorithm::R2hpdhg;  // a GPU request defaults to r2hpdhg
    set_host_threads(options.threads);
    s = solve_impl(model, resolved);
    gate(model, resolved, s);
    if (was_auto) {
      char buf[96];
      const double work = static_cast<double>(std::max(model.num_rows, 1)) * static_cast<double>(model.nnz());
      std::snprintf(buf, sizeof buf, "auto: r2hpdhg (rows*nnz = %.2g)", work);
      s.message = s.message.empty() ? buf : std::string(buf) + "; " + s.message;
    }
    if (s.engine.empty()) s.engine = resolved.algorithm == Algorithm::Pdlp ? "pdlp" : "r2hpdhg";
    return s;
  } catch (const std::bad_alloc&) {
    Solution s;
    s.status = Status::NotSolved;
    s.message = "out of memory";
    s.model_fingerprint = model.fingerprint_hex();
    return s;
  } catch (const std::exception& e) {
    Solution s;
    s.status = Status::NumericalError;
    s.message = std::string("internal error: ") + e.what();
    s.model_fingerprint = model.fingerprint_hex();
    return s;
  }
}

}  // namespace ps26119::amd
