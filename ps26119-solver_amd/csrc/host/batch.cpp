// batch.cpp — scenarios sharing one matrix, solved in lockstep on the GPU (reference
// src/pdhg/batch.cpp:129-441; include/ps26119/batch.h:10-33). Decisions per scenario on the host
// (restarts, PID primal weight, termination, compaction) exactly as the reference; the
// iterations, residuals and KKT sums of all live scenarios on the device (psh_batch_*).
// Additions the reference does not have: the GPU, an optional warm start of every scenario
// from Options::warm_x / warm_y / warm_primal_weight (e.g. the base model's solution), and
// precision = mixed: every slot in increment form (fp32 matrix and increments around fp64
// anchors rebuilt at its restarts); the whole batch moves to fp64 with a restart of every slot
// when some scenario's increment form stalls (RefineGuard).
#include "batch.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <new>
#include <stdexcept>

#include "engine.h"
#include "la.h"
#include "verify.h"

namespace ps26119::amd {

namespace {

// A scenario's five vectors (original space and sense), the base model's where the scenario
// leaves one empty, with what keeps it from being solved: `defect` (a vector of the wrong
// length, a non-finite cost, a NaN bound) or `crossing` (a bound range that is empty, so the
// scenario is infeasible as given).
struct ScenarioVectors {
  std::vector<double> c, cl, cu, rl, ru;
  std::string defect, crossing;
};

ScenarioVectors overlay(const Model& base, const Scenario& s) {
  ScenarioVectors v;
  const std::size_t n = static_cast<std::size_t>(base.num_cols), m = static_cast<std::size_t>(base.num_rows);
  const struct {
    const std::vector<double>& given;
    const std::vector<double>& fallback;
    std::vector<double>& into;
    std::size_t len;
    const char* name;
  } parts[] = {{s.obj, base.obj, v.c, n, "objective"},
               {s.col_lower, base.col_lower, v.cl, n, "column lower bounds"},
               {s.col_upper, base.col_upper, v.cu, n, "column upper bounds"},
               {s.row_lower, base.row_lower, v.rl, m, "row lower bounds"},
               {s.row_upper, base.row_upper, v.ru, m, "row upper bounds"}};
  for (const auto& p : parts) {
    if (!p.given.empty() && p.given.size() != p.len) {
      v.defect = std::string("scenario ") + p.name + " have " + std::to_string(p.given.size()) + " entries, not " +
                 std::to_string(p.len);
      return v;
    }
    p.into = p.given.empty() ? p.fallback : p.given;
  }
  char buf[160];
  for (std::size_t j = 0; j < n; ++j)
    if (!std::isfinite(v.c[j]) || std::isnan(v.cl[j]) || std::isnan(v.cu[j])) {
      std::snprintf(buf, sizeof buf, "scenario column %zu has a non-finite cost or a NaN bound", j);
      v.defect = buf;
      return v;
    }
  for (std::size_t i = 0; i < m; ++i)
    if (std::isnan(v.rl[i]) || std::isnan(v.ru[i])) {
      std::snprintf(buf, sizeof buf, "scenario row %zu has a NaN bound", i);
      v.defect = buf;
      return v;
    }
  for (std::size_t j = 0; j < n && v.crossing.empty(); ++j)
    if (v.cl[j] > v.cu[j]) std::snprintf(buf, sizeof buf, "column %zu of the scenario", j), v.crossing = buf;
  for (std::size_t i = 0; i < m && v.crossing.empty(); ++i)
    if (v.rl[i] > v.ru[i]) std::snprintf(buf, sizeof buf, "row %zu of the scenario", i), v.crossing = buf;
  return v;
}

// Why the whole batch cannot run (every scenario then gets NotSolved with it), or "".
std::string batch_refusal(const Model& base, const Options& options, EngineOptions& eo) {
  if (auto err = base.validate(); !err.empty()) return "the shared model is invalid: " + err;
  if (auto bad = apply_engine_params(options, eo); !bad.empty()) return "unknown engine parameter '" + bad + "'";
  if (eo.max_epoch_gap_only == 2) return "hip_max_epoch_gap_only = 2 is implemented for single solves only";
  if (!options.use_gpu) return "the MI355X backend serves GPU solves: set use_gpu = 1";
  if ((!options.warm_x.empty() && static_cast<int>(options.warm_x.size()) != base.num_cols) ||
      (!options.warm_y.empty() && static_cast<int>(options.warm_y.size()) != base.num_rows))
    return "a warm start vector does not match the shared model's size";
  return {};
}

}  // namespace

std::vector<Solution> solve_batch(const Model& base, const std::vector<Scenario>& scenarios, const Options& options) {
  const auto t0 = std::chrono::steady_clock::now();
  auto elapsed = [&] { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); };
  const int K = static_cast<int>(scenarios.size());
  std::vector<Solution> out(K);
  const int m = base.num_rows, n = base.num_cols;
  {
    Model w = base;  // fingerprint of the model each scenario solves
    for (int k = 0; k < K; ++k) {
      const Scenario& s = scenarios[k];
      w.obj = s.obj.empty() ? base.obj : s.obj;
      w.col_lower = s.col_lower.empty() ? base.col_lower : s.col_lower;
      w.col_upper = s.col_upper.empty() ? base.col_upper : s.col_upper;
      w.row_lower = s.row_lower.empty() ? base.row_lower : s.row_lower;
      w.row_upper = s.row_upper.empty() ? base.row_upper : s.row_upper;
      out[k].engine = "r2hpdhg-batch";
      out[k].precision = options.precision == Precision::Mixed ? "mixed" : "fp64";
      out[k].model_fingerprint = w.fingerprint_hex();
    }
  }
  if (K == 0) return out;
  EngineOptions eo = engine_options_from(options);
  if (const std::string why = batch_refusal(base, options, eo); !why.empty()) {
    for (auto& s : out) s.status = Status::NotSolved, s.message = why;
    return out;
  }
  const std::vector<double>& warm_x = options.warm_x;
  const std::vector<double>& warm_y = options.warm_y;
  set_host_threads(options.threads);

  std::vector<ScenarioVectors> D(K);
  for (int k = 0; k < K; ++k) {
    D[k] = overlay(base, scenarios[k]);
    if (!D[k].defect.empty()) {
      out[k].status = Status::NotSolved;
      out[k].message = D[k].defect;
    } else if (!D[k].crossing.empty()) {
      out[k].status = Status::Infeasible;
      out[k].message = "bounds cross: " + D[k].crossing + " has a lower bound above its upper bound";
    }
  }

  try {
    ScalingOptions so = eo.scaling;
    so.bound_objective_rescaling = false;  // per scenario below (reference batch.cpp:182-221)
    std::string err;
    std::unique_ptr<Device> dev = Device::open(eo.device, err);
    if (!dev) {
      for (auto& r : out) r.status = Status::NotSolved, r.message = err;
      return out;
    }
    ScaledProblem sp;
    scale_and_load(*dev, base, so, eo.device_scaling != 0, sp);
    psh_ctx* c = dev->ctx();
    if (eo.timing) dev->check(psh_timing(c, 1), "timing");
    dev->check(psh_graphs(c, eo.graphs), "graphs");
    const bool empty = n == 0 || sp.A.nnz() == 0;
    const double norm = empty ? 0.0 : estimate_norm2(*dev, n, eo.cpu_order ? PSH_RED_CPU : PSH_RED_TREE);
    const double eta = norm > 0 ? 0.998 / norm : 1.0;
    const double setup_shared = elapsed();
    const int mode = eo.cpu_order ? PSH_RED_CPU : PSH_RED_TREE;
    const double rho = eo.reflection;
    const int Kc = eo.check_every;

    std::vector<int> valid;
    for (int k = 0; k < K; ++k)
      if (out[k].message.empty() && out[k].status == Status::NotSolved) valid.push_back(k);
    // groups of at most PSH_BATCH_MAX_W scenarios share one device batch
    for (std::size_t g0 = 0; g0 < valid.size(); g0 += PSH_BATCH_MAX_W) {
      std::vector<int> sid(valid.begin() + g0, valid.begin() + std::min(valid.size(), g0 + PSH_BATCH_MAX_W));
      int W = static_cast<int>(sid.size());
      std::vector<double> bs(W, 1.0), os(W, 1.0);
      const std::size_t NW = static_cast<std::size_t>(n) * W, MW = static_cast<std::size_t>(m) * W;
      std::vector<double> C(NW), L(NW), U(NW), RL(MW), RU(MW), CM(NW), CLO(NW), CUO(NW), RLO(MW), RUO(MW);
      for (int t = 0; t < W; ++t) {
        const ScenarioVectors& d = D[sid[t]];
        if (eo.scaling.bound_objective_rescaling) {
          double b2 = 0, c2 = 0;
          for (int i = 0; i < m; ++i) {
            const double lo = std::isfinite(d.rl[i]) ? std::fabs(d.rl[i]) : 0.0;
            const double up = std::isfinite(d.ru[i]) ? std::fabs(d.ru[i]) : 0.0;
            const double b = sp.row_scale[i] * std::max(lo, up);
            b2 += b * b;
          }
          for (int j = 0; j < n; ++j) {
            const double cc = sp.col_scale[j] * d.c[j];
            c2 += cc * cc;
          }
          bs[t] = 1.0 / (std::sqrt(b2) + 1.0);
          os[t] = 1.0 / (std::sqrt(c2) + 1.0);
        }
        for (int j = 0; j < n; ++j) {
          const std::size_t q = static_cast<std::size_t>(j) * W + t;
          C[q] = os[t] * sp.col_scale[j] * base.sense * d.c[j];
          L[q] = bs[t] * d.cl[j] / sp.col_scale[j];
          U[q] = bs[t] * d.cu[j] / sp.col_scale[j];
          CM[q] = base.sense * d.c[j];
          CLO[q] = d.cl[j];
          CUO[q] = d.cu[j];
        }
        for (int i = 0; i < m; ++i) {
          const std::size_t q = static_cast<std::size_t>(i) * W + t;
          RL[q] = bs[t] * sp.row_scale[i] * d.rl[i];
          RU[q] = bs[t] * sp.row_scale[i] * d.ru[i];
          RLO[q] = d.rl[i];
          RUO[q] = d.ru[i];
        }
      }
      dev->check(psh_batch_begin(c, W, C.data(), L.data(), U.data(), RL.data(), RU.data(), CM.data(), CLO.data(),
                                 CUO.data(), RLO.data(), RUO.data(), bs.data(), os.data()),
                 "batch");
      // initial iterate: x = proj(0), y = 0 (reference batch.cpp:238-242), or the warm point
      std::vector<double> X(NW), Y(MW, 0.0);
      for (int t = 0; t < W; ++t) {
        const ScenarioVectors& d = D[sid[t]];
        for (int j = 0; j < n; ++j) {
          const std::size_t q = static_cast<std::size_t>(j) * W + t;
          if (!warm_x.empty()) {
            const double xo = std::min(std::max(warm_x[j], d.cl[j]), d.cu[j]);
            X[q] = bs[t] * xo / sp.col_scale[j];
          } else {
            X[q] = 0.0 < L[q] ? L[q] : (0.0 > U[q] ? U[q] : 0.0);
          }
        }
        if (!warm_y.empty())
          for (int i = 0; i < m; ++i) {
            double v = base.sense * warm_y[i];
            if (!std::isfinite(d.rl[i])) v = std::min(v, 0.0);
            if (!std::isfinite(d.ru[i])) v = std::max(v, 0.0);
            Y[static_cast<std::size_t>(i) * W + t] = os[t] * v / sp.row_scale[i];
          }
      }
      dev->check(psh_batch_set_iterate(c, X.data(), Y.data()), "batch");
      bool incr = options.precision == Precision::Mixed && eo.refine != 0;
      if (incr) dev->check(psh_batch_delta(c, 1), "batch precision");
      std::vector<RefineGuard> guard(K);
      std::vector<std::int64_t> promoted_at(K, -1);

      const double w0 = options.warm_primal_weight > 0 ? options.warm_primal_weight : 1.0;
      std::vector<PidState> pid(K);
      for (auto& q : pid) q.omega = q.best_omega = w0;
      std::vector<double> r0(K, -1.0), r_last(K, kInf), fast_s(K, -1.0);
      std::vector<std::int64_t> inner(K, 0), fast_it(K, -1);
      std::int64_t it = 0;
      double last_log = 0.0;
      std::vector<double> tau, sigma, stats, x, ym;
      std::vector<long long> in0;
      std::vector<int> mask;
      const int SC = PSH_BSTAT_COUNT;
      while (W > 0) {
        tau.resize(W);
        sigma.resize(W);
        in0.resize(W);
        for (int t = 0; t < W; ++t) {
          tau[t] = eta / pid[sid[t]].omega;
          sigma[t] = eta * pid[sid[t]].omega;
          in0[t] = inner[sid[t]];
        }
        dev->check(psh_batch_chunk(c, Kc, tau.data(), sigma.data(), in0.data(), rho, mode), "batch iteration");
        for (int t = 0; t < W; ++t) inner[sid[t]] += Kc;
        it += Kc;
        stats.resize(static_cast<std::size_t>(W) * SC);
        const bool want_ray = it >= 4 * Kc;  // as the single solve (reference r2hpdhg.cpp:93-111)
        dev->check(psh_batch_check(c, mode, want_ray ? 1 : 0, stats.data()), "batch check");
        const bool over_iter = it >= eo.iteration_limit, over_time = elapsed() > eo.time_limit;
        std::vector<int> keep;
        mask.assign(W, 0);
        bool promote = false;  // increment form stalled for some scenario: the batch goes to fp64
        for (int t = 0; t < W; ++t) {
          const int k = sid[t];
          const double* st = stats.data() + static_cast<std::size_t>(t) * SC;
          const double om = pid[k].omega;
          auto resid = [&](double px, double py, double cr) { return std::sqrt(std::max(om * px + py / om + 2.0 * eta * cr, 0.0)); };
          if (in0[t] == 0) r0[k] = resid(st[PSH_BSTAT_R0_PX], st[PSH_BSTAT_R0_PY], st[PSH_BSTAT_R0_CROSS]);
          KktStats kk;
          kk.primal_residual = std::sqrt(st[PSH_BSTAT_RP2]);
          kk.dual_residual = std::sqrt(st[PSH_BSTAT_RD2]);
          kk.primal_obj = st[PSH_BSTAT_P];
          kk.dual_obj = st[PSH_BSTAT_D];
          kk.b_norm = std::sqrt(st[PSH_BSTAT_B2]);
          kk.c_norm = std::sqrt(st[PSH_BSTAT_C2]);
          kk.primal_max_rel = st[PSH_BSTAT_PMAX];
          kk.dual_max_rel = st[PSH_BSTAT_DMAX] / (1.0 + st[PSH_BSTAT_CINF]);
          auto finish = [&](Status status, const std::string& msg) {
            Solution& s = out[k];
            x.assign(n, 0.0);
            ym.assign(m, 0.0);
            dev->check(psh_batch_point(c, t, x.data(), ym.data()), "batch point");
            s.status = status;
            s.message = msg;
            s.x = x;
            s.y.resize(m);
            for (int i = 0; i < m; ++i) s.y[i] = base.sense * ym[i];
            s.z = D[k].c;
            parallel_for(n, [&](std::int64_t b, std::int64_t e) {
              for (std::int64_t j = b; j < e; ++j)
                for (int p = base.col_start[j]; p < base.col_start[j + 1]; ++p) s.z[j] -= base.value[p] * s.y[base.row_index[p]];
            });
            s.row_activity = base.row_activity(x);
            double obj = base.obj_offset;
            for (int j = 0; j < n; ++j) obj += D[k].c[j] * x[j];
            s.objective = obj;
            s.dual_objective = base.sense * kk.dual_obj + base.obj_offset;
            s.primal_residual = kk.rel_primal();
            s.dual_residual = kk.rel_dual();
            s.gap = kk.rel_gap();
            s.iterations = it;
            s.seconds = elapsed();
            s.setup_seconds = setup_shared;
            s.iterations_to_fast = fast_it[k];
            s.seconds_to_fast = fast_s[k];
            s.primal_weight = pid[k].omega;
            if (options.precision == Precision::Mixed && eo.refine != 0) {
              s.message = s.message.empty() ? "fp32 increments on fp64 anchors" : s.message;
              if (promoted_at[k] >= 0)
                s.message += "; fp32 -> fp64 at iteration " + std::to_string(promoted_at[k]);
            }
          };
          if (!kk.finite()) {
            finish(Status::NumericalError, "non-finite iterate");
            continue;
          }
          if (fast_it[k] < 0 && kk.rel_kkt() <= tol::kFirstOrderFast) fast_it[k] = it, fast_s[k] = elapsed();
          if (incr && guard[k].stalled(kk.rel_kkt(), it, eo)) promote = true;
          if (kk.converged(eo.tolerance)) {
            finish(Status::Optimal, "");
            continue;
          }
          if (over_iter || over_time) {
            finish(over_iter ? Status::IterationLimit : Status::TimeLimit, "");
            continue;
          }
          if (want_ray) {  // per-scenario ray tests, each certificate re-verified on its own model
            psh_ray ry;
            std::memcpy(&ry, st + PSH_BSTAT_RAY_RMAX, sizeof ry);
            double robj = 0, rviol = 0;
            const Status v = ray_decision(ry, robj, rviol);
            if (v != Status::NotSolved) {
              Model w = base;
              w.obj = D[k].c;
              w.col_lower = D[k].cl;
              w.col_upper = D[k].cu;
              w.row_lower = D[k].rl;
              w.row_upper = D[k].ru;
              std::vector<double> dr, pr;
              if (v == Status::Infeasible) dr.assign(m, 0.0);
              else pr.assign(n, 0.0);
              dev->check(psh_batch_ray_vectors(c, t, v == Status::Infeasible ? dr.data() : nullptr,
                                               v == Status::Unbounded ? pr.data() : nullptr),
                         "batch ray vectors");
              bool ok = false;
              char buf[200];
              if (v == Status::Infeasible) {
                ok = check_infeasibility_certificate(w, dr).passed;
                std::snprintf(buf, sizeof buf, "primal infeasible: dual ray certificate (objective %f, violation %f)",
                              robj, rviol);
              } else {
                x.assign(n, 0.0);
                ym.assign(m, 0.0);
                dev->check(psh_batch_point(c, t, x.data(), ym.data()), "batch point");
                ok = check_unboundedness_certificate(w, x, pr).passed;
                std::snprintf(buf, sizeof buf, "unbounded: primal ray certificate (c'd %f, violation %f)", robj, rviol);
              }
              if (ok) {
                finish(v, buf);
                Solution& so = out[k];
                so.objective = so.dual_objective = std::numeric_limits<double>::quiet_NaN();
                if (v == Status::Infeasible) so.dual_ray = std::move(dr);
                else so.primal_ray = std::move(pr);
                continue;
              }
            }
          }
          keep.push_back(t);
          const double r = resid(st[PSH_BSTAT_PX], st[PSH_BSTAT_PY], st[PSH_BSTAT_CROSS]);
          const bool restart = restart_due(eo, r, r0[k], r_last[k], inner[k], it) || increment_epoch_full(eo, incr, inner[k]) ||
                               epoch_cap_reached(eo, inner[k], kk);
          r_last[k] = r;
          if (!restart) continue;
          pid_update(pid[k], eo, std::sqrt(st[PSH_BSTAT_PIDX]), std::sqrt(st[PSH_BSTAT_PIDY]), kk.rel_primal(),
                     kk.rel_dual());
          mask[t] = 1;
          inner[k] = 0;
          r_last[k] = kInf;
        }
        if (promote) {  // every live slot restarts at its x̂, ŷ in fp64 (PID as at a restart)
          dev->check(psh_batch_delta(c, 0), "batch precision");
          incr = false;
          for (int t : keep) {
            const int k = sid[t];
            if (!mask[t]) {
              const double* st = stats.data() + static_cast<std::size_t>(t) * SC;
              KktStats kk;
              kk.primal_residual = std::sqrt(st[PSH_BSTAT_RP2]);
              kk.dual_residual = std::sqrt(st[PSH_BSTAT_RD2]);
              kk.primal_obj = st[PSH_BSTAT_P];
              kk.dual_obj = st[PSH_BSTAT_D];
              kk.b_norm = std::sqrt(st[PSH_BSTAT_B2]);
              kk.c_norm = std::sqrt(st[PSH_BSTAT_C2]);
              pid_update(pid[k], eo, std::sqrt(st[PSH_BSTAT_PIDX]), std::sqrt(st[PSH_BSTAT_PIDY]), kk.rel_primal(),
                         kk.rel_dual());
              mask[t] = 1;
              inner[k] = 0;
              r_last[k] = kInf;
            }
            promoted_at[k] = it;
          }
        }
        dev->check(psh_batch_restart(c, mask.data()), "batch restart");
        if (static_cast<int>(keep.size()) != W) {
          dev->check(psh_batch_compact(c, keep.data(), static_cast<int>(keep.size())), "batch compaction");
          std::vector<int> ns;
          for (int t : keep) ns.push_back(sid[t]);
          sid.swap(ns);
          W = static_cast<int>(sid.size());
        }
        if (eo.verbosity >= 2 || (eo.verbosity == 1 && elapsed() - last_log >= 1.0)) {  // 1: at most once a second
          last_log = elapsed();
          std::fprintf(stderr, "batch it %lld  t %.2fs  running %d/%d\n", static_cast<long long>(it), last_log, W, K);
        }
      }
    }
  } catch (const std::bad_alloc&) {
    for (auto& s : out)
      if (s.status == Status::NotSolved && s.message.empty()) s.message = "out of memory";
  } catch (const std::exception& e) {
    for (auto& s : out)
      if (s.status == Status::NotSolved && s.message.empty()) s.status = Status::NumericalError, s.message = e.what();
  }
  return out;
}

}  // namespace ps26119::amd
