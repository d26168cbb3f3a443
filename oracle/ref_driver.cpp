// ref_driver.cpp — our extern "C" harness around the REFERENCE solver core, compiled together
// with the reference's own sources by oracle/ref.mk into oracle/_ref/libps26119_ref.so.
// Test infrastructure only (see orc_abi.h). Every function calls the reference's public C++
// entry points; nothing here re-implements reference logic.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "core/safe_bound.h"
#include "io/lpm_reader.h"
#include "io/solution_writer.h"
#include "core/solution_checker.h"
#include "la/csr.h"
#include "pdhg/backend.h"
#include "pdhg/scaling.h"
#include "ps26119/batch.h"
#include "ps26119/solve.h"
#include "../oracle/orc_abi.h"

using namespace ps26119;

namespace {

Model to_model(const orc_lp* p) {
  Model m;
  m.num_rows = p->m;
  m.num_cols = p->n;
  m.sense = p->sense;
  m.obj_offset = p->offset;
  m.obj.assign(p->c, p->c + p->n);
  m.col_lower.assign(p->col_lower, p->col_lower + p->n);
  m.col_upper.assign(p->col_upper, p->col_upper + p->n);
  m.row_lower.assign(p->row_lower, p->row_lower + p->m);
  m.row_upper.assign(p->row_upper, p->row_upper + p->m);
  m.col_start.assign(p->col_start, p->col_start + p->n + 1);
  const int nnz = p->col_start[p->n];
  m.row_index.assign(p->row_index, p->row_index + nnz);
  m.value.assign(p->value, p->value + nnz);
  return m;
}

}  // namespace

extern "C" {

int ref_solve(const orc_lp* lp, const orc_opts* o, orc_result* r, double* x, double* y, double* z, double* dual_ray,
              double* primal_ray) {
  const Model m = to_model(lp);
  Options opt;
  opt.algorithm = o->algorithm == 2 ? Algorithm::Pdlp : Algorithm::R2hpdhg;
  opt.precision = o->precision ? Precision::Mixed : Precision::Fp64;
  opt.presolve = o->presolve != 0;
  opt.threads = o->threads;
  if (o->check_every > 0) opt.termination_check_every = o->check_every;
  if (o->tolerance > 0) opt.tolerance = o->tolerance;
  if (o->time_limit > 0) opt.time_limit = o->time_limit;
  if (o->iteration_limit > 0) opt.iteration_limit = o->iteration_limit;
  if (o->warm_x) opt.warm_x.assign(o->warm_x, o->warm_x + lp->n);
  if (o->warm_y) opt.warm_y.assign(o->warm_y, o->warm_y + lp->m);
  opt.warm_primal_weight = o->warm_primal_weight;
  opt.use_gpu = o->use_gpu != 0;
  for (int k = 0; k < o->n_params; ++k) opt.engine_params.emplace_back(o->param_names[k], o->param_values[k]);
  const Solution s = solve(m, opt);
  std::memset(r, 0, sizeof *r);
  r->status = static_cast<int>(s.status);
  r->iterations = s.iterations;
  r->iterations_to_fast = s.iterations_to_fast;
  r->objective = s.objective;
  r->dual_objective = s.dual_objective;
  r->primal_residual = s.primal_residual;
  r->dual_residual = s.dual_residual;
  r->gap = s.gap;
  r->certified_bound = s.certified_bound;
  r->primal_weight = s.primal_weight;
  r->seconds = s.seconds;
  r->setup_seconds = s.setup_seconds;
  r->check_primal = s.check_primal;
  r->check_dual = s.check_dual;
  r->check_gap = s.check_gap;
  r->check = s.check == "PASS" ? 1 : (s.check == "FAIL" ? 0 : -1);
  std::snprintf(r->engine, sizeof r->engine, "%s", s.engine.c_str());
  std::snprintf(r->message, sizeof r->message, "%s", s.message.c_str());
  auto put = [](double* dst, const std::vector<double>& v, int len) {
    if (dst && static_cast<int>(v.size()) == len) std::memcpy(dst, v.data(), sizeof(double) * len);
  };
  put(x, s.x, lp->n);
  put(y, s.y, lp->m);
  put(z, s.z, lp->n);
  put(dual_ray, s.dual_ray, lp->m);
  put(primal_ray, s.primal_ray, lp->n);
  return r->status;
}

// Scaling factors and scaled values of Ã (CSR order) and Ãᵀ (CSR order of the transpose).
// scalars = {bound_scale, obj_scale, log10_range, geometric_mean_applied}.
int ref_scale(const orc_lp* lp, int geo_iters, double geo_min_range, int ruiz, int pock_chambolle, int bor,
              double* row_scale, double* col_scale, double* scalars, double* a_vals, double* at_vals, double* c_scaled,
              double* bounds_scaled /* l(n) u(n) rl(m) ru(m) */) {
  const Model m = to_model(lp);
  pdhg::ScalingOptions so;
  so.geometric_mean_iterations = geo_iters;
  so.geometric_mean_min_log10_range = geo_min_range;
  so.ruiz_iterations = ruiz;
  so.pock_chambolle = pock_chambolle != 0;
  so.bound_objective_rescaling = bor != 0;
  const pdhg::ScaledProblem sp = pdhg::make_scaled_problem(m, so);
  std::memcpy(row_scale, sp.row_scale.data(), sizeof(double) * sp.m);
  std::memcpy(col_scale, sp.col_scale.data(), sizeof(double) * sp.n);
  scalars[0] = sp.bound_scale;
  scalars[1] = sp.obj_scale;
  scalars[2] = sp.log10_range;
  scalars[3] = sp.geometric_mean_applied ? 1.0 : 0.0;
  if (a_vals) std::memcpy(a_vals, sp.A.val.data(), sizeof(double) * sp.A.val.size());
  if (at_vals) std::memcpy(at_vals, sp.At.val.data(), sizeof(double) * sp.At.val.size());
  if (c_scaled) std::memcpy(c_scaled, sp.c.data(), sizeof(double) * sp.n);
  if (bounds_scaled) {
    std::memcpy(bounds_scaled, sp.col_lower.data(), sizeof(double) * sp.n);
    std::memcpy(bounds_scaled + sp.n, sp.col_upper.data(), sizeof(double) * sp.n);
    std::memcpy(bounds_scaled + 2 * sp.n, sp.row_lower.data(), sizeof(double) * sp.m);
    std::memcpy(bounds_scaled + 2 * sp.n + sp.m, sp.row_upper.data(), sizeof(double) * sp.m);
  }
  return 0;
}

// ‖Ã‖₂ estimate of the default-scaled problem (the step size is η = 0.998 / this).
double ref_norm2(const orc_lp* lp) {
  const Model m = to_model(lp);
  const pdhg::ScaledProblem sp = pdhg::make_scaled_problem(m, pdhg::ScalingOptions{});
  return la::estimate_norm2(sp.A, sp.At);
}

// KKT statistics on the original model of (x, y_min). out: primal_residual, dual_residual,
// primal_obj, dual_obj, b_norm, c_norm, primal_max_rel, dual_max_rel, rel_kkt, converged(1e-8).
int ref_kkt(const orc_lp* lp, const double* x, const double* ymin, double* out) {
  const Model m = to_model(lp);
  const la::Csr<double> A = la::csr_from_model(m), At = la::csr_transpose_from_model(m);
  const pdhg::LpView v{m.num_rows, m.num_cols, m.sense, m.obj.data(), m.col_lower.data(), m.col_upper.data(),
                       m.row_lower.data(), m.row_upper.data()};
  const pdhg::KktStats k = pdhg::kkt_general(A, At, v, std::vector<double>(x, x + m.num_cols),
                                             std::vector<double>(ymin, ymin + m.num_rows));
  const double vals[10] = {k.primal_residual, k.dual_residual, k.primal_obj, k.dual_obj, k.b_norm, k.c_norm,
                           k.primal_max_rel, k.dual_max_rel, k.rel_kkt(), k.converged(1e-8) ? 1.0 : 0.0};
  std::memcpy(out, vals, sizeof vals);
  return 0;
}

// The reference CPU backend driven through a fixed op sequence on the default-scaled problem:
// x ← sin(1+j), y ← cos(2+i); `iters` r²HPDHG steps (spmv_t, r2h_primal, spmv, r2h_dual with
// w = (k+1)/(k+2)); then primal_step, dual_step, axpby(0.3, xh, −1.7, xb).
// Outputs x, xh, xb (n) and y, yh, yb (m); red = {dot(x,xh), norm2(y), diff_norm2(x,x0)} then the
// 8 KKT fields of kkt(xh, yh).
int ref_backend_ops(const orc_lp* lp, int precision, int iters, double tau, double sigma, double rho, double* ox,
                    double* oxh, double* oxb, double* oy, double* oyh, double* oyb, double* red) {
  const Model m = to_model(lp);
  const pdhg::ScaledProblem sp = pdhg::make_scaled_problem(m, pdhg::ScalingOptions{});
  auto b = pdhg::make_cpu_backend();
  b->setup(sp);
  b->set_precision(precision ? Precision::Mixed : Precision::Fp64);
  const int x = b->create(pdhg::Space::Primal), x0 = b->create(pdhg::Space::Primal), xh = b->create(pdhg::Space::Primal);
  const int xb = b->create(pdhg::Space::Primal), g = b->create(pdhg::Space::Primal);
  const int y = b->create(pdhg::Space::Dual), y0 = b->create(pdhg::Space::Dual), yh = b->create(pdhg::Space::Dual);
  const int yb = b->create(pdhg::Space::Dual), ax = b->create(pdhg::Space::Dual);
  std::vector<double> xi(sp.n), yi(sp.m);
  for (int j = 0; j < sp.n; ++j) xi[j] = std::sin(1.0 + j);
  for (int i = 0; i < sp.m; ++i) yi[i] = std::cos(2.0 + i);
  b->upload(x, xi);
  b->upload(y, yi);
  b->copy(x0, x);
  b->copy(y0, y);
  for (int it = 0; it < iters; ++it) {
    const double w = (it + 1.0) / (it + 2.0);
    b->spmv_t(y, g);
    b->r2h_primal(x, x0, g, tau, w, rho, xh, xb);
    b->spmv(xb, ax);
    b->r2h_dual(y, y0, ax, sigma, w, rho, yh, yb);
  }
  b->primal_step(x, g, 0.5 * tau, xb);
  b->dual_step(y, ax, 1.25 * sigma, yb);
  b->axpby(0.3, xh, -1.7, xb);
  std::vector<double> v;
  const int ids[6] = {x, xh, xb, y, yh, yb};
  double* outs[6] = {ox, oxh, oxb, oy, oyh, oyb};
  for (int t = 0; t < 6; ++t) {
    b->download(ids[t], v);
    std::memcpy(outs[t], v.data(), sizeof(double) * v.size());
  }
  red[0] = b->dot(x, xh);
  red[1] = b->norm2(y);
  red[2] = b->diff_norm2(x, x0);
  const pdhg::KktStats k = b->kkt(xh, yh);
  const double kv[8] = {k.primal_residual, k.dual_residual, k.primal_obj, k.dual_obj,
                        k.b_norm,          k.c_norm,        k.primal_max_rel, k.dual_max_rel};
  std::memcpy(red + 3, kv, sizeof kv);
  return 0;
}

// Batched scenarios through the reference's solve_batch (src/pdhg/batch.cpp:129-441).
// results[k], xs[k·n ..], ys[k·m ..] per scenario.
int ref_solve_batch(const orc_lp* lp, int K, const orc_scenario* sc, const orc_opts* o, orc_result* results, double* xs,
                    double* ys) {
  const Model base = to_model(lp);
  std::vector<Scenario> scen(K);
  auto take = [](const double* p, int len) { return p ? std::vector<double>(p, p + len) : std::vector<double>{}; };
  for (int k = 0; k < K; ++k) {
    scen[k].obj = take(sc[k].obj, lp->n);
    scen[k].col_lower = take(sc[k].col_lower, lp->n);
    scen[k].col_upper = take(sc[k].col_upper, lp->n);
    scen[k].row_lower = take(sc[k].row_lower, lp->m);
    scen[k].row_upper = take(sc[k].row_upper, lp->m);
  }
  Options opt;
  opt.threads = o->threads;
  if (o->check_every > 0) opt.termination_check_every = o->check_every;
  if (o->tolerance > 0) opt.tolerance = o->tolerance;
  if (o->time_limit > 0) opt.time_limit = o->time_limit;
  if (o->iteration_limit > 0) opt.iteration_limit = o->iteration_limit;
  for (int k = 0; k < o->n_params; ++k) opt.engine_params.emplace_back(o->param_names[k], o->param_values[k]);
  const std::vector<Solution> out = solve_batch(base, scen, opt);
  for (int k = 0; k < K; ++k) {
    const Solution& s = out[k];
    orc_result* r = results + k;
    std::memset(r, 0, sizeof *r);
    r->status = static_cast<int>(s.status);
    r->iterations = s.iterations;
    r->iterations_to_fast = s.iterations_to_fast;
    r->objective = s.objective;
    r->dual_objective = s.dual_objective;
    r->primal_residual = s.primal_residual;
    r->dual_residual = s.dual_residual;
    r->gap = s.gap;
    r->certified_bound = s.certified_bound;
    r->primal_weight = s.primal_weight;
    r->seconds = s.seconds;
    r->check = -1;
    std::snprintf(r->message, sizeof r->message, "%s", s.message.c_str());
    if (xs && static_cast<int>(s.x.size()) == lp->n) std::memcpy(xs + static_cast<size_t>(k) * lp->n, s.x.data(), sizeof(double) * lp->n);
    if (ys && static_cast<int>(s.y.size()) == lp->m) std::memcpy(ys + static_cast<size_t>(k) * lp->m, s.y.data(), sizeof(double) * lp->m);
  }
  return 0;
}

double ref_certified_bound(const orc_lp* lp, const double* y, int use_implied) {
  const Model m = to_model(lp);
  return certified_dual_bound(m, std::vector<double>(y, y + lp->m), use_implied != 0).bound;
}

// out: max_primal_violation, max_dual_violation, relative_gap, primal_objective, dual_objective, passed.
int ref_check_solution(const orc_lp* lp, const double* x, const double* y, double* out) {
  const Model m = to_model(lp);
  const CheckReport r = check_solution(m, std::vector<double>(x, x + lp->n), std::vector<double>(y, y + lp->m));
  out[0] = r.max_primal_violation;
  out[1] = r.max_dual_violation;
  out[2] = r.relative_gap;
  out[3] = r.primal_objective;
  out[4] = r.dual_objective;
  out[5] = r.passed() ? 1.0 : 0.0;
  return 0;
}

// The reference's writer on the model read by the reference's .lpm reader and the given
// contents (the fingerprint is left for the writer to compute from the model). 0 = written,
// -1 = the model does not read, -2 = unknown status, -3 = write failed.
int ref_write_solution(const char* lpm_path, const orc_solution_file* f, const char* out_path) {
  Model m;
  if (!io::read_lpm(lpm_path, m).ok) return -1;
  Solution s;
  if (!status_from_string(f->status, s.status)) return -2;
  s.engine = f->engine;
  s.precision = f->precision;
  s.message = f->message;
  s.check = f->check;
  s.objective = f->objective;
  s.dual_objective = f->dual_objective;
  s.certified_bound = f->certified_bound;
  s.primal_residual = f->primal_residual;
  s.dual_residual = f->dual_residual;
  s.gap = f->gap;
  s.iterations = f->iterations;
  s.seconds = f->seconds;
  s.setup_seconds = f->setup_seconds;
  s.primal_weight = f->primal_weight;
  s.iterations_to_fast = f->iterations_to_fast;
  s.seconds_to_fast = f->seconds_to_fast;
  s.check_primal = f->check_primal;
  s.check_dual = f->check_dual;
  s.check_gap = f->check_gap;
  const int n = m.num_cols, mm = m.num_rows;
  if (f->x) s.x.assign(f->x, f->x + n);
  if (f->z) s.z.assign(f->z, f->z + n);
  if (f->activity) s.row_activity.assign(f->activity, f->activity + mm);
  if (f->y) s.y.assign(f->y, f->y + mm);
  if (f->dual_ray) s.dual_ray.assign(f->dual_ray, f->dual_ray + f->dual_ray_len);
  if (f->primal_ray) s.primal_ray.assign(f->primal_ray, f->primal_ray + f->primal_ray_len);
  std::string err;
  return io::write_solution(out_path, m, s, err) ? 0 : -3;
}

}  // extern "C"
