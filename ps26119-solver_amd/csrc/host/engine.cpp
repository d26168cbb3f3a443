// engine.cpp — r²HPDHG on the MI355X. The driver follows the reference loop decision for
// decision (src/pdhg/r2hpdhg.cpp:11-162, src/pdhg/engine.cpp:56-307); what changes is where the
// work runs: each block of K iterations is one psh_chunk (2 fused kernels per iteration, no host
// round trip), and each check is one psh_check (residual, KKT, ray-test scalars and PID
// distances in one device pass, one ~350-byte copy back). Only decisions run on the host.
#include "engine.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <new>
#include <stdexcept>

#include "la.h"

namespace ps26119::amd {

// ------------------------------------------------------------------ options
EngineOptions engine_options_from(const Options& o) {
  EngineOptions e;
  e.tolerance = o.tolerance;
  e.time_limit = o.time_limit;
  e.iteration_limit = o.iteration_limit;
  e.check_every = o.termination_check_every > 0 ? o.termination_check_every : 64;
  e.precision = o.precision;
  e.verbosity = o.verbosity;
  e.scaling.ruiz_iterations = o.ruiz_iterations;
  e.scaling.pock_chambolle = o.pock_chambolle;
  e.warm_x = o.warm_x;
  e.warm_y = o.warm_y;
  e.warm_primal_weight = o.warm_primal_weight;
  e.kkt_b_norm = o.kkt_b_norm;
  e.kkt_c_norm = o.kkt_c_norm;
  e.kkt_obj_shift = o.kkt_obj_shift;
  return e;
}

std::string apply_engine_params(const Options& o, EngineOptions& e) {  // reference engine.cpp:34-53
  for (const auto& [name, v] : o.engine_params) {
    if (name == "reflection") e.reflection = v;
    else if (name == "restart_sufficient") e.restart_sufficient = v;
    else if (name == "restart_necessary") e.restart_necessary = v;
    else if (name == "restart_artificial") e.restart_artificial = v;
    else if (name == "pid_kp") e.pid_kp = v;
    else if (name == "pid_ki") e.pid_ki = v;
    else if (name == "pid_kd") e.pid_kd = v;
    else if (name == "pid_integral_decay") e.pid_integral_decay = v;
    else if (name == "pid_max_log_step") e.pid_max_log_step = v;
    else if (name == "pdlp_restart_necessary") e.pdlp_restart_necessary = v;
    else if (name == "pdlp_primal_weight_smoothing") e.pdlp_primal_weight_smoothing = v;
    else if (name == "bound_objective_rescaling") e.scaling.bound_objective_rescaling = v != 0;
    else if (name == "geometric_mean_iterations") e.scaling.geometric_mean_iterations = static_cast<int>(v);
    else if (name == "geometric_mean_min_log10_range") e.scaling.geometric_mean_min_log10_range = v;
    else if (name == "ruiz_iterations") e.scaling.ruiz_iterations = static_cast<int>(v);
    else if (name == "pock_chambolle") e.scaling.pock_chambolle = v != 0;
    else if (name == "hip_cpu_order") e.cpu_order = v != 0 ? 1 : 0;
    else if (name == "hip_tree_reductions") e.tree = v != 0 ? 1 : 0;
    else if (name == "hip_device") e.device = static_cast<int>(v);
    else if (name == "hip_timing") e.timing = v != 0 ? 1 : 0;
    else if (name == "hip_graphs") e.graphs = v != 0 ? 1 : 0;
    else if (name == "hip_sliced") e.sliced = v >= 2 ? 2 : (v != 0 ? 1 : 0);
    else if (name == "hip_device_scaling") e.device_scaling = v != 0 ? 1 : 0;
    else if (name == "hip_refine") e.refine = v != 0 ? 1 : 0;
    else if (name == "hip_refine_patience") e.refine_patience = std::max(1, static_cast<int>(v));
    else if (name == "hip_bf16_until") e.bf16_until = v > 0 ? v : 0.0;
    else if (name == "hip_refine_max_epoch") e.refine_max_epoch = v > 0 ? static_cast<std::int64_t>(v) : 0;
    else if (name == "hip_max_epoch") e.max_epoch = v > 0 ? static_cast<std::int64_t>(v) : 0;
    else if (name == "hip_max_epoch_gap_only") e.max_epoch_gap_only = v >= 2 ? 2 : (v != 0 ? 1 : 0);
    else return name;
  }
  if (e.refine < 0) e.refine = e.cpu_order ? 0 : 1;
  return {};
}

int reduction_mode(const EngineOptions& o) {
  return o.cpu_order ? PSH_RED_CPU : (o.tree ? PSH_RED_TREE : PSH_RED_DECISION);
}

// ------------------------------------------------------------------ KKT
double KktStats::rel_gap() const {
  return std::fabs(primal_obj - dual_obj) / (1.0 + std::fabs(primal_obj + obj_shift) + std::fabs(dual_obj + obj_shift));
}
double KktStats::rel_kkt() const { return std::max({rel_primal(), rel_dual(), rel_gap()}); }
bool KktStats::converged(double eps) const {
  if (rel_kkt() > eps) return false;
  if (eps >= tol::kVerifyPrimal) return true;
  return primal_max_rel <= tol::kVerifyPrimal && dual_max_rel <= tol::kVerifyDual && rel_gap() <= tol::kVerifyGap;
}
bool KktStats::finite() const {
  return std::isfinite(primal_residual) && std::isfinite(dual_residual) && std::isfinite(primal_obj) && std::isfinite(dual_obj);
}

KktStats kkt_on_original(const ScaledProblem& sp, const std::vector<double>& x, const std::vector<double>& y) {
  const Model& M = *sp.original;
  const int m = sp.m, n = sp.n;
  const double sense = M.sense;
  // A·x by the model's columns: every row's products are added in increasing column order, the
  // order of a row-major product over Ã's layout (same bits, no copy of the matrix); Aᵀy by
  // columns in storage order
  const std::vector<double> ax = M.row_activity(x);
  std::vector<double> aty(n);
  column_products(M, y.data(), aty.data());
  double rp2 = 0, rd2 = 0, b2 = 0, c2 = 0, p = 0, d = 0, cinf = 0, pmax = 0, dmax = 0;
  auto elem_viol = [](double v, double lo, double up) {
    if (v < lo) return (lo - v) / (1.0 + std::fabs(lo));
    if (v > up) return (v - up) / (1.0 + std::fabs(up));
    return 0.0;
  };
  for (int i = 0; i < m; ++i) {
    const double lo = M.row_lower[i], up = M.row_upper[i];
    const double viol = ax[i] < lo ? lo - ax[i] : (ax[i] > up ? ax[i] - up : 0.0);
    rp2 += viol * viol;
    pmax = std::max(pmax, elem_viol(ax[i], lo, up));
    const double bl = std::isfinite(lo) ? std::fabs(lo) : 0.0, bu = std::isfinite(up) ? std::fabs(up) : 0.0;
    b2 += std::max(bl, bu) * std::max(bl, bu);
    if (y[i] > 0) {
      if (std::isfinite(lo)) d += y[i] * lo;
      else rd2 += y[i] * y[i], dmax = std::max(dmax, y[i]);
    } else if (y[i] < 0) {
      if (std::isfinite(up)) d += y[i] * up;
      else rd2 += y[i] * y[i], dmax = std::max(dmax, -y[i]);
    }
  }
  for (int j = 0; j < n; ++j) {
    const double cj = sense * M.obj[j];
    c2 += cj * cj;
    cinf = std::max(cinf, std::fabs(cj));
    p += cj * x[j];
    pmax = std::max(pmax, elem_viol(x[j], M.col_lower[j], M.col_upper[j]));
    const double lam = cj - aty[j];
    const double lo = M.col_lower[j], up = M.col_upper[j];
    double z;
    if (std::isfinite(lo) && std::isfinite(up)) z = lam;
    else if (std::isfinite(lo)) z = std::max(lam, 0.0);
    else if (std::isfinite(up)) z = std::min(lam, 0.0);
    else z = 0.0;
    rd2 += (lam - z) * (lam - z);
    dmax = std::max(dmax, std::fabs(lam - z));
    if (z > 0) d += z * lo;
    else if (z < 0) d += z * up;
  }
  KktStats k;
  k.primal_residual = std::sqrt(rp2);
  k.dual_residual = std::sqrt(rd2);
  k.primal_obj = p;
  k.dual_obj = d;
  k.b_norm = std::sqrt(b2);
  k.c_norm = std::sqrt(c2);
  k.primal_max_rel = pmax;
  k.dual_max_rel = dmax / (1.0 + cinf);
  return k;
}

static KktStats from_device(const psh_kkt& d) {
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

// ------------------------------------------------------------------ device
std::unique_ptr<Device> Device::open(int device, std::string& error) {
  psh_ctx* c = nullptr;
  const int rc = psh_create(device, &c);
  if (rc != PSH_OK) {
    error = rc == PSH_E_NO_DEVICE ? "HIP backend unavailable: no usable MI355X/gfx950 device is visible"
                                  : "HIP backend unavailable: device context creation failed";
    return nullptr;
  }
  return std::unique_ptr<Device>(new Device(c));
}

Device::~Device() { psh_destroy(ctx_); }

void Device::check(int rc, const char* what) const {
  if (rc == PSH_OK) return;
  if (rc == PSH_E_OOM) throw std::bad_alloc();
  throw std::runtime_error(std::string("HIP backend: ") + what + ": " + psh_error(ctx_));
}

void Device::load(const ScaledProblem& sp) {
  std::vector<double> cmin;
  const psh_problem p = problem_of(sp, cmin);
  check(psh_load(ctx_, &p), "upload");
}

psh_problem problem_of(const ScaledProblem& sp, std::vector<double>& cmin) {
  const Model& M = *sp.original;
  cmin.resize(sp.n);
  for (int j = 0; j < sp.n; ++j) cmin[j] = M.sense * M.obj[j];
  psh_problem p{};
  p.m = sp.m;
  p.n = sp.n;
  p.nnz = sp.A.nnz();
  p.a_ptr = sp.A.ptr.data();
  p.a_col = sp.A.col.data();
  p.a_val = sp.A.val.data();
  p.a_val_orig = sp.a_orig.data();
  p.at_ptr = sp.At.ptr.data();
  p.at_col = sp.At.col.data();
  p.ato_col = M.row_index.data();
  p.at_val = sp.At.val.data();
  p.ato_val = M.value.data();
  p.c = sp.c.data();
  p.l = sp.col_lower.data();
  p.u = sp.col_upper.data();
  p.rl = sp.row_lower.data();
  p.ru = sp.row_upper.data();
  p.row_scale = sp.row_scale.data();
  p.col_scale = sp.col_scale.data();
  p.bound_scale = sp.bound_scale;
  p.obj_scale = sp.obj_scale;
  p.cl_o = M.col_lower.data();
  p.cu_o = M.col_upper.data();
  p.rl_o = M.row_lower.data();
  p.ru_o = M.row_upper.data();
  p.c_min = cmin.data();
  return p;
}

void scale_and_load(Device& dev, const Model& model, const ScalingOptions& so, bool on_device, ScaledProblem& sp) {
  if (!on_device) {
    sp = make_scaled_problem(model, so);
    dev.load(sp);
    return;
  }
  sp = make_unscaled_problem(model, so);
  dev.load(sp);
  dev.check(psh_scale(dev.ctx(), sp.geometric_mean_applied ? so.geometric_mean_iterations : 0, so.ruiz_iterations,
                      so.pock_chambolle ? 1 : 0, sp.row_scale.data(), sp.col_scale.data()),
            "scaling");
  finish_scaled_problem(sp, model, so);
  dev.check(psh_set_vectors(dev.ctx(), sp.c.data(), sp.col_lower.data(), sp.col_upper.data(), sp.row_lower.data(),
                            sp.row_upper.data(), sp.bound_scale, sp.obj_scale),
            "scaling");
  std::vector<double>().swap(sp.A.val);  // the scaled values live on the device only
  std::vector<double>().swap(sp.At.val);
}

Status ray_decision(const psh_ray& t, double& obj, double& viol) {
  if (t.rmax > 0) {
    const double scale = std::max(t.rmax, t.dual_viol);
    obj = t.dual_obj / scale;
    viol = t.dual_viol / scale;
    if (obj > 0 && t.dual_obj > tol::kFirstOrderInfeasible * t.dual_mag && viol <= tol::kFirstOrderInfeasible * obj)
      return Status::Infeasible;
  }
  if (t.dmax > 0) {
    const double scale = std::max(t.dmax, t.primal_viol);
    obj = t.primal_cd / scale;
    viol = t.primal_viol / scale;
    if (obj < 0 && -t.primal_cd > tol::kFirstOrderInfeasible * t.primal_cmag && viol <= tol::kFirstOrderInfeasible * -obj)
      return Status::Unbounded;
  }
  return Status::NotSolved;
}

bool restart_due(const EngineOptions& o, double r, double r0, double r_last, std::int64_t inner, std::int64_t it) {
  const bool ratios_ok = r0 > 0 && std::isfinite(r0) && r > 0;
  return (ratios_ok && (r <= o.restart_sufficient * r0 || (r <= o.restart_necessary * r0 && r > r_last))) ||
         static_cast<double>(inner) >= o.restart_artificial * static_cast<double>(it);
}

void pid_update(PidState& p, const EngineOptions& o, double dxn, double dyn, double rel_primal, double rel_dual) {
  const bool both_pos = rel_dual > 0 && rel_primal > 0;
  const double ratio = both_pos ? rel_dual / rel_primal : 1.0;
  if (dxn > 1e-16 && dyn > 1e-16 && dxn < 1e12 && dyn < 1e12 && ratio > 1e-8 && ratio < 1e8) {
    const double e = std::log(dyn) - std::log(dxn) - std::log(p.omega);
    p.integral = o.pid_integral_decay * p.integral + e;
    const double step = o.pid_kp * e + o.pid_ki * p.integral + o.pid_kd * (e - p.last);
    p.omega *= std::exp(std::clamp(step, -o.pid_max_log_step, o.pid_max_log_step));
    p.last = e;
  } else {
    p.omega = p.best_omega;
    p.integral = p.last = 0.0;
  }
  const double balance = std::fabs(std::log10(ratio));
  if (both_pos && balance < p.best_balance) {
    p.best_balance = balance;
    p.best_omega = p.omega;
  }
}

// ------------------------------------------------------------------ ‖Ã‖₂
namespace {

// Largest eigenvalue of the symmetric tridiagonal (a, b) by Sturm-count bisection
// (reference src/la/csr.cpp:44-69).
double tridiag_max_eig(const std::vector<double>& a, const std::vector<double>& b, int k) {
  double lo = 1e300, hi = -1e300;
  for (int i = 0; i < k; ++i) {
    const double r = (i > 0 ? std::fabs(b[i - 1]) : 0.0) + (i + 1 < k ? std::fabs(b[i]) : 0.0);
    lo = std::min(lo, a[i] - r);
    hi = std::max(hi, a[i] + r);
  }
  auto count_below = [&](double x) {
    int c = 0;
    double d = 1.0;
    for (int i = 0; i < k; ++i) {
      const double bb = i > 0 ? b[i - 1] * b[i - 1] : 0.0;
      d = (a[i] - x) - (i > 0 ? bb / d : 0.0);
      if (d == 0.0) d = -1e-300;
      if (d < 0) ++c;
    }
    return c;
  };
  for (int it = 0; it < 200 && hi - lo > 1e-15 * std::max(1.0, std::fabs(hi)); ++it) {
    const double mid = 0.5 * (lo + hi);
    if (count_below(mid) == k) hi = mid;
    else lo = mid;
  }
  return hi;
}

}  // namespace

namespace {

// Lanczos driver over abstract products (reference src/la/csr.cpp:86-132): begin(v0) loads the
// unit start vector, step(β_prev, α, β²) does w = ÃᵀÃv, α = wᵀv, w ← w − (αv + β_prev v_prev),
// β² = Σw²; advance(β) sets v_prev ← v, v ← w/β.
template <class Begin, class Step, class Advance>
double lanczos_norm(int n, int max_iter, double rel_tol, Begin&& begin, Step&& step, Advance&& advance) {
  if (n == 0) return 0.0;
  const int K = std::min(std::max(max_iter, 20), 400);
  std::vector<double> v(n);
  std::uint64_t state = 0x9E3779B97F4A7C15ULL;  // fixed xorshift64* start → uniform(−1, 1)
  double nv = 0;
  for (int j = 0; j < n; ++j) {
    state ^= state >> 12, state ^= state << 25, state ^= state >> 27;
    v[j] = static_cast<double>((state * 0x2545F4914F6CDD1DULL) >> 11) / 9007199254740992.0 * 2.0 - 1.0;
    nv += v[j] * v[j];
  }
  nv = std::sqrt(nv);
  for (double& e : v) e /= nv;
  begin(v);
  std::vector<double> alpha, beta;
  double theta = 0, theta_old = -1;
  int stagnant = 0;
  for (int k = 0; k < K; ++k) {
    double al = 0, bsq = 0;
    step(beta.empty() ? 0.0 : beta.back(), al, bsq);
    const double bn = std::sqrt(bsq);
    alpha.push_back(al);
    theta = tridiag_max_eig(alpha, beta, static_cast<int>(alpha.size()));
    if (bn <= 1e-14 * std::max(1.0, theta)) break;
    beta.push_back(bn);
    if (theta - theta_old <= rel_tol * theta) {
      if (++stagnant >= 5 && k >= 20) break;
    } else {
      stagnant = 0;
    }
    theta_old = theta;
    advance(bn);
  }
  return std::sqrt(std::max(theta, 0.0)) * (1.0 + 1e-4);
}

}  // namespace

double estimate_norm2(Device& dev, int n, int mode, int max_iter, double rel_tol) {
  psh_ctx* c = dev.ctx();
  if (mode != PSH_RED_TREE) {
    // The reference's sequential dots (csr.cpp:117-123) on a host core, which runs a dependent
    // fp64 add chain about 4x faster than one device lane; the products stay on the device.
    std::vector<double> v, vp(n, 0.0), w(n);
    return lanczos_norm(
        n, max_iter, rel_tol, [&](const std::vector<double>& v0) { v = v0; },
        [&](double bp, double& al, double& bsq) {
          dev.check(psh_lanczos_product(c, v.data(), w.data()), "lanczos");
          al = 0.0;
          for (int j = 0; j < n; ++j) al += w[j] * v[j];
          bsq = 0.0;
          for (int j = 0; j < n; ++j) {
            w[j] -= al * v[j] + bp * vp[j];
            bsq += w[j] * w[j];
          }
        },
        [&](double bn) {
          for (int j = 0; j < n; ++j) {
            vp[j] = v[j];
            v[j] = w[j] / bn;
          }
        });
  }
  return lanczos_norm(
      n, max_iter, rel_tol, [&](const std::vector<double>& v) { dev.check(psh_lanczos_begin(c, v.data()), "lanczos"); },
      [&](double bp, double& al, double& bsq) { dev.check(psh_lanczos_step(c, bp, mode, &al, &bsq), "lanczos"); },
      [&](double bn) { dev.check(psh_lanczos_advance(c, bn), "lanczos"); });
}

double estimate_norm2_host(const ScaledProblem& sp, int max_iter, double rel_tol) {
  const int n = sp.n;
  std::vector<double> v, vp(n, 0.0), w(n), t(sp.m);
  return lanczos_norm(
      n, max_iter, rel_tol, [&](const std::vector<double>& v0) { v = v0; },
      [&](double bp, double& al, double& bsq) {
        multiply(sp.A, v.data(), t.data());
        multiply(sp.At, t.data(), w.data());
        al = 0.0;
        for (int j = 0; j < n; ++j) al += w[j] * v[j];
        bsq = 0.0;
        for (int j = 0; j < n; ++j) {
          w[j] -= al * v[j] + bp * vp[j];
          bsq += w[j] * w[j];
        }
      },
      [&](double bn) {
        for (int j = 0; j < n; ++j) {
          vp[j] = v[j];
          v[j] = w[j] / bn;
        }
      });
}

// ------------------------------------------------------------------ engine
namespace {

class Context {  // reference EngineContext (engine.cpp:195-307)
 public:
  Context(const Model& model, const EngineOptions& opt)
      : model_(model), opt_(opt), t0_(std::chrono::steady_clock::now()) {
    if (static_cast<long long>(model.value.size()) >= (1LL << 31)) {
      error_ = "matrix has 2^31 or more nonzeros (beyond the 32-bit device index range)";
      return;
    }
    dev_ = Device::open(opt.device, error_);
    if (!dev_) return;
    const double t_open = elapsed();
    dev_->check(psh_sliced(dev_->ctx(), opt.sliced), "sliced");  // governs the phased build at load
    scale_and_load(*dev_, model, opt.scaling, opt.device_scaling != 0, sp_);
    if (opt.timing) dev_->check(psh_timing(dev_->ctx(), 1), "timing");
    dev_->check(psh_graphs(dev_->ctx(), opt.graphs), "graphs");
    dev_->check(psh_sync(dev_->ctx()), "load");
    const double t_load = elapsed();
    const bool empty = sp_.n == 0 || sp_.A.nnz() == 0;
    const double norm = empty ? 0.0 : estimate_norm2(*dev_, sp_.n, reduction_mode(opt));
    eta_ = norm > 0 ? 0.998 / norm : 1.0;
    dev_->check(psh_sync(dev_->ctx()), "setup");
    setup_seconds_ = elapsed();
    if (opt.verbosity >= 1)
      std::fprintf(stderr, "setup: device open %.3fs, scaling and load (%s) %.3fs, norm %.3fs\n", t_open,
                   opt.device_scaling ? "sweeps on the device" : "host", t_load - t_open, setup_seconds_ - t_load);
  }
  bool ok() const { return dev_ != nullptr; }
  const std::string& error() const { return error_; }
  Device& dev() { return *dev_; }
  psh_ctx* ctx() { return dev_->ctx(); }
  const ScaledProblem& sp() const { return sp_; }
  double eta() const { return eta_; }
  int mode() const { return reduction_mode(opt_); }
  double elapsed() const { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0_).count(); }
  bool out_of_time() const { return elapsed() > opt_.time_limit; }
  double initial_primal_weight() const {
    if (opt_.warm_primal_weight > 0) return opt_.warm_primal_weight;
    if (opt_.scaling.bound_objective_rescaling) return 1.0;
    double c2 = 0, b2 = 0;
    for (double v : sp_.c) c2 += v * v;
    for (int i = 0; i < sp_.m; ++i) {
      const double lo = std::isfinite(sp_.row_lower[i]) ? std::fabs(sp_.row_lower[i]) : 0.0;
      const double up = std::isfinite(sp_.row_upper[i]) ? std::fabs(sp_.row_upper[i]) : 0.0;
      b2 += std::max(lo, up) * std::max(lo, up);
    }
    return (std::sqrt(c2) + 1.0) / (std::sqrt(b2) + 1.0);
  }
  KktStats adjust(KktStats k) const {
    if (opt_.kkt_b_norm >= 0) k.b_norm = opt_.kkt_b_norm;
    if (opt_.kkt_c_norm >= 0) k.c_norm = opt_.kkt_c_norm;
    k.obj_shift = opt_.kkt_obj_shift;
    return k;
  }
  KktStats device_kkt(int xs, int ys) {
    psh_kkt d{};
    dev_->check(psh_kkt_of(ctx(), xs, ys, mode(), &d), "kkt");
    return adjust(from_device(d));
  }
  bool record(const KktStats& k, std::int64_t it) {
    const double e = k.rel_kkt();
    if (fast_iterations_ < 0 && e <= tol::kFirstOrderFast) {
      fast_iterations_ = it;
      fast_seconds_ = elapsed();
    }
    // verbosity 1: a progress line at most once a second; 2: every check
    if (opt_.verbosity >= 2 || (opt_.verbosity == 1 && elapsed() - last_log_ >= 1.0)) {
      last_log_ = elapsed();
      std::fprintf(stderr, "r2hpdhg-hip it %8lld  t %7.2fs  pobj %+.10e  dobj %+.10e  rp %.2e  rd %.2e  gap %.2e  %s\n",
                   static_cast<long long>(it), last_log_, k.primal_obj, k.dual_obj, k.rel_primal(), k.rel_dual(),
                   k.rel_gap(), psh_is_fp32(ctx()) ? "fp32" : "fp64");
    }
    return k.converged(opt_.tolerance);
  }
  void note_primal_weight(double w) { primal_weight_ = w; }

  // Ray-test verdict from the device scalars (reference engine.cpp:82-184).
  Status ray_verdict(const psh_ray& t, bool drift, const KktStats& cur, std::string& msg) {
    char buf[200];
    double obj = 0, viol = 0;
    const Status st = ray_decision(t, obj, viol);
    if (st == Status::Infeasible) {
      dual_ray_.assign(sp_.m, 0.0);
      dev_->check(psh_ray_vectors(ctx(), drift ? 1 : 0, dual_ray_.data(), nullptr), "ray vectors");
      std::snprintf(buf, sizeof buf, "primal infeasible: dual ray certificate (objective %f, violation %f)", obj, viol);
      msg = buf;
    } else if (st == Status::Unbounded) {
      const bool feasible = cur.rel_primal() <= tol::kVerifyPrimal && cur.primal_max_rel <= tol::kVerifyPrimal;
      primal_ray_.assign(sp_.n, 0.0);
      dev_->check(psh_ray_vectors(ctx(), drift ? 1 : 0, nullptr, primal_ray_.data()), "ray vectors");
      std::snprintf(buf, sizeof buf, "unbounded: primal ray certificate (c'd %f, violation %f)%s", obj, viol,
                    feasible ? " at a primal-feasible iterate" : "; iterate not primal-feasible yet");
      msg = buf;
    }
    return st;
  }

  // Warm start (reference engine.cpp:237-266); returns true when one was applied.
  bool apply_warm_start(bool mixed_now, bool& promote) {
    promote = false;
    if (opt_.warm_x.empty() && opt_.warm_y.empty()) return false;
    const Model& M = model_;
    std::vector<double> xo, yo(sp_.m, 0.0);
    if (!opt_.warm_x.empty()) {
      xo.resize(sp_.n);
      for (int j = 0; j < sp_.n; ++j) xo[j] = std::min(std::max(opt_.warm_x[j], M.col_lower[j]), M.col_upper[j]);
    } else {
      std::vector<double> xs(sp_.n);
      dev_->check(psh_download(ctx(), PSH_X, xs.data()), "download");
      sp_.unscale_primal(xs, xo);
    }
    if (!opt_.warm_y.empty())
      for (int i = 0; i < sp_.m; ++i) {
        double v = M.sense * opt_.warm_y[i];
        if (!std::isfinite(M.row_lower[i])) v = std::min(v, 0.0);
        if (!std::isfinite(M.row_upper[i])) v = std::max(v, 0.0);
        yo[i] = v;
      }
    // precision decided BEFORE the upload (an fp32 copy would lose the warm point's accuracy)
    const double rel = adjust(kkt_on_original(sp_, xo, yo)).rel_kkt();
    if (mixed_now && rel <= tol::kMixedPrecisionSwitch && opt_.tolerance < tol::kMixedPrecisionSwitch) promote = true;
    if (promote) dev_->check(psh_set_fp32(ctx(), 0), "precision");
    std::vector<double> xs(sp_.n), ys(sp_.m);
    for (int j = 0; j < sp_.n; ++j) xs[j] = sp_.bound_scale * xo[j] / sp_.col_scale[j];
    for (int i = 0; i < sp_.m; ++i) ys[i] = sp_.obj_scale * yo[i] / sp_.row_scale[i];
    dev_->check(psh_upload(ctx(), PSH_X, xs.data()), "upload");
    dev_->check(psh_upload(ctx(), PSH_Y, ys.data()), "upload");
    return true;
  }

  void set_engine_name(const char* name) { engine_name_ = name; }
  void loop_started() { loop_t0_ = elapsed(); }
  Solution finish(Status status, int xs_slot, int ys_slot, std::int64_t iterations, const std::string& message) {
    const double loop_end = elapsed();
    Solution sol;
    sol.iterate_seconds = loop_t0_ >= 0 ? loop_end - loop_t0_ : 0.0;
    sol.engine = engine_name_;
    sol.precision = to_string(opt_.precision);
    sol.iterations = iterations;
    sol.message = message;
    std::vector<double> xs(sp_.n), ys(sp_.m), y_min;
    dev_->check(psh_download(ctx(), xs_slot, xs.data()), "download");
    dev_->check(psh_download(ctx(), ys_slot, ys.data()), "download");
    sp_.unscale_primal(xs, sol.x);
    sp_.unscale_dual(ys, y_min);
    const double t_download = elapsed();
    const KktStats k = adjust(kkt_on_original(sp_, sol.x, y_min));
    const double t_kkt = elapsed();
    if (!k.finite()) status = Status::NumericalError;
    if (status == Status::Optimal && !k.converged(opt_.tolerance)) status = Status::NumericalError;
    const double sense = model_.sense;
    sol.y.resize(sp_.m);
    for (int i = 0; i < sp_.m; ++i) sol.y[i] = sense * y_min[i];
    sol.z = model_.obj;  // z = c − Aᵀy (original sense)
    parallel_for(sp_.n, [&](std::int64_t b, std::int64_t e) {
      for (std::int64_t j = b; j < e; ++j)
        for (int p = model_.col_start[j]; p < model_.col_start[j + 1]; ++p)
          sol.z[j] -= model_.value[p] * sol.y[model_.row_index[p]];
    });
    {
      Csr A;
      A.rows = sp_.m;
      A.cols = sp_.n;
      A.ptr = sp_.A.ptr;
      A.col = sp_.A.col;
      A.val = sp_.a_orig;
      sol.row_activity.resize(sp_.m);
      multiply(A, sol.x.data(), sol.row_activity.data());
    }
    sol.objective = model_.objective_value(sol.x);
    if (opt_.verbosity >= 1)
      std::fprintf(stderr, "finish: download and unscale %.3fs, KKT on the original %.3fs, z and activities %.3fs\n",
                   t_download - loop_end, t_kkt - t_download, elapsed() - t_kkt);
    sol.dual_objective = sense * k.dual_obj + model_.obj_offset;
    sol.primal_residual = k.rel_primal();
    sol.dual_residual = k.rel_dual();
    sol.gap = k.rel_gap();
    sol.status = status;
    sol.iterations_to_fast = fast_iterations_;
    sol.seconds_to_fast = fast_seconds_;
    sol.seconds = elapsed();
    sol.setup_seconds = setup_seconds_;
    sol.primal_weight = primal_weight_;
    if (status == Status::Infeasible) sol.dual_ray = dual_ray_;
    if (status == Status::Unbounded) sol.primal_ray = primal_ray_;
    if (status == Status::Infeasible || status == Status::Unbounded) sol.objective = sol.dual_objective = kNaN;
    if (opt_.timing)
      for (int k = 0; k < 3; ++k) dev_->check(psh_timing_get(ctx(), k, &sol.kernel_launches[k], &sol.kernel_ms[k]), "timing");
    sol.device_bytes = static_cast<double>(psh_device_bytes(ctx()));
    return sol;
  }

 private:
  const Model& model_;
  const EngineOptions& opt_;
  ScaledProblem sp_;
  std::unique_ptr<Device> dev_;
  std::string error_;
  double eta_ = 1.0, setup_seconds_ = 0, primal_weight_ = kNaN, fast_seconds_ = -1, last_log_ = 0;
  std::int64_t fast_iterations_ = -1;
  std::chrono::steady_clock::time_point t0_;
  std::vector<double> dual_ray_, primal_ray_;
  const char* engine_name_ = "r2hpdhg";
  double loop_t0_ = -1;
};

// Mixed precision. With refine = 0 the reference's policy (engine.cpp:56-79): fp32 storage until
// fp32 has done what it can, then fp64. With refine = 1 (the default outside CPU-order mode)
// the increment form: fp32 matrix and increments around fp64 anchors that every restart
// rebuilds, to the full tolerance; fp64 takes over only if progress stops for `refine_patience`
// checks once the relative KKT error is below 1e-5, or at any level for a quarter of the
// iterations so far (RefineGuard). The bf16 variant (bf16_until > 0) starts the
// increment form with bf16 matrix values and moves to fp32 values at bf16_until or on a stall of
// 10 checks.
class PrecisionPolicy {
 public:
  PrecisionPolicy(Context& ctx, const EngineOptions& o)
      : ctx_(ctx), opt_(o), mixed_(o.precision == Precision::Mixed), refine_(mixed_ && o.refine != 0),
        target_(o.tolerance), bf16_until_(refine_ ? o.bf16_until : 0.0) {
    if (refine_) ctx_.dev().check(psh_delta(ctx_.ctx(), bf16_until_ > 0 ? 2 : 1), "precision");
    else ctx_.dev().check(psh_set_fp32(ctx_.ctx(), mixed_ ? 1 : 0), "precision");
  }
  bool fp32() { return psh_is_fp32(ctx_.ctx()) != 0; }
  bool refining() const { return refine_ && !left_; }
  bool on_check(double rel, std::int64_t it) {
    if (refine_) {
      if (left_) return false;
      if (bf16_until_ > 0) {
        if (rel < 0.9 * best_) best_ = rel, stale_ = 0;
        else ++stale_;
        if (rel <= bf16_until_ || stale_ >= 10) {
          ctx_.dev().check(psh_delta(ctx_.ctx(), 1), "precision");
          bf16_until_ = 0;
          guard_ = RefineGuard{};
        }
        return false;
      }
      if (guard_.stalled(rel, it, opt_)) {
        ctx_.dev().check(psh_delta(ctx_.ctx(), 0), "precision");
        left_ = true;
        return true;
      }
      return false;
    }
    if (!mixed_ || !fp32()) return false;
    if (rel < 0.9 * best_) best_ = rel, stale_ = 0;
    else ++stale_;
    const bool reached = rel <= tol::kMixedPrecisionSwitch && target_ < tol::kMixedPrecisionSwitch;
    if (reached || stale_ >= 10) {
      ctx_.dev().check(psh_set_fp32(ctx_.ctx(), 0), "precision");
      return true;
    }
    return false;
  }

 private:
  Context& ctx_;
  const EngineOptions& opt_;
  bool mixed_, refine_, left_ = false;
  double target_, bf16_until_, best_ = 1e300;
  int stale_ = 0;
  RefineGuard guard_;
};

}  // namespace

Solution solve_r2hpdhg(const Model& model, const EngineOptions& opt) {
  Context ctx(model, opt);
  if (!ctx.ok()) {
    Solution s;
    s.engine = "r2hpdhg";
    s.status = Status::NotSolved;
    s.message = ctx.error();
    return s;
  }
  Device& dev = ctx.dev();
  psh_ctx* c = ctx.ctx();
  PrecisionPolicy policy(ctx, opt);
  const double eta = ctx.eta(), rho = opt.reflection;
  const int mode = ctx.mode();
  PidState pid;
  pid.omega = pid.best_omega = ctx.initial_primal_weight();

  // x = proj_X(0), y = 0 (reference r2hpdhg.cpp:31-34)
  dev.check(psh_fill(c, PSH_DX, 0.0), "fill");
  dev.check(psh_fill(c, PSH_ATY, 0.0), "fill");
  dev.check(psh_primal_step(c, PSH_DX, PSH_ATY, 0.0, PSH_X), "init");
  dev.check(psh_fill(c, PSH_Y, 0.0), "fill");
  bool promoted_by_warm = false;
  const bool warm = ctx.apply_warm_start(policy.fp32(), promoted_by_warm);
  ctx.note_primal_weight(pid.omega);
  if (warm) {
    const KktStats kw = ctx.device_kkt(PSH_X, PSH_Y);
    if (kw.finite() && ctx.record(kw, 0)) return ctx.finish(Status::Optimal, PSH_X, PSH_Y, 0, "warm start already optimal");
  }
  dev.check(psh_copy(c, PSH_X0, PSH_X), "copy");
  dev.check(psh_copy(c, PSH_Y0, PSH_Y), "copy");
  dev.check(psh_copy(c, PSH_XH, PSH_X), "copy");
  dev.check(psh_copy(c, PSH_YH, PSH_Y), "copy");

  auto residual = [&](double px, double py, double cross) {
    const double r2 = pid.omega * px + py / pid.omega + 2.0 * eta * cross;
    return std::sqrt(std::max(r2, 0.0));
  };
  const int K = opt.check_every;
  std::int64_t it = 0, inner = 0;
  double r0 = -1.0, r_last = kInf;
  double gap_best = kInf;  // hip_max_epoch_gap_only = 2: the cap waits for the gap to stall
  std::int64_t gap_best_it = 0;
  std::string msg = warm ? "warm start" : "";
  if (policy.refining())
    msg += std::string(msg.empty() ? "" : "; ") + "fp32 increments on fp64 anchors" +
           (opt.bf16_until > 0 ? " (bf16 matrix values first)" : "");
  ctx.loop_started();

  for (;;) {
    psh_chunk_args ch{};
    ch.iterations = K;
    ch.inner0 = inner;
    ch.tau = eta / pid.omega;
    ch.sigma = eta * pid.omega;
    ch.rho = rho;
    ch.eta = eta;
    ch.omega = pid.omega;
    ch.want_hat_last = 1;
    ch.mode = mode;
    dev.check(psh_chunk(c, &ch), "iteration");
    inner += K;
    it += K;

    psh_check_args ca{};
    ca.omega = pid.omega;
    ca.eta = eta;
    ca.want_ray = it >= 4 * K ? 1 : 0;
    ca.want_drift = it >= 4 * K && inner >= K && (it / K) % 4 == 0 ? 1 : 0;
    ca.want_pid = mode == PSH_RED_TREE ? 1 : 0;
    ca.mode = mode;
    psh_stats st{};
    dev.check(psh_check(c, &ca, &st), "check");
    if (st.have_r0) r0 = residual(st.r0_px, st.r0_py, st.r0_cross);  // residual at the epoch's anchor
    const double r = residual(st.px, st.py, st.cross);
    const KktStats k = ctx.adjust(from_device(st.kkt));
    if (!k.finite() || !std::isfinite(r)) return ctx.finish(Status::NumericalError, PSH_XH, PSH_YH, it, "non-finite iterate");
    if (ctx.record(k, it)) return ctx.finish(Status::Optimal, PSH_XH, PSH_YH, it, msg);
    if (it >= opt.iteration_limit) return ctx.finish(Status::IterationLimit, PSH_XH, PSH_YH, it, msg);
    if (ctx.out_of_time()) return ctx.finish(Status::TimeLimit, PSH_XH, PSH_YH, it, msg);
    if (it >= 4 * K) {  // ray tests on T(z) − z, then on the drift since the anchor
      std::string why;
      Status verdict = ctx.ray_verdict(st.ray, false, k, why);
      if (verdict == Status::NotSolved && ca.want_drift) {
        verdict = ctx.ray_verdict(st.drift, true, k, why);
        if (verdict != Status::NotSolved) why += " (drift since the restart anchor)";
      }
      if (verdict != Status::NotSolved) return ctx.finish(verdict, PSH_XH, PSH_YH, it, why);
    }
    const bool promoted = policy.on_check(k.rel_kkt(), it);
    if (promoted) msg += std::string(msg.empty() ? "" : "; ") + "fp32 -> fp64 at iteration " + std::to_string(it);
    if (promoted) gap_best = kInf;  // fp64 starts its own record of the gap's progress
    if (k.rel_gap() < 0.9 * gap_best) gap_best = k.rel_gap(), gap_best_it = it;
    const bool cap_ok = opt.max_epoch_gap_only != 2 || it - gap_best_it >= opt.max_epoch;
    const bool restart = promoted || restart_due(opt, r, r0, r_last, inner, it) ||
                         increment_epoch_full(opt, policy.refining(), inner) || (cap_ok && epoch_cap_reached(opt, inner, k));
    r_last = r;
    if (!restart) continue;

    // PID primal weight (reference r2hpdhg.cpp:127-148)
    double dxn = st.dx_norm, dyn = st.dy_norm;
    if (mode != PSH_RED_TREE || promoted) {
      dev.check(psh_diff_norm2(c, PSH_XH, PSH_X0, mode, &dxn), "pid");
      dev.check(psh_diff_norm2(c, PSH_YH, PSH_Y0, mode, &dyn), "pid");
    }
    pid_update(pid, opt, dxn, dyn, k.rel_primal(), k.rel_dual());
    ctx.note_primal_weight(pid.omega);
    if (opt.verbosity >= 2)
      std::fprintf(stderr, "  restart at %lld (inner %lld): r/r0 %.3f  omega %.4e  dx %.3e dy %.3e\n",
                   static_cast<long long>(it), static_cast<long long>(inner), r / r0, pid.omega, dxn, dyn);
    dev.check(psh_restart(c), "restart");
    inner = 0;
    r_last = kInf;
  }
}

Solution solve_pdlp(const Model& model, const EngineOptions& opt_in) {
  EngineOptions opt = opt_in;
  opt.refine = 0;  // the increment form lives in the fused chunk; PDLP keeps the reference's policy
  Context ctx(model, opt);
  ctx.set_engine_name("pdlp");
  if (!ctx.ok()) {
    Solution s;
    s.engine = "pdlp";
    s.status = Status::NotSolved;
    s.message = ctx.error();
    return s;
  }
  Device& dev = ctx.dev();
  psh_ctx* c = ctx.ctx();
  PrecisionPolicy policy(ctx, opt);
  const double eta = ctx.eta();
  const int mode = ctx.mode();
  double omega = ctx.initial_primal_weight();
  // iterate, next iterate, averages, restart anchors, work vectors (reference pdlp.cpp:24-28)
  int x = PSH_X, xn = PSH_XH, xa = PSH_XB, x0 = PSH_X0, xbar = PSH_DX, aty = PSH_ATY;
  int y = PSH_Y, yn = PSH_YH, ya = PSH_YB, y0 = PSH_Y0, ax = PSH_AX;
  auto ok = [&](int rc, const char* what) { dev.check(rc, what); };
  ok(psh_fill(c, xbar, 0.0), "fill");
  ok(psh_fill(c, aty, 0.0), "fill");
  ok(psh_primal_step(c, xbar, aty, 0.0, x), "init");
  ok(psh_fill(c, y, 0.0), "fill");
  bool promoted_by_warm = false;
  const bool warm = ctx.apply_warm_start(policy.fp32(), promoted_by_warm);
  ctx.note_primal_weight(omega);
  if (warm) {
    const KktStats kw = ctx.device_kkt(x, y);
    if (kw.finite() && ctx.record(kw, 0)) return ctx.finish(Status::Optimal, x, y, 0, "warm start already optimal");
  }
  ok(psh_copy(c, x0, x), "copy");
  ok(psh_copy(c, y0, y), "copy");
  ok(psh_copy(c, xa, x), "copy");
  ok(psh_copy(c, ya, y), "copy");
  const KktStats k0 = ctx.device_kkt(x, y);
  ctx.record(k0, 0);
  double restart_kkt = k0.rel_kkt(), last_candidate = 1e300;
  std::int64_t it = 0, inner = 0;
  const int K = opt.check_every;
  std::string msg = warm ? "warm start" : "";
  ctx.loop_started();
  for (;;) {
    const double tau = eta / omega, sigma = eta * omega;
    for (int s = 0; s < K; ++s) {
      ok(psh_spmv_t(c, y, aty), "iteration");
      ok(psh_primal_step(c, x, aty, tau, xn), "iteration");
      ok(psh_copy(c, xbar, xn), "iteration");
      ok(psh_axpby(c, -1.0, x, 2.0, xbar), "iteration");  // x̄ = 2x⁺ − x
      ok(psh_spmv(c, xbar, ax), "iteration");
      ok(psh_dual_step(c, y, ax, sigma, yn), "iteration");
      std::swap(x, xn);
      std::swap(y, yn);
      ++inner;
      const double wa = 1.0 / static_cast<double>(inner);
      ok(psh_axpby(c, wa, x, 1.0 - wa, xa), "iteration");
      ok(psh_axpby(c, wa, y, 1.0 - wa, ya), "iteration");
      ++it;
      if (s + 1 < K && ctx.out_of_time()) {
        const KktStats kt = ctx.device_kkt(x, y);
        if (kt.finite() && ctx.record(kt, it)) return ctx.finish(Status::Optimal, x, y, it, msg);
        return ctx.finish(Status::TimeLimit, x, y, it, msg);
      }
    }
    const KktStats kc = ctx.device_kkt(x, y), ka = ctx.device_kkt(xa, ya);
    const bool use_avg = ka.finite() && ka.rel_kkt() < kc.rel_kkt();
    const KktStats& kk = use_avg ? ka : kc;
    const int cx = use_avg ? xa : x, cy = use_avg ? ya : y;
    if (!kk.finite()) return ctx.finish(Status::NumericalError, x, y, it, "non-finite KKT statistics");
    if (ctx.record(kk, it)) return ctx.finish(Status::Optimal, cx, cy, it, msg);
    if (it >= opt.iteration_limit) return ctx.finish(Status::IterationLimit, cx, cy, it, msg);
    if (ctx.out_of_time()) return ctx.finish(Status::TimeLimit, cx, cy, it, msg);
    if (it >= 4 * K) {  // Δ = z_k − z_{k−1}; xn / yn hold z_{k−1} after the swap
      ok(psh_axpby(c, 1.0, x, -1.0, xn), "ray");
      ok(psh_axpby(c, 1.0, y, -1.0, yn), "ray");
      std::string why;
      psh_ray ray{};
      ok(psh_ray_of(c, xn, yn, &ray), "ray");
      Status st = ctx.ray_verdict(ray, false, kc, why);
      if (st == Status::NotSolved && inner >= K && (it / K) % 4 == 0) {  // the drift since the anchor
        ok(psh_copy(c, xn, x), "ray");
        ok(psh_axpby(c, -1.0, x0, 1.0, xn), "ray");
        ok(psh_copy(c, yn, y), "ray");
        ok(psh_axpby(c, -1.0, y0, 1.0, yn), "ray");
        ok(psh_ray_of(c, xn, yn, &ray), "ray");
        st = ctx.ray_verdict(ray, false, kc, why);
        if (st != Status::NotSolved) why += " (drift since the restart anchor)";
      }
      if (st != Status::NotSolved) return ctx.finish(st, x, y, it, why);
    }
    const double e = kk.rel_kkt();
    const bool promoted = policy.on_check(e, it);
    if (promoted) msg += std::string(msg.empty() ? "" : "; ") + "fp32 -> fp64 at iteration " + std::to_string(it);
    const bool restart = promoted || e <= 0.2 * restart_kkt ||
                         (e <= opt.pdlp_restart_necessary * restart_kkt && e > last_candidate) ||
                         static_cast<double>(inner) >= 0.36 * static_cast<double>(it);
    last_candidate = e;
    if (restart) {
      double dx = 0, dy = 0;
      ok(psh_diff_norm2(c, cx, x0, mode, &dx), "pid");
      ok(psh_diff_norm2(c, cy, y0, mode, &dy), "pid");
      if (dx > 1e-10 && dy > 1e-10) {
        const double th = opt.pdlp_primal_weight_smoothing;
        omega = std::exp(th * std::log(dy / dx) + (1.0 - th) * std::log(omega));
        ctx.note_primal_weight(omega);
      }
      if (cx != x) ok(psh_copy(c, x, cx), "restart");
      if (cy != y) ok(psh_copy(c, y, cy), "restart");
      ok(psh_copy(c, x0, x), "restart");
      ok(psh_copy(c, y0, y), "restart");
      ok(psh_copy(c, xa, x), "restart");
      ok(psh_copy(c, ya, y), "restart");
      inner = 0;
      restart_kkt = e;
      last_candidate = 1e300;
    }
  }
}

}  // namespace ps26119::amd
