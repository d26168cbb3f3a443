// engine.h — the r²HPDHG engine on the MI355X (reference src/pdhg/r2hpdhg.{h,cpp},
// engine.{h,cpp}): same restarts, PID primal weight, precision policy, ray tests, warm start
// and finish, with the iteration and the per-check reductions on the GPU.
#pragma once

#include <algorithm>
#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "contract.h"
#include "ps26119_hip.h"
#include "scaling.h"

namespace ps26119::amd {

struct EngineOptions {  // reference src/pdhg/engine.h:27-48
  double tolerance = tol::kFirstOrderHigh;
  std::vector<double> warm_x, warm_y;
  double warm_primal_weight = 0.0;
  double kkt_b_norm = -1, kkt_c_norm = -1, kkt_obj_shift = 0;
  double time_limit = 3600.0;
  std::int64_t iteration_limit = 100'000'000;
  int check_every = 64;
  Precision precision = Precision::Fp64;
  int verbosity = 0;
  ScalingOptions scaling;
  double reflection = 1.0;
  double restart_sufficient = 0.2, restart_necessary = 0.5, restart_artificial = 0.36;
  double pid_kp = 0.99, pid_ki = 0.01, pid_kd = 0.0, pid_integral_decay = 0.3;
  double pid_max_log_step = 2.302585092994046;
  double pdlp_restart_necessary = 0.8;        // PDLP engine (reference engine.h:46-47)
  double pdlp_primal_weight_smoothing = 0.5;  // θ
  // MI355X knobs (engine_params names hip_cpu_order, hip_tree_reductions, hip_device):
  //   cpu_order = 1: every reduction in the reference CPU order (bit-identical trajectory and KKT
  //   figures to the reference CPU backend; the KKT sums are long sequential loops).
  //   Default (both 0): the sums that restart and primal-weight decisions read (residual dots,
  //   PID distances, Lanczos dots) in the reference order and the KKT sums as device trees: the
  //   same η, restarts, ω and iteration counts as the reference; KKT figures within rounding.
  //   tree = 1: every sum a fixed-shape device tree (and, at large m, the column-sliced dual
  //   half step): a deterministic trajectory of its own.
  int cpu_order = 0;
  int tree = 0;
  int device = 0;
  int timing = 0;  // hip_timing: per-kernel HIP event timing into the Solution
  int graphs = 1;  // hip_graphs: replay captured chunk graphs (off while timing)
  int sliced = 1;  // hip_sliced: phased products for large fp64 solves (0 off, 1 auto, 2 at any size)
  int device_scaling = 1;  // hip_device_scaling: the scaling sweeps on the device (same bits as the host's)
  // hip_refine: mixed precision as fp32 increments on fp64 anchors (default) instead of the
  // reference's fp32-then-fp64 switch; CPU-order mode (hip_cpu_order = 1) defaults it to 0.
  int refine = -1;
  int refine_patience = 40;  // hip_refine_patience: stalled checks before fp64 takes over
  // hip_refine_max_epoch: longest epoch of the increment form (0 = no limit). The Halpern term
  // (1 − w)·δ with w = (k+1)/(k+2) falls below fp32 resolution of the increments once k nears 10^6,
  // so a longer epoch loses its anchoring; the epoch then ends with a restart.
  std::int64_t refine_max_epoch = 1 << 16;
  // hip_max_epoch: longest epoch in any precision (0 = the reference's schedule alone). Not in the
  // reference: on the random LPs with a planted optimum the artificial restart (inner ≥ 0.36·it)
  // lets epochs grow with the iteration count and the relative gap stalls near 3e-7 at 1M rows;
  // shorter epochs keep closing it. hip_max_epoch_gap_only = 1 applies the cap only while the
  // relative gap is the largest KKT term, so long epochs return once feasibility is what is left;
  // 2 (single solves) also waits until the gap has not improved by 10% for max_epoch iterations.
  std::int64_t max_epoch = 0;
  int max_epoch_gap_only = 0;
  // hip_bf16_until: > 0 runs the increment form's SpMVs with bf16 matrix values until the relative
  // KKT error reaches this value or stops improving for 10 checks, then with fp32 values (0: off).
  double bf16_until = 0.0;
};

EngineOptions engine_options_from(const Options& o);
// The psh reduction mode (PSH_RED_*) the options ask for.
int reduction_mode(const EngineOptions& o);
// Applies Options::engine_params; returns the first unknown name or "".
std::string apply_engine_params(const Options& o, EngineOptions& e);

// KKT statistics (reference src/pdhg/termination.h:35-60), fp64, original problem, min form.
struct KktStats {
  double primal_residual = 0, dual_residual = 0, primal_obj = 0, dual_obj = 0, b_norm = 0, c_norm = 0;
  double primal_max_rel = 0, dual_max_rel = 0, obj_shift = 0;
  double rel_primal() const { return primal_residual / (1.0 + b_norm); }
  double rel_dual() const { return dual_residual / (1.0 + c_norm); }
  double rel_gap() const;
  double rel_kkt() const;
  bool converged(double eps) const;
  bool finite() const;
};

// Host fp64 KKT of an original-space point (x, y_min) — the exact reference computation
// (reference src/pdhg/cpu_backend.cpp:226-284), used for warm starts and the final report.
KktStats kkt_on_original(const ScaledProblem& sp, const std::vector<double>& x, const std::vector<double>& y_min);

// RAII device context for one solve. Errors become exceptions: std::bad_alloc for device
// out-of-memory (solve() reports NotSolved "out of memory"), std::runtime_error otherwise.
class Device {
 public:
  static std::unique_ptr<Device> open(int device, std::string& error);
  ~Device();
  psh_ctx* ctx() const { return ctx_; }
  void check(int rc, const char* what) const;
  void load(const ScaledProblem& sp);

 private:
  explicit Device(psh_ctx* c) : ctx_(c) {}
  psh_ctx* ctx_;
};

// ‖Ã‖₂ by Lanczos on ÃᵀÃ with the device products (reference src/la/csr.cpp:86-132).
double estimate_norm2(Device& dev, int n, int mode, int max_iter = 400, double rel_tol = 1e-12);
// The same iteration with host products and sequential dots (serving setup, on host threads).
double estimate_norm2_host(const ScaledProblem& sp, int max_iter = 400, double rel_tol = 1e-12);

// The psh_problem view of a scaled problem; `cmin` receives sense·c and must outlive the view.
psh_problem problem_of(const ScaledProblem& sp, std::vector<double>& cmin);

// Builds the scaled problem of `model` and loads it on `dev`: the structure on the host and the
// scaling sweeps on the device (psh_scale) when `on_device`, else everything on the host. On
// return sp holds the factors and the scaled vectors; with on_device its matrix values are not
// kept on the host (only the structure and a_orig).
void scale_and_load(Device& dev, const Model& model, const ScalingOptions& so, bool on_device, ScaledProblem& sp);

// Restart test of r²HPDHG (reference src/pdhg/r2hpdhg.cpp:118-122): sufficient, necessary or
// artificial restart. r0 < 0 means no epoch-start residual yet.
bool restart_due(const EngineOptions& o, double r, double r0, double r_last, std::int64_t inner, std::int64_t it);
// When the increment form hands over to fp64: (a) the relative KKT error is below 1e-5 and has not
// improved by 10% for refine_patience checks (fp32's floor), or (b) at any level, it has not
// improved by 10% over the last quarter of the iterations (at least 16·patience checks): on a
// badly conditioned LP the fp32 increments stop refining well above the tolerance.
struct RefineGuard {
  double best = 1e300;
  std::int64_t it_best = 0;  // iteration of the last 10% improvement
  int stale = 0;             // checks since then
  bool stalled(double rel, std::int64_t it, const EngineOptions& o) {
    if (rel < 0.9 * best) {
      best = rel;
      it_best = it;
      stale = 0;
      return false;
    }
    ++stale;
    const std::int64_t window =
        std::max<std::int64_t>(16LL * o.refine_patience * o.check_every, it / 4);
    return (rel <= 1e-5 && stale >= o.refine_patience) || it - it_best >= window;
  }
};

// The epoch limits: the increment form's (refine_max_epoch) and the optional one of any
// precision (max_epoch, gated on the gap leading the KKT terms when max_epoch_gap_only is set).
inline bool increment_epoch_full(const EngineOptions& o, bool increment_form, std::int64_t inner) {
  return increment_form && o.refine_max_epoch > 0 && inner >= o.refine_max_epoch;
}
inline bool epoch_cap_reached(const EngineOptions& o, std::int64_t inner, const KktStats& k) {
  if (o.max_epoch <= 0 || inner < o.max_epoch) return false;
  return !o.max_epoch_gap_only || k.rel_gap() >= std::max(k.rel_primal(), k.rel_dual());
}

// PID primal weight with damping and the best-balance fallback (reference r2hpdhg.cpp:127-148).
struct PidState {
  double omega = 1.0, best_omega = 1.0, best_balance = 1e300, integral = 0.0, last = 0.0;
};
void pid_update(PidState& p, const EngineOptions& o, double dx_norm, double dy_norm, double rel_primal,
                double rel_dual);

// Ray-test decision on the device scalars of one direction (reference engine.cpp:82-184):
// Infeasible (dual ray), Unbounded (primal ray) or NotSolved; obj and viol are the certificate's
// scaled objective and violation.
Status ray_decision(const psh_ray& t, double& obj, double& viol);

Solution solve_r2hpdhg(const Model& model, const EngineOptions& opt);
// PDLP-style restarted PDHG with iterate averaging on the same device operations (reference
// src/pdhg/pdlp.cpp:9-126): unfused Backend ops, KKT of the current and the averaged point.
Solution solve_pdlp(const Model& model, const EngineOptions& opt);

}  // namespace ps26119::amd
