/* ps26119_amd.h — extensions of the C API that the reference does not have (it has no C entry
 * points for these). Same model arguments as ps26119_solve_lp. No function throws.
 */
#ifndef PS26119_AMD_H
#define PS26119_AMD_H

#include "ps26119/ps26119.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Named engine knobs (the reference's Options::engine_params, options.h:70-82 there), e.g.
 * "restart_sufficient", "pid_kp", ..., plus the MI355X ones: "hip_cpu_order" (1 = every
 * reduction in the reference CPU order, bit-identical trajectory and KKT figures; by default the
 * sums that restart and primal-weight decisions read follow the reference order and the KKT
 * sums are device trees, which keeps the reference's iteration counts), "hip_tree_reductions"
 * (1 = every sum a device tree: a deterministic trajectory of its own), "hip_timing" (1 = per-kernel
 * HIP event timing into ps26119_amd_stats), "hip_graphs" (1 = default: replay captured HIP
 * graphs of the iteration chunks; timing runs launch eagerly), "hip_device", "hip_sliced"
 * (1 = default: phased products for large fp64 solves with scattered gathers, the same bits as
 * the row tiles; 0 = row tiles only; 2 = phased at any size), "hip_refine"
 * (mixed precision as fp32 increments on fp64 anchors; default 1, 0 in CPU-order mode: the
 * reference's fp32-then-fp64 policy), "hip_refine_patience" (stalled checks before fp64 takes
 * over, default 40), "hip_refine_max_epoch" (longest epoch of the increment form, default
 * 65536; 0 = no limit), "hip_max_epoch" (longest epoch in any precision: a restart to the
 * averaged point is forced when the epoch reaches it; default 0 = the reference's restart schedule
 * alone), "hip_max_epoch_gap_only" (1 = hip_max_epoch applies only while the relative gap is the
 * largest KKT term; 2 = also only once the gap has not improved by 10% for hip_max_epoch
 * iterations, single solves only, experimental), "hip_bf16_until" (> 0: the increment form starts with bf16 matrix values
 * and moves to fp32 values at this relative KKT error or on a stall; default 0 = off),
 * "hip_device_scaling" (1 = default: the scaling sweeps run on the device, same bits as the host).
 * "warm_primal_weight" is not an engine knob but the reference's Options::warm_primal_weight
 * (options.h there), which the C struct lacks: the first primal weight of a warm-started solve or
 * batch (0 = the default rule; a re-solve passes the earlier solve's stats->primal_weight).
 * presolve: 1 (default) / 0. */
typedef struct {
  int presolve;
  int check_every;            /* iterations between checks (0 = 64) */
  int n_params;
  const char* const* param_names;
  const double* param_values;
} ps26119_amd_options;

typedef struct {
  double setup_seconds;       /* scaling, upload, ‖Ã‖₂ */
  long long iterations_to_fast;
  double seconds_to_fast;
  double primal_weight;
  double check_primal, check_dual, check_gap;
  long long kernel_launches[3]; /* hip_timing: fused primal, fused dual, check pass */
  double kernel_ms[3];
  double device_bytes;          /* device memory of the solve */
  double iterate_seconds;       /* wall time of the iteration loop (setup and finish excluded) */
} ps26119_amd_stats;

void ps26119_amd_default_options(ps26119_amd_options* o);

/* ps26119_solve_lp_ex with the extension options and statistics (either may be NULL). */
int ps26119_amd_solve_lp(int num_rows, int num_cols, int sense, double obj_offset, const double* c,
                         const double* col_lower, const double* col_upper, const double* row_lower,
                         const double* row_upper, const int* col_start, const int* row_index,
                         const double* value, const ps26119_options* opt, const ps26119_amd_options* ext,
                         ps26119_result* result, ps26119_amd_stats* stats, double* x, double* y, double* z,
                         double* dual_ray, double* primal_ray);

/* One scenario of a batch: any array left NULL keeps the base model's. */
typedef struct {
  const double* obj;        /* n */
  const double* col_lower;  /* n */
  const double* col_upper;  /* n */
  const double* row_lower;  /* m */
  const double* row_upper;  /* m */
} ps26119_amd_scenario;

/* Scenarios sharing the base model's constraint matrix, solved in lockstep on the GPU (replaces
 * the reference's C++-only ps26119::solve_batch, include/ps26119/batch.h:32-33 there, which runs
 * on the CPU). opt->use_gpu must be 1; opt->warm_x / warm_y, when set, start every scenario from
 * that point. results[k] (required), xs[k·n + j] and ys[k·m + i] (may be NULL) receive scenario k.
 * Returns PS26119_OPTIMAL when every scenario is optimal, else the first other status. */
int ps26119_amd_solve_batch(int num_rows, int num_cols, int sense, double obj_offset, const double* c,
                            const double* col_lower, const double* col_upper, const double* row_lower,
                            const double* row_upper, const int* col_start, const int* row_index,
                            const double* value, int num_scenarios, const ps26119_amd_scenario* scenarios,
                            const ps26119_options* opt, const ps26119_amd_options* ext, ps26119_result* results,
                            double* xs, double* ys);

/* ---- serving: continuous batching of independent LP solves on one GPU (no counterpart in the
 * reference). Requests are prepared on host threads, admitted while the HBM byte budget has room,
 * and advanced together as one packed block-diagonal system; each finishes on its own schedule.
 * Per request: fp64 r²HPDHG without presolve, the reference's restart / PID rules, KKT on the
 * original model, certified bound and in-process check; the result of a request does not depend
 * on the other requests in the mix. */
typedef struct ps26119_amd_server ps26119_amd_server;
typedef struct {
  int device;
  double budget_bytes;   /* HBM arena (0 = 8 GiB) */
  int max_instances;     /* live at once (0 = 256) */
  int check_every;       /* 0 = 64 */
  int setup_threads;     /* host workers for preparation and finish (0 = 4) */
  int eager_launches;    /* 1: launch each chunk's kernels directly instead of replaying captured graphs */
  int timing;            /* 1: time every chunk with HIP events (ps26119_amd_server_kernel_time) */
} ps26119_amd_server_options;

/* NULL on failure, with the reason in error (may be NULL). */
ps26119_amd_server* ps26119_amd_server_create(const ps26119_amd_server_options* o, char* error, int error_len);
/* Abandons requests still in flight. */
void ps26119_amd_server_destroy(ps26119_amd_server* s);
/* Queues a solve (tolerance, iteration_limit, time_limit and verbosity of opt are honoured;
 * ext may carry engine parameters). Returns the request id (>= 0) or -1 for invalid arguments. */
long long ps26119_amd_server_submit(ps26119_amd_server* s, int num_rows, int num_cols, int sense, double obj_offset,
                                    const double* c, const double* col_lower, const double* col_upper,
                                    const double* row_lower, const double* row_upper, const int* col_start,
                                    const int* row_index, const double* value, const ps26119_options* opt,
                                    const ps26119_amd_options* ext);
/* Waits for request `id` (timeout_s < 0: forever). Returns 0 when done (result, x, y, z filled as
 * in ps26119_solve_lp; times = {submitted, admitted, done, latency} in seconds since the server
 * started, may be NULL), 1 on timeout, -1 for an unknown id. */
int ps26119_amd_server_wait(ps26119_amd_server* s, long long id, double timeout_s, ps26119_result* result, double* x,
                            double* y, double* z, double* times);
int ps26119_amd_server_in_flight(ps26119_amd_server* s);
/* With timing = 1: the served iterations so far (each one primal and one dual half step over the
 * live instances), their total device milliseconds (HIP events around every chunk on the serving
 * stream) and the algorithmic bytes they moved. Returns 0, or -1 when the server was not created
 * with timing. */
int ps26119_amd_server_kernel_time(ps26119_amd_server* s, long long* iterations, double* total_ms, double* bytes);

/* Number of visible GPUs the backend can use (0 when none). */
int ps26119_amd_device_count(void);

/* ---- host stages of the path, callable without a GPU (verification / tests) ---- */

/* Scaling of the model with the default options: row_scale (m), col_scale (n),
 * scalars = {bound_scale, obj_scale, log10_range, geometric_mean_applied}, the scaled values of
 * Ã in CSR order and of Ãᵀ in CSR order (nnz each; may be NULL). */
int ps26119_amd_scale(int num_rows, int num_cols, int sense, const double* c, const double* col_lower,
                      const double* col_upper, const double* row_lower, const double* row_upper,
                      const int* col_start, const int* row_index, const double* value, double* row_scale,
                      double* col_scale, double* scalars, double* a_vals, double* at_vals);

/* Rounding-proof bound from y (reference src/core/safe_bound.h semantics). */
double ps26119_amd_certified_bound(int num_rows, int num_cols, int sense, double obj_offset, const double* c,
                                   const double* col_lower, const double* col_upper, const double* row_lower,
                                   const double* row_upper, const int* col_start, const int* row_index,
                                   const double* value, const double* y, int use_implied_bounds);

/* In-process check: out = {max_primal_violation, max_dual_violation, relative_gap,
 * primal_objective, dual_objective, passed}. */
int ps26119_amd_check_solution(int num_rows, int num_cols, int sense, double obj_offset, const double* c,
                               const double* col_lower, const double* col_upper, const double* row_lower,
                               const double* row_upper, const int* col_start, const int* row_index,
                               const double* value, const double* x, const double* y, double* out);

/* Presolve summary: out = {outcome (0 reduced, 1 infeasible, 2 unchanged), removed_rows,
 * removed_cols, reduced_rows, reduced_cols, reduced_nnz, reduced_offset}. */
int ps26119_amd_presolve_info(int num_rows, int num_cols, int sense, double obj_offset, const double* c,
                              const double* col_lower, const double* col_upper, const double* row_lower,
                              const double* row_upper, const int* col_start, const int* row_index,
                              const double* value, double* out);

/* The reduced model itself: info as ps26119_presolve_info's out; when info[0] = 0 (reduced) the
 * kept rows and columns (reduced → original index), the reduced row and column bounds and the
 * reduced CSC (col_start: reduced_cols + 1, row_index / value: reduced_nnz). Arrays sized for the
 * original model suffice; any may be NULL. */
int ps26119_amd_presolve_model(int num_rows, int num_cols, int sense, double obj_offset, const double* c,
                               const double* col_lower, const double* col_upper, const double* row_lower,
                               const double* row_upper, const int* col_start, const int* row_index,
                               const double* value, double* info, int* kept_rows, int* kept_cols, double* rl,
                               double* ru, double* cl, double* cu, int* r_col_start, int* r_row_index,
                               double* r_value);

#ifdef __cplusplus
}
#endif

#endif /* PS26119_AMD_H */
