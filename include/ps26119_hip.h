/* ps26119_hip.h — the thin C-ABI layer between the C++ host code of the MI355X r²HPDHG backend
 * and its hand-written CDNA4 (gfx950) HIP kernels. Built as libps26119_hip.so. Plain C types
 * only: no HIP, no C++ and no torch types cross this boundary.
 *
 * It replaces the reference's CUDA backend object (reference src/gpu/cuda_backend.cu:423-670,
 * behind the virtual interface src/pdhg/backend.h:36-69). The reference compiles its backend
 * into the solver library, so there is no reference C-ABI for this layer; the calls below are
 * what the reference's Backend methods map to, plus fused entry points (psh_chunk, psh_check)
 * that the MI355X driver uses to keep the whole iteration on the device.
 *
 * Conventions
 *  - Every function returns 0 on success or a negative PSH_E_* code; psh_error(ctx) gives the
 *    text of the last error on that context (thread-compatible: one context per solve).
 *  - Vectors live on the device in fixed slots (PSH_X ... PSH_AX), in the working precision
 *    (fp64, or fp32 after psh_set_fp32(ctx, 1)); host copies are always fp64 arrays.
 *  - Reductions accumulate in fp64. mode PSH_RED_TREE = fixed-shape tree (deterministic, any
 *    size); PSH_RED_CPU = the reference CPU order (64 fixed chunks for dot / norms, exactly
 *    src/la/parallel.h:169-188; single sequential loops for the KKT sums,
 *    src/pdhg/cpu_backend.cpp:240-274, and the Lanczos dots, src/la/csr.cpp:117-123),
 *    bit-identical to the CPU backend; PSH_RED_DECISION = the reference order for every sum a
 *    restart or primal-weight decision reads (residual dots, PID distances, Lanczos dots) and
 *    trees for the KKT and ray-test sums, so the trajectory is the CPU backend's while a check
 *    costs about what a tree check costs.
 */
#ifndef PS26119_HIP_H
#define PS26119_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  PSH_OK = 0,
  PSH_E_NO_DEVICE = -1,   /* no usable GPU (reported like the reference's "backend unavailable") */
  PSH_E_OOM = -2,         /* device allocation failed (reference: std::bad_alloc → "out of memory") */
  PSH_E_ARG = -3,         /* bad argument / size mismatch */
  PSH_E_HIP = -4,         /* HIP runtime error (reference: std::runtime_error → NumericalError) */
  PSH_E_STATE = -5        /* call out of order (e.g. before psh_load) */
};

enum { PSH_RED_TREE = 0, PSH_RED_CPU = 1, PSH_RED_DECISION = 2 };

/* Vector slots. Primal (size n): X, X0, XH, XB, DX, ATY. Dual (size m): Y, Y0, YH, YB, DY, AX. */
enum { PSH_X = 0, PSH_X0, PSH_XH, PSH_XB, PSH_DX, PSH_ATY, PSH_Y, PSH_Y0, PSH_YH, PSH_YB, PSH_DY, PSH_AX, PSH_NSLOT };

typedef struct psh_ctx psh_ctx;

/* The scaled problem (reference src/pdhg/scaling.h:48-67) plus the original data the KKT
 * check needs. CSR with int32 row pointers and column indices (nnz < 2^31).
 *   A  (m×n): a_ptr, a_col, a_val (scaled Ã), a_val_orig (A, same structure);
 *   At (n×m): at_ptr, at_col, at_val (scaled Ãᵀ, rows sorted);
 *   At_orig (n×m): at_ptr, ato_col (may equal at_col), ato_val (the model's CSC values).
 * Vectors: c, l, u (scaled, n); rl, ru (scaled, m); row_scale (m), col_scale (n);
 * original bounds cl_o, cu_o (n), rl_o, ru_o (m); original objective in min form c_min = sense·c. */
typedef struct {
  int m, n;
  long long nnz;
  const int *a_ptr, *a_col;
  const double *a_val, *a_val_orig;
  const int *at_ptr, *at_col, *ato_col;
  const double *at_val, *ato_val;
  const double *c, *l, *u, *rl, *ru;
  const double *row_scale, *col_scale;
  double bound_scale, obj_scale;
  const double *cl_o, *cu_o, *rl_o, *ru_o, *c_min;
} psh_problem;

/* KKT statistics of an iterate on the ORIGINAL problem (reference src/pdhg/termination.h:35-60). */
typedef struct {
  double primal_residual, dual_residual, primal_obj, dual_obj, b_norm, c_norm, primal_max_rel, dual_max_rel;
} psh_kkt;

/* Ray-test scalars (reference src/pdhg/engine.cpp:82-155) for a scaled direction (Δx, Δy). */
typedef struct {
  double rmax, dual_obj, dual_viol, dual_mag;     /* dual ray r from Δy */
  double dmax, primal_cd, primal_viol, primal_cmag; /* primal ray d from Δx */
} psh_ray;

/* One chunk of fused r²HPDHG iterations (reference src/pdhg/r2hpdhg.cpp:66-85). */
typedef struct {
  int iterations;      /* K */
  long long inner0;    /* Halpern index k of the first iteration (w = (k+1)/(k+2)) */
  double tau, sigma, rho, eta, omega;
  int want_hat_last;   /* write x̂, ŷ, ȳ on the last iteration (a check follows) */
  int mode;            /* PSH_RED_* for the epoch-start residual r⁰ */
} psh_chunk_args;

/* What a check computes on the device, in one pass and one host round trip. */
typedef struct {
  double omega, eta;   /* for r = sqrt(ω‖Δx‖² + ‖Δy‖²/ω + 2η Δyᵀ Ã Δx) */
  int want_ray;        /* ray-test scalars of (Δx, Δy) = T(z) − z */
  int want_drift;      /* ray-test scalars of the drift (x̂ − x0, ŷ − y0) */
  int want_pid;        /* ‖x̂ − x0‖, ‖ŷ − y0‖ */
  int mode;            /* PSH_RED_* for the KKT sums and the residual / PID norms */
} psh_check_args;

typedef struct {
  double px, py, cross;      /* residual sums at this check */
  double r0_px, r0_py, r0_cross; /* residual sums after the first step of the last epoch start (if any) */
  int have_r0;
  psh_kkt kkt;
  psh_ray ray, drift;
  double dx_norm, dy_norm;   /* ‖x̂ − x0‖, ‖ŷ − y0‖ (want_pid) */
  int finite;                /* all of the above finite */
} psh_stats;

/* ---- device and context */
int psh_device_count(void);
int psh_mem_info(int device, size_t* free_bytes, size_t* total_bytes);
int psh_create(int device, psh_ctx** out);
void psh_destroy(psh_ctx* ctx);
const char* psh_error(const psh_ctx* ctx);
int psh_sync(psh_ctx* ctx);
size_t psh_device_bytes(const psh_ctx* ctx); /* device memory held by the context */

/* ---- problem and precision */
int psh_load(psh_ctx* ctx, const psh_problem* p);
/* Scaling on the device (reference src/pdhg/scaling.cpp:37-142), for a problem loaded with the
 * ORIGINAL values in a_val / at_val (at_val in Ãᵀ's order) and unit scales: geo_iterations
 * geometric-mean sweeps (0 = none), ruiz_iterations ∞-norm sweeps, then one Pock–Chambolle sweep
 * if asked, with the reference's operations in its order, so factors and values are the bits of
 * the host scaling. Leaves Ã, Ãᵀ scaled in place and the row / column factors in the device's
 * scale vectors, and writes them to row_scale (m) and col_scale (n). Once per context, before any
 * precision change. */
int psh_scale(psh_ctx* ctx, int geo_iterations, int ruiz_iterations, int pock_chambolle, double* row_scale,
              double* col_scale);
/* Replaces the scaled c, l, u (n), rl, ru (m) and the bound / objective scalars (after psh_scale). */
int psh_set_vectors(psh_ctx* ctx, const double* c, const double* l, const double* u, const double* rl, const double* ru,
                    double bound_scale, double obj_scale);
int psh_set_fp32(psh_ctx* ctx, int on); /* working precision: converts every slot (reference set_precision) */
int psh_is_fp32(const psh_ctx* ctx);

/* ---- the reference Backend operations (src/pdhg/backend.h:46-65), on slots */
int psh_upload(psh_ctx* ctx, int slot, const double* host);
int psh_download(psh_ctx* ctx, int slot, double* host);
int psh_fill(psh_ctx* ctx, int slot, double a);
int psh_copy(psh_ctx* ctx, int dst, int src);
int psh_axpby(psh_ctx* ctx, double a, int x, double b, int y);
int psh_spmv(psh_ctx* ctx, int x, int out);   /* out ← Ã x   */
int psh_spmv_t(psh_ctx* ctx, int y, int out); /* out ← Ãᵀ y  */
int psh_dot(psh_ctx* ctx, int a, int b, int mode, double* out);
int psh_diff_norm2(psh_ctx* ctx, int a, int b, int mode, double* out);
int psh_primal_step(psh_ctx* ctx, int x, int aty, double tau, int xhat);
int psh_dual_step(psh_ctx* ctx, int y, int ax, double sigma, int yhat);
int psh_r2h_primal(psh_ctx* ctx, int x, int x0, int aty, double tau, double w, double rho, int xhat, int xbar);
int psh_r2h_dual(psh_ctx* ctx, int y, int y0, int ax, double sigma, double w, double rho, int yhat, int ybar);
int psh_kkt_of(psh_ctx* ctx, int x_scaled, int y_scaled, int mode, psh_kkt* out);
/* Ray-test scalars of the direction in slots (dxs, dys). */
int psh_ray_of(psh_ctx* ctx, int dx_slot, int dy_slot, psh_ray* out);
/* The ray vectors of the last check / psh_ray_of in the original space: dual ray r (m) and
 * primal ray d (n) (reference src/pdhg/engine.cpp:90-97, 132-139); drift = 1 for the rays of
 * the drift direction. Either output may be NULL. */
int psh_ray_vectors(psh_ctx* ctx, int drift, double* dual_ray, double* primal_ray);

/* ---- fused fast path */
/* `iterations` r²HPDHG steps, two fused launches each (reference src/pdhg/r2hpdhg.cpp:66-85).
 * τ, σ and the epoch counter are written to device memory first, so every launch of a chunk
 * has fixed arguments: unless per-kernel timing is on or psh_graphs(ctx, 0) was called, the
 * launches are captured once per (precision, length, epoch start, last-step hats, mode) into a
 * HIP graph and replayed. */
int psh_chunk(psh_ctx* ctx, const psh_chunk_args* a);
int psh_graphs(psh_ctx* ctx, int on); /* 1 (default): replay captured chunk graphs */
/* Phased products for fp64 chunks: persistent workgroups sweep column slices of Ã (Ãᵀ) in order,
 * so the workgroups of an XCD gather from one L2-sized slice of x̄ (y) at a time while every row's
 * running sum keeps the CSR order: the same bits as the row-tiled kernels, in every reduction
 * mode. The fp64 check pass sums the original matrix over the same layouts (its values copied
 * into their order), again with the row tiles' bits. 1 (default): built at psh_load for a matrix
 * whose gathered vector exceeds 4 MiB and whose row tiles gather from scattered cache lines (at
 * least half a distinct 128-byte line per nonzero; banded or staircase matrices reuse lines and
 * keep the tiles), used by fp64 chunks and checks; 0: row tiles only; 2: built at any size
 * (tests). Call before psh_load for 0 and 2 to govern the build; afterwards 0 / nonzero only
 * switches the use of what was built. */
int psh_sliced(psh_ctx* ctx, int on);
/* Mixed precision in increment form (no counterpart in the reference; its mixed mode is
 * psh_set_fp32): the iterate is the fp64 anchor in X0 / Y0 plus an fp32 increment, the matrix
 * is read in fp32, and the anchor constants are rebuilt in fp64 at every restart, so the fp32
 * rounding is relative to the increment and each epoch refines the last. psh_chunk writes
 * x̂, x̄, ŷ, ȳ in fp64 on its hat steps, so psh_check / psh_restart work unchanged. Off
 * writes the current point into X / Y. Exclusive with psh_set_fp32(ctx, 1). `on` = 2 is the
 * variant with bf16 matrix values in the increment SpMVs (fp64 products and row sums, fp32
 * increments); 1 and 2 may be switched while on. */
int psh_delta(psh_ctx* ctx, int on);
int psh_check(psh_ctx* ctx, const psh_check_args* a, psh_stats* out);
int psh_restart(psh_ctx* ctx); /* x ← x̂, x0 ← x̂, y ← ŷ, y0 ← ŷ */

/* ---- ‖Ã‖₂ by Lanczos on ÃᵀÃ (reference src/la/csr.cpp:86-132): the host runs the
 * tridiagonal eigenvalue test, the device runs the products and the dots.
 * psh_lanczos_begin uploads the unit start vector v (n). psh_lanczos_step computes
 * w = ÃᵀÃ v, α = wᵀv, w ← w − (α v + β_prev v_prev), β = ‖w‖ (sum of squares returned
 * before the square root, as the host takes it). psh_lanczos_advance sets v_prev ← v, v ← w/β. */
int psh_lanczos_begin(psh_ctx* ctx, const double* v0);
int psh_lanczos_step(psh_ctx* ctx, double beta_prev, int mode, double* alpha, double* beta_sq);
int psh_lanczos_advance(psh_ctx* ctx, double beta);
/* w = ÃᵀÃ v for a host vector v (n), returned in w (n): the two products of one Lanczos step
 * (CSR order, the bits of the host products), for drivers that take the step's dots on the host
 * in the reference's sequential order. */
int psh_lanczos_product(psh_ctx* ctx, const double* v, double* w);

/* ---- batched scenarios sharing Ã (reference src/pdhg/batch.cpp:129-441), on a context that has
 * the shared scaled problem loaded (psh_load). Block vectors are interleaved: entry (row r,
 * slot t) at r·W + t, W ≤ PSH_BATCH_MAX_W slots. Scaled scenario data: c, l, u (n·W), rl, ru
 * (m·W); original-space data for the KKT: c_min = sense·c, cl, cu (n·W), rl, ru (m·W); per-slot
 * bound / objective scalars bs, os (W). */
enum { PSH_BATCH_MAX_W = 256 };
/* Per-slot statistics written by psh_batch_check, PSH_BSTAT_COUNT doubles per slot:
 * residual sums, KKT sums and maxima (termination.h), PID sums, residual sums of the last
 * epoch start, and (want_ray) the ray-test scalars of T(z) − z in psh_ray's order. */
enum {
  PSH_BSTAT_PX, PSH_BSTAT_PY, PSH_BSTAT_CROSS, PSH_BSTAT_RP2, PSH_BSTAT_B2, PSH_BSTAT_D, PSH_BSTAT_RD2,
  PSH_BSTAT_C2, PSH_BSTAT_P, PSH_BSTAT_PIDX, PSH_BSTAT_PIDY, PSH_BSTAT_PMAX, PSH_BSTAT_DMAX, PSH_BSTAT_CINF,
  PSH_BSTAT_R0_PX, PSH_BSTAT_R0_PY, PSH_BSTAT_R0_CROSS,
  PSH_BSTAT_RAY_RMAX, PSH_BSTAT_RAY_DUAL_OBJ, PSH_BSTAT_RAY_DUAL_VIOL, PSH_BSTAT_RAY_DUAL_MAG,
  PSH_BSTAT_RAY_DMAX, PSH_BSTAT_RAY_PRIMAL_CD, PSH_BSTAT_RAY_PRIMAL_VIOL, PSH_BSTAT_RAY_PRIMAL_CMAG,
  PSH_BSTAT_COUNT
};
int psh_batch_begin(psh_ctx* ctx, int W, const double* c, const double* l, const double* u, const double* rl,
                    const double* ru, const double* c_min, const double* cl_o, const double* cu_o,
                    const double* rl_o, const double* ru_o, const double* bs, const double* os);
/* x (n·W) and y (m·W), scaled: x = x0 = x̂, y = y0 = ŷ. */
int psh_batch_set_iterate(psh_ctx* ctx, const double* x, const double* y);
/* `iterations` lockstep r²HPDHG steps; per-slot tau, sigma (W) and Halpern index of the first step
 * inner0 (W). When some slot starts an epoch (inner0 = 0) the residual sums of every slot after
 * the first step are kept for the next psh_batch_check (reference batch.cpp:347-354). */
int psh_batch_chunk(psh_ctx* ctx, int iterations, const double* tau, const double* sigma, const long long* inner0,
                    double rho, int mode);
/* Per-slot statistics into out (W · PSH_BSTAT_COUNT); mode PSH_RED_* (CPU order = the reference's
 * sequential per-slot loops). want_ray: also the ray-test scalars of every slot (an addition:
 * the reference's batch has no infeasibility detection), the single solve's tests of
 * engine.cpp:82-184 there on the direction T(z) − z. */
int psh_batch_check(psh_ctx* ctx, int mode, int want_ray, double* out);
/* The unscaled dual ray (m) and primal ray (n) of slot t from the last want_ray check (either
 * pointer may be NULL). */
int psh_batch_ray_vectors(psh_ctx* ctx, int t, double* dual_ray, double* primal_ray);
/* x ← x̂, x0 ← x̂, y ← ŷ, y0 ← ŷ for the slots with mask[t] != 0. */
int psh_batch_restart(psh_ctx* ctx, const int* mask);
/* Mixed precision in increment form for every slot (no counterpart in the reference, whose batch
 * is fp64 only): each slot's iterate is its fp64 anchor x0 / y0 plus an fp32 increment, Ã is
 * read in fp32, and a slot's anchor constants are rebuilt in fp64 (one fp64 SpMM pass for the
 * re-anchored slots) at the chunk after its restart; x̂, x̄, ŷ, ȳ are written in fp64 on the
 * hat steps, so psh_batch_check / _restart / _point work unchanged. On: every slot anchors at
 * its x0 / y0 (call it where x = x0, y = y0: after psh_batch_set_iterate or a restart of all
 * slots). Off writes the current points into x / y. */
int psh_batch_delta(psh_ctx* ctx, int on);
/* Keeps the slots keep[0..W2) in that order (finished scenarios leave the batch). */
int psh_batch_compact(psh_ctx* ctx, const int* keep, int W2);
/* Unscaled point of slot t: x = clip(C x̂ / bs, l, u) (n), y_min = R ŷ / os (m). */
int psh_batch_point(psh_ctx* ctx, int t, double* x, double* y_min);

/* ---- allocation-free dot product. No counterpart in the reference: it is for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). out[0] ← Σ a[i]·b[i] for device arrays a, b of n doubles and a device output out, on
 * `stream` (a hipStream_t; NULL = the null stream), with psh_dot's reduction in `mode` (tree, or
 * the 64 fixed chunks of the CPU order for PSH_RED_CPU / PSH_RED_DECISION), so the result has
 * psh_dot's bits. Allocates nothing: its scratch (512 partial sums) is a static device array of
 * this library, one per device, whose address is looked up on the first call on a device. Every
 * call on a device shares that array and the last call leaves its partials there: calls that
 * overlap on one device (different streams, or host threads) race. 0 <= n <= PSH_NOALLOC_MAX_N,
 * else PSH_E_ARG; errors are described by psh_noalloc_error() (the calling thread's last one).
 * Asynchronous: out is written when the stream reaches the call. */
enum { PSH_NOALLOC_MAX_N = 2147483647 };
int psh_dot_noalloc(const double* a, const double* b, long long n, int mode, double* out, void* stream);
const char* psh_noalloc_error(void);

/* ---- per-kernel timing (HIP events on the context's stream), for the benchmark. */
int psh_timing(psh_ctx* ctx, int on);
/* kernel classes: 0 fused primal (Ãᵀy + projection/reflection/Halpern), 1 fused dual (Ãx̄ + ...),
 * 2 check pass; returns launches and total milliseconds. */
int psh_timing_get(psh_ctx* ctx, int kernel_class, long long* launches, double* total_ms);

/* ---- serving: independent LPs of different sizes as one packed block-diagonal system
 * (no counterpart in the reference). An arena of `budget_bytes` of device memory holds every
 * live instance (problem, original data, iterates); one launch of each fused half step advances
 * all instances that a chunk lists; per-instance sums are reduced inside each tile and then over
 * the instance's own tiles in a fixed order, so an instance's trajectory does not depend on the
 * others. Instances are added / removed between chunks; chunks replay a captured graph while the
 * instance set is unchanged. Host order per round: chunk → check → restart → remove / add. */
typedef struct psh_server psh_server;
typedef struct {
  int id;
  double tau, sigma;
  long long inner0;  /* Halpern index of the chunk's first iteration; 0 = the chunk begins an epoch */
  double rho;        /* reflection ρ of the instance (reference src/pdhg/engine.h:40, 1 = full) */
} psh_serve_param;
int psh_server_create(int device, size_t budget_bytes, int max_instances, psh_server** out);
void psh_server_destroy(psh_server* s);
const char* psh_server_error(const psh_server* s);
/* Instance flags. PSH_SERVE_INCREMENT: also hold fp32 values and the increment-form state, and
 * start the instance in increment form (mixed precision: fp32 matrix and increments around the
 * fp64 anchor x0 / y0, rebuilt in fp64 at the chunk after each restart; x̂, x̄, ŷ, ȳ are still
 * written in fp64 on the hat steps, so check / restart / download are unchanged). */
enum { PSH_SERVE_INCREMENT = 1 };
size_t psh_server_bytes(int m, int n, long long nnz, int flags); /* arena bytes of one instance */
size_t psh_server_free_bytes(const psh_server* s);               /* largest free block of the arena */
int psh_server_add(psh_server* s, const psh_problem* p, int flags, int* id); /* PSH_E_OOM when it does not fit */
/* 1: increment form (needs PSH_SERVE_INCREMENT storage), 0: fp64, from the next chunk. Switch
 * where x = x0 and y = y0 (after psh_server_set_iterate or a restart of the instance). */
int psh_server_set_precision(psh_server* s, int id, int mixed);
int psh_server_remove(psh_server* s, int id);
int psh_server_set_iterate(psh_server* s, int id, const double* x, const double* y);
int psh_server_chunk(psh_server* s, int iterations, const psh_serve_param* params, int count);
/* 1 (default): a chunk replays a graph captured while the instance set is unchanged; 0: every
 * chunk launches its kernels directly (per-kernel profiling, or callers that avoid graphs). */
int psh_server_graphs(psh_server* s, int on);
/* 1: HIP events on the server's stream around every chunk (a replayed graph has no host gaps
 * between its launches); psh_server_timing_get returns the iterations timed, their total
 * milliseconds and the algorithmic bytes of their two half steps (per iteration and active
 * instance: 24·nnz + 4(m+n+2) + 48m + 64n in fp64, 16·nnz + 4(m+n+2) + 24m + 28n in increment form). */
int psh_server_timing(psh_server* s, int on);
int psh_server_timing_get(psh_server* s, long long* iterations, double* total_ms, double* bytes);
/* want_ray: also psh_stats.ray (T(z) − z) of every listed instance, and its rays for
 * psh_server_ray_vectors (unscaled; dual ray m, primal ray n). */
int psh_server_check(psh_server* s, const int* ids, int count, int want_ray, psh_stats* out);
int psh_server_ray_vectors(psh_server* s, int id, double* dual_ray, double* primal_ray);
int psh_server_restart(psh_server* s, const int* ids, int count);
int psh_server_download(psh_server* s, int id, double* xh, double* yh);

#ifdef __cplusplus
}
#endif
#endif
