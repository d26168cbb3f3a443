/* orc_abi.h — the extern "C" surface shared by the CPU checkers of this repository:
 *   oracle/liboracle.so            our CPU restatement of the reference r²HPDHG path (prefix orc_)
 *   oracle/_ref/libps26119_ref.so  the reference's own sources + oracle/ref_driver.cpp (prefix ref_)
 *   oracle/_ref/libps26119_ref_hip.so  the same with the MI355X backend answering the reference's
 *                                  GPU hook (oracle/ref_hip_hook.cpp), for use_gpu = 1 (prefix ref_)
 * All are TEST INFRASTRUCTURE: only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
 * leg load them, and only as the checker (ref_hip: as the caller of the backend under test). The
 * product never links or calls any of them.
 *
 * Model layout is the reference contract (include/ps26119/model.h of the reference): CSC A,
 * infinity = HUGE_VAL, sense +1 min / -1 max, objective value cᵀx + offset.
 */
#ifndef PS26119_ORC_ABI_H
#define PS26119_ORC_ABI_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int m, n, sense;
  double offset;
  const double *c, *col_lower, *col_upper, *row_lower, *row_upper;
  const int *col_start, *row_index;
  const double* value;
} orc_lp;

typedef struct {
  int algorithm;  /* 3 = r2hpdhg (the only one the restatement implements) */
  int precision;  /* 0 fp64, 1 mixed */
  int presolve;   /* 0/1 */
  int threads;    /* reference pool size; results never depend on it */
  int check_every;
  double tolerance, time_limit;
  long long iteration_limit;
  const double *warm_x, *warm_y;
  double warm_primal_weight;
  int n_params;
  const char* const* param_names;
  const double* param_values;
  int use_gpu;    /* reference build with the MI355X backend hooked in (_ref/libps26119_ref_hip.so) only */
} orc_opts;

typedef struct {
  int status;
  long long iterations, iterations_to_fast;
  double objective, dual_objective, primal_residual, dual_residual, gap;
  double certified_bound, primal_weight, seconds, setup_seconds;
  double check_primal, check_dual, check_gap;
  int check; /* 1 PASS, 0 FAIL, -1 none */
  char engine[32];
  char message[512];
} orc_result;

/* One scenario of a batch sharing the base matrix; NULL vectors mean "same as the base". */
typedef struct {
  const double *obj, *col_lower, *col_upper, *row_lower, *row_upper;
} orc_scenario;

/* A solution file's contents (the reference format, src/io/solution_writer.h of the reference),
 * to be written again by the reference's own writer (ref_write_solution). Strings are
 * NUL-terminated; a vector pointer is NULL when the file has no such block. */
typedef struct {
  const char *status, *engine, *precision, *message, *check;
  double objective, dual_objective, certified_bound, primal_residual, dual_residual, gap;
  long long iterations;
  double seconds, setup_seconds, primal_weight;
  long long iterations_to_fast;
  double seconds_to_fast, check_primal, check_dual, check_gap;
  const double *x, *z;           /* COLUMNS (n) */
  const double *activity, *y;    /* ROWS (m) */
  long long dual_ray_len, primal_ray_len;
  const double *dual_ray, *primal_ray;
} orc_solution_file;

#ifdef __cplusplus
}
#endif
#endif
