/* ps26119.h — C API of the MI355X r²HPDHG backend. Drop-in for the reference's C API
 * (reference include/ps26119/ps26119.h:19-94): the same functions, enums and struct layouts,
 * so C callers and the reference's ctypes binding (python/ps26119/__init__.py:29-40 there)
 * work unchanged against libps26119.so from this repository.
 *
 *   minimise / maximise  cᵀx + obj_offset  s.t.  row_lower ≤ A x ≤ row_upper,
 *                                                col_lower ≤ x ≤ col_upper
 * A in CSC (col_start[n+1], row_index[nnz], value[nnz]); infinity = HUGE_VAL.
 * Duals: z = c − Aᵀy with HiGHS signs. No function throws; every error is a status code.
 *
 * This library serves the GPU path of the reference (use_gpu = 1, algorithm r²HPDHG or Auto):
 * other algorithms and CPU requests return PS26119_NOT_SOLVED with the reason in
 * result->message. See ps26119_amd.h for the extensions (statistics, batches, serving).
 */
#ifndef PS26119_C_API_H
#define PS26119_C_API_H

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes — same values and order as the reference (ps26119.h:19-28 there). */
enum {
  PS26119_OPTIMAL = 0,
  PS26119_INFEASIBLE = 1,
  PS26119_UNBOUNDED = 2,
  PS26119_ITERATION_LIMIT = 3,
  PS26119_TIME_LIMIT = 4,
  PS26119_NUMERICAL_ERROR = 5,
  PS26119_NOT_SOLVED = 6,
  PS26119_INVALID_ARGUMENT = 7
};

enum { PS26119_ALG_AUTO = 0, PS26119_ALG_ORACLE = 1, PS26119_ALG_PDLP = 2, PS26119_ALG_R2HPDHG = 3, PS26119_ALG_SIMPLEX = 4 };
enum { PS26119_PREC_FP64 = 0, PS26119_PREC_MIXED = 1 };

typedef struct {
  int algorithm;             /* PS26119_ALG_* */
  int precision;             /* PS26119_PREC_* */
  int use_gpu;               /* 1 = run on the MI355X */
  double tolerance;          /* relative KKT target (0 = default 1e-8) */
  double time_limit;         /* seconds (0 = default); negative or NaN → PS26119_INVALID_ARGUMENT */
  long long iteration_limit; /* 0 = default; negative → PS26119_INVALID_ARGUMENT */
  int verbosity;
  int threads;               /* host threads for setup / verification (0 = all cores) */
  const double* warm_x;      /* warm start x (num_cols) or NULL */
  const double* warm_y;      /* warm start y (num_rows) or NULL */
} ps26119_options;

typedef struct {
  int status;
  double objective;
  double dual_objective;
  double primal_residual;
  double dual_residual;
  double gap;
  double certified_bound;    /* rounding-proof bound from y (lower for MIN, upper for MAX) */
  long long iterations;
  double seconds;
  char engine[16];
  char message[160];
  int check;                 /* in-process verification: 1 PASS, 0 FAIL, -1 n/a */
} ps26119_result;

const char* ps26119_version(void);
void ps26119_default_options(ps26119_options* opt);

int ps26119_solve_lp(int num_rows, int num_cols, int sense, double obj_offset, const double* c,
                     const double* col_lower, const double* col_upper, const double* row_lower,
                     const double* row_upper, const int* col_start, const int* row_index,
                     const double* value, const ps26119_options* opt, ps26119_result* result,
                     double* x, double* y, double* z);

int ps26119_solve_lp_ex(int num_rows, int num_cols, int sense, double obj_offset, const double* c,
                        const double* col_lower, const double* col_upper, const double* row_lower,
                        const double* row_upper, const int* col_start, const int* row_index,
                        const double* value, const ps26119_options* opt, ps26119_result* result,
                        double* x, double* y, double* z, double* dual_ray, double* primal_ray);

/* MPS input is not read by this library (the reference's MPS reader is not part of the GPU
 * path): always PS26119_INVALID_ARGUMENT with the reason in result->message. */
int ps26119_solve_mps(const char* path, const ps26119_options* opt, ps26119_result* result,
                      const char* solution_path);

#ifdef __cplusplus
}
#endif

#endif /* PS26119_C_API_H */
