// contract.h — the solver contract types, with the reference's field names and meanings
// (reference include/ps26119/model.h:22-68, options.h:15-82, solution.h:17-84,
// tolerances.h:7-46). They must stay alike: the C API, the Python binding and callers of the
// reference read and write exactly these fields. Own code.
#pragma once

#include <cstdint>
#include <limits>
#include <string>
#include <utility>
#include <vector>

namespace ps26119 {

inline constexpr double kInf = std::numeric_limits<double>::infinity();
inline constexpr double kNaN = std::numeric_limits<double>::quiet_NaN();

namespace tol {  // reference include/ps26119/tolerances.h
inline constexpr double kFirstOrderFast = 1e-4;
inline constexpr double kFirstOrderHigh = 1e-8;
inline constexpr double kFirstOrderInfeasible = 1e-8;
inline constexpr double kMixedPrecisionSwitch = 1e-6;
inline constexpr double kVerifyPrimal = 1e-6;
inline constexpr double kVerifyDual = 1e-6;
inline constexpr double kVerifyGap = 1e-6;
inline constexpr double kVerifyRay = 1e-8;
}  // namespace tol

// minimise / maximise cᵀx + obj_offset  s.t.  row_lower ≤ A x ≤ row_upper, col_lower ≤ x ≤ col_upper
// A in CSC (col_start n+1, row_index/value nnz); infinity = +inf.
struct Model {
  std::string name;
  int num_rows = 0;
  int num_cols = 0;
  int sense = 1;
  double obj_offset = 0.0;
  std::vector<double> obj;
  std::vector<double> col_lower, col_upper;
  std::vector<double> row_lower, row_upper;
  std::vector<int> col_start;
  std::vector<int> row_index;
  std::vector<double> value;
  std::vector<std::uint8_t> is_integer;

  std::size_t nnz() const { return value.size(); }
  std::uint64_t fingerprint() const;
  std::string fingerprint_hex() const;
  std::string validate() const;
  std::string crossed_bounds() const;
  double objective_value(const std::vector<double>& x) const;
  std::vector<double> row_activity(const std::vector<double>& x) const;
};

enum class Status { Optimal, Infeasible, Unbounded, IterationLimit, TimeLimit, NumericalError, NotSolved };
enum class Algorithm { Auto, Oracle, Pdlp, R2hpdhg, Simplex };
enum class Precision { Fp64, Mixed };

const char* to_string(Status s);
const char* to_string(Algorithm a);
const char* to_string(Precision p);

struct Options {
  Algorithm algorithm = Algorithm::Auto;
  Precision precision = Precision::Fp64;
  bool use_gpu = false;
  double tolerance = tol::kFirstOrderHigh;
  double time_limit = 3600.0;
  std::int64_t iteration_limit = 100'000'000;
  int verbosity = 0;
  double kkt_b_norm = -1, kkt_c_norm = -1, kkt_obj_shift = 0;
  bool presolve = true;
  bool relax_integrality = false;
  int threads = 1;
  int ruiz_iterations = 10;
  bool pock_chambolle = true;
  int termination_check_every = 64;
  std::vector<double> warm_x, warm_y;
  double warm_primal_weight = 0.0;
  std::vector<std::pair<std::string, double>> engine_params;
};

struct Solution {
  Status status = Status::NotSolved;
  std::vector<double> x, row_activity, y, z;
  double objective = kNaN, dual_objective = kNaN;
  double primal_residual = kNaN, dual_residual = kNaN, gap = kNaN;
  std::int64_t iterations = 0;
  double seconds = 0.0, setup_seconds = 0.0;
  std::int64_t iterations_to_fast = -1;
  double seconds_to_fast = -1.0;
  double certified_bound = kNaN;
  double primal_weight = kNaN;
  std::string check;
  double check_primal = kNaN, check_dual = kNaN, check_gap = kNaN;
  std::vector<double> dual_ray, primal_ray;
  std::string engine, precision, model_fingerprint, message;
  // MI355X extras (not in the reference's Solution): per-kernel timing when requested, and
  // the device memory the solve held.
  long long kernel_launches[3] = {0, 0, 0};
  double kernel_ms[3] = {0, 0, 0};
  double device_bytes = 0;
  double iterate_seconds = 0;  // wall time of the iteration loop (chunks, checks, restarts)
};

// Shared-matrix scenario (reference include/ps26119/batch.h:23-29); empty = same as the base.
struct Scenario {
  std::vector<double> obj;
  std::vector<double> col_lower, col_upper;
  std::vector<double> row_lower, row_upper;
};

}  // namespace ps26119
