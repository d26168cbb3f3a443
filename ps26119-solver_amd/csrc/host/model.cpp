// model.cpp — Model helpers and status strings (reference src/core/model.cpp:53-180,
// src/core/status.cpp:9-74). The fingerprint is the reference's 64-bit FNV-1a over the same
// canonical byte stream (sizes, sense, offset, vectors, per-column entries sorted by row, the
// integrality flags; -0.0 hashed as 0.0, one NaN pattern), so a model hashes alike in both.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>

#include "contract.h"

namespace ps26119 {

namespace {

struct Fnv {
  std::uint64_t h = 0xcbf29ce484222325ULL;
  void bytes(const void* p, std::size_t n) {
    const auto* b = static_cast<const unsigned char*>(p);
    for (std::size_t i = 0; i < n; ++i) h = (h ^ b[i]) * 0x100000001b3ULL;
  }
  void u64(std::uint64_t v) { bytes(&v, 8); }
  void i64(std::int64_t v) { bytes(&v, 8); }
  void f64(double v) {
    std::uint64_t bits = 0x7ff8000000000000ULL;
    if (!std::isnan(v)) {
      if (v == 0.0) v = 0.0;
      std::memcpy(&bits, &v, 8);
    }
    u64(bits);
  }
  void vec(const std::vector<double>& v) {
    u64(v.size());
    for (double d : v) f64(d);
  }
};

}  // namespace

std::uint64_t Model::fingerprint() const {
  Fnv f;
  f.i64(num_rows);
  f.i64(num_cols);
  f.i64(sense);
  f.f64(obj_offset);
  f.vec(obj);
  f.vec(col_lower);
  f.vec(col_upper);
  f.vec(row_lower);
  f.vec(row_upper);
  f.u64(value.size());
  std::vector<std::pair<int, double>> e;
  for (int j = 0; j < num_cols; ++j) {
    e.clear();
    for (int k = col_start[j]; k < col_start[j + 1]; ++k) e.emplace_back(row_index[k], value[k]);
    std::sort(e.begin(), e.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
    f.i64(static_cast<std::int64_t>(e.size()));
    for (const auto& [i, v] : e) {
      f.i64(i);
      f.f64(v);
    }
  }
  for (int j = 0; j < num_cols; ++j) f.i64(j < static_cast<int>(is_integer.size()) && is_integer[j] ? 1 : 0);
  return f.h;
}

std::string Model::fingerprint_hex() const {
  char b[17];
  std::snprintf(b, sizeof b, "%016llx", static_cast<unsigned long long>(fingerprint()));
  return b;
}

// The first defect of the model contract found, in the order: sizes and sense, then column by
// column (pointers, objective, entries), then the bound vectors; "" when the model is usable.
std::string Model::validate() const {
  const std::size_t n = num_cols < 0 ? 0 : static_cast<std::size_t>(num_cols);
  const std::size_t m = num_rows < 0 ? 0 : static_cast<std::size_t>(num_rows);
  const struct {
    bool bad;
    const char* what;
  } shape[] = {
      {num_rows < 0 || num_cols < 0, "a dimension is negative"},
      {sense != 1 && sense != -1, "sense is neither +1 (min) nor -1 (max)"},
      {!std::isfinite(obj_offset), "the objective offset is not finite"},
      {obj.size() != n, "the objective does not have num_cols entries"},
      {col_lower.size() != n || col_upper.size() != n, "a column bound vector does not have num_cols entries"},
      {row_lower.size() != m || row_upper.size() != m, "a row bound vector does not have num_rows entries"},
      {col_start.size() != n + 1, "col_start does not have num_cols + 1 entries"},
      {!col_start.empty() && col_start[0] != 0, "col_start does not begin at 0"},
      {row_index.size() != value.size(), "row_index and value differ in length"},
      {col_start.size() == n + 1 && static_cast<std::size_t>(col_start[n]) != value.size(),
       "col_start does not end at the number of entries"},
      {!is_integer.empty() && is_integer.size() != n, "is_integer does not have num_cols entries"},
  };
  for (const auto& c : shape)
    if (c.bad) return c.what;
  char buf[160];
  std::vector<int> last_col(m, -1);  // the column that last used each row: duplicate detection
  for (std::size_t j = 0; j < n; ++j) {
    const int col = static_cast<int>(j);
    if (col_start[j + 1] < col_start[j]) {
      std::snprintf(buf, sizeof buf, "col_start decreases at column %d", col);
      return buf;
    }
    if (!std::isfinite(obj[j])) {
      std::snprintf(buf, sizeof buf, "column %d has a non-finite objective coefficient", col);
      return buf;
    }
    for (int k = col_start[j]; k < col_start[j + 1]; ++k) {
      const int i = row_index[k];
      if (i < 0 || static_cast<std::size_t>(i) >= m)
        std::snprintf(buf, sizeof buf, "column %d names row %d, outside [0, %d)", col, i, num_rows);
      else if (last_col[i] == col)
        std::snprintf(buf, sizeof buf, "column %d has two entries in row %d", col, i);
      else if (!std::isfinite(value[k]))
        std::snprintf(buf, sizeof buf, "entry (%d, %d) is not finite", i, col);
      else {
        last_col[i] = col;
        continue;
      }
      return buf;
    }
  }
  const struct {
    const std::vector<double>& lo;
    const std::vector<double>& up;
    const char* kind;
  } ranges[] = {{col_lower, col_upper, "column"}, {row_lower, row_upper, "row"}};
  for (const auto& r : ranges)
    for (std::size_t k = 0; k < r.lo.size(); ++k) {
      const char* flaw = std::isnan(r.lo[k]) || std::isnan(r.up[k]) ? "a NaN bound"
                         : r.lo[k] == kInf                          ? "a lower bound of +inf"
                         : r.up[k] == -kInf                         ? "an upper bound of -inf"
                                                                    : nullptr;
      if (flaw) {
        std::snprintf(buf, sizeof buf, "%s %zu has %s", r.kind, k, flaw);
        return buf;
      }
    }
  return {};
}

// The first column, then row, whose lower bound exceeds its upper bound; "" when none does.
std::string Model::crossed_bounds() const {
  const struct {
    const std::vector<double>& lo;
    const std::vector<double>& up;
    const char* kind;
  } ranges[] = {{col_lower, col_upper, "column"}, {row_lower, row_upper, "row"}};
  for (const auto& r : ranges) {
    const std::size_t len = std::min(r.lo.size(), r.up.size());
    for (std::size_t k = 0; k < len; ++k)
      if (r.lo[k] > r.up[k]) {
        char buf[160];
        std::snprintf(buf, sizeof buf, "%s %zu needs at least %.17g and at most %.17g", r.kind, k, r.lo[k], r.up[k]);
        return buf;
      }
  }
  return {};
}

double Model::objective_value(const std::vector<double>& x) const {
  double v = obj_offset;
  for (int j = 0; j < num_cols; ++j) v += obj[j] * x[j];
  return v;
}

std::vector<double> Model::row_activity(const std::vector<double>& x) const {
  std::vector<double> ax(num_rows, 0.0);
  for (int j = 0; j < num_cols; ++j)
    for (int k = col_start[j]; k < col_start[j + 1]; ++k) ax[row_index[k]] += value[k] * x[j];
  return ax;
}

const char* to_string(Status s) {
  switch (s) {
    case Status::Optimal: return "Optimal";
    case Status::Infeasible: return "Infeasible";
    case Status::Unbounded: return "Unbounded";
    case Status::IterationLimit: return "IterationLimit";
    case Status::TimeLimit: return "TimeLimit";
    case Status::NumericalError: return "NumericalError";
    case Status::NotSolved: return "NotSolved";
  }
  return "NotSolved";
}

const char* to_string(Algorithm a) {
  switch (a) {
    case Algorithm::Auto: return "auto";
    case Algorithm::Oracle: return "oracle";
    case Algorithm::Pdlp: return "pdlp";
    case Algorithm::R2hpdhg: return "r2hpdhg";
    case Algorithm::Simplex: return "simplex";
  }
  return "auto";
}

const char* to_string(Precision p) { return p == Precision::Mixed ? "mixed" : "fp64"; }

}  // namespace ps26119
