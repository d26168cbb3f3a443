// model.cpp — Model helpers and status strings (reference src/core/model.cpp:53-180,
// src/core/status.cpp:9-74). The fingerprint is the reference's 64-bit FNV-1a over the same
// canonical byte stream (sizes, sense, offset, vectors, per-column entries sorted by row, the
// integrality flags; -0.0 hashed as 0.0, one NaN pattern), so a model hashes alike in both.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <sstream>

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

std::string Model::validate() const {
  const auto n = static_cast<std::size_t>(num_cols), m = static_cast<std::size_t>(num_rows);
  if (num_rows < 0 || num_cols < 0) return "negative dimension";
  if (sense != 1 && sense != -1) return "sense must be +1 or -1";
  if (!std::isfinite(obj_offset)) return "obj_offset is not finite";
  if (obj.size() != n) return "obj has wrong size";
  if (col_lower.size() != n || col_upper.size() != n) return "column bounds have wrong size";
  if (row_lower.size() != m || row_upper.size() != m) return "row bounds have wrong size";
  if (col_start.size() != n + 1) return "col_start must have num_cols+1 entries";
  if (col_start[0] != 0) return "col_start[0] must be 0";
  if (row_index.size() != value.size()) return "row_index and value differ in size";
  if (static_cast<std::size_t>(col_start[n]) != value.size()) return "col_start[n] != nnz";
  if (!is_integer.empty() && is_integer.size() != n) return "is_integer has wrong size";
  std::vector<int> seen(m, -1);
  std::ostringstream err;
  for (std::size_t j = 0; j < n; ++j) {
    if (col_start[j + 1] < col_start[j]) {
      err << "col_start not monotone at column " << j;
      return err.str();
    }
    if (!std::isfinite(obj[j])) {
      err << "objective coefficient of column " << j << " is not finite";
      return err.str();
    }
    for (int k = col_start[j]; k < col_start[j + 1]; ++k) {
      const int i = row_index[k];
      if (i < 0 || static_cast<std::size_t>(i) >= m) {
        err << "row index out of range in column " << j;
        return err.str();
      }
      if (seen[i] == static_cast<int>(j)) {
        err << "duplicate entry (" << i << "," << j << ")";
        return err.str();
      }
      seen[i] = static_cast<int>(j);
      if (!std::isfinite(value[k])) {
        err << "matrix entry (" << i << "," << j << ") is not finite";
        return err.str();
      }
    }
  }
  auto bounds = [](const std::vector<double>& lo, const std::vector<double>& up, const char* what) -> std::string {
    for (std::size_t k = 0; k < lo.size(); ++k) {
      std::ostringstream e;
      if (std::isnan(lo[k]) || std::isnan(up[k])) e << what << " bound " << k << " is NaN";
      else if (lo[k] == kInf) e << what << " lower bound " << k << " is +inf";
      else if (up[k] == -kInf) e << what << " upper bound " << k << " is -inf";
      if (!e.str().empty()) return e.str();
    }
    return {};
  };
  if (auto s = bounds(col_lower, col_upper, "column"); !s.empty()) return s;
  return bounds(row_lower, row_upper, "row");
}

std::string Model::crossed_bounds() const {
  auto first = [](const std::vector<double>& lo, const std::vector<double>& up, const char* what) -> std::string {
    for (std::size_t k = 0; k < lo.size() && k < up.size(); ++k)
      if (lo[k] > up[k]) {
        std::ostringstream e;
        e << what << " " << k << " has lower bound " << lo[k] << " > upper bound " << up[k];
        return e.str();
      }
    return {};
  };
  if (auto s = first(col_lower, col_upper, "column"); !s.empty()) return s;
  return first(row_lower, row_upper, "row");
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
