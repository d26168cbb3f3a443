// presolve.cpp — see presolve.h. A Reducer keeps working copies of the bounds and, per row and
// per column, the number of entries still alive; each round collects the rows (then the
// columns) a reduction applies to, in index order, and applies them one after another, because
// a bound set by one row is what the next row on the same column compares against.
#include "presolve.h"

#include <algorithm>
#include <cmath>

#include "la.h"
#include "verify.h"

namespace ps26119::amd {

namespace {

constexpr double kPresolveTol = 1e-9;  // relative slack of presolve's own feasibility tests

double slack(double bound) { return kPresolveTol * (1.0 + std::fabs(bound)); }

class Reducer {
 public:
  explicit Reducer(const Model& M)
      : M_(M), rows_(transpose(csc_as_transpose(M))), rl_(M.row_lower), ru_(M.row_upper), cl_(M.col_lower),
        cu_(M.col_upper), row_live_(M.num_rows), col_live_(M.num_cols), row_gone_(M.num_rows, 0),
        col_gone_(M.num_cols, 0), constant_(M.obj_offset) {
    for (int i = 0; i < M.num_rows; ++i) row_live_[i] = rows_.ptr[i + 1] - rows_.ptr[i];
    for (int j = 0; j < M.num_cols; ++j) col_live_[j] = M.col_start[j + 1] - M.col_start[j];
  }

  Presolved run() {
    for (;;) {
      bool rows_changed = false, cols_changed = false;
      if (!row_pass(rows_changed)) return std::move(out_);
      cols_changed = column_pass();
      if (!rows_changed && !cols_changed) break;
    }
    if (out_.rows_dropped == 0 && out_.cols_dropped == 0) {
      out_.outcome = Presolved::Outcome::NothingToDo;
      return std::move(out_);
    }
    assemble();
    out_.outcome = Presolved::Outcome::Reduced;
    return std::move(out_);
  }

 private:
  // Rows with at most one live entry, in index order. Returns false when a row proves the model
  // infeasible (out_ then says which).
  bool row_pass(bool& changed) {
    std::vector<int> candidates;
    for (int i = 0; i < M_.num_rows; ++i)
      if (!row_gone_[i] && row_live_[i] <= 1) candidates.push_back(i);
    for (int i : candidates) {
      if (row_live_[i] == 0) {
        if (!(0.0 >= rl_[i] - slack(rl_[i]) && 0.0 <= ru_[i] + slack(ru_[i])))
          return prove_infeasible("row " + std::to_string(i) + " has no entries and a range that excludes 0");
        row_gone_[i] = 1;
        ++out_.rows_dropped;
        out_.log.push_back({ReductionStep::Kind::DropEmptyRow, i, -1, 0.0, 0.0, false, false});
        changed = true;
        continue;
      }
      int col = -1;
      double a = 0.0;
      for (int p = rows_.ptr[i]; p < rows_.ptr[i + 1]; ++p)
        if (!col_gone_[rows_.col[p]]) col = rows_.col[p], a = rows_.val[p];
      if (col < 0 || a == 0.0) continue;
      // the row's range divided by its coefficient (infinite ends stay infinite)
      const double from = a > 0 ? rl_[i] / a : ru_[i] / a;
      const double to = a > 0 ? ru_[i] / a : rl_[i] / a;
      const bool lower_moves = from > cl_[col], upper_moves = to < cu_[col];
      double lower = lower_moves ? from : cl_[col];
      double upper = upper_moves ? to : cu_[col];
      if (lower > upper) {
        if (lower - upper > slack(lower))
          return prove_infeasible("row " + std::to_string(i) + " bounds column " + std::to_string(col) +
                                  " outside the column's own range");
        lower = upper = 0.5 * (lower + upper);  // within rounding: collapse to the midpoint
      }
      cl_[col] = lower;
      cu_[col] = upper;
      out_.log.push_back({ReductionStep::Kind::RowToBound, i, col, a, 0.0, lower_moves, upper_moves});
      out_.bounds_tightened += (lower_moves ? 1 : 0) + (upper_moves ? 1 : 0);
      row_gone_[i] = 1;
      --col_live_[col];
      ++out_.rows_dropped;
      changed = true;
    }
    return true;
  }

  // Fixed columns and columns with no live entry, in index order.
  bool column_pass() {
    std::vector<std::pair<int, double>> fixes;
    for (int j = 0; j < M_.num_cols; ++j) {
      if (col_gone_[j]) continue;
      if (cl_[j] == cu_[j]) {
        fixes.emplace_back(j, cl_[j]);
      } else if (col_live_[j] == 0) {
        const double v = best_bound(j);
        if (std::isfinite(v)) fixes.emplace_back(j, v);  // an unbounded direction is the engine's to find
      }
    }
    for (const auto& [j, v] : fixes) substitute(j, v);
    return !fixes.empty();
  }

  // The value an entry-free column takes at an optimum: its bound in the improving direction.
  double best_bound(int j) const {
    const double c = M_.sense * M_.obj[j];
    if (c > 0) return cl_[j];
    if (c < 0) return cu_[j];
    return std::min(std::max(0.0, cl_[j]), cu_[j]);
  }

  void substitute(int j, double v) {
    for (int p = M_.col_start[j]; p < M_.col_start[j + 1]; ++p) {
      const int i = M_.row_index[p];
      if (row_gone_[i]) continue;
      const double moved = M_.value[p] * v;
      if (std::isfinite(rl_[i])) rl_[i] -= moved;
      if (std::isfinite(ru_[i])) ru_[i] -= moved;
      --row_live_[i];
    }
    constant_ += M_.obj[j] * v;
    out_.log.push_back({ReductionStep::Kind::SubstituteColumn, -1, j, 0.0, v, false, false});
    col_gone_[j] = 1;
    ++out_.cols_dropped;
  }

  bool prove_infeasible(std::string why) {
    out_.outcome = Presolved::Outcome::ProvedInfeasible;
    out_.reason = std::move(why);
    return false;
  }

  // The reduced model: live rows and columns in their original order.
  void assemble() {
    std::vector<int> renumber(M_.num_rows, -1);
    for (int i = 0; i < M_.num_rows; ++i)
      if (!row_gone_[i]) {
        renumber[i] = static_cast<int>(out_.kept_rows.size());
        out_.kept_rows.push_back(i);
      }
    for (int j = 0; j < M_.num_cols; ++j)
      if (!col_gone_[j]) out_.kept_cols.push_back(j);
    Model& R = out_.reduced;
    R.name = M_.name;
    R.sense = M_.sense;
    R.obj_offset = constant_;
    R.num_rows = static_cast<int>(out_.kept_rows.size());
    R.num_cols = static_cast<int>(out_.kept_cols.size());
    R.col_start.assign(1, 0);
    for (int j : out_.kept_cols) {
      R.obj.push_back(M_.obj[j]);
      R.col_lower.push_back(cl_[j]);
      R.col_upper.push_back(cu_[j]);
      if (!M_.is_integer.empty()) R.is_integer.push_back(M_.is_integer[j]);
      for (int p = M_.col_start[j]; p < M_.col_start[j + 1]; ++p)
        if (renumber[M_.row_index[p]] >= 0) {
          R.row_index.push_back(renumber[M_.row_index[p]]);
          R.value.push_back(M_.value[p]);
        }
      R.col_start.push_back(static_cast<int>(R.value.size()));
    }
    for (int i : out_.kept_rows) {
      R.row_lower.push_back(rl_[i]);
      R.row_upper.push_back(ru_[i]);
    }
  }

  const Model& M_;
  const Csr rows_;  // A by rows, to find a singleton row's column
  std::vector<double> rl_, ru_, cl_, cu_;
  std::vector<int> row_live_, col_live_;  // entries in live columns / live rows
  std::vector<unsigned char> row_gone_, col_gone_;
  double constant_;
  Presolved out_;
};

// Σ over column j of a_ij·v_i subtracted from `start`, in the column's storage order.
double minus_column_dot(const Model& M, int j, const std::vector<double>& v, double start) {
  for (int p = M.col_start[j]; p < M.col_start[j + 1]; ++p) start -= M.value[p] * v[M.row_index[p]];
  return start;
}

}  // namespace

Presolved presolve(const Model& model) { return Reducer(model).run(); }

Solution postsolve(const Model& M, const Presolved& p, const Solution& red) {
  Solution s = red;
  const int m = M.num_rows, n = M.num_cols;
  const bool have_x = red.x.size() == p.kept_cols.size();
  const bool have_y = red.y.size() == p.kept_rows.size();
  s.x.assign(n, 0.0);
  s.y.assign(m, 0.0);
  if (have_x)
    for (std::size_t k = 0; k < p.kept_cols.size(); ++k) s.x[p.kept_cols[k]] = red.x[k];
  if (have_y)
    for (std::size_t k = 0; k < p.kept_rows.size(); ++k) s.y[p.kept_rows[k]] = red.y[k];
  // Undo in reverse: a substituted column gets its value back; a row that became a column bound
  // takes over the column's reduced cost when the column rests on a bound that row set (the
  // newest such row first owns it).
  for (auto it = p.log.rbegin(); it != p.log.rend(); ++it) {
    if (it->kind == ReductionStep::Kind::SubstituteColumn) {
      s.x[it->col] = it->value;
    } else if (it->kind == ReductionStep::Kind::RowToBound) {
      const double z = minus_column_dot(M, it->col, s.y, M.obj[it->col]);
      const double zmin = M.sense * z;
      if ((zmin > 0 && it->sets_lower) || (zmin < 0 && it->sets_upper)) s.y[it->row] = z / it->coef;
    }
  }
  s.z = M.obj;
  parallel_for(n, [&](std::int64_t b, std::int64_t e) {
    for (std::int64_t j = b; j < e; ++j) s.z[j] = minus_column_dot(M, static_cast<int>(j), s.y, M.obj[j]);
  });
  s.row_activity = activity(M, s.x);
  s.objective = M.objective_value(s.x);
  if (!have_x) {  // no point came back: no objective either
    s.x.clear();
    s.row_activity.clear();
    s.objective = kNaN;
  }
  if (!have_y) {
    s.y.clear();
    s.z.clear();
  }
  return s;
}

std::vector<double> lift_dual_ray(const Model& M, const Presolved& p, const std::vector<double>& red) {
  if (red.size() != p.kept_rows.size()) return {};
  std::vector<double> ray(M.num_rows, 0.0);
  for (std::size_t k = 0; k < p.kept_rows.size(); ++k) ray[p.kept_rows[k]] = red[k];
  // a Farkas vector is a dual with c = 0: the same recovery with λ_j = −(Aᵀr)_j
  for (auto it = p.log.rbegin(); it != p.log.rend(); ++it) {
    if (it->kind != ReductionStep::Kind::RowToBound) continue;
    const double lam = minus_column_dot(M, it->col, ray, 0.0);
    if ((lam > 0 && it->sets_lower) || (lam < 0 && it->sets_upper)) ray[it->row] = lam / it->coef;
  }
  return ray;
}

}  // namespace ps26119::amd
