// presolve.cpp — see presolve.h. Row access comes from a CSR built once; the reduction
// passes themselves are sequential because each reduction changes what the next one sees.
#include "presolve.h"

#include <algorithm>
#include <cmath>

#include "la.h"
#include "verify.h"

namespace ps26119::amd {

namespace {
constexpr double kFeasTol = 1e-9;
bool within(double v, double lo, double up) {
  return v >= lo - kFeasTol * (1 + std::fabs(lo)) && v <= up + kFeasTol * (1 + std::fabs(up));
}
}  // namespace

PresolveResult presolve(const Model& M) {
  PresolveResult r;
  const int m = M.num_rows, n = M.num_cols;
  const Csr A = transpose(csc_as_transpose(M));
  std::vector<double> rl = M.row_lower, ru = M.row_upper, cl = M.col_lower, cu = M.col_upper;
  std::vector<char> row_alive(m, 1), col_alive(n, 1);
  std::vector<int> row_cnt(m), col_cnt(n);
  for (int i = 0; i < m; ++i) row_cnt[i] = A.ptr[i + 1] - A.ptr[i];
  for (int j = 0; j < n; ++j) col_cnt[j] = M.col_start[j + 1] - M.col_start[j];
  r.fixed_value.assign(n, kNaN);
  double offset = M.obj_offset;
  auto infeasible = [&](const std::string& why) {
    r.outcome = PresolveResult::Outcome::Infeasible;
    r.message = why;
    return r;
  };
  bool changed = true;
  while (changed) {
    changed = false;
    for (int i = 0; i < m; ++i) {  // R1 and R4
      if (!row_alive[i]) continue;
      if (row_cnt[i] == 0) {
        if (!within(0.0, rl[i], ru[i])) return infeasible("empty row " + std::to_string(i) + " excludes 0");
        row_alive[i] = 0;
        ++r.removed_rows;
        changed = true;
      } else if (row_cnt[i] == 1) {
        int j = -1;
        double a = 0;
        for (int p = A.ptr[i]; p < A.ptr[i + 1]; ++p)
          if (col_alive[A.col[p]]) j = A.col[p], a = A.val[p];
        if (j < 0 || a == 0.0) continue;
        const double lo = a > 0 ? rl[i] / a : ru[i] / a;
        const double hi = a > 0 ? ru[i] / a : rl[i] / a;
        const bool set_lower = lo > cl[j], set_upper = hi < cu[j];
        double nl = set_lower ? lo : cl[j], nu = set_upper ? hi : cu[j];
        if (nl > nu) {
          if (nl - nu > kFeasTol * (1 + std::fabs(nl)))
            return infeasible("singleton row " + std::to_string(i) + " conflicts with the bounds of its column");
          nl = nu = 0.5 * (nl + nu);
        }
        cl[j] = nl;
        cu[j] = nu;
        r.singletons.push_back({i, j, a, set_lower, set_upper});
        r.tightened_bounds += set_lower + set_upper;
        row_alive[i] = 0;
        --col_cnt[j];
        ++r.removed_rows;
        changed = true;
      }
    }
    for (int j = 0; j < n; ++j) {  // R2 and R3
      if (!col_alive[j]) continue;
      double v;
      if (cl[j] == cu[j]) {
        v = cl[j];
      } else if (col_cnt[j] == 0) {
        const double c = M.sense * M.obj[j];
        v = c > 0 ? cl[j] : (c < 0 ? cu[j] : std::min(std::max(0.0, cl[j]), cu[j]));
        if (!std::isfinite(v)) continue;
      } else {
        continue;
      }
      for (int p = M.col_start[j]; p < M.col_start[j + 1]; ++p) {
        const int i = M.row_index[p];
        if (!row_alive[i]) continue;
        const double t = M.value[p] * v;
        if (std::isfinite(rl[i])) rl[i] -= t;
        if (std::isfinite(ru[i])) ru[i] -= t;
        --row_cnt[i];
      }
      offset += M.obj[j] * v;
      r.fixed_value[j] = v;
      col_alive[j] = 0;
      ++r.removed_cols;
      changed = true;
    }
  }
  if (r.removed_rows == 0 && r.removed_cols == 0) {
    r.outcome = PresolveResult::Outcome::Unchanged;
    return r;
  }
  Model& R = r.reduced;
  std::vector<int> new_row(m, -1);
  for (int i = 0; i < m; ++i)
    if (row_alive[i]) new_row[i] = static_cast<int>(r.row_map.size()), r.row_map.push_back(i);
  for (int j = 0; j < n; ++j)
    if (col_alive[j]) r.col_map.push_back(j);
  R.name = M.name;
  R.sense = M.sense;
  R.obj_offset = offset;
  R.num_rows = static_cast<int>(r.row_map.size());
  R.num_cols = static_cast<int>(r.col_map.size());
  R.col_start.reserve(R.num_cols + 1);
  R.col_start.push_back(0);
  for (int j : r.col_map) {
    R.obj.push_back(M.obj[j]);
    R.col_lower.push_back(cl[j]);
    R.col_upper.push_back(cu[j]);
    for (int p = M.col_start[j]; p < M.col_start[j + 1]; ++p) {
      const int i = M.row_index[p];
      if (new_row[i] >= 0) R.row_index.push_back(new_row[i]), R.value.push_back(M.value[p]);
    }
    R.col_start.push_back(static_cast<int>(R.value.size()));
    if (!M.is_integer.empty()) R.is_integer.push_back(M.is_integer[j]);
  }
  for (int i : r.row_map) {
    R.row_lower.push_back(rl[i]);
    R.row_upper.push_back(ru[i]);
  }
  r.outcome = PresolveResult::Outcome::Reduced;
  return r;
}

Solution postsolve(const Model& M, const PresolveResult& r, const Solution& red) {
  Solution s = red;
  const int m = M.num_rows, n = M.num_cols;
  s.x.assign(n, 0.0);
  s.y.assign(m, 0.0);
  const bool have_x = red.x.size() == r.col_map.size();
  const bool have_y = red.y.size() == r.row_map.size();
  for (int j = 0; j < n; ++j)
    if (!std::isnan(r.fixed_value[j])) s.x[j] = r.fixed_value[j];
  if (have_x)
    for (std::size_t c = 0; c < r.col_map.size(); ++c) s.x[r.col_map[c]] = red.x[c];
  if (have_y)
    for (std::size_t q = 0; q < r.row_map.size(); ++q) s.y[r.row_map[q]] = red.y[q];
  // a singleton row takes the reduced cost of its column when the column sits on the bound
  // that row created (newest first)
  for (auto it = r.singletons.rbegin(); it != r.singletons.rend(); ++it) {
    const int j = it->col;
    double z = M.obj[j];
    for (int p = M.col_start[j]; p < M.col_start[j + 1]; ++p) z -= M.value[p] * s.y[M.row_index[p]];
    const double zm = M.sense * z;
    if ((zm > 0 && it->set_lower) || (zm < 0 && it->set_upper)) s.y[it->row] = z / it->a;
  }
  s.z = M.obj;
  parallel_for(n, [&](std::int64_t b, std::int64_t e) {
    for (std::int64_t j = b; j < e; ++j)
      for (int p = M.col_start[j]; p < M.col_start[j + 1]; ++p) s.z[j] -= M.value[p] * s.y[M.row_index[p]];
  });
  s.row_activity = activity(M, s.x);
  s.objective = M.objective_value(s.x);
  if (!have_x) {
    s.x.clear();
    s.row_activity.clear();
    s.objective = kNaN;
  }
  if (!have_y) s.y.clear(), s.z.clear();
  return s;
}

std::vector<double> postsolve_farkas(const Model& M, const PresolveResult& r, const std::vector<double>& red) {
  std::vector<double> ray(M.num_rows, 0.0);
  if (red.size() != r.row_map.size()) return {};
  for (std::size_t q = 0; q < r.row_map.size(); ++q) ray[r.row_map[q]] = red[q];
  for (auto it = r.singletons.rbegin(); it != r.singletons.rend(); ++it) {
    const int j = it->col;
    double lam = 0.0;
    for (int p = M.col_start[j]; p < M.col_start[j + 1]; ++p) lam -= M.value[p] * ray[M.row_index[p]];
    if ((lam > 0 && it->set_lower) || (lam < 0 && it->set_upper)) ray[it->row] = lam / it->a;
  }
  return ray;
}

}  // namespace ps26119::amd
