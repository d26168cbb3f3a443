// scaling.cpp — see scaling.h.
#include "scaling.h"

#include <cmath>

namespace ps26119::amd {

namespace {

double inv_sqrt_or_one(double v) { return v > 0.0 && std::isfinite(v) ? 1.0 / std::sqrt(v) : 1.0; }

// Transpose of the structure of `a` with, for every transposed entry, its index in `a`.
void transpose_structure(const Csr& a, Csr& t, std::vector<int>& perm) {
  t.rows = a.cols;
  t.cols = a.rows;
  const std::int64_t nz = a.nnz();
  t.ptr.assign(static_cast<std::size_t>(t.rows) + 1, 0);
  for (std::int64_t k = 0; k < nz; ++k) ++t.ptr[a.col[k] + 1];
  for (int r = 0; r < t.rows; ++r) t.ptr[r + 1] += t.ptr[r];
  t.col.resize(nz);
  perm.resize(nz);
  std::vector<int> next(t.ptr.begin(), t.ptr.end() - 1);
  for (int r = 0; r < a.rows; ++r)
    for (int k = a.ptr[r]; k < a.ptr[r + 1]; ++k) {
      const int q = next[a.col[k]]++;
      t.col[q] = r;
      perm[q] = k;
    }
}

// Structure, original values and the scaling decisions (no sweeps yet).
ScaledProblem scaled_structure(const Model& model, const ScalingOptions& opt, std::vector<int>& perm) {
  ScaledProblem sp;
  const int m = model.num_rows, n = model.num_cols;
  sp.m = m;
  sp.n = n;
  sp.sense = model.sense;
  sp.original = &model;
  sp.A = transpose(csc_as_transpose(model));  // CSR of A, columns sorted
  sp.a_orig = sp.A.val;
  transpose_structure(sp.A, sp.At, perm);
  sp.row_scale.assign(m, 1.0);
  sp.col_scale.assign(n, 1.0);
  {
    double amin = kInf, amax = 0.0;
    for (double v : sp.a_orig)
      if (v != 0.0) amin = std::min(amin, std::fabs(v)), amax = std::max(amax, std::fabs(v));
    if (amax > 0) sp.log10_range = std::log10(amax / amin);
  }
  sp.geometric_mean_applied =
      opt.geometric_mean_iterations > 0 && sp.log10_range >= opt.geometric_mean_min_log10_range;
  return sp;
}

void at_values(ScaledProblem& sp, const std::vector<int>& perm) {
  sp.At.val.resize(perm.size());
  parallel_for(static_cast<std::int64_t>(perm.size()), [&](std::int64_t b, std::int64_t e) {
    for (std::int64_t q = b; q < e; ++q) sp.At.val[q] = sp.A.val[perm[q]];
  });
}

}  // namespace

ScaledProblem make_unscaled_problem(const Model& model, const ScalingOptions& opt) {
  std::vector<int> perm;
  ScaledProblem sp = scaled_structure(model, opt, perm);
  at_values(sp, perm);
  finish_scaled_problem(sp, model, ScalingOptions{opt.geometric_mean_iterations, opt.geometric_mean_min_log10_range,
                                                  opt.ruiz_iterations, opt.pock_chambolle, false});
  return sp;
}

ScaledProblem make_scaled_problem(const Model& model, const ScalingOptions& opt) {
  std::vector<int> perm;
  ScaledProblem sp = scaled_structure(model, opt, perm);
  const int m = sp.m, n = sp.n;
  const Csr& A = sp.A;
  const Csr& At = sp.At;
  std::vector<double>& val = sp.A.val;

  std::vector<double> rf(m), cf(n);
  if (sp.geometric_mean_applied) {
    // factors from the ORIGINAL entries each sweep, rows then columns, applied once at the end
    std::vector<double> r(m, 1.0), c(n, 1.0);
    auto clampf = [](double v) { return std::min(std::max(v, 1e-20), 1e20); };
    const std::vector<double>& a0 = sp.a_orig;
    for (int it = 0; it < opt.geometric_mean_iterations; ++it) {
      parallel_for(m, [&](std::int64_t b, std::int64_t e) {
        for (std::int64_t i = b; i < e; ++i) {
          double lo = kInf, hi = 0.0;
          for (int k = A.ptr[i]; k < A.ptr[i + 1]; ++k) {
            const double v = std::fabs(a0[k]) * c[A.col[k]];
            if (v > 0 && std::isfinite(v)) lo = std::min(lo, v), hi = std::max(hi, v);
          }
          if (hi > 0) r[i] = clampf(1.0 / (std::sqrt(lo) * std::sqrt(hi)));
        }
      });
      parallel_for(n, [&](std::int64_t b, std::int64_t e) {
        for (std::int64_t j = b; j < e; ++j) {
          double lo = kInf, hi = 0.0;
          for (int q = At.ptr[j]; q < At.ptr[j + 1]; ++q) {
            const double v = std::fabs(a0[perm[q]]) * r[At.col[q]];
            if (v > 0 && std::isfinite(v)) lo = std::min(lo, v), hi = std::max(hi, v);
          }
          if (hi > 0) c[j] = clampf(1.0 / (std::sqrt(lo) * std::sqrt(hi)));
        }
      });
    }
    parallel_for(m, [&](std::int64_t b, std::int64_t e) {
      for (std::int64_t i = b; i < e; ++i)
        for (int k = A.ptr[i]; k < A.ptr[i + 1]; ++k) val[k] *= r[i] * c[A.col[k]];
    });
    sp.row_scale = r;
    sp.col_scale = c;
  }

  auto sweep = [&](bool inf_norm) {
    parallel_for(m, [&](std::int64_t b, std::int64_t e) {
      for (std::int64_t i = b; i < e; ++i) {
        double s = 0.0;
        for (int k = A.ptr[i]; k < A.ptr[i + 1]; ++k) {
          const double a = std::fabs(val[k]);
          s = inf_norm ? std::max(s, a) : s + a;
        }
        rf[i] = inv_sqrt_or_one(s);
      }
    });
    parallel_for(n, [&](std::int64_t b, std::int64_t e) {
      for (std::int64_t j = b; j < e; ++j) {
        double s = 0.0;
        for (int q = At.ptr[j]; q < At.ptr[j + 1]; ++q) {
          const double a = std::fabs(val[perm[q]]);
          s = inf_norm ? std::max(s, a) : s + a;
        }
        cf[j] = inv_sqrt_or_one(s);
      }
    });
    parallel_for(m, [&](std::int64_t b, std::int64_t e) {
      for (std::int64_t i = b; i < e; ++i) {
        for (int k = A.ptr[i]; k < A.ptr[i + 1]; ++k) val[k] *= rf[i] * cf[A.col[k]];
        sp.row_scale[i] *= rf[i];
      }
    });
    for (int j = 0; j < n; ++j) sp.col_scale[j] *= cf[j];
  };
  for (int it = 0; it < opt.ruiz_iterations; ++it) sweep(true);
  if (opt.pock_chambolle) sweep(false);

  finish_scaled_problem(sp, model, opt);
  at_values(sp, perm);
  return sp;
}

void finish_scaled_problem(ScaledProblem& sp, const Model& model, const ScalingOptions& opt) {
  const int m = sp.m, n = sp.n;
  sp.c.resize(n);
  sp.col_lower.resize(n);
  sp.col_upper.resize(n);
  for (int j = 0; j < n; ++j) {
    sp.c[j] = sp.col_scale[j] * model.sense * model.obj[j];
    sp.col_lower[j] = model.col_lower[j] / sp.col_scale[j];
    sp.col_upper[j] = model.col_upper[j] / sp.col_scale[j];
  }
  sp.row_lower.resize(m);
  sp.row_upper.resize(m);
  for (int i = 0; i < m; ++i) {
    sp.row_lower[i] = sp.row_scale[i] * model.row_lower[i];
    sp.row_upper[i] = sp.row_scale[i] * model.row_upper[i];
  }
  if (opt.bound_objective_rescaling) {
    double b2 = 0.0, c2 = 0.0;
    for (int i = 0; i < m; ++i) {
      const double lo = std::isfinite(sp.row_lower[i]) ? std::fabs(sp.row_lower[i]) : 0.0;
      const double up = std::isfinite(sp.row_upper[i]) ? std::fabs(sp.row_upper[i]) : 0.0;
      const double b = std::max(lo, up);
      b2 += b * b;
    }
    for (double v : sp.c) c2 += v * v;
    sp.bound_scale = 1.0 / (std::sqrt(b2) + 1.0);
    sp.obj_scale = 1.0 / (std::sqrt(c2) + 1.0);
    for (double& v : sp.row_lower) v *= sp.bound_scale;
    for (double& v : sp.row_upper) v *= sp.bound_scale;
    for (double& v : sp.col_lower) v *= sp.bound_scale;
    for (double& v : sp.col_upper) v *= sp.bound_scale;
    for (double& v : sp.c) v *= sp.obj_scale;
  }
}

void ScaledProblem::unscale_primal(const std::vector<double>& xs, std::vector<double>& x) const {
  x.resize(n);
  for (int j = 0; j < n; ++j)
    x[j] = std::min(std::max(col_scale[j] * xs[j] / bound_scale, original->col_lower[j]), original->col_upper[j]);
}

void ScaledProblem::unscale_dual(const std::vector<double>& ys, std::vector<double>& y) const {
  y.resize(m);
  for (int i = 0; i < m; ++i) y[i] = row_scale[i] * ys[i] / obj_scale;
}

}  // namespace ps26119::amd
