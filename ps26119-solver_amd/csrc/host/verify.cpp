// verify.cpp — see verify.h.
#include "verify.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>

#include "la.h"

namespace ps26119::amd {

// ------------------------------------------------------------------ double-double
// Error-free transformations (Knuth two-sum, fma two-product) and the accurate dd addition
// (Hida, Li, Bailey, "Library for double-double and quad-double arithmetic").
namespace {

struct dd {
  double hi = 0, lo = 0;
  dd() = default;
  dd(double h) : hi(h) {}  // NOLINT: exact
  dd(double h, double l) : hi(h), lo(l) {}
  double to_double() const { return hi + lo; }
};

inline void two_sum(double a, double b, double& s, double& e) {
  s = a + b;
  const double bb = s - a;
  e = (a - (s - bb)) + (b - bb);
}
inline void quick_two_sum(double a, double b, double& s, double& e) {
  s = a + b;
  e = b - (s - a);
}
inline dd add(const dd& a, const dd& b) {
  double s1, s2, t1, t2;
  two_sum(a.hi, b.hi, s1, s2);
  two_sum(a.lo, b.lo, t1, t2);
  s2 += t1;
  quick_two_sum(s1, s2, s1, s2);
  s2 += t2;
  dd r;
  quick_two_sum(s1, s2, r.hi, r.lo);
  return r;
}
inline dd add(const dd& a, double b) {
  double s1, s2;
  two_sum(a.hi, b, s1, s2);
  s2 += a.lo;
  dd r;
  quick_two_sum(s1, s2, r.hi, r.lo);
  return r;
}
inline dd neg(const dd& a) { return {-a.hi, -a.lo}; }
inline dd sub(const dd& a, const dd& b) { return add(a, neg(b)); }
inline dd exact_product(double a, double b) {
  dd r;
  r.hi = a * b;
  r.lo = std::fma(a, b, -r.hi);
  return r;
}
inline dd mul(const dd& a, const dd& b) {
  double p1 = a.hi * b.hi, p2 = std::fma(a.hi, b.hi, -p1);
  p2 += a.hi * b.lo + a.lo * b.hi;
  dd r;
  quick_two_sum(p1, p2, r.hi, r.lo);
  return r;
}
inline dd div(const dd& a, const dd& b) {  // three partial quotients, remainders in dd
  const double q1 = a.hi / b.hi;
  if (!std::isfinite(q1)) return dd(q1);
  dd r = sub(a, mul(b, dd(q1)));
  const double q2 = r.hi / b.hi;
  r = sub(r, mul(b, dd(q2)));
  const double q3 = r.hi / b.hi;
  dd q;
  quick_two_sum(q1, q2, q.hi, q.lo);
  return add(q, q3);
}
inline bool less(const dd& a, const dd& b) { return a.hi < b.hi || (a.hi == b.hi && a.lo < b.lo); }

// Error model of the certified bound: a dd addition errs by ≤ 2·2⁻¹⁰⁴ of its magnitude, an
// exact product by an absolute 2⁻⁹⁶⁰ when it underflows; both doubled for margin.
constexpr double kRel = 2.0 * 4.930380657631324e-32;
constexpr double kAbs = 2.0 * 1.0e-289;

double lower_double(const dd& v, double err) {
  if (err == 0.0 && v.lo == 0.0) return v.hi;
  return std::nextafter(sub(v, dd(err)).to_double(), -kInf);
}
double upper_double(const dd& v, double err) {
  if (err == 0.0 && v.lo == 0.0) return v.hi;
  return std::nextafter(add(v, dd(err)).to_double(), kInf);
}
bool underflows(const dd& p) { return p.hi != 0.0 && std::fabs(p.hi) < 1e-280; }

// Exact sum of doubles as a nonoverlapping expansion (Shewchuk 1997, GROW-EXPANSION with zero
// elimination); false on non-finite input or when the expansion grows past 64 parts.
bool grow_expansion(std::vector<double>& e, double b, std::vector<double>& tmp) {
  if (b == 0.0) return true;
  if (!std::isfinite(b)) return false;
  tmp.clear();
  double q = b;
  for (double ei : e) {
    double s, err;
    two_sum(q, ei, s, err);
    q = s;
    if (err != 0.0) tmp.push_back(err);
  }
  if (!std::isfinite(q)) return false;
  if (q != 0.0) tmp.push_back(q);
  e.swap(tmp);
  return e.size() <= 64;
}

// min over v ∈ [lo, hi] of z·v as an exact product; `inf` when it is −∞.
dd min_term(double z, double lo, double hi, bool& inf) {
  if (z > 0) {
    if (!std::isfinite(lo)) return inf = true, dd(0.0);
    return exact_product(z, lo);
  }
  if (z < 0) {
    if (!std::isfinite(hi)) return inf = true, dd(0.0);
    return exact_product(z, hi);
  }
  return dd(0.0);
}

struct ColTerm {
  dd t;
  bool inf = false;
};

SafeBound bound_with(const Model& M, const std::vector<double>& y, const std::vector<double>& lo_b,
                     const std::vector<double>& up_b) {
  SafeBound r;
  const int m = M.num_rows, n = M.num_cols;
  const double sense = M.sense;
  if (static_cast<int>(y.size()) != m) return r;
  // Column terms are independent: evaluate them in parallel, accumulate in column order below.
  std::vector<ColTerm> col(n);
  parallel_for(n, [&](std::int64_t b, std::int64_t e) {
    std::vector<double> ex, tmp;
    ex.reserve(16);
    tmp.reserve(17);
    for (std::int64_t j = b; j < e; ++j) {
      dd z = sense * M.obj[j];
      double mag = 0.0;
      int cnt = 0;
      bool uf = false;
      for (int p = M.col_start[j]; p < M.col_start[j + 1]; ++p) {
        const dd prod = exact_product(M.value[p], sense * y[M.row_index[p]]);
        if (prod.hi == 0.0 && prod.lo == 0.0 && M.value[p] * y[M.row_index[p]] == 0.0) continue;
        z = sub(z, prod);
        mag += std::fabs(prod.to_double()) + std::fabs(z.to_double());
        uf = uf || underflows(prod);
        ++cnt;
      }
      const double ez = cnt == 0 ? 0.0 : (cnt + 2) * (kRel * mag * 1.0000001 + (uf ? kAbs : 0.0));
      double zlo = lower_double(z, ez), zhi = upper_double(z, ez);
      if (cnt > 0 && !uf) {
        ex.clear();
        bool ok = grow_expansion(ex, sense * M.obj[j], tmp);
        for (int p = M.col_start[j]; ok && p < M.col_start[j + 1]; ++p) {
          const dd prod = exact_product(M.value[p], sense * y[M.row_index[p]]);
          ok = grow_expansion(ex, -prod.hi, tmp) && grow_expansion(ex, -prod.lo, tmp);
        }
        if (ok) {
          if (ex.empty()) {
            zlo = zhi = 0.0;
          } else if (ex.size() == 1) {
            zlo = zhi = ex[0];
          } else {
            dd v = 0.0;
            for (double c : ex) v = add(v, dd(c));
            const double ev = static_cast<double>(ex.size() + 2) * kRel * std::fabs(v.to_double()) * 1.0000001;
            zlo = std::max(zlo, lower_double(v, ev));
            zhi = std::min(zhi, upper_double(v, ev));
          }
        }
      }
      bool i1 = false, i2 = false;
      const dd tlo = min_term(zlo, lo_b[j], up_b[j], i1), thi = min_term(zhi, lo_b[j], up_b[j], i2);
      col[j].inf = i1 || i2;
      col[j].t = less(tlo, thi) ? tlo : thi;
    }
  });
  dd sum = 0.0;
  double abs_sum = 0.0;
  int count = 0;
  for (int i = 0; i < m; ++i) {
    bool inf = false;
    const dd t = min_term(sense * y[i], M.row_lower[i], M.row_upper[i], inf);
    if (inf) {
      ++r.unbounded_row_terms;
      continue;
    }
    sum = add(sum, t);
    abs_sum += std::fabs(t.to_double());
    ++count;
  }
  for (int j = 0; j < n; ++j) {
    if (col[j].inf) {
      ++r.unbounded_col_terms;
      continue;
    }
    sum = add(sum, col[j].t);
    abs_sum += std::fabs(col[j].t.to_double());
    ++count;
  }
  if (r.unbounded_row_terms > 0 || r.unbounded_col_terms > 0) {
    r.finite = false;
    r.bound = sense > 0 ? -kInf : kInf;
    return r;
  }
  const double err = (count + 2) * (kRel * abs_sum * 1.0000001 + kAbs) + kAbs;
  r.finite = true;
  if (sense > 0) r.bound = lower_double(add(sum, dd(M.obj_offset)), err + kRel * std::fabs(M.obj_offset) * 4);
  else r.bound = upper_double(add(neg(sum), dd(M.obj_offset)), err + kRel * std::fabs(M.obj_offset) * 4);
  return r;
}

// Bounds implied by the rows (feasibility-based tightening), loosened by 1e-9·scale and rounded
// outward; rows are revisited only when one of their columns changed.
// Returns false when `cancel` stopped it (lo and up are then incomplete).
bool implied_bounds(const Model& M, std::vector<double>& lo, std::vector<double>& up,
                    const std::atomic<bool>* cancel = nullptr, int max_passes = 20) {
  lo = M.col_lower;
  up = M.col_upper;
  const Csr A = transpose(csc_as_transpose(M));
  std::vector<char> dirty(M.num_rows, 1);
  auto touch = [&](int j) {
    for (int k = M.col_start[j]; k < M.col_start[j + 1]; ++k) dirty[M.row_index[k]] = 1;
  };
  constexpr double kSlack = 1e-9;
  auto loosen_up = [&](double v, double s) { return std::nextafter(v + kSlack * (s + std::fabs(v)), kInf); };
  auto loosen_down = [&](double v, double s) { return std::nextafter(v - kSlack * (s + std::fabs(v)), -kInf); };
  for (int pass = 0; pass < max_passes; ++pass) {
    bool changed = false;
    for (int i = 0; i < M.num_rows; ++i) {
      if (!dirty[i]) continue;
      if (cancel && (i & 4095) == 0 && cancel->load(std::memory_order_relaxed)) return false;
      dirty[i] = 0;
      const bool has_up = std::isfinite(M.row_upper[i]), has_lo = std::isfinite(M.row_lower[i]);
      if (!has_up && !has_lo) continue;
      dd amin = 0.0, amax = 0.0;
      double mag = 0.0;
      int nmin = 0, nmax = 0, jmin = -1, jmax = -1;
      for (int p = A.ptr[i]; p < A.ptr[i + 1]; ++p) {
        const int j = A.col[p];
        const double a = A.val[p];
        const double bmin = a > 0 ? lo[j] : up[j], bmax = a > 0 ? up[j] : lo[j];
        if (std::isfinite(bmin)) {
          const dd t = exact_product(a, bmin);
          amin = add(amin, t);
          mag += std::fabs(t.to_double());
        } else {
          ++nmin, jmin = j;
        }
        if (std::isfinite(bmax)) {
          const dd t = exact_product(a, bmax);
          amax = add(amax, t);
          mag += std::fabs(t.to_double());
        } else {
          ++nmax, jmax = j;
        }
      }
      for (int p = A.ptr[i]; p < A.ptr[i + 1]; ++p) {
        const int j = A.col[p];
        const double a = A.val[p];
        const double scale = mag / std::fabs(a);
        if (has_up && (nmin == 0 || (nmin == 1 && jmin == j))) {
          dd rest = amin;
          if (nmin == 0) rest = sub(rest, exact_product(a, a > 0 ? lo[j] : up[j]));
          const double v = div(sub(dd(M.row_upper[i]), rest), dd(a)).to_double();
          if (a > 0) {
            const double nb = loosen_up(v, scale);
            if (nb < up[j] - 1e-7 * (1 + std::fabs(nb))) up[j] = nb, changed = true, touch(j);
          } else {
            const double nb = loosen_down(v, scale);
            if (nb > lo[j] + 1e-7 * (1 + std::fabs(nb))) lo[j] = nb, changed = true, touch(j);
          }
        }
        if (has_lo && (nmax == 0 || (nmax == 1 && jmax == j))) {
          dd rest = amax;
          if (nmax == 0) rest = sub(rest, exact_product(a, a > 0 ? up[j] : lo[j]));
          const double v = div(sub(dd(M.row_lower[i]), rest), dd(a)).to_double();
          if (a > 0) {
            const double nb = loosen_down(v, scale);
            if (nb > lo[j] + 1e-7 * (1 + std::fabs(nb))) lo[j] = nb, changed = true, touch(j);
          } else {
            const double nb = loosen_up(v, scale);
            if (nb < up[j] - 1e-7 * (1 + std::fabs(nb))) up[j] = nb, changed = true, touch(j);
          }
        }
      }
    }
    if (!changed) break;
  }
  return true;
}

}  // namespace

std::vector<double> activity(const Model& m, const std::vector<double>& x) {
  // by columns: each row's products come in increasing column order, as in a row-major product
  // over the transpose (same bits) without building it
  return m.row_activity(x);
}

std::string CheckReport::summary() const {
  char buf[256];
  std::snprintf(buf, sizeof buf, "primal viol %.1e  dual viol %.1e  compl %.1e  gap %.1e  => %s", max_primal_violation,
                max_dual_violation, max_complementarity, relative_gap, passed() ? "PASS" : "FAIL");
  return buf;
}

CheckReport check_solution(const Model& lp, const std::vector<double>& x, const std::vector<double>& row_dual) {
  CheckReport rep;
  const int m = lp.num_rows, n = lp.num_cols;
  const double sense = lp.sense;
  auto reject = [&rep]() {
    rep.max_primal_violation = rep.max_dual_violation = rep.relative_gap = kInf;
    rep.primal_ok = rep.dual_ok = rep.gap_ok = false;
    return rep;
  };
  if (static_cast<int>(x.size()) != n || static_cast<int>(row_dual.size()) != m) return reject();
  for (double v : x)
    if (!std::isfinite(v)) return reject();
  for (double v : row_dual)
    if (!std::isfinite(v)) return reject();
  auto bound_violation = [](double v, double lo, double up) {
    double r = 0.0;
    if (lo != -kInf) r = std::max(r, (lo - v) / (1.0 + std::fabs(lo)));
    if (up != kInf) r = std::max(r, (v - up) / (1.0 + std::fabs(up)));
    return r;
  };
  const std::vector<double> ax = activity(lp, x);
  for (int i = 0; i < m; ++i)
    rep.max_primal_violation = std::max(rep.max_primal_violation, bound_violation(ax[i], lp.row_lower[i], lp.row_upper[i]));
  for (int j = 0; j < n; ++j)
    rep.max_primal_violation = std::max(rep.max_primal_violation, bound_violation(x[j], lp.col_lower[j], lp.col_upper[j]));
  std::vector<double> y(m), aty(n, 0.0);
  for (int i = 0; i < m; ++i) y[i] = sense * row_dual[i];
  parallel_for(n, [&](std::int64_t b, std::int64_t e) {
    for (std::int64_t j = b; j < e; ++j) {
      double s = 0.0;
      for (int k = lp.col_start[j]; k < lp.col_start[j + 1]; ++k) s += lp.value[k] * y[lp.row_index[k]];
      aty[j] = s;
    }
  });
  double dual_obj = 0.0, primal_obj = 0.0, c_scale = 1.0;
  for (int j = 0; j < n; ++j) c_scale = std::max(c_scale, std::fabs(lp.obj[j]));
  // multiplier `mu` on a quantity with bounds [lo, up] and value v: wrong-sign part,
  // objective contribution and complementarity (solution_checker.cpp:23-39)
  auto term = [](double mu, double lo, double up, double v, double& viol, double& contrib, double& compl_) {
    viol = contrib = compl_ = 0.0;
    if (mu > 0.0) {
      if (lo == -kInf) viol = mu;
      else contrib = mu * lo, compl_ = mu * std::fabs(v - lo);
    } else if (mu < 0.0) {
      if (up == kInf) viol = -mu;
      else contrib = mu * up, compl_ = -mu * std::fabs(up - v);
    }
  };
  for (int i = 0; i < m; ++i) {
    double v, ct, cp;
    term(y[i], lp.row_lower[i], lp.row_upper[i], ax[i], v, ct, cp);
    rep.max_dual_violation = std::max(rep.max_dual_violation, v / c_scale);
    rep.max_complementarity = std::max(rep.max_complementarity, cp);
    dual_obj += ct;
  }
  for (int j = 0; j < n; ++j) {
    const double c = sense * lp.obj[j];
    double v, ct, cp;
    term(c - aty[j], lp.col_lower[j], lp.col_upper[j], x[j], v, ct, cp);
    rep.max_dual_violation = std::max(rep.max_dual_violation, v / c_scale);
    rep.max_complementarity = std::max(rep.max_complementarity, cp);
    dual_obj += ct;
    primal_obj += c * x[j];
  }
  rep.primal_objective = sense * primal_obj + lp.obj_offset;
  rep.dual_objective = sense * dual_obj + lp.obj_offset;
  rep.relative_gap = std::fabs(rep.primal_objective - rep.dual_objective) /
                     (1.0 + std::fabs(rep.primal_objective) + std::fabs(rep.dual_objective));
  if (!std::isfinite(rep.primal_objective) || !std::isfinite(rep.dual_objective)) return reject();
  rep.primal_ok = rep.max_primal_violation <= tol::kVerifyPrimal;
  rep.dual_ok = rep.max_dual_violation <= tol::kVerifyDual;
  rep.gap_ok = rep.relative_gap <= tol::kVerifyGap;
  return rep;
}

ImpliedBoundsJob::ImpliedBoundsJob(const Model& model) : model_(model) {
  thread_ = std::thread([this] {
    try {
      ok_ = implied_bounds(model_, lo_, up_, &cancel_);
    } catch (...) {
      ok_ = false;
    }
  });
}

ImpliedBoundsJob::~ImpliedBoundsJob() {
  cancel_.store(true, std::memory_order_relaxed);
  if (thread_.joinable()) thread_.join();
}

bool ImpliedBoundsJob::wait(const std::vector<double>** lo, const std::vector<double>** up) {
  if (thread_.joinable()) thread_.join();
  *lo = &lo_;
  *up = &up_;
  return ok_;
}

SafeBound certified_dual_bound(const Model& M, const std::vector<double>& y_in, bool use_implied_bounds,
                               ImpliedBoundsJob* implied) {
  if (static_cast<int>(y_in.size()) != M.num_rows) return SafeBound{};
  // Any multipliers give a valid bound: zero those pointing at an infinite row bound.
  std::vector<double> y = y_in;
  for (int i = 0; i < M.num_rows; ++i) {
    const double ym = M.sense * y[i];
    if ((ym > 0 && !std::isfinite(M.row_lower[i])) || (ym < 0 && !std::isfinite(M.row_upper[i]))) y[i] = 0.0;
  }
  SafeBound r = bound_with(M, y, M.col_lower, M.col_upper);
  if (use_implied_bounds && r.unbounded_col_terms > 0 && r.unbounded_row_terms == 0) {
    const std::vector<double>*lo = nullptr, *up = nullptr;
    if (implied && &implied->model() == &M && implied->wait(&lo, &up)) {
      r = bound_with(M, y, *lo, *up);
    } else {
      std::vector<double> lo_own, up_own;
      implied_bounds(M, lo_own, up_own);
      r = bound_with(M, y, lo_own, up_own);
    }
  }
  return r;
}

CertificateCheck check_infeasibility_certificate(const Model& model, const std::vector<double>& r) {
  CertificateCheck c;
  if (static_cast<int>(r.size()) != model.num_rows || model.num_rows == 0) {
    c.detail = "no dual ray";
    return c;
  }
  c.present = true;
  for (double v : r)
    if (!std::isfinite(v)) {
      c.detail = "dual ray is not finite";
      return c;
    }
  // Stage 1, rigorous: the certified bound of the zero-objective model is L₀(r).
  Model zero = model;
  zero.sense = 1;
  zero.obj_offset = 0.0;
  std::fill(zero.obj.begin(), zero.obj.end(), 0.0);
  zero.is_integer.clear();
  const SafeBound sb = certified_dual_bound(zero, r);
  char buf[200];
  if (sb.finite && sb.bound > 0.0) {
    c.passed = c.rigorous = true;
    c.measure = sb.bound;
    std::snprintf(buf, sizeof buf, "Farkas certificate: certified L0(r) = %.3g > 0 (rounding-proof)", sb.bound);
    c.detail = buf;
    return c;
  }
  // Stage 2: tolerance test, normalised like the engines' ray test.
  const int m = model.num_rows, n = model.num_cols;
  double rmax = 0.0, obj = 0.0, viol = 0.0, mag = 0.0;
  for (int i = 0; i < m; ++i) {
    rmax = std::max(rmax, std::fabs(r[i]));
    if (r[i] > 0) {
      if (std::isfinite(model.row_lower[i])) obj += r[i] * model.row_lower[i], mag += std::fabs(r[i] * model.row_lower[i]);
      else viol = std::max(viol, r[i]);
    } else if (r[i] < 0) {
      if (std::isfinite(model.row_upper[i])) obj += r[i] * model.row_upper[i], mag += std::fabs(r[i] * model.row_upper[i]);
      else viol = std::max(viol, -r[i]);
    }
  }
  std::vector<double> atr(n, 0.0), atr_abs(n, 0.0);
  for (int j = 0; j < n; ++j)
    for (int k = model.col_start[j]; k < model.col_start[j + 1]; ++k) {
      atr[j] += model.value[k] * r[model.row_index[k]];
      atr_abs[j] += std::fabs(model.value[k] * r[model.row_index[k]]);
    }
  for (int j = 0; j < n; ++j) {
    const double lam = -atr[j];
    if (lam > 0) {
      if (std::isfinite(model.col_lower[j])) obj += lam * model.col_lower[j], mag += atr_abs[j] * std::fabs(model.col_lower[j]);
      else viol = std::max(viol, lam);
    } else if (lam < 0) {
      if (std::isfinite(model.col_upper[j])) obj += lam * model.col_upper[j], mag += atr_abs[j] * std::fabs(model.col_upper[j]);
      else viol = std::max(viol, -lam);
    }
  }
  const double scale = std::max(rmax, viol);
  c.measure = scale > 0 ? obj / scale : 0.0;
  c.violation = scale > 0 ? viol / scale : 0.0;
  c.passed = std::isfinite(c.measure) && c.measure > 0 && obj > tol::kVerifyRay * mag && c.violation <= tol::kVerifyRay * c.measure;
  std::snprintf(buf, sizeof buf,
                "Farkas certificate: rigorous L0(r) = %.3g; tolerance test L0/s = %.3g, violation/s = %.3g, "
                "L0/sum|terms| = %.3g (%s)",
                sb.bound, c.measure, c.violation, mag > 0 ? obj / mag : 0.0,
                c.passed ? "infeasibility shown within tolerance, not rounding-proof" : "FAIL");
  c.detail = buf;
  return c;
}

CertificateCheck check_unboundedness_certificate(const Model& model, const std::vector<double>& x,
                                                 const std::vector<double>& ray) {
  CertificateCheck c;
  const int m = model.num_rows, n = model.num_cols;
  if (static_cast<int>(ray.size()) != n || static_cast<int>(x.size()) != n || n == 0) {
    c.detail = "no primal ray / point";
    return c;
  }
  c.present = true;
  double dmax = 0.0, cmax = 0.0;
  for (int j = 0; j < n; ++j) {
    if (!std::isfinite(ray[j]) || !std::isfinite(x[j])) {
      c.detail = "ray or point is not finite";
      return c;
    }
    dmax = std::max(dmax, std::fabs(ray[j]));
    cmax = std::max(cmax, std::fabs(model.obj[j]));
  }
  if (dmax == 0.0) {
    c.detail = "zero ray";
    return c;
  }
  const CheckReport rep = check_solution(model, x, std::vector<double>(m, 0.0));
  std::vector<double> d(n);
  double cd = 0.0, viol = 0.0;
  for (int j = 0; j < n; ++j) {
    d[j] = ray[j] / dmax;
    cd += model.sense * model.obj[j] * d[j];
    if (std::isfinite(model.col_lower[j])) viol = std::max(viol, -d[j]);
    if (std::isfinite(model.col_upper[j])) viol = std::max(viol, d[j]);
  }
  const std::vector<double> ad = activity(model, d);
  for (int i = 0; i < m; ++i) {
    if (std::isfinite(model.row_lower[i])) viol = std::max(viol, -ad[i]);
    if (std::isfinite(model.row_upper[i])) viol = std::max(viol, ad[i]);
  }
  c.measure = -cd;
  c.violation = viol;
  c.passed = rep.primal_ok && -cd >= tol::kVerifyRay * std::max(1.0, cmax) && viol <= tol::kVerifyRay * (-cd);
  char buf[200];
  std::snprintf(buf, sizeof buf, "unbounded ray: -c'd = %.3g, cone violation %.3g, point primal violation %.3g (%s)", -cd,
                viol, rep.max_primal_violation, c.passed ? "unboundedness proven" : "FAIL");
  c.detail = buf;
  return c;
}

}  // namespace ps26119::amd
