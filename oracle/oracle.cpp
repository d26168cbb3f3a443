// oracle.cpp — CPU restatement of the reference r²HPDHG path. TEST INFRASTRUCTURE ONLY:
// it is the checker for the MI355X backend (tests/, smoke(), bench's CPU-baseline leg);
// nothing in the product links or calls it.
//
// It follows the reference CPU path operation for operation, so that on the same input it
// produces the same bits (single-threaded; the reference's results do not depend on its
// thread count):
//   CSR build / transpose ........ src/la/csr.cpp:9-36
//   scaling ...................... src/pdhg/scaling.cpp:37-155
//   ‖Ã‖₂ Lanczos ................. src/la/csr.cpp:44-132
//   64-chunk reductions .......... src/la/parallel.h:150-188
//   backend kernels (fp64 / fp32). src/pdhg/cpu_backend.cpp:41-187, csr.h:30-43
//   KKT statistics ............... src/pdhg/cpu_backend.cpp:226-284, termination.h:35-60
//   r²HPDHG driver ............... src/pdhg/r2hpdhg.cpp:11-162
//   engine context, precision policy, ray test, warm start, finish
//                                  src/pdhg/engine.cpp:56-307
//   presolve / postsolve ......... src/core/presolve.cpp:22-209
//   solve_direct / solve_impl .... src/core/solve.cpp:266-609 (LP, r²HPDHG only)
//   gate, check, certified bound . src/core/solve.cpp:119-185, solution_checker.cpp:61-133,
//                                  safe_bound.cpp:16-195, implied_bounds.cpp:11-111,
//                                  certificates.cpp:13-135
// Pinning: tests/test_oracle.py compares every function here bit for bit (or to the stated
// tolerance) with oracle/_ref (the reference sources built by oracle/ref.mk), with the golden
// fixtures in tests/golden/, and with closed forms (hand LPs, planted optima).
//
// Build: g++ -O2 -ffp-contract=off (x86-64 baseline has no FMA, matching the reference build).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <limits>
#include <memory>
#include <string>
#include <vector>

#include "orc_abi.h"

namespace orc {

constexpr double kInf = std::numeric_limits<double>::infinity();
constexpr double kNaN = std::numeric_limits<double>::quiet_NaN();

enum Status { Optimal, Infeasible, Unbounded, IterationLimit, TimeLimit, NumericalError, NotSolved };

struct Lp {  // the model contract (reference include/ps26119/model.h:22-68)
  int m = 0, n = 0, sense = 1;
  double offset = 0;
  std::vector<double> c, cl, cu, rl, ru;
  std::vector<int> cs, ri;
  std::vector<double> v;
  std::vector<double> activity(const std::vector<double>& x) const {  // model.cpp:175-180
    std::vector<double> ax(m, 0.0);
    for (int j = 0; j < n; ++j)
      for (int k = cs[j]; k < cs[j + 1]; ++k) ax[ri[k]] += v[k] * x[j];
    return ax;
  }
  double objective(const std::vector<double>& x) const {  // model.cpp:169-173
    double s = offset;
    for (int j = 0; j < n; ++j) s += c[j] * x[j];
    return s;
  }
};

Lp from_abi(const orc_lp* p) {
  Lp l;
  l.m = p->m;
  l.n = p->n;
  l.sense = p->sense;
  l.offset = p->offset;
  l.c.assign(p->c, p->c + p->n);
  l.cl.assign(p->col_lower, p->col_lower + p->n);
  l.cu.assign(p->col_upper, p->col_upper + p->n);
  l.rl.assign(p->row_lower, p->row_lower + p->m);
  l.ru.assign(p->row_upper, p->row_upper + p->m);
  l.cs.assign(p->col_start, p->col_start + p->n + 1);
  const int nnz = p->col_start[p->n];
  l.ri.assign(p->row_index, p->row_index + nnz);
  l.v.assign(p->value, p->value + nnz);
  return l;
}

// ------------------------------------------------------------------ sparse matrices
template <class T>
struct Csr {
  int rows = 0, cols = 0;
  std::vector<std::int64_t> ptr;
  std::vector<int> idx;
  std::vector<T> val;
  // Row-sequential product with a double accumulator (csr.h:30-43).
  template <class X, class Out>
  void mul(const X* x, Out* out) const {
    for (int i = 0; i < rows; ++i) {
      double s = 0;
      for (std::int64_t k = ptr[i]; k < ptr[i + 1]; ++k) s += static_cast<double>(val[k]) * static_cast<double>(x[idx[k]]);
      out[i] = static_cast<Out>(s);
    }
  }
  template <class U>
  Csr<U> as() const {
    Csr<U> r;
    r.rows = rows;
    r.cols = cols;
    r.ptr = ptr;
    r.idx = idx;
    r.val.assign(val.begin(), val.end());
    return r;
  }
};

// Transpose by a counting sort that visits source rows in order, so the output rows list
// their entries by increasing column (csr.cpp:19-36).
Csr<double> transpose(const Csr<double>& a) {
  Csr<double> t;
  t.rows = a.cols;
  t.cols = a.rows;
  t.ptr.assign(static_cast<std::size_t>(t.rows) + 1, 0);
  const std::int64_t nnz = a.ptr.empty() ? 0 : a.ptr.back();
  for (std::int64_t k = 0; k < nnz; ++k) ++t.ptr[a.idx[k] + 1];
  for (int r = 0; r < t.rows; ++r) t.ptr[r + 1] += t.ptr[r];
  t.idx.resize(nnz);
  t.val.resize(nnz);
  std::vector<std::int64_t> fill(t.ptr.begin(), t.ptr.end() - 1);
  for (int r = 0; r < a.rows; ++r)
    for (std::int64_t k = a.ptr[r]; k < a.ptr[r + 1]; ++k) {
      const std::int64_t q = fill[a.idx[k]]++;
      t.idx[q] = r;
      t.val[q] = a.val[k];
    }
  return t;
}

Csr<double> csc_as_transpose(const Lp& l) {  // csr.cpp:9-17: the CSC arrays ARE Aᵀ in CSR
  Csr<double> t;
  t.rows = l.n;
  t.cols = l.m;
  t.ptr.assign(l.cs.begin(), l.cs.end());
  t.idx = l.ri;
  t.val = l.v;
  return t;
}

// ------------------------------------------------------------------ reductions
// Fixed 64-chunk sum (parallel.h:150-188): chunk c covers [n·c/C, n·(c+1)/C), C = min(64, max(1,n)),
// summed sequentially, partials combined in chunk order.
template <class G>
double chunk_sum(std::int64_t n, G&& g) {
  const int chunks = static_cast<int>(std::min<std::int64_t>(64, std::max<std::int64_t>(1, n)));
  double total = 0;
  for (int c = 0; c < chunks; ++c) {
    const std::int64_t b = n * c / chunks, e = n * (c + 1) / chunks;
    double s = 0;
    for (std::int64_t i = b; i < e; ++i) s += g(i);
    total += s;
  }
  return total;
}

// ------------------------------------------------------------------ scaling
struct ScaleOpt {  // scaling.h:32-46
  int geo_iters = 12;
  double geo_min_range = 4.5;
  int ruiz = 10;
  bool pc = true;
  bool bor = true;
};

struct Scaled {  // scaling.h:48-67
  int m = 0, n = 0;
  Csr<double> A, At, A0, At0;
  std::vector<double> c, l, u, rl, ru, rs, cs;
  double bs = 1, os = 1, log10_range = 0;
  bool geo = false;
  const Lp* orig = nullptr;
  void unscale_x(const std::vector<double>& xs, std::vector<double>& x) const {  // scaling.cpp:144-150
    x.resize(n);
    for (int j = 0; j < n; ++j) x[j] = std::min(std::max(cs[j] * xs[j] / bs, orig->cl[j]), orig->cu[j]);
  }
  void unscale_y(const std::vector<double>& ys, std::vector<double>& y) const {  // scaling.cpp:152-155
    y.resize(m);
    for (int i = 0; i < m; ++i) y[i] = rs[i] * ys[i] / os;
  }
};

static void scale_entries(Csr<double>& A, const std::vector<double>& r, const std::vector<double>& c) {
  for (int i = 0; i < A.rows; ++i)
    for (std::int64_t k = A.ptr[i]; k < A.ptr[i + 1]; ++k) A.val[k] *= r[i] * c[A.idx[k]];
}

static double inv_sqrt_or_one(double v) { return v > 0.0 && std::isfinite(v) ? 1.0 / std::sqrt(v) : 1.0; }

Scaled scale(const Lp& l, const ScaleOpt& o) {  // scaling.cpp:37-142
  Scaled s;
  s.m = l.m;
  s.n = l.n;
  s.orig = &l;
  s.At0 = csc_as_transpose(l);
  s.A0 = transpose(s.At0);
  s.A = s.A0;
  s.rs.assign(l.m, 1.0);
  s.cs.assign(l.n, 1.0);
  double amin = kInf, amax = 0;
  for (double v : s.A0.val)
    if (v != 0.0) amin = std::min(amin, std::fabs(v)), amax = std::max(amax, std::fabs(v));
  s.log10_range = amax > 0 ? std::log10(amax / amin) : 0.0;
  s.geo = o.geo_iters > 0 && s.log10_range >= o.geo_min_range;
  if (s.geo) {
    std::vector<double> r(l.m, 1.0), c(l.n, 1.0), lo, hi;
    auto clampf = [](double v) { return std::min(std::max(v, 1e-20), 1e20); };
    for (int it = 0; it < o.geo_iters; ++it) {
      lo.assign(l.m, kInf);
      hi.assign(l.m, 0.0);
      for (int i = 0; i < l.m; ++i)
        for (std::int64_t k = s.A0.ptr[i]; k < s.A0.ptr[i + 1]; ++k) {
          const double v = std::fabs(s.A0.val[k]) * c[s.A0.idx[k]];
          if (v > 0 && std::isfinite(v)) lo[i] = std::min(lo[i], v), hi[i] = std::max(hi[i], v);
        }
      for (int i = 0; i < l.m; ++i)
        if (hi[i] > 0) r[i] = clampf(1.0 / (std::sqrt(lo[i]) * std::sqrt(hi[i])));
      lo.assign(l.n, kInf);
      hi.assign(l.n, 0.0);
      for (int i = 0; i < l.m; ++i)
        for (std::int64_t k = s.A0.ptr[i]; k < s.A0.ptr[i + 1]; ++k) {
          const int j = s.A0.idx[k];
          const double v = std::fabs(s.A0.val[k]) * r[i];
          if (v > 0 && std::isfinite(v)) lo[j] = std::min(lo[j], v), hi[j] = std::max(hi[j], v);
        }
      for (int j = 0; j < l.n; ++j)
        if (hi[j] > 0) c[j] = clampf(1.0 / (std::sqrt(lo[j]) * std::sqrt(hi[j])));
    }
    scale_entries(s.A, r, c);
    s.rs = r;
    s.cs = c;
  }
  std::vector<double> rn(l.m), cn(l.n), rf(l.m), cf(l.n);
  auto sweep = [&](bool inf_norm) {
    std::fill(rn.begin(), rn.end(), 0.0);
    std::fill(cn.begin(), cn.end(), 0.0);
    for (int i = 0; i < l.m; ++i)
      for (std::int64_t k = s.A.ptr[i]; k < s.A.ptr[i + 1]; ++k) {
        const double a = std::fabs(s.A.val[k]);
        const int j = s.A.idx[k];
        if (inf_norm) rn[i] = std::max(rn[i], a), cn[j] = std::max(cn[j], a);
        else rn[i] += a, cn[j] += a;
      }
    for (int i = 0; i < l.m; ++i) rf[i] = inv_sqrt_or_one(rn[i]);
    for (int j = 0; j < l.n; ++j) cf[j] = inv_sqrt_or_one(cn[j]);
    scale_entries(s.A, rf, cf);
    for (int i = 0; i < l.m; ++i) s.rs[i] *= rf[i];
    for (int j = 0; j < l.n; ++j) s.cs[j] *= cf[j];
  };
  for (int it = 0; it < o.ruiz; ++it) sweep(true);
  if (o.pc) sweep(false);
  s.c.resize(l.n);
  s.l.resize(l.n);
  s.u.resize(l.n);
  for (int j = 0; j < l.n; ++j) {
    s.c[j] = s.cs[j] * l.sense * l.c[j];
    s.l[j] = l.cl[j] / s.cs[j];
    s.u[j] = l.cu[j] / s.cs[j];
  }
  s.rl.resize(l.m);
  s.ru.resize(l.m);
  for (int i = 0; i < l.m; ++i) {
    s.rl[i] = s.rs[i] * l.rl[i];
    s.ru[i] = s.rs[i] * l.ru[i];
  }
  if (o.bor) {
    double b2 = 0, c2 = 0;
    for (int i = 0; i < l.m; ++i) {
      const double a = std::isfinite(s.rl[i]) ? std::fabs(s.rl[i]) : 0.0, b = std::isfinite(s.ru[i]) ? std::fabs(s.ru[i]) : 0.0;
      const double mx = std::max(a, b);
      b2 += mx * mx;
    }
    for (double v : s.c) c2 += v * v;
    s.bs = 1.0 / (std::sqrt(b2) + 1.0);
    s.os = 1.0 / (std::sqrt(c2) + 1.0);
    for (double& v : s.rl) v *= s.bs;
    for (double& v : s.ru) v *= s.bs;
    for (double& v : s.l) v *= s.bs;
    for (double& v : s.u) v *= s.bs;
    for (double& v : s.c) v *= s.os;
  }
  s.At = transpose(s.A);
  return s;
}

// ------------------------------------------------------------------ ‖A‖₂ (Lanczos on AᵀA)
static double top_eig_tridiag(const std::vector<double>& a, const std::vector<double>& b, int k) {  // csr.cpp:44-69
  double lo = 1e300, hi = -1e300;
  for (int i = 0; i < k; ++i) {
    const double r = (i > 0 ? std::fabs(b[i - 1]) : 0.0) + (i + 1 < k ? std::fabs(b[i]) : 0.0);
    lo = std::min(lo, a[i] - r);
    hi = std::max(hi, a[i] + r);
  }
  auto below = [&](double x) {
    int cnt = 0;
    double d = 1.0;
    for (int i = 0; i < k; ++i) {
      const double bb = i > 0 ? b[i - 1] * b[i - 1] : 0.0;
      d = (a[i] - x) - (i > 0 ? bb / d : 0.0);
      if (d == 0.0) d = -1e-300;
      if (d < 0) ++cnt;
    }
    return cnt;
  };
  for (int it = 0; it < 200 && hi - lo > 1e-15 * std::max(1.0, std::fabs(hi)); ++it) {
    const double mid = 0.5 * (lo + hi);
    if (below(mid) == k) hi = mid;
    else lo = mid;
  }
  return hi;
}

double norm2(const Csr<double>& a, const Csr<double>& at, int max_iter = 400, double rel_tol = 1e-12) {  // csr.cpp:86-132
  const int n = a.cols;
  if (n == 0 || a.ptr.back() == 0) return 0.0;
  const int K = std::min(std::max(max_iter, 20), 400);
  std::vector<double> v(n), vp(n, 0.0), w(n), t(a.rows);
  std::uint64_t st = 0x9E3779B97F4A7C15ULL;
  double nv = 0;
  for (int j = 0; j < n; ++j) {
    st ^= st >> 12, st ^= st << 25, st ^= st >> 27;
    v[j] = static_cast<double>((st * 0x2545F4914F6CDD1DULL) >> 11) / 9007199254740992.0 * 2.0 - 1.0;
    nv += v[j] * v[j];
  }
  nv = std::sqrt(nv);
  for (double& e : v) e /= nv;
  std::vector<double> al_, be_;
  double theta = 0, theta_old = -1;
  int flat = 0;
  for (int k = 0; k < K; ++k) {
    a.mul(v.data(), t.data());
    at.mul(t.data(), w.data());
    double al = 0;
    for (int j = 0; j < n; ++j) al += w[j] * v[j];
    const double bprev = be_.empty() ? 0.0 : be_.back();
    double bn = 0;
    for (int j = 0; j < n; ++j) {
      w[j] -= al * v[j] + bprev * vp[j];
      bn += w[j] * w[j];
    }
    bn = std::sqrt(bn);
    al_.push_back(al);
    theta = top_eig_tridiag(al_, be_, static_cast<int>(al_.size()));
    if (bn <= 1e-14 * std::max(1.0, theta)) break;
    be_.push_back(bn);
    if (theta - theta_old <= rel_tol * theta) {
      if (++flat >= 5 && k >= 20) break;
    } else {
      flat = 0;
    }
    theta_old = theta;
    for (int j = 0; j < n; ++j) {
      vp[j] = v[j];
      v[j] = w[j] / bn;
    }
  }
  return std::sqrt(std::max(theta, 0.0)) * (1.0 + 1e-4);
}

// ------------------------------------------------------------------ KKT statistics
struct Kkt {  // termination.h:35-60
  double rp = 0, rd = 0, pobj = 0, dobj = 0, bn = 0, cn = 0, pmax = 0, dmax = 0, shift = 0;
  double rel_p() const { return rp / (1.0 + bn); }
  double rel_d() const { return rd / (1.0 + cn); }
  double rel_gap() const { return std::fabs(pobj - dobj) / (1.0 + std::fabs(pobj + shift) + std::fabs(dobj + shift)); }
  double rel() const { return std::max({rel_p(), rel_d(), rel_gap()}); }
  bool converged(double eps) const {
    if (rel() > eps) return false;
    if (eps >= 1e-6) return true;
    return pmax <= 1e-6 && dmax <= 1e-6 && rel_gap() <= 1e-6;
  }
  bool finite() const { return std::isfinite(rp) && std::isfinite(rd) && std::isfinite(pobj) && std::isfinite(dobj); }
};

struct View {
  int m, n, sense;
  const double *c, *cl, *cu, *rl, *ru;
};

Kkt kkt_of(const Csr<double>& A, const Csr<double>& At, const View& M, const std::vector<double>& x,
           const std::vector<double>& y) {  // cpu_backend.cpp:226-284
  Kkt k;
  std::vector<double> ax(M.m), aty(M.n);
  A.mul(x.data(), ax.data());
  At.mul(y.data(), aty.data());
  double rp2 = 0, rd2 = 0, b2 = 0, c2 = 0, p = 0, d = 0, cinf = 0, pmax = 0, dmax = 0;
  auto ev = [](double val, double lo, double up) {
    if (val < lo) return (lo - val) / (1.0 + std::fabs(lo));
    if (val > up) return (val - up) / (1.0 + std::fabs(up));
    return 0.0;
  };
  for (int i = 0; i < M.m; ++i) {
    const double lo = M.rl[i], up = M.ru[i];
    const double viol = ax[i] < lo ? lo - ax[i] : (ax[i] > up ? ax[i] - up : 0.0);
    rp2 += viol * viol;
    pmax = std::max(pmax, ev(ax[i], lo, up));
    const double bl = std::isfinite(lo) ? std::fabs(lo) : 0.0, bu = std::isfinite(up) ? std::fabs(up) : 0.0;
    b2 += std::max(bl, bu) * std::max(bl, bu);
    if (y[i] > 0) {
      if (std::isfinite(lo)) d += y[i] * lo;
      else rd2 += y[i] * y[i], dmax = std::max(dmax, y[i]);
    } else if (y[i] < 0) {
      if (std::isfinite(up)) d += y[i] * up;
      else rd2 += y[i] * y[i], dmax = std::max(dmax, -y[i]);
    }
  }
  for (int j = 0; j < M.n; ++j) {
    const double cj = M.sense * M.c[j];
    c2 += cj * cj;
    cinf = std::max(cinf, std::fabs(cj));
    p += cj * x[j];
    pmax = std::max(pmax, ev(x[j], M.cl[j], M.cu[j]));
    const double lam = cj - aty[j];
    const double lo = M.cl[j], up = M.cu[j];
    double z;
    if (std::isfinite(lo) && std::isfinite(up)) z = lam;
    else if (std::isfinite(lo)) z = std::max(lam, 0.0);
    else if (std::isfinite(up)) z = std::min(lam, 0.0);
    else z = 0.0;
    rd2 += (lam - z) * (lam - z);
    dmax = std::max(dmax, std::fabs(lam - z));
    if (z > 0) d += z * lo;
    else if (z < 0) d += z * up;
  }
  k.rp = std::sqrt(rp2);
  k.rd = std::sqrt(rd2);
  k.pobj = p;
  k.dobj = d;
  k.bn = std::sqrt(b2);
  k.cn = std::sqrt(c2);
  k.pmax = pmax;
  k.dmax = dmax / (1.0 + cinf);
  return k;
}

Kkt kkt_original(const Scaled& s, const std::vector<double>& x, const std::vector<double>& y) {
  const Lp& l = *s.orig;
  const View v{l.m, l.n, l.sense, l.c.data(), l.cl.data(), l.cu.data(), l.rl.data(), l.ru.data()};
  return kkt_of(s.A0, s.At0, v, x, y);
}

// ------------------------------------------------------------------ backend (fp64 / fp32)
template <class T>
static T clampv(T v, T lo, T hi) { return v < lo ? lo : (v > hi ? hi : v); }

template <class T>
struct Data {
  Csr<T> A, At;
  std::vector<T> c, l, u, rl, ru;
  std::vector<std::vector<T>> vec;
};

class Backend {  // cpu_backend.cpp:31-187, one working precision at a time
 public:
  explicit Backend(const Scaled& s) : s_(s) {
    load(d_);
    load(f_);
  }
  bool mixed() const { return mixed_; }
  void set_mixed(bool mx) {  // cpu_backend.cpp:50-59
    if (mx == mixed_) return;
    if (mx)
      for (std::size_t k = 0; k < d_.vec.size(); ++k) f_.vec[k].assign(d_.vec[k].begin(), d_.vec[k].end());
    else
      for (std::size_t k = 0; k < f_.vec.size(); ++k) d_.vec[k].assign(f_.vec[k].begin(), f_.vec[k].end());
    mixed_ = mx;
  }
  int create(bool primal) {
    const std::size_t len = primal ? s_.n : s_.m;
    d_.vec.emplace_back(mixed_ ? 0 : len, 0.0);
    f_.vec.emplace_back(mixed_ ? len : 0, 0.0f);
    len_.push_back(len);
    return static_cast<int>(len_.size()) - 1;
  }
  void upload(int h, const std::vector<double>& src) {
    on([&](auto& D) { D.vec[h].assign(src.begin(), src.end()); });
  }
  void download(int h, std::vector<double>& dst) {
    on([&](auto& D) { dst.assign(D.vec[h].begin(), D.vec[h].end()); });
  }
  void fill(int h, double a) {
    on([&](auto& D) {
      using T = typename std::decay_t<decltype(D.c)>::value_type;
      D.vec[h].assign(len_[h], static_cast<T>(a));
    });
  }
  void copy(int dst, int src) {
    on([&](auto& D) { D.vec[dst] = D.vec[src]; });
  }
  void axpby(double a, int x, double b, int y) {
    on([&](auto& D) {
      using T = typename std::decay_t<decltype(D.c)>::value_type;
      const T ta = static_cast<T>(a), tb = static_cast<T>(b);
      auto& Y = D.vec[y];
      const auto& X = D.vec[x];
      for (std::size_t i = 0; i < Y.size(); ++i) Y[i] = ta * X[i] + tb * Y[i];
    });
  }
  void spmv(int x, int out) {
    on([&](auto& D) { D.A.mul(D.vec[x].data(), D.vec[out].data()); });
  }
  void spmv_t(int y, int out) {
    on([&](auto& D) { D.At.mul(D.vec[y].data(), D.vec[out].data()); });
  }
  double dot(int a, int b) {
    double r = 0;
    on([&](auto& D) {
      const auto &A = D.vec[a], &B = D.vec[b];
      r = chunk_sum(static_cast<std::int64_t>(A.size()),
                    [&](std::int64_t i) { return static_cast<double>(A[i]) * static_cast<double>(B[i]); });
    });
    return r;
  }
  double diff_norm2(int a, int b) {
    double r = 0;
    on([&](auto& D) {
      const auto &A = D.vec[a], &B = D.vec[b];
      r = chunk_sum(static_cast<std::int64_t>(A.size()), [&](std::int64_t i) {
        const double t = static_cast<double>(A[i]) - static_cast<double>(B[i]);
        return t * t;
      });
    });
    return std::sqrt(r);
  }
  void primal_step(int x, int aty, double tau, int xh) {
    on([&](auto& D) {
      using T = typename std::decay_t<decltype(D.c)>::value_type;
      const T t = static_cast<T>(tau);
      const auto &X = D.vec[x], &G = D.vec[aty];
      auto& H = D.vec[xh];
      for (std::size_t j = 0; j < X.size(); ++j) H[j] = clampv<T>(X[j] - t * (D.c[j] - G[j]), D.l[j], D.u[j]);
    });
  }
  void dual_step(int y, int ax, double sigma, int yh) {
    on([&](auto& D) {
      using T = typename std::decay_t<decltype(D.c)>::value_type;
      const T sg = static_cast<T>(sigma);
      const auto &Y = D.vec[y], &AX = D.vec[ax];
      auto& H = D.vec[yh];
      for (std::size_t i = 0; i < Y.size(); ++i) {
        const T q = AX[i] - Y[i] / sg;
        H[i] = Y[i] - sg * AX[i] + sg * clampv<T>(q, D.rl[i], D.ru[i]);
      }
    });
  }
  void r2h_primal(int x, int x0, int aty, double tau, double w, double rho, int xh, int xb) {
    on([&](auto& D) {
      using T = typename std::decay_t<decltype(D.c)>::value_type;
      const T t = static_cast<T>(tau), tw = static_cast<T>(w), a = static_cast<T>(2 * rho), b = static_cast<T>(1 - 2 * rho);
      auto& X = D.vec[x];
      const auto &X0 = D.vec[x0], &G = D.vec[aty];
      auto &H = D.vec[xh], &B = D.vec[xb];
      for (std::size_t j = 0; j < X.size(); ++j) {
        const T h = clampv<T>(X[j] - t * (D.c[j] - G[j]), D.l[j], D.u[j]);
        H[j] = h;
        B[j] = 2 * h - X[j];
        X[j] = tw * (a * h + b * X[j]) + (1 - tw) * X0[j];
      }
    });
  }
  void r2h_dual(int y, int y0, int ax, double sigma, double w, double rho, int yh, int yb) {
    on([&](auto& D) {
      using T = typename std::decay_t<decltype(D.c)>::value_type;
      const T sg = static_cast<T>(sigma), tw = static_cast<T>(w), a = static_cast<T>(2 * rho), b = static_cast<T>(1 - 2 * rho);
      auto& Y = D.vec[y];
      const auto &Y0 = D.vec[y0], &AX = D.vec[ax];
      auto& H = D.vec[yh];
      T* B = yb >= 0 ? D.vec[yb].data() : nullptr;
      for (std::size_t i = 0; i < Y.size(); ++i) {
        const T q = AX[i] - Y[i] / sg;
        const T h = Y[i] - sg * AX[i] + sg * clampv<T>(q, D.rl[i], D.ru[i]);
        H[i] = h;
        if (B) B[i] = 2 * h - Y[i];
        Y[i] = tw * (a * h + b * Y[i]) + (1 - tw) * Y0[i];
      }
    });
  }
  Kkt kkt(int xs, int ys) {
    std::vector<double> a, b, x, y;
    download(xs, a);
    download(ys, b);
    s_.unscale_x(a, x);
    s_.unscale_y(b, y);
    return kkt_original(s_, x, y);
  }

 private:
  template <class T>
  void load(Data<T>& D) {
    D.A = s_.A.as<T>();
    D.At = s_.At.as<T>();
    D.c.assign(s_.c.begin(), s_.c.end());
    D.l.assign(s_.l.begin(), s_.l.end());
    D.u.assign(s_.u.begin(), s_.u.end());
    D.rl.assign(s_.rl.begin(), s_.rl.end());
    D.ru.assign(s_.ru.begin(), s_.ru.end());
  }
  template <class F>
  void on(F&& f) {
    if (mixed_) f(f_);
    else f(d_);
  }
  const Scaled& s_;
  bool mixed_ = false;
  Data<double> d_;
  Data<float> f_;
  std::vector<std::size_t> len_;
};

// ------------------------------------------------------------------ engine options and context
struct Opt {  // engine.h:27-48 (r²HPDHG part)
  double tol = 1e-8, time_limit = 3600.0;
  std::int64_t iter_limit = 100000000;
  int K = 64;
  bool mixed = false;
  std::vector<double> warm_x, warm_y;
  double warm_w = 0;
  double b_norm = -1, c_norm = -1, obj_shift = 0;
  ScaleOpt sc;
  double rho = 1.0, r_suff = 0.2, r_nec = 0.5, r_art = 0.36;
  double kp = 0.99, ki = 0.01, kd = 0.0, idecay = 0.3, max_log_step = 2.302585092994046;
};

// Applies named knobs (engine.cpp:34-53); returns the first unknown name.
std::string apply_params(Opt& e, const std::vector<std::pair<std::string, double>>& ps) {
  for (const auto& [name, v] : ps) {
    if (name == "reflection") e.rho = v;
    else if (name == "restart_sufficient") e.r_suff = v;
    else if (name == "restart_necessary") e.r_nec = v;
    else if (name == "restart_artificial") e.r_art = v;
    else if (name == "pid_kp") e.kp = v;
    else if (name == "pid_ki") e.ki = v;
    else if (name == "pid_kd") e.kd = v;
    else if (name == "pid_integral_decay") e.idecay = v;
    else if (name == "pid_max_log_step") e.max_log_step = v;
    else if (name == "bound_objective_rescaling") e.sc.bor = v != 0;
    else if (name == "geometric_mean_iterations") e.sc.geo_iters = static_cast<int>(v);
    else if (name == "geometric_mean_min_log10_range") e.sc.geo_min_range = v;
    else if (name == "ruiz_iterations") e.sc.ruiz = static_cast<int>(v);
    else if (name == "pock_chambolle") e.sc.pc = v != 0;
    else return name;
  }
  return {};
}

struct Sol {
  Status status = NotSolved;
  std::vector<double> x, y, z, row_activity, dual_ray, primal_ray;
  double objective = kNaN, dual_objective = kNaN, rp = kNaN, rd = kNaN, gap = kNaN;
  std::int64_t iterations = 0, it_fast = -1;
  double seconds = 0, setup_seconds = 0, primal_weight = kNaN, certified = kNaN;
  int check = -1;
  double check_p = kNaN, check_d = kNaN, check_g = kNaN;
  std::string engine = "r2hpdhg", message;
};

struct Precision {  // engine.cpp:56-79, 186-192
  Backend& b;
  bool mixed;
  double target, best = 1e300;
  int stale = 0;
  Precision(Backend& be, const Opt& o) : b(be), mixed(o.mixed), target(o.tol) { b.set_mixed(mixed); }
  bool on_check(double rel) {
    if (!mixed || !b.mixed()) return false;
    if (rel < 0.9 * best) best = rel, stale = 0;
    else ++stale;
    if ((rel <= 1e-6 && target < 1e-6) || stale >= 10) {
      b.set_mixed(false);
      return true;
    }
    return false;
  }
  void on_warm(double rel) {
    if (mixed && b.mixed() && rel <= 1e-6 && target < 1e-6) b.set_mixed(false);
  }
};

struct Ray {
  bool pinf = false, dinf = false;
  double dobj = 0, dviol = 0, pobj = 0, pviol = 0;
  std::vector<double> r, d;
};

Ray ray_test(const Scaled& s, const std::vector<double>& dxs, const std::vector<double>& dys) {  // engine.cpp:82-155
  const Lp& M = *s.orig;
  Ray t;
  std::vector<double> r(s.m), g(s.n);
  double rmax = 0;
  for (int i = 0; i < s.m; ++i) {
    double v = s.rs[i] * dys[i] / s.os;
    if (!std::isfinite(M.rl[i])) v = std::min(v, 0.0);
    if (!std::isfinite(M.ru[i])) v = std::max(v, 0.0);
    r[i] = v;
    rmax = std::max(rmax, std::fabs(v));
  }
  if (rmax > 0) {
    s.At0.mul(r.data(), g.data());
    std::vector<double> gabs(s.n, 0.0);
    for (int j = 0; j < s.n; ++j)
      for (int k = M.cs[j]; k < M.cs[j + 1]; ++k) gabs[j] += std::fabs(M.v[k] * r[M.ri[k]]);
    double obj = 0, viol = 0, mag = 0;
    for (int i = 0; i < s.m; ++i) {
      const double term = r[i] > 0 ? M.rl[i] * r[i] : (r[i] < 0 ? M.ru[i] * r[i] : 0.0);
      obj += term;
      mag += std::fabs(term);
    }
    for (int j = 0; j < s.n; ++j) {
      const double lam = -g[j], lo = M.cl[j], up = M.cu[j];
      if (lam > 0) {
        if (std::isfinite(lo)) obj += lo * lam, mag += gabs[j] * std::fabs(lo);
        else viol = std::max(viol, lam);
      } else if (lam < 0) {
        if (std::isfinite(up)) obj += up * lam, mag += gabs[j] * std::fabs(up);
        else viol = std::max(viol, -lam);
      }
    }
    const double sc = std::max(rmax, viol);
    t.dobj = obj / sc;
    t.dviol = viol / sc;
    t.pinf = t.dobj > 0 && obj > 1e-8 * mag && t.dviol <= 1e-8 * t.dobj;
    t.r = r;
  }
  std::vector<double> d(s.n), ad(s.m);
  double dmax = 0;
  for (int j = 0; j < s.n; ++j) {
    double v = s.cs[j] * dxs[j] / s.bs;
    if (std::isfinite(M.cl[j])) v = std::max(v, 0.0);
    if (std::isfinite(M.cu[j])) v = std::min(v, 0.0);
    d[j] = v;
    dmax = std::max(dmax, std::fabs(v));
  }
  if (dmax > 0) {
    s.A0.mul(d.data(), ad.data());
    double cd = 0, viol = 0, cmag = 0;
    for (int j = 0; j < s.n; ++j) cd += M.sense * M.c[j] * d[j], cmag += std::fabs(M.c[j] * d[j]);
    for (int i = 0; i < s.m; ++i) {
      if (std::isfinite(M.rl[i])) viol = std::max(viol, -ad[i]);
      if (std::isfinite(M.ru[i])) viol = std::max(viol, ad[i]);
    }
    const double sc = std::max(dmax, viol);
    t.pobj = cd / sc;
    t.pviol = viol / sc;
    t.dinf = t.pobj < 0 && -cd > 1e-8 * cmag && t.pviol <= 1e-8 * -t.pobj;
    t.d = d;
  }
  return t;
}

class Context {  // engine.cpp:195-307
 public:
  Context(const Lp& l, const Opt& o) : l_(l), o_(o), t0_(std::chrono::steady_clock::now()) {
    s_ = scale(l, o.sc);
    b_ = std::make_unique<Backend>(s_);
    const double nrm = norm2(s_.A, s_.At);
    eta_ = nrm > 0 ? 0.998 / nrm : 1.0;
    setup_ = elapsed();
  }
  Backend& b() { return *b_; }
  const Scaled& s() const { return s_; }
  double eta() const { return eta_; }
  double w0() const {
    if (o_.warm_w > 0) return o_.warm_w;
    if (o_.sc.bor) return 1.0;
    double c2 = 0, b2 = 0;
    for (double v : s_.c) c2 += v * v;
    for (int i = 0; i < s_.m; ++i) {
      const double a = std::isfinite(s_.rl[i]) ? std::fabs(s_.rl[i]) : 0.0, b = std::isfinite(s_.ru[i]) ? std::fabs(s_.ru[i]) : 0.0;
      b2 += std::max(a, b) * std::max(a, b);
    }
    return (std::sqrt(c2) + 1.0) / (std::sqrt(b2) + 1.0);
  }
  double elapsed() const { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0_).count(); }
  bool out_of_time() const { return elapsed() > o_.time_limit; }
  Kkt adjust(Kkt k) const {
    if (o_.b_norm >= 0) k.bn = o_.b_norm;
    if (o_.c_norm >= 0) k.cn = o_.c_norm;
    k.shift = o_.obj_shift;
    return k;
  }
  Kkt kkt(int xs, int ys) { return adjust(b_->kkt(xs, ys)); }
  bool record(const Kkt& k, std::int64_t it) {
    if (fast_it_ < 0 && k.rel() <= 1e-4) fast_it_ = it;
    return k.converged(o_.tol);
  }
  void note_w(double w) { w_ = w; }
  Status infeasibility(int dx, int dy, const Kkt& cur, std::string& msg) {  // engine.cpp:157-184
    std::vector<double> a, b;
    b_->download(dx, a);
    b_->download(dy, b);
    Ray t = ray_test(s_, a, b);
    char buf[160];
    if (t.pinf) {
      dual_ray_ = t.r;
      std::snprintf(buf, sizeof buf, "primal infeasible: dual ray certificate (objective %f, violation %f)", t.dobj, t.dviol);
      msg = buf;
      return Infeasible;
    }
    if (t.dinf) {
      const bool feas = cur.rel_p() <= 1e-6 && cur.pmax <= 1e-6;
      primal_ray_ = t.d;
      std::snprintf(buf, sizeof buf, "unbounded: primal ray certificate (c'd %f, violation %f)%s", t.pobj, t.pviol,
                    feas ? " at a primal-feasible iterate" : "; iterate not primal-feasible yet");
      msg = buf;
      return Unbounded;
    }
    return NotSolved;
  }
  bool warm_start(int x, int y, Precision& pol) {  // engine.cpp:237-266
    if (o_.warm_x.empty() && o_.warm_y.empty()) return false;
    std::vector<double> xo, yo(s_.m, 0.0);
    if (!o_.warm_x.empty()) {
      xo.resize(s_.n);
      for (int j = 0; j < s_.n; ++j) xo[j] = std::min(std::max(o_.warm_x[j], l_.cl[j]), l_.cu[j]);
    } else {
      std::vector<double> xs;
      b_->download(x, xs);
      s_.unscale_x(xs, xo);
    }
    if (!o_.warm_y.empty())
      for (int i = 0; i < s_.m; ++i) {
        double v = l_.sense * o_.warm_y[i];
        if (!std::isfinite(l_.rl[i])) v = std::min(v, 0.0);
        if (!std::isfinite(l_.ru[i])) v = std::max(v, 0.0);
        yo[i] = v;
      }
    pol.on_warm(adjust(kkt_original(s_, xo, yo)).rel());
    std::vector<double> xs(s_.n), ys(s_.m);
    for (int j = 0; j < s_.n; ++j) xs[j] = s_.bs * xo[j] / s_.cs[j];
    for (int i = 0; i < s_.m; ++i) ys[i] = s_.os * yo[i] / s_.rs[i];
    b_->upload(x, xs);
    b_->upload(y, ys);
    return true;
  }
  Sol finish(Status st, int xs, int ys, std::int64_t it, const std::string& msg) {  // engine.cpp:268-307
    Sol s;
    s.iterations = it;
    s.message = msg;
    std::vector<double> a, b, ym;
    b_->download(xs, a);
    b_->download(ys, b);
    s_.unscale_x(a, s.x);
    s_.unscale_y(b, ym);
    const Kkt k = adjust(kkt_original(s_, s.x, ym));
    if (!k.finite()) st = NumericalError;
    if (st == Optimal && !k.converged(o_.tol)) st = NumericalError;
    s.y.resize(s_.m);
    for (int i = 0; i < s_.m; ++i) s.y[i] = l_.sense * ym[i];
    s.z = l_.c;
    for (int j = 0; j < s_.n; ++j)
      for (int p = l_.cs[j]; p < l_.cs[j + 1]; ++p) s.z[j] -= l_.v[p] * s.y[l_.ri[p]];
    s.row_activity = l_.activity(s.x);
    s.objective = l_.objective(s.x);
    s.dual_objective = l_.sense * k.dobj + l_.offset;
    s.rp = k.rel_p();
    s.rd = k.rel_d();
    s.gap = k.rel_gap();
    s.status = st;
    s.it_fast = fast_it_;
    s.seconds = elapsed();
    s.setup_seconds = setup_;
    s.primal_weight = w_;
    if (st == Infeasible) s.dual_ray = dual_ray_;
    if (st == Unbounded) s.primal_ray = primal_ray_;
    if (st == Infeasible || st == Unbounded) s.objective = s.dual_objective = kNaN;
    return s;
  }

 private:
  const Lp& l_;
  const Opt& o_;
  Scaled s_;
  std::unique_ptr<Backend> b_;
  double eta_ = 1, setup_ = 0, w_ = kNaN;
  std::int64_t fast_it_ = -1;
  std::chrono::steady_clock::time_point t0_;
  std::vector<double> dual_ray_, primal_ray_;
};

Sol r2hpdhg(const Lp& l, const Opt& o) {  // r2hpdhg.cpp:11-162
  Context ctx(l, o);
  Backend& b = ctx.b();
  Precision pol(b, o);
  const double eta = ctx.eta(), rho = o.rho;
  double w = ctx.w0();
  int x = b.create(true), x0 = b.create(true), xh = b.create(true), xb = b.create(true), aty = b.create(true),
      dx = b.create(true);
  int y = b.create(false), y0 = b.create(false), yh = b.create(false), yb = b.create(false), ax = b.create(false),
      dy = b.create(false);
  b.fill(dx, 0.0);
  b.fill(aty, 0.0);
  b.primal_step(dx, aty, 0.0, x);
  b.fill(y, 0.0);
  const bool warm = ctx.warm_start(x, y, pol);
  ctx.note_w(w);
  if (warm) {
    const Kkt kw = ctx.kkt(x, y);
    if (kw.finite() && ctx.record(kw, 0)) return ctx.finish(Optimal, x, y, 0, "warm start already optimal");
  }
  b.copy(x0, x);
  b.copy(y0, y);
  b.copy(xh, x);
  b.copy(yh, y);
  double integ = 0, last_e = 0, best_w = w, best_bal = 1e300;
  auto residual = [&]() {
    b.copy(dx, xb);
    b.axpby(-1.0, xh, 1.0, dx);
    b.copy(dy, yb);
    b.axpby(-1.0, yh, 1.0, dy);
    b.spmv(dx, ax);
    const double px = b.dot(dx, dx), py = b.dot(dy, dy), cr = b.dot(dy, ax);
    const double r2 = w * px + py / w + 2.0 * eta * cr;
    return std::sqrt(std::max(r2, 0.0));
  };
  const int K = o.K;
  std::int64_t it = 0, inner = 0;
  double r0 = -1.0, r_last = kInf;
  std::string msg = warm ? "warm start" : "";
  for (;;) {
    const double tau = eta / w, sigma = eta * w;
    for (int s = 0; s < K; ++s) {
      const double hw = static_cast<double>(inner + 1) / static_cast<double>(inner + 2);
      const bool need = (s == K - 1) || inner == 0;
      b.spmv_t(y, aty);
      b.r2h_primal(x, x0, aty, tau, hw, rho, xh, xb);
      b.spmv(xb, ax);
      b.r2h_dual(y, y0, ax, sigma, hw, rho, yh, need ? yb : -1);
      if (inner == 0) r0 = residual();
      ++inner;
      ++it;
      if (s + 1 < K && ctx.out_of_time()) {
        const Kkt kt = ctx.kkt(xh, yh);
        if (kt.finite() && ctx.record(kt, it)) return ctx.finish(Optimal, xh, yh, it, msg);
        return ctx.finish(TimeLimit, xh, yh, it, msg);
      }
    }
    const double r = residual();
    const Kkt k = ctx.kkt(xh, yh);
    if (!k.finite() || !std::isfinite(r)) return ctx.finish(NumericalError, xh, yh, it, "non-finite iterate");
    if (ctx.record(k, it)) return ctx.finish(Optimal, xh, yh, it, msg);
    if (it >= o.iter_limit) return ctx.finish(IterationLimit, xh, yh, it, msg);
    if (ctx.out_of_time()) return ctx.finish(TimeLimit, xh, yh, it, msg);
    if (it >= 4 * K) {
      std::string why;
      Status st = ctx.infeasibility(dx, dy, k, why);
      if (st == NotSolved && inner >= K && (it / K) % 4 == 0) {
        b.copy(dx, xh);
        b.axpby(-1.0, x0, 1.0, dx);
        b.copy(dy, yh);
        b.axpby(-1.0, y0, 1.0, dy);
        st = ctx.infeasibility(dx, dy, k, why);
        if (st != NotSolved) why += " (drift since the restart anchor)";
      }
      if (st != NotSolved) return ctx.finish(st, xh, yh, it, why);
    }
    const bool promoted = pol.on_check(k.rel());
    if (promoted) msg += std::string(msg.empty() ? "" : "; ") + "fp32 -> fp64 at iteration " + std::to_string(it);
    const bool ok = r0 > 0 && std::isfinite(r0) && r > 0;
    const bool restart = promoted || (ok && (r <= o.r_suff * r0 || (r <= o.r_nec * r0 && r > r_last))) ||
                         static_cast<double>(inner) >= o.r_art * static_cast<double>(it);
    r_last = r;
    if (!restart) continue;
    const double dxn = b.diff_norm2(xh, x0), dyn = b.diff_norm2(yh, y0);
    const bool both = k.rel_d() > 0 && k.rel_p() > 0;
    const double ratio = both ? k.rel_d() / k.rel_p() : 1.0;
    if (dxn > 1e-16 && dyn > 1e-16 && dxn < 1e12 && dyn < 1e12 && ratio > 1e-8 && ratio < 1e8) {
      const double e = std::log(dyn) - std::log(dxn) - std::log(w);
      integ = o.idecay * integ + e;
      const double step = o.kp * e + o.ki * integ + o.kd * (e - last_e);
      w *= std::exp(std::clamp(step, -o.max_log_step, o.max_log_step));
      last_e = e;
    } else {
      w = best_w;
      integ = last_e = 0.0;
    }
    const double bal = std::fabs(std::log10(ratio));
    if (both && bal < best_bal) best_bal = bal, best_w = w;
    ctx.note_w(w);
    b.copy(x, xh);
    b.copy(x0, xh);
    b.copy(y, yh);
    b.copy(y0, yh);
    inner = 0;
    r_last = kInf;
  }
}

// ------------------------------------------------------------------ presolve (R1–R4)
struct Pre {  // presolve.h:32-48
  int outcome = 2;  // 0 reduced, 1 infeasible, 2 unchanged
  Lp red;
  std::string msg;
  int rows_removed = 0, cols_removed = 0;
  std::vector<int> row_map, col_map;
  std::vector<double> fixed;
  struct One {
    int row, col;
    double a;
    bool lo, up;
  };
  std::vector<One> ones;
};

Pre presolve(const Lp& M) {  // presolve.cpp:22-156 (LP only)
  Pre r;
  const int m = M.m, n = M.n;
  const Csr<double> A = transpose(csc_as_transpose(M));
  std::vector<double> rl = M.rl, ru = M.ru, cl = M.cl, cu = M.cu;
  std::vector<char> ra(m, 1), ca(n, 1);
  std::vector<int> rc(m), cc(n);
  for (int i = 0; i < m; ++i) rc[i] = static_cast<int>(A.ptr[i + 1] - A.ptr[i]);
  for (int j = 0; j < n; ++j) cc[j] = M.cs[j + 1] - M.cs[j];
  r.fixed.assign(n, kNaN);
  double offset = M.offset;
  const double ft = 1e-9;
  auto within = [&](double v, double lo, double up) {
    return v >= lo - ft * (1 + std::fabs(lo)) && v <= up + ft * (1 + std::fabs(up));
  };
  auto infeasible = [&](const std::string& why) {
    r.outcome = 1;
    r.msg = why;
    return r;
  };
  bool changed = true;
  while (changed) {
    changed = false;
    for (int i = 0; i < m; ++i) {
      if (!ra[i]) continue;
      if (rc[i] == 0) {
        if (!within(0.0, rl[i], ru[i])) return infeasible("empty row " + std::to_string(i) + " excludes 0");
        ra[i] = 0;
        ++r.rows_removed;
        changed = true;
      } else if (rc[i] == 1) {
        int j = -1;
        double a = 0;
        for (std::int64_t p = A.ptr[i]; p < A.ptr[i + 1]; ++p)
          if (ca[A.idx[p]]) j = A.idx[p], a = A.val[p];
        if (j < 0 || a == 0.0) continue;
        const double lo = a > 0 ? rl[i] / a : ru[i] / a;
        const double hi = a > 0 ? ru[i] / a : rl[i] / a;
        const bool sl = lo > cl[j], su = hi < cu[j];
        double nl = sl ? lo : cl[j], nu = su ? hi : cu[j];
        if (nl > nu) {
          if (nl - nu > ft * (1 + std::fabs(nl)))
            return infeasible("singleton row " + std::to_string(i) + " conflicts with the bounds of its column");
          nl = nu = 0.5 * (nl + nu);
        }
        cl[j] = nl;
        cu[j] = nu;
        r.ones.push_back({i, j, a, sl, su});
        ra[i] = 0;
        --cc[j];
        ++r.rows_removed;
        changed = true;
      }
    }
    for (int j = 0; j < n; ++j) {
      if (!ca[j]) continue;
      double v;
      if (cl[j] == cu[j]) {
        v = cl[j];
      } else if (cc[j] == 0) {
        const double c = M.sense * M.c[j];
        v = c > 0 ? cl[j] : (c < 0 ? cu[j] : std::min(std::max(0.0, cl[j]), cu[j]));
        if (!std::isfinite(v)) continue;
      } else {
        continue;
      }
      for (int p = M.cs[j]; p < M.cs[j + 1]; ++p) {
        const int i = M.ri[p];
        if (!ra[i]) continue;
        const double t = M.v[p] * v;
        if (std::isfinite(rl[i])) rl[i] -= t;
        if (std::isfinite(ru[i])) ru[i] -= t;
        --rc[i];
      }
      offset += M.c[j] * v;
      r.fixed[j] = v;
      ca[j] = 0;
      ++r.cols_removed;
      changed = true;
    }
  }
  if (r.rows_removed == 0 && r.cols_removed == 0) {
    r.outcome = 2;
    return r;
  }
  Lp& R = r.red;
  std::vector<int> nr(m, -1);
  for (int i = 0; i < m; ++i)
    if (ra[i]) nr[i] = static_cast<int>(r.row_map.size()), r.row_map.push_back(i);
  for (int j = 0; j < n; ++j)
    if (ca[j]) r.col_map.push_back(j);
  R.sense = M.sense;
  R.offset = offset;
  R.m = static_cast<int>(r.row_map.size());
  R.n = static_cast<int>(r.col_map.size());
  R.cs.push_back(0);
  for (int j : r.col_map) {
    R.c.push_back(M.c[j]);
    R.cl.push_back(cl[j]);
    R.cu.push_back(cu[j]);
    for (int p = M.cs[j]; p < M.cs[j + 1]; ++p)
      if (nr[M.ri[p]] >= 0) R.ri.push_back(nr[M.ri[p]]), R.v.push_back(M.v[p]);
    R.cs.push_back(static_cast<int>(R.v.size()));
  }
  for (int i : r.row_map) R.rl.push_back(rl[i]), R.ru.push_back(ru[i]);
  r.outcome = 0;
  return r;
}

Sol postsolve(const Lp& M, const Pre& r, const Sol& red) {  // presolve.cpp:158-193
  Sol s = red;
  const int m = M.m, n = M.n;
  s.x.assign(n, 0.0);
  s.y.assign(m, 0.0);
  const bool hx = red.x.size() == r.col_map.size(), hy = red.y.size() == r.row_map.size();
  for (int j = 0; j < n; ++j)
    if (!std::isnan(r.fixed[j])) s.x[j] = r.fixed[j];
  if (hx)
    for (std::size_t c = 0; c < r.col_map.size(); ++c) s.x[r.col_map[c]] = red.x[c];
  if (hy)
    for (std::size_t q = 0; q < r.row_map.size(); ++q) s.y[r.row_map[q]] = red.y[q];
  for (auto it = r.ones.rbegin(); it != r.ones.rend(); ++it) {
    const int j = it->col;
    double z = M.c[j];
    for (int p = M.cs[j]; p < M.cs[j + 1]; ++p) z -= M.v[p] * s.y[M.ri[p]];
    const double zm = M.sense * z;
    if ((zm > 0 && it->lo) || (zm < 0 && it->up)) s.y[it->row] = z / it->a;
  }
  s.z = M.c;
  for (int j = 0; j < n; ++j)
    for (int p = M.cs[j]; p < M.cs[j + 1]; ++p) s.z[j] -= M.v[p] * s.y[M.ri[p]];
  s.row_activity = M.activity(s.x);
  s.objective = M.objective(s.x);
  if (!hx) s.x.clear(), s.row_activity.clear(), s.objective = kNaN;
  if (!hy) s.y.clear(), s.z.clear();
  return s;
}

std::vector<double> postsolve_farkas(const Lp& M, const Pre& r, const std::vector<double>& red) {  // presolve.cpp:195-207
  std::vector<double> ray(M.m, 0.0);
  if (red.size() != r.row_map.size()) return {};
  for (std::size_t q = 0; q < r.row_map.size(); ++q) ray[r.row_map[q]] = red[q];
  for (auto it = r.ones.rbegin(); it != r.ones.rend(); ++it) {
    const int j = it->col;
    double lam = 0.0;
    for (int p = M.cs[j]; p < M.cs[j + 1]; ++p) lam -= M.v[p] * ray[M.ri[p]];
    if ((lam > 0 && it->lo) || (lam < 0 && it->up)) ray[it->row] = lam / it->a;
  }
  return ray;
}

// ------------------------------------------------------------------ double-double + checks
struct dd {
  double hi = 0, lo = 0;
  dd() = default;
  dd(double h) : hi(h) {}
  dd(double h, double l) : hi(h), lo(l) {}
  double d() const { return hi + lo; }
};
static inline void two_sum(double a, double b, double& s, double& e) {
  s = a + b;
  const double bb = s - a;
  e = (a - (s - bb)) + (b - bb);
}
static inline void fast_two_sum(double a, double b, double& s, double& e) {
  s = a + b;
  e = b - (s - a);
}
static inline dd operator+(const dd& a, const dd& b) {  // dd.h:81-92 (Hida–Li–Bailey accurate add)
  double s1, s2, t1, t2;
  two_sum(a.hi, b.hi, s1, s2);
  two_sum(a.lo, b.lo, t1, t2);
  s2 += t1;
  fast_two_sum(s1, s2, s1, s2);
  s2 += t2;
  dd r;
  fast_two_sum(s1, s2, r.hi, r.lo);
  return r;
}
static inline dd neg(const dd& a) { return {-a.hi, -a.lo}; }
static inline dd operator-(const dd& a, const dd& b) { return a + neg(b); }
static inline dd mul_exact(double a, double b) {
  dd r;
  r.hi = a * b;
  r.lo = std::fma(a, b, -r.hi);
  return r;
}
static inline dd ddmul(const dd& a, const dd& b) {  // dd.h:108-115
  double p1 = a.hi * b.hi, p2 = std::fma(a.hi, b.hi, -p1);
  p2 += a.hi * b.lo + a.lo * b.hi;
  dd r;
  fast_two_sum(p1, p2, r.hi, r.lo);
  return r;
}
static inline dd dddiv(const dd& a, const dd& b) {  // dd.h:135-146
  const double q1 = a.hi / b.hi;
  if (!std::isfinite(q1)) return dd(q1);
  dd r = a - ddmul(b, dd(q1));
  const double q2 = r.hi / b.hi;
  r = r - ddmul(b, dd(q2));
  const double q3 = r.hi / b.hi;
  dd q;
  fast_two_sum(q1, q2, q.hi, q.lo);
  // q + q3 with a double operand (dd.h:94-101)
  double s1, s2;
  two_sum(q.hi, q3, s1, s2);
  s2 += q.lo;
  dd out;
  fast_two_sum(s1, s2, out.hi, out.lo);
  return out;
}
static inline bool ddless(const dd& a, const dd& b) { return a.hi < b.hi || (a.hi == b.hi && a.lo < b.lo); }
static inline dd add_d(const dd& a, double b) {  // dd.h:94-101
  double s1, s2;
  two_sum(a.hi, b, s1, s2);
  s2 += a.lo;
  dd r;
  fast_two_sum(s1, s2, r.hi, r.lo);
  return r;
}

struct Check {  // solution_checker.h:33-47
  double pviol = 0, dviol = 0, pobj = 0, dobj = 0, gap = 0;
  bool ok_p = false, ok_d = false, ok_g = false;
  bool passed() const { return ok_p && ok_d && ok_g; }
};

Check check_solution(const Lp& lp, const std::vector<double>& x, const std::vector<double>& yin) {  // solution_checker.cpp:61-133
  Check c;
  auto reject = [&]() {
    c.pviol = c.dviol = c.gap = kInf;
    c.ok_p = c.ok_d = c.ok_g = false;
    return c;
  };
  if (static_cast<int>(x.size()) != lp.n || static_cast<int>(yin.size()) != lp.m) return reject();
  for (double v : x)
    if (!std::isfinite(v)) return reject();
  for (double v : yin)
    if (!std::isfinite(v)) return reject();
  auto bviol = [](double val, double lo, double up) {
    double v = 0.0;
    if (lo != -kInf) v = std::max(v, (lo - val) / (1.0 + std::fabs(lo)));
    if (up != kInf) v = std::max(v, (val - up) / (1.0 + std::fabs(up)));
    return v;
  };
  const std::vector<double> ax = lp.activity(x);
  for (int i = 0; i < lp.m; ++i) c.pviol = std::max(c.pviol, bviol(ax[i], lp.rl[i], lp.ru[i]));
  for (int j = 0; j < lp.n; ++j) c.pviol = std::max(c.pviol, bviol(x[j], lp.cl[j], lp.cu[j]));
  std::vector<double> y(lp.m), aty(lp.n, 0.0);
  for (int i = 0; i < lp.m; ++i) y[i] = lp.sense * yin[i];
  for (int j = 0; j < lp.n; ++j)
    for (int k = lp.cs[j]; k < lp.cs[j + 1]; ++k) aty[j] += lp.v[k] * y[lp.ri[k]];
  double dsum = 0, psum = 0, cscale = 1.0;
  for (int j = 0; j < lp.n; ++j) cscale = std::max(cscale, std::fabs(lp.c[j]));
  auto mult = [&](double mu, double lo, double up, double& contrib) {
    double viol = 0;
    contrib = 0;
    if (mu > 0.0) {
      if (lo == -kInf) viol = mu;
      else contrib = mu * lo;
    } else if (mu < 0.0) {
      if (up == kInf) viol = -mu;
      else contrib = mu * up;
    }
    return viol;
  };
  for (int i = 0; i < lp.m; ++i) {
    double ct;
    c.dviol = std::max(c.dviol, mult(y[i], lp.rl[i], lp.ru[i], ct) / cscale);
    dsum += ct;
  }
  for (int j = 0; j < lp.n; ++j) {
    const double cj = lp.sense * lp.c[j], dj = cj - aty[j];
    double ct;
    c.dviol = std::max(c.dviol, mult(dj, lp.cl[j], lp.cu[j], ct) / cscale);
    dsum += ct;
    psum += cj * x[j];
  }
  c.pobj = lp.sense * psum + lp.offset;
  c.dobj = lp.sense * dsum + lp.offset;
  c.gap = std::fabs(c.pobj - c.dobj) / (1.0 + std::fabs(c.pobj) + std::fabs(c.dobj));
  if (!std::isfinite(c.pobj) || !std::isfinite(c.dobj)) return reject();
  c.ok_p = c.pviol <= 1e-6;
  c.ok_d = c.dviol <= 1e-6;
  c.ok_g = c.gap <= 1e-6;
  return c;
}

// Rounding-proof Lagrangian bound (safe_bound.cpp:16-195).
constexpr double kRel = 2.0 * 4.930380657631324e-32, kAbs = 2.0 * 1.0e-289;
static double round_down(const dd& v, double err) {
  if (err == 0.0 && v.lo == 0.0) return v.hi;
  return std::nextafter((v - dd(err)).d(), -kInf);
}
static double round_up(const dd& v, double err) {
  if (err == 0.0 && v.lo == 0.0) return v.hi;
  return std::nextafter((v + dd(err)).d(), kInf);
}
static bool tiny(const dd& p) { return p.hi != 0.0 && std::fabs(p.hi) < 1e-280; }
static bool grow(std::vector<double>& e, double b, std::vector<double>& tmp) {  // Shewchuk GROW-EXPANSION
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
struct SafeAcc {
  dd sum;
  double abs_sum = 0;
  int count = 0;
  void add(const dd& t) {
    sum = sum + t;
    abs_sum += std::fabs(t.d());
    ++count;
  }
  double error() const { return (count + 2) * (kRel * abs_sum * 1.0000001 + kAbs); }
};
static dd min_term(double z, double lo, double hi, bool& inf) {
  if (z > 0) {
    if (!std::isfinite(lo)) return inf = true, dd(0.0);
    return mul_exact(z, lo);
  }
  if (z < 0) {
    if (!std::isfinite(hi)) return inf = true, dd(0.0);
    return mul_exact(z, hi);
  }
  return dd(0.0);
}

struct Bound {
  bool finite = false;
  double bound = 0;
  int row_inf = 0, col_inf = 0;
};

static Bound bound_with(const Lp& M, const std::vector<double>& y, const std::vector<double>& lo_b,
                        const std::vector<double>& up_b) {
  Bound r;
  const double sense = M.sense;
  SafeAcc L;
  for (int i = 0; i < M.m; ++i) {
    bool inf = false;
    const dd t = min_term(sense * y[i], M.rl[i], M.ru[i], inf);
    if (inf) ++r.row_inf;
    else L.add(t);
  }
  std::vector<double> ex, tmp;
  for (int j = 0; j < M.n; ++j) {
    dd z = sense * M.c[j];
    double mag = 0.0;
    int cnt = 0;
    bool uf = false;
    for (int p = M.cs[j]; p < M.cs[j + 1]; ++p) {
      const dd prod = mul_exact(M.v[p], sense * y[M.ri[p]]);
      if (prod.hi == 0.0 && prod.lo == 0.0 && M.v[p] * y[M.ri[p]] == 0.0) continue;
      z = z - prod;
      mag += std::fabs(prod.d()) + std::fabs(z.d());
      uf = uf || tiny(prod);
      ++cnt;
    }
    const double ez = cnt == 0 ? 0.0 : (cnt + 2) * (kRel * mag * 1.0000001 + (uf ? kAbs : 0.0));
    double zlo = round_down(z, ez), zhi = round_up(z, ez);
    if (cnt > 0 && !uf) {
      ex.clear();
      bool ok = grow(ex, sense * M.c[j], tmp);
      for (int p = M.cs[j]; ok && p < M.cs[j + 1]; ++p) {
        const dd prod = mul_exact(M.v[p], sense * y[M.ri[p]]);
        ok = grow(ex, -prod.hi, tmp) && grow(ex, -prod.lo, tmp);
      }
      if (ok) {
        if (ex.empty()) {
          zlo = zhi = 0.0;
        } else if (ex.size() == 1) {
          zlo = zhi = ex[0];
        } else {
          dd v = 0.0;
          for (double c : ex) v = v + dd(c);
          const double ev = static_cast<double>(ex.size() + 2) * kRel * std::fabs(v.d()) * 1.0000001;
          zlo = std::max(zlo, round_down(v, ev));
          zhi = std::min(zhi, round_up(v, ev));
        }
      }
    }
    bool i1 = false, i2 = false;
    const dd tlo = min_term(zlo, lo_b[j], up_b[j], i1), thi = min_term(zhi, lo_b[j], up_b[j], i2);
    if (i1 || i2) {
      ++r.col_inf;
      continue;
    }
    L.add(ddless(tlo, thi) ? tlo : thi);
  }
  if (r.row_inf > 0 || r.col_inf > 0) {
    r.finite = false;
    r.bound = sense > 0 ? -kInf : kInf;
    return r;
  }
  const double err = L.error() + kAbs;
  r.finite = true;
  if (sense > 0) r.bound = round_down(L.sum + dd(M.offset), err + kRel * std::fabs(M.offset) * 4);
  else r.bound = round_up(neg(L.sum) + dd(M.offset), err + kRel * std::fabs(M.offset) * 4);
  return r;
}

// Bounds implied by the rows, rounded outward (implied_bounds.cpp:11-111).
static void implied(const Lp& M, std::vector<double>& lo, std::vector<double>& up, int passes = 20) {
  lo = M.cl;
  up = M.cu;
  const Csr<double> A = transpose(csc_as_transpose(M));
  std::vector<char> dirty(M.m, 1);
  auto touch = [&](int j) {
    for (int k = M.cs[j]; k < M.cs[j + 1]; ++k) dirty[M.ri[k]] = 1;
  };
  const double slack = 1e-9;
  auto loosen_up = [&](double v, double s) { return std::nextafter(v + slack * (s + std::fabs(v)), kInf); };
  auto loosen_down = [&](double v, double s) { return std::nextafter(v - slack * (s + std::fabs(v)), -kInf); };
  for (int pass = 0; pass < passes; ++pass) {
    bool changed = false;
    for (int i = 0; i < M.m; ++i) {
      if (!dirty[i]) continue;
      dirty[i] = 0;
      const bool hu = std::isfinite(M.ru[i]), hl = std::isfinite(M.rl[i]);
      if (!hu && !hl) continue;
      dd amin = 0.0, amax = 0.0;
      double mag = 0.0;
      int nmin = 0, nmax = 0, jmin = -1, jmax = -1;
      for (std::int64_t p = A.ptr[i]; p < A.ptr[i + 1]; ++p) {
        const int j = A.idx[p];
        const double a = A.val[p];
        const double bmin = a > 0 ? lo[j] : up[j], bmax = a > 0 ? up[j] : lo[j];
        if (std::isfinite(bmin)) {
          const dd t = mul_exact(a, bmin);
          amin = amin + t;
          mag += std::fabs(t.d());
        } else {
          ++nmin, jmin = j;
        }
        if (std::isfinite(bmax)) {
          const dd t = mul_exact(a, bmax);
          amax = amax + t;
          mag += std::fabs(t.d());
        } else {
          ++nmax, jmax = j;
        }
      }
      for (std::int64_t p = A.ptr[i]; p < A.ptr[i + 1]; ++p) {
        const int j = A.idx[p];
        const double a = A.val[p];
        const double sc = mag / std::fabs(a);
        if (hu && (nmin == 0 || (nmin == 1 && jmin == j))) {
          const double bmin = a > 0 ? lo[j] : up[j];
          dd rest = amin;
          if (nmin == 0) rest = rest - mul_exact(a, bmin);
          const double v = dddiv(dd(M.ru[i]) - rest, dd(a)).d();
          if (a > 0) {
            const double nb = loosen_up(v, sc);
            if (nb < up[j] - 1e-7 * (1 + std::fabs(nb))) up[j] = nb, changed = true, touch(j);
          } else {
            const double nb = loosen_down(v, sc);
            if (nb > lo[j] + 1e-7 * (1 + std::fabs(nb))) lo[j] = nb, changed = true, touch(j);
          }
        }
        if (hl && (nmax == 0 || (nmax == 1 && jmax == j))) {
          const double bmax = a > 0 ? up[j] : lo[j];
          dd rest = amax;
          if (nmax == 0) rest = rest - mul_exact(a, bmax);
          const double v = dddiv(dd(M.rl[i]) - rest, dd(a)).d();
          if (a > 0) {
            const double nb = loosen_down(v, sc);
            if (nb > lo[j] + 1e-7 * (1 + std::fabs(nb))) lo[j] = nb, changed = true, touch(j);
          } else {
            const double nb = loosen_up(v, sc);
            if (nb < up[j] - 1e-7 * (1 + std::fabs(nb))) up[j] = nb, changed = true, touch(j);
          }
        }
      }
    }
    if (!changed) break;
  }
}

Bound certified_bound(const Lp& M, const std::vector<double>& yin, bool use_implied) {  // safe_bound.cpp:178-195
  if (static_cast<int>(yin.size()) != M.m) return Bound{};
  std::vector<double> y = yin;
  for (int i = 0; i < M.m; ++i) {
    const double ym = M.sense * y[i];
    if ((ym > 0 && !std::isfinite(M.rl[i])) || (ym < 0 && !std::isfinite(M.ru[i]))) y[i] = 0.0;
  }
  Bound r = bound_with(M, y, M.cl, M.cu);
  if (use_implied && r.col_inf > 0 && r.row_inf == 0) {
    std::vector<double> lo, up;
    implied(M, lo, up);
    r = bound_with(M, y, lo, up);
  }
  return r;
}

// Farkas / ray certificate checks (certificates.cpp:13-135). Returns 1 pass, 0 fail, -1 absent.
int check_farkas(const Lp& M, const std::vector<double>& r) {
  if (static_cast<int>(r.size()) != M.m || M.m == 0) return -1;
  for (double v : r)
    if (!std::isfinite(v)) return 0;
  Lp zero = M;
  zero.sense = 1;
  zero.offset = 0.0;
  std::fill(zero.c.begin(), zero.c.end(), 0.0);
  const Bound sb = certified_bound(zero, r, true);
  if (sb.finite && sb.bound > 0.0) return 1;
  double rmax = 0, obj = 0, viol = 0, mag = 0;
  for (int i = 0; i < M.m; ++i) {
    rmax = std::max(rmax, std::fabs(r[i]));
    if (r[i] > 0) {
      if (std::isfinite(M.rl[i])) obj += r[i] * M.rl[i], mag += std::fabs(r[i] * M.rl[i]);
      else viol = std::max(viol, r[i]);
    } else if (r[i] < 0) {
      if (std::isfinite(M.ru[i])) obj += r[i] * M.ru[i], mag += std::fabs(r[i] * M.ru[i]);
      else viol = std::max(viol, -r[i]);
    }
  }
  std::vector<double> atr(M.n, 0.0), atra(M.n, 0.0);
  for (int j = 0; j < M.n; ++j)
    for (int k = M.cs[j]; k < M.cs[j + 1]; ++k) atr[j] += M.v[k] * r[M.ri[k]], atra[j] += std::fabs(M.v[k] * r[M.ri[k]]);
  for (int j = 0; j < M.n; ++j) {
    const double lam = -atr[j];
    if (lam > 0) {
      if (std::isfinite(M.cl[j])) obj += lam * M.cl[j], mag += atra[j] * std::fabs(M.cl[j]);
      else viol = std::max(viol, lam);
    } else if (lam < 0) {
      if (std::isfinite(M.cu[j])) obj += lam * M.cu[j], mag += atra[j] * std::fabs(M.cu[j]);
      else viol = std::max(viol, -lam);
    }
  }
  const double sc = std::max(rmax, viol);
  const double meas = sc > 0 ? obj / sc : 0.0, vv = sc > 0 ? viol / sc : 0.0;
  return std::isfinite(meas) && meas > 0 && obj > 1e-8 * mag && vv <= 1e-8 * meas ? 1 : 0;
}

int check_ray(const Lp& M, const std::vector<double>& x, const std::vector<double>& ray) {
  if (static_cast<int>(ray.size()) != M.n || static_cast<int>(x.size()) != M.n || M.n == 0) return -1;
  double dmax = 0, cmax = 0;
  for (int j = 0; j < M.n; ++j) {
    if (!std::isfinite(ray[j]) || !std::isfinite(x[j])) return 0;
    dmax = std::max(dmax, std::fabs(ray[j]));
    cmax = std::max(cmax, std::fabs(M.c[j]));
  }
  if (dmax == 0.0) return 0;
  const Check rep = check_solution(M, x, std::vector<double>(M.m, 0.0));
  std::vector<double> d(M.n);
  double cd = 0, viol = 0;
  for (int j = 0; j < M.n; ++j) {
    d[j] = ray[j] / dmax;
    cd += M.sense * M.c[j] * d[j];
    if (std::isfinite(M.cl[j])) viol = std::max(viol, -d[j]);
    if (std::isfinite(M.cu[j])) viol = std::max(viol, d[j]);
  }
  const std::vector<double> ad = M.activity(d);
  for (int i = 0; i < M.m; ++i) {
    if (std::isfinite(M.rl[i])) viol = std::max(viol, -ad[i]);
    if (std::isfinite(M.ru[i])) viol = std::max(viol, ad[i]);
  }
  return rep.ok_p && -cd >= 1e-8 * std::max(1.0, cmax) && viol <= 1e-8 * (-cd) ? 1 : 0;
}

// ------------------------------------------------------------------ dispatcher
struct SolveOpt {
  Opt e;
  bool presolve = true;
  std::vector<std::pair<std::string, double>> params;
};

static std::string crossed(const Lp& l) {
  for (int j = 0; j < l.n; ++j)
    if (l.cl[j] > l.cu[j]) return "column " + std::to_string(j);
  for (int i = 0; i < l.m; ++i)
    if (l.rl[i] > l.ru[i]) return "row " + std::to_string(i);
  return {};
}

Sol solve_direct(const Lp& l, const SolveOpt& so, bool certify) {  // solve.cpp:469-609 (r²HPDHG, LP)
  const auto t0 = std::chrono::steady_clock::now();
  Sol s;
  if ((!so.e.warm_x.empty() && static_cast<int>(so.e.warm_x.size()) != l.n) ||
      (!so.e.warm_y.empty() && static_cast<int>(so.e.warm_y.size()) != l.m)) {
    s.message = "warm start vectors have the wrong size";
    return s;
  }
  for (const auto* v : {&so.e.warm_x, &so.e.warm_y})
    for (double a : *v)
      if (!std::isfinite(a)) {
        s.message = "warm start contains non-finite values";
        return s;
      }
  if (auto why = crossed(l); !why.empty()) {
    s.status = Infeasible;
    s.message = "bounds cross: " + why;
    return s;
  }
  Opt eo = so.e;
  if (auto bad = apply_params(eo, so.params); !bad.empty()) {
    s.message = "unknown engine parameter '" + bad + "'";
    return s;
  }
  s = r2hpdhg(l, eo);
  if (s.status == Unbounded && check_ray(l, s.x, s.primal_ray) != 1) {
    Lp feas = l;
    std::fill(feas.c.begin(), feas.c.end(), 0.0);
    SolveOpt fo = so;
    fo.presolve = false;
    fo.e.warm_x.clear();
    fo.e.warm_y.clear();
    fo.e.time_limit = std::max(0.0, so.e.time_limit - std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
    fo.e.iter_limit = std::max<std::int64_t>(0, so.e.iter_limit - s.iterations);
    Sol f = solve_direct(feas, fo, true);
    const std::int64_t total = s.iterations + f.iterations;
    if (f.status == Optimal) {
      s.x = f.x;
      s.row_activity = l.activity(s.x);
      s.message += "; feasible point from a zero-objective solve";
    } else if (f.status == Infeasible) {
      f.message = "unbounded direction found, but the model is infeasible: " + f.message;
      f.primal_ray.clear();
      s = f;
    } else {
      s.status = f.status;
      s.primal_ray.clear();
    }
    s.iterations = total;
  }
  if (certify && static_cast<int>(s.y.size()) == l.m && l.m + l.n > 0)
    s.certified = certified_bound(l, s.y, true).bound;
  s.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return s;
}

static bool passes_original(const Lp& l, const Sol& s, double tol) {
  if (static_cast<int>(s.x.size()) != l.n || static_cast<int>(s.y.size()) != l.m) return false;
  Scaled tmp;  // only A0/At0 and the original are needed for the KKT evaluation
  tmp.orig = &l;
  tmp.At0 = csc_as_transpose(l);
  tmp.A0 = transpose(tmp.At0);
  std::vector<double> ym(s.y);
  for (double& v : ym) v *= l.sense;
  return kkt_original(tmp, s.x, ym).converged(tol);
}

Sol solve_impl(const Lp& l, const SolveOpt& so) {  // solve.cpp:266-467 (LP)
  if (!so.presolve) return solve_direct(l, so, true);
  if ((!so.e.warm_x.empty() && static_cast<int>(so.e.warm_x.size()) != l.n) ||
      (!so.e.warm_y.empty() && static_cast<int>(so.e.warm_y.size()) != l.m))
    return solve_direct(l, so, true);
  for (const auto* v : {&so.e.warm_x, &so.e.warm_y})
    for (double a : *v)
      if (!std::isfinite(a)) return solve_direct(l, so, true);
  if (auto why = crossed(l); !why.empty()) {
    Sol s;
    s.status = Infeasible;
    s.message = "bounds cross: " + why;
    return s;
  }
  const auto t0 = std::chrono::steady_clock::now();
  Pre pr = presolve(l);
  if (pr.outcome == 2) return solve_direct(l, so, true);
  SolveOpt inner = so;
  inner.presolve = false;
  {
    double b2 = 0, c2 = 0;
    for (int i = 0; i < l.m; ++i) {
      const double a = std::isfinite(l.rl[i]) ? std::fabs(l.rl[i]) : 0.0, b = std::isfinite(l.ru[i]) ? std::fabs(l.ru[i]) : 0.0;
      b2 += std::max(a, b) * std::max(a, b);
    }
    for (double c : l.c) c2 += c * c;
    inner.e.b_norm = std::sqrt(b2);
    inner.e.c_norm = std::sqrt(c2);
    inner.e.obj_shift = l.sense * (pr.red.offset - l.offset);
  }
  auto elapsed = [&] { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); };
  auto budget = [&](SolveOpt o, std::int64_t used) {
    o.e.time_limit = std::max(0.0, so.e.time_limit - elapsed());
    o.e.iter_limit = std::max<std::int64_t>(0, so.e.iter_limit - used);
    return o;
  };
  auto on_original = [&](const Sol& why) {
    SolveOpt d = budget(so, why.iterations);
    d.presolve = false;
    Sol s = solve_direct(l, d, true);
    s.iterations += why.iterations;
    s.seconds = elapsed();
    return s;
  };
  if (pr.outcome == 1) return on_original(Sol{});
  if (!so.e.warm_x.empty()) {
    inner.e.warm_x.clear();
    for (int j : pr.col_map) inner.e.warm_x.push_back(so.e.warm_x[j]);
  }
  if (!so.e.warm_y.empty()) {
    inner.e.warm_y.clear();
    for (int i : pr.row_map) inner.e.warm_y.push_back(so.e.warm_y[i]);
  }
  Sol red;
  if (pr.red.n == 0 && pr.red.m == 0) {
    red.status = Optimal;
    red.engine = "presolve";
  } else {
    red = solve_direct(pr.red, budget(inner, 0), false);
  }
  if (red.status == Infeasible && static_cast<int>(red.dual_ray.size()) == pr.red.m) {
    std::vector<double> r = postsolve_farkas(l, pr, red.dual_ray);
    if (!r.empty() && check_farkas(l, r) == 1) {
      Sol s = red;
      s.x.clear();
      s.y.clear();
      s.z.clear();
      s.row_activity.clear();
      s.dual_ray = r;
      s.seconds = elapsed();
      return s;
    }
  }
  if (red.status == Unbounded && static_cast<int>(red.primal_ray.size()) == pr.red.n &&
      static_cast<int>(red.x.size()) == pr.red.n) {
    std::vector<double> d(l.n, 0.0);
    for (int j = 0; j < pr.red.n; ++j) d[pr.col_map[j]] = red.primal_ray[j];
    Sol s = postsolve(l, pr, red);
    if (check_ray(l, s.x, d) == 1) {
      s.status = Unbounded;
      s.primal_ray = d;
      s.dual_ray.clear();
      s.y.clear();
      s.z.clear();
      s.objective = s.dual_objective = kNaN;
      s.seconds = elapsed();
      return s;
    }
  }
  if (red.status == Infeasible || red.status == Unbounded) return on_original(red);
  Sol post = postsolve(l, pr, red);
  if (post.status == Optimal && !passes_original(l, post, so.e.tol)) {
    SolveOpt tight = inner;
    tight.e.tol = std::max(so.e.tol * 1e-2, 1e-13);
    tight.e.warm_x = red.x;
    tight.e.warm_y = red.y;
    const Sol red2 = solve_direct(pr.red, budget(tight, red.iterations), true);
    Sol post2 = postsolve(l, pr, red2);
    post2.iterations += red.iterations;
    if (post2.status == Optimal && passes_original(l, post2, so.e.tol)) {
      post = post2;
    } else {
      Sol cold = solve_direct(l, budget(inner, red.iterations + red2.iterations), true);
      cold.iterations += red.iterations + red2.iterations;
      post = cold;
    }
  } else if (post.status == Optimal && static_cast<int>(post.x.size()) != l.n) {
    post.status = NumericalError;
  }
  if (static_cast<int>(post.y.size()) == l.m && l.m + l.n > 0) post.certified = certified_bound(l, post.y, true).bound;
  post.seconds = elapsed();
  return post;
}

void gate(const Lp& l, double tol, Sol& s) {  // solve.cpp:119-185 (LP, first-order engine)
  if (s.status == Infeasible || s.status == Unbounded) {
    s.objective = s.dual_objective = kNaN;
    if (s.status == Infeasible && !crossed(l).empty()) {
      s.check = 1;
      return;
    }
    const int c = s.status == Infeasible ? check_farkas(l, s.dual_ray) : check_ray(l, s.x, s.primal_ray);
    if (c == -1) {
      s.check = -1;
    } else if (c == 1) {
      s.check = 1;
    } else {
      s.check = 0;
      s.status = NumericalError;
    }
    return;
  }
  if (s.status != Optimal) return;
  if (static_cast<int>(s.x.size()) != l.n || static_cast<int>(s.y.size()) != l.m) {
    s.status = NumericalError;
    return;
  }
  const Check c = check_solution(l, s.x, s.y);
  s.check = c.passed() ? 1 : 0;
  s.check_p = c.pviol;
  s.check_d = c.dviol;
  s.check_g = c.gap;
  if (std::isnan(s.dual_objective)) s.dual_objective = c.dobj;
  if (c.passed()) return;
  if (tol < 1e-6) s.status = NumericalError;
}

// ------------------------------------------------------------------ batched scenarios
// Lockstep r²HPDHG over K scenarios sharing A (reference src/pdhg/batch.cpp:129-441): shared
// scaling without the bound/objective scalars, per-scenario bs/os, interleaved n×W layout,
// SpMM, per-scenario restarts / PID / termination, compaction of finished scenarios.
struct Scen {
  std::vector<double> c, cl, cu, rl, ru;
};

static void spmm(const Csr<double>& M, const std::vector<double>& X, std::vector<double>& Y, int W) {  // batch.cpp:40-91
  std::vector<double> acc(W);
  for (int i = 0; i < M.rows; ++i) {
    std::fill(acc.begin(), acc.end(), 0.0);
    for (std::int64_t p = M.ptr[i]; p < M.ptr[i + 1]; ++p) {
      const double a = M.val[p];
      const double* x = X.data() + static_cast<std::size_t>(M.idx[p]) * W;
      for (int k = 0; k < W; ++k) acc[k] += a * x[k];
    }
    std::copy(acc.begin(), acc.end(), Y.begin() + static_cast<std::size_t>(i) * W);
  }
}

static void compact(std::vector<double>& v, std::size_t rows, int W, const std::vector<int>& keep) {
  const int W2 = static_cast<int>(keep.size());
  std::vector<double> nv(rows * W2);
  for (std::size_t r = 0; r < rows; ++r)
    for (int t = 0; t < W2; ++t) nv[r * W2 + t] = v[r * W + keep[t]];
  v.swap(nv);
}

std::vector<Sol> solve_batch(const Lp& base, const std::vector<Scen>& S, const Opt& eo) {
  const auto t0 = std::chrono::steady_clock::now();
  auto elapsed = [&] { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); };
  const int K = static_cast<int>(S.size());
  std::vector<Sol> out(K);
  if (K == 0) return out;
  const int m = base.m, n = base.n;
  std::vector<Scen> D(K);
  for (int k = 0; k < K; ++k) {
    auto pick = [](const std::vector<double>& v, const std::vector<double>& d, std::size_t len, bool& ok) {
      if (v.empty()) return d;
      ok = ok && v.size() == len;
      return v;
    };
    bool ok = true;
    D[k].c = pick(S[k].c, base.c, n, ok);
    D[k].cl = pick(S[k].cl, base.cl, n, ok);
    D[k].cu = pick(S[k].cu, base.cu, n, ok);
    D[k].rl = pick(S[k].rl, base.rl, m, ok);
    D[k].ru = pick(S[k].ru, base.ru, m, ok);
    for (int j = 0; ok && j < n; ++j) ok = std::isfinite(D[k].c[j]) && !std::isnan(D[k].cl[j]) && !std::isnan(D[k].cu[j]);
    for (int i = 0; ok && i < m; ++i) ok = !std::isnan(D[k].rl[i]) && !std::isnan(D[k].ru[i]);
    bool crossed = false;
    for (int j = 0; ok && !crossed && j < n; ++j) crossed = D[k].cl[j] > D[k].cu[j];
    for (int i = 0; ok && !crossed && i < m; ++i) crossed = D[k].rl[i] > D[k].ru[i];
    if (!ok) {
      out[k].status = NotSolved;
      out[k].message = "scenario has wrong vector sizes or non-finite data";
    } else if (crossed) {
      out[k].status = Infeasible;
      out[k].message = "bounds cross: scenario has a lower bound above its upper bound";
    }
  }
  ScaleOpt so = eo.sc;
  so.bor = false;
  const Scaled sp = scale(base, so);
  const double nrm = norm2(sp.A, sp.At);
  const double eta = nrm > 0 ? 0.998 / nrm : 1.0;
  const double setup = elapsed();
  const std::size_t NK = static_cast<std::size_t>(n) * K, MK = static_cast<std::size_t>(m) * K;
  std::vector<double> C(NK), L(NK), U(NK), RL(MK), RU(MK), bs(K, 1.0), os(K, 1.0);
  for (int k = 0; k < K; ++k) {
    if (!out[k].message.empty()) continue;
    if (eo.sc.bor) {
      double b2 = 0, c2 = 0;
      for (int i = 0; i < m; ++i) {
        const double lo = std::isfinite(D[k].rl[i]) ? std::fabs(D[k].rl[i]) : 0.0;
        const double up = std::isfinite(D[k].ru[i]) ? std::fabs(D[k].ru[i]) : 0.0;
        const double b = sp.rs[i] * std::max(lo, up);
        b2 += b * b;
      }
      for (int j = 0; j < n; ++j) {
        const double c = sp.cs[j] * D[k].c[j];
        c2 += c * c;
      }
      bs[k] = 1.0 / (std::sqrt(b2) + 1.0);
      os[k] = 1.0 / (std::sqrt(c2) + 1.0);
    }
    for (int j = 0; j < n; ++j) {
      const std::size_t q = static_cast<std::size_t>(j) * K + k;
      C[q] = os[k] * sp.cs[j] * base.sense * D[k].c[j];
      L[q] = bs[k] * D[k].cl[j] / sp.cs[j];
      U[q] = bs[k] * D[k].cu[j] / sp.cs[j];
    }
    for (int i = 0; i < m; ++i) {
      const std::size_t q = static_cast<std::size_t>(i) * K + k;
      RL[q] = bs[k] * sp.rs[i] * D[k].rl[i];
      RU[q] = bs[k] * sp.rs[i] * D[k].ru[i];
    }
  }
  std::vector<int> sid;
  for (int k = 0; k < K; ++k)
    if (out[k].message.empty()) sid.push_back(k);
  {
    std::vector<int> keep(sid.begin(), sid.end());
    for (auto* v : {&C, &L, &U}) compact(*v, n, K, keep);
    for (auto* v : {&RL, &RU}) compact(*v, m, K, keep);
  }
  int W = static_cast<int>(sid.size());
  std::vector<double> X(static_cast<std::size_t>(n) * W), X0, XH, XB(X.size()), ATY(X.size()), DX(X.size());
  std::vector<double> Y(static_cast<std::size_t>(m) * W, 0.0), Y0, YH, YB(Y.size()), AX(Y.size()), DY(Y.size()), ADX(Y.size());
  for (std::size_t q = 0; q < X.size(); ++q) X[q] = clampv<double>(0.0, L[q], U[q]);
  X0 = X;
  XH = X;
  Y0 = Y;
  YH = Y;
  std::vector<double> omega(K, 1.0), best_w(K, 1.0), best_bal(K, 1e300), integ(K, 0.0), last_e(K, 0.0);
  std::vector<double> r0(K, -1.0), r_last(K, kInf);
  std::vector<std::int64_t> inner(K, 0), fast(K, -1);
  const double rho = eo.rho, a2 = 2 * rho, b2c = 1 - 2 * rho;
  const int Kc = eo.K;
  std::int64_t it = 0;
  std::vector<double> ct, cg, cw, resid, xk, yk;
  auto unscale = [&](int t, std::vector<double>& x, std::vector<double>& y) {
    const int k = sid[t];
    x.resize(n);
    y.resize(m);
    for (int j = 0; j < n; ++j)
      x[j] = clampv<double>(sp.cs[j] * XH[static_cast<std::size_t>(j) * W + t] / bs[k], D[k].cl[j], D[k].cu[j]);
    for (int i = 0; i < m; ++i) y[i] = sp.rs[i] * YH[static_cast<std::size_t>(i) * W + t] / os[k];
  };
  auto finish = [&](int k, Status st, const Kkt& kk, const std::vector<double>& x, const std::vector<double>& ym) {
    Sol& s = out[k];
    s.status = st;
    s.x = x;
    s.y.resize(m);
    for (int i = 0; i < m; ++i) s.y[i] = base.sense * ym[i];
    s.z = D[k].c;
    for (int j = 0; j < n; ++j)
      for (int p = base.cs[j]; p < base.cs[j + 1]; ++p) s.z[j] -= base.v[p] * s.y[base.ri[p]];
    double obj = base.offset;
    for (int j = 0; j < n; ++j) obj += D[k].c[j] * x[j];
    s.objective = obj;
    s.dual_objective = base.sense * kk.dobj + base.offset;
    s.rp = kk.rel_p();
    s.rd = kk.rel_d();
    s.gap = kk.rel_gap();
    s.iterations = it;
    s.seconds = elapsed();
    s.setup_seconds = setup;
    s.it_fast = fast[k];
    s.primal_weight = omega[k];
    s.engine = "r2hpdhg-batch";
  };
  auto residuals = [&]() {
    for (std::size_t q = 0; q < X.size(); ++q) DX[q] = XB[q] - XH[q];
    for (std::size_t q = 0; q < Y.size(); ++q) DY[q] = YB[q] - YH[q];
    spmm(sp.A, DX, ADX, W);
    std::vector<double> px(W, 0.0), py(W, 0.0), cr(W, 0.0);
    for (std::size_t q = 0; q < X.size(); ++q) px[q % W] += DX[q] * DX[q];
    for (std::size_t q = 0; q < Y.size(); ++q) py[q % W] += DY[q] * DY[q], cr[q % W] += DY[q] * ADX[q];
    resid.assign(W, 0.0);
    for (int t = 0; t < W; ++t) {
      const double om = omega[sid[t]];
      resid[t] = std::sqrt(std::max(om * px[t] + py[t] / om + 2.0 * eta * cr[t], 0.0));
    }
  };
  while (W > 0) {
    for (int step = 0; step < Kc; ++step) {
      ct.resize(W);
      cg.resize(W);
      cw.resize(W);
      for (int t = 0; t < W; ++t) {
        const int k = sid[t];
        ct[t] = eta / omega[k];
        cg[t] = eta * omega[k];
        cw[t] = static_cast<double>(inner[k] + 1) / static_cast<double>(inner[k] + 2);
      }
      spmm(sp.At, Y, ATY, W);
      for (int j = 0; j < n; ++j)
        for (int t = 0; t < W; ++t) {
          const std::size_t q = static_cast<std::size_t>(j) * W + t;
          const double h = clampv<double>(X[q] - ct[t] * (C[q] - ATY[q]), L[q], U[q]);
          XH[q] = h;
          XB[q] = 2 * h - X[q];
          X[q] = cw[t] * (a2 * h + b2c * X[q]) + (1 - cw[t]) * X0[q];
        }
      spmm(sp.A, XB, AX, W);
      for (int i = 0; i < m; ++i)
        for (int t = 0; t < W; ++t) {
          const std::size_t q = static_cast<std::size_t>(i) * W + t;
          const double sg = cg[t];
          const double v = AX[q] - Y[q] / sg;
          const double h = Y[q] - sg * AX[q] + sg * clampv<double>(v, RL[q], RU[q]);
          YH[q] = h;
          YB[q] = 2 * h - Y[q];
          Y[q] = cw[t] * (a2 * h + b2c * Y[q]) + (1 - cw[t]) * Y0[q];
        }
      bool any_new = false;
      for (int t = 0; t < W; ++t) any_new = any_new || inner[sid[t]] == 0;
      if (any_new) {
        residuals();
        for (int t = 0; t < W; ++t)
          if (inner[sid[t]] == 0) r0[sid[t]] = resid[t];
      }
      for (int t = 0; t < W; ++t) ++inner[sid[t]];
    }
    it += Kc;
    residuals();
    const bool over_iter = it >= eo.iter_limit, over_time = elapsed() > eo.time_limit;
    std::vector<int> keep;
    for (int t = 0; t < W; ++t) {
      const int k = sid[t];
      unscale(t, xk, yk);
      const View lv{m, n, base.sense, D[k].c.data(), D[k].cl.data(), D[k].cu.data(), D[k].rl.data(), D[k].ru.data()};
      const Kkt kk = kkt_of(sp.A0, sp.At0, lv, xk, yk);
      if (!kk.finite()) {
        finish(k, NumericalError, kk, xk, yk);
        out[k].message = "non-finite iterate";
        continue;
      }
      if (fast[k] < 0 && kk.rel() <= 1e-4) fast[k] = it;
      if (kk.converged(eo.tol)) {
        finish(k, Optimal, kk, xk, yk);
        continue;
      }
      if (over_iter || over_time) {
        finish(k, over_iter ? IterationLimit : TimeLimit, kk, xk, yk);
        continue;
      }
      keep.push_back(t);
      const double r = resid[t];
      const bool ok = r0[k] > 0 && std::isfinite(r0[k]) && r > 0;
      const bool restart = (ok && (r <= eo.r_suff * r0[k] || (r <= eo.r_nec * r0[k] && r > r_last[k]))) ||
                           static_cast<double>(inner[k]) >= eo.r_art * static_cast<double>(it);
      r_last[k] = r;
      if (!restart) continue;
      double dxn = 0, dyn = 0;
      for (int j = 0; j < n; ++j) {
        const std::size_t q = static_cast<std::size_t>(j) * W + t;
        dxn += (XH[q] - X0[q]) * (XH[q] - X0[q]);
      }
      for (int i = 0; i < m; ++i) {
        const std::size_t q = static_cast<std::size_t>(i) * W + t;
        dyn += (YH[q] - Y0[q]) * (YH[q] - Y0[q]);
      }
      dxn = std::sqrt(dxn);
      dyn = std::sqrt(dyn);
      const bool both = kk.rel_d() > 0 && kk.rel_p() > 0;
      const double ratio = both ? kk.rel_d() / kk.rel_p() : 1.0;
      if (dxn > 1e-16 && dyn > 1e-16 && dxn < 1e12 && dyn < 1e12 && ratio > 1e-8 && ratio < 1e8) {
        const double e = std::log(dyn) - std::log(dxn) - std::log(omega[k]);
        integ[k] = eo.idecay * integ[k] + e;
        const double stp = eo.kp * e + eo.ki * integ[k] + eo.kd * (e - last_e[k]);
        omega[k] *= std::exp(std::clamp(stp, -eo.max_log_step, eo.max_log_step));
        last_e[k] = e;
      } else {
        omega[k] = best_w[k];
        integ[k] = last_e[k] = 0.0;
      }
      if (both && std::fabs(std::log10(ratio)) < best_bal[k]) {
        best_bal[k] = std::fabs(std::log10(ratio));
        best_w[k] = omega[k];
      }
      for (int j = 0; j < n; ++j) {
        const std::size_t q = static_cast<std::size_t>(j) * W + t;
        X[q] = X0[q] = XH[q];
      }
      for (int i = 0; i < m; ++i) {
        const std::size_t q = static_cast<std::size_t>(i) * W + t;
        Y[q] = Y0[q] = YH[q];
      }
      inner[k] = 0;
      r_last[k] = kInf;
    }
    if (static_cast<int>(keep.size()) != W) {
      for (auto* v : {&C, &L, &U, &X, &X0, &XH, &XB, &ATY, &DX}) compact(*v, n, W, keep);
      for (auto* v : {&RL, &RU, &Y, &Y0, &YH, &YB, &AX, &DY, &ADX}) compact(*v, m, W, keep);
      std::vector<int> ns;
      for (int t : keep) ns.push_back(sid[t]);
      sid.swap(ns);
      W = static_cast<int>(sid.size());
    }
  }
  return out;
}

}  // namespace orc

// ================================================================== extern "C" surface
using namespace orc;

extern "C" {

int orc_solve(const orc_lp* lp, const orc_opts* o, orc_result* r, double* x, double* y, double* z, double* dual_ray,
              double* primal_ray) {
  const Lp l = from_abi(lp);
  SolveOpt so;
  so.presolve = o->presolve != 0;
  so.e.mixed = o->precision != 0;
  if (o->check_every > 0) so.e.K = o->check_every;
  if (o->tolerance > 0) so.e.tol = o->tolerance;
  if (o->time_limit > 0) so.e.time_limit = o->time_limit;
  if (o->iteration_limit > 0) so.e.iter_limit = o->iteration_limit;
  if (o->warm_x) so.e.warm_x.assign(o->warm_x, o->warm_x + lp->n);
  if (o->warm_y) so.e.warm_y.assign(o->warm_y, o->warm_y + lp->m);
  so.e.warm_w = o->warm_primal_weight;
  for (int k = 0; k < o->n_params; ++k) so.params.emplace_back(o->param_names[k], o->param_values[k]);
  Sol s = solve_impl(l, so);
  gate(l, so.e.tol, s);
  std::memset(r, 0, sizeof *r);
  r->status = s.status;
  r->iterations = s.iterations;
  r->iterations_to_fast = s.it_fast;
  r->objective = s.objective;
  r->dual_objective = s.dual_objective;
  r->primal_residual = s.rp;
  r->dual_residual = s.rd;
  r->gap = s.gap;
  r->certified_bound = s.certified;
  r->primal_weight = s.primal_weight;
  r->seconds = s.seconds;
  r->setup_seconds = s.setup_seconds;
  r->check_primal = s.check_p;
  r->check_dual = s.check_d;
  r->check_gap = s.check_g;
  r->check = s.check;
  std::snprintf(r->engine, sizeof r->engine, "%s", s.engine.c_str());
  std::snprintf(r->message, sizeof r->message, "%s", s.message.c_str());
  auto put = [](double* dst, const std::vector<double>& v, int len) {
    if (dst && static_cast<int>(v.size()) == len) std::memcpy(dst, v.data(), sizeof(double) * len);
  };
  put(x, s.x, lp->n);
  put(y, s.y, lp->m);
  put(z, s.z, lp->n);
  put(dual_ray, s.dual_ray, lp->m);
  put(primal_ray, s.primal_ray, lp->n);
  return r->status;
}

// The oracle's presolve result, in ps26119_amd_presolve_model's layout (test comparison).
int orc_presolve(const orc_lp* lp, double* info, int* kept_rows, int* kept_cols, double* rl, double* ru, double* cl,
                 double* cu, int* r_col_start, int* r_row_index, double* r_value) {
  const Lp l = from_abi(lp);
  const Pre r = presolve(l);
  info[0] = r.outcome;
  info[1] = r.rows_removed;
  info[2] = r.cols_removed;
  info[3] = r.red.m;
  info[4] = r.red.n;
  info[5] = static_cast<double>(r.red.v.size());
  info[6] = r.red.offset;
  if (r.outcome != 0) return 0;
  auto put = [](auto* dst, const auto& src) {
    if (dst && !src.empty()) std::copy(src.begin(), src.end(), dst);
  };
  put(kept_rows, r.row_map);
  put(kept_cols, r.col_map);
  put(rl, r.red.rl);
  put(ru, r.red.ru);
  put(cl, r.red.cl);
  put(cu, r.red.cu);
  put(r_col_start, r.red.cs);
  put(r_row_index, r.red.ri);
  put(r_value, r.red.v);
  return 0;
}

int orc_scale(const orc_lp* lp, int geo_iters, double geo_min_range, int ruiz, int pc, int bor, double* row_scale,
              double* col_scale, double* scalars, double* a_vals, double* at_vals, double* c_scaled, double* bounds) {
  const Lp l = from_abi(lp);
  ScaleOpt o;
  o.geo_iters = geo_iters;
  o.geo_min_range = geo_min_range;
  o.ruiz = ruiz;
  o.pc = pc != 0;
  o.bor = bor != 0;
  const Scaled s = scale(l, o);
  std::memcpy(row_scale, s.rs.data(), sizeof(double) * s.m);
  std::memcpy(col_scale, s.cs.data(), sizeof(double) * s.n);
  scalars[0] = s.bs;
  scalars[1] = s.os;
  scalars[2] = s.log10_range;
  scalars[3] = s.geo ? 1.0 : 0.0;
  if (a_vals) std::memcpy(a_vals, s.A.val.data(), sizeof(double) * s.A.val.size());
  if (at_vals) std::memcpy(at_vals, s.At.val.data(), sizeof(double) * s.At.val.size());
  if (c_scaled) std::memcpy(c_scaled, s.c.data(), sizeof(double) * s.n);
  if (bounds) {
    std::memcpy(bounds, s.l.data(), sizeof(double) * s.n);
    std::memcpy(bounds + s.n, s.u.data(), sizeof(double) * s.n);
    std::memcpy(bounds + 2 * s.n, s.rl.data(), sizeof(double) * s.m);
    std::memcpy(bounds + 2 * s.n + s.m, s.ru.data(), sizeof(double) * s.m);
  }
  return 0;
}

double orc_norm2(const orc_lp* lp) {
  const Lp l = from_abi(lp);
  const Scaled s = scale(l, ScaleOpt{});
  return norm2(s.A, s.At);
}

int orc_kkt(const orc_lp* lp, const double* x, const double* ymin, double* out) {
  const Lp l = from_abi(lp);
  Scaled tmp;
  tmp.orig = &l;
  tmp.At0 = csc_as_transpose(l);
  tmp.A0 = transpose(tmp.At0);
  const Kkt k = kkt_original(tmp, std::vector<double>(x, x + l.n), std::vector<double>(ymin, ymin + l.m));
  const double v[10] = {k.rp, k.rd, k.pobj, k.dobj, k.bn, k.cn, k.pmax, k.dmax, k.rel(), k.converged(1e-8) ? 1.0 : 0.0};
  std::memcpy(out, v, sizeof v);
  return 0;
}

int orc_backend_ops(const orc_lp* lp, int precision, int iters, double tau, double sigma, double rho, double* ox,
                    double* oxh, double* oxb, double* oy, double* oyh, double* oyb, double* red) {
  const Lp l = from_abi(lp);
  const Scaled s = scale(l, ScaleOpt{});
  Backend b(s);
  b.set_mixed(precision != 0);
  const int x = b.create(true), x0 = b.create(true), xh = b.create(true), xb = b.create(true), g = b.create(true);
  const int y = b.create(false), y0 = b.create(false), yh = b.create(false), yb = b.create(false), ax = b.create(false);
  std::vector<double> xi(s.n), yi(s.m);
  for (int j = 0; j < s.n; ++j) xi[j] = std::sin(1.0 + j);
  for (int i = 0; i < s.m; ++i) yi[i] = std::cos(2.0 + i);
  b.upload(x, xi);
  b.upload(y, yi);
  b.copy(x0, x);
  b.copy(y0, y);
  for (int it = 0; it < iters; ++it) {
    const double w = (it + 1.0) / (it + 2.0);
    b.spmv_t(y, g);
    b.r2h_primal(x, x0, g, tau, w, rho, xh, xb);
    b.spmv(xb, ax);
    b.r2h_dual(y, y0, ax, sigma, w, rho, yh, yb);
  }
  b.primal_step(x, g, 0.5 * tau, xb);
  b.dual_step(y, ax, 1.25 * sigma, yb);
  b.axpby(0.3, xh, -1.7, xb);
  std::vector<double> v;
  const int ids[6] = {x, xh, xb, y, yh, yb};
  double* outs[6] = {ox, oxh, oxb, oy, oyh, oyb};
  for (int t = 0; t < 6; ++t) {
    b.download(ids[t], v);
    std::memcpy(outs[t], v.data(), sizeof(double) * v.size());
  }
  red[0] = b.dot(x, xh);
  red[1] = std::sqrt(b.dot(y, y));
  red[2] = b.diff_norm2(x, x0);
  const Kkt k = b.kkt(xh, yh);
  const double kv[8] = {k.rp, k.rd, k.pobj, k.dobj, k.bn, k.cn, k.pmax, k.dmax};
  std::memcpy(red + 3, kv, sizeof kv);
  return 0;
}

int orc_solve_batch(const orc_lp* lp, int K, const orc_scenario* sc, const orc_opts* o, orc_result* results, double* xs,
                    double* ys) {
  const Lp base = from_abi(lp);
  std::vector<Scen> S(K);
  auto take = [](const double* p, int len) { return p ? std::vector<double>(p, p + len) : std::vector<double>{}; };
  for (int k = 0; k < K; ++k) {
    S[k].c = take(sc[k].obj, lp->n);
    S[k].cl = take(sc[k].col_lower, lp->n);
    S[k].cu = take(sc[k].col_upper, lp->n);
    S[k].rl = take(sc[k].row_lower, lp->m);
    S[k].ru = take(sc[k].row_upper, lp->m);
  }
  Opt e;
  if (o->check_every > 0) e.K = o->check_every;
  if (o->tolerance > 0) e.tol = o->tolerance;
  if (o->time_limit > 0) e.time_limit = o->time_limit;
  if (o->iteration_limit > 0) e.iter_limit = o->iteration_limit;
  std::vector<std::pair<std::string, double>> ps;
  for (int k = 0; k < o->n_params; ++k) ps.emplace_back(o->param_names[k], o->param_values[k]);
  apply_params(e, ps);
  const std::vector<Sol> out = solve_batch(base, S, e);
  for (int k = 0; k < K; ++k) {
    const Sol& s = out[k];
    orc_result* r = results + k;
    std::memset(r, 0, sizeof *r);
    r->status = s.status;
    r->iterations = s.iterations;
    r->iterations_to_fast = s.it_fast;
    r->objective = s.objective;
    r->dual_objective = s.dual_objective;
    r->primal_residual = s.rp;
    r->dual_residual = s.rd;
    r->gap = s.gap;
    r->certified_bound = s.certified;
    r->primal_weight = s.primal_weight;
    r->seconds = s.seconds;
    r->check = -1;
    std::snprintf(r->message, sizeof r->message, "%s", s.message.c_str());
    if (xs && static_cast<int>(s.x.size()) == lp->n) std::memcpy(xs + static_cast<size_t>(k) * lp->n, s.x.data(), sizeof(double) * lp->n);
    if (ys && static_cast<int>(s.y.size()) == lp->m) std::memcpy(ys + static_cast<size_t>(k) * lp->m, s.y.data(), sizeof(double) * lp->m);
  }
  return 0;
}

double orc_certified_bound(const orc_lp* lp, const double* y, int use_implied) {
  const Lp l = from_abi(lp);
  return certified_bound(l, std::vector<double>(y, y + l.m), use_implied != 0).bound;
}

int orc_check_solution(const orc_lp* lp, const double* x, const double* y, double* out) {
  const Lp l = from_abi(lp);
  const Check c = check_solution(l, std::vector<double>(x, x + l.n), std::vector<double>(y, y + l.m));
  out[0] = c.pviol;
  out[1] = c.dviol;
  out[2] = c.gap;
  out[3] = c.pobj;
  out[4] = c.dobj;
  out[5] = c.passed() ? 1.0 : 0.0;
  return 0;
}

}  // extern "C"
