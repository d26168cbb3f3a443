// la.cpp — see la.h.
#include "la.h"

#include <atomic>
#include <cstring>

namespace ps26119::amd {

namespace {
std::atomic<int> g_threads{0};
}

int host_threads() {
  int t = g_threads.load();
  if (t <= 0) t = static_cast<int>(std::max(1u, std::thread::hardware_concurrency()));
  return std::min(t, 64);
}

void set_host_threads(int t) { g_threads.store(t < 0 ? 0 : t); }

void parallel_for(std::int64_t n, const std::function<void(std::int64_t, std::int64_t)>& f, std::int64_t min_piece) {
  if (n <= 0) return;
  const int t = static_cast<int>(std::min<std::int64_t>(host_threads(), (n + min_piece - 1) / min_piece));
  if (t <= 1) {
    f(0, n);
    return;
  }
  std::vector<std::thread> pool;
  pool.reserve(t - 1);
  for (int k = 1; k < t; ++k) pool.emplace_back([&, k] { f(n * k / t, n * (k + 1) / t); });
  f(0, n / t);
  for (auto& th : pool) th.join();
}

Csr csc_as_transpose(const Model& m) {
  Csr t;
  t.rows = m.num_cols;
  t.cols = m.num_rows;
  t.ptr = m.col_start;
  t.col = m.row_index;
  t.val = m.value;
  return t;
}

Csr transpose(const Csr& a) {
  Csr t;
  t.rows = a.cols;
  t.cols = a.rows;
  const std::int64_t nz = a.nnz();
  t.ptr.assign(static_cast<std::size_t>(t.rows) + 1, 0);
  for (std::int64_t k = 0; k < nz; ++k) ++t.ptr[a.col[k] + 1];
  for (int r = 0; r < t.rows; ++r) t.ptr[r + 1] += t.ptr[r];
  t.col.resize(nz);
  t.val.resize(nz);
  std::vector<int> next(t.ptr.begin(), t.ptr.end() - 1);
  for (int r = 0; r < a.rows; ++r)
    for (int k = a.ptr[r]; k < a.ptr[r + 1]; ++k) {
      const int q = next[a.col[k]]++;
      t.col[q] = r;
      t.val[q] = a.val[k];
    }
  return t;
}

void multiply(const Csr& a, const double* x, double* out) {
  parallel_for(a.rows, [&](std::int64_t b, std::int64_t e) {
    for (std::int64_t i = b; i < e; ++i) {
      double s = 0;
      for (int k = a.ptr[i]; k < a.ptr[i + 1]; ++k) s += a.val[k] * x[a.col[k]];
      out[i] = s;
    }
  });
}

}  // namespace ps26119::amd
