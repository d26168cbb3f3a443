// la.cpp — see la.h.
#include "la.h"

#include <unistd.h>

#include <atomic>
#include <condition_variable>
#include <cstring>
#include <mutex>
#include <thread>

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

namespace {

// Parked workers shared by every host loop of the process, started on first use and never
// stopped (the pool outlives static destruction on purpose). One loop owns the pool at a time;
// a loop that finds it owned (by another thread, or by its own caller when nested) or that runs
// in a forked child (whose copy of the pool has no threads) does its pieces on the calling
// thread, which gives the same results because every index is still visited exactly once.
class WorkerPool {
 public:
  static WorkerPool& get() {
    static WorkerPool* pool = new WorkerPool();
    return *pool;
  }

  bool try_run(int pieces, std::int64_t n, const std::function<void(std::int64_t, std::int64_t)>& f) {
    if (::getpid() != pid_) return false;
    std::unique_lock<std::mutex> own(owner_, std::try_to_lock);
    if (!own.owns_lock()) return false;
    {
      std::lock_guard<std::mutex> g(m_);
      while (started_ < pieces - 1) {
        std::thread([this] { work(); }).detach();
        ++started_;
      }
      job_ = &f;
      n_ = n;
      pieces_ = pieces;
      next_ = 1;  // piece 0 is the caller's
      left_ = pieces - 1;
    }
    wake_.notify_all();
    f(0, n / pieces);
    std::unique_lock<std::mutex> g(m_);
    done_.wait(g, [this] { return left_ == 0; });
    job_ = nullptr;
    return true;
  }

 private:
  WorkerPool() : pid_(::getpid()) {}

  void work() {
    std::unique_lock<std::mutex> g(m_);
    for (;;) {
      wake_.wait(g, [this] { return job_ != nullptr && next_ < pieces_; });
      const int k = next_++;
      const auto* f = job_;
      const std::int64_t n = n_;
      const int p = pieces_;
      g.unlock();
      (*f)(n * k / p, n * (k + 1) / p);
      g.lock();
      if (--left_ == 0) done_.notify_one();
    }
  }

  const pid_t pid_;
  std::mutex owner_;  // held by the loop that uses the pool
  std::mutex m_;      // guards everything below
  std::condition_variable wake_, done_;
  int started_ = 0;
  const std::function<void(std::int64_t, std::int64_t)>* job_ = nullptr;
  std::int64_t n_ = 0;
  int pieces_ = 0, next_ = 0, left_ = 0;
};

}  // namespace

void parallel_for(std::int64_t n, const std::function<void(std::int64_t, std::int64_t)>& f, std::int64_t min_piece) {
  if (n <= 0) return;
  const int t = static_cast<int>(std::min<std::int64_t>(host_threads(), (n + min_piece - 1) / min_piece));
  if (t <= 1 || !WorkerPool::get().try_run(t, n, f)) f(0, n);
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

void column_products(const Model& m, const double* y, double* out) {
  parallel_for(m.num_cols, [&](std::int64_t b, std::int64_t e) {
    for (std::int64_t j = b; j < e; ++j) {
      double s = 0;
      for (int k = m.col_start[j]; k < m.col_start[j + 1]; ++k) s += m.value[k] * y[m.row_index[k]];
      out[j] = s;
    }
  });
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
