// psh_core.hip — device context, problem upload, vector slots, the reference Backend operations
// (src/pdhg/backend.h:46-65 of the reference), deterministic reductions and the Lanczos pieces.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <type_traits>

#include "psh_internal.h"

using psh::kBlock;

namespace psh {

int fail(psh_ctx* c, int code, const std::string& what) {
  if (c) c->err = what;
  return code;
}

int hip_fail(psh_ctx* c, hipError_t e, const char* where) {
  const int code = (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) ? PSH_E_OOM : PSH_E_HIP;
  return fail(c, code, std::string(where) + ": " + hipGetErrorString(e));
}

std::recursive_mutex& api_mutex() {
  static std::recursive_mutex m;
  return m;
}

int dev_alloc(psh_ctx* c, void** p, size_t bytes) {
  std::lock_guard<std::recursive_mutex> lock(api_mutex());
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    *p = nullptr;
    return fail(c, PSH_E_OOM, "out of memory: device allocation of " + std::to_string(bytes) + " bytes failed");
  }
  c->bytes += bytes;
  return PSH_OK;
}

void dev_free(void* p) {
  if (!p) return;
  std::lock_guard<std::recursive_mutex> lock(api_mutex());
  (void)hipFree(p);
}

// Greedy row tiles: ≤ kBlock rows and ≤ kTileNnz nonzeros, or one row that alone exceeds kTileNnz.
std::vector<int> build_tiles(const int* ptr, int rows) {
  std::vector<int> t;
  t.reserve(rows / 64 + 2);
  t.push_back(0);
  int r = 0;
  while (r < rows) {
    long long nz = 0;
    int cnt = 0;
    while (r < rows && cnt < kBlock) {
      const long long len = ptr[r + 1] - ptr[r];
      if (cnt > 0 && nz + len > kTileNnz) break;
      nz += len;
      ++cnt;
      ++r;
      if (nz > kTileNnz) break;
    }
    t.push_back(r);
  }
  return t;
}

bool is_primal_slot(int s) { return s >= PSH_X && s <= PSH_ATY; }
long long slot_len(const psh_ctx* c, int s) { return is_primal_slot(s) ? c->n : c->m; }

// Event pairs are recorded around timed launches and settled lazily (only pairs the GPU has
// already passed are read), so timing never makes the host wait for the device.
static void settle_done(psh_ctx* c) {
  std::size_t k = 0;
  for (; k < c->pending.size(); ++k) {
    auto& p = c->pending[k];
    if (hipEventQuery(p.b) != hipSuccess) break;
    float ms = 0;
    (void)hipEventElapsedTime(&ms, p.a, p.b);
    c->t_ms[p.cls] += ms;
    c->t_launch[p.cls] += 1;
    c->ev_pool.push_back(p.a);
    c->ev_pool.push_back(p.b);
  }
  if (k > 0) c->pending.erase(c->pending.begin(), c->pending.begin() + static_cast<std::ptrdiff_t>(k));
}

static hipEvent_t take_event(psh_ctx* c) {
  if (c->ev_pool.empty()) {
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
  }
  hipEvent_t e = c->ev_pool.back();
  c->ev_pool.pop_back();
  return e;
}

void time_begin(psh_ctx* c, int, hipEvent_t* a) {
  *a = nullptr;
  if (!c->timing) return;
  settle_done(c);
  *a = take_event(c);
  if (*a) (void)hipEventRecord(*a, c->stream);
}

void time_end(psh_ctx* c, int cls, hipEvent_t a) {
  if (!a) return;
  hipEvent_t b = take_event(c);
  if (!b) {
    c->ev_pool.push_back(a);
    return;
  }
  (void)hipEventRecord(b, c->stream);
  c->pending.push_back({cls, a, b});
}

// ------------------------------------------------------------------ reductions
__global__ void __launch_bounds__(kBlock) k_tree1(const RedJob* __restrict__ jobs, double* __restrict__ partials) {
  const RedJob J = jobs[blockIdx.y];
  double* out = partials + static_cast<size_t>(blockIdx.y) * kTreeBlocks + blockIdx.x;
  if (J.op == 0 && J.mode != 0) return;
  const long long N = J.na + J.nb;
  const long long lo = N * blockIdx.x / kTreeBlocks, hi = N * (blockIdx.x + 1) / kTreeBlocks;
  double acc = 0.0;
  for (long long i = lo + threadIdx.x; i < hi; i += kBlock) {
    const double v = i < J.na ? J.a[i] : J.b[i - J.na];
    if (J.op == 0) acc += v;
    else acc = fmax(acc, v);
  }
  const double r = J.op == 0 ? block_tree_sum(acc) : block_tree_max(acc);
  if (threadIdx.x == 0) *out = r;
}

__global__ void __launch_bounds__(kBlock) k_tree2(const RedJob* __restrict__ jobs, const double* __restrict__ partials,
                                                  double* __restrict__ results, int first) {
  const RedJob J = jobs[blockIdx.x];
  if (J.op == 0 && J.mode != 0) return;
  const double* p = partials + static_cast<size_t>(blockIdx.x) * kTreeBlocks;
  double acc = 0.0;
  for (int i = threadIdx.x; i < kTreeBlocks; i += kBlock) acc = J.op == 0 ? acc + p[i] : fmax(acc, p[i]);
  const double r = J.op == 0 ? block_tree_sum(acc) : block_tree_max(acc);
  if (threadIdx.x == 0) results[first + blockIdx.x] = r;
}

// CPU-order sums: mode 1 = 64 fixed chunks (la/parallel.h:169-188), mode 2 = one sequential loop
// over a then b (cpu_backend.cpp:240-274, csr.cpp:117-123). Grid (64, jobs), one wavefront per
// (chunk, job); chunk sums go to partials[job][chunk] and k_ordered_final adds them in order.
__global__ void __launch_bounds__(64) k_ordered(const RedJob* __restrict__ jobs, double* __restrict__ partials) {
  const RedJob J = jobs[blockIdx.y];
  if (J.op != 0 || J.mode == 0) return;
  __shared__ double stage[kSeqBatch];
  ordered_chunk<false>(J, partials + static_cast<size_t>(blockIdx.y) * kTreeBlocks, static_cast<int>(blockIdx.x), stage);
}

__global__ void __launch_bounds__(64) k_ordered_final(const RedJob* __restrict__ jobs,
                                                      const double* __restrict__ partials, double* __restrict__ results,
                                                      int first) {
  const RedJob J = jobs[blockIdx.x];
  if (J.op != 0 || J.mode == 0 || threadIdx.x != 0) return;
  const double* p = partials + static_cast<size_t>(blockIdx.x) * kTreeBlocks;
  if (J.mode == 2) {
    results[first + blockIdx.x] = p[0];
    return;
  }
  const long long n = J.na;
  const int chunks = static_cast<int>(n < 1 ? 1 : (n > 64 ? 64 : n));
  double s = 0.0;
  for (int k = 0; k < chunks; ++k) s += p[k];
  results[first + blockIdx.x] = s;
}

// parts: 1 = tree jobs, 2 = ordered jobs (chunk sums and their combine), 4 = only the combine
static int launch_reduce(psh_ctx* c, const RedJob* dev, const RedJob* host, int njobs, int first, int parts = 3) {
  bool tree = false, ordered = false;
  for (int k = 0; k < njobs; ++k) {
    if (host[k].op == 1 || host[k].mode == 0) tree = true;
    else ordered = true;
  }
  if (tree && (parts & 1)) {
    hipLaunchKernelGGL(k_tree1, dim3(kTreeBlocks, njobs), dim3(kBlock), 0, c->stream, dev, c->partials);
    hipLaunchKernelGGL(k_tree2, dim3(njobs), dim3(kBlock), 0, c->stream, dev, c->partials, c->results, first);
  }
  // ordered jobs use their own rows of `partials` (a job is either a tree job or an ordered one)
  if (ordered && (parts & 2))
    hipLaunchKernelGGL(k_ordered, dim3(64, njobs), dim3(64), 0, c->stream, dev, c->partials);
  if (ordered && (parts & 6))
    hipLaunchKernelGGL(k_ordered_final, dim3(njobs), dim3(64), 0, c->stream, dev, c->partials, c->results, first);
  PSH_LAUNCHED(c, "reduction kernels");
  return PSH_OK;
}

int reduce_table_part(psh_ctx* c, const JobTable& t, int first, int parts) {
  if (t.host.empty()) return PSH_OK;
  return launch_reduce(c, t.dev, t.host.data(), static_cast<int>(t.host.size()), first, parts);
}

ChunkSums chunk_sums_of(const psh_ctx* c, const JobTable& t) {
  ChunkSums cs;
  cs.jobs = t.dev;
  cs.partials = c->partials;
  int count = 0;
  for (std::size_t k = 0; k < t.host.size(); ++k) {
    const RedJob& J = t.host[k];
    if (J.op != 0 || J.mode == 0) continue;
    if (J.mode != 1 || k >= 32) {
      cs.blocks = -1;
      return cs;
    }
    cs.mask |= 1u << k;
    ++count;
  }
  cs.blocks = count * 64 / kChainsPerBlock;
  return cs;
}

int reduce(psh_ctx* c, const RedJob* jobs, int njobs, int first) {
  if (njobs <= 0) return PSH_OK;
  if (njobs > kMaxJobs) return fail(c, PSH_E_ARG, "too many reductions in one call");
  std::memcpy(c->jobs_host, jobs, sizeof(RedJob) * njobs);
  PSH_TRY(c, hipMemcpyAsync(c->jobs_dev, c->jobs_host, sizeof(RedJob) * njobs, hipMemcpyHostToDevice, c->stream));
  int rc = launch_reduce(c, c->jobs_dev, jobs, njobs, first);
  if (rc) return rc;
  // The pinned job buffer is reused by the next call: make sure this copy has been consumed.
  PSH_TRY(c, hipStreamSynchronize(c->stream));
  return PSH_OK;
}

int reduce_table(psh_ctx* c, const JobTable& t, int first) {
  if (t.host.empty()) return PSH_OK;
  return launch_reduce(c, t.dev, t.host.data(), static_cast<int>(t.host.size()), first);
}

// Sum jobs of a check pass. CPU-order mode follows the reference: 64-chunk sums for the
// residual dots and the PID distances (parallel_sum), single loops for the KKT sums with the
// row part followed by the column part in one accumulator (cpu_backend.cpp:240-274). Decision
// mode keeps the 64-chunk order for the sums the restart and primal-weight decisions read and
// takes the KKT sums as trees. The ray sums are trees in every mode.
const JobTable* check_table(psh_ctx* c, int mode, unsigned flags, int* err) {
  const int key = (mode << 8) | static_cast<int>(flags);
  auto it = c->tables.find(key);
  if (it != c->tables.end()) return &it->second;
  JobTable t;
  const long long m = c->m, n = c->n, L = c->term_len;
  auto T = [&](int k) { return c->terms + static_cast<size_t>(k) * L; };
  const bool cpu = mode == PSH_RED_CPU;
  const int chunk = mode != PSH_RED_TREE ? 1 : 0, seq = cpu ? 2 : 0;
  // Every slot R_PX..R_CMAG2 gets a job (unused ones sum nothing) so results keep their index.
  auto job = [&](int k, long long na, int md, int k2 = -1, long long nb = 0) {
    t.host.push_back(RedJob{T(k), k2 >= 0 ? T(k2) : nullptr, na, nb, 0, md});
  };
  const bool res = flags & F_RES, kkt = flags & F_KKT, ray = flags & F_RAY, dr = flags & F_DRIFT, pid = flags & F_PID;
  job(T_DX2, res ? n : 0, chunk);
  job(T_DY2, res ? m : 0, chunk);
  job(T_CROSS, res ? m : 0, chunk);
  job(T_RP2, kkt ? m : 0, seq);
  job(T_B2, kkt ? m : 0, seq);
  job(T_DROW, kkt ? m : 0, seq, T_DCOL, kkt ? n : 0);
  job(T_RD2ROW, kkt ? m : 0, seq, T_RD2COL, kkt ? n : 0);
  job(T_C2, kkt ? n : 0, seq);
  job(T_P, kkt ? n : 0, seq);
  job(T_PIDX, pid ? n : 0, chunk);
  job(T_PIDY, pid ? m : 0, chunk);
  job(T_ROBJ_ROW, ray ? m : 0, 0, T_ROBJ_COL, ray ? n : 0);
  job(T_RMAG_ROW, ray ? m : 0, 0, T_RMAG_COL, ray ? n : 0);
  job(T_CD, ray ? n : 0, 0);
  job(T_CMAG, ray ? n : 0, 0);
  job(T_ROBJ_ROW2, dr ? m : 0, 0, T_ROBJ_COL2, dr ? n : 0);
  job(T_RMAG_ROW2, dr ? m : 0, 0, T_RMAG_COL2, dr ? n : 0);
  job(T_CD2, dr ? n : 0, 0);
  job(T_CMAG2, dr ? n : 0, 0);
  if (alloc_n(c, &t.dev, static_cast<long long>(t.host.size()))) {
    *err = PSH_E_OOM;
    return nullptr;
  }
  hipError_t e = hipMemcpyAsync(t.dev, t.host.data(), sizeof(RedJob) * t.host.size(), hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);  // t.host moves into the map below
  if (e != hipSuccess) {
    *err = hip_fail(c, e, "reduction table upload");
    return nullptr;
  }
  return &c->tables.emplace(key, std::move(t)).first->second;
}

int fetch_results(psh_ctx* c, int count, double* out) {
  PSH_TRY(c, hipMemcpyAsync(c->host_results, c->results, sizeof(double) * count, hipMemcpyDeviceToHost, c->stream));
  PSH_TRY(c, hipStreamSynchronize(c->stream));
  std::memcpy(out, c->host_results, sizeof(double) * count);
  return PSH_OK;
}

// ------------------------------------------------------------------ elementwise kernels
template <class S, class D>
__global__ void k_convert(const S* __restrict__ a, D* __restrict__ b, long long n) {
  const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) b[i] = static_cast<D>(a[i]);
}

// fp64 → fp32 (round to nearest) → bf16 (round to nearest, ties to even); values are finite.
__global__ void k_to_bf16(const double* __restrict__ a, Bf16* __restrict__ b, long long n) {
  const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const unsigned u = __float_as_uint(static_cast<float>(a[i]));
  b[i].bits = static_cast<unsigned short>((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}

__global__ void k_gather_vals(const double* __restrict__ a, const int* __restrict__ src, double* __restrict__ b,
                              long long n) {
  const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) b[i] = a[src[i]];
}

// ---- scaling sweeps on the device (reference src/pdhg/scaling.cpp:37-142; the host restatement
// csrc/host/scaling.cpp does the same operations in the same order, so the factors and values
// are the same bits). One thread per row of Ã or of Ãᵀ (a column of A).
// Geometric-mean factor of a row from |a|·f over its nonzeros: 1/(√lo·√hi), clamped.
__global__ void k_sc_geo(const int* __restrict__ ptr, const int* __restrict__ col, const double* __restrict__ a0,
                         const double* __restrict__ f, double* __restrict__ out, int rows) {
  const int i = static_cast<int>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= rows) return;
  double lo = HUGE_VAL, hi = 0.0;
  for (int k = ptr[i]; k < ptr[i + 1]; ++k) {
    const double v = fabs(a0[k]) * f[col[k]];
    if (v > 0 && isfinite(v)) {
      lo = v < lo ? v : lo;
      hi = hi < v ? v : hi;
    }
  }
  if (hi > 0) {
    const double g = 1.0 / (sqrt(lo) * sqrt(hi));
    const double ge = g < 1e-20 ? 1e-20 : g;
    out[i] = 1e20 < ge ? 1e20 : ge;
  }
}

// Ruiz (∞-norm) or Pock–Chambolle (1-norm) factor of a row: 1/√‖row‖, 1 for an empty or non-finite row.
__global__ void k_sc_norm(const int* __restrict__ ptr, const double* __restrict__ val, int inf_norm,
                          double* __restrict__ out, int rows) {
  const int i = static_cast<int>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= rows) return;
  double s = 0.0;
  for (int k = ptr[i]; k < ptr[i + 1]; ++k) {
    const double a = fabs(val[k]);
    s = inf_norm ? (s < a ? a : s) : s + a;
  }
  out[i] = s > 0.0 && isfinite(s) ? 1.0 / sqrt(s) : 1.0;
}

// val ← val·(frow_i·fcol_j) for every nonzero (row i, column j) of the CSR.
__global__ void k_sc_apply(const int* __restrict__ ptr, const int* __restrict__ col, double* __restrict__ val,
                           const double* __restrict__ frow, const double* __restrict__ fcol, int rows) {
  const int i = static_cast<int>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= rows) return;
  const double fr = frow[i];
  for (int k = ptr[i]; k < ptr[i + 1]; ++k) val[k] *= fr * fcol[col[k]];
}

__global__ void k_sc_mul(double* __restrict__ v, const double* __restrict__ f, long long n) {
  const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) v[i] *= f[i];
}

template <class T>
__global__ void k_fill(T* __restrict__ v, T a, long long n) {
  const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) v[i] = a;
}

template <class T>
__global__ void k_axpby(T a, const T* __restrict__ x, T b, T* __restrict__ y, long long n) {
  const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) y[i] = a * x[i] + b * y[i];
}

// Terms of a dot product / squared distance, converted to fp64 first (cpu_backend.cpp:92-111).
template <class T>
__global__ void k_dot_terms(const T* __restrict__ a, const T* __restrict__ b, double* __restrict__ t, long long n,
                            int diff) {
  const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double u = static_cast<double>(a[i]), v = static_cast<double>(b[i]);
  if (diff) {
    const double d = u - v;
    t[i] = d * d;
  } else {
    t[i] = u * v;
  }
}

template <class T>
__global__ void k_primal_step(const T* __restrict__ x, const T* __restrict__ g, const T* __restrict__ c,
                              const T* __restrict__ l, const T* __restrict__ u, T tau, T* __restrict__ h, long long n) {
  const long long j = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (j < n) h[j] = clampv<T>(x[j] - tau * (c[j] - g[j]), l[j], u[j]);
}

template <class T>
__global__ void k_dual_step(const T* __restrict__ y, const T* __restrict__ ax, const T* __restrict__ rl,
                            const T* __restrict__ ru, T sg, T* __restrict__ h, long long m) {
  const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const T q = ax[i] - y[i] / sg;
  h[i] = y[i] - sg * ax[i] + sg * clampv<T>(q, rl[i], ru[i]);
}

template <class T>
__global__ void k_r2h_primal_ew(T* __restrict__ x, const T* __restrict__ x0, const T* __restrict__ g,
                                const T* __restrict__ c, const T* __restrict__ l, const T* __restrict__ u, T tau, T tw,
                                T a, T b, T* __restrict__ xh, T* __restrict__ xb, long long n) {
  const long long j = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const T xj = x[j];
  const T h = clampv<T>(xj - tau * (c[j] - g[j]), l[j], u[j]);
  xh[j] = h;
  xb[j] = 2 * h - xj;
  x[j] = tw * (a * h + b * xj) + (1 - tw) * x0[j];
}

template <class T>
__global__ void k_r2h_dual_ew(T* __restrict__ y, const T* __restrict__ y0, const T* __restrict__ ax,
                              const T* __restrict__ rl, const T* __restrict__ ru, T sg, T tw, T a, T b,
                              T* __restrict__ yh, T* __restrict__ yb, long long m) {
  const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const T yi = y[i];
  const T q = ax[i] - yi / sg;
  const T h = yi - sg * ax[i] + sg * clampv<T>(q, rl[i], ru[i]);
  yh[i] = h;
  if (yb) yb[i] = 2 * h - yi;
  y[i] = tw * (a * h + b * yi) + (1 - tw) * y0[i];
}

// Plain tiled SpMV: out = M·x, accumulated in fp64 row-sequentially, stored in T.
template <class V, class T>
struct EpiStore {
  T* out;
  __device__ void operator()(int row, bool has, double s) {
    if (has) out[row] = static_cast<T>(s);
  }
};

template <class V, class T>
__global__ void __launch_bounds__(kBlock) k_spmv(const int* __restrict__ tiles, int ntiles, const int* __restrict__ ptr,
                                                 const int* __restrict__ col, const V* __restrict__ val,
                                                 const T* __restrict__ x, T* __restrict__ out) {
  EpiStore<V, T> e{out};
  tiled_rows<V, T>(tiles, ntiles, ptr, col, val, x, e);
}

// Lanczos update: w ← w − (α v + β_prev v_prev), term = w² (csr.cpp:121-125).
__global__ void k_lanczos_update(double* __restrict__ w, const double* __restrict__ v, const double* __restrict__ vp,
                                 double al, double bp, double* __restrict__ t, long long n) {
  const long long j = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const double nw = w[j] - (al * v[j] + bp * vp[j]);
  w[j] = nw;
  t[j] = nw * nw;
}

__global__ void k_lanczos_advance(double* __restrict__ v, double* __restrict__ vp, const double* __restrict__ w,
                                  double bn, long long n) {
  const long long j = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (j >= n) return;
  vp[j] = v[j];
  v[j] = w[j] / bn;
}

}  // namespace psh

using namespace psh;

namespace {

template <class F>
int with_t(psh_ctx* c, F&& f) {
  if (c->fp32) return f(static_cast<float*>(nullptr));
  return f(static_cast<double*>(nullptr));
}

template <class T>
T* slot_ptr(psh_ctx* c, int s) {
  if constexpr (sizeof(T) == 8) return c->vd[s];
  else return c->vf[s];
}

int check_slot(psh_ctx* c, int s) {
  if (!c->loaded) return fail(c, PSH_E_STATE, "no problem loaded");
  if (s < 0 || s >= PSH_NSLOT) return fail(c, PSH_E_ARG, "bad vector slot " + std::to_string(s));
  return PSH_OK;
}

}  // namespace

// fp32 copies of Ã's and Ãᵀ's values (shared by fp32 mode and the increment form).
int psh::ensure_fp32_matrix(psh_ctx* c) {
  if (c->A.valf) return PSH_OK;
  int rc;
  if ((rc = alloc_n(c, &c->A.valf, c->nnz)) || (rc = alloc_n(c, &c->At.valf, c->nnz))) return rc;
  if (c->nnz > 0) {
    hipLaunchKernelGGL((k_convert<double, float>), dim3(grid_for(c->nnz)), dim3(kBlock), 0, c->stream, c->A.val,
                       c->A.valf, c->nnz);
    hipLaunchKernelGGL((k_convert<double, float>), dim3(grid_for(c->nnz)), dim3(kBlock), 0, c->stream, c->At.val,
                       c->At.valf, c->nnz);
  }
  PSH_LAUNCHED(c, "fp32 matrix");
  return PSH_OK;
}

// bf16 copies of Ã's and Ãᵀ's values (the increment form's bf16 variant).
int psh::ensure_bf16_matrix(psh_ctx* c) {
  if (c->A.valh) return PSH_OK;
  int rc;
  if ((rc = alloc_n(c, &c->A.valh, c->nnz)) || (rc = alloc_n(c, &c->At.valh, c->nnz))) return rc;
  if (c->nnz > 0) {
    hipLaunchKernelGGL(k_to_bf16, dim3(grid_for(c->nnz)), dim3(kBlock), 0, c->stream, c->A.val, c->A.valh, c->nnz);
    hipLaunchKernelGGL(k_to_bf16, dim3(grid_for(c->nnz)), dim3(kBlock), 0, c->stream, c->At.val, c->At.valh, c->nnz);
  }
  PSH_LAUNCHED(c, "bf16 matrix");
  return PSH_OK;
}

namespace {

int ensure_fp32(psh_ctx* c) {
  if (c->vf[0]) return PSH_OK;
  int rc;
  for (int s = 0; s < PSH_NSLOT; ++s)
    if ((rc = alloc_n(c, &c->vf[s], slot_len(c, s)))) return rc;
  if ((rc = ensure_fp32_matrix(c)) || (rc = alloc_n(c, &c->cf, c->n)) ||
      (rc = alloc_n(c, &c->lf, c->n)) || (rc = alloc_n(c, &c->uf, c->n)) || (rc = alloc_n(c, &c->rlf, c->m)) ||
      (rc = alloc_n(c, &c->ruf, c->m)))
    return rc;
  auto conv = [&](const double* a, float* b, long long n) {
    if (n > 0) hipLaunchKernelGGL((k_convert<double, float>), dim3(grid_for(n)), dim3(kBlock), 0, c->stream, a, b, n);
  };
  conv(c->c, c->cf, c->n);
  conv(c->l, c->lf, c->n);
  conv(c->u, c->uf, c->n);
  conv(c->rl, c->rlf, c->m);
  conv(c->ru, c->ruf, c->m);
  PSH_LAUNCHED(c, "fp32 copies");
  return PSH_OK;
}

template <class T>
struct Vecs {
  const T *c, *l, *u, *rl, *ru;
};
template <class T>
Vecs<T> vecs(psh_ctx* c) {
  if constexpr (sizeof(T) == 8) return {c->c, c->l, c->u, c->rl, c->ru};
  else return {c->cf, c->lf, c->uf, c->rlf, c->ruf};
}
template <class T>
const T* mat_val(const DevMat& M) {
  if constexpr (sizeof(T) == 8) return M.val;
  else return M.valf;
}

int spmv_impl(psh_ctx* c, const DevMat& M, int xs, int out) {
  return with_t(c, [&](auto* tag) -> int {
    using T = std::remove_pointer_t<decltype(tag)>;
    if (M.ntiles > 0)
      hipLaunchKernelGGL((k_spmv<T, T>), dim3(grid_for_tiles(M.ntiles)), dim3(kBlock), 0, c->stream, M.tiles, M.ntiles,
                         M.ptr, M.col, mat_val<T>(M), slot_ptr<T>(c, xs), slot_ptr<T>(c, out));
    PSH_LAUNCHED(c, "spmv");
    return PSH_OK;
  });
}

}  // namespace

namespace {

// Distinct 128-byte lines of the gathered vector per entry, sampled over every 8th block of 256
// rows (the row tiles' size): ≈ 1 when a tile's gathers fall at random over a large vector, a few
// hundredths for banded or staircase structure (multi-period planning models), whose tiles keep
// reusing the same lines from L2.
double gather_lines_per_entry(const int* ptr, const int* col, int rows) {
  long long lines = 0, entries = 0;
  std::vector<int> v;
  for (int b0 = 0; b0 < rows; b0 += 8 * 256) {
    const int b1 = std::min(rows, b0 + 256);
    v.assign(col + ptr[b0], col + ptr[b1]);
    for (int& x : v) x >>= 4;
    std::sort(v.begin(), v.end());
    lines += std::unique(v.begin(), v.end()) - v.begin();
    entries += ptr[b1] - ptr[b0];
  }
  return entries > 0 ? static_cast<double>(lines) / static_cast<double>(entries) : 0.0;
}

// The phased layout of a CSR (struct Phased), built when the gathered vector (cols fp64 values)
// is larger than an XCD's L2 share (kPhMinBytes) and the gathers are scattered (at least
// kPhMinLinesPerEntry distinct lines per entry), or when `force` is set: S phases (slices of
// about 2 MiB, at least 8), K rows per thread so that the workgroups stay resident together
// (about 1024 of 512 threads), J product slots per thread for the largest (workgroup, phase)
// block. Left unbuilt (the row tiles serve) when a block or a row's entries in one phase do not
// fit (more than kPhMaxJ·kPhBlock entries, or 255 in one row). `orig_dev` (may be null): the
// original values in the same CSR order, copied into the phased order for the check pass.
int build_phased(psh_ctx* c, Phased& P, const int* ptr, const int* col, const double* vals_dev,
                 const double* orig_dev, int rows, int cols, bool force) {
  const long long nz = rows > 0 ? ptr[rows] : 0;
  const long long bytes = static_cast<long long>(cols) * 8;
  if (nz == 0 || rows == 0 || cols == 0 || (!force && bytes <= kPhMinBytes)) return PSH_OK;
  if (!force && gather_lines_per_entry(ptr, col, rows) < kPhMinLinesPerEntry) return PSH_OK;
  for (int r = 0; r < rows; ++r)
    for (int q = ptr[r] + 1; q < ptr[r + 1]; ++q)
      if (col[q] < col[q - 1]) return PSH_OK;  // a row with unsorted columns: row tiles only
  const int S = bytes <= (16LL << 20) ? 8 : static_cast<int>(std::min<long long>(32, (bytes + (2LL << 20) - 1) >> 21));
  const long long per_k = 1024LL * kPhBlock;
  const int K = rows <= per_k ? 1 : (rows <= 2 * per_k ? 2 : 4);
  const int R = K * kPhBlock;
  const int G = (rows + R - 1) / R;
  const long long width = (static_cast<long long>(cols) + S - 1) / S;
  std::vector<int> blk(static_cast<size_t>(G) * S + 1, 0), pcol, psrc;
  std::vector<unsigned char> cnt(static_cast<size_t>(G) * S * R, 0);
  pcol.reserve(static_cast<size_t>(nz));
  psrc.reserve(static_cast<size_t>(nz));
  std::vector<int> cur(ptr, ptr + rows);
  int maxblk = 0;
  for (int g = 0; g < G; ++g) {
    const int r0 = g * R, r1 = std::min(rows, r0 + R);
    for (int p = 0; p < S; ++p) {
      const size_t bp = static_cast<size_t>(g) * S + p;
      const int start = static_cast<int>(pcol.size());
      for (int r = r0; r < r1; ++r) {
        int k = 0;
        for (int q = cur[r]; q < ptr[r + 1] && col[q] / width == p; ++q, ++k) {
          pcol.push_back(col[q]);
          psrc.push_back(q);
        }
        cur[r] += k;
        if (k > 255) return PSH_OK;
        cnt[bp * R + (r - r0)] = static_cast<unsigned char>(k);
      }
      blk[bp + 1] = static_cast<int>(pcol.size());
      maxblk = std::max(maxblk, blk[bp + 1] - start);
      if (maxblk > kPhMaxJ * kPhBlock) return PSH_OK;
    }
  }
  if (static_cast<long long>(pcol.size()) != nz) return fail(c, PSH_E_ARG, "phased layout does not cover the matrix");
  const int J = maxblk <= 2 * kPhBlock ? 2 : (maxblk <= 4 * kPhBlock ? 4 : 8);
  int rc;
  auto put = [&](auto** dst, const auto* src, long long count) -> int {
    using T = std::remove_const_t<std::remove_pointer_t<decltype(src)>>;
    int r = alloc_n(c, dst, count);
    if (r) return r;
    PSH_TRY(c, hipMemcpyAsync(*dst, src, sizeof(T) * count, hipMemcpyHostToDevice, c->stream));
    return PSH_OK;
  };
  if ((rc = put(&P.blk, blk.data(), static_cast<long long>(blk.size()))) ||
      (rc = put(&P.cnt, cnt.data(), static_cast<long long>(cnt.size()))) || (rc = put(&P.col, pcol.data(), nz)) ||
      (rc = put(&P.src, psrc.data(), nz)) || (rc = alloc_n(c, &P.val, nz)) ||
      (orig_dev && (rc = alloc_n(c, &P.valo, nz))))
    return rc;
  hipLaunchKernelGGL(k_gather_vals, dim3(grid_for(nz)), dim3(kBlock), 0, c->stream, vals_dev, P.src, P.val, nz);
  if (orig_dev)
    hipLaunchKernelGGL(k_gather_vals, dim3(grid_for(nz)), dim3(kBlock), 0, c->stream, orig_dev, P.src, P.valo, nz);
  PSH_LAUNCHED(c, "phased values");
  PSH_TRY(c, hipStreamSynchronize(c->stream));  // the staging vectors go out of scope
  P.S = S;
  P.K = K;
  P.J = J;
  P.G = G;
  P.R = R;
  return PSH_OK;
}

// The phased copies take the current (scaled) values of Ã and Ãᵀ.
int regather_phased(psh_ctx* c) {
  if (c->Pa.S > 0)
    hipLaunchKernelGGL(k_gather_vals, dim3(grid_for(c->nnz)), dim3(kBlock), 0, c->stream, c->A.val, c->Pa.src, c->Pa.val,
                       c->nnz);
  if (c->Pat.S > 0)
    hipLaunchKernelGGL(k_gather_vals, dim3(grid_for(c->nnz)), dim3(kBlock), 0, c->stream, c->At.val, c->Pat.src,
                       c->Pat.val, c->nnz);
  PSH_LAUNCHED(c, "phased values");
  return PSH_OK;
}

}  // namespace

extern "C" {

int psh_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  return n;
}

int psh_mem_info(int device, size_t* free_bytes, size_t* total_bytes) {
  if (hipSetDevice(device) != hipSuccess) return PSH_E_NO_DEVICE;
  return hipMemGetInfo(free_bytes, total_bytes) == hipSuccess ? PSH_OK : PSH_E_HIP;
}

int psh_create(int device, psh_ctx** out) {
  if (!out) return PSH_E_ARG;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
    (void)hipGetLastError();
    return PSH_E_NO_DEVICE;
  }
  if (device < 0 || device >= n) return PSH_E_NO_DEVICE;
  if (hipSetDevice(device) != hipSuccess) return PSH_E_NO_DEVICE;
  auto* c = new psh_ctx;
  c->device = device;
  std::unique_lock<std::recursive_mutex> lock(api_mutex());
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return PSH_E_HIP;
  }
  const bool pinned_ok = hipHostMalloc(reinterpret_cast<void**>(&c->host_results), sizeof(double) * 128) == hipSuccess &&
                         hipHostMalloc(reinterpret_cast<void**>(&c->jobs_host), sizeof(RedJob) * kMaxJobs) == hipSuccess;
  lock.unlock();
  if (!pinned_ok) {
    psh_destroy(c);
    return PSH_E_OOM;
  }
  int rc;
  if ((rc = alloc_n(c, &c->results, 128)) || (rc = alloc_n(c, &c->jobs_dev, kMaxJobs)) ||
      (rc = alloc_n(c, &c->partials, static_cast<long long>(kTreeBlocks) * kMaxJobs)) || (rc = alloc_n(c, &c->r0_dev, 4)) ||
      (rc = alloc_n(c, &c->prm, 4))) {
    psh_destroy(c);
    return rc;
  }
  *out = c;
  return PSH_OK;
}

void psh_destroy(psh_ctx* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  psh::destroy_batch(c);
  for (auto& kv : c->graph_cache) (void)hipGraphExecDestroy(kv.second);
  {
    DeltaState& d = c->delta;
    void* dp[] = {d.dx, d.dxh, d.dxb, d.G, d.LX, d.UX, d.dy, d.dyh, d.dyb, d.BL, d.BU, d.KY};
    for (void* p : dp) dev_free(p);
  }
  void* ptrs[] = {c->A.ptr, c->A.col, c->A.val, c->A.valf, c->A.tiles, c->At.ptr, c->At.col, c->At.val, c->At.valf,
                  c->At.tiles, c->A.valh, c->At.valh, c->a_orig, c->ato_alias ? nullptr : c->ato_col, c->ato_val, c->c, c->l, c->u, c->rl,
                  c->ru, c->cf, c->lf, c->uf, c->rlf, c->ruf, c->rs, c->cs, c->clo, c->cuo, c->rlo, c->ruo, c->cmin,
                  c->xo, c->yo, c->rv, c->dv, c->rv2, c->dv2, c->terms, c->partials, c->results, c->jobs_dev, c->lv,
                  c->lvp, c->lw, c->lt, c->r0_dev, c->prm, c->mpart_m, c->mpart_n, c->Pa.blk, c->Pa.cnt,
                  c->Pa.col, c->Pa.src, c->Pa.val, c->Pa.valo, c->Pat.blk, c->Pat.cnt, c->Pat.col, c->Pat.src,
                  c->Pat.val, c->Pat.valo};
  for (void* p : ptrs) dev_free(p);
  for (int s = 0; s < PSH_NSLOT; ++s) dev_free(c->vd[s]), dev_free(c->vf[s]);
  for (auto& kv : c->tables) dev_free(kv.second.dev);
  {
    std::lock_guard<std::recursive_mutex> lock(api_mutex());
    if (c->host_results) (void)hipHostFree(c->host_results);
    if (c->jobs_host) (void)hipHostFree(c->jobs_host);
  }
  for (auto& p : c->pending) c->ev_pool.push_back(p.a), c->ev_pool.push_back(p.b);
  for (hipEvent_t e : c->ev_pool) (void)hipEventDestroy(e);
  if (c->stream) {
    std::lock_guard<std::recursive_mutex> lock(api_mutex());
    (void)hipStreamDestroy(c->stream);
  }
  delete c;
}

const char* psh_error(const psh_ctx* c) { return c ? c->err.c_str() : "no context"; }

int psh_sync(psh_ctx* c) {
  if (!c) return PSH_E_ARG;
  PSH_TRY(c, hipStreamSynchronize(c->stream));
  return PSH_OK;
}

size_t psh_device_bytes(const psh_ctx* c) { return c ? c->bytes : 0; }

int psh_load(psh_ctx* c, const psh_problem* p) {
  if (!c || !p) return PSH_E_ARG;
  if (c->loaded) return fail(c, PSH_E_STATE, "problem already loaded on this context");
  if (p->m < 0 || p->n < 0 || p->nnz < 0 || p->nnz >= (1LL << 31)) return fail(c, PSH_E_ARG, "bad problem sizes");
  PSH_TRY(c, hipSetDevice(c->device));
  c->m = p->m;
  c->n = p->n;
  c->nnz = p->nnz;
  c->bs = p->bound_scale;
  c->os = p->obj_scale;
  const int m = p->m, n = p->n;
  const long long nz = p->nnz;
  if (p->a_ptr[m] != nz || p->at_ptr[n] != nz) return fail(c, PSH_E_ARG, "row pointers do not end at nnz");
  int rc;
  auto up = [&](auto** dst, const auto* src, long long count) -> int {
    using T = std::remove_const_t<std::remove_pointer_t<decltype(src)>>;
    int r = alloc_n(c, dst, count);
    if (r) return r;
    if (count > 0) PSH_TRY(c, hipMemcpyAsync(*dst, src, sizeof(T) * count, hipMemcpyHostToDevice, c->stream));
    return PSH_OK;
  };
  c->A.rows = m;
  c->At.rows = n;
  if ((rc = up(&c->A.ptr, p->a_ptr, m + 1LL)) || (rc = up(&c->A.col, p->a_col, nz)) || (rc = up(&c->A.val, p->a_val, nz)) ||
      (rc = up(&c->a_orig, p->a_val_orig, nz)) || (rc = up(&c->At.ptr, p->at_ptr, n + 1LL)) ||
      (rc = up(&c->At.col, p->at_col, nz)) || (rc = up(&c->At.val, p->at_val, nz)) || (rc = up(&c->ato_val, p->ato_val, nz)))
    return rc;
  c->ato_alias = p->ato_col == p->at_col || (nz > 0 && std::memcmp(p->ato_col, p->at_col, sizeof(int) * nz) == 0);
  if (c->ato_alias) c->ato_col = c->At.col;
  else if ((rc = up(&c->ato_col, p->ato_col, nz))) return rc;
  if ((rc = up(&c->c, p->c, n)) || (rc = up(&c->l, p->l, n)) || (rc = up(&c->u, p->u, n)) || (rc = up(&c->rl, p->rl, m)) ||
      (rc = up(&c->ru, p->ru, m)) || (rc = up(&c->rs, p->row_scale, m)) || (rc = up(&c->cs, p->col_scale, n)) ||
      (rc = up(&c->clo, p->cl_o, n)) || (rc = up(&c->cuo, p->cu_o, n)) || (rc = up(&c->rlo, p->rl_o, m)) ||
      (rc = up(&c->ruo, p->ru_o, m)) || (rc = up(&c->cmin, p->c_min, n)))
    return rc;
  c->boxed = true;
  for (int j = 0; j < n && c->boxed; ++j) c->boxed = std::isfinite(p->cl_o[j]) && std::isfinite(p->cu_o[j]);
  const std::vector<int> ta = build_tiles(p->a_ptr, m), tt = build_tiles(p->at_ptr, n);
  c->A.ntiles = static_cast<int>(ta.size()) - 1;
  c->At.ntiles = static_cast<int>(tt.size()) - 1;
  if ((rc = up(&c->A.tiles, ta.data(), static_cast<long long>(ta.size()))) ||
      (rc = up(&c->At.tiles, tt.data(), static_cast<long long>(tt.size()))))
    return rc;
  // The check pass sums the original Aᵀ in the model's CSC order, which is Ãᵀ's only when the
  // column indices agree; otherwise its column pass keeps the row tiles.
  if (c->sliced &&
      ((rc = build_phased(c, c->Pa, p->a_ptr, p->a_col, c->A.val, c->a_orig, m, n, c->sliced == 2)) ||
       (rc = build_phased(c, c->Pat, p->at_ptr, p->at_col, c->At.val, c->ato_alias ? c->ato_val : nullptr, n, m,
                          c->sliced == 2))))
    return rc;
  for (int s = 0; s < PSH_NSLOT; ++s)
    if ((rc = alloc_n(c, &c->vd[s], slot_len(c, s)))) return rc;
  const long long mx = std::max<long long>(std::max(m, n), 1);
  c->term_len = mx;
  if ((rc = alloc_n(c, &c->xo, n)) || (rc = alloc_n(c, &c->yo, m)) || (rc = alloc_n(c, &c->rv, m)) ||
      (rc = alloc_n(c, &c->dv, n)) || (rc = alloc_n(c, &c->rv2, m)) || (rc = alloc_n(c, &c->dv2, n)) ||
      (rc = alloc_n(c, &c->terms, mx * kTerms)))
    return rc;
  // per-workgroup maxima: the elementwise passes run grid_for(rows) workgroups, the row / column
  // passes one per tile (padded to the XCD count)
  c->mpart_rows_m = std::max(m > 0 ? grid_for(m) : 0, c->A.ntiles > 0 ? grid_for_tiles(c->A.ntiles) : 0);
  c->mpart_rows_n = std::max(n > 0 ? grid_for(n) : 0, c->At.ntiles > 0 ? grid_for_tiles(c->At.ntiles) : 0);
  {
    const long long pm = static_cast<long long>(std::max(c->mpart_rows_m, 1)) * kMaxCount;
    const long long pn = static_cast<long long>(std::max(c->mpart_rows_n, 1)) * kMaxCount;
    if ((rc = alloc_n(c, &c->mpart_m, pm)) || (rc = alloc_n(c, &c->mpart_n, pn))) return rc;
    hipLaunchKernelGGL((k_fill<double>), dim3(grid_for(pm)), dim3(kBlock), 0, c->stream, c->mpart_m, 0.0, pm);
    hipLaunchKernelGGL((k_fill<double>), dim3(grid_for(pn)), dim3(kBlock), 0, c->stream, c->mpart_n, 0.0, pn);
  }
  PSH_TRY(c, hipStreamSynchronize(c->stream));  // host arrays may be freed after return
  for (int s = 0; s < PSH_NSLOT; ++s)
    if (slot_len(c, s) > 0)
      hipLaunchKernelGGL((k_fill<double>), dim3(grid_for(slot_len(c, s))), dim3(kBlock), 0, c->stream, c->vd[s], 0.0,
                         slot_len(c, s));
  PSH_LAUNCHED(c, "slot init");
  c->loaded = true;
  return PSH_OK;
}

int psh_scale(psh_ctx* c, int geo_iterations, int ruiz_iterations, int pock_chambolle, double* row_scale,
              double* col_scale) {
  if (!c || !c->loaded) return c ? fail(c, PSH_E_STATE, "no problem loaded") : PSH_E_ARG;
  if (c->scaled) return fail(c, PSH_E_STATE, "the problem is already scaled");
  if (c->delta.on || c->fp32 || c->A.valf || c->A.valh) return fail(c, PSH_E_STATE, "scale before any precision change");
  if (geo_iterations < 0 || ruiz_iterations < 0) return fail(c, PSH_E_ARG, "negative sweep count");
  PSH_TRY(c, hipSetDevice(c->device));
  const int m = c->m, n = c->n;
  double *r = nullptr, *f = nullptr;
  int rc;
  if ((rc = alloc_n(c, &r, std::max(m, 1))) || (rc = alloc_n(c, &f, std::max(n, 1)))) {
    dev_free(r);
    return rc;
  }
  const int gm = grid_for(std::max(m, 1)), gn = grid_for(std::max(n, 1));
  auto ones = [&](double* v, long long len) {
    if (len > 0) hipLaunchKernelGGL((k_fill<double>), dim3(grid_for(len)), dim3(kBlock), 0, c->stream, v, 1.0, len);
  };
  ones(c->rs, m);
  ones(c->cs, n);
  if (geo_iterations > 0) {
    // factors from the ORIGINAL values every sweep (Ã's order: a_orig; Ãᵀ's order: At.val, not
    // yet modified), rows then columns, applied once at the end; the factors are the scales
    ones(r, m);
    ones(f, n);
    for (int it = 0; it < geo_iterations; ++it) {
      if (m > 0) hipLaunchKernelGGL(k_sc_geo, dim3(gm), dim3(kBlock), 0, c->stream, c->A.ptr, c->A.col, c->a_orig, f, r, m);
      if (n > 0) hipLaunchKernelGGL(k_sc_geo, dim3(gn), dim3(kBlock), 0, c->stream, c->At.ptr, c->At.col, c->At.val, r, f, n);
    }
    if (m > 0) hipLaunchKernelGGL(k_sc_apply, dim3(gm), dim3(kBlock), 0, c->stream, c->A.ptr, c->A.col, c->A.val, r, f, m);
    if (n > 0) hipLaunchKernelGGL(k_sc_apply, dim3(gn), dim3(kBlock), 0, c->stream, c->At.ptr, c->At.col, c->At.val, f, r, n);
    if (m > 0) hipLaunchKernelGGL(k_sc_mul, dim3(gm), dim3(kBlock), 0, c->stream, c->rs, r, static_cast<long long>(m));
    if (n > 0) hipLaunchKernelGGL(k_sc_mul, dim3(gn), dim3(kBlock), 0, c->stream, c->cs, f, static_cast<long long>(n));
  }
  auto sweep = [&](int inf_norm) {
    if (m > 0) hipLaunchKernelGGL(k_sc_norm, dim3(gm), dim3(kBlock), 0, c->stream, c->A.ptr, c->A.val, inf_norm, r, m);
    if (n > 0) hipLaunchKernelGGL(k_sc_norm, dim3(gn), dim3(kBlock), 0, c->stream, c->At.ptr, c->At.val, inf_norm, f, n);
    if (m > 0) hipLaunchKernelGGL(k_sc_apply, dim3(gm), dim3(kBlock), 0, c->stream, c->A.ptr, c->A.col, c->A.val, r, f, m);
    if (n > 0) hipLaunchKernelGGL(k_sc_apply, dim3(gn), dim3(kBlock), 0, c->stream, c->At.ptr, c->At.col, c->At.val, f, r, n);
    if (m > 0) hipLaunchKernelGGL(k_sc_mul, dim3(gm), dim3(kBlock), 0, c->stream, c->rs, r, static_cast<long long>(m));
    if (n > 0) hipLaunchKernelGGL(k_sc_mul, dim3(gn), dim3(kBlock), 0, c->stream, c->cs, f, static_cast<long long>(n));
  };
  for (int it = 0; it < ruiz_iterations; ++it) sweep(1);
  if (pock_chambolle) sweep(0);
  PSH_LAUNCHED(c, "scaling sweeps");
  if ((rc = regather_phased(c))) return rc;
  if (row_scale && m > 0) PSH_TRY(c, hipMemcpyAsync(row_scale, c->rs, sizeof(double) * m, hipMemcpyDeviceToHost, c->stream));
  if (col_scale && n > 0) PSH_TRY(c, hipMemcpyAsync(col_scale, c->cs, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
  PSH_TRY(c, hipStreamSynchronize(c->stream));
  dev_free(r);
  dev_free(f);
  c->scaled = true;
  return PSH_OK;
}

int psh_set_vectors(psh_ctx* c, const double* cc, const double* l, const double* u, const double* rl, const double* ru,
                    double bound_scale, double obj_scale) {
  if (!c || !c->loaded) return c ? fail(c, PSH_E_STATE, "no problem loaded") : PSH_E_ARG;
  if (c->fp32) return fail(c, PSH_E_STATE, "set the vectors before any precision change");
  const size_t n = static_cast<size_t>(c->n), m = static_cast<size_t>(c->m);
  auto put = [&](double* dst, const double* src, size_t len) -> int {
    if (len == 0) return PSH_OK;
    if (!src) return fail(c, PSH_E_ARG, "null vector");
    PSH_TRY(c, hipMemcpyAsync(dst, src, sizeof(double) * len, hipMemcpyHostToDevice, c->stream));
    return PSH_OK;
  };
  int rc;
  if ((rc = put(c->c, cc, n)) || (rc = put(c->l, l, n)) || (rc = put(c->u, u, n)) || (rc = put(c->rl, rl, m)) ||
      (rc = put(c->ru, ru, m)))
    return rc;
  c->bs = bound_scale;
  c->os = obj_scale;
  PSH_TRY(c, hipStreamSynchronize(c->stream));  // the host arrays may go away
  return PSH_OK;
}

int psh_set_fp32(psh_ctx* c, int on) {
  if (!c || !c->loaded) return c ? fail(c, PSH_E_STATE, "no problem loaded") : PSH_E_ARG;
  const bool want = on != 0;
  if (want && c->delta.on) return fail(c, PSH_E_STATE, "fp32 storage and the increment form are exclusive");
  if (want == c->fp32) return PSH_OK;
  int rc;
  if (want && (rc = ensure_fp32(c))) return rc;
  for (int s = 0; s < PSH_NSLOT; ++s) {
    const long long len = slot_len(c, s);
    if (len <= 0) continue;
    if (want) hipLaunchKernelGGL((k_convert<double, float>), dim3(grid_for(len)), dim3(kBlock), 0, c->stream, c->vd[s], c->vf[s], len);
    else hipLaunchKernelGGL((k_convert<float, double>), dim3(grid_for(len)), dim3(kBlock), 0, c->stream, c->vf[s], c->vd[s], len);
  }
  PSH_LAUNCHED(c, "precision switch");
  c->fp32 = want;
  return PSH_OK;
}

int psh_is_fp32(const psh_ctx* c) { return c && c->fp32 ? 1 : 0; }

int psh_upload(psh_ctx* c, int s, const double* host) {
  if (!c) return PSH_E_ARG;
  int rc = check_slot(c, s);
  if (rc) return rc;
  const long long len = slot_len(c, s);
  if (len == 0) return PSH_OK;
  PSH_TRY(c, hipMemcpyAsync(c->vd[s], host, sizeof(double) * len, hipMemcpyHostToDevice, c->stream));
  if (c->fp32) hipLaunchKernelGGL((k_convert<double, float>), dim3(grid_for(len)), dim3(kBlock), 0, c->stream, c->vd[s], c->vf[s], len);
  PSH_LAUNCHED(c, "upload");
  PSH_TRY(c, hipStreamSynchronize(c->stream));
  return PSH_OK;
}

int psh_download(psh_ctx* c, int s, double* host) {
  if (!c) return PSH_E_ARG;
  int rc = check_slot(c, s);
  if (rc) return rc;
  const long long len = slot_len(c, s);
  if (len == 0) return PSH_OK;
  const double* src = c->vd[s];
  if (c->fp32) {
    hipLaunchKernelGGL((k_convert<float, double>), dim3(grid_for(len)), dim3(kBlock), 0, c->stream, c->vf[s], c->terms, len);
    PSH_LAUNCHED(c, "download convert");
    src = c->terms;
  }
  PSH_TRY(c, hipMemcpyAsync(host, src, sizeof(double) * len, hipMemcpyDeviceToHost, c->stream));
  PSH_TRY(c, hipStreamSynchronize(c->stream));
  return PSH_OK;
}

int psh_fill(psh_ctx* c, int s, double a) {
  if (!c) return PSH_E_ARG;
  int rc = check_slot(c, s);
  if (rc) return rc;
  const long long len = slot_len(c, s);
  if (len == 0) return PSH_OK;
  return with_t(c, [&](auto* tag) -> int {
    using T = std::remove_pointer_t<decltype(tag)>;
    hipLaunchKernelGGL((k_fill<T>), dim3(grid_for(len)), dim3(kBlock), 0, c->stream, slot_ptr<T>(c, s), static_cast<T>(a), len);
    PSH_LAUNCHED(c, "fill");
    return PSH_OK;
  });
}

int psh_copy(psh_ctx* c, int dst, int src) {
  if (!c) return PSH_E_ARG;
  int rc;
  if ((rc = check_slot(c, dst)) || (rc = check_slot(c, src))) return rc;
  if (is_primal_slot(dst) != is_primal_slot(src)) return fail(c, PSH_E_ARG, "copy between primal and dual slots");
  const long long len = slot_len(c, dst);
  if (len == 0 || dst == src) return PSH_OK;
  return with_t(c, [&](auto* tag) -> int {
    using T = std::remove_pointer_t<decltype(tag)>;
    PSH_TRY(c, hipMemcpyAsync(slot_ptr<T>(c, dst), slot_ptr<T>(c, src), sizeof(T) * len, hipMemcpyDeviceToDevice, c->stream));
    return PSH_OK;
  });
}

int psh_axpby(psh_ctx* c, double a, int x, double b, int y) {
  if (!c) return PSH_E_ARG;
  int rc;
  if ((rc = check_slot(c, x)) || (rc = check_slot(c, y))) return rc;
  if (is_primal_slot(x) != is_primal_slot(y)) return fail(c, PSH_E_ARG, "axpby between primal and dual slots");
  const long long len = slot_len(c, y);
  if (len == 0) return PSH_OK;
  return with_t(c, [&](auto* tag) -> int {
    using T = std::remove_pointer_t<decltype(tag)>;
    hipLaunchKernelGGL((k_axpby<T>), dim3(grid_for(len)), dim3(kBlock), 0, c->stream, static_cast<T>(a), slot_ptr<T>(c, x),
                       static_cast<T>(b), slot_ptr<T>(c, y), len);
    PSH_LAUNCHED(c, "axpby");
    return PSH_OK;
  });
}

int psh_spmv(psh_ctx* c, int x, int out) {
  if (!c) return PSH_E_ARG;
  int rc;
  if ((rc = check_slot(c, x)) || (rc = check_slot(c, out))) return rc;
  if (!is_primal_slot(x) || is_primal_slot(out)) return fail(c, PSH_E_ARG, "spmv needs a primal input and a dual output");
  return spmv_impl(c, c->A, x, out);
}

int psh_spmv_t(psh_ctx* c, int y, int out) {
  if (!c) return PSH_E_ARG;
  int rc;
  if ((rc = check_slot(c, y)) || (rc = check_slot(c, out))) return rc;
  if (is_primal_slot(y) || !is_primal_slot(out)) return fail(c, PSH_E_ARG, "spmv_t needs a dual input and a primal output");
  return spmv_impl(c, c->At, y, out);
}

static int dot_like(psh_ctx* c, int a, int b, int mode, int diff, double* out) {
  if (!c || !out) return PSH_E_ARG;
  int rc;
  if ((rc = check_slot(c, a)) || (rc = check_slot(c, b))) return rc;
  if (is_primal_slot(a) != is_primal_slot(b)) return fail(c, PSH_E_ARG, "reduction over slots of different spaces");
  const long long len = slot_len(c, a);
  if (len == 0) {
    *out = 0.0;
    return PSH_OK;
  }
  rc = with_t(c, [&](auto* tag) -> int {
    using T = std::remove_pointer_t<decltype(tag)>;
    hipLaunchKernelGGL((k_dot_terms<T>), dim3(grid_for(len)), dim3(kBlock), 0, c->stream, slot_ptr<T>(c, a),
                       slot_ptr<T>(c, b), c->terms, len, diff);
    PSH_LAUNCHED(c, "dot terms");
    return PSH_OK;
  });
  if (rc) return rc;
  RedJob j{c->terms, nullptr, len, 0, 0, mode != PSH_RED_TREE ? 1 : 0};
  if ((rc = reduce(c, &j, 1, 0))) return rc;
  return fetch_results(c, 1, out);
}

int psh_dot(psh_ctx* c, int a, int b, int mode, double* out) { return dot_like(c, a, b, mode, 0, out); }

int psh_diff_norm2(psh_ctx* c, int a, int b, int mode, double* out) {
  int rc = dot_like(c, a, b, mode, 1, out);
  if (rc == PSH_OK) *out = std::sqrt(*out);
  return rc;
}

int psh_primal_step(psh_ctx* c, int x, int aty, double tau, int xh) {
  if (!c) return PSH_E_ARG;
  int rc;
  if ((rc = check_slot(c, x)) || (rc = check_slot(c, aty)) || (rc = check_slot(c, xh))) return rc;
  if (c->n == 0) return PSH_OK;
  return with_t(c, [&](auto* tag) -> int {
    using T = std::remove_pointer_t<decltype(tag)>;
    const Vecs<T> v = vecs<T>(c);
    hipLaunchKernelGGL((k_primal_step<T>), dim3(grid_for(c->n)), dim3(kBlock), 0, c->stream, slot_ptr<T>(c, x),
                       slot_ptr<T>(c, aty), v.c, v.l, v.u, static_cast<T>(tau), slot_ptr<T>(c, xh), (long long)c->n);
    PSH_LAUNCHED(c, "primal_step");
    return PSH_OK;
  });
}

int psh_dual_step(psh_ctx* c, int y, int ax, double sigma, int yh) {
  if (!c) return PSH_E_ARG;
  int rc;
  if ((rc = check_slot(c, y)) || (rc = check_slot(c, ax)) || (rc = check_slot(c, yh))) return rc;
  if (c->m == 0) return PSH_OK;
  return with_t(c, [&](auto* tag) -> int {
    using T = std::remove_pointer_t<decltype(tag)>;
    const Vecs<T> v = vecs<T>(c);
    hipLaunchKernelGGL((k_dual_step<T>), dim3(grid_for(c->m)), dim3(kBlock), 0, c->stream, slot_ptr<T>(c, y),
                       slot_ptr<T>(c, ax), v.rl, v.ru, static_cast<T>(sigma), slot_ptr<T>(c, yh), (long long)c->m);
    PSH_LAUNCHED(c, "dual_step");
    return PSH_OK;
  });
}

int psh_r2h_primal(psh_ctx* c, int x, int x0, int aty, double tau, double w, double rho, int xh, int xb) {
  if (!c) return PSH_E_ARG;
  int rc;
  for (int s : {x, x0, aty, xh, xb})
    if ((rc = check_slot(c, s))) return rc;
  if (c->n == 0) return PSH_OK;
  return with_t(c, [&](auto* tag) -> int {
    using T = std::remove_pointer_t<decltype(tag)>;
    const Vecs<T> v = vecs<T>(c);
    hipLaunchKernelGGL((k_r2h_primal_ew<T>), dim3(grid_for(c->n)), dim3(kBlock), 0, c->stream, slot_ptr<T>(c, x),
                       slot_ptr<T>(c, x0), slot_ptr<T>(c, aty), v.c, v.l, v.u, static_cast<T>(tau), static_cast<T>(w),
                       static_cast<T>(2 * rho), static_cast<T>(1 - 2 * rho), slot_ptr<T>(c, xh), slot_ptr<T>(c, xb),
                       (long long)c->n);
    PSH_LAUNCHED(c, "r2h_primal");
    return PSH_OK;
  });
}

int psh_r2h_dual(psh_ctx* c, int y, int y0, int ax, double sigma, double w, double rho, int yh, int yb) {
  if (!c) return PSH_E_ARG;
  int rc;
  for (int s : {y, y0, ax, yh})
    if ((rc = check_slot(c, s))) return rc;
  if (yb >= 0 && (rc = check_slot(c, yb))) return rc;
  if (c->m == 0) return PSH_OK;
  return with_t(c, [&](auto* tag) -> int {
    using T = std::remove_pointer_t<decltype(tag)>;
    const Vecs<T> v = vecs<T>(c);
    hipLaunchKernelGGL((k_r2h_dual_ew<T>), dim3(grid_for(c->m)), dim3(kBlock), 0, c->stream, slot_ptr<T>(c, y),
                       slot_ptr<T>(c, y0), slot_ptr<T>(c, ax), v.rl, v.ru, static_cast<T>(sigma), static_cast<T>(w),
                       static_cast<T>(2 * rho), static_cast<T>(1 - 2 * rho), slot_ptr<T>(c, yh),
                       yb >= 0 ? slot_ptr<T>(c, yb) : static_cast<T*>(nullptr), (long long)c->m);
    PSH_LAUNCHED(c, "r2h_dual");
    return PSH_OK;
  });
}

int psh_lanczos_begin(psh_ctx* c, const double* v0) {
  if (!c || !v0) return PSH_E_ARG;
  if (!c->loaded) return fail(c, PSH_E_STATE, "no problem loaded");
  int rc;
  if (!c->lv && ((rc = alloc_n(c, &c->lv, c->n)) || (rc = alloc_n(c, &c->lvp, c->n)) || (rc = alloc_n(c, &c->lw, c->n)) ||
                 (rc = alloc_n(c, &c->lt, c->m))))
    return rc;
  if (c->n == 0) return PSH_OK;
  PSH_TRY(c, hipMemcpyAsync(c->lv, v0, sizeof(double) * c->n, hipMemcpyHostToDevice, c->stream));
  hipLaunchKernelGGL((k_fill<double>), dim3(grid_for(c->n)), dim3(kBlock), 0, c->stream, c->lvp, 0.0, (long long)c->n);
  PSH_LAUNCHED(c, "lanczos init");
  PSH_TRY(c, hipStreamSynchronize(c->stream));
  return PSH_OK;
}

int psh_lanczos_step(psh_ctx* c, double beta_prev, int mode, double* alpha, double* beta_sq) {
  if (!c || !alpha || !beta_sq) return PSH_E_ARG;
  if (!c->lv) return fail(c, PSH_E_STATE, "psh_lanczos_begin not called");
  const long long n = c->n;
  if (c->A.ntiles > 0)
    hipLaunchKernelGGL((k_spmv<double, double>), dim3(grid_for_tiles(c->A.ntiles)), dim3(kBlock), 0, c->stream,
                       c->A.tiles, c->A.ntiles, c->A.ptr, c->A.col, c->A.val, c->lv, c->lt);
  if (c->At.ntiles > 0)
    hipLaunchKernelGGL((k_spmv<double, double>), dim3(grid_for_tiles(c->At.ntiles)), dim3(kBlock), 0, c->stream,
                       c->At.tiles, c->At.ntiles, c->At.ptr, c->At.col, c->At.val, c->lt, c->lw);
  hipLaunchKernelGGL((k_dot_terms<double>), dim3(grid_for(n)), dim3(kBlock), 0, c->stream, c->lw, c->lv, c->terms, n, 0);
  PSH_LAUNCHED(c, "lanczos products");
  // α = wᵀv as one sequential loop in CPU and decision modes (csr.cpp:117), tree otherwise
  RedJob j{c->terms, nullptr, n, 0, 0, mode != PSH_RED_TREE ? 2 : 0};
  int rc;
  if ((rc = reduce(c, &j, 1, 0)) || (rc = fetch_results(c, 1, alpha))) return rc;
  hipLaunchKernelGGL(k_lanczos_update, dim3(grid_for(n)), dim3(kBlock), 0, c->stream, c->lw, c->lv, c->lvp, *alpha,
                     beta_prev, c->terms, n);
  PSH_LAUNCHED(c, "lanczos update");
  if ((rc = reduce(c, &j, 1, 0)) || (rc = fetch_results(c, 1, beta_sq))) return rc;
  return PSH_OK;
}

int psh_lanczos_product(psh_ctx* c, const double* v, double* w) {
  if (!c || !v || !w) return PSH_E_ARG;
  if (!c->loaded) return fail(c, PSH_E_STATE, "no problem loaded");
  int rc;
  if (!c->lv && ((rc = alloc_n(c, &c->lv, c->n)) || (rc = alloc_n(c, &c->lvp, c->n)) || (rc = alloc_n(c, &c->lw, c->n)) ||
                 (rc = alloc_n(c, &c->lt, c->m))))
    return rc;
  if (c->n == 0) return PSH_OK;
  PSH_TRY(c, hipMemcpyAsync(c->lv, v, sizeof(double) * c->n, hipMemcpyHostToDevice, c->stream));
  if (c->m > 0 && c->A.ntiles > 0)
    hipLaunchKernelGGL((k_spmv<double, double>), dim3(grid_for_tiles(c->A.ntiles)), dim3(kBlock), 0, c->stream,
                       c->A.tiles, c->A.ntiles, c->A.ptr, c->A.col, c->A.val, c->lv, c->lt);
  else if (c->m > 0)
    hipLaunchKernelGGL((k_fill<double>), dim3(grid_for(c->m)), dim3(kBlock), 0, c->stream, c->lt, 0.0, (long long)c->m);
  if (c->At.ntiles > 0)
    hipLaunchKernelGGL((k_spmv<double, double>), dim3(grid_for_tiles(c->At.ntiles)), dim3(kBlock), 0, c->stream,
                       c->At.tiles, c->At.ntiles, c->At.ptr, c->At.col, c->At.val, c->lt, c->lw);
  else
    hipLaunchKernelGGL((k_fill<double>), dim3(grid_for(c->n)), dim3(kBlock), 0, c->stream, c->lw, 0.0, (long long)c->n);
  PSH_LAUNCHED(c, "lanczos products");
  PSH_TRY(c, hipMemcpyAsync(w, c->lw, sizeof(double) * c->n, hipMemcpyDeviceToHost, c->stream));
  PSH_TRY(c, hipStreamSynchronize(c->stream));
  return PSH_OK;
}

int psh_lanczos_advance(psh_ctx* c, double beta) {
  if (!c || !c->lv) return PSH_E_ARG;
  if (c->n == 0) return PSH_OK;
  hipLaunchKernelGGL(k_lanczos_advance, dim3(grid_for(c->n)), dim3(kBlock), 0, c->stream, c->lv, c->lvp, c->lw, beta,
                     (long long)c->n);
  PSH_LAUNCHED(c, "lanczos advance");
  return PSH_OK;
}

int psh_timing(psh_ctx* c, int on) {
  if (!c) return PSH_E_ARG;
  c->timing = on != 0;
  for (int k = 0; k < 3; ++k) c->t_launch[k] = 0, c->t_ms[k] = 0;
  return PSH_OK;
}

int psh_timing_get(psh_ctx* c, int cls, long long* launches, double* total_ms) {
  if (!c || cls < 0 || cls > 2) return PSH_E_ARG;
  PSH_TRY(c, hipStreamSynchronize(c->stream));
  for (auto& p : c->pending) {
    float ms = 0;
    (void)hipEventElapsedTime(&ms, p.a, p.b);
    c->t_ms[p.cls] += ms;
    c->t_launch[p.cls] += 1;
    c->ev_pool.push_back(p.a);
    c->ev_pool.push_back(p.b);
  }
  c->pending.clear();
  if (launches) *launches = c->t_launch[cls];
  if (total_ms) *total_ms = c->t_ms[cls];
  return PSH_OK;
}

}  // extern "C"
