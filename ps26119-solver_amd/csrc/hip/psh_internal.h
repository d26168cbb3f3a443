// psh_internal.h — device-side state and helpers shared by the HIP translation units of
// libps26119_hip.so. Not part of the public C-ABI (include/ps26119_hip.h).
//
// Design (MI355X / gfx950):
//  * Sparse products are row-tiled: a tile is ≤ 256 consecutive rows holding ≤ kTileNnz
//    nonzeros (or a single longer row). One 256-thread workgroup per tile stages the tile's
//    products val·x[col] in LDS with coalesced loads of col/val and parallel gathers, then each
//    thread sums ITS row sequentially in CSR order — the reference CPU order (csr.h:36-39), so
//    fp64 results are bit-identical to the CPU backend — and runs the fused epilogue.
//  * Workgroup → tile mapping is XCD-aware: consecutive tiles go to the same XCD (dispatch is
//    round-robin over the 8 XCDs), so gathers of neighbouring tiles share that XCD's L2.
//  * All arithmetic on the iterate follows the reference expressions operation for operation,
//    compiled with -ffp-contract=off (no FMA contraction), so it matches the CPU bits.
#pragma once

#include <type_traits>
#include <utility>

#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "ps26119_hip.h"

namespace psh {

constexpr int kBlock = 256;      // threads per workgroup (4 wave64)
constexpr int kTileNnz = 2048;   // products staged in LDS per tile pass (16 KiB of fp64)
constexpr int kXcd = 8;          // XCDs per MI355X
constexpr int kTreeBlocks = 512; // fixed first-level block count of tree reductions

// A bf16 matrix value: the top 16 bits of an fp32, rounded to nearest even from the fp64 value's
// fp32 rounding. Products with it are formed in fp64 like every other value type.
struct Bf16 {
  unsigned short bits;
  __device__ explicit operator double() const {
    return static_cast<double>(__uint_as_float(static_cast<unsigned>(bits) << 16));
  }
};

struct DevMat {
  int rows = 0;
  int* ptr = nullptr;     // rows+1
  int* col = nullptr;     // nnz
  double* val = nullptr;  // nnz, fp64
  float* valf = nullptr;  // nnz, fp32 copy (allocated on first switch to fp32)
  Bf16* valh = nullptr;   // nnz, bf16 copy (increment form with bf16 values only)
  int ntiles = 0;
  int* tiles = nullptr;   // ntiles+1 tile start rows
};

// A CSR's entries regrouped for the phased product (fp64 iteration kernels of large problems).
// Persistent workgroup g owns the rows [g·R, g·R + R), R = K·kPhBlock; phase p holds the entries
// of those rows whose column lies in slice p ([p·w, (p+1)·w), w = ⌈cols / S⌉), rows in order and
// each row's entries in CSR order. Sweeping the phases in order keeps every row's running sum
// in CSR order (columns are sorted in each row), so the sums are the bits of the row-tiled
// product, while the workgroups of an XCD gather from one slice of the vector at a time, which
// stays in that XCD's L2. Values are gathered from the CSR's values on the device (`src`).
struct Phased {
  int S = 0;                      // phases (0: not built)
  int K = 0, J = 0, G = 0, R = 0; // rows per thread, product slots per thread, workgroups, rows per workgroup
  int* blk = nullptr;             // G·S + 1 entry starts of the (workgroup, phase) blocks
  unsigned char* cnt = nullptr;   // G·S·R: entries of each row in each phase, thread-major rows
  int* col = nullptr;             // nnz
  int* src = nullptr;             // nnz: index of each entry in the CSR
  double* val = nullptr;          // nnz
  double* valo = nullptr;         // nnz: the original (unscaled) values, for the check pass
};
constexpr int kPhBlock = 512;                    // threads per phased workgroup
constexpr long long kPhMinBytes = 4LL << 20;     // gathered vectors up to this size stay in L2 anyway
// Below this many distinct 128-byte lines of the gathered vector per entry the row tiles reuse
// lines from L2 and beat the phased layout (a staircase planning LP at 4.2 MB gathered: 0.025
// lines per entry, 41k iterations/s on tiles against 28k phased; a random LP: 0.99 lines per
// entry, 7.75k phased against 5.29k on tiles).
constexpr double kPhMinLinesPerEntry = 0.5;
constexpr int kPhMaxJ = 8;                       // ≤ kPhMaxJ·kPhBlock entries per (workgroup, phase)

// A reduction job: Σ or max over t[0..na) followed by u[0..nb) (u optional).
struct RedJob {
  const double* a;
  const double* b;
  long long na, nb;
  int op;    // 0 sum, 1 max
  int mode;  // 0 tree, 1 64-chunk CPU order, 2 single sequential loop
};

struct JobTable {
  RedJob* dev = nullptr;
  std::vector<RedJob> host;
};

// The 64-chunk sums of a job table that a column pass carries: one wavefront per (job, chunk),
// in workgroups placed ahead of the column tiles of the same launch. The chunk sums are
// latency-bound sequential chains that read only what the passes before the column pass wrote,
// so they run beside the column tiles without a second stream.
struct ChunkSums {
  const RedJob* jobs = nullptr;  // the table on the device
  double* partials = nullptr;    // chunk k of job j goes to partials[j * kTreeBlocks + k]
  unsigned mask = 0;             // bit j: job j of the table is a 64-chunk sum
  int blocks = 0;                // workgroups ahead of the column tiles (16 per job: 4 chains each)
};
constexpr int kChainsPerBlock = 4;  // chains per workgroup of a column pass that carries ChunkSums

// Per-element term arrays of one check pass (each c->term_len doubles in c->terms).
enum Term {
  T_DX2, T_DY2, T_CROSS,                               // fixed-point residual
  T_RP2, T_B2, T_DROW, T_RD2ROW,                       // KKT, row part
  T_C2, T_P, T_DCOL, T_RD2COL,                         // KKT, column part
  T_PIDX, T_PIDY,                                      // ‖x̂ − x0‖², ‖ŷ − y0‖²
  T_ROBJ_ROW, T_RMAG_ROW, T_ROBJ_COL, T_RMAG_COL, T_CD, T_CMAG,        // ray test of T(z) − z
  T_ROBJ_ROW2, T_RMAG_ROW2, T_ROBJ_COL2, T_RMAG_COL2, T_CD2, T_CMAG2,  // ray test of the drift
  kTermCount
};
// Result slots in c->results.
enum Result {
  R_PX, R_PY, R_CROSS, R_RP2, R_B2, R_D, R_RD2, R_C2, R_P, R_PIDX, R_PIDY,
  R_ROBJ, R_RMAG, R_CD, R_CMAG, R_ROBJ2, R_RMAG2, R_CD2, R_CMAG2, kSumCount,
  R_MAX0 = 24,  // maxima, stored as fp64 bit patterns by atomicMax
  R_R0 = 40,    // residual sums of the last epoch start (px, py, cross)
  kResultCount = 44
};
enum Max { M_PMAX, M_DMAX, M_CINF, M_RMAX, M_DVIOL, M_DMAXRAY, M_PVIOL, M_RMAX2, M_DVIOL2, M_DMAXRAY2, M_PVIOL2, kMaxCount };
enum Flag : unsigned { F_RES = 1, F_KKT = 2, F_RAY = 4, F_DRIFT = 8, F_PID = 16 };

// Mixed precision in increment form: fp32 δ around the fp64 anchors in slots X0 / Y0.
struct DeltaState {
  bool on = false, dirty = true;
  bool bf16 = false;   // Ã's values of the increment SpMVs in bf16 (else fp32)
  double sigma = 0.0;  // σ the dual constants were built with
  float *dx = nullptr, *dxh = nullptr, *dxb = nullptr, *G = nullptr, *LX = nullptr, *UX = nullptr;
  float *dy = nullptr, *dyh = nullptr, *dyb = nullptr, *BL = nullptr, *BU = nullptr, *KY = nullptr;
};

// Device arrays of a batch of scenarios (interleaved block vectors, W slots of `cap`); plain
// data, passed to the batch kernels by value.
struct BatchArrays {
  int W = 0, cap = 0;
  long long rows_max = 0;
  double *c = nullptr, *l = nullptr, *u = nullptr, *rl = nullptr, *ru = nullptr;
  double *cmin = nullptr, *clo = nullptr, *cuo = nullptr, *rlo = nullptr, *ruo = nullptr;
  double *bs = nullptr, *os = nullptr;
  double *x = nullptr, *x0 = nullptr, *xh = nullptr, *xb = nullptr, *y = nullptr, *y0 = nullptr, *yh = nullptr,
         *yb = nullptr, *dx = nullptr, *dy = nullptr, *adx = nullptr, *xo = nullptr, *yo = nullptr;
  double *rv = nullptr, *dv = nullptr;  // unscaled dual (m·cap) and primal (n·cap) rays of the check
  double *tau = nullptr, *sigma = nullptr;
  long long* inner0 = nullptr;
  int *mask = nullptr, *keep = nullptr;
  double* terms = nullptr;     // per-element terms, kBTerms arrays of rows_max·cap
  double* partials = nullptr;  // tree partials
  double* stats = nullptr;     // cap · PSH_BSTAT_COUNT
  double* tmp = nullptr;       // compaction buffer, rows_max·cap
  // increment form (psh_batch_delta): fp32 δ of every slot around its fp64 anchor x0 / y0, and
  // the anchor constants G, LX, UX (n·cap) and BL, BU, KY (m·cap) of the single-solve form
  float *ix = nullptr, *ixh = nullptr, *ixb = nullptr, *G = nullptr, *LX = nullptr, *UX = nullptr;
  float *iy = nullptr, *iyh = nullptr, *iyb = nullptr, *BL = nullptr, *BU = nullptr, *KY = nullptr;
  int* anchor = nullptr;       // per-slot codes of the next anchor pass (1: KY only, 2: all)
};

struct BatchState : BatchArrays {
  char* pinned = nullptr;      // host staging: params | mask | keep | stats | anchor codes
  bool have_r0 = false;
  unsigned pin_pending = 0;                    // regions of `pinned` a queued copy may still read
  std::map<long long, hipGraphExec_t> graphs;  // captured chunks
  double graphs_rho = 1.0;                     // the reflection they were captured with
  bool delta = false;
  std::vector<double> ky_sigma;  // σ each slot's KY was built with
  std::vector<char> dirty;       // slots to re-anchor before the next chunk
};

}  // namespace psh

struct psh_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  int m = 0, n = 0;
  long long nnz = 0;
  bool loaded = false, fp32 = false;
  bool scaled = false;  // psh_scale ran
  size_t bytes = 0;

  psh::DevMat A, At;             // scaled Ã (m×n) and Ãᵀ (n×m)
  psh::Phased Pa, Pat;           // Ã and Ãᵀ regrouped for the phased product (large problems)
  int sliced = 1;                // 0: row tiles only; 1: phased where built (auto); 2: build at any size
  double* a_orig = nullptr;      // original A values, Ã's structure
  int* ato_col = nullptr;        // original Aᵀ (CSC) row indices (aliases At.col when equal)
  bool ato_alias = false;
  double* ato_val = nullptr;

  double *c = nullptr, *l = nullptr, *u = nullptr, *rl = nullptr, *ru = nullptr;
  float *cf = nullptr, *lf = nullptr, *uf = nullptr, *rlf = nullptr, *ruf = nullptr;
  double *rs = nullptr, *cs = nullptr;
  double bs = 1, os = 1;
  double *clo = nullptr, *cuo = nullptr, *rlo = nullptr, *ruo = nullptr, *cmin = nullptr;

  double* vd[PSH_NSLOT] = {};
  float* vf[PSH_NSLOT] = {};

  // check-pass scratch
  double *xo = nullptr, *yo = nullptr;          // unscaled iterate
  double *rv = nullptr, *dv = nullptr;          // dual / primal ray of (Δx, Δy)
  double *rv2 = nullptr, *dv2 = nullptr;        // rays of the drift
  double* terms = nullptr;                      // kTerms arrays of max(m, n)
  double *mpart_m = nullptr, *mpart_n = nullptr; // per-workgroup maxima of the check passes
  int mpart_rows_m = 0, mpart_rows_n = 0;        // workgroup slots in each
  long long term_len = 0;
  bool boxed = false;                           // every column has two finite bounds: the primal ray is 0
  double* partials = nullptr;                   // tree first level
  double* results = nullptr;                    // device results
  double* host_results = nullptr;               // pinned
  psh::RedJob* jobs_dev = nullptr;
  psh::RedJob* jobs_host = nullptr;             // pinned

  // Lanczos
  double *lv = nullptr, *lvp = nullptr, *lw = nullptr, *lt = nullptr;

  // epoch-start residual sums kept on the device (in `results`) until the next check
  double* r0_dev = nullptr;
  bool have_r0 = false;
  // persistent reduction tables, keyed by (mode, check flags)
  std::map<int, psh::JobTable> tables;

  psh::BatchState* batch = nullptr;

  psh::DeltaState delta;

  // chunk parameters {τ, σ, inner₀} on the device, and captured chunk graphs
  double* prm = nullptr;
  bool graphs = true;
  std::map<long long, hipGraphExec_t> graph_cache;
  double graph_rho = 1.0;  // the reflection the cached chunk graphs were captured with

  // timing
  bool timing = false;
  std::vector<hipEvent_t> ev_pool;
  struct Pending { int cls; hipEvent_t a, b; };
  std::vector<Pending> pending;
  long long t_launch[3] = {0, 0, 0};
  double t_ms[3] = {0, 0, 0};
};

namespace psh {

constexpr int kTerms = kTermCount;  // term arrays in the scratch area
constexpr int kMaxJobs = 64;

int fail(psh_ctx* c, int code, const std::string& what);
int hip_fail(psh_ctx* c, hipError_t e, const char* where);
#define PSH_TRY(ctx, call)                                   \
  do {                                                       \
    hipError_t e_ = (call);                                  \
    if (e_ != hipSuccess) return psh::hip_fail(ctx, e_, #call); \
  } while (0)
#define PSH_LAUNCHED(ctx, where)                                   \
  do {                                                             \
    hipError_t e_ = hipGetLastError();                             \
    if (e_ != hipSuccess) return psh::hip_fail(ctx, e_, where);    \
  } while (0)

// Several host threads may drive their own contexts at once. Graph capture (even thread-local)
// must not overlap allocation and release calls of other threads, which synchronise the device:
// captures and every hipMalloc / hipFree / hipHostMalloc / hipHostFree of this library take this
// lock. Copies always go through a context's own stream, never the legacy stream.
std::recursive_mutex& api_mutex();

int dev_alloc(psh_ctx* c, void** p, size_t bytes);
template <class T>
int alloc_n(psh_ctx* c, T** p, long long count) {
  return dev_alloc(c, reinterpret_cast<void**>(p), static_cast<size_t>(count > 0 ? count : 1) * sizeof(T));
}
void dev_free(void* p);

std::vector<int> build_tiles(const int* ptr, int rows);
inline int grid_for_tiles(int ntiles) { return ((ntiles + kXcd - 1) / kXcd) * kXcd; }
inline int grid_for(long long n) {
  long long g = (n + kBlock - 1) / kBlock;
  return static_cast<int>(g < 1 ? 1 : g);
}
bool is_primal_slot(int s);
long long slot_len(const psh_ctx* c, int s);

// Runs `njobs` reductions (sums in the job's mode, maxes as trees); results land in
// c->results[first .. first+njobs) on the device (not synchronised).
int reduce(psh_ctx* c, const RedJob* jobs, int njobs, int first);
// Same, with a table already on the device (no upload, no synchronisation).
int reduce_table(psh_ctx* c, const JobTable& t, int first);
// Parts of a table's reductions: 1 = the tree jobs, 2 = the 64-chunk / sequential sum jobs,
// 4 = only the final in-order combine of the sum jobs' chunk partials (the chunks having been
// summed by a column pass that carried them, see ChunkSums).
int reduce_table_part(psh_ctx* c, const JobTable& t, int first, int parts);
// The 64-chunk sums (mode 1) of a table, carried by a column pass (blocks = -1 when the table
// has a job that cannot be carried: a single-loop sum or an index beyond the mask).
ChunkSums chunk_sums_of(const psh_ctx* c, const JobTable& t);
// The table for a check pass with `flags` in reduction `mode` (created on first use).
const JobTable* check_table(psh_ctx* c, int mode, unsigned flags, int* err);
// Copies `count` device results to the host (synchronises the stream).
int fetch_results(psh_ctx* c, int count, double* out);

// fp32 copies of the matrix values (psh_core.hip)
int ensure_fp32_matrix(psh_ctx* c);
int ensure_bf16_matrix(psh_ctx* c);

// frees the batch state of a context (psh_batch.hip)
void destroy_batch(psh_ctx* c);

// timing helpers
void time_begin(psh_ctx* c, int cls, hipEvent_t* a);
void time_end(psh_ctx* c, int cls, hipEvent_t a);

}  // namespace psh

// ------------------------------------------------------------------ device helpers
namespace psh {

// Workgroup → tile: the dispatcher deals workgroups round-robin to the 8 XCDs, so give each XCD
// a contiguous run of tiles.
__device__ inline int tile_of_block(int bid, int nblocks) {
  const int per = nblocks / kXcd;
  return (bid % kXcd) * per + bid / kXcd;
}

template <class T>
__device__ inline T clampv(T v, T lo, T hi) {
  return v < lo ? lo : (v > hi ? hi : v);
}

// Row-tiled sparse product with an fp64 row-sequential accumulator (CSR order) and a fused
// epilogue epi(row, has_row, sum) run by every thread of the tile's workgroup.
// The rows [r0, r1) of one tile (the whole workgroup calls it; r1 − r0 ≤ blockDim.x).
// Loads of data that every launch streams once (matrix values and indices, the epilogue's
// vectors), marked non-temporal so that the L2 keeps the randomly gathered vector.
template <class T>
__device__ inline T ld_stream(const T* p) {
  return __builtin_nontemporal_load(p);
}
template <>
__device__ inline Bf16 ld_stream<Bf16>(const Bf16* p) {
  return Bf16{__builtin_nontemporal_load(&p->bits)};
}

// Device data (matrices and vectors) lives in global memory. A pointer that was itself read from
// memory (a table of instances or jobs) reaches the kernel as a generic one, whose loads also
// count on the LDS counter and may alias LDS stores; as_global tells the compiler the address
// space, so such loads can be issued ahead of LDS stores and waited for on their own counter.
template <class T>
__device__ inline const T* as_global(const T* p) {
  return (const T*)(const __attribute__((address_space(1))) T*)p;
}

// An epilogue may load its row's operands ahead of the product loop: prefetch(row, has).
template <class E, class = void>
struct has_prefetch : std::false_type {};
template <class E>
struct has_prefetch<E, std::void_t<decltype(std::declval<E&>().prefetch(0, true))>> : std::true_type {};

// A thread's products of one LDS pass over entries [cb, ce) of a tile (ce − cb ≤ kTileNnz): entry
// cb + threadIdx.x + j·kBlock gives v[j] = val·x[col], for the slots j of [J0, J0 + NJ). Every
// load is issued unconditionally, at entry cb for a slot past ce (a valid entry whose product is
// not stored), so that the loads and gathers of all NJ slots are in flight together instead of
// one slot's after another's.
template <int J0, int NJ, class V, class X>
__device__ inline void tile_products(int cb, int ce, const int* __restrict__ col, const V* __restrict__ val,
                                     const X* __restrict__ x, double (&v)[kTileNnz / kBlock]) {
  int c[NJ];
#pragma unroll
  for (int j = 0; j < NJ; ++j) {
    const int p = cb + static_cast<int>(threadIdx.x) + (J0 + j) * kBlock;
    c[j] = ld_stream(col + (p < ce ? p : cb));
  }
#pragma unroll
  for (int j = 0; j < NJ; ++j) {
    const int p = cb + static_cast<int>(threadIdx.x) + (J0 + j) * kBlock;
    v[J0 + j] = static_cast<double>(ld_stream(val + (p < ce ? p : cb))) * static_cast<double>(x[c[j]]);
  }
}

// All of a pass's slots in two batches of half the slots: the second only when the pass has entries
// there (a uniform test), so that passes of short rows do not issue loads for empty slots.
template <class V, class X>
__device__ inline void tile_products(int cb, int ce, const int* __restrict__ col, const V* __restrict__ val,
                                     const X* __restrict__ x, double (&v)[kTileNnz / kBlock]) {
  constexpr int kHalf = kTileNnz / kBlock / 2;
  tile_products<0, kHalf>(cb, ce, col, val, x, v);
  if (cb + kHalf * kBlock < ce) tile_products<kHalf, kHalf>(cb, ce, col, val, x, v);
}

template <class V, class X, class Epi>
__device__ inline void tile_rows(int r0, int r1, const int* __restrict__ ptr, const int* __restrict__ col,
                                 const V* __restrict__ val, const X* __restrict__ x, Epi& epi) {
  __shared__ double prod[kTileNnz];
  ptr = as_global(ptr), col = as_global(col), val = as_global(val), x = as_global(x);
  const int p0 = ptr[r0], p1 = ptr[r1];
  const int row = r0 + static_cast<int>(threadIdx.x);
  const bool has = row < r1;
  if constexpr (has_prefetch<Epi>::value) epi.prefetch(row, has);
  const int rb = has ? ptr[row] : 0, re = has ? ptr[row + 1] : 0;
  double s = 0.0;
  for (int cb = p0; cb < p1; cb += kTileNnz) {
    const int ce = p1 - cb > kTileNnz ? cb + kTileNnz : p1;
    double v[kTileNnz / kBlock];
    tile_products(cb, ce, col, val, x, v);
#pragma unroll
    for (int j = 0; j < kTileNnz / kBlock; ++j) {
      const int p = cb + static_cast<int>(threadIdx.x) + j * kBlock;
      if (p < ce) prod[p - cb] = v[j];
    }
    __syncthreads();
    if (has) {
      const int b = rb > cb ? rb : cb, e = re < ce ? re : ce;
      for (int p = b; p < e; ++p) s += prod[p - cb];
    }
    __syncthreads();
  }
  epi(row, has, s);
}

// Row sums of one tile with the products staged in LDS (CSR order, as tile_rows): s = Σ a·x and,
// with `abs_too`, sa = Σ |a·x|. The whole workgroup calls it.
// tile_sums_in stages the products in `prod` (kTileNnz doubles of LDS that the caller owns).
template <class V, class X>
__device__ inline void tile_sums_in(double* __restrict__ prod, int r0, int r1, const int* __restrict__ ptr,
                                    const int* __restrict__ col, const V* __restrict__ val, const X* __restrict__ x,
                                    bool abs_too, double& s, double& sa) {
  ptr = as_global(ptr), col = as_global(col), val = as_global(val), x = as_global(x);
  const int p0 = ptr[r0], p1 = ptr[r1];
  const int row = r0 + static_cast<int>(threadIdx.x);
  const bool has = row < r1;
  const int rb = has ? ptr[row] : 0, re = has ? ptr[row + 1] : 0;
  s = 0.0;
  sa = 0.0;
  for (int cb = p0; cb < p1; cb += kTileNnz) {
    const int ce = p1 - cb > kTileNnz ? cb + kTileNnz : p1;
    double v[kTileNnz / kBlock];
    tile_products(cb, ce, col, val, x, v);
#pragma unroll
    for (int j = 0; j < kTileNnz / kBlock; ++j) {
      const int p = cb + static_cast<int>(threadIdx.x) + j * kBlock;
      if (p < ce) prod[p - cb] = v[j];
    }
    __syncthreads();
    if (has) {
      const int b = rb > cb ? rb : cb, e = re < ce ? re : ce;
      for (int p = b; p < e; ++p) s += prod[p - cb];
      if (abs_too)
        for (int p = b; p < e; ++p) sa += fabs(prod[p - cb]);
    }
    __syncthreads();
  }
}

template <class V, class X>
__device__ inline void tile_sums(int r0, int r1, const int* __restrict__ ptr, const int* __restrict__ col,
                                 const V* __restrict__ val, const X* __restrict__ x, bool abs_too, double& s,
                                 double& sa) {
  __shared__ double prod[kTileNnz];
  tile_sums_in(prod, r0, r1, ptr, col, val, x, abs_too, s, sa);
}

// ---- reduction shapes shared by the context reductions and the allocation-free entry point
// Fixed-shape sum (max) of one value per thread over a kBlock-thread workgroup: pairwise in LDS,
// halving strides. Every thread calls it; the result is valid in thread 0.
__device__ inline double block_tree_sum(double acc) {
  __shared__ double red[kBlock];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = kBlock / 2; w > 0; w >>= 1) {
    if (static_cast<int>(threadIdx.x) < w) red[threadIdx.x] = red[threadIdx.x] + red[threadIdx.x + w];
    __syncthreads();
  }
  return red[0];
}
__device__ inline double block_tree_max(double acc) {
  __shared__ double red[kBlock];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = kBlock / 2; w > 0; w >>= 1) {
    if (static_cast<int>(threadIdx.x) < w) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + w]);
    __syncthreads();
  }
  return red[0];
}

// Orders the LDS accesses of one wavefront around this point: a wavefront's LDS operations are
// performed in the order it issues them, so keeping the compiler from moving loads and stores
// across it is enough for the lanes of one wavefront to see each other's stores.
__device__ inline void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// s + t(b) + t(b+1) + ... + t(e−1), strictly left to right, by one wavefront: the lanes form 512
// consecutive terms (coalesced loads) into an LDS stage, issue the loads of the next 512, and
// meanwhile every lane adds the staged terms in order (a uniform dependent chain, so the lanes
// agree). The last batch is padded with +0.0: the running sum is never −0.0 (it starts at +0.0
// or at a sum that started there, and x + (−x) rounds to +0.0), so adding +0.0 leaves it
// unchanged bit for bit. `stage` holds kSeqBatch doubles of LDS. WaveOnly: the stage belongs to
// this wavefront alone (one of several in a workgroup, each with its own chain), so only the
// wavefront synchronises; otherwise the workgroup is this one wavefront and uses its barrier.
constexpr int kSeqBatch = 512;
template <bool WaveOnly = false, class Get>
__device__ inline double wave_seq_sum(Get get, long long b, long long e, double s, double* __restrict__ stage) {
  const int lane = static_cast<int>(threadIdx.x) & 63;
  auto sync = [] {
    if constexpr (WaveOnly) wave_lds_sync();
    else __syncthreads();
  };
  double v[kSeqBatch / 64];
#pragma unroll
  for (int u = 0; u < kSeqBatch / 64; ++u) {
    const long long i = b + u * 64 + lane;
    v[u] = i < e ? get(i) : 0.0;
  }
  for (long long base = b; base < e; base += kSeqBatch) {
#pragma unroll
    for (int u = 0; u < kSeqBatch / 64; ++u) stage[u * 64 + lane] = v[u];
    sync();
    const long long next = base + kSeqBatch;
#pragma unroll
    for (int u = 0; u < kSeqBatch / 64; ++u) {
      const long long i = next + u * 64 + lane;
      v[u] = i < e ? get(i) : 0.0;
    }
#pragma unroll 16
    for (int k = 0; k < kSeqBatch; ++k) s += stage[k];
    sync();
  }
  return s;
}

// Chunk `k` of a CPU-order sum job by one wavefront (la/parallel.h:169-188): mode 1 splits a into
// min(64, max(1, na)) chunks [na·k/chunks, na·(k+1)/chunks) and writes the chunk's sequential
// sum to row[k]; mode 2 is one sequential loop over a then b (cpu_backend.cpp:240-274), run by
// chunk 0 into row[0]. The in-order combine of the chunk sums is k_ordered_final's.
// The terms are read through global-address-space pointers: the table's pointers are plain
// (generic) ones, whose loads count on the LDS counter too, so every wait for the staged terms
// would also wait for the next batch's loads from memory.
using GlobalDouble = const __attribute__((address_space(1))) double;
template <bool WaveOnly>
__device__ inline void ordered_chunk(const RedJob& J, double* __restrict__ row, int k, double* __restrict__ stage) {
  const bool lead = (threadIdx.x & 63) == 0;
  GlobalDouble* const a = (GlobalDouble*)J.a;
  auto ga = [a](long long i) { return a[i]; };
  if (J.mode == 2) {
    if (k != 0) return;
    GlobalDouble* const b = (GlobalDouble*)J.b;
    double s = wave_seq_sum<WaveOnly>(ga, 0, J.na, 0.0, stage);
    s = wave_seq_sum<WaveOnly>([b](long long i) { return b[i]; }, 0, J.nb, s, stage);
    if (lead) row[0] = s;
    return;
  }
  const long long n = J.na;
  const int chunks = static_cast<int>(n < 1 ? 1 : (n > 64 ? 64 : n));
  if (k >= chunks) return;
  const double s = wave_seq_sum<WaveOnly>(ga, n * k / chunks, n * (k + 1) / chunks, 0.0, stage);
  if (lead) row[k] = s;
}

// One sweep of the phased product (struct Phased) by workgroup g: thread t accumulates the rows
// t·K .. t·K + K − 1 of the workgroup in registers across the phases (products staged in `prod`,
// J·kPhBlock doubles of LDS, each row's products added in order), and with ABS the sums of the
// products' magnitudes in the same order. `wtot` is kPhBlock / 64 ints of LDS. The whole
// workgroup calls it.
template <int K, int J, bool ABS, class V, class X>
__device__ inline void phased_accumulate(const Phased& P, int g, const V* __restrict__ val, const X* __restrict__ x,
                                         double* __restrict__ prod, int* __restrict__ wtot, double (&acc)[K],
                                         double (&acca)[K]) {
  const int t = static_cast<int>(threadIdx.x), lane = t & 63, w = t >> 6;
#pragma unroll
  for (int k = 0; k < K; ++k) acc[k] = 0.0, acca[k] = 0.0;
  for (int p = 0; p < P.S; ++p) {
    const int bp = g * P.S + p;
    const int b0 = P.blk[bp], ne = P.blk[bp + 1] - b0;
    unsigned cn[K];
    const unsigned char* cp = P.cnt + static_cast<size_t>(bp) * P.R + t * K;
    if constexpr (K == 4) {
      const unsigned v = *reinterpret_cast<const unsigned*>(cp);
#pragma unroll
      for (int k = 0; k < 4; ++k) cn[k] = (v >> (8 * k)) & 255u;
    } else if constexpr (K == 2) {
      const unsigned short v = *reinterpret_cast<const unsigned short*>(cp);
      cn[0] = v & 255u;
      cn[1] = static_cast<unsigned>(v) >> 8;
    } else {
#pragma unroll
      for (int k = 0; k < K; ++k) cn[k] = cp[k];
    }
    int mine = 0;
#pragma unroll
    for (int k = 0; k < K; ++k) mine += static_cast<int>(cn[k]);
    int inc = mine;  // inclusive scan of the entry counts over the wavefront
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const int v = __shfl_up(inc, o);
      if (lane >= o) inc += v;
    }
    if (lane == 63) wtot[w] = inc;
#pragma unroll
    for (int j = 0; j < J; ++j) {
      const int q = t + j * kPhBlock;
      if (q < ne) prod[q] = static_cast<double>(ld_stream(val + b0 + q)) * static_cast<double>(x[ld_stream(P.col + b0 + q)]);
    }
    __syncthreads();
    int off = inc - mine;
    for (int k = 0; k < w; ++k) off += wtot[k];
#pragma unroll
    for (int k = 0; k < K; ++k)
      for (unsigned c = 0; c < cn[k]; ++c) {
        const double v = prod[off++];
        acc[k] += v;
        if constexpr (ABS) acca[k] += fabs(v);
      }
    __syncthreads();
  }
}

// A sweep's sums in the epilogue's row order: out[k] is the sum of the workgroup's row
// k·kPhBlock + t (so the epilogue's accesses are coalesced), through `part` (K·kPhBlock doubles
// of LDS). The whole workgroup calls it.
template <int K>
__device__ inline void phased_transpose(const double (&acc)[K], double* __restrict__ part, double (&out)[K]) {
  const int t = static_cast<int>(threadIdx.x);
#pragma unroll
  for (int k = 0; k < K; ++k) part[t * K + k] = acc[k];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < K; ++k) out[k] = part[k * kPhBlock + t];
  __syncthreads();
}

// The phased product with the epilogue of the row-tiled kernels: one sweep (phased_accumulate),
// then the epilogue runs on rows k·kPhBlock + t. The whole workgroup calls it.
template <int K, int J, class V, class X, class Epi>
__device__ inline void phased_rows(const Phased& P, int rows, const V* __restrict__ val, const X* __restrict__ x,
                                   Epi& epi) {
  __shared__ double prod[J * kPhBlock];
  __shared__ double part[K * kPhBlock];
  __shared__ int wtot[kPhBlock / 64];
  const int t = static_cast<int>(threadIdx.x), g = static_cast<int>(blockIdx.x);
  double acc[K], unused[K];
  phased_accumulate<K, J, false>(P, g, val, x, prod, wtot, acc, unused);
#pragma unroll
  for (int k = 0; k < K; ++k) part[t * K + k] = acc[k];
  __syncthreads();
  const int r0 = g * P.R;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int row = r0 + k * kPhBlock + t;
    const bool has = row < rows;
    if constexpr (has_prefetch<Epi>::value) epi.prefetch(row, has);
    epi(row, has, part[k * kPhBlock + t]);
  }
}

// Calls f(K, J) with K, J as compile-time constants for a built Phased (K ∈ {1,2,4}, J ∈ {2,4,8}).
template <class F>
inline void phased_dispatch(const Phased& P, F&& f) {
  auto by_j = [&](auto kc) {
    if (P.J <= 2) f(kc, std::integral_constant<int, 2>{});
    else if (P.J <= 4) f(kc, std::integral_constant<int, 4>{});
    else f(kc, std::integral_constant<int, 8>{});
  };
  if (P.K == 1) by_j(std::integral_constant<int, 1>{});
  else if (P.K == 2) by_j(std::integral_constant<int, 2>{});
  else by_j(std::integral_constant<int, 4>{});
}

template <class V, class X, class Epi>
__device__ inline void tiled_rows(const int* __restrict__ tiles, int ntiles, const int* __restrict__ ptr,
                                  const int* __restrict__ col, const V* __restrict__ val, const X* __restrict__ x,
                                  Epi& epi) {
  const int tile = tile_of_block(blockIdx.x, gridDim.x);
  if (tile >= ntiles) return;
  tile_rows<V, X, Epi>(tiles[tile], tiles[tile + 1], ptr, col, val, x, epi);
}

}  // namespace psh
