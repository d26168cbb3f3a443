// psh_noalloc.hip — psh_dot_noalloc: the dot product of psh_dot for callers that cannot allocate
// device memory (code under HIP graph capture, tight serving loops, harnesses that forbid
// hipMalloc). The operands and the output are the caller's device arrays; the only scratch, the
// partial sums of the reduction, lives in a file-scope device array whose address is looked up
// once per device. Products are formed inside the reduction kernels (no per-element term array),
// and the reduction shapes are psh_dot's, so the result has psh_dot's bits in each mode.
#include <cstdio>
#include <mutex>

#include "psh_internal.h"

using namespace psh;

namespace {

// kTreeBlocks first-level partials of a tree reduction, or the 64 chunk sums of the CPU order.
constexpr int kScratchDoubles = kTreeBlocks;
__device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;
std::mutex g_mutex;
double* g_scratch[kMaxDevices] = {};  // per device: the address of `scratch` on that device
thread_local char g_error[200] = "";

int na_fail(int code, const char* what) {
  std::snprintf(g_error, sizeof g_error, "%s", what);
  return code;
}

// The device address of `scratch` on the current device (looked up on the first call there).
int scratch_of(int device, double** out) {
  std::lock_guard<std::mutex> lock(g_mutex);
  if (!g_scratch[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      std::snprintf(g_error, sizeof g_error, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
      return PSH_E_HIP;
    }
    g_scratch[device] = static_cast<double*>(p);
  }
  *out = g_scratch[device];
  return PSH_OK;
}

// Tree mode: block x sums a[i]·b[i] over [n·x/512, n·(x+1)/512) with the stride-256 thread loop and
// the fixed LDS tree of k_tree1 (psh_dot's terms are the same rounded products).
__global__ void __launch_bounds__(kBlock) k_na_tree1(const double* __restrict__ a, const double* __restrict__ b,
                                                     long long n, double* __restrict__ part) {
  const long long lo = n * blockIdx.x / kTreeBlocks, hi = n * (blockIdx.x + 1) / kTreeBlocks;
  double acc = 0.0;
  for (long long i = lo + threadIdx.x; i < hi; i += kBlock) acc += a[i] * b[i];
  const double r = block_tree_sum(acc);
  if (threadIdx.x == 0) part[blockIdx.x] = r;
}

__global__ void __launch_bounds__(kBlock) k_na_tree2(const double* __restrict__ part, double* __restrict__ out) {
  double acc = 0.0;
  for (int i = threadIdx.x; i < kTreeBlocks; i += kBlock) acc += part[i];
  const double r = block_tree_sum(acc);
  if (threadIdx.x == 0) out[0] = r;
}

// CPU order: 64 fixed chunks summed left to right (la/parallel.h:169-188), one wavefront each.
__global__ void __launch_bounds__(64) k_na_chunks(const double* __restrict__ a, const double* __restrict__ b,
                                                  long long n, double* __restrict__ part) {
  __shared__ double stage[kSeqBatch];
  const int chunks = static_cast<int>(n < 1 ? 1 : (n > 64 ? 64 : n));
  const int c = static_cast<int>(blockIdx.x);
  if (c >= chunks) return;
  const long long lo = n * c / chunks, hi = n * (c + 1) / chunks;
  const double s = wave_seq_sum([&](long long i) { return a[i] * b[i]; }, lo, hi, 0.0, stage);
  if (threadIdx.x == 0) part[c] = s;
}

__global__ void k_na_chunks_final(const double* __restrict__ part, long long n, double* __restrict__ out) {
  if (threadIdx.x != 0) return;
  const int chunks = static_cast<int>(n < 1 ? 1 : (n > 64 ? 64 : n));
  double s = 0.0;
  for (int k = 0; k < chunks; ++k) s += part[k];
  out[0] = s;
}

}  // namespace

extern "C" {

int psh_dot_noalloc(const double* a, const double* b, long long n, int mode, double* out, void* stream) {
  if (n < 0 || n > PSH_NOALLOC_MAX_N)
    return na_fail(PSH_E_ARG, "psh_dot_noalloc: n must be in [0, PSH_NOALLOC_MAX_N]");
  if (!out || (n > 0 && (!a || !b))) return na_fail(PSH_E_ARG, "psh_dot_noalloc: null array");
  if (mode != PSH_RED_TREE && mode != PSH_RED_CPU && mode != PSH_RED_DECISION)
    return na_fail(PSH_E_ARG, "psh_dot_noalloc: unknown reduction mode");
  int device = 0;
  if (hipGetDevice(&device) != hipSuccess || device < 0 || device >= kMaxDevices)
    return na_fail(PSH_E_NO_DEVICE, "psh_dot_noalloc: no current HIP device");
  double* part = nullptr;
  int rc = scratch_of(device, &part);
  if (rc) return rc;
  const hipStream_t s = static_cast<hipStream_t>(stream);
  if (mode == PSH_RED_TREE) {
    hipLaunchKernelGGL(k_na_tree1, dim3(kTreeBlocks), dim3(kBlock), 0, s, a, b, n, part);
    hipLaunchKernelGGL(k_na_tree2, dim3(1), dim3(kBlock), 0, s, part, out);
  } else {
    hipLaunchKernelGGL(k_na_chunks, dim3(64), dim3(64), 0, s, a, b, n, part);
    hipLaunchKernelGGL(k_na_chunks_final, dim3(1), dim3(64), 0, s, part, n, out);
  }
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    std::snprintf(g_error, sizeof g_error, "psh_dot_noalloc launch: %s", hipGetErrorString(e));
    return PSH_E_HIP;
  }
  return PSH_OK;
}

const char* psh_noalloc_error(void) { return g_error; }

}  // extern "C"
