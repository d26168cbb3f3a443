// psh_fused.hip — the fused r²HPDHG iteration and the one-pass check.
//
// Per iteration two kernels (reference loop src/pdhg/r2hpdhg.cpp:68-74):
//   k_iter_primal  tiles of Ãᵀ: aty_j = Σ ã_ij y_i (fp64, CSR order), then in the epilogue
//                  x̂ = proj(x − τ(c − aty)), x̄ = 2x̂ − x, x ← w(2ρx̂ + (1−2ρ)x) + (1−w)x0
//                  (cpu_backend.cpp:140-157); aty never goes to memory.
//   k_iter_dual    tiles of Ã: ax_i = Σ ã_ij x̄_j, then ŷ, ȳ and the Halpern update of y
//                  (cpu_backend.cpp:159-178); ax never goes to memory.
// x̂, ŷ and ȳ are stored only on the iterations whose values are read afterwards (the first
// step of an epoch, whose residual gives r⁰, and the last step before a check); the values
// themselves are the reference's.
//
// A check runs four kernels over the vectors and the original matrix (KKT of the unscaled
// point, termination.h:35-60; fixed-point residual, r2hpdhg.cpp:50-59; ray-test scalars,
// engine.cpp:82-155; PID distances, r2hpdhg.cpp:127) that write per-element terms, then the
// deterministic reducer, then ONE copy of ~44 doubles to the host.
#include <cmath>
#include <type_traits>

#include "psh_internal.h"

using namespace psh;

namespace {

// ------------------------------------------------------------------ iteration kernels
// Step sizes and the epoch counter live in device memory (`prm` = {τ, σ, inner₀}, written by
// k_set_params before each chunk), so a chunk's launches have fixed arguments and can be
// replayed from a captured graph. The Halpern weight of step s is (k+1)/(k+2), k = inner₀ + s,
// computed in fp64 (IEEE division, the host's bits) and then rounded to T like the host value.
__device__ inline double halpern_weight(const double* prm, int step) {
  const double k = prm[2] + static_cast<double>(step);
  return (k + 1.0) / (k + 2.0);
}

__global__ void k_set_params(double* prm, double tau, double sigma, double inner0) {
  if (threadIdx.x == 0) {
    prm[0] = tau;
    prm[1] = sigma;
    prm[2] = inner0;
  }
}

template <class T>
struct PrimalArgs {
  T* x;
  const T* x0;
  const T *c, *l, *u;
  T* xh;
  T* xb;
  const double* prm;
  int step;
  T a, b;
  int write_hat;
};

// The epilogues load their row's operands in prefetch(), which tile_rows calls before the
// product loop, so those loads are in flight while the gathers are; operator() then only computes
// and stores (the same expressions, so the same bits).
template <class T>
struct EpiPrimal {
  PrimalArgs<T> p;
  T tau, tw;
  T xj{}, cj{}, lj{}, uj{}, x0j{};
  __device__ explicit EpiPrimal(const PrimalArgs<T>& a)
      : p(a), tau(static_cast<T>(a.prm[0])), tw(static_cast<T>(halpern_weight(a.prm, a.step))) {}
  __device__ void prefetch(int j, bool has) {
    if (!has) return;
    xj = ld_stream(p.x + j);
    cj = ld_stream(p.c + j);
    lj = ld_stream(p.l + j);
    uj = ld_stream(p.u + j);
    x0j = ld_stream(p.x0 + j);
  }
  __device__ void operator()(int j, bool has, double s) {
    if (!has) return;
    const T g = static_cast<T>(s);
    const T h = clampv<T>(xj - tau * (cj - g), lj, uj);
    if (p.write_hat) p.xh[j] = h;
    p.xb[j] = 2 * h - xj;
    p.x[j] = tw * (p.a * h + p.b * xj) + (1 - tw) * x0j;
  }
};

template <class T>
__global__ void __launch_bounds__(kBlock) k_iter_primal(const int* __restrict__ tiles, int ntiles, const int* __restrict__ ptr,
                                                        const int* __restrict__ col, const T* __restrict__ val,
                                                        const T* __restrict__ y, PrimalArgs<T> p) {
  EpiPrimal<T> e(p);
  tiled_rows<T, T>(tiles, ntiles, ptr, col, val, y, e);
}

template <class T>
struct DualArgs {
  T* y;
  const T* y0;
  const T *rl, *ru;
  T* yh;
  T* yb;
  const double* prm;
  int step;
  T a, b;
  int write_hat;
};

template <class T>
struct EpiDual {
  DualArgs<T> p;
  T sg, tw;
  T yi{}, rli{}, rui{}, y0i{};
  __device__ explicit EpiDual(const DualArgs<T>& a)
      : p(a), sg(static_cast<T>(a.prm[1])), tw(static_cast<T>(halpern_weight(a.prm, a.step))) {}
  __device__ void prefetch(int i, bool has) {
    if (!has) return;
    yi = ld_stream(p.y + i);
    rli = ld_stream(p.rl + i);
    rui = ld_stream(p.ru + i);
    y0i = ld_stream(p.y0 + i);
  }
  __device__ void operator()(int i, bool has, double s) {
    if (!has) return;
    const T ax = static_cast<T>(s);
    const T q = ax - yi / sg;
    const T h = yi - sg * ax + sg * clampv<T>(q, rli, rui);
    if (p.write_hat) {
      p.yh[i] = h;
      p.yb[i] = 2 * h - yi;
    }
    p.y[i] = tw * (p.a * h + p.b * yi) + (1 - tw) * y0i;
  }
};

template <class T>
__global__ void __launch_bounds__(kBlock) k_iter_dual(const int* __restrict__ tiles, int ntiles, const int* __restrict__ ptr,
                                                      const int* __restrict__ col, const T* __restrict__ val,
                                                      const T* __restrict__ xb, DualArgs<T> p) {
  EpiDual<T> e(p);
  tiled_rows<T, T>(tiles, ntiles, ptr, col, val, xb, e);
}

// The same two half steps over the phased layouts (large problems; same sums, same epilogues).
template <class T, int K, int J>
__global__ void __launch_bounds__(kPhBlock) k_ph_primal(Phased P, int rows, const T* __restrict__ val,
                                                        const T* __restrict__ y, PrimalArgs<T> p) {
  EpiPrimal<T> e(p);
  phased_rows<K, J, T, T>(P, rows, val, y, e);
}

template <class T, int K, int J>
__global__ void __launch_bounds__(kPhBlock) k_ph_dual(Phased P, int rows, const T* __restrict__ val,
                                                      const T* __restrict__ xb, DualArgs<T> p) {
  EpiDual<T> e(p);
  phased_rows<K, J, T, T>(P, rows, val, xb, e);
}

// ------------------------------------------------------------------ increment form (mixed precision)
// The iterate is held as an fp64 anchor plus an fp32 increment, z = z₀ + δ, where z₀ is the
// epoch's restart point — which is also the Halpern anchor, so the anchor term of the averaging
// vanishes in δ. Per epoch the anchor terms are folded into fp32 constants computed in fp64:
//   G = c̃ − Ãᵀȳ₀, LX = l̃ − x₀, UX = ũ − x₀;   BL = Ãx₀ − r̃l, BU = Ãx₀ − r̃u, KY = ȳ₀/σ.
// Primal:  δx̂ = clamp(δx − τ(G − Ãᵀδy), LX, UX),  δx̄ = 2δx̂ − δx,  δx ← w(2ρδx̂ + (1−2ρ)δx)
// Dual:    a = Ãδx̄,  t = clamp(KY + δy/σ, BU + a, BL + a)  (= Ãx̄ − proj(Ãx̄ − y/σ)),
//          δŷ = δy − σt,  δȳ = 2δŷ − δy,  δy ← w(2ρδŷ + (1−2ρ)δy)
// which is the r²HPDHG step translated by z₀ (reference cpu_backend.cpp:140-178). Rounding is
// relative to |δ|, and every restart re-anchors in fp64, so the fp32 error does not
// accumulate: each epoch refines the last. Ã's values are read in fp32 (half the bytes).
struct DeltaP {
  float *dx, *dxh, *dxb;
  const float *G, *LX, *UX;
  const double* prm;
  int step;
  float a, b;
  int write_hat;
};

struct EpiDeltaP {
  DeltaP p;
  float tau, tw;
  float xj = 0.0f, gj = 0.0f, lj = 0.0f, uj = 0.0f;
  __device__ explicit EpiDeltaP(const DeltaP& a)
      : p(a), tau(static_cast<float>(a.prm[0])), tw(static_cast<float>(halpern_weight(a.prm, a.step))) {}
  __device__ void prefetch(int j, bool has) {
    if (!has) return;
    xj = ld_stream(p.dx + j);
    gj = ld_stream(p.G + j);
    lj = ld_stream(p.LX + j);
    uj = ld_stream(p.UX + j);
  }
  __device__ void operator()(int j, bool has, double s) {
    if (!has) return;
    const float g = static_cast<float>(s);
    const float h = clampv<float>(xj - tau * (gj - g), lj, uj);
    if (p.write_hat) p.dxh[j] = h;
    p.dxb[j] = 2 * h - xj;
    p.dx[j] = tw * (p.a * h + p.b * xj);
  }
};

// V = float, or Bf16 for the variant with bf16 matrix values (fp64 products and sums either way).
template <class V>
__global__ void __launch_bounds__(kBlock) k_delta_primal(const int* __restrict__ tiles, int ntiles, const int* __restrict__ ptr,
                                                         const int* __restrict__ col, const V* __restrict__ val,
                                                         const float* __restrict__ dy, DeltaP p) {
  EpiDeltaP e(p);
  tiled_rows<V, float>(tiles, ntiles, ptr, col, val, dy, e);
}

struct DeltaD {
  float *dy, *dyh, *dyb;
  const float *BL, *BU, *KY;
  const double* prm;
  int step;
  float a, b;
  int write_hat;
};

struct EpiDeltaD {
  DeltaD p;
  float sg, tw;
  float yi = 0.0f, kyi = 0.0f, bui = 0.0f, bli = 0.0f;
  __device__ explicit EpiDeltaD(const DeltaD& a)
      : p(a), sg(static_cast<float>(a.prm[1])), tw(static_cast<float>(halpern_weight(a.prm, a.step))) {}
  __device__ void prefetch(int i, bool has) {
    if (!has) return;
    yi = ld_stream(p.dy + i);
    kyi = ld_stream(p.KY + i);
    bui = ld_stream(p.BU + i);
    bli = ld_stream(p.BL + i);
  }
  __device__ void operator()(int i, bool has, double s) {
    if (!has) return;
    const float ax = static_cast<float>(s);
    const float t = clampv<float>(kyi + yi / sg, bui + ax, bli + ax);
    const float h = yi - sg * t;
    if (p.write_hat) {
      p.dyh[i] = h;
      p.dyb[i] = 2 * h - yi;
    }
    p.dy[i] = tw * (p.a * h + p.b * yi);
  }
};

template <class V>
__global__ void __launch_bounds__(kBlock) k_delta_dual(const int* __restrict__ tiles, int ntiles, const int* __restrict__ ptr,
                                                       const int* __restrict__ col, const V* __restrict__ val,
                                                       const float* __restrict__ dxb, DeltaD p) {
  EpiDeltaD e(p);
  tiled_rows<V, float>(tiles, ntiles, ptr, col, val, dxb, e);
}

// Anchor constants of an epoch (fp64 arithmetic, one rounding to fp32); resets δ when asked.
__global__ void k_delta_anchor_n(const double* __restrict__ c, const double* __restrict__ l, const double* __restrict__ u,
                                 const double* __restrict__ x0, const double* __restrict__ aty, float* __restrict__ G,
                                 float* __restrict__ LX, float* __restrict__ UX, float* __restrict__ dx,
                                 float* __restrict__ dxh, float* __restrict__ dxb, long long n, int reset) {
  const long long j = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const double x = x0[j];
  G[j] = static_cast<float>(c[j] - aty[j]);
  LX[j] = static_cast<float>(l[j] - x);
  UX[j] = static_cast<float>(u[j] - x);
  if (reset) dx[j] = dxh[j] = dxb[j] = 0.0f;
}

__global__ void k_delta_anchor_m(const double* __restrict__ rl, const double* __restrict__ ru, const double* __restrict__ ax,
                                 const double* __restrict__ y0, double sigma, float* __restrict__ BL,
                                 float* __restrict__ BU, float* __restrict__ KY, float* __restrict__ dy,
                                 float* __restrict__ dyh, float* __restrict__ dyb, long long m, int reset) {
  const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= m) return;
  BL[i] = static_cast<float>(ax[i] - rl[i]);
  BU[i] = static_cast<float>(ax[i] - ru[i]);
  KY[i] = static_cast<float>(y0[i] / sigma);
  if (reset) dy[i] = dyh[i] = dyb[i] = 0.0f;
}

// fp64 points from anchor + increment: ẑ, z̄ (and z when `cur` is set) into the fp64 slots.
__global__ void k_delta_point(const double* __restrict__ z0, const float* __restrict__ dh, const float* __restrict__ db,
                              const float* __restrict__ dc, double* __restrict__ zh, double* __restrict__ zb,
                              double* __restrict__ zc, long long n) {
  const long long j = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const double a = z0[j];
  zh[j] = a + static_cast<double>(dh[j]);
  zb[j] = a + static_cast<double>(db[j]);
  if (zc) zc[j] = a + static_cast<double>(dc[j]);
}

template <class T>
__global__ void k_restart(T* __restrict__ x, T* __restrict__ x0, const T* __restrict__ xh, long long n) {
  const long long j = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const T v = xh[j];
  x[j] = v;
  x0[j] = v;
}

// ------------------------------------------------------------------ check kernels
__device__ inline double wave_max(double v) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
  return v;
}

// Maximum of non-negative values over the workgroup (NaN ignored, as std::max(acc, NaN) does),
// stored as partial row `row` (the workgroup's index among the pass's workgroups) for quantity
// q; k_max_final combines the partials. Every thread of the NT-thread workgroup must reach this
// call (no atomics: one store per workgroup).
template <int NT = kBlock>
__device__ inline void block_max_into(double* part, int q, double v, int row) {
  __shared__ double wm[NT / 64];
  v = v >= 0.0 ? v : 0.0;
  v = wave_max(v);
  if ((threadIdx.x & 63) == 0) wm[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double mx = wm[0];
    for (int k = 1; k < NT / 64; ++k) mx = fmax(mx, wm[k]);
    part[static_cast<size_t>(row) * kMaxCount + q] = mx;
  }
  __syncthreads();
}

// A pass of `used` workgroups: workgroup g writes zeros into the partial rows g + used, g + 2·used,
// … below `rows` of its quantity q, so that maxima an earlier check on the same context left there
// with a launch of more workgroups (the row tiles) do not count.
__device__ inline void clear_max_rows(double* part, int q, int g, int used, int rows) {
  if (threadIdx.x != 0) return;
  for (int r = used + g; r < rows; r += used) part[static_cast<size_t>(r) * kMaxCount + q] = 0.0;
}
__device__ inline double nonneg(double v) { return v >= 0.0 ? v : 0.0; }  // NaN → 0, as block_max_into

__device__ inline double dmaxd(double a, double b) { return a < b ? b : a; }  // std::max
__device__ inline double dmind(double a, double b) { return b < a ? b : a; }  // std::min

__device__ inline double elem_viol(double v, double lo, double up) {  // cpu_backend.cpp:248-252
  if (v < lo) return (lo - v) / (1.0 + fabs(lo));
  if (v > up) return (v - up) / (1.0 + fabs(up));
  return 0.0;
}

template <class T>
struct CheckArgs {
  unsigned flags;
  int m, n;
  int pray;         // form A·d for the primal rays (0 when every column is boxed: d is then 0)
  // primal side
  const T *xh, *xb, *x0;
  T* dx;            // F_RES: Δx = x̄ − x̂ is written here; otherwise the ray direction is read here
  // dual side
  const T *yh, *yb, *y0;
  T* dy;
  T* ax;            // Ã·Δx (stored like the reference's work vector)
  // matrices
  const int *a_ptr, *a_col, *at_ptr, *ato_col;
  const T* a_val;   // scaled Ã (working precision)
  const double *a_orig, *ato_val;
  // scaling and original data
  const double *rs, *cs, *clo, *cuo, *rlo, *ruo, *cmin;
  double bs, os;
  // scratch
  double *xo, *yo, *rv, *dv, *rv2, *dv2;
  double* terms;
  long long tl;
  double *mpart_m, *mpart_n;  // per-workgroup maxima of the row-side (m) and column-side (n) passes
};

template <class T>
__device__ inline double* term(const CheckArgs<T>& a, int k) {
  return a.terms + static_cast<size_t>(k) * a.tl;
}

// Primal ray from a scaled direction (engine.cpp:132-139).
__device__ inline double primal_ray_entry(double dxs, double cs, double bs, double lo, double up) {
  double v = cs * dxs / bs;
  if (isfinite(lo)) v = dmaxd(v, 0.0);
  if (isfinite(up)) v = dmind(v, 0.0);
  return v;
}
// Dual ray from a scaled direction (engine.cpp:90-97).
__device__ inline double dual_ray_entry(double dys, double rs, double os, double lo, double up) {
  double v = rs * dys / os;
  if (!isfinite(lo)) v = dmind(v, 0.0);
  if (!isfinite(up)) v = dmaxd(v, 0.0);
  return v;
}

template <class T>
__global__ void __launch_bounds__(kBlock) k_check_n(CheckArgs<T> a) {
  const long long j = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  const bool in = j < a.n;
  double dmax_ray = 0, dmax_ray2 = 0;
  if (in) {
    T d = 0;
    if (a.flags & F_RES) {
      d = a.xb[j] - a.xh[j];
      a.dx[j] = d;
      term(a, T_DX2)[j] = static_cast<double>(d) * static_cast<double>(d);
    } else if (a.flags & F_RAY) {
      d = a.dx[j];
    }
    if (a.flags & F_KKT)  // unscale_primal (scaling.cpp:144-150)
      a.xo[j] = dmind(dmaxd(a.cs[j] * static_cast<double>(a.xh[j]) / a.bs, a.clo[j]), a.cuo[j]);
    if (a.flags & F_RAY) {
      const double v = primal_ray_entry(static_cast<double>(d), a.cs[j], a.bs, a.clo[j], a.cuo[j]);
      a.dv[j] = v;
      dmax_ray = fabs(v);
      term(a, T_CD)[j] = a.cmin[j] * v;
      term(a, T_CMAG)[j] = fabs(a.cmin[j] * v);
    }
    if (a.flags & F_DRIFT) {
      const T dd = a.xh[j] - a.x0[j];
      const double v = primal_ray_entry(static_cast<double>(dd), a.cs[j], a.bs, a.clo[j], a.cuo[j]);
      a.dv2[j] = v;
      dmax_ray2 = fabs(v);
      term(a, T_CD2)[j] = a.cmin[j] * v;
      term(a, T_CMAG2)[j] = fabs(a.cmin[j] * v);
    }
    if (a.flags & F_PID) {
      const double t = static_cast<double>(a.xh[j]) - static_cast<double>(a.x0[j]);
      term(a, T_PIDX)[j] = t * t;
    }
  }
  if (a.flags & F_RAY) block_max_into(a.mpart_n, M_DMAXRAY, dmax_ray, blockIdx.x);
  if (a.flags & F_DRIFT) block_max_into(a.mpart_n, M_DMAXRAY2, dmax_ray2, blockIdx.x);
}

template <class T>
__global__ void __launch_bounds__(kBlock) k_check_m(CheckArgs<T> a) {
  const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  const bool in = i < a.m;
  double rmax = 0, rmax2 = 0;
  if (in) {
    T d = 0;
    if (a.flags & F_RES) {
      d = a.yb[i] - a.yh[i];
      a.dy[i] = d;
      term(a, T_DY2)[i] = static_cast<double>(d) * static_cast<double>(d);
    } else if (a.flags & F_RAY) {
      d = a.dy[i];
    }
    if (a.flags & F_KKT) a.yo[i] = a.rs[i] * static_cast<double>(a.yh[i]) / a.os;  // scaling.cpp:152-155
    const double lo = a.rlo[i], up = a.ruo[i];
    if (a.flags & F_RAY) {
      const double r = dual_ray_entry(static_cast<double>(d), a.rs[i], a.os, lo, up);
      a.rv[i] = r;
      rmax = fabs(r);
      const double t = r > 0 ? lo * r : (r < 0 ? up * r : 0.0);
      term(a, T_ROBJ_ROW)[i] = t;
      term(a, T_RMAG_ROW)[i] = fabs(t);
    }
    if (a.flags & F_DRIFT) {
      const T dd = a.yh[i] - a.y0[i];
      const double r = dual_ray_entry(static_cast<double>(dd), a.rs[i], a.os, lo, up);
      a.rv2[i] = r;
      rmax2 = fabs(r);
      const double t = r > 0 ? lo * r : (r < 0 ? up * r : 0.0);
      term(a, T_ROBJ_ROW2)[i] = t;
      term(a, T_RMAG_ROW2)[i] = fabs(t);
    }
    if (a.flags & F_PID) {
      const double t = static_cast<double>(a.yh[i]) - static_cast<double>(a.y0[i]);
      term(a, T_PIDY)[i] = t * t;
    }
  }
  if (a.flags & F_RAY) block_max_into(a.mpart_m, M_RMAX, rmax, blockIdx.x);
  if (a.flags & F_DRIFT) block_max_into(a.mpart_m, M_RMAX2, rmax2, blockIdx.x);
}

// Dual-ray column terms of λ = −Aᵀr (engine.cpp:98-119).
__device__ inline void ray_col_terms(double g, double gabs, double lo, double up, double& obj, double& mag, double& viol) {
  const double lam = -g;
  obj = 0.0;
  mag = 0.0;
  if (lam > 0) {
    if (isfinite(lo)) obj = lo * lam, mag = gabs * fabs(lo);
    else viol = dmaxd(viol, lam);
  } else if (lam < 0) {
    if (isfinite(up)) obj = up * lam, mag = gabs * fabs(up);
    else viol = dmaxd(viol, -lam);
  }
}

// Row i of the row pass, one part at a time from that part's row sum s: F_RES (Ã·Δx)_i, the
// residual's cross term; F_KKT (A·x)_i, the row part of kkt_general (cpu_backend.cpp:254-270);
// F_RAY / F_DRIFT (A·d)_i, the primal-ray test against the finite row bounds
// (engine.cpp:140-145). lo, up: the row's original bounds (loaded by the caller ahead of the
// sums). The row's maxima are folded into mx.
struct RowMax {
  double pmax = 0.0, dmax = 0.0, pviol = 0.0, pviol2 = 0.0;
};
template <class T>
__device__ inline void check_row(const CheckArgs<T>& a, unsigned part, int i, double lo, double up, double s,
                                 RowMax& mx) {
  if (part == F_RES) {  // Ã·Δx (fp64 accumulate, stored in T), cross term Δyᵀ(ÃΔx)
    const T ax = static_cast<T>(s);
    a.ax[i] = ax;
    term(a, T_CROSS)[i] = static_cast<double>(a.dy[i]) * static_cast<double>(ax);
    return;
  }
  if (part == F_KKT) {
    const double viol = s < lo ? lo - s : (s > up ? s - up : 0.0);
    term(a, T_RP2)[i] = viol * viol;
    mx.pmax = fmax(mx.pmax, nonneg(elem_viol(s, lo, up)));
    const double bl = isfinite(lo) ? fabs(lo) : 0.0, bu = isfinite(up) ? fabs(up) : 0.0;
    term(a, T_B2)[i] = dmaxd(bl, bu) * dmaxd(bl, bu);
    const double y = a.yo[i];
    double dt = 0.0, rd = 0.0, dm = 0.0;
    if (y > 0) {
      if (isfinite(lo)) dt = y * lo;
      else rd = y * y, dm = y;
    } else if (y < 0) {
      if (isfinite(up)) dt = y * up;
      else rd = y * y, dm = -y;
    }
    mx.dmax = fmax(mx.dmax, nonneg(dm));
    term(a, T_DROW)[i] = dt;
    term(a, T_RD2ROW)[i] = rd;
    return;
  }
  double v = 0.0;  // F_RAY or F_DRIFT
  if (isfinite(lo)) v = dmaxd(v, -s);
  if (isfinite(up)) v = dmaxd(v, s);
  if (part == F_RAY) mx.pviol = fmax(mx.pviol, nonneg(v));
  else mx.pviol2 = fmax(mx.pviol2, nonneg(v));
}

// The row pass's maxima of an NT-thread workgroup into partial row `row`; with used < rows the
// rows past the launch's workgroups are cleared (clear_max_rows).
template <int NT, class T>
__device__ inline void row_maxima(const CheckArgs<T>& a, const RowMax& mx, int row, int used, int rows) {
  auto put = [&](int q, double v) {
    block_max_into<NT>(a.mpart_m, q, v, row);
    clear_max_rows(a.mpart_m, q, row, used, rows);
  };
  if (a.flags & F_KKT) {
    put(M_PMAX, mx.pmax);
    put(M_DMAX, mx.dmax);
  }
  if (a.flags & F_RAY) put(M_PVIOL, mx.pviol);
  if (a.flags & F_DRIFT) put(M_PVIOL2, mx.pviol2);
}

// Column j of the column pass, one part at a time from that part's sums: F_KKT s = (Aᵀy)_j, the
// column part of kkt_general (cpu_backend.cpp:271-282); F_RAY / F_DRIFT s = (Aᵀr)_j and
// sa = Σ_i |a_ij r_i|, the dual-ray column terms. lo, up: the column's original bounds (loaded
// by the caller ahead of the sums). The column's maxima are folded into mx.
struct ColMax {
  double pmax = 0.0, dmax = 0.0, cinf = 0.0, dviol = 0.0, dviol2 = 0.0;
};
template <class T>
__device__ inline void check_col(const CheckArgs<T>& a, unsigned part, int j, double lo, double up, double s,
                                 double sa, ColMax& mx) {
  if (part == F_KKT) {
    const double cj = a.cmin[j], xj = a.xo[j];
    term(a, T_C2)[j] = cj * cj;
    mx.cinf = fmax(mx.cinf, nonneg(fabs(cj)));
    term(a, T_P)[j] = cj * xj;
    mx.pmax = fmax(mx.pmax, nonneg(elem_viol(xj, lo, up)));
    const double lam = cj - s;
    double z;
    if (isfinite(lo) && isfinite(up)) z = lam;
    else if (isfinite(lo)) z = dmaxd(lam, 0.0);
    else if (isfinite(up)) z = dmind(lam, 0.0);
    else z = 0.0;
    term(a, T_RD2COL)[j] = (lam - z) * (lam - z);
    mx.dmax = fmax(mx.dmax, nonneg(fabs(lam - z)));
    term(a, T_DCOL)[j] = z > 0 ? z * lo : (z < 0 ? z * up : 0.0);
    return;
  }
  double obj, mag, v = 0.0;
  ray_col_terms(s, sa, lo, up, obj, mag, v);
  if (part == F_RAY) {
    term(a, T_ROBJ_COL)[j] = obj;
    term(a, T_RMAG_COL)[j] = mag;
    mx.dviol = fmax(mx.dviol, nonneg(v));
  } else {
    term(a, T_ROBJ_COL2)[j] = obj;
    term(a, T_RMAG_COL2)[j] = mag;
    mx.dviol2 = fmax(mx.dviol2, nonneg(v));
  }
}

template <int NT, class T>
__device__ inline void col_maxima(const CheckArgs<T>& a, const ColMax& mx, int row, int used, int rows) {
  auto put = [&](int q, double v) {
    block_max_into<NT>(a.mpart_n, q, v, row);
    clear_max_rows(a.mpart_n, q, row, used, rows);
  };
  if (a.flags & F_KKT) {
    put(M_PMAX, mx.pmax);
    put(M_DMAX, mx.dmax);
    put(M_CINF, mx.cinf);
  }
  if (a.flags & F_RAY) put(M_DVIOL, mx.dviol);
  if (a.flags & F_DRIFT) put(M_DVIOL2, mx.dviol2);
}

// The chains of `cs` carried by workgroup blockIdx.x < cs.blocks of a column pass: wavefront
// w < kChainsPerBlock sums one (job, chunk) in its own kSeqBatch doubles of `lds`.
__device__ inline void carry_chains(const ChunkSums& cs, double* lds) {
  const int w = static_cast<int>(threadIdx.x) >> 6;
  if (w >= kChainsPerBlock) return;
  const int chain = static_cast<int>(blockIdx.x) * kChainsPerBlock + w;
  int job = -1;
  for (int b = 0, k = chain >> 6; b < 32; ++b) {
    if (!((cs.mask >> b) & 1u)) continue;
    if (k == 0) {
      job = b;
      break;
    }
    --k;
  }
  if (job < 0) return;
  const RedJob J = cs.jobs[job];
  ordered_chunk<true>(J, cs.partials + static_cast<size_t>(job) * kTreeBlocks, chain & 63, lds + w * kSeqBatch);
}

// Tiles of A (Ã and the original A share the structure): residual, KKT row part and ray tests.
template <class T>
__global__ void __launch_bounds__(kBlock) k_check_rows(CheckArgs<T> a, const int* __restrict__ tiles, int ntiles) {
  const int tile = tile_of_block(blockIdx.x, gridDim.x);
  const bool valid = tile < ntiles;  // padding workgroups still write their (zero) maxima
  const int r0 = valid ? tiles[tile] : 0, r1 = valid ? tiles[tile + 1] : 0;
  const int i = r0 + static_cast<int>(threadIdx.x);
  const bool in = i < r1;
  const double lo = in ? a.rlo[i] : 0.0, up = in ? a.ruo[i] : 0.0;
  double s_res = 0.0, s_kkt = 0.0, s_ray = 0.0, s_dr = 0.0, sa;
  if (a.flags & F_RES) tile_sums<T, T>(r0, r1, a.a_ptr, a.a_col, a.a_val, a.dx, false, s_res, sa);
  if (a.flags & F_KKT) tile_sums<double, double>(r0, r1, a.a_ptr, a.a_col, a.a_orig, a.xo, false, s_kkt, sa);
  if ((a.flags & F_RAY) && a.pray) tile_sums<double, double>(r0, r1, a.a_ptr, a.a_col, a.a_orig, a.dv, false, s_ray, sa);
  if ((a.flags & F_DRIFT) && a.pray)
    tile_sums<double, double>(r0, r1, a.a_ptr, a.a_col, a.a_orig, a.dv2, false, s_dr, sa);
  RowMax mx;
  if (in) {
    if (a.flags & F_RES) check_row(a, F_RES, i, lo, up, s_res, mx);
    if (a.flags & F_KKT) check_row(a, F_KKT, i, lo, up, s_kkt, mx);
    if ((a.flags & F_RAY) && a.pray) check_row(a, F_RAY, i, lo, up, s_ray, mx);
    if ((a.flags & F_DRIFT) && a.pray) check_row(a, F_DRIFT, i, lo, up, s_dr, mx);
  }
  row_maxima<kBlock>(a, mx, blockIdx.x, gridDim.x, gridDim.x);
}

// The row pass over the phased layout of Ã (P.val: Ã, P.valo: A, in the layout's order): the
// tiles' sums in the same order, one sweep per product, then the rows k·kPhBlock + t of each
// workgroup. `rows_max`: the partial rows of the row-side maxima.
template <int K, int J>
__global__ void __launch_bounds__(kPhBlock) k_check_rows_ph(CheckArgs<double> a, Phased P, int rows_max) {
  __shared__ double prod[J * kPhBlock];
  __shared__ double part[K * kPhBlock];
  __shared__ int wtot[kPhBlock / 64];
  const int g = static_cast<int>(blockIdx.x), t = static_cast<int>(threadIdx.x);
  double sum[K], acc[K], unused[K], lo[K], up[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {  // the rows' bounds, loaded ahead of the sweeps
    const int i = g * P.R + k * kPhBlock + t;
    lo[k] = i < a.m ? a.rlo[i] : 0.0;
    up[k] = i < a.m ? a.ruo[i] : 0.0;
  }
  RowMax mx;
  // one sweep per part, each followed by that part's row work (only one part's sums stay live)
  auto sweep = [&](unsigned p, const double* val, const double* x) {
    phased_accumulate<K, J, false>(P, g, val, x, prod, wtot, acc, unused);
    phased_transpose<K>(acc, part, sum);
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int i = g * P.R + k * kPhBlock + t;
      if (i < a.m) check_row(a, p, i, lo[k], up[k], sum[k], mx);
    }
  };
  if (a.flags & F_RES) sweep(F_RES, P.val, a.dx);
  if (a.flags & F_KKT) sweep(F_KKT, P.valo, a.xo);
  if ((a.flags & F_RAY) && a.pray) sweep(F_RAY, P.valo, a.dv);
  if ((a.flags & F_DRIFT) && a.pray) sweep(F_DRIFT, P.valo, a.dv2);
  row_maxima<kPhBlock>(a, mx, g, P.G, rows_max);
}

// Tiles of the original Aᵀ (the model's CSC columns; Ãᵀ's tiles), behind cs.blocks workgroups
// that carry the chunk sums of `cs` in the LDS buffer the tiles stage products in. cs.blocks is a
// multiple of kXcd, so the tiles keep their XCD assignment. The chains' unrolled sums take the
// kernel from 8 to 6 wavefronts per SIMD; holding it at 8 slows the chains more than it speeds
// the column tiles (measured when the 1M-row LP's column pass ran on these tiles: 1.5% more time
// per iteration).
template <class T>
__global__ void __launch_bounds__(kBlock) k_check_cols(CheckArgs<T> a, const int* __restrict__ tiles, int ntiles,
                                                       ChunkSums cs) {
  static_assert(kChainsPerBlock * kSeqBatch <= kTileNnz && kChainsPerBlock <= kBlock / 64, "chain stages");
  __shared__ double lds[kTileNnz];
  if (static_cast<int>(blockIdx.x) < cs.blocks) {
    carry_chains(cs, lds);
    return;
  }
  const int prow = static_cast<int>(blockIdx.x) - cs.blocks;  // this workgroup's row of the max partials
  const int tile = tile_of_block(prow, static_cast<int>(gridDim.x) - cs.blocks);
  const bool valid = tile < ntiles;
  const int r0 = valid ? tiles[tile] : 0, r1 = valid ? tiles[tile + 1] : 0;
  const int j = r0 + static_cast<int>(threadIdx.x);
  const bool in = j < r1;
  const double lo = in ? a.clo[j] : 0.0, up = in ? a.cuo[j] : 0.0;
  double s_kkt = 0.0, s_ray = 0.0, a_ray = 0.0, s_dr = 0.0, a_dr = 0.0, sa;
  if (a.flags & F_KKT) tile_sums_in<double, double>(lds, r0, r1, a.at_ptr, a.ato_col, a.ato_val, a.yo, false, s_kkt, sa);
  if (a.flags & F_RAY) tile_sums_in<double, double>(lds, r0, r1, a.at_ptr, a.ato_col, a.ato_val, a.rv, true, s_ray, a_ray);
  if (a.flags & F_DRIFT)
    tile_sums_in<double, double>(lds, r0, r1, a.at_ptr, a.ato_col, a.ato_val, a.rv2, true, s_dr, a_dr);
  ColMax mx;
  if (in) {
    if (a.flags & F_KKT) check_col(a, F_KKT, j, lo, up, s_kkt, 0.0, mx);
    if (a.flags & F_RAY) check_col(a, F_RAY, j, lo, up, s_ray, a_ray, mx);
    if (a.flags & F_DRIFT) check_col(a, F_DRIFT, j, lo, up, s_dr, a_dr, mx);
  }
  const int used = static_cast<int>(gridDim.x) - cs.blocks;
  col_maxima<kBlock>(a, mx, prow, used, used);
}

// The column pass over the phased layout of Ãᵀ (P.valo: the model's CSC values, in the layout's
// order; built only when the CSC's row order is Ãᵀ's), behind cs.blocks workgroups that carry the
// chunk sums of `cs` (4 of the 8 wavefronts each). `rows_max`: the partial rows of the
// column-side maxima.
template <int K, int J>
__global__ void __launch_bounds__(kPhBlock) k_check_cols_ph(CheckArgs<double> a, Phased P, int rows_max, ChunkSums cs) {
  constexpr int kProd = (J + K) * kPhBlock, kStage = kChainsPerBlock * kSeqBatch;
  __shared__ double lds[kProd > kStage ? kProd : kStage];
  __shared__ int wtot[kPhBlock / 64];
  if (static_cast<int>(blockIdx.x) < cs.blocks) {
    carry_chains(cs, lds);
    return;
  }
  double* const prod = lds;
  double* const part = lds + J * kPhBlock;
  const int g = static_cast<int>(blockIdx.x) - cs.blocks, t = static_cast<int>(threadIdx.x);
  double sum[K], asum[K], acc[K], acca[K];
  ColMax mx;
  // one sweep per part, each followed by that part's column work
  auto sweep = [&](unsigned p, const double* x, auto abs_tag) {
    constexpr bool kAbs = decltype(abs_tag)::value;
    phased_accumulate<K, J, kAbs>(P, g, P.valo, x, prod, wtot, acc, acca);
    phased_transpose<K>(acc, part, sum);
    if constexpr (kAbs) phased_transpose<K>(acca, part, asum);
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int j = g * P.R + k * kPhBlock + t;
      if (j < a.n) check_col(a, p, j, a.clo[j], a.cuo[j], sum[k], kAbs ? asum[k] : 0.0, mx);
    }
  };
  if (a.flags & F_KKT) sweep(F_KKT, a.yo, std::false_type{});
  if (a.flags & F_RAY) sweep(F_RAY, a.rv, std::true_type{});
  if (a.flags & F_DRIFT) sweep(F_DRIFT, a.rv2, std::true_type{});
  col_maxima<kPhBlock>(a, mx, g, P.G, rows_max);
}

// Maxima of the per-workgroup partials of both sides into out[0 .. kMaxCount).
__global__ void __launch_bounds__(kBlock) k_max_final(const double* __restrict__ pm, int nbm, const double* __restrict__ pn,
                                                      int nbn, double* __restrict__ out) {
  __shared__ double red[kBlock];
  const int q = blockIdx.x;
  double v = 0.0;
  for (int b = threadIdx.x; b < nbm; b += kBlock) v = fmax(v, pm[static_cast<size_t>(b) * kMaxCount + q]);
  for (int b = threadIdx.x; b < nbn; b += kBlock) v = fmax(v, pn[static_cast<size_t>(b) * kMaxCount + q]);
  red[threadIdx.x] = v;
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if (static_cast<int>(threadIdx.x) < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) out[q] = red[0];
}

__global__ void k_zero(double* p, int n) {
  if (static_cast<int>(threadIdx.x) < n) p[threadIdx.x] = 0.0;
}

template <class T>
T* sp(psh_ctx* c, int s) {
  if constexpr (sizeof(T) == 8) return c->vd[s];
  else return c->vf[s];
}

template <class T>
CheckArgs<T> make_args(psh_ctx* c, unsigned flags, int xh, int xb, int x0, int dx, int yh, int yb, int y0, int dy) {
  CheckArgs<T> a{};
  a.flags = flags;
  a.m = c->m;
  a.n = c->n;
  a.pray = c->boxed ? 0 : 1;
  a.xh = sp<T>(c, xh);
  a.xb = sp<T>(c, xb);
  a.x0 = sp<T>(c, x0);
  a.dx = sp<T>(c, dx);
  a.yh = sp<T>(c, yh);
  a.yb = sp<T>(c, yb);
  a.y0 = sp<T>(c, y0);
  a.dy = sp<T>(c, dy);
  a.ax = sp<T>(c, PSH_AX);
  a.a_ptr = c->A.ptr;
  a.a_col = c->A.col;
  a.at_ptr = c->At.ptr;
  a.ato_col = c->ato_col;
  if constexpr (sizeof(T) == 8) a.a_val = c->A.val;
  else a.a_val = c->A.valf;
  a.a_orig = c->a_orig;
  a.ato_val = c->ato_val;
  a.rs = c->rs;
  a.cs = c->cs;
  a.clo = c->clo;
  a.cuo = c->cuo;
  a.rlo = c->rlo;
  a.ruo = c->ruo;
  a.cmin = c->cmin;
  a.bs = c->bs;
  a.os = c->os;
  a.xo = c->xo;
  a.yo = c->yo;
  a.rv = c->rv;
  a.dv = c->dv;
  a.rv2 = c->rv2;
  a.dv2 = c->dv2;
  a.terms = c->terms;
  a.tl = c->term_len;
  a.mpart_m = c->mpart_m;
  a.mpart_n = c->mpart_n;
  return a;
}

// Launches the four check kernels for `flags`; the dual-side pass runs before the row pass
// because the row pass reads Δy, ŷ (unscaled) and the dual ray. The column pass carries the
// chunk sums of `cs` when cs.blocks > 0 (they read only what the passes before it write).
template <class T>
int launch_check(psh_ctx* c, const CheckArgs<T>& a, bool maxes, const ChunkSums& cs = ChunkSums{}) {
  const unsigned nf = a.flags & (F_RES | F_KKT | F_RAY | F_DRIFT | F_PID);
  const unsigned mf = nf;
  if (nf && c->n > 0) hipLaunchKernelGGL((k_check_n<T>), dim3(grid_for(c->n)), dim3(kBlock), 0, c->stream, a);
  if (mf && c->m > 0) hipLaunchKernelGGL((k_check_m<T>), dim3(grid_for(c->m)), dim3(kBlock), 0, c->stream, a);
  // fp64 checks use the phased layouts where they are built (the original values in P.valo)
  bool ph_rows = false, ph_cols = false;
  if constexpr (sizeof(T) == 8) {
    ph_rows = c->sliced && c->Pa.S > 0 && c->Pa.valo;
    ph_cols = c->sliced && c->Pat.S > 0 && c->Pat.valo;
  }
  if ((a.flags & (F_RES | F_KKT | F_RAY | F_DRIFT)) && c->A.ntiles > 0) {
    if constexpr (sizeof(T) == 8) {
      if (ph_rows)
        phased_dispatch(c->Pa, [&](auto kc, auto jc) {
          hipLaunchKernelGGL((k_check_rows_ph<decltype(kc)::value, decltype(jc)::value>), dim3(c->Pa.G), dim3(kPhBlock),
                             0, c->stream, a, c->Pa, c->mpart_rows_m);
        });
    }
    if (!ph_rows)
      hipLaunchKernelGGL((k_check_rows<T>), dim3(grid_for_tiles(c->A.ntiles)), dim3(kBlock), 0, c->stream, a,
                         c->A.tiles, c->A.ntiles);
  }
  const bool cols = (a.flags & (F_KKT | F_RAY | F_DRIFT)) && c->At.ntiles > 0;
  ChunkSums carry = cs;
  if (carry.blocks < 0) carry.blocks = 0;
  if constexpr (sizeof(T) == 8) {
    if (cols && ph_cols)
      phased_dispatch(c->Pat, [&](auto kc, auto jc) {
        hipLaunchKernelGGL((k_check_cols_ph<decltype(kc)::value, decltype(jc)::value>), dim3(carry.blocks + c->Pat.G),
                           dim3(kPhBlock), 0, c->stream, a, c->Pat, c->mpart_rows_n, carry);
      });
  }
  if ((cols && !ph_cols) || (!cols && carry.blocks > 0))
    hipLaunchKernelGGL((k_check_cols<T>), dim3(carry.blocks + (cols ? grid_for_tiles(c->At.ntiles) : 0)), dim3(kBlock),
                       0, c->stream, a, c->At.tiles, cols ? c->At.ntiles : 0, carry);
  if (maxes)
    hipLaunchKernelGGL(k_max_final, dim3(kMaxCount), dim3(kBlock), 0, c->stream, c->mpart_m, c->mpart_rows_m,
                       c->mpart_n, c->mpart_rows_n, c->results + R_MAX0);
  PSH_LAUNCHED(c, "check kernels");
  return PSH_OK;
}

template <class F>
int with_t(psh_ctx* c, F&& f) {
  if (c->fp32) return f(static_cast<float*>(nullptr));
  return f(static_cast<double*>(nullptr));
}

double max_of(const double* r, int k) { return r[R_MAX0 + k]; }

void fill_kkt(const double* r, psh_kkt* k) {  // termination.h quantities from the sums
  k->primal_residual = std::sqrt(r[R_RP2]);
  k->dual_residual = std::sqrt(r[R_RD2]);
  k->primal_obj = r[R_P];
  k->dual_obj = r[R_D];
  k->b_norm = std::sqrt(r[R_B2]);
  k->c_norm = std::sqrt(r[R_C2]);
  k->primal_max_rel = max_of(r, M_PMAX);
  k->dual_max_rel = max_of(r, M_DMAX) / (1.0 + max_of(r, M_CINF));
}

void fill_ray(const double* r, int second, psh_ray* o) {
  o->rmax = max_of(r, second ? M_RMAX2 : M_RMAX);
  o->dual_obj = r[second ? R_ROBJ2 : R_ROBJ];
  o->dual_mag = r[second ? R_RMAG2 : R_RMAG];
  o->dual_viol = max_of(r, second ? M_DVIOL2 : M_DVIOL);
  o->dmax = max_of(r, second ? M_DMAXRAY2 : M_DMAXRAY);
  o->primal_cd = r[second ? R_CD2 : R_CD];
  o->primal_cmag = r[second ? R_CMAG2 : R_CMAG];
  o->primal_viol = max_of(r, second ? M_PVIOL2 : M_PVIOL);
}

// Anchor constants for the increment form: fp64 products of the anchors in X0 / Y0 (into the
// AX / ATY slots), then the fp32 constants; δ restarts from 0 when the anchor moved.
int delta_prepare(psh_ctx* c, double sigma) {
  DeltaState& d = c->delta;
  if (!d.dirty && sigma == d.sigma) return PSH_OK;
  int rc;
  if ((rc = psh_spmv_t(c, PSH_Y0, PSH_ATY)) || (rc = psh_spmv(c, PSH_X0, PSH_AX))) return rc;
  const int reset = d.dirty ? 1 : 0;
  if (c->n > 0)
    hipLaunchKernelGGL(k_delta_anchor_n, dim3(grid_for(c->n)), dim3(kBlock), 0, c->stream, c->c, c->l, c->u,
                       c->vd[PSH_X0], c->vd[PSH_ATY], d.G, d.LX, d.UX, d.dx, d.dxh, d.dxb, (long long)c->n, reset);
  if (c->m > 0)
    hipLaunchKernelGGL(k_delta_anchor_m, dim3(grid_for(c->m)), dim3(kBlock), 0, c->stream, c->rl, c->ru,
                       c->vd[PSH_AX], c->vd[PSH_Y0], sigma, d.BL, d.BU, d.KY, d.dy, d.dyh, d.dyb, (long long)c->m, reset);
  PSH_LAUNCHED(c, "increment-form anchor");
  d.dirty = false;
  d.sigma = sigma;
  return PSH_OK;
}

// x̂, x̄, ŷ, ȳ (and x, y when `cur`) in fp64 from the anchors and the increments.
int delta_points(psh_ctx* c, bool cur) {
  DeltaState& d = c->delta;
  if (c->n > 0)
    hipLaunchKernelGGL(k_delta_point, dim3(grid_for(c->n)), dim3(kBlock), 0, c->stream, c->vd[PSH_X0], d.dxh, d.dxb,
                       d.dx, c->vd[PSH_XH], c->vd[PSH_XB], cur ? c->vd[PSH_X] : nullptr, (long long)c->n);
  if (c->m > 0)
    hipLaunchKernelGGL(k_delta_point, dim3(grid_for(c->m)), dim3(kBlock), 0, c->stream, c->vd[PSH_Y0], d.dyh, d.dyb,
                       d.dy, c->vd[PSH_YH], c->vd[PSH_YB], cur ? c->vd[PSH_Y] : nullptr, (long long)c->m);
  PSH_LAUNCHED(c, "increment-form points");
  return PSH_OK;
}

// Residual sums at the epoch's anchor (after its first step) into R_R0, on the device.
template <class T>
int epoch_residual(psh_ctx* c, int mode) {
  const CheckArgs<T> a = make_args<T>(c, F_RES, PSH_XH, PSH_XB, PSH_X0, PSH_DX, PSH_YH, PSH_YB, PSH_Y0, PSH_DY);
  int r;
  if ((r = launch_check<T>(c, a, false))) return r;
  int err = PSH_OK;
  const JobTable* t = check_table(c, mode, F_RES, &err);
  if (!t) return err;
  JobTable r0;
  r0.dev = t->dev;
  r0.host.assign(t->host.begin(), t->host.begin() + 3);
  return reduce_table(c, r0, R_R0);
}

int check_slots(psh_ctx* c) {
  if (!c) return PSH_E_ARG;
  if (!c->loaded) return fail(c, PSH_E_STATE, "no problem loaded");
  return PSH_OK;
}

}  // namespace

extern "C" {

int psh_chunk(psh_ctx* c, const psh_chunk_args* g) {
  int rc = check_slots(c);
  if (rc) return rc;
  if (!g || g->iterations <= 0) return fail(c, PSH_E_ARG, "chunk needs at least one iteration");
  if (g->inner0 < 0 || g->inner0 > (1LL << 52)) return fail(c, PSH_E_ARG, "epoch counter out of range");
  const bool epoch_start = g->inner0 == 0;
  const bool hat_last = g->want_hat_last != 0;
  if (epoch_start) {  // the residual's reduction table must exist before any capture
    int err = PSH_OK;
    if (!check_table(c, g->mode, F_RES, &err)) return err;
  }
  const bool delta = c->delta.on;
  // phased products for fp64 chunks of large problems (the same sums as the row tiles)
  const bool ph_primal = c->sliced && c->Pat.S > 0 && !c->fp32 && !delta;
  const bool ph_dual = c->sliced && c->Pa.S > 0 && !c->fp32 && !delta;
  if (delta && c->nnz > 0 &&
      !(c->delta.bf16 ? c->A.valh && c->At.valh : c->A.valf && c->At.valf))  // the launches' value arrays
    return fail(c, PSH_E_STATE, "increment form without its matrix values");
  if (delta && (rc = delta_prepare(c, g->sigma))) return rc;
  hipLaunchKernelGGL(k_set_params, dim3(1), dim3(64), 0, c->stream, c->prm, g->tau, g->sigma,
                     static_cast<double>(g->inner0));
  PSH_LAUNCHED(c, "chunk parameters");
  auto enqueue_delta = [&]() -> int {
    DeltaState& d = c->delta;
    const float a2 = static_cast<float>(2 * g->rho), b2 = static_cast<float>(1 - 2 * g->rho);
    for (int s = 0; s < g->iterations; ++s) {
      const bool first = s == 0 && epoch_start;
      const bool last_hat = s == g->iterations - 1 && hat_last;
      const int hat = first || last_hat ? 1 : 0;
      const DeltaP pa{d.dx, d.dxh, d.dxb, d.G, d.LX, d.UX, c->prm, s, a2, b2, hat};
      const DeltaD da{d.dy, d.dyh, d.dyb, d.BL, d.BU, d.KY, c->prm, s, a2, b2, hat};
      hipEvent_t ev;
      if (c->At.ntiles > 0) {
        time_begin(c, 0, &ev);
        if (d.bf16)
          hipLaunchKernelGGL(k_delta_primal<Bf16>, dim3(grid_for_tiles(c->At.ntiles)), dim3(kBlock), 0, c->stream,
                             c->At.tiles, c->At.ntiles, c->At.ptr, c->At.col, c->At.valh, d.dy, pa);
        else
          hipLaunchKernelGGL(k_delta_primal<float>, dim3(grid_for_tiles(c->At.ntiles)), dim3(kBlock), 0, c->stream,
                             c->At.tiles, c->At.ntiles, c->At.ptr, c->At.col, c->At.valf, d.dy, pa);
        time_end(c, 0, ev);
      }
      if (c->A.ntiles > 0) {
        time_begin(c, 1, &ev);
        if (d.bf16)
          hipLaunchKernelGGL(k_delta_dual<Bf16>, dim3(grid_for_tiles(c->A.ntiles)), dim3(kBlock), 0, c->stream,
                             c->A.tiles, c->A.ntiles, c->A.ptr, c->A.col, c->A.valh, d.dxb, da);
        else
          hipLaunchKernelGGL(k_delta_dual<float>, dim3(grid_for_tiles(c->A.ntiles)), dim3(kBlock), 0, c->stream,
                             c->A.tiles, c->A.ntiles, c->A.ptr, c->A.col, c->A.valf, d.dxb, da);
        time_end(c, 1, ev);
      }
      PSH_LAUNCHED(c, "increment-form iteration kernels");
      int r;
      if (hat && (r = delta_points(c, false))) return r;
      if (first && (r = epoch_residual<double>(c, g->mode))) return r;
    }
    return PSH_OK;
  };
  // The launches of one chunk depend only on (precision, length, epoch start, last-step hats,
  // reduction mode): replay a captured graph unless per-kernel timing is on.
  auto enqueue = [&]() -> int {
    return with_t(c, [&](auto* tag) -> int {
      using T = std::remove_pointer_t<decltype(tag)>;
      const T* at_val;
      const T* a_val;
      if constexpr (sizeof(T) == 8) at_val = c->At.val, a_val = c->A.val;
      else at_val = c->At.valf, a_val = c->A.valf;
      const T *cc, *ll, *uu, *rl, *ru;
      if constexpr (sizeof(T) == 8) cc = c->c, ll = c->l, uu = c->u, rl = c->rl, ru = c->ru;
      else cc = c->cf, ll = c->lf, uu = c->uf, rl = c->rlf, ru = c->ruf;
      const T a2 = static_cast<T>(2 * g->rho), b2 = static_cast<T>(1 - 2 * g->rho);
      for (int s = 0; s < g->iterations; ++s) {
        const bool first = s == 0 && epoch_start;
        const int hat = first || (s == g->iterations - 1 && hat_last) ? 1 : 0;
        PrimalArgs<T> pa{sp<T>(c, PSH_X), sp<T>(c, PSH_X0), cc, ll, uu, sp<T>(c, PSH_XH), sp<T>(c, PSH_XB), c->prm, s,
                         a2, b2, hat};
        DualArgs<T> da{sp<T>(c, PSH_Y), sp<T>(c, PSH_Y0), rl, ru, sp<T>(c, PSH_YH), sp<T>(c, PSH_YB), c->prm, s,
                       a2, b2, hat};
        hipEvent_t ev;
        if (c->At.ntiles > 0) {
          time_begin(c, 0, &ev);
          bool done = false;
          if constexpr (sizeof(T) == 8) {
            if (ph_primal) {
              phased_dispatch(c->Pat, [&](auto kc, auto jc) {
                hipLaunchKernelGGL((k_ph_primal<T, decltype(kc)::value, decltype(jc)::value>), dim3(c->Pat.G),
                                   dim3(kPhBlock), 0, c->stream, c->Pat, c->n, c->Pat.val, sp<T>(c, PSH_Y), pa);
              });
              done = true;
            }
          }
          if (!done)
            hipLaunchKernelGGL((k_iter_primal<T>), dim3(grid_for_tiles(c->At.ntiles)), dim3(kBlock), 0, c->stream,
                               c->At.tiles, c->At.ntiles, c->At.ptr, c->At.col, at_val, sp<T>(c, PSH_Y), pa);
          time_end(c, 0, ev);
        }
        if (c->A.ntiles > 0) {
          time_begin(c, 1, &ev);
          bool done = false;
          if constexpr (sizeof(T) == 8) {
            if (ph_dual) {
              phased_dispatch(c->Pa, [&](auto kc, auto jc) {
                hipLaunchKernelGGL((k_ph_dual<T, decltype(kc)::value, decltype(jc)::value>), dim3(c->Pa.G),
                                   dim3(kPhBlock), 0, c->stream, c->Pa, c->m, c->Pa.val, sp<T>(c, PSH_XB), da);
              });
              done = true;
            }
          }
          if (!done)
            hipLaunchKernelGGL((k_iter_dual<T>), dim3(grid_for_tiles(c->A.ntiles)), dim3(kBlock), 0, c->stream,
                               c->A.tiles, c->A.ntiles, c->A.ptr, c->A.col, a_val, sp<T>(c, PSH_XB), da);
          time_end(c, 1, ev);
        }
        PSH_LAUNCHED(c, "iteration kernels");
        if (first) {  // residual at the epoch's anchor (r2hpdhg.cpp:75), kept on the device
          const int r = epoch_residual<T>(c, g->mode);
          if (r) return r;
        }
      }
      return PSH_OK;
    });
  };
  if (c->timing || !c->graphs) {
    if ((rc = delta ? enqueue_delta() : enqueue())) return rc;
  } else {
    const long long key = (static_cast<long long>(g->iterations) << 9) | (c->fp32 ? 1 : 0) | (epoch_start ? 2 : 0) |
                          (hat_last ? 4 : 0) | (static_cast<long long>(g->mode & 7) << 3) | (delta ? 1LL << 6 : 0) |
                          (ph_primal || ph_dual ? 1LL << 7 : 0) | (delta && c->delta.bf16 ? 1LL << 8 : 0);
    if (g->rho != c->graph_rho) {  // ρ is an argument of the captured launches: recapture
      PSH_TRY(c, hipStreamSynchronize(c->stream));
      for (auto& kv : c->graph_cache) (void)hipGraphExecDestroy(kv.second);
      c->graph_cache.clear();
      c->graph_rho = g->rho;
    }
    auto it = c->graph_cache.find(key);
    if (it == c->graph_cache.end()) {
      std::lock_guard<std::recursive_mutex> capture_lock(api_mutex());  // no device-wide calls meanwhile
      PSH_TRY(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
      const int erc = delta ? enqueue_delta() : enqueue();
      hipGraph_t graph = nullptr;
      const hipError_t ec = hipStreamEndCapture(c->stream, &graph);
      if (erc) {
        if (graph) (void)hipGraphDestroy(graph);
        return erc;
      }
      if (ec != hipSuccess) return hip_fail(c, ec, "hipStreamEndCapture");
      hipGraphExec_t exec = nullptr;
      const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      if (ei != hipSuccess) return hip_fail(c, ei, "hipGraphInstantiate");
      it = c->graph_cache.emplace(key, exec).first;
    }
    PSH_TRY(c, hipGraphLaunch(it->second, c->stream));
  }
  if (epoch_start) c->have_r0 = true;
  return PSH_OK;
}

int psh_graphs(psh_ctx* c, int on) {
  if (!c) return PSH_E_ARG;
  c->graphs = on != 0;
  return PSH_OK;
}

int psh_sliced(psh_ctx* c, int on) {
  if (!c) return PSH_E_ARG;
  if (on < 0 || on > 2) return fail(c, PSH_E_ARG, "psh_sliced: on must be 0, 1 or 2");
  c->sliced = on;
  return PSH_OK;
}

int psh_delta(psh_ctx* c, int on) {
  int rc = check_slots(c);
  if (rc) return rc;
  if (on < 0 || on > 2) return fail(c, PSH_E_ARG, "psh_delta: on must be 0, 1 or 2");
  DeltaState& d = c->delta;
  const bool want = on != 0;
  if (want && d.on) {  // switching the value type only: its copy of Ã must exist first
    if ((rc = on == 2 ? ensure_bf16_matrix(c) : ensure_fp32_matrix(c))) return rc;
    d.bf16 = on == 2;
    return PSH_OK;
  }
  if (want == d.on) return PSH_OK;
  if (want) {
    if (c->fp32) return fail(c, PSH_E_STATE, "fp32 storage and the increment form are exclusive");
    if ((rc = on == 2 ? ensure_bf16_matrix(c) : ensure_fp32_matrix(c))) return rc;
    d.bf16 = on == 2;
    if (!d.dx) {
      for (float** v : {&d.dx, &d.dxh, &d.dxb, &d.G, &d.LX, &d.UX})
        if ((rc = alloc_n(c, v, c->n))) return rc;
      for (float** v : {&d.dy, &d.dyh, &d.dyb, &d.BL, &d.BU, &d.KY})
        if ((rc = alloc_n(c, v, c->m))) return rc;
    }
    d.on = true;
    d.dirty = true;  // anchors are read from X0 / Y0 at the next chunk
    return PSH_OK;
  }
  // leaving: the current point, x̂, x̄, ŷ, ȳ back into the fp64 slots (only valid after a chunk)
  if (!d.dirty && (rc = delta_points(c, true))) return rc;
  d.on = false;
  return PSH_OK;
}

int psh_check(psh_ctx* c, const psh_check_args* g, psh_stats* out) {
  int rc = check_slots(c);
  if (rc) return rc;
  if (!g || !out) return fail(c, PSH_E_ARG, "null check arguments");
  unsigned flags = F_RES | F_KKT;
  if (g->want_ray) flags |= F_RAY;
  if (g->want_drift) flags |= F_DRIFT;
  if (g->want_pid) flags |= F_PID;
  int err = PSH_OK;
  const JobTable* t = check_table(c, g->mode, flags, &err);
  if (!t) return err;
  // In decision mode the 64-chunk sums (residual, PID distances) read only what the elementwise
  // passes and the row pass write; they are latency-bound (one sequential chain per chunk), so
  // the column pass carries them and only their in-order combine runs after it.
  ChunkSums cs;
  if (g->mode == PSH_RED_DECISION) cs = chunk_sums_of(c, *t);
  const bool carried = cs.blocks > 0;
  rc = with_t(c, [&](auto* tag) -> int {
    using T = std::remove_pointer_t<decltype(tag)>;
    const CheckArgs<T> a = make_args<T>(c, flags, PSH_XH, PSH_XB, PSH_X0, PSH_DX, PSH_YH, PSH_YB, PSH_Y0, PSH_DY);
    hipEvent_t ev;
    time_begin(c, 2, &ev);
    int r = launch_check<T>(c, a, true, carried ? cs : ChunkSums{});
    time_end(c, 2, ev);
    return r;
  });
  if (rc) return rc;
  if ((rc = carried ? reduce_table_part(c, *t, 0, 1 | 4) : reduce_table(c, *t, 0))) return rc;
  double r[kResultCount];
  if ((rc = fetch_results(c, kResultCount, r))) return rc;
  out->px = r[R_PX];
  out->py = r[R_PY];
  out->cross = r[R_CROSS];
  out->have_r0 = c->have_r0 ? 1 : 0;
  out->r0_px = r[R_R0];
  out->r0_py = r[R_R0 + 1];
  out->r0_cross = r[R_R0 + 2];
  c->have_r0 = false;
  fill_kkt(r, &out->kkt);
  fill_ray(r, 0, &out->ray);
  fill_ray(r, 1, &out->drift);
  out->dx_norm = std::sqrt(r[R_PIDX]);
  out->dy_norm = std::sqrt(r[R_PIDY]);
  bool fin = true;
  for (int k = 0; k < kSumCount; ++k) fin = fin && std::isfinite(r[k]);
  out->finite = fin ? 1 : 0;
  return PSH_OK;
}

int psh_kkt_of(psh_ctx* c, int xs, int ys, int mode, psh_kkt* out) {
  int rc = check_slots(c);
  if (rc) return rc;
  if (!out || xs < 0 || xs >= PSH_NSLOT || ys < 0 || ys >= PSH_NSLOT || !is_primal_slot(xs) || is_primal_slot(ys))
    return fail(c, PSH_E_ARG, "kkt needs a primal and a dual slot");
  rc = with_t(c, [&](auto* tag) -> int {
    using T = std::remove_pointer_t<decltype(tag)>;
    const CheckArgs<T> a = make_args<T>(c, F_KKT, xs, PSH_XB, PSH_X0, PSH_DX, ys, PSH_YB, PSH_Y0, PSH_DY);
    return launch_check<T>(c, a, true);
  });
  if (rc) return rc;
  int err = PSH_OK;
  const JobTable* t = check_table(c, mode, F_KKT, &err);
  if (!t) return err;
  if ((rc = reduce_table(c, *t, 0))) return rc;
  double r[kResultCount];
  if ((rc = fetch_results(c, kResultCount, r))) return rc;
  fill_kkt(r, out);
  return PSH_OK;
}

int psh_ray_of(psh_ctx* c, int dxs, int dys, psh_ray* out) {
  int rc = check_slots(c);
  if (rc) return rc;
  if (!out || dxs < 0 || dxs >= PSH_NSLOT || dys < 0 || dys >= PSH_NSLOT || !is_primal_slot(dxs) || is_primal_slot(dys))
    return fail(c, PSH_E_ARG, "ray test needs a primal and a dual slot");
  rc = with_t(c, [&](auto* tag) -> int {
    using T = std::remove_pointer_t<decltype(tag)>;
    const CheckArgs<T> a = make_args<T>(c, F_RAY, PSH_XH, PSH_XB, PSH_X0, dxs, PSH_YH, PSH_YB, PSH_Y0, dys);
    return launch_check<T>(c, a, true);
  });
  if (rc) return rc;
  int err = PSH_OK;
  const JobTable* t = check_table(c, PSH_RED_TREE, F_RAY, &err);
  if (!t) return err;
  if ((rc = reduce_table(c, *t, 0))) return rc;
  double r[kResultCount];
  if ((rc = fetch_results(c, kResultCount, r))) return rc;
  fill_ray(r, 0, out);
  return PSH_OK;
}

int psh_ray_vectors(psh_ctx* c, int drift, double* dual_ray, double* primal_ray) {
  int rc = check_slots(c);
  if (rc) return rc;
  if (dual_ray && c->m > 0)
    PSH_TRY(c, hipMemcpyAsync(dual_ray, drift ? c->rv2 : c->rv, sizeof(double) * c->m, hipMemcpyDeviceToHost, c->stream));
  if (primal_ray && c->n > 0)
    PSH_TRY(c, hipMemcpyAsync(primal_ray, drift ? c->dv2 : c->dv, sizeof(double) * c->n, hipMemcpyDeviceToHost, c->stream));
  PSH_TRY(c, hipStreamSynchronize(c->stream));
  return PSH_OK;
}

int psh_restart(psh_ctx* c) {
  int rc = check_slots(c);
  if (rc) return rc;
  if (c->delta.on) c->delta.dirty = true;  // the fp64 slots hold the new anchor
  return with_t(c, [&](auto* tag) -> int {
    using T = std::remove_pointer_t<decltype(tag)>;
    if (c->n > 0)
      hipLaunchKernelGGL((k_restart<T>), dim3(grid_for(c->n)), dim3(kBlock), 0, c->stream, sp<T>(c, PSH_X),
                         sp<T>(c, PSH_X0), sp<T>(c, PSH_XH), (long long)c->n);
    if (c->m > 0)
      hipLaunchKernelGGL((k_restart<T>), dim3(grid_for(c->m)), dim3(kBlock), 0, c->stream, sp<T>(c, PSH_Y),
                         sp<T>(c, PSH_Y0), sp<T>(c, PSH_YH), (long long)c->m);
    PSH_LAUNCHED(c, "restart");
    return PSH_OK;
  });
}

}  // extern "C"
