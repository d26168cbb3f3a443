// psh_batch.hip — lockstep r²HPDHG over W scenarios sharing Ã (reference src/pdhg/batch.cpp).
//
// Block vectors are interleaved, entry (row r, slot t) at r·W + t, so one thread per (row, slot)
// with consecutive threads on consecutive slots turns every gather of the SpMM into a contiguous
// W-element read, and each nonzero of Ã / Ãᵀ is fetched once per iteration for all scenarios.
// Each thread sums its (row, slot) product sequentially in CSR order and runs the fused update —
// the reference's per-scenario expressions (batch.cpp:319-346), so the bits are the reference's.
// Per-slot sums (residual, KKT, PID) go through a segmented reducer: fixed-shape trees, or one
// sequential loop per slot in the reference order (batch.cpp:293-305, 390-400; kkt_general).
#include <algorithm>
#include <cmath>
#include <cstring>

#include "psh_internal.h"

using namespace psh;

namespace {

// Batch term arrays.
enum BTerm {
  B_DX2, B_DY2, B_CROSS, B_RP2, B_B2, B_DROW, B_RD2ROW, B_C2, B_P, B_DCOL, B_RD2COL, B_PIDX, B_PIDY,
  B_PMAXR, B_PMAXC, B_DMAXR, B_DMAXC, B_CINF,
  // ray tests (engine.cpp:82-184 of the reference, per slot)
  B_ROBJR, B_RMAGR, B_ROBJC, B_RMAGC, B_CD, B_CMAG, B_RMAX, B_DVIOL, B_DMAXRAY, B_PVIOL, kBTerms
};
constexpr int kBJobs = 24;
constexpr int kBTreeBlocks = 512;  // level-1 blocks per job: enough workgroups for 200k rows × 64 slots

struct BArgs {
  int m, n, W, it, hat;
  double rho2, rho1;  // 2ρ, 1 − 2ρ
  const int *a_ptr, *a_col, *at_ptr, *at_col, *ato_col;
  const double *a_val, *at_val, *a_orig, *ato_val;
  const float *a_valf, *at_valf;  // increment form
  const double *rs, *cs;
  BatchArrays b;
};

__device__ inline double clampd(double v, double lo, double hi) { return v < lo ? lo : (v > hi ? hi : v); }
__device__ inline double dmaxd(double a, double b) { return a < b ? b : a; }
__device__ inline double dmind(double a, double b) { return b < a ? b : a; }

__device__ inline double* bterm(const BArgs& a, int k) {
  return a.b.terms + static_cast<size_t>(k) * static_cast<size_t>(a.b.rows_max) * a.b.cap;
}

__device__ inline double halpern_w(long long inner0, int it) {
  const long long k = inner0 + it;
  return static_cast<double>(k + 1) / static_cast<double>(k + 2);
}

__global__ void __launch_bounds__(kBlock) k_bprimal(BArgs a) {
  const long long g = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  const long long j = g / a.W;
  if (j >= a.n) return;
  const int t = static_cast<int>(g - j * a.W);
  double s = 0.0;
  for (int p = a.at_ptr[j]; p < a.at_ptr[j + 1]; ++p) s += a.at_val[p] * a.b.y[static_cast<size_t>(a.at_col[p]) * a.W + t];
  const size_t q = static_cast<size_t>(g);
  const double w = halpern_w(a.b.inner0[t], a.it);
  const double xq = a.b.x[q];
  const double h = clampd(xq - a.b.tau[t] * (a.b.c[q] - s), a.b.l[q], a.b.u[q]);
  if (a.hat) a.b.xh[q] = h;
  a.b.xb[q] = 2 * h - xq;
  a.b.x[q] = w * (a.rho2 * h + a.rho1 * xq) + (1 - w) * a.b.x0[q];
}

__global__ void __launch_bounds__(kBlock) k_bdual(BArgs a) {
  const long long g = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  const long long i = g / a.W;
  if (i >= a.m) return;
  const int t = static_cast<int>(g - i * a.W);
  double s = 0.0;
  for (int p = a.a_ptr[i]; p < a.a_ptr[i + 1]; ++p) s += a.a_val[p] * a.b.xb[static_cast<size_t>(a.a_col[p]) * a.W + t];
  const size_t q = static_cast<size_t>(g);
  const double w = halpern_w(a.b.inner0[t], a.it);
  const double sg = a.b.sigma[t];
  const double yq = a.b.y[q];
  const double v = s - yq / sg;
  const double h = yq - sg * s + sg * clampd(v, a.b.rl[q], a.b.ru[q]);
  if (a.hat) {
    a.b.yh[q] = h;
    a.b.yb[q] = 2 * h - yq;
  }
  a.b.y[q] = w * (a.rho2 * h + a.rho1 * yq) + (1 - w) * a.b.y0[q];
}

// fp64 steps for an even batch width, two adjacent slots per thread with 16-byte loads (same
// arithmetic per slot as k_bprimal / k_bdual, so the same bits).
__global__ void __launch_bounds__(kBlock) k_bprimal2(BArgs a) {
  const long long g = 2 * (static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x);
  const long long j = g / a.W;
  if (j >= a.n) return;
  const int t = static_cast<int>(g - j * a.W);
  // the row's operands first, so their loads are in flight during the gathers (same values)
  const size_t q = static_cast<size_t>(g);
  const double2 xq = *reinterpret_cast<const double2*>(a.b.x + q);
  const double2 c = *reinterpret_cast<const double2*>(a.b.c + q);
  const double2 l = *reinterpret_cast<const double2*>(a.b.l + q);
  const double2 u = *reinterpret_cast<const double2*>(a.b.u + q);
  const double2 x0 = *reinterpret_cast<const double2*>(a.b.x0 + q);
  const double w0 = halpern_w(a.b.inner0[t], a.it), w1 = halpern_w(a.b.inner0[t + 1], a.it);
  const double tau0 = a.b.tau[t], tau1 = a.b.tau[t + 1];
  double s0 = 0.0, s1 = 0.0;
  for (int p = a.at_ptr[j]; p < a.at_ptr[j + 1]; ++p) {
    const double v = a.at_val[p];
    const double2 y = *reinterpret_cast<const double2*>(a.b.y + static_cast<size_t>(a.at_col[p]) * a.W + t);
    s0 += v * y.x;
    s1 += v * y.y;
  }
  double2 h, xb, xn;
  h.x = clampd(xq.x - tau0 * (c.x - s0), l.x, u.x);
  h.y = clampd(xq.y - tau1 * (c.y - s1), l.y, u.y);
  xb.x = 2 * h.x - xq.x;
  xb.y = 2 * h.y - xq.y;
  xn.x = w0 * (a.rho2 * h.x + a.rho1 * xq.x) + (1 - w0) * x0.x;
  xn.y = w1 * (a.rho2 * h.y + a.rho1 * xq.y) + (1 - w1) * x0.y;
  if (a.hat) *reinterpret_cast<double2*>(a.b.xh + q) = h;
  *reinterpret_cast<double2*>(a.b.xb + q) = xb;
  *reinterpret_cast<double2*>(a.b.x + q) = xn;
}

__global__ void __launch_bounds__(kBlock) k_bdual2(BArgs a) {
  const long long g = 2 * (static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x);
  const long long i = g / a.W;
  if (i >= a.m) return;
  const int t = static_cast<int>(g - i * a.W);
  // the row's operands first, so their loads are in flight during the gathers (same values)
  const size_t q = static_cast<size_t>(g);
  const double2 yq = *reinterpret_cast<const double2*>(a.b.y + q);
  const double2 rl = *reinterpret_cast<const double2*>(a.b.rl + q);
  const double2 ru = *reinterpret_cast<const double2*>(a.b.ru + q);
  const double2 y0 = *reinterpret_cast<const double2*>(a.b.y0 + q);
  const double w0 = halpern_w(a.b.inner0[t], a.it), w1 = halpern_w(a.b.inner0[t + 1], a.it);
  const double sg0 = a.b.sigma[t], sg1 = a.b.sigma[t + 1];
  double s0 = 0.0, s1 = 0.0;
  for (int p = a.a_ptr[i]; p < a.a_ptr[i + 1]; ++p) {
    const double v = a.a_val[p];
    const double2 x = *reinterpret_cast<const double2*>(a.b.xb + static_cast<size_t>(a.a_col[p]) * a.W + t);
    s0 += v * x.x;
    s1 += v * x.y;
  }
  double2 h, yb, yn;
  h.x = yq.x - sg0 * s0 + sg0 * clampd(s0 - yq.x / sg0, rl.x, ru.x);
  h.y = yq.y - sg1 * s1 + sg1 * clampd(s1 - yq.y / sg1, rl.y, ru.y);
  yb.x = 2 * h.x - yq.x;
  yb.y = 2 * h.y - yq.y;
  yn.x = w0 * (a.rho2 * h.x + a.rho1 * yq.x) + (1 - w0) * y0.x;
  yn.y = w1 * (a.rho2 * h.y + a.rho1 * yq.y) + (1 - w1) * y0.y;
  if (a.hat) {
    *reinterpret_cast<double2*>(a.b.yh + q) = h;
    *reinterpret_cast<double2*>(a.b.yb + q) = yb;
  }
  *reinterpret_cast<double2*>(a.b.y + q) = yn;
}

// ---- increment form (mixed precision): the single-solve form of psh_fused.hip per slot.
// Anchor constants of slot t in fp64 from its anchor (x0, y0): code 2 rebuilds every constant
// and clears δ, code 1 only KY = y0/σ (σ changed without a restart).
__global__ void __launch_bounds__(kBlock) k_bdanchor_n(BArgs a) {
  const long long g = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  const long long j = g / a.W;
  if (j >= a.n) return;
  const int t = static_cast<int>(g - j * a.W);
  if (a.b.anchor[t] != 2) return;
  const size_t q = static_cast<size_t>(g);
  double s = 0.0;
  for (int p = a.at_ptr[j]; p < a.at_ptr[j + 1]; ++p) s += a.at_val[p] * a.b.y0[static_cast<size_t>(a.at_col[p]) * a.W + t];
  const double x = a.b.x0[q];
  a.b.G[q] = static_cast<float>(a.b.c[q] - s);
  a.b.LX[q] = static_cast<float>(a.b.l[q] - x);
  a.b.UX[q] = static_cast<float>(a.b.u[q] - x);
  a.b.ix[q] = a.b.ixh[q] = a.b.ixb[q] = 0.0f;
}

__global__ void __launch_bounds__(kBlock) k_bdanchor_m(BArgs a) {
  const long long g = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  const long long i = g / a.W;
  if (i >= a.m) return;
  const int t = static_cast<int>(g - i * a.W);
  const int code = a.b.anchor[t];
  if (code == 0) return;
  const size_t q = static_cast<size_t>(g);
  a.b.KY[q] = static_cast<float>(a.b.y0[q] / a.b.sigma[t]);
  if (code != 2) return;
  double s = 0.0;
  for (int p = a.a_ptr[i]; p < a.a_ptr[i + 1]; ++p) s += a.a_val[p] * a.b.x0[static_cast<size_t>(a.a_col[p]) * a.W + t];
  a.b.BL[q] = static_cast<float>(s - a.b.rl[q]);
  a.b.BU[q] = static_cast<float>(s - a.b.ru[q]);
  a.b.iy[q] = a.b.iyh[q] = a.b.iyb[q] = 0.0f;
}

// δx̂ = clamp(δx − τ(G − Ãᵀδy), LX, UX), δx̄ = 2δx̂ − δx, δx ← w(2ρδx̂ + (1−2ρ)δx); fp32 values and
// increments, products and row sums in fp64 (CSR order).
__global__ void __launch_bounds__(kBlock) k_bdprimal(BArgs a) {
  const long long g = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  const long long j = g / a.W;
  if (j >= a.n) return;
  const int t = static_cast<int>(g - j * a.W);
  double s = 0.0;
  for (int p = a.at_ptr[j]; p < a.at_ptr[j + 1]; ++p)
    s += static_cast<double>(a.at_valf[p]) * static_cast<double>(a.b.iy[static_cast<size_t>(a.at_col[p]) * a.W + t]);
  const size_t q = static_cast<size_t>(g);
  const float w = static_cast<float>(halpern_w(a.b.inner0[t], a.it));
  const float tau = static_cast<float>(a.b.tau[t]);
  const float xq = a.b.ix[q];
  float h = xq - tau * (a.b.G[q] - static_cast<float>(s));
  h = h < a.b.LX[q] ? a.b.LX[q] : (h > a.b.UX[q] ? a.b.UX[q] : h);
  if (a.hat) a.b.ixh[q] = h;
  a.b.ixb[q] = 2 * h - xq;
  a.b.ix[q] = w * (static_cast<float>(a.rho2) * h + static_cast<float>(a.rho1) * xq);
}

// t = clamp(KY + δy/σ, BU + Ãδx̄, BL + Ãδx̄), δŷ = δy − σt, δȳ = 2δŷ − δy, δy ← w(2ρδŷ + (1−2ρ)δy).
__global__ void __launch_bounds__(kBlock) k_bddual(BArgs a) {
  const long long g = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  const long long i = g / a.W;
  if (i >= a.m) return;
  const int t = static_cast<int>(g - i * a.W);
  double s = 0.0;
  for (int p = a.a_ptr[i]; p < a.a_ptr[i + 1]; ++p)
    s += static_cast<double>(a.a_valf[p]) * static_cast<double>(a.b.ixb[static_cast<size_t>(a.a_col[p]) * a.W + t]);
  const size_t q = static_cast<size_t>(g);
  const float w = static_cast<float>(halpern_w(a.b.inner0[t], a.it));
  const float sg = static_cast<float>(a.b.sigma[t]);
  const float ax = static_cast<float>(s);
  const float yq = a.b.iy[q];
  const float lo = a.b.BU[q] + ax, hi = a.b.BL[q] + ax;
  float v = a.b.KY[q] + yq / sg;
  v = v < lo ? lo : (v > hi ? hi : v);
  const float h = yq - sg * v;
  if (a.hat) {
    a.b.iyh[q] = h;
    a.b.iyb[q] = 2 * h - yq;
  }
  a.b.iy[q] = w * (static_cast<float>(a.rho2) * h + static_cast<float>(a.rho1) * yq);
}

// The same two steps for a batch width divisible by P, P adjacent slots per thread with 4P-byte
// loads (a wave then covers P rows of 64 slots): same arithmetic per slot, so the same bits.
template <int P>
struct Floats;
template <>
struct Floats<2> {
  using V = float2;
};
template <>
struct Floats<4> {
  using V = float4;
};
template <int P>
__device__ inline void ldf(const float* p, float (&v)[P]) {
  *reinterpret_cast<typename Floats<P>::V*>(v) = *reinterpret_cast<const typename Floats<P>::V*>(p);
}
template <int P>
__device__ inline void stf(float* p, const float (&v)[P]) {
  *reinterpret_cast<typename Floats<P>::V*>(p) = *reinterpret_cast<const typename Floats<P>::V*>(v);
}

template <int P>
__global__ void __launch_bounds__(kBlock) k_bdprimalP(BArgs a) {
  const long long g = P * (static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x);
  const long long j = g / a.W;
  if (j >= a.n) return;
  const int t = static_cast<int>(g - j * a.W);
  // the row's operands first, so their loads are in flight during the gathers (same values)
  const size_t q = static_cast<size_t>(g);
  float xq[P], G[P], LX[P], UX[P], h[P], xb[P], xn[P], w[P], tau[P];
  ldf<P>(a.b.ix + q, xq);
  ldf<P>(a.b.G + q, G);
  ldf<P>(a.b.LX + q, LX);
  ldf<P>(a.b.UX + q, UX);
#pragma unroll
  for (int e = 0; e < P; ++e) {
    w[e] = static_cast<float>(halpern_w(a.b.inner0[t + e], a.it));
    tau[e] = static_cast<float>(a.b.tau[t + e]);
  }
  double s[P];
#pragma unroll
  for (int e = 0; e < P; ++e) s[e] = 0.0;
  for (int p = a.at_ptr[j]; p < a.at_ptr[j + 1]; ++p) {
    const double v = static_cast<double>(a.at_valf[p]);
    float y[P];
    ldf<P>(a.b.iy + static_cast<size_t>(a.at_col[p]) * a.W + t, y);
#pragma unroll
    for (int e = 0; e < P; ++e) s[e] += v * static_cast<double>(y[e]);
  }
  const float r2 = static_cast<float>(a.rho2), r1 = static_cast<float>(a.rho1);
#pragma unroll
  for (int e = 0; e < P; ++e) {
    const float v = xq[e] - tau[e] * (G[e] - static_cast<float>(s[e]));
    h[e] = v < LX[e] ? LX[e] : (v > UX[e] ? UX[e] : v);
    xb[e] = 2 * h[e] - xq[e];
    xn[e] = w[e] * (r2 * h[e] + r1 * xq[e]);
  }
  if (a.hat) stf<P>(a.b.ixh + q, h);
  stf<P>(a.b.ixb + q, xb);
  stf<P>(a.b.ix + q, xn);
}

template <int P>
__global__ void __launch_bounds__(kBlock) k_bddualP(BArgs a) {
  const long long g = P * (static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x);
  const long long i = g / a.W;
  if (i >= a.m) return;
  const int t = static_cast<int>(g - i * a.W);
  // the row's operands first, so their loads are in flight during the gathers (same values)
  const size_t q = static_cast<size_t>(g);
  float yq[P], KY[P], BL[P], BU[P], h[P], yb[P], yn[P], w[P], sgs[P];
  ldf<P>(a.b.iy + q, yq);
  ldf<P>(a.b.KY + q, KY);
  ldf<P>(a.b.BL + q, BL);
  ldf<P>(a.b.BU + q, BU);
#pragma unroll
  for (int e = 0; e < P; ++e) {
    w[e] = static_cast<float>(halpern_w(a.b.inner0[t + e], a.it));
    sgs[e] = static_cast<float>(a.b.sigma[t + e]);
  }
  double s[P];
#pragma unroll
  for (int e = 0; e < P; ++e) s[e] = 0.0;
  for (int p = a.a_ptr[i]; p < a.a_ptr[i + 1]; ++p) {
    const double v = static_cast<double>(a.a_valf[p]);
    float x[P];
    ldf<P>(a.b.ixb + static_cast<size_t>(a.a_col[p]) * a.W + t, x);
#pragma unroll
    for (int e = 0; e < P; ++e) s[e] += v * static_cast<double>(x[e]);
  }
  const float r2 = static_cast<float>(a.rho2), r1 = static_cast<float>(a.rho1);
#pragma unroll
  for (int e = 0; e < P; ++e) {
    const float sg = sgs[e];
    const float ax = static_cast<float>(s[e]);
    const float lo = BU[e] + ax, hi = BL[e] + ax;
    float v = KY[e] + yq[e] / sg;
    v = v < lo ? lo : (v > hi ? hi : v);
    h[e] = yq[e] - sg * v;
    yb[e] = 2 * h[e] - yq[e];
    yn[e] = w[e] * (r2 * h[e] + r1 * yq[e]);
  }
  if (a.hat) {
    stf<P>(a.b.iyh + q, h);
    stf<P>(a.b.iyb + q, yb);
  }
  stf<P>(a.b.iy + q, yn);
}

// fp64 points from anchor + increment: ẑ and z̄ (hat steps), and z itself when `cur` is set.
__global__ void k_bdpoint(const double* __restrict__ z0, const float* __restrict__ ih, const float* __restrict__ ib,
                          const float* __restrict__ ic, double* __restrict__ zh, double* __restrict__ zb,
                          double* __restrict__ zc, long long len) {
  const long long g = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (g >= len) return;
  const double a = z0[g];
  zh[g] = a + static_cast<double>(ih[g]);
  zb[g] = a + static_cast<double>(ib[g]);
  if (zc) zc[g] = a + static_cast<double>(ic[g]);
}

// ---- residual and check passes (one thread per (row, slot))
// Primal and dual ray entries from a scaled direction (reference engine.cpp:132-139, 90-97).
__device__ inline double bprimal_ray(double d, double cs, double bs, double lo, double up) {
  double v = cs * d / bs;
  if (isfinite(lo)) v = dmaxd(v, 0.0);
  if (isfinite(up)) v = dmind(v, 0.0);
  return v;
}
__device__ inline double bdual_ray(double d, double rs, double os, double lo, double up) {
  double v = rs * d / os;
  if (!isfinite(lo)) v = dmind(v, 0.0);
  if (!isfinite(up)) v = dmaxd(v, 0.0);
  return v;
}

__global__ void __launch_bounds__(kBlock) k_bcheck_n(BArgs a, int kkt, int ray) {
  const long long g = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  const long long j = g / a.W;
  if (j >= a.n) return;
  const int t = static_cast<int>(g - j * a.W);
  const size_t q = static_cast<size_t>(g);
  const double d = a.b.xb[q] - a.b.xh[q];
  a.b.dx[q] = d;
  bterm(a, B_DX2)[q] = d * d;
  if (kkt) {
    a.b.xo[q] = clampd(a.cs[j] * a.b.xh[q] / a.b.bs[t], a.b.clo[q], a.b.cuo[q]);
    const double e = a.b.xh[q] - a.b.x0[q];
    bterm(a, B_PIDX)[q] = e * e;
  }
  if (ray) {  // primal ray on T(z) − z (engine.cpp:132-139): c·d and |c·d| sums, max |d|
    const double v = bprimal_ray(d, a.cs[j], a.b.bs[t], a.b.clo[q], a.b.cuo[q]);
    a.b.dv[q] = v;
    bterm(a, B_DMAXRAY)[q] = fabs(v);
    bterm(a, B_CD)[q] = a.b.cmin[q] * v;
    bterm(a, B_CMAG)[q] = fabs(a.b.cmin[q] * v);
  }
}

__global__ void __launch_bounds__(kBlock) k_bcheck_m(BArgs a, int kkt, int ray) {
  const long long g = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  const long long i = g / a.W;
  if (i >= a.m) return;
  const int t = static_cast<int>(g - i * a.W);
  const size_t q = static_cast<size_t>(g);
  const double d = a.b.yb[q] - a.b.yh[q];
  a.b.dy[q] = d;
  bterm(a, B_DY2)[q] = d * d;
  if (kkt) {
    a.b.yo[q] = a.rs[i] * a.b.yh[q] / a.b.os[t];
    const double e = a.b.yh[q] - a.b.y0[q];
    bterm(a, B_PIDY)[q] = e * e;
  }
  if (ray) {  // dual ray on T(z) − z (engine.cpp:90-97): row part of its objective, max |r|
    const double lo = a.b.rlo[q], up = a.b.ruo[q];
    const double r = bdual_ray(d, a.rs[i], a.b.os[t], lo, up);
    a.b.rv[q] = r;
    bterm(a, B_RMAX)[q] = fabs(r);
    const double v = r > 0 ? lo * r : (r < 0 ? up * r : 0.0);
    bterm(a, B_ROBJR)[q] = v;
    bterm(a, B_RMAGR)[q] = fabs(v);
  }
}

__device__ inline double elem_viol(double v, double lo, double up) {
  if (v < lo) return (lo - v) / (1.0 + fabs(lo));
  if (v > up) return (v - up) / (1.0 + fabs(up));
  return 0.0;
}

__global__ void __launch_bounds__(kBlock) k_bcheck_rows(BArgs a, int kkt, int ray) {
  const long long g = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  const long long i = g / a.W;
  if (i >= a.m) return;
  const int t = static_cast<int>(g - i * a.W);
  const size_t q = static_cast<size_t>(g);
  const int b = a.a_ptr[i], e = a.a_ptr[i + 1];
  double s = 0.0;
  for (int p = b; p < e; ++p) s += a.a_val[p] * a.b.dx[static_cast<size_t>(a.a_col[p]) * a.W + t];
  a.b.adx[q] = s;
  bterm(a, B_CROSS)[q] = a.b.dy[q] * s;
  if (ray) {  // A·d against the finite row bounds (engine.cpp:140-145)
    double ad = 0.0;
    for (int p = b; p < e; ++p) ad += a.a_orig[p] * a.b.dv[static_cast<size_t>(a.a_col[p]) * a.W + t];
    double pv = 0.0;
    if (isfinite(a.b.rlo[q])) pv = dmaxd(pv, -ad);
    if (isfinite(a.b.ruo[q])) pv = dmaxd(pv, ad);
    bterm(a, B_PVIOL)[q] = pv;
  }
  if (!kkt) return;
  double ax = 0.0;
  for (int p = b; p < e; ++p) ax += a.a_orig[p] * a.b.xo[static_cast<size_t>(a.a_col[p]) * a.W + t];
  const double lo = a.b.rlo[q], up = a.b.ruo[q];
  const double viol = ax < lo ? lo - ax : (ax > up ? ax - up : 0.0);
  bterm(a, B_RP2)[q] = viol * viol;
  bterm(a, B_PMAXR)[q] = elem_viol(ax, lo, up);
  const double bl = isfinite(lo) ? fabs(lo) : 0.0, bu = isfinite(up) ? fabs(up) : 0.0;
  bterm(a, B_B2)[q] = dmaxd(bl, bu) * dmaxd(bl, bu);
  const double y = a.b.yo[q];
  double dt = 0.0, rd = 0.0, dm = 0.0;
  if (y > 0) {
    if (isfinite(lo)) dt = y * lo;
    else rd = y * y, dm = y;
  } else if (y < 0) {
    if (isfinite(up)) dt = y * up;
    else rd = y * y, dm = -y;
  }
  bterm(a, B_DROW)[q] = dt;
  bterm(a, B_RD2ROW)[q] = rd;
  bterm(a, B_DMAXR)[q] = dm;
}

__global__ void __launch_bounds__(kBlock) k_bcheck_cols(BArgs a, int ray) {
  const long long g = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  const long long j = g / a.W;
  if (j >= a.n) return;
  const int t = static_cast<int>(g - j * a.W);
  const size_t q = static_cast<size_t>(g);
  if (ray) {  // column terms of λ = −Aᵀr (engine.cpp:98-119)
    double gs = 0.0, ga = 0.0;
    for (int p = a.at_ptr[j]; p < a.at_ptr[j + 1]; ++p) {
      const double v = a.ato_val[p] * a.b.rv[static_cast<size_t>(a.ato_col[p]) * a.W + t];
      gs += v;
      ga += fabs(v);
    }
    const double lam = -gs, lo = a.b.clo[q], up = a.b.cuo[q];
    double obj = 0.0, mag = 0.0, viol = 0.0;
    if (lam > 0) {
      if (isfinite(lo)) obj = lo * lam, mag = ga * fabs(lo);
      else viol = lam;
    } else if (lam < 0) {
      if (isfinite(up)) obj = up * lam, mag = ga * fabs(up);
      else viol = -lam;
    }
    bterm(a, B_ROBJC)[q] = obj;
    bterm(a, B_RMAGC)[q] = mag;
    bterm(a, B_DVIOL)[q] = viol;
  }
  double s = 0.0;
  for (int p = a.at_ptr[j]; p < a.at_ptr[j + 1]; ++p) s += a.ato_val[p] * a.b.yo[static_cast<size_t>(a.ato_col[p]) * a.W + t];
  const double cj = a.b.cmin[q], xj = a.b.xo[q], lo = a.b.clo[q], up = a.b.cuo[q];
  bterm(a, B_C2)[q] = cj * cj;
  bterm(a, B_CINF)[q] = fabs(cj);
  bterm(a, B_P)[q] = cj * xj;
  bterm(a, B_PMAXC)[q] = elem_viol(xj, lo, up);
  const double lam = cj - s;
  double z;
  if (isfinite(lo) && isfinite(up)) z = lam;
  else if (isfinite(lo)) z = dmaxd(lam, 0.0);
  else if (isfinite(up)) z = dmind(lam, 0.0);
  else z = 0.0;
  bterm(a, B_RD2COL)[q] = (lam - z) * (lam - z);
  bterm(a, B_DMAXC)[q] = fabs(lam - z);
  bterm(a, B_DCOL)[q] = z > 0 ? z * lo : (z < 0 ? z * up : 0.0);
}

// ---- segmented reducer over interleaved arrays: job → per-slot value at out[t·stride + slot]
struct BJob {
  int ta, tb;               // term indices (tb < 0: none)
  long long ra, rb;         // rows of each
  int op;                   // 0 sum, 1 max
  int stat;                 // PSH_BSTAT_* index
};

struct BRed {
  BJob jobs[kBJobs];
  int njobs, W, mode;
  const double* terms;
  long long stride;  // rows_max · cap
  double* partials;
  double* stats;
};

__device__ inline double bval(const BRed& r, const BJob& J, long long row, int t) {
  const int k = row < J.ra ? J.ta : J.tb;
  const long long rr = row < J.ra ? row : row - J.ra;
  return r.terms[static_cast<size_t>(k) * r.stride + static_cast<size_t>(rr) * r.W + t];
}

// Tree, level 1: block b of job j covers rows [R·b/B, R·(b+1)/B); threads are (slot, group).
__global__ void __launch_bounds__(kBlock) k_bred1(BRed r) {
  const BJob J = r.jobs[blockIdx.y];
  if (J.op == 0 && r.mode != 0) return;
  const int W = r.W, G = kBlock / W;
  const int t = threadIdx.x % W, grp = threadIdx.x / W;
  const long long R = J.ra + J.rb;
  const long long lo = R * blockIdx.x / kBTreeBlocks, hi = R * (blockIdx.x + 1) / kBTreeBlocks;
  double acc = 0.0;
  if (grp < G)
    for (long long row = lo + grp; row < hi; row += G) {
      const double v = bval(r, J, row, t);
      acc = J.op == 0 ? acc + v : fmax(acc, v);
    }
  __shared__ double red[kBlock];
  red[threadIdx.x] = acc;
  __syncthreads();
  if (grp == 0) {
    for (int k = 1; k < G; ++k) acc = J.op == 0 ? acc + red[k * W + t] : fmax(acc, red[k * W + t]);
    r.partials[(static_cast<size_t>(blockIdx.y) * kBTreeBlocks + blockIdx.x) * W + t] = acc;
  }
}

__global__ void k_bred2(BRed r) {
  const int j = blockIdx.x, t = threadIdx.x;
  const BJob J = r.jobs[j];
  if (t >= r.W || (J.op == 0 && r.mode != 0)) return;
  double acc = 0.0;
  for (int b = 0; b < kBTreeBlocks; ++b) {
    const double v = r.partials[(static_cast<size_t>(j) * kBTreeBlocks + b) * r.W + t];
    acc = J.op == 0 ? acc + v : fmax(acc, v);
  }
  r.stats[static_cast<size_t>(t) * PSH_BSTAT_COUNT + J.stat] = acc;
}

// CPU order: one sequential loop per slot (lanes of a wave read consecutive addresses).
__global__ void k_bred_seq(BRed r) {
  const int j = blockIdx.x, t = threadIdx.x;
  const BJob J = r.jobs[j];
  if (t >= r.W || J.op != 0 || r.mode == 0) return;
  double s = 0.0;
  const long long R = J.ra + J.rb;
  for (long long row = 0; row < R; ++row) s += bval(r, J, row, t);
  r.stats[static_cast<size_t>(t) * PSH_BSTAT_COUNT + J.stat] = s;
}

__global__ void k_bcolumn(const double* __restrict__ src, long long rows, int W, int t, double* __restrict__ dst) {
  const long long r = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (r < rows) dst[r] = src[r * W + t];
}

template <class T>
__global__ void k_bgather(const T* __restrict__ src, T* __restrict__ dst, long long rows, int W, int W2,
                          const int* __restrict__ keep) {
  const long long g = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (g >= rows * W2) return;
  const long long r = g / W2;
  const int t2 = static_cast<int>(g - r * W2);
  dst[g] = src[r * W + keep[t2]];
}

__global__ void k_brestart(double* __restrict__ x, double* __restrict__ x0, const double* __restrict__ xh,
                           const int* __restrict__ mask, long long rows, int W) {
  const long long g = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (g >= rows * W) return;
  const int t = static_cast<int>(g % W);
  if (!mask[t]) return;
  x[g] = xh[g];
  x0[g] = xh[g];
}

__global__ void k_bpoint(const double* __restrict__ xh, const double* __restrict__ yh, const double* __restrict__ cs,
                         const double* __restrict__ rs, const double* __restrict__ clo, const double* __restrict__ cuo,
                         double bs, double os, int t, int W, int m, int n, double* __restrict__ x, double* __restrict__ y) {
  const long long g = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (g < n) {
    const size_t q = static_cast<size_t>(g) * W + t;
    x[g] = clampd(cs[g] * xh[q] / bs, clo[q], cuo[q]);
  }
  if (g < m) y[g] = rs[g] * yh[static_cast<size_t>(g) * W + t] / os;
}

BArgs make_bargs(psh_ctx* c, int it, int hat, double rho) {
  BArgs a{};
  a.m = c->m;
  a.n = c->n;
  a.W = c->batch->W;
  a.it = it;
  a.hat = hat;
  a.rho2 = 2 * rho;
  a.rho1 = 1 - 2 * rho;
  a.a_ptr = c->A.ptr;
  a.a_col = c->A.col;
  a.at_ptr = c->At.ptr;
  a.at_col = c->At.col;
  a.ato_col = c->ato_col;
  a.a_val = c->A.val;
  a.at_val = c->At.val;
  a.a_valf = c->A.valf;
  a.at_valf = c->At.valf;
  a.a_orig = c->a_orig;
  a.ato_val = c->ato_val;
  a.rs = c->rs;
  a.cs = c->cs;
  a.b = *c->batch;
  return a;
}

int grid_rows(long long rows, int W) { return grid_for(rows * W); }

void free_batch(psh_ctx* c) {
  BatchState* b = c->batch;
  if (!b) return;
  void* ptrs[] = {b->c, b->l, b->u, b->rl, b->ru, b->cmin, b->clo, b->cuo, b->rlo, b->ruo, b->bs, b->os, b->x, b->x0,
                  b->xh, b->xb, b->y, b->y0, b->yh, b->yb, b->dx, b->dy, b->adx, b->xo, b->yo, b->rv, b->dv, b->tau, b->sigma,
                  b->inner0, b->mask, b->keep, b->terms, b->partials, b->stats, b->tmp, b->ix, b->ixh, b->ixb,
                  b->G, b->LX, b->UX, b->iy, b->iyh, b->iyb, b->BL, b->BU, b->KY, b->anchor};
  for (void* p : ptrs) dev_free(p);
  for (auto& kv : b->graphs) (void)hipGraphExecDestroy(kv.second);
  if (b->pinned) {
    std::lock_guard<std::recursive_mutex> lock(api_mutex());
    (void)hipHostFree(b->pinned);
  }
  delete b;
  c->batch = nullptr;
}

// Pinned staging layout (bytes): params (tau, sigma, inner0) | mask | keep | stats | anchor codes.
constexpr size_t kPinParams = 0;
constexpr size_t kPinMask = 3 * 8 * PSH_BATCH_MAX_W;
constexpr size_t kPinKeep = kPinMask + 4 * PSH_BATCH_MAX_W;
constexpr size_t kPinStats = kPinKeep + 4 * PSH_BATCH_MAX_W;
constexpr size_t kPinAnchor = kPinStats + 8 * PSH_BATCH_MAX_W * PSH_BSTAT_COUNT;
constexpr size_t kPinBytes = kPinAnchor + 4 * PSH_BATCH_MAX_W;

int batch_reduce(psh_ctx* c, int mode, bool kkt, bool only_residual, int r0, bool ray) {
  BatchState& b = *c->batch;
  BRed r{};
  r.W = b.W;
  r.mode = mode;
  r.terms = b.terms;
  r.stride = b.rows_max * b.cap;
  r.partials = b.partials;
  r.stats = b.stats;
  const long long m = c->m, n = c->n;
  int k = 0;
  auto add = [&](int ta, long long ra, int op, int stat, int tb = -1, long long rb = 0) {
    r.jobs[k++] = BJob{ta, tb < 0 ? ta : tb, ra, rb, op, stat};
  };
  add(B_DX2, n, 0, r0 ? PSH_BSTAT_R0_PX : PSH_BSTAT_PX);
  add(B_DY2, m, 0, r0 ? PSH_BSTAT_R0_PY : PSH_BSTAT_PY);
  add(B_CROSS, m, 0, r0 ? PSH_BSTAT_R0_CROSS : PSH_BSTAT_CROSS);
  if (kkt && !only_residual) {
    add(B_RP2, m, 0, PSH_BSTAT_RP2);
    add(B_B2, m, 0, PSH_BSTAT_B2);
    add(B_DROW, m, 0, PSH_BSTAT_D, B_DCOL, n);
    add(B_RD2ROW, m, 0, PSH_BSTAT_RD2, B_RD2COL, n);
    add(B_C2, n, 0, PSH_BSTAT_C2);
    add(B_P, n, 0, PSH_BSTAT_P);
    add(B_PIDX, n, 0, PSH_BSTAT_PIDX);
    add(B_PIDY, m, 0, PSH_BSTAT_PIDY);
    add(B_PMAXR, m, 1, PSH_BSTAT_PMAX, B_PMAXC, n);
    add(B_DMAXR, m, 1, PSH_BSTAT_DMAX, B_DMAXC, n);
    add(B_CINF, n, 1, PSH_BSTAT_CINF);
  }
  if (ray) {  // psh_ray's scalars
    add(B_RMAX, m, 1, PSH_BSTAT_RAY_RMAX);
    add(B_ROBJR, m, 0, PSH_BSTAT_RAY_DUAL_OBJ, B_ROBJC, n);
    add(B_RMAGR, m, 0, PSH_BSTAT_RAY_DUAL_MAG, B_RMAGC, n);
    add(B_DVIOL, n, 1, PSH_BSTAT_RAY_DUAL_VIOL);
    add(B_DMAXRAY, n, 1, PSH_BSTAT_RAY_DMAX);
    add(B_CD, n, 0, PSH_BSTAT_RAY_PRIMAL_CD);
    add(B_CMAG, n, 0, PSH_BSTAT_RAY_PRIMAL_CMAG);
    add(B_PVIOL, m, 1, PSH_BSTAT_RAY_PRIMAL_VIOL);
  }
  r.njobs = k;
  hipLaunchKernelGGL(k_bred1, dim3(kBTreeBlocks, k), dim3(kBlock), 0, c->stream, r);
  hipLaunchKernelGGL(k_bred2, dim3(k), dim3(PSH_BATCH_MAX_W), 0, c->stream, r);
  if (mode != 0) hipLaunchKernelGGL(k_bred_seq, dim3(k), dim3(PSH_BATCH_MAX_W), 0, c->stream, r);
  PSH_LAUNCHED(c, "batch reductions");
  return PSH_OK;
}

int batch_residual_pass(psh_ctx* c, bool kkt, int mode, bool r0, bool ray = false) {
  BatchState& b = *c->batch;
  const BArgs a = make_bargs(c, 0, 0, 1.0);
  const int k = kkt ? 1 : 0, rr = ray ? 1 : 0;
  // the dual ray (rv) must exist before the column pass reads it; the primal ray (dv) before the
  // row pass: n-pass, m-pass, rows, cols
  if (c->n > 0) hipLaunchKernelGGL(k_bcheck_n, dim3(grid_rows(c->n, b.W)), dim3(kBlock), 0, c->stream, a, k, rr);
  if (c->m > 0) {
    hipLaunchKernelGGL(k_bcheck_m, dim3(grid_rows(c->m, b.W)), dim3(kBlock), 0, c->stream, a, k, rr);
    hipLaunchKernelGGL(k_bcheck_rows, dim3(grid_rows(c->m, b.W)), dim3(kBlock), 0, c->stream, a, k, rr);
  }
  if ((kkt || ray) && c->n > 0)
    hipLaunchKernelGGL(k_bcheck_cols, dim3(grid_rows(c->n, b.W)), dim3(kBlock), 0, c->stream, a, rr);
  PSH_LAUNCHED(c, "batch check kernels");
  return batch_reduce(c, mode, kkt, !kkt, r0 ? 1 : 0, ray);
}

// Each pinned staging region (1 params, 2 mask, 4 keep, 8 anchor codes) is reused by the next
// call of the same kind: wait for a copy that may still read it.
int pinned_free(psh_ctx* c, unsigned region) {
  BatchState& b = *c->batch;
  if (!(b.pin_pending & region)) return PSH_OK;
  PSH_TRY(c, hipStreamSynchronize(c->stream));
  b.pin_pending = 0;
  return PSH_OK;
}

// Increment form, before a chunk: rebuild the anchor constants of re-anchored slots (code 2) and
// KY of slots whose σ changed (code 1).
int batch_anchor(psh_ctx* c, const double* sigma, double rho) {
  BatchState& b = *c->batch;
  bool any = false, full = false;
  int rc = pinned_free(c, 8);
  if (rc) return rc;
  int* pin = reinterpret_cast<int*>(b.pinned + kPinAnchor);
  for (int t = 0; t < b.W; ++t) {
    pin[t] = b.dirty[t] ? 2 : (sigma[t] != b.ky_sigma[t] ? 1 : 0);
    any = any || pin[t] != 0;
    full = full || pin[t] == 2;
    b.dirty[t] = 0;
    b.ky_sigma[t] = sigma[t];
  }
  if (!any) return PSH_OK;
  PSH_TRY(c, hipMemcpyAsync(b.anchor, pin, sizeof(int) * b.W, hipMemcpyHostToDevice, c->stream));
  b.pin_pending |= 8;
  const BArgs a = make_bargs(c, 0, 0, rho);
  if (full && c->n > 0) hipLaunchKernelGGL(k_bdanchor_n, dim3(grid_rows(c->n, b.W)), dim3(kBlock), 0, c->stream, a);
  if (c->m > 0) hipLaunchKernelGGL(k_bdanchor_m, dim3(grid_rows(c->m, b.W)), dim3(kBlock), 0, c->stream, a);
  PSH_LAUNCHED(c, "batch anchors");
  return PSH_OK;
}

// Increment form: x̂, x̄, ŷ, ȳ (and x, y with `cur`) of every slot in fp64.
int batch_points(psh_ctx* c, bool cur) {
  BatchState& b = *c->batch;
  const long long NW = static_cast<long long>(c->n) * b.W, MW = static_cast<long long>(c->m) * b.W;
  if (NW > 0)
    hipLaunchKernelGGL(k_bdpoint, dim3(grid_for(NW)), dim3(kBlock), 0, c->stream, b.x0, b.ixh, b.ixb, b.ix, b.xh, b.xb,
                       cur ? b.x : nullptr, NW);
  if (MW > 0)
    hipLaunchKernelGGL(k_bdpoint, dim3(grid_for(MW)), dim3(kBlock), 0, c->stream, b.y0, b.iyh, b.iyb, b.iy, b.yh, b.yb,
                       cur ? b.y : nullptr, MW);
  PSH_LAUNCHED(c, "batch points");
  return PSH_OK;
}

int need_batch(psh_ctx* c) {
  if (!c) return PSH_E_ARG;
  if (!c->batch) return fail(c, PSH_E_STATE, "psh_batch_begin not called");
  return PSH_OK;
}

}  // namespace

extern "C" {

int psh_batch_begin(psh_ctx* c, int W, const double* cc, const double* l, const double* u, const double* rl,
                    const double* ru, const double* c_min, const double* cl_o, const double* cu_o, const double* rl_o,
                    const double* ru_o, const double* bs, const double* os) {
  if (!c) return PSH_E_ARG;
  if (!c->loaded) return fail(c, PSH_E_STATE, "no problem loaded");
  if (W < 1 || W > PSH_BATCH_MAX_W) return fail(c, PSH_E_ARG, "batch width out of range");
  free_batch(c);
  c->batch = new BatchState;
  BatchState& b = *c->batch;
  b.W = b.cap = W;
  const long long m = c->m, n = c->n;
  b.rows_max = std::max<long long>(std::max(m, n), 1);
  const long long NW = n * W, MW = m * W, RW = b.rows_max * W;
  int rc;
  auto up = [&](double** dst, const double* src, long long count) -> int {
    int r = alloc_n(c, dst, count);
    if (r) return r;
    if (count > 0 && src) PSH_TRY(c, hipMemcpyAsync(*dst, src, sizeof(double) * count, hipMemcpyHostToDevice, c->stream));
    return PSH_OK;
  };
  if ((rc = up(&b.c, cc, NW)) || (rc = up(&b.l, l, NW)) || (rc = up(&b.u, u, NW)) || (rc = up(&b.rl, rl, MW)) ||
      (rc = up(&b.ru, ru, MW)) || (rc = up(&b.cmin, c_min, NW)) || (rc = up(&b.clo, cl_o, NW)) ||
      (rc = up(&b.cuo, cu_o, NW)) || (rc = up(&b.rlo, rl_o, MW)) || (rc = up(&b.ruo, ru_o, MW)) ||
      (rc = up(&b.bs, bs, W)) || (rc = up(&b.os, os, W)))
    return rc;
  for (double** v : {&b.x, &b.x0, &b.xh, &b.xb, &b.dx, &b.xo, &b.dv})
    if ((rc = alloc_n(c, v, NW))) return rc;
  for (double** v : {&b.y, &b.y0, &b.yh, &b.yb, &b.dy, &b.adx, &b.yo, &b.rv})
    if ((rc = alloc_n(c, v, MW))) return rc;
  if ((rc = alloc_n(c, &b.tau, W)) || (rc = alloc_n(c, &b.sigma, W)) || (rc = alloc_n(c, &b.inner0, W)) ||
      (rc = alloc_n(c, &b.mask, W)) || (rc = alloc_n(c, &b.keep, W)) || (rc = alloc_n(c, &b.terms, RW * kBTerms)) ||
      (rc = alloc_n(c, &b.partials, static_cast<long long>(kBTreeBlocks) * kBJobs * W)) ||
      (rc = alloc_n(c, &b.stats, static_cast<long long>(W) * PSH_BSTAT_COUNT)) ||
      (rc = alloc_n(c, &b.tmp, std::max(RW, m + n))) ||  // gather staging, and one slot's x | y_min
      (rc = alloc_n(c, &b.anchor, W)))
    return rc;
  b.dirty.assign(W, 1);
  b.ky_sigma.assign(W, 0.0);
  {
    std::lock_guard<std::recursive_mutex> lock(api_mutex());
    if (hipHostMalloc(reinterpret_cast<void**>(&b.pinned), kPinBytes) != hipSuccess)
      return fail(c, PSH_E_OOM, "pinned staging");
  }
  PSH_TRY(c, hipStreamSynchronize(c->stream));
  return PSH_OK;
}

int psh_batch_set_iterate(psh_ctx* c, const double* x, const double* y) {
  int rc = need_batch(c);
  if (rc) return rc;
  BatchState& b = *c->batch;
  const size_t NW = static_cast<size_t>(c->n) * b.W, MW = static_cast<size_t>(c->m) * b.W;
  for (double* d : {b.x, b.x0, b.xh})
    if (NW) PSH_TRY(c, hipMemcpyAsync(d, x, sizeof(double) * NW, hipMemcpyHostToDevice, c->stream));
  for (double* d : {b.y, b.y0, b.yh})
    if (MW) PSH_TRY(c, hipMemcpyAsync(d, y, sizeof(double) * MW, hipMemcpyHostToDevice, c->stream));
  PSH_TRY(c, hipStreamSynchronize(c->stream));
  return PSH_OK;
}

int psh_batch_chunk(psh_ctx* c, int iterations, const double* tau, const double* sigma, const long long* inner0, double rho,
                    int mode) {
  int rc = need_batch(c);
  if (rc) return rc;
  BatchState& b = *c->batch;
  const int W = b.W;
  if (iterations < 1 || W < 1) return fail(c, PSH_E_ARG, "empty batch chunk");
  if ((rc = pinned_free(c, 1))) return rc;
  char* pin = b.pinned + kPinParams;
  std::memcpy(pin, tau, 8 * W);
  std::memcpy(pin + 8 * W, sigma, 8 * W);
  std::memcpy(pin + 16 * W, inner0, 8 * W);
  PSH_TRY(c, hipMemcpyAsync(b.tau, pin, 8 * W, hipMemcpyHostToDevice, c->stream));
  PSH_TRY(c, hipMemcpyAsync(b.sigma, pin + 8 * W, 8 * W, hipMemcpyHostToDevice, c->stream));
  PSH_TRY(c, hipMemcpyAsync(b.inner0, pin + 16 * W, 8 * W, hipMemcpyHostToDevice, c->stream));
  b.pin_pending |= 1;
  bool any_new = false;
  for (int t = 0; t < W; ++t) any_new = any_new || inner0[t] == 0;
  const bool delta = b.delta;
  if (delta && (rc = batch_anchor(c, sigma, rho))) return rc;
  // Launch arguments depend only on (length, epoch start in any slot, width, mode): replay a
  // captured graph unless per-kernel timing is on.
  auto enqueue = [&]() -> int {
    for (int s = 0; s < iterations; ++s) {
      const int hat = (s == 0 && any_new) || s == iterations - 1 ? 1 : 0;
      const BArgs a = make_bargs(c, s, hat, rho);
      hipEvent_t ev;
      if (c->n > 0) {
        time_begin(c, 0, &ev);
        if (delta && W % 4 == 0)
          hipLaunchKernelGGL(k_bdprimalP<4>, dim3(grid_for(c->n * static_cast<long long>(W / 4))), dim3(kBlock), 0,
                             c->stream, a);
        else if (delta && W % 2 == 0)
          hipLaunchKernelGGL(k_bdprimalP<2>, dim3(grid_for(c->n * static_cast<long long>(W / 2))), dim3(kBlock), 0,
                             c->stream, a);
        else if (delta) hipLaunchKernelGGL(k_bdprimal, dim3(grid_rows(c->n, W)), dim3(kBlock), 0, c->stream, a);
        else if (W % 2 == 0)
          hipLaunchKernelGGL(k_bprimal2, dim3(grid_for(c->n * static_cast<long long>(W / 2))), dim3(kBlock), 0,
                             c->stream, a);
        else hipLaunchKernelGGL(k_bprimal, dim3(grid_rows(c->n, W)), dim3(kBlock), 0, c->stream, a);
        time_end(c, 0, ev);
      }
      if (c->m > 0) {
        time_begin(c, 1, &ev);
        if (delta && W % 4 == 0)
          hipLaunchKernelGGL(k_bddualP<4>, dim3(grid_for(c->m * static_cast<long long>(W / 4))), dim3(kBlock), 0,
                             c->stream, a);
        else if (delta && W % 2 == 0)
          hipLaunchKernelGGL(k_bddualP<2>, dim3(grid_for(c->m * static_cast<long long>(W / 2))), dim3(kBlock), 0,
                             c->stream, a);
        else if (delta) hipLaunchKernelGGL(k_bddual, dim3(grid_rows(c->m, W)), dim3(kBlock), 0, c->stream, a);
        else if (W % 2 == 0)
          hipLaunchKernelGGL(k_bdual2, dim3(grid_for(c->m * static_cast<long long>(W / 2))), dim3(kBlock), 0,
                             c->stream, a);
        else hipLaunchKernelGGL(k_bdual, dim3(grid_rows(c->m, W)), dim3(kBlock), 0, c->stream, a);
        time_end(c, 1, ev);
      }
      PSH_LAUNCHED(c, "batch iteration");
      if (delta && hat) {
        const int r = batch_points(c, false);
        if (r) return r;
      }
      if (s == 0 && any_new) {  // residuals of every slot after the first step (reference batch.cpp:347-354)
        const int r = batch_residual_pass(c, false, mode, true);
        if (r) return r;
      }
    }
    return PSH_OK;
  };
  if (c->timing || !c->graphs) {
    if ((rc = enqueue())) return rc;
  } else {
    const long long key = (static_cast<long long>(iterations) << 17) | (static_cast<long long>(W) << 5) |
                          (any_new ? 1 : 0) | (static_cast<long long>(mode & 7) << 1) | (delta ? 16 : 0);
    if (rho != b.graphs_rho) {  // ρ is an argument of the captured launches: recapture
      PSH_TRY(c, hipStreamSynchronize(c->stream));
      for (auto& kv : b.graphs) (void)hipGraphExecDestroy(kv.second);
      b.graphs.clear();
      b.graphs_rho = rho;
    }
    auto it = b.graphs.find(key);
    if (it == b.graphs.end()) {
      std::lock_guard<std::recursive_mutex> capture_lock(api_mutex());  // no device-wide calls meanwhile
      PSH_TRY(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
      const int erc = enqueue();
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
      it = b.graphs.emplace(key, exec).first;
    }
    PSH_TRY(c, hipGraphLaunch(it->second, c->stream));
  }
  if (any_new) b.have_r0 = true;
  return PSH_OK;
}

int psh_batch_check(psh_ctx* c, int mode, int want_ray, double* out) {
  int rc = need_batch(c);
  if (rc) return rc;
  BatchState& b = *c->batch;
  hipEvent_t ev;
  time_begin(c, 2, &ev);
  rc = batch_residual_pass(c, true, mode, false, want_ray != 0);
  time_end(c, 2, ev);
  if (rc) return rc;
  double* pin = reinterpret_cast<double*>(b.pinned + kPinStats);
  const size_t bytes = sizeof(double) * b.W * PSH_BSTAT_COUNT;
  PSH_TRY(c, hipMemcpyAsync(pin, b.stats, bytes, hipMemcpyDeviceToHost, c->stream));
  PSH_TRY(c, hipStreamSynchronize(c->stream));
  std::memcpy(out, pin, bytes);
  b.pin_pending = 0;
  b.have_r0 = false;
  return PSH_OK;
}

int psh_batch_restart(psh_ctx* c, const int* mask) {
  int rc = need_batch(c);
  if (rc) return rc;
  BatchState& b = *c->batch;
  if ((rc = pinned_free(c, 2))) return rc;
  int* pin = reinterpret_cast<int*>(b.pinned + kPinMask);
  std::memcpy(pin, mask, sizeof(int) * b.W);
  PSH_TRY(c, hipMemcpyAsync(b.mask, pin, sizeof(int) * b.W, hipMemcpyHostToDevice, c->stream));
  b.pin_pending |= 2;
  if (c->n > 0)
    hipLaunchKernelGGL(k_brestart, dim3(grid_rows(c->n, b.W)), dim3(kBlock), 0, c->stream, b.x, b.x0, b.xh, b.mask,
                       (long long)c->n, b.W);
  if (c->m > 0)
    hipLaunchKernelGGL(k_brestart, dim3(grid_rows(c->m, b.W)), dim3(kBlock), 0, c->stream, b.y, b.y0, b.yh, b.mask,
                       (long long)c->m, b.W);
  PSH_LAUNCHED(c, "batch restart");
  for (int t = 0; t < b.W; ++t)
    if (mask[t]) b.dirty[t] = 1;  // increment form: new anchors at the next chunk
  return PSH_OK;
}

int psh_batch_delta(psh_ctx* c, int on) {
  int rc = need_batch(c);
  if (rc) return rc;
  BatchState& b = *c->batch;
  const bool want = on != 0;
  if (want == b.delta) return PSH_OK;
  if (want) {
    if ((rc = ensure_fp32_matrix(c))) return rc;
    if (!b.ix) {
      const long long NW = static_cast<long long>(c->n) * b.cap, MW = static_cast<long long>(c->m) * b.cap;
      for (float** v : {&b.ix, &b.ixh, &b.ixb, &b.G, &b.LX, &b.UX})
        if ((rc = alloc_n(c, v, NW))) return rc;
      for (float** v : {&b.iy, &b.iyh, &b.iyb, &b.BL, &b.BU, &b.KY})
        if ((rc = alloc_n(c, v, MW))) return rc;
    }
    std::fill(b.dirty.begin(), b.dirty.end(), 1);  // every slot anchors at x0 / y0 with δ = 0
    b.delta = true;
    return PSH_OK;
  }
  if ((rc = batch_points(c, true))) return rc;  // the current points back into the fp64 arrays
  b.delta = false;
  return PSH_OK;
}

int psh_batch_compact(psh_ctx* c, const int* keep, int W2) {
  int rc = need_batch(c);
  if (rc) return rc;
  BatchState& b = *c->batch;
  if (W2 < 0 || W2 > b.W) return fail(c, PSH_E_ARG, "bad compaction");
  for (int k = 0; k < W2; ++k)
    if (keep[k] < 0 || keep[k] >= b.W) return fail(c, PSH_E_ARG, "bad compaction index");
  if (W2 == b.W) return PSH_OK;
  if ((rc = pinned_free(c, 4))) return rc;
  int* pin = reinterpret_cast<int*>(b.pinned + kPinKeep);
  std::memcpy(pin, keep, sizeof(int) * W2);
  PSH_TRY(c, hipMemcpyAsync(b.keep, pin, sizeof(int) * std::max(W2, 1), hipMemcpyHostToDevice, c->stream));
  const long long m = c->m, n = c->n;
  auto squeeze = [&](auto* v, long long rows) -> int {
    using T = std::remove_pointer_t<decltype(v)>;
    if (rows * W2 == 0) return PSH_OK;
    T* tmp = reinterpret_cast<T*>(b.tmp);  // rows_max·cap doubles hold rows·W2 values of either type
    hipLaunchKernelGGL((k_bgather<T>), dim3(grid_for(rows * W2)), dim3(kBlock), 0, c->stream, v, tmp, rows, b.W, W2,
                       b.keep);
    PSH_TRY(c, hipMemcpyAsync(v, tmp, sizeof(T) * rows * W2, hipMemcpyDeviceToDevice, c->stream));
    return PSH_OK;
  };
  for (double* v : {b.c, b.l, b.u, b.cmin, b.clo, b.cuo, b.x, b.x0, b.xh, b.xb})
    if ((rc = squeeze(v, n))) return rc;
  for (double* v : {b.rl, b.ru, b.rlo, b.ruo, b.y, b.y0, b.yh, b.yb})
    if ((rc = squeeze(v, m))) return rc;
  if ((rc = squeeze(b.bs, 1)) || (rc = squeeze(b.os, 1))) return rc;
  if (b.ix) {  // increment-form state (kept while off: it is rebuilt on the next switch anyway)
    for (float* v : {b.ix, b.ixh, b.ixb, b.G, b.LX, b.UX})
      if ((rc = squeeze(v, n))) return rc;
    for (float* v : {b.iy, b.iyh, b.iyb, b.BL, b.BU, b.KY})
      if ((rc = squeeze(v, m))) return rc;
  }
  PSH_LAUNCHED(c, "batch compaction");
  std::vector<char> dirty(W2);
  std::vector<double> kys(W2);
  for (int k = 0; k < W2; ++k) dirty[k] = b.dirty[keep[k]], kys[k] = b.ky_sigma[keep[k]];
  b.dirty.swap(dirty);
  b.ky_sigma.swap(kys);
  b.W = W2;
  PSH_TRY(c, hipStreamSynchronize(c->stream));
  b.pin_pending = 0;
  return PSH_OK;
}

int psh_batch_ray_vectors(psh_ctx* c, int t, double* dual_ray, double* primal_ray) {
  int rc = need_batch(c);
  if (rc) return rc;
  BatchState& b = *c->batch;
  if (t < 0 || t >= b.W) return fail(c, PSH_E_ARG, "bad slot");
  const long long m = c->m, n = c->n;
  // column t of the interleaved rays into tmp: dual ray | primal ray
  if (m > 0)
    hipLaunchKernelGGL(k_bcolumn, dim3(grid_for(m)), dim3(kBlock), 0, c->stream, b.rv, m, b.W, t, b.tmp);
  if (n > 0)
    hipLaunchKernelGGL(k_bcolumn, dim3(grid_for(n)), dim3(kBlock), 0, c->stream, b.dv, n, b.W, t, b.tmp + m);
  PSH_LAUNCHED(c, "batch ray vectors");
  if (dual_ray && m > 0) PSH_TRY(c, hipMemcpyAsync(dual_ray, b.tmp, sizeof(double) * m, hipMemcpyDeviceToHost, c->stream));
  if (primal_ray && n > 0)
    PSH_TRY(c, hipMemcpyAsync(primal_ray, b.tmp + m, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
  PSH_TRY(c, hipStreamSynchronize(c->stream));
  return PSH_OK;
}

int psh_batch_point(psh_ctx* c, int t, double* x, double* y_min) {
  int rc = need_batch(c);
  if (rc) return rc;
  BatchState& b = *c->batch;
  if (t < 0 || t >= b.W) return fail(c, PSH_E_ARG, "bad slot");
  double bs = 1, os = 1;
  PSH_TRY(c, hipMemcpyAsync(&bs, b.bs + t, 8, hipMemcpyDeviceToHost, c->stream));
  PSH_TRY(c, hipMemcpyAsync(&os, b.os + t, 8, hipMemcpyDeviceToHost, c->stream));
  PSH_TRY(c, hipStreamSynchronize(c->stream));
  const long long rows = std::max(c->m, c->n);
  if (rows > 0)
    hipLaunchKernelGGL(k_bpoint, dim3(grid_for(rows)), dim3(kBlock), 0, c->stream, b.xh, b.yh, c->cs, c->rs, b.clo, b.cuo,
                       bs, os, t, b.W, c->m, c->n, b.tmp, b.tmp + c->n);
  PSH_LAUNCHED(c, "batch point");
  if (c->n > 0) PSH_TRY(c, hipMemcpyAsync(x, b.tmp, sizeof(double) * c->n, hipMemcpyDeviceToHost, c->stream));
  if (c->m > 0) PSH_TRY(c, hipMemcpyAsync(y_min, b.tmp + c->n, sizeof(double) * c->m, hipMemcpyDeviceToHost, c->stream));
  PSH_TRY(c, hipStreamSynchronize(c->stream));
  return PSH_OK;
}

}  // extern "C"

namespace psh {
void destroy_batch(psh_ctx* c) { free_batch(c); }
}  // namespace psh
