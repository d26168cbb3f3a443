// psh_serve.hip — many independent LPs on one GPU as one packed block-diagonal system.
//
// Every live instance keeps its scaled problem, original data and iterates in one arena of
// device memory (a fixed HBM byte budget). The row tiles of all instances are concatenated into
// two tile tables (rows of Ãᵀ for the primal half step, rows of Ã for the dual half step), so one
// launch of each fused kernel advances every live instance; a per-instance parameter table
// carries τ, σ, the epoch counter and the restart flags. Per-instance sums (fixed-point
// residual, KKT on the original problem, PID distances) are reduced first inside each tile in a
// fixed order and then over the instance's own tiles in tile order, so an instance's numbers do
// not depend on which other instances share the GPU: its trajectory is the same alone and in
// any mix. The per-iteration arithmetic is that of k_iter_primal / k_iter_dual (reference
// src/pdhg/cpu_backend.cpp:140-178); the check terms are those of kkt_general
// (cpu_backend.cpp:226-284) and the residual of r2hpdhg.cpp:50-59.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <vector>

#include "psh_internal.h"

using namespace psh;

namespace {

// Per-instance reduced quantities.
enum SQ {
  Q_PX, Q_PY, Q_CROSS, Q_RP2, Q_B2, Q_D, Q_RD2, Q_C2, Q_P, Q_PIDX, Q_PIDY, Q_PMAX, Q_DMAX, Q_CINF, Q_R0PX, Q_R0PY,
  Q_R0CROSS,
  // ray tests on T(z) − z (reference engine.cpp:82-155): dual ray r, primal ray d
  Q_RMAX, Q_ROBJ, Q_RMAG, Q_DVIOL, Q_DMAXRAY, Q_CD, Q_CMAG, Q_PVIOL, kQ
};
static_assert(kQ <= 32, "quantity masks are 32-bit");
__host__ __device__ constexpr bool q_is_max(int q) {
  return q == Q_PMAX || q == Q_DMAX || q == Q_CINF || q == Q_RMAX || q == Q_DVIOL || q == Q_DMAXRAY || q == Q_PVIOL;
}
constexpr unsigned kRayMask = (1u << Q_RMAX) | (1u << Q_ROBJ) | (1u << Q_RMAG) | (1u << Q_DVIOL) | (1u << Q_DMAXRAY) |
                              (1u << Q_CD) | (1u << Q_CMAG) | (1u << Q_PVIOL);

struct Inst {
  int m, n;
  const int *a_ptr, *a_col, *at_ptr, *at_col, *ato_col;
  const double *a_val, *at_val, *a_orig, *ato_val;
  const double *c, *l, *u, *rl, *ru, *rs, *cs, *clo, *cuo, *rlo, *ruo, *cmin;
  double *x, *x0, *xh, *xb, *dx, *xo, *y, *y0, *yh, *yb, *dy, *yo;
  double *rv, *dv;  // unscaled dual (m) and primal (n) rays of the last ray check
  // increment form (PSH_SERVE_INCREMENT): fp32 values, δx, δx̄, δy and the anchor constants
  const float *a_valf, *at_valf;
  float *ix, *ixb, *G, *LX, *UX, *iy, *BL, *BU, *KY;
  double bs, os;
  int pt0, pt1, dt0, dt1;  // this instance's tiles in the primal / dual tables
};

struct Param {
  double tau, sigma, inner0;
  double a2, b2;  // 2ρ and 1 − 2ρ of the instance's reflection ρ
  int active, restart;
  int mixed;   // 1: this chunk runs the instance in increment form
  int anchor;  // anchor pass before the chunk: 1 KY only (σ changed), 2 every constant and δ = 0
};

struct Tile {
  int inst, r0, r1, pad;
};

__device__ inline double halpern_w(double inner0, int step) {
  const double k = inner0 + static_cast<double>(step);
  return (k + 1.0) / (k + 2.0);
}
__device__ inline double dmaxd(double a, double b) { return a < b ? b : a; }
__device__ inline double dmind(double a, double b) { return b < a ? b : a; }

// ------------------------------------------------------------------ iteration kernels
// Each epilogue loads its row's operands in prefetch(), which tile_rows calls before the product
// loop, so the loads overlap the gathers; operator() computes the same expressions.
struct SrvEpiPrimal {
  const Inst* d;
  double tau, tw, a, b;
  int hat;
  double xj = 0.0, cj = 0.0, lj = 0.0, uj = 0.0, x0j = 0.0;
  __device__ void prefetch(int j, bool has) {
    if (!has) return;
    xj = d->x[j];
    cj = d->c[j];
    lj = d->l[j];
    uj = d->u[j];
    x0j = d->x0[j];
  }
  __device__ void operator()(int j, bool has, double s) const {
    if (!has) return;
    const double h = clampv<double>(xj - tau * (cj - s), lj, uj);
    if (hat) d->xh[j] = h;
    d->xb[j] = 2 * h - xj;
    d->x[j] = tw * (a * h + b * xj) + (1 - tw) * x0j;
  }
};

struct SrvEpiDual {
  const Inst* d;
  double sg, tw, a, b;
  int hat;
  double yi = 0.0, rli = 0.0, rui = 0.0, y0i = 0.0;
  __device__ void prefetch(int i, bool has) {
    if (!has) return;
    yi = d->y[i];
    rli = d->rl[i];
    rui = d->ru[i];
    y0i = d->y0[i];
  }
  __device__ void operator()(int i, bool has, double s) const {
    if (!has) return;
    const double q = s - yi / sg;
    const double h = yi - sg * s + sg * clampv<double>(q, rli, rui);
    if (hat) {
      d->yh[i] = h;
      d->yb[i] = 2 * h - yi;
    }
    d->y[i] = tw * (a * h + b * yi) + (1 - tw) * y0i;
  }
};

// Increment form (psh_fused.hip's, per instance): z = z₀ + δ with the anchor z₀ in x0 / y0; on
// hat steps the epilogue also writes x̂, x̄ (ŷ, ȳ) in fp64, so the check and restart kernels read
// fp64 points in either form.
struct SrvEpiDeltaP {
  const Inst* d;
  float tau, tw, a, b;
  int hat;
  float xj = 0.0f, gj = 0.0f, lj = 0.0f, uj = 0.0f;
  __device__ void prefetch(int j, bool has) {
    if (!has) return;
    xj = d->ix[j];
    gj = d->G[j];
    lj = d->LX[j];
    uj = d->UX[j];
  }
  __device__ void operator()(int j, bool has, double s) const {
    if (!has) return;
    const float h = clampv<float>(xj - tau * (gj - static_cast<float>(s)), lj, uj);
    const float xb = 2 * h - xj;
    d->ixb[j] = xb;
    d->ix[j] = tw * (a * h + b * xj);
    if (hat) {
      const double a0 = d->x0[j];
      d->xh[j] = a0 + static_cast<double>(h);
      d->xb[j] = a0 + static_cast<double>(xb);
    }
  }
};

struct SrvEpiDeltaD {
  const Inst* d;
  float sg, tw, a, b;
  int hat;
  float yi = 0.0f, kyi = 0.0f, bui = 0.0f, bli = 0.0f;
  __device__ void prefetch(int i, bool has) {
    if (!has) return;
    yi = d->iy[i];
    kyi = d->KY[i];
    bui = d->BU[i];
    bli = d->BL[i];
  }
  __device__ void operator()(int i, bool has, double s) const {
    if (!has) return;
    const float ax = static_cast<float>(s);
    const float t = clampv<float>(kyi + yi / sg, bui + ax, bli + ax);
    const float h = yi - sg * t;
    d->iy[i] = tw * (a * h + b * yi);
    if (hat) {
      const double a0 = d->y0[i];
      d->yh[i] = a0 + static_cast<double>(h);
      d->yb[i] = a0 + static_cast<double>(2 * h - yi);
    }
  }
};

__device__ inline void srv_primal(const Tile* __restrict__ tiles, int ntiles, const Inst* __restrict__ I,
                                  const Param* __restrict__ P, int step, int last) {
  const int t = tile_of_block(blockIdx.x, gridDim.x);
  if (t >= ntiles) return;
  const Tile T = tiles[t];
  const Param q = P[T.inst];
  if (!q.active) return;
  const Inst* d = I + T.inst;
  const int hat = (step == 0 && q.inner0 == 0.0) || last ? 1 : 0;
  if (q.mixed) {
    SrvEpiDeltaP e{d, static_cast<float>(q.tau), static_cast<float>(halpern_w(q.inner0, step)), static_cast<float>(q.a2),
                   static_cast<float>(q.b2), hat};
    tile_rows<float, float>(T.r0, T.r1, d->at_ptr, d->at_col, d->at_valf, d->iy, e);
    return;
  }
  SrvEpiPrimal e{d, q.tau, halpern_w(q.inner0, step), q.a2, q.b2, hat};
  tile_rows<double, double>(T.r0, T.r1, d->at_ptr, d->at_col, d->at_val, d->y, e);
}

__global__ void __launch_bounds__(kBlock) k_srv_primal(const Tile* __restrict__ tiles, int ntiles, const Inst* __restrict__ I,
                                                       const Param* __restrict__ P, int step, int last) {
  srv_primal(tiles, ntiles, I, P, step, last);
}

__global__ void __launch_bounds__(kBlock) k_srv_dual(const Tile* __restrict__ tiles, int ntiles, const Inst* __restrict__ I,
                                                     const Param* __restrict__ P, int step, int last) {
  const int t = tile_of_block(blockIdx.x, gridDim.x);
  if (t >= ntiles) return;
  const Tile T = tiles[t];
  const Param q = P[T.inst];
  if (!q.active) return;
  const Inst* d = I + T.inst;
  const int hat = (step == 0 && q.inner0 == 0.0) || last ? 1 : 0;
  if (q.mixed) {
    SrvEpiDeltaD e{d, static_cast<float>(q.sigma), static_cast<float>(halpern_w(q.inner0, step)),
                   static_cast<float>(q.a2), static_cast<float>(q.b2), hat};
    tile_rows<float, float>(T.r0, T.r1, d->a_ptr, d->a_col, d->a_valf, d->ixb, e);
    return;
  }
  SrvEpiDual e{d, q.sigma, halpern_w(q.inner0, step), q.a2, q.b2, hat};
  tile_rows<double, double>(T.r0, T.r1, d->a_ptr, d->a_col, d->a_val, d->xb, e);
}

// Anchor constants of the increment form in fp64 (one rounding to fp32), per instance tile:
// primal tiles G = c̃ − Ãᵀy₀, LX = l̃ − x₀, UX = ũ − x₀ and δx = δx̄ = 0 (code 2); dual tiles
// BL = Ãx₀ − r̃l, BU = Ãx₀ − r̃u and δy = 0 (code 2), KY = y₀/σ (codes 1 and 2).
__global__ void __launch_bounds__(kBlock) k_srv_anchor_p(const Tile* __restrict__ tiles, int ntiles,
                                                         const Inst* __restrict__ I, const Param* __restrict__ P) {
  const int t = tile_of_block(blockIdx.x, gridDim.x);
  if (t >= ntiles) return;
  const Tile T = tiles[t];
  const Param q = P[T.inst];
  if (!q.active || q.anchor != 2) return;
  const Inst* d = I + T.inst;
  struct Epi {
    const Inst* d;
    __device__ void operator()(int j, bool has, double s) const {
      if (!has) return;
      const double x = d->x0[j];
      d->G[j] = static_cast<float>(d->c[j] - s);
      d->LX[j] = static_cast<float>(d->l[j] - x);
      d->UX[j] = static_cast<float>(d->u[j] - x);
      d->ix[j] = d->ixb[j] = 0.0f;
    }
  } e{d};
  tile_rows<double, double>(T.r0, T.r1, d->at_ptr, d->at_col, d->at_val, d->y0, e);
}

__global__ void __launch_bounds__(kBlock) k_srv_anchor_d(const Tile* __restrict__ tiles, int ntiles,
                                                         const Inst* __restrict__ I, const Param* __restrict__ P) {
  const int t = tile_of_block(blockIdx.x, gridDim.x);
  if (t >= ntiles) return;
  const Tile T = tiles[t];
  const Param q = P[T.inst];
  if (!q.active || q.anchor == 0) return;
  const Inst* d = I + T.inst;
  if (q.anchor == 1) {
    const int i = T.r0 + static_cast<int>(threadIdx.x);
    if (i < T.r1) d->KY[i] = static_cast<float>(d->y0[i] / q.sigma);
    return;
  }
  struct Epi {
    const Inst* d;
    double sigma;
    __device__ void operator()(int i, bool has, double s) const {
      if (!has) return;
      d->BL[i] = static_cast<float>(s - d->rl[i]);
      d->BU[i] = static_cast<float>(s - d->ru[i]);
      d->KY[i] = static_cast<float>(d->y0[i] / sigma);
      d->iy[i] = 0.0f;
    }
  } e{d, q.sigma};
  tile_rows<double, double>(T.r0, T.r1, d->a_ptr, d->a_col, d->a_val, d->x0, e);
}

__global__ void k_srv_to_f32(const double* __restrict__ a, float* __restrict__ b, long long n) {
  const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) b[i] = static_cast<float>(a[i]);
}

// ------------------------------------------------------------------ check kernels
// Fixed-order workgroup sum / max of one value per thread; the result is valid in thread 0.
__device__ double wg_reduce(double v, bool is_max, double* red) {
  red[threadIdx.x] = v;
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if (static_cast<int>(threadIdx.x) < s) {
      const double o = red[threadIdx.x + s];
      red[threadIdx.x] = is_max ? dmaxd(red[threadIdx.x], o) : red[threadIdx.x] + o;
    }
    __syncthreads();
  }
  const double r = red[0];
  __syncthreads();
  return r;
}

__device__ inline double elem_viol(double v, double lo, double up) {
  if (v < lo) return (lo - v) / (1.0 + fabs(lo));
  if (v > up) return (v - up) / (1.0 + fabs(up));
  return 0.0;
}

// Pass 1 (primal tiles, elementwise): Δx = x̄ − x̂; with kkt, the unscaled x and the PID term.
// r0 = 1: the epoch-start residual of instances whose chunk began an epoch.
__global__ void __launch_bounds__(kBlock) k_srv_cols1(const Tile* __restrict__ tiles, int ntiles, const Inst* __restrict__ I,
                                                      const Param* __restrict__ P, int kkt, int r0, int ray,
                                                      double* __restrict__ part) {
  __shared__ double red[kBlock];
  const int t = tile_of_block(blockIdx.x, gridDim.x);
  if (t >= ntiles) return;
  const Tile T = tiles[t];
  const Param q = P[T.inst];
  if (!q.active || (r0 && q.inner0 != 0.0)) return;
  const Inst* d = I + T.inst;
  const int j = T.r0 + static_cast<int>(threadIdx.x);
  double dx2 = 0, pid = 0, dmr = 0, cd = 0, cmag = 0;
  if (j < T.r1) {
    const double dd = d->xb[j] - d->xh[j];
    d->dx[j] = dd;
    dx2 = dd * dd;
    if (kkt) {
      d->xo[j] = clampv<double>(d->cs[j] * d->xh[j] / d->bs, d->clo[j], d->cuo[j]);
      const double e = d->xh[j] - d->x0[j];
      pid = e * e;
    }
    if (ray) {  // primal ray d = C·Δx/bs kept inside its sign constraints (engine.cpp:132-139)
      double v = d->cs[j] * dd / d->bs;
      if (isfinite(d->clo[j])) v = dmaxd(v, 0.0);
      if (isfinite(d->cuo[j])) v = dmind(v, 0.0);
      d->dv[j] = v;
      dmr = fabs(v);
      cd = d->cmin[j] * v;
      cmag = fabs(cd);
    }
  }
  double* out = part + static_cast<size_t>(t) * kQ;
  const double s1 = wg_reduce(dx2, false, red);
  if (threadIdx.x == 0) out[r0 ? Q_R0PX : Q_PX] = s1;
  if (kkt) {
    const double s2 = wg_reduce(pid, false, red);
    if (threadIdx.x == 0) out[Q_PIDX] = s2;
  }
  if (ray) {
    double v = wg_reduce(dmr, true, red);
    if (threadIdx.x == 0) out[Q_DMAXRAY] = v;
    v = wg_reduce(cd, false, red);
    if (threadIdx.x == 0) out[Q_CD] = v;
    v = wg_reduce(cmag, false, red);
    if (threadIdx.x == 0) out[Q_CMAG] = v;
  }
}

// Pass 2 (dual tiles): Δy, Ã·Δx and the cross term; with kkt, the unscaled y, A·x on the
// original values and the row terms of the KKT statistics.
__global__ void __launch_bounds__(kBlock) k_srv_rows(const Tile* __restrict__ tiles, int ntiles, const Inst* __restrict__ I,
                                                     const Param* __restrict__ P, int kkt, int r0, int ray,
                                                     double* __restrict__ part) {
  __shared__ double prod[kTileNnz];
  __shared__ double prod2[kTileNnz];
  __shared__ double red[kBlock];
  const int t = tile_of_block(blockIdx.x, gridDim.x);
  if (t >= ntiles) return;
  const Tile T = tiles[t];
  const Param q = P[T.inst];
  if (!q.active || (r0 && q.inner0 != 0.0)) return;
  const Inst* d = I + T.inst;
  const int i = T.r0 + static_cast<int>(threadIdx.x);
  const bool has = i < T.r1;
  const int p0 = d->a_ptr[T.r0], p1 = d->a_ptr[T.r1];
  const int rb = has ? d->a_ptr[i] : 0, re = has ? d->a_ptr[i + 1] : 0;
  double sdx = 0.0, sxo = 0.0;
  for (int cb = p0; cb < p1; cb += kTileNnz) {
    const int ce = p1 - cb > kTileNnz ? cb + kTileNnz : p1;
    for (int p = cb + static_cast<int>(threadIdx.x); p < ce; p += kBlock) {
      const int cj = d->a_col[p];
      prod[p - cb] = d->a_val[p] * d->dx[cj];
      if (kkt) prod2[p - cb] = d->a_orig[p] * d->xo[cj];
    }
    __syncthreads();
    if (has) {
      const int b = rb > cb ? rb : cb, e = re < ce ? re : ce;
      for (int p = b; p < e; ++p) sdx += prod[p - cb];
      if (kkt)
        for (int p = b; p < e; ++p) sxo += prod2[p - cb];
    }
    __syncthreads();
  }
  double sdv = 0.0, sa_unused = 0.0;  // A·d of the primal ray (original values)
  if (ray) tile_sums<double, double>(T.r0, T.r1, d->a_ptr, d->a_col, d->a_orig, d->dv, false, sdv, sa_unused);
  double dy2 = 0, cross = 0, pid = 0, rp2 = 0, b2 = 0, dt = 0, rd = 0, pmax = 0, dmax = 0;
  double rmax = 0, robj = 0, rmag = 0, pviol = 0;
  if (has) {
    const double dd = d->yb[i] - d->yh[i];
    d->dy[i] = dd;
    dy2 = dd * dd;
    cross = dd * sdx;
    if (ray) {  // dual ray r = R·Δy/os kept inside its sign constraints (engine.cpp:90-97), row terms
      const double lo = d->rlo[i], up = d->ruo[i];
      double r = d->rs[i] * dd / d->os;
      if (!isfinite(lo)) r = dmind(r, 0.0);
      if (!isfinite(up)) r = dmaxd(r, 0.0);
      d->rv[i] = r;
      rmax = fabs(r);
      robj = r > 0 ? lo * r : (r < 0 ? up * r : 0.0);
      rmag = fabs(robj);
      if (isfinite(lo)) pviol = dmaxd(pviol, -sdv);  // A·d against the finite row bounds (engine.cpp:140-145)
      if (isfinite(up)) pviol = dmaxd(pviol, sdv);
    }
    if (kkt) {
      const double y = d->rs[i] * d->yh[i] / d->os;
      d->yo[i] = y;
      const double e = d->yh[i] - d->y0[i];
      pid = e * e;
      const double lo = d->rlo[i], up = d->ruo[i];
      const double viol = sxo < lo ? lo - sxo : (sxo > up ? sxo - up : 0.0);
      rp2 = viol * viol;
      pmax = elem_viol(sxo, lo, up);
      const double bl = isfinite(lo) ? fabs(lo) : 0.0, bu = isfinite(up) ? fabs(up) : 0.0;
      b2 = dmaxd(bl, bu) * dmaxd(bl, bu);
      if (y > 0) {
        if (isfinite(lo)) dt = y * lo;
        else rd = y * y, dmax = y;
      } else if (y < 0) {
        if (isfinite(up)) dt = y * up;
        else rd = y * y, dmax = -y;
      }
    }
  }
  double* out = part + static_cast<size_t>(t) * kQ;
  double v;
  v = wg_reduce(dy2, false, red);
  if (threadIdx.x == 0) out[r0 ? Q_R0PY : Q_PY] = v;
  v = wg_reduce(cross, false, red);
  if (threadIdx.x == 0) out[r0 ? Q_R0CROSS : Q_CROSS] = v;
  if (ray) {
    v = wg_reduce(rmax, true, red);
    if (threadIdx.x == 0) out[Q_RMAX] = v;
    v = wg_reduce(robj, false, red);
    if (threadIdx.x == 0) out[Q_ROBJ] = v;
    v = wg_reduce(rmag, false, red);
    if (threadIdx.x == 0) out[Q_RMAG] = v;
    v = wg_reduce(pviol, true, red);
    if (threadIdx.x == 0) out[Q_PVIOL] = v;
  }
  if (!kkt) return;
  v = wg_reduce(pid, false, red);
  if (threadIdx.x == 0) out[Q_PIDY] = v;
  v = wg_reduce(rp2, false, red);
  if (threadIdx.x == 0) out[Q_RP2] = v;
  v = wg_reduce(b2, false, red);
  if (threadIdx.x == 0) out[Q_B2] = v;
  v = wg_reduce(dt, false, red);
  if (threadIdx.x == 0) out[Q_D] = v;
  v = wg_reduce(rd, false, red);
  if (threadIdx.x == 0) out[Q_RD2] = v;
  v = wg_reduce(pmax, true, red);
  if (threadIdx.x == 0) out[Q_PMAX] = v;
  v = wg_reduce(dmax, true, red);
  if (threadIdx.x == 0) out[Q_DMAX] = v;
}

// Pass 3 (primal tiles): Aᵀ·y on the original values and the column terms of the KKT statistics.
__global__ void __launch_bounds__(kBlock) k_srv_cols2(const Tile* __restrict__ tiles, int ntiles, const Inst* __restrict__ I,
                                                      const Param* __restrict__ P, int ray, double* __restrict__ part) {
  __shared__ double red[kBlock];
  const int t = tile_of_block(blockIdx.x, gridDim.x);
  if (t >= ntiles) return;
  const Tile T = tiles[t];
  const Param q = P[T.inst];
  if (!q.active) return;
  const Inst* d = I + T.inst;
  struct Cols {
    const Inst* d;
    double c2 = 0, cinf = 0, p = 0, pmax = 0, rd = 0, dmax = 0, dc = 0;
    __device__ void operator()(int j, bool has, double s) {
      if (!has) return;
      const double cj = d->cmin[j], xj = d->xo[j], lo = d->clo[j], up = d->cuo[j];
      c2 = cj * cj;
      cinf = fabs(cj);
      p = cj * xj;
      pmax = elem_viol(xj, lo, up);
      const double lam = cj - s;
      double z;
      if (isfinite(lo) && isfinite(up)) z = lam;
      else if (isfinite(lo)) z = dmaxd(lam, 0.0);
      else if (isfinite(up)) z = dmind(lam, 0.0);
      else z = 0.0;
      rd = (lam - z) * (lam - z);
      dmax = fabs(lam - z);
      dc = z > 0 ? z * lo : (z < 0 ? z * up : 0.0);
    }
  } e{d};
  tile_rows<double, double>(T.r0, T.r1, d->at_ptr, d->ato_col, d->ato_val, d->yo, e);
  double* out = part + static_cast<size_t>(t) * kQ;
  double v;
  v = wg_reduce(e.c2, false, red);
  if (threadIdx.x == 0) out[Q_C2] = v;
  v = wg_reduce(e.p, false, red);
  if (threadIdx.x == 0) out[Q_P] = v;
  v = wg_reduce(e.rd, false, red);
  if (threadIdx.x == 0) out[Q_RD2] = v;
  v = wg_reduce(e.dc, false, red);
  if (threadIdx.x == 0) out[Q_D] = v;
  v = wg_reduce(e.pmax, true, red);
  if (threadIdx.x == 0) out[Q_PMAX] = v;
  v = wg_reduce(e.dmax, true, red);
  if (threadIdx.x == 0) out[Q_DMAX] = v;
  v = wg_reduce(e.cinf, true, red);
  if (threadIdx.x == 0) out[Q_CINF] = v;
  if (!ray) return;
  // column terms of λ = −Aᵀr for the dual ray (engine.cpp:98-119)
  double g, ga;
  tile_sums<double, double>(T.r0, T.r1, d->at_ptr, d->ato_col, d->ato_val, d->rv, true, g, ga);
  double obj = 0, mag = 0, viol = 0;
  const int j = T.r0 + static_cast<int>(threadIdx.x);
  if (j < T.r1) {
    const double lam = -g, lo = d->clo[j], up = d->cuo[j];
    if (lam > 0) {
      if (isfinite(lo)) obj = lo * lam, mag = ga * fabs(lo);
      else viol = lam;
    } else if (lam < 0) {
      if (isfinite(up)) obj = up * lam, mag = ga * fabs(up);
      else viol = -lam;
    }
  }
  v = wg_reduce(obj, false, red);
  if (threadIdx.x == 0) out[Q_ROBJ] = v;
  v = wg_reduce(mag, false, red);
  if (threadIdx.x == 0) out[Q_RMAG] = v;
  v = wg_reduce(viol, true, red);
  if (threadIdx.x == 0) out[Q_DVIOL] = v;
}

// Per (instance, quantity): the instance's dual-tile partials then its primal-tile partials, in
// tile order (the reference's row part, then column part). Only quantities in `qmask`.
__global__ void k_srv_reduce(const Inst* __restrict__ I, const Param* __restrict__ P, int ninst, unsigned qmask,
                             const double* __restrict__ part_p, const double* __restrict__ part_d,
                             double* __restrict__ stats) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  const int k = g / kQ, q = g - k * kQ;
  if (k >= ninst || !(qmask & (1u << q)) || !P[k].active) return;
  const Inst& d = I[k];
  const bool mx = q_is_max(q);
  const bool from_d = q == Q_PY || q == Q_CROSS || q == Q_RP2 || q == Q_B2 || q == Q_D || q == Q_RD2 ||
                      q == Q_PIDY || q == Q_PMAX || q == Q_DMAX || q == Q_R0PY || q == Q_R0CROSS || q == Q_RMAX ||
                      q == Q_ROBJ || q == Q_RMAG || q == Q_PVIOL;
  const bool from_p = q == Q_PX || q == Q_D || q == Q_RD2 || q == Q_C2 || q == Q_P || q == Q_PIDX || q == Q_PMAX ||
                      q == Q_DMAX || q == Q_CINF || q == Q_R0PX || q == Q_ROBJ || q == Q_RMAG || q == Q_DVIOL ||
                      q == Q_DMAXRAY || q == Q_CD || q == Q_CMAG;
  double s = 0.0;
  if (from_d)
    for (int t = d.dt0; t < d.dt1; ++t) {
      const double v = part_d[static_cast<size_t>(t) * kQ + q];
      s = mx ? dmaxd(s, v) : s + v;
    }
  if (from_p)
    for (int t = d.pt0; t < d.pt1; ++t) {
      const double v = part_p[static_cast<size_t>(t) * kQ + q];
      s = mx ? dmaxd(s, v) : s + v;
    }
  stats[static_cast<size_t>(k) * kQ + q] = s;
}

// Restart (x ← x̂, x0 ← x̂ and the dual likewise) of the flagged instances.
__global__ void __launch_bounds__(kBlock) k_srv_restart(const Tile* __restrict__ tiles, int ntiles, const Inst* __restrict__ I,
                                                        const Param* __restrict__ P, int dual) {
  const int t = tile_of_block(blockIdx.x, gridDim.x);
  if (t >= ntiles) return;
  const Tile T = tiles[t];
  if (!P[T.inst].active || !P[T.inst].restart) return;
  const Inst* d = I + T.inst;
  const int r = T.r0 + static_cast<int>(threadIdx.x);
  if (r >= T.r1) return;
  if (dual) {
    const double v = d->yh[r];
    d->y[r] = v;
    d->y0[r] = v;
  } else {
    const double v = d->xh[r];
    d->x[r] = v;
    d->x0[r] = v;
  }
}

constexpr size_t kAlign = 256;
size_t align_up(size_t v) { return (v + kAlign - 1) / kAlign * kAlign; }

}  // namespace

struct psh_server {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  char* arena = nullptr;
  size_t budget = 0;
  std::map<size_t, size_t> free_blocks;  // offset → size
  struct Slot {
    bool used = false;
    size_t off = 0, bytes = 0;
    int m = 0, n = 0;
    long long nnz = 0;
    std::vector<int> tA, tAt;  // tile row starts of Ã and Ãᵀ
    Inst inst{};
    bool incr = false;         // increment-form storage present
    bool mixed = false;        // running in increment form
    bool dirty = true;         // anchor constants to rebuild before the next chunk
    double ky_sigma = 0.0;     // σ the constant KY was built with
  };
  std::vector<Slot> slots;
  int max_inst = 0;
  Inst* d_inst = nullptr;
  Param* d_param = nullptr;
  Param* h_param = nullptr;  // pinned
  Tile *d_ptiles = nullptr, *d_dtiles = nullptr;
  int ntp = 0, ntd = 0, cap_tiles_p = 0, cap_tiles_d = 0;
  double *d_part_p = nullptr, *d_part_d = nullptr;
  double* d_stats = nullptr;
  double* h_stats = nullptr;  // pinned
  bool tables_dirty = true;
  bool param_pending = false;
  std::map<long long, hipGraphExec_t> graphs;  // by chunk length
  bool use_graphs = true;
  // timing of whole chunks (HIP events around each chunk's launches on the serving stream, which
  // with graph replay carry no host gaps) with the chunk's iterations and algorithmic bytes
  bool timing = false;
  struct Timed { hipEvent_t a, b; int iterations; double bytes; };
  std::vector<Timed> timed;
  std::vector<hipEvent_t> ev_pool;
  long long t_launches = 0;
  double t_ms = 0.0, t_bytes = 0.0;
};

namespace {

int sfail(psh_server* s, int code, const std::string& what) {
  s->err = what;
  return code;
}
int shfail(psh_server* s, hipError_t e, const char* where) {
  s->err = std::string(where) + ": " + hipGetErrorString(e);
  (void)hipGetLastError();
  return e == hipErrorOutOfMemory ? PSH_E_OOM : PSH_E_HIP;
}
#define SRV_TRY(s, call)                                   \
  do {                                                     \
    hipError_t e_ = (call);                                \
    if (e_ != hipSuccess) return shfail(s, e_, #call);     \
  } while (0)
#define SRV_LAUNCHED(s, where)                              \
  do {                                                      \
    hipError_t e_ = hipGetLastError();                      \
    if (e_ != hipSuccess) return shfail(s, e_, where);      \
  } while (0)

// Arena layout of one instance: int arrays then double arrays, each 256-byte aligned.
struct Layout {
  size_t a_ptr, a_col, at_ptr, at_col, ato_col;
  size_t a_val, at_val, a_orig, ato_val;
  size_t c, l, u, cs, clo, cuo, cmin, x, x0, xh, xb, dx, xo, dv;
  size_t rl, ru, rs, rlo, ruo, y, y0, yh, yb, dy, yo, rv;
  size_t a_valf, at_valf, ix, ixb, G, LX, UX, iy, BL, BU, KY;  // with PSH_SERVE_INCREMENT
  size_t total;
};

Layout layout_of(long long m, long long n, long long nnz, int flags) {
  Layout L{};
  size_t o = 0;
  auto take = [&](size_t bytes) {
    const size_t at = o;
    o += align_up(std::max<size_t>(bytes, 1));
    return at;
  };
  L.a_ptr = take(4 * (m + 1));
  L.a_col = take(4 * nnz);
  L.at_ptr = take(4 * (n + 1));
  L.at_col = take(4 * nnz);
  L.ato_col = take(4 * nnz);
  L.a_val = take(8 * nnz);
  L.at_val = take(8 * nnz);
  L.a_orig = take(8 * nnz);
  L.ato_val = take(8 * nnz);
  for (size_t* f : {&L.c, &L.l, &L.u, &L.cs, &L.clo, &L.cuo, &L.cmin, &L.x, &L.x0, &L.xh, &L.xb, &L.dx, &L.xo, &L.dv})
    *f = take(8 * n);
  for (size_t* f : {&L.rl, &L.ru, &L.rs, &L.rlo, &L.ruo, &L.y, &L.y0, &L.yh, &L.yb, &L.dy, &L.yo, &L.rv})
    *f = take(8 * m);
  if (flags & PSH_SERVE_INCREMENT) {
    L.a_valf = take(4 * nnz);
    L.at_valf = take(4 * nnz);
    for (size_t* f : {&L.ix, &L.ixb, &L.G, &L.LX, &L.UX}) *f = take(4 * n);
    for (size_t* f : {&L.iy, &L.BL, &L.BU, &L.KY}) *f = take(4 * m);
  }
  L.total = o;
  return L;
}

bool arena_alloc(psh_server* s, size_t bytes, size_t* off) {
  for (auto it = s->free_blocks.begin(); it != s->free_blocks.end(); ++it)
    if (it->second >= bytes) {
      *off = it->first;
      const size_t rest = it->second - bytes;
      const size_t roff = it->first + bytes;
      s->free_blocks.erase(it);
      if (rest > 0) s->free_blocks[roff] = rest;
      return true;
    }
  return false;
}

void arena_free(psh_server* s, size_t off, size_t bytes) {
  auto it = s->free_blocks.emplace(off, bytes).first;
  auto next = std::next(it);
  if (next != s->free_blocks.end() && it->first + it->second == next->first) {
    it->second += next->second;
    s->free_blocks.erase(next);
  }
  if (it != s->free_blocks.begin()) {
    auto prev = std::prev(it);
    if (prev->first + prev->second == it->first) {
      prev->second += it->second;
      s->free_blocks.erase(it);
    }
  }
}

void drop_graphs(psh_server* s) {
  for (auto& kv : s->graphs) (void)hipGraphExecDestroy(kv.second);
  s->graphs.clear();
}

int wait_param(psh_server* s) {  // the pinned parameter table is reused by the next upload
  if (!s->param_pending) return PSH_OK;
  SRV_TRY(s, hipStreamSynchronize(s->stream));
  s->param_pending = false;
  return PSH_OK;
}

int upload_params(psh_server* s) {
  SRV_TRY(s, hipMemcpyAsync(s->d_param, s->h_param, sizeof(Param) * s->max_inst, hipMemcpyHostToDevice, s->stream));
  s->param_pending = true;
  return PSH_OK;
}

// Rebuilds the concatenated tile tables and the descriptors' tile ranges.
int rebuild_tables(psh_server* s) {
  if (!s->tables_dirty) return PSH_OK;
  std::vector<Tile> tp, td;
  for (int k = 0; k < s->max_inst; ++k) {
    auto& sl = s->slots[k];
    if (!sl.used) continue;
    sl.inst.pt0 = static_cast<int>(tp.size());
    for (size_t t = 0; t + 1 < sl.tAt.size(); ++t) tp.push_back(Tile{k, sl.tAt[t], sl.tAt[t + 1], 0});
    sl.inst.pt1 = static_cast<int>(tp.size());
    sl.inst.dt0 = static_cast<int>(td.size());
    for (size_t t = 0; t + 1 < sl.tA.size(); ++t) td.push_back(Tile{k, sl.tA[t], sl.tA[t + 1], 0});
    sl.inst.dt1 = static_cast<int>(td.size());
  }
  SRV_TRY(s, hipStreamSynchronize(s->stream));  // tables and partials may be in use
  s->param_pending = false;
  auto grow = [&](Tile** dt, double** part, int* cap, size_t need) -> int {
    if (static_cast<int>(need) <= *cap) return PSH_OK;
    const int nc = std::max<int>(static_cast<int>(need), std::max(1024, 2 * *cap));
    std::lock_guard<std::recursive_mutex> lock(psh::api_mutex());
    if (*dt) (void)hipFree(*dt);
    if (*part) (void)hipFree(*part);
    *dt = nullptr;
    *part = nullptr;
    SRV_TRY(s, hipMalloc(reinterpret_cast<void**>(dt), sizeof(Tile) * nc));
    SRV_TRY(s, hipMalloc(reinterpret_cast<void**>(part), sizeof(double) * kQ * nc));
    *cap = nc;
    drop_graphs(s);  // captured pointers changed
    return PSH_OK;
  };
  int rc;
  if ((rc = grow(&s->d_ptiles, &s->d_part_p, &s->cap_tiles_p, tp.size())) ||
      (rc = grow(&s->d_dtiles, &s->d_part_d, &s->cap_tiles_d, td.size())))
    return rc;
  if (!tp.empty())
    SRV_TRY(s, hipMemcpyAsync(s->d_ptiles, tp.data(), sizeof(Tile) * tp.size(), hipMemcpyHostToDevice, s->stream));
  if (!td.empty())
    SRV_TRY(s, hipMemcpyAsync(s->d_dtiles, td.data(), sizeof(Tile) * td.size(), hipMemcpyHostToDevice, s->stream));
  std::vector<Inst> desc(s->max_inst);
  for (int k = 0; k < s->max_inst; ++k)
    if (s->slots[k].used) desc[k] = s->slots[k].inst;
  SRV_TRY(s, hipMemcpyAsync(s->d_inst, desc.data(), sizeof(Inst) * s->max_inst, hipMemcpyHostToDevice, s->stream));
  SRV_TRY(s, hipStreamSynchronize(s->stream));  // the host tables go out of scope
  s->ntp = static_cast<int>(tp.size());
  s->ntd = static_cast<int>(td.size());
  s->tables_dirty = false;
  drop_graphs(s);  // tile counts changed
  return PSH_OK;
}

int grid_tiles(int ntiles) { return std::max(kXcd, grid_for_tiles(ntiles)); }

hipEvent_t take_event(psh_server* s) {
  if (!s->ev_pool.empty()) {
    hipEvent_t e = s->ev_pool.back();
    s->ev_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  return hipEventCreate(&e) == hipSuccess ? e : nullptr;
}

// Adds the timed chunks to the totals (synchronises).
int collect_timing(psh_server* s) {
  if (s->timed.empty()) return PSH_OK;
  SRV_TRY(s, hipStreamSynchronize(s->stream));
  for (auto& t : s->timed) {
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, t.a, t.b) == hipSuccess) {
      s->t_ms += ms;
      s->t_launches += t.iterations;
      s->t_bytes += t.bytes;
    }
    s->ev_pool.push_back(t.a);
    s->ev_pool.push_back(t.b);
  }
  s->timed.clear();
  return PSH_OK;
}

}  // namespace

extern "C" {

size_t psh_server_bytes(int m, int n, long long nnz, int flags) { return layout_of(m, n, nnz, flags).total; }

int psh_server_create(int device, size_t budget_bytes, int max_instances, psh_server** out) {
  if (!out || max_instances < 1 || budget_bytes < kAlign) return PSH_E_ARG;
  *out = nullptr;
  int nd = 0;
  if (hipGetDeviceCount(&nd) != hipSuccess || nd <= 0 || device < 0 || device >= nd) {
    (void)hipGetLastError();
    return PSH_E_NO_DEVICE;
  }
  if (hipSetDevice(device) != hipSuccess) return PSH_E_NO_DEVICE;
  auto* s = new psh_server;
  s->device = device;
  s->max_inst = max_instances;
  s->slots.resize(max_instances);
  s->budget = budget_bytes / kAlign * kAlign;
  auto bad = [&](hipError_t e) {
    (void)hipGetLastError();
    const int rc = e == hipErrorOutOfMemory ? PSH_E_OOM : PSH_E_HIP;
    psh_server_destroy(s);
    return rc;
  };
  hipError_t e;
  std::unique_lock<std::recursive_mutex> lock(psh::api_mutex());
  if ((e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking)) != hipSuccess) return bad(e);
  if ((e = hipMalloc(reinterpret_cast<void**>(&s->arena), s->budget)) != hipSuccess) return bad(e);
  if ((e = hipMalloc(reinterpret_cast<void**>(&s->d_inst), sizeof(Inst) * max_instances)) != hipSuccess) return bad(e);
  if ((e = hipMalloc(reinterpret_cast<void**>(&s->d_param), sizeof(Param) * max_instances)) != hipSuccess) return bad(e);
  if ((e = hipMalloc(reinterpret_cast<void**>(&s->d_stats), sizeof(double) * kQ * max_instances)) != hipSuccess)
    return bad(e);
  if ((e = hipHostMalloc(reinterpret_cast<void**>(&s->h_param), sizeof(Param) * max_instances)) != hipSuccess)
    return bad(e);
  if ((e = hipHostMalloc(reinterpret_cast<void**>(&s->h_stats), sizeof(double) * kQ * max_instances)) != hipSuccess)
    return bad(e);
  std::memset(s->h_param, 0, sizeof(Param) * max_instances);
  lock.unlock();
  if ((e = hipMemcpyAsync(s->d_param, s->h_param, sizeof(Param) * max_instances, hipMemcpyHostToDevice, s->stream)) !=
          hipSuccess ||
      (e = hipStreamSynchronize(s->stream)) != hipSuccess)
    return bad(e);
  s->free_blocks[0] = s->budget;
  *out = s;
  return PSH_OK;
}

void psh_server_destroy(psh_server* s) {
  if (!s) return;
  (void)hipSetDevice(s->device);
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  drop_graphs(s);
  for (auto& t : s->timed) s->ev_pool.push_back(t.a), s->ev_pool.push_back(t.b);
  for (hipEvent_t e : s->ev_pool) (void)hipEventDestroy(e);
  std::lock_guard<std::recursive_mutex> lock(psh::api_mutex());
  for (void* p : {static_cast<void*>(s->arena), static_cast<void*>(s->d_inst), static_cast<void*>(s->d_param),
                  static_cast<void*>(s->d_stats), static_cast<void*>(s->d_ptiles), static_cast<void*>(s->d_dtiles),
                  static_cast<void*>(s->d_part_p), static_cast<void*>(s->d_part_d)})
    if (p) (void)hipFree(p);
  if (s->h_param) (void)hipHostFree(s->h_param);
  if (s->h_stats) (void)hipHostFree(s->h_stats);
  if (s->stream) (void)hipStreamDestroy(s->stream);
  delete s;
}

const char* psh_server_error(const psh_server* s) { return s ? s->err.c_str() : "no server"; }

size_t psh_server_free_bytes(const psh_server* s) {
  if (!s) return 0;
  size_t best = 0;
  for (const auto& kv : s->free_blocks) best = std::max(best, kv.second);
  return best;
}

int psh_server_add(psh_server* s, const psh_problem* p, int flags, int* id) {
  if (!s || !p || !id) return PSH_E_ARG;
  *id = -1;
  if (p->m < 0 || p->n < 0 || p->nnz < 0 || p->nnz >= (1LL << 31)) return sfail(s, PSH_E_ARG, "bad problem sizes");
  if (flags & ~PSH_SERVE_INCREMENT) return sfail(s, PSH_E_ARG, "unknown instance flags");
  int k = 0;
  while (k < s->max_inst && s->slots[k].used) ++k;
  if (k == s->max_inst) return sfail(s, PSH_E_OOM, "all instance slots are in use");
  const Layout L = layout_of(p->m, p->n, p->nnz, flags);
  size_t off = 0;
  if (!arena_alloc(s, L.total, &off)) return sfail(s, PSH_E_OOM, "HBM budget exhausted");
  SRV_TRY(s, hipSetDevice(s->device));
  auto& sl = s->slots[k];
  sl = psh_server::Slot{};
  sl.used = true;
  sl.off = off;
  sl.bytes = L.total;
  sl.m = p->m;
  sl.n = p->n;
  sl.nnz = p->nnz;
  char* base = s->arena + off;
  Inst& d = sl.inst;
  d.m = p->m;
  d.n = p->n;
  auto I = [&](size_t o) { return reinterpret_cast<int*>(base + o); };
  auto D = [&](size_t o) { return reinterpret_cast<double*>(base + o); };
  d.a_ptr = I(L.a_ptr);
  d.a_col = I(L.a_col);
  d.at_ptr = I(L.at_ptr);
  d.at_col = I(L.at_col);
  d.ato_col = I(L.ato_col);
  d.a_val = D(L.a_val);
  d.at_val = D(L.at_val);
  d.a_orig = D(L.a_orig);
  d.ato_val = D(L.ato_val);
  d.c = D(L.c), d.l = D(L.l), d.u = D(L.u), d.cs = D(L.cs), d.clo = D(L.clo), d.cuo = D(L.cuo), d.cmin = D(L.cmin);
  d.x = D(L.x), d.x0 = D(L.x0), d.xh = D(L.xh), d.xb = D(L.xb), d.dx = D(L.dx), d.xo = D(L.xo), d.dv = D(L.dv);
  d.rl = D(L.rl), d.ru = D(L.ru), d.rs = D(L.rs), d.rlo = D(L.rlo), d.ruo = D(L.ruo);
  d.y = D(L.y), d.y0 = D(L.y0), d.yh = D(L.yh), d.yb = D(L.yb), d.dy = D(L.dy), d.yo = D(L.yo), d.rv = D(L.rv);
  if (flags & PSH_SERVE_INCREMENT) {
    auto F = [&](size_t o) { return reinterpret_cast<float*>(base + o); };
    d.a_valf = F(L.a_valf);
    d.at_valf = F(L.at_valf);
    d.ix = F(L.ix), d.ixb = F(L.ixb), d.G = F(L.G), d.LX = F(L.LX), d.UX = F(L.UX);
    d.iy = F(L.iy), d.BL = F(L.BL), d.BU = F(L.BU), d.KY = F(L.KY);
    sl.incr = sl.mixed = true;
  }
  d.bs = p->bound_scale;
  d.os = p->obj_scale;
  const long long m = p->m, n = p->n, nnz = p->nnz;
  auto cp = [&](const void* dst, const void* src, size_t bytes) -> int {
    if (bytes == 0) return PSH_OK;
    if (!src) return sfail(s, PSH_E_ARG, "null problem array");
    SRV_TRY(s, hipMemcpyAsync(const_cast<void*>(dst), src, bytes, hipMemcpyHostToDevice, s->stream));
    return PSH_OK;
  };
  int rc;
  if ((rc = cp(d.a_ptr, p->a_ptr, 4 * (m + 1))) || (rc = cp(d.a_col, p->a_col, 4 * nnz)) ||
      (rc = cp(d.at_ptr, p->at_ptr, 4 * (n + 1))) || (rc = cp(d.at_col, p->at_col, 4 * nnz)) ||
      (rc = cp(d.ato_col, p->ato_col ? p->ato_col : p->at_col, 4 * nnz)) || (rc = cp(d.a_val, p->a_val, 8 * nnz)) ||
      (rc = cp(d.at_val, p->at_val, 8 * nnz)) || (rc = cp(d.a_orig, p->a_val_orig, 8 * nnz)) ||
      (rc = cp(d.ato_val, p->ato_val, 8 * nnz)) || (rc = cp(d.c, p->c, 8 * n)) || (rc = cp(d.l, p->l, 8 * n)) ||
      (rc = cp(d.u, p->u, 8 * n)) || (rc = cp(d.cs, p->col_scale, 8 * n)) || (rc = cp(d.clo, p->cl_o, 8 * n)) ||
      (rc = cp(d.cuo, p->cu_o, 8 * n)) || (rc = cp(d.cmin, p->c_min, 8 * n)) || (rc = cp(d.rl, p->rl, 8 * m)) ||
      (rc = cp(d.ru, p->ru, 8 * m)) || (rc = cp(d.rs, p->row_scale, 8 * m)) || (rc = cp(d.rlo, p->rl_o, 8 * m)) ||
      (rc = cp(d.ruo, p->ru_o, 8 * m))) {
    arena_free(s, sl.off, sl.bytes);
    sl = psh_server::Slot{};
    return rc;
  }
  if (sl.incr && nnz > 0) {
    hipLaunchKernelGGL(k_srv_to_f32, dim3(grid_for(nnz)), dim3(kBlock), 0, s->stream, d.a_val,
                       const_cast<float*>(d.a_valf), nnz);
    hipLaunchKernelGGL(k_srv_to_f32, dim3(grid_for(nnz)), dim3(kBlock), 0, s->stream, d.at_val,
                       const_cast<float*>(d.at_valf), nnz);
    SRV_LAUNCHED(s, "fp32 values");
  }
  SRV_TRY(s, hipStreamSynchronize(s->stream));  // the caller's arrays may go away
  sl.tA = build_tiles(p->a_ptr, p->m);
  sl.tAt = build_tiles(p->at_ptr, p->n);
  s->tables_dirty = true;
  *id = k;
  return PSH_OK;
}

int psh_server_remove(psh_server* s, int id) {
  if (!s || id < 0 || id >= s->max_inst || !s->slots[id].used) return PSH_E_ARG;
  int rc;
  if ((rc = wait_param(s))) return rc;
  SRV_TRY(s, hipStreamSynchronize(s->stream));
  s->h_param[id] = Param{};
  arena_free(s, s->slots[id].off, s->slots[id].bytes);
  s->slots[id] = psh_server::Slot{};
  s->tables_dirty = true;
  return upload_params(s);
}

// Start point (scaled): x, y; x0 = x̂ = x, y0 = ŷ = y.
int psh_server_set_iterate(psh_server* s, int id, const double* x, const double* y) {
  if (!s || id < 0 || id >= s->max_inst || !s->slots[id].used) return PSH_E_ARG;
  const Inst& d = s->slots[id].inst;
  for (double* dst : {d.x, d.x0, d.xh, d.xb})
    if (d.n > 0) SRV_TRY(s, hipMemcpyAsync(dst, x, 8 * static_cast<size_t>(d.n), hipMemcpyHostToDevice, s->stream));
  for (double* dst : {d.y, d.y0, d.yh, d.yb})
    if (d.m > 0) SRV_TRY(s, hipMemcpyAsync(dst, y, 8 * static_cast<size_t>(d.m), hipMemcpyHostToDevice, s->stream));
  SRV_TRY(s, hipStreamSynchronize(s->stream));
  s->slots[id].dirty = true;
  return PSH_OK;
}

int psh_server_set_precision(psh_server* s, int id, int mixed) {
  if (!s || id < 0 || id >= s->max_inst || !s->slots[id].used) return PSH_E_ARG;
  auto& sl = s->slots[id];
  if (mixed && !sl.incr) return sfail(s, PSH_E_STATE, "instance was added without increment-form storage");
  if (sl.mixed != (mixed != 0)) sl.dirty = true;
  sl.mixed = mixed != 0;
  return PSH_OK;
}

// `count` iterations for the listed instances (others pause). params[k] = {id, τ, σ, inner0}.
int psh_server_chunk(psh_server* s, int iterations, const psh_serve_param* params, int count) {
  if (!s || iterations < 1 || count < 0 || (count > 0 && !params)) return PSH_E_ARG;
  int rc;
  if ((rc = wait_param(s))) return rc;
  for (int k = 0; k < s->max_inst; ++k) s->h_param[k].active = 0, s->h_param[k].restart = 0;
  for (int k = 0; k < count; ++k) {
    const int id = params[k].id;
    if (id < 0 || id >= s->max_inst || !s->slots[id].used) return sfail(s, PSH_E_ARG, "bad instance id");
    if (params[k].inner0 < 0 || params[k].inner0 > (1LL << 52)) return sfail(s, PSH_E_ARG, "epoch counter out of range");
    if (!std::isfinite(params[k].rho)) return sfail(s, PSH_E_ARG, "reflection is not finite");
    Param& q = s->h_param[id];
    q.tau = params[k].tau;
    q.sigma = params[k].sigma;
    q.inner0 = static_cast<double>(params[k].inner0);
    q.a2 = 2 * params[k].rho;
    q.b2 = 1 - 2 * params[k].rho;
    q.active = 1;
  }
  bool any_anchor = false;
  for (int k = 0; k < count; ++k) {
    const int id = params[k].id;
    auto& sl = s->slots[id];
    Param& q = s->h_param[id];
    q.mixed = sl.mixed ? 1 : 0;
    q.anchor = !sl.mixed ? 0 : (sl.dirty ? 2 : (q.sigma != sl.ky_sigma ? 1 : 0));
    any_anchor = any_anchor || q.anchor != 0;
    if (sl.mixed) sl.dirty = false, sl.ky_sigma = q.sigma;
  }
  if ((rc = rebuild_tables(s))) return rc;
  if ((rc = upload_params(s))) return rc;
  if (count == 0 || (s->ntp == 0 && s->ntd == 0)) return PSH_OK;
  if (any_anchor) {
    if (s->ntp > 0)
      hipLaunchKernelGGL(k_srv_anchor_p, dim3(grid_tiles(s->ntp)), dim3(kBlock), 0, s->stream, s->d_ptiles, s->ntp,
                         s->d_inst, s->d_param);
    if (s->ntd > 0)
      hipLaunchKernelGGL(k_srv_anchor_d, dim3(grid_tiles(s->ntd)), dim3(kBlock), 0, s->stream, s->d_dtiles, s->ntd,
                         s->d_inst, s->d_param);
    SRV_LAUNCHED(s, "serving anchors");
  }
  auto enqueue = [&]() -> int {
    for (int st = 0; st < iterations; ++st) {
      const int last = st == iterations - 1 ? 1 : 0;
      if (s->ntp > 0)
        hipLaunchKernelGGL(k_srv_primal, dim3(grid_tiles(s->ntp)), dim3(kBlock), 0, s->stream, s->d_ptiles, s->ntp,
                           s->d_inst, s->d_param, st, last);
      if (s->ntd > 0)
        hipLaunchKernelGGL(k_srv_dual, dim3(grid_tiles(s->ntd)), dim3(kBlock), 0, s->stream, s->d_dtiles, s->ntd,
                           s->d_inst, s->d_param, st, last);
      if (st == 0) {  // residual after the first step of instances that begin an epoch (r2hpdhg.cpp:75)
        if (s->ntp > 0)
          hipLaunchKernelGGL(k_srv_cols1, dim3(grid_tiles(s->ntp)), dim3(kBlock), 0, s->stream, s->d_ptiles, s->ntp,
                             s->d_inst, s->d_param, 0, 1, 0, s->d_part_p);
        if (s->ntd > 0)
          hipLaunchKernelGGL(k_srv_rows, dim3(grid_tiles(s->ntd)), dim3(kBlock), 0, s->stream, s->d_dtiles, s->ntd,
                             s->d_inst, s->d_param, 0, 1, 0, s->d_part_d);
        const unsigned mask = (1u << Q_R0PX) | (1u << Q_R0PY) | (1u << Q_R0CROSS);
        hipLaunchKernelGGL(k_srv_reduce, dim3((s->max_inst * kQ + kBlock - 1) / kBlock), dim3(kBlock), 0, s->stream,
                           s->d_inst, s->d_param, s->max_inst, mask, s->d_part_p, s->d_part_d, s->d_stats);
      }
    }
    SRV_LAUNCHED(s, "serving iteration kernels");
    return PSH_OK;
  };
  psh_server::Timed tm{nullptr, nullptr, iterations, 0.0};
  if (s->timing) {  // events around the chunk's launches; bytes of its iterations over the active instances
    for (int k = 0; k < s->max_inst; ++k) {
      if (!s->h_param[k].active) continue;
      const auto& sl = s->slots[k];
      const double m = sl.m, n = sl.n, nz = static_cast<double>(sl.nnz);
      const double pb = sl.mixed ? 8 * nz + 4 * (n + 1) + 4 * m + 24 * n : 12 * nz + 4 * (n + 1) + 8 * m + 56 * n;
      const double db = sl.mixed ? 8 * nz + 4 * (m + 1) + 4 * n + 20 * m : 12 * nz + 4 * (m + 1) + 8 * n + 40 * m;
      tm.bytes += iterations * (pb + db);
    }
    if (s->timed.size() >= 4096 && (rc = collect_timing(s))) return rc;
    tm.a = take_event(s);
    tm.b = take_event(s);
    if (!tm.a || !tm.b) return sfail(s, PSH_E_HIP, "timing events could not be created");
    SRV_TRY(s, hipEventRecord(tm.a, s->stream));
  }
  auto timed_end = [&]() -> int {
    if (!s->timing) return PSH_OK;
    SRV_TRY(s, hipEventRecord(tm.b, s->stream));
    s->timed.push_back(tm);
    return PSH_OK;
  };
  if (!s->use_graphs) {
    if ((rc = enqueue())) return rc;
    return timed_end();
  }

  const long long key = iterations;
  auto it = s->graphs.find(key);
  if (it == s->graphs.end()) {
    std::lock_guard<std::recursive_mutex> capture_lock(psh::api_mutex());  // no device-wide calls meanwhile
    SRV_TRY(s, hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
    const int erc = enqueue();
    hipGraph_t graph = nullptr;
    const hipError_t ec = hipStreamEndCapture(s->stream, &graph);
    if (erc) {
      if (graph) (void)hipGraphDestroy(graph);
      return erc;
    }
    if (ec != hipSuccess) return shfail(s, ec, "hipStreamEndCapture");
    hipGraphExec_t exec = nullptr;
    const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ei != hipSuccess) return shfail(s, ei, "hipGraphInstantiate");
    it = s->graphs.emplace(key, exec).first;
  }
  SRV_TRY(s, hipGraphLaunch(it->second, s->stream));
  return timed_end();
}

int psh_server_graphs(psh_server* s, int on) {
  if (!s) return PSH_E_ARG;
  s->use_graphs = on != 0;
  return PSH_OK;
}

int psh_server_timing(psh_server* s, int on) {
  if (!s) return PSH_E_ARG;
  s->timing = on != 0;
  return PSH_OK;
}

int psh_server_timing_get(psh_server* s, long long* iterations, double* total_ms, double* bytes) {
  if (!s) return PSH_E_ARG;
  int rc;
  if ((rc = collect_timing(s))) return rc;
  if (iterations) *iterations = s->t_launches;
  if (total_ms) *total_ms = s->t_ms;
  if (bytes) *bytes = s->t_bytes;
  return PSH_OK;
}

// Check of the instances active in the last chunk: out[k] for ids[k] (residual sums, r⁰ sums
// when that chunk began the instance's epoch, KKT on the original problem, PID distances, and
// with want_ray the ray-test scalars of T(z) − z).
int psh_server_check(psh_server* s, const int* ids, int count, int want_ray, psh_stats* out) {
  if (!s || count < 0 || (count > 0 && (!ids || !out))) return PSH_E_ARG;
  if (s->tables_dirty) return sfail(s, PSH_E_STATE, "instance set changed since the last chunk");
  const int ray = want_ray ? 1 : 0;
  // the primal ray (cols1) is read by the row pass, the dual ray (rows) by the column pass
  if (s->ntp > 0)
    hipLaunchKernelGGL(k_srv_cols1, dim3(grid_tiles(s->ntp)), dim3(kBlock), 0, s->stream, s->d_ptiles, s->ntp, s->d_inst,
                       s->d_param, 1, 0, ray, s->d_part_p);
  if (s->ntd > 0)
    hipLaunchKernelGGL(k_srv_rows, dim3(grid_tiles(s->ntd)), dim3(kBlock), 0, s->stream, s->d_dtiles, s->ntd, s->d_inst,
                       s->d_param, 1, 0, ray, s->d_part_d);
  if (s->ntp > 0)
    hipLaunchKernelGGL(k_srv_cols2, dim3(grid_tiles(s->ntp)), dim3(kBlock), 0, s->stream, s->d_ptiles, s->ntp, s->d_inst,
                       s->d_param, ray, s->d_part_p);
  unsigned mask = ((1u << kQ) - 1) & ~((1u << Q_R0PX) | (1u << Q_R0PY) | (1u << Q_R0CROSS));
  if (!ray) mask &= ~kRayMask;
  hipLaunchKernelGGL(k_srv_reduce, dim3((s->max_inst * kQ + kBlock - 1) / kBlock), dim3(kBlock), 0, s->stream, s->d_inst,
                     s->d_param, s->max_inst, mask, s->d_part_p, s->d_part_d, s->d_stats);
  SRV_LAUNCHED(s, "serving check kernels");
  SRV_TRY(s, hipMemcpyAsync(s->h_stats, s->d_stats, sizeof(double) * kQ * s->max_inst, hipMemcpyDeviceToHost, s->stream));
  SRV_TRY(s, hipStreamSynchronize(s->stream));
  s->param_pending = false;
  for (int k = 0; k < count; ++k) {
    const int id = ids[k];
    if (id < 0 || id >= s->max_inst || !s->slots[id].used || !s->h_param[id].active)
      return sfail(s, PSH_E_ARG, "check of an instance that did not run");
    const double* q = s->h_stats + static_cast<size_t>(id) * kQ;
    psh_stats& o = out[k];
    std::memset(&o, 0, sizeof o);
    o.px = q[Q_PX];
    o.py = q[Q_PY];
    o.cross = q[Q_CROSS];
    o.have_r0 = s->h_param[id].inner0 == 0.0 ? 1 : 0;
    if (o.have_r0) o.r0_px = q[Q_R0PX], o.r0_py = q[Q_R0PY], o.r0_cross = q[Q_R0CROSS];
    o.kkt.primal_residual = std::sqrt(q[Q_RP2]);
    o.kkt.dual_residual = std::sqrt(q[Q_RD2]);
    o.kkt.primal_obj = q[Q_P];
    o.kkt.dual_obj = q[Q_D];
    o.kkt.b_norm = std::sqrt(q[Q_B2]);
    o.kkt.c_norm = std::sqrt(q[Q_C2]);
    o.kkt.primal_max_rel = q[Q_PMAX];
    o.kkt.dual_max_rel = q[Q_DMAX] / (1.0 + q[Q_CINF]);
    o.dx_norm = std::sqrt(q[Q_PIDX]);
    o.dy_norm = std::sqrt(q[Q_PIDY]);
    if (ray) {
      o.ray.rmax = q[Q_RMAX];
      o.ray.dual_obj = q[Q_ROBJ];
      o.ray.dual_viol = q[Q_DVIOL];
      o.ray.dual_mag = q[Q_RMAG];
      o.ray.dmax = q[Q_DMAXRAY];
      o.ray.primal_cd = q[Q_CD];
      o.ray.primal_viol = q[Q_PVIOL];
      o.ray.primal_cmag = q[Q_CMAG];
    }
    bool fin = true;
    for (int j = 0; j < kQ; ++j) fin = fin && (std::isfinite(q[j]) || (!ray && ((kRayMask >> j) & 1u)));
    o.finite = fin ? 1 : 0;
  }
  return PSH_OK;
}

int psh_server_restart(psh_server* s, const int* ids, int count) {
  if (!s || count < 0 || (count > 0 && !ids)) return PSH_E_ARG;
  if (count == 0) return PSH_OK;
  if (s->tables_dirty) return sfail(s, PSH_E_STATE, "instance set changed since the last chunk");
  int rc;
  if ((rc = wait_param(s))) return rc;
  for (int k = 0; k < count; ++k) {
    const int id = ids[k];
    if (id < 0 || id >= s->max_inst || !s->slots[id].used) return sfail(s, PSH_E_ARG, "bad instance id");
    s->h_param[id].restart = 1;
    s->slots[id].dirty = true;  // increment form: new anchors at the next chunk
  }
  if ((rc = upload_params(s))) return rc;
  if (s->ntp > 0)
    hipLaunchKernelGGL(k_srv_restart, dim3(grid_tiles(s->ntp)), dim3(kBlock), 0, s->stream, s->d_ptiles, s->ntp, s->d_inst,
                       s->d_param, 0);
  if (s->ntd > 0)
    hipLaunchKernelGGL(k_srv_restart, dim3(grid_tiles(s->ntd)), dim3(kBlock), 0, s->stream, s->d_dtiles, s->ntd, s->d_inst,
                       s->d_param, 1);
  SRV_LAUNCHED(s, "serving restart");
  return PSH_OK;  // the flags stay set in the pinned table until the next chunk clears them
}

int psh_server_ray_vectors(psh_server* s, int id, double* dual_ray, double* primal_ray) {
  if (!s || id < 0 || id >= s->max_inst || !s->slots[id].used) return PSH_E_ARG;
  const Inst& d = s->slots[id].inst;
  if (dual_ray && d.m > 0)
    SRV_TRY(s, hipMemcpyAsync(dual_ray, d.rv, 8 * static_cast<size_t>(d.m), hipMemcpyDeviceToHost, s->stream));
  if (primal_ray && d.n > 0)
    SRV_TRY(s, hipMemcpyAsync(primal_ray, d.dv, 8 * static_cast<size_t>(d.n), hipMemcpyDeviceToHost, s->stream));
  SRV_TRY(s, hipStreamSynchronize(s->stream));
  s->param_pending = false;
  return PSH_OK;
}

// x̂ (n) and ŷ (m), scaled, of one instance.
int psh_server_download(psh_server* s, int id, double* xh, double* yh) {
  if (!s || id < 0 || id >= s->max_inst || !s->slots[id].used) return PSH_E_ARG;
  const Inst& d = s->slots[id].inst;
  if (xh && d.n > 0) SRV_TRY(s, hipMemcpyAsync(xh, d.xh, 8 * static_cast<size_t>(d.n), hipMemcpyDeviceToHost, s->stream));
  if (yh && d.m > 0) SRV_TRY(s, hipMemcpyAsync(yh, d.yh, 8 * static_cast<size_t>(d.m), hipMemcpyDeviceToHost, s->stream));
  SRV_TRY(s, hipStreamSynchronize(s->stream));
  s->param_pending = false;
  return PSH_OK;
}

}  // extern "C"
