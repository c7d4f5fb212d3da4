// Device-resident conjugate gradient pieces for the normal equations
// (P^T P + lam (I + sum_k D_k^T D_k)) x = P^T y that lo.acg solves in the
// reference (srt.py:225, models.py:85-94; lo is not vendored).
//
// D_k is the forward difference along axis k with the numpy.diff shape
// convention (output shorter by one), so D_k^T D_k x at index m along k is
//   [m >= 1] (x_m - x_{m-1}) - [m <= n-2] (x_{m+1} - x_m).
// With a per-voxel weight w (the height prior of models.py:85-104, where each
// D_k is composed with diag(r^2)), the prior term is ||D_k diag(w) x||^2 and
// its normal-equation operator diag(w) D_k^T D_k diag(w) x.
// Dot products are reduced in a fixed order (per-block partials, then one
// block), so a CG run is reproducible given reproducible operator outputs.
#include <hip/hip_runtime.h>

#include "siddon_common.h"
#include "siddon_host.h"

namespace tg {
namespace {

constexpr int kRedBlocks = 1024;  // max partials, lives at scal[16 .. 16 + kRedBlocks)
constexpr int kScalPartials = 16;

// k / d for k < 2^32 from the divisor's double reciprocal rd = 1 / d: k * rd
// is within one of the quotient, and one test of the remainder corrects it.
__device__ __forceinline__ uint32_t udiv_rcp(uint32_t k, uint32_t d, double rd) {
  const uint32_t q = (uint32_t)((double)k * rd);
  const int64_t rem = (int64_t)k - (int64_t)q * (int64_t)d;
  return rem < 0 ? q - 1u : (rem >= (int64_t)d ? q + 1u : q);
}

// A C-order array of shape n: its strides and, for arrays of at most 2^32
// elements, the divisors and reciprocals that turn a flat index into
// coordinates with 32-bit arithmetic (the integer divisions, 64-bit ones above
// all, are emulated at many times the cost).  Computed once per thread.
struct GridDiv {
  int64_t n[4], st[4];
  uint32_t s[3];
  double r[3];
  bool small;
};

__device__ __forceinline__ GridDiv grid_div(const int64_t n[4]) {
  GridDiv d;
  for (int a = 0; a < 4; a++) d.n[a] = n[a];
  d.st[3] = 1;
  d.st[2] = n[3];
  d.st[1] = n[2] * n[3];
  d.st[0] = n[1] * n[2] * n[3];
  d.small = n[0] * d.st[0] <= (int64_t)0xffffffff;
  for (int a = 0; a < 3; a++) {
    d.s[a] = (uint32_t)d.st[a];
    d.r[a] = 1.0 / (double)(d.s[a] ? d.s[a] : 1u);
  }
  return d;
}

// Coordinates m of flat index k.
__device__ __forceinline__ void grid_pos(const GridDiv &d, int64_t k, int64_t m[4]) {
  if (d.small) {
    const uint32_t kk = (uint32_t)k;
    const uint32_t a = udiv_rcp(kk, d.s[0], d.r[0]), r0 = kk - a * d.s[0];
    const uint32_t b = udiv_rcp(r0, d.s[1], d.r[1]), r1 = r0 - b * d.s[1];
    const uint32_t c = udiv_rcp(r1, d.s[2], d.r[2]);
    m[0] = a;
    m[1] = b;
    m[2] = c;
    m[3] = r1 - c * d.s[2];
  } else {
    m[0] = k / d.st[0];
    m[1] = (k / d.st[1]) % d.n[1];
    m[2] = (k / d.st[2]) % d.n[2];
    m[3] = k % d.n[3];
  }
}

// Global-memory pointers: loads through them are global, not flat, loads
// (pointers read from a table are generic otherwise).
template <typename T>
using gptr = __attribute__((address_space(1))) T *;

// (identity ? x_k : 0) + sum_k D_k^T D_k x at flat index k of the array d
// describes, or with the voxel weight w (not null)
// (identity ? x_k : 0) + w_k sum_k D_k^T D_k (w x).  P: the pointer type.
// Every operand is loaded (a neighbour outside the array at the element
// itself) before any is used, so the loads are in flight together; only the
// sums depend on the bounds.
template <typename T, typename P>
__device__ __forceinline__ T reg_at(const GridDiv &d, uint32_t axes, int identity, P x, P w,
                                    int64_t k) {
  int64_t m[4];
  grid_pos(d, k, m);
  int64_t lo[4], hi[4];
#pragma unroll
  for (int ax = 0; ax < 4; ax++) {
    lo[ax] = m[ax] >= 1 ? k - d.st[ax] : k;
    hi[ax] = m[ax] <= d.n[ax] - 2 ? k + d.st[ax] : k;
  }
  if (!w) {
    const T xc = x[k];
    T xl[4], xh[4];
#pragma unroll
    for (int ax = 0; ax < 4; ax++) {
      xl[ax] = (axes & (1u << ax)) ? x[lo[ax]] : (T)0;
      xh[ax] = (axes & (1u << ax)) ? x[hi[ax]] : (T)0;
    }
    T acc = identity ? xc : (T)0;
#pragma unroll
    for (int ax = 0; ax < 4; ax++) {
      if (!(axes & (1u << ax))) continue;
      if (m[ax] >= 1) acc += xc - xl[ax];
      if (m[ax] <= d.n[ax] - 2) acc -= xh[ax] - xc;
    }
    return acc;
  }
  const T xc = x[k], wc = w[k];
  T xl[4], xh[4], wl[4], wh[4];
#pragma unroll
  for (int ax = 0; ax < 4; ax++) {
    const bool on = (axes & (1u << ax)) != 0;
    xl[ax] = on ? x[lo[ax]] : (T)0;
    xh[ax] = on ? x[hi[ax]] : (T)0;
    wl[ax] = on ? w[lo[ax]] : (T)0;
    wh[ax] = on ? w[hi[ax]] : (T)0;
  }
  const T acc = identity ? xc : (T)0;
  const T uc = wc * xc;
  T dd = (T)0;
#pragma unroll
  for (int ax = 0; ax < 4; ax++) {
    if (!(axes & (1u << ax))) continue;
    if (m[ax] >= 1) dd += uc - wl[ax] * xl[ax];
    if (m[ax] <= d.n[ax] - 2) dd -= wh[ax] * xh[ax] - uc;
  }
  return acc + wc * dd;
}

template <typename T>
__global__ __launch_bounds__(256) void k_regularize(int64_t n0, int64_t n1, int64_t n2,
                                                    int64_t n3, uint32_t axes, T lam,
                                                    int identity, const T *__restrict__ x,
                                                    T *__restrict__ out,
                                                    const uint8_t *__restrict__ mask,
                                                    const T *__restrict__ w) {
  const int64_t total = n0 * n1 * n2 * n3;
  const int64_t n[4] = {n0, n1, n2, n3};
  const GridDiv gd = grid_div(n);
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < total;
       k += (int64_t)gridDim.x * blockDim.x) {
    if (mask && mask[k]) continue;
    out[k] = out[k] + lam * reg_at<T>(gd, axes, identity, x, w, k);
  }
}

// Smoothness prior sum_k w_k ||D_k x||^2 and its gradient 2 sum_k w_k D_k^T D_k x
// in one pass: per-block partial sums of the value (double) at
// scal[kScalPartials + blockIdx.x], the gradient (in T) at grad (if not null).
// With a voxel weight vw (not null) the differences are those of u = vw x and
// the gradient is 2 vw sum_k w_k D_k^T D_k u.
template <typename T>
__global__ __launch_bounds__(256) void k_smooth(int64_t n0, int64_t n1, int64_t n2, int64_t n3,
                                                double w0, double w1, double w2, double w3,
                                                const T *__restrict__ x, T *__restrict__ grad,
                                                double *scal, const T *__restrict__ vw) {
  __shared__ double lds[256];
  const int64_t total = n0 * n1 * n2 * n3;
  const int64_t n[4] = {n0, n1, n2, n3};
  const GridDiv gd = grid_div(n);
  const int64_t *st = gd.st;
  const double w[4] = {w0, w1, w2, w3};
  double v = 0.0;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < total;
       k += (int64_t)gridDim.x * blockDim.x) {
    int64_t m[4];
    grid_pos(gd, k, m);
    const T xc = vw ? vw[k] * x[k] : x[k];
    // both neighbours of every axis loaded at once (at k where outside)
    T un[4], up[4];
#pragma unroll
    for (int ax = 0; ax < 4; ax++) {
      const int64_t kn = m[ax] <= n[ax] - 2 ? k + st[ax] : k;
      const int64_t kp = m[ax] >= 1 ? k - st[ax] : k;
      un[ax] = w[ax] != 0.0 ? (vw ? vw[kn] * x[kn] : x[kn]) : (T)0;
      up[ax] = w[ax] != 0.0 ? (vw ? vw[kp] * x[kp] : x[kp]) : (T)0;
    }
    T g = (T)0;
#pragma unroll
    for (int ax = 0; ax < 4; ax++) {
      if (w[ax] == 0.0) continue;
      const T wt = (T)w[ax];
      if (m[ax] <= n[ax] - 2) {  // the difference x[m + e] - x[m] belongs to m
        const T d = un[ax] - xc;
        v += w[ax] * ((double)d * (double)d);
        g -= wt * d;
      }
      if (m[ax] >= 1) g += wt * (xc - up[ax]);
    }
    if (grad) grad[k] = vw ? (T)2 * (vw[k] * g) : (T)2 * g;
  }
  lds[threadIdx.x] = v;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) lds[threadIdx.x] += lds[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) scal[kScalPartials + blockIdx.x] = lds[0];
}

// Per-block partial sums of a*b (double), written to partials[blockIdx.x].
template <typename T>
__device__ void block_dot_partial(int64_t n, const T *a, const T *b, double *partials) {
  __shared__ double lds[256];
  double s = 0.0;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n;
       k += (int64_t)gridDim.x * blockDim.x)
    s += (double)a[k] * (double)b[k];
  lds[threadIdx.x] = s;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) lds[threadIdx.x] += lds[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) partials[blockIdx.x] = lds[0];
}

template <typename T>
__global__ __launch_bounds__(256) void k_dot_partial(int64_t n, const T *a, const T *b,
                                                     double *scal) {
  block_dot_partial(n, a, b, scal + kScalPartials);
}

// One block: scal[slot] = sum of nblocks partials, in index order.
__global__ __launch_bounds__(256) void k_dot_final(double *scal, int nblocks, int slot) {
  __shared__ double lds[256];
  double s = 0.0;
  for (int k = threadIdx.x; k < nblocks; k += 256) s += scal[kScalPartials + k];
  lds[threadIdx.x] = s;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) lds[threadIdx.x] += lds[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) scal[slot] = lds[0];
}

// alpha = scal[0] / scal[1]; x += alpha p; r -= alpha q; partial r.r
template <typename T>
__global__ __launch_bounds__(256) void k_cg_update(int64_t n, T *x, T *r, const T *p, const T *q,
                                                   double *scal) {
  const double den = scal[1];
  const double alpha_d = den != 0.0 ? scal[0] / den : 0.0;
  const T alpha = (T)alpha_d;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n;
       k += (int64_t)gridDim.x * blockDim.x) {
    x[k] = x[k] + alpha * p[k];
    r[k] = r[k] - alpha * q[k];
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) scal[3] = alpha_d;
  __syncthreads();
  block_dot_partial(n, (const T *)r, (const T *)r, scal + kScalPartials);
}

// beta = scal[2] / scal[0]; p = r + beta p
template <typename T>
__global__ __launch_bounds__(256) void k_cg_direction(int64_t n, const T *r, T *p, double *scal) {
  const double den = scal[0];
  const double beta_d = den != 0.0 ? scal[2] / den : 0.0;
  const T beta = (T)beta_d;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n;
       k += (int64_t)gridDim.x * blockDim.x)
    p[k] = r[k] + beta * p[k];
}

// scal[4] = beta; scal[0] = scal[2] (after k_cg_direction has read both)
__global__ void k_cg_shift(double *scal) {
  if (threadIdx.x == 0) {
    scal[4] = scal[0] != 0.0 ? scal[2] / scal[0] : 0.0;
    scal[0] = scal[2];
  }
}

// ---- segmented (multi-request) CG kernels --------------------------------
// One launch serves every request of a batch: block b belongs to the segment
// whose block0 <= b < block0 + blocks, found by binary search (uniform per
// block).  Each block covers kSegElems consecutive elements of its segment.
constexpr int kSegElems = 1024;

__device__ __forceinline__ int find_segment(const tg_cg_seg *segs, int nseg, int64_t b) {
  int lo = 0, hi = nseg - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (segs[mid].block0 <= b) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

__device__ __forceinline__ GridDiv seg_grid(const tg_cg_seg &g) {
  const int64_t n[4] = {g.shape[0], g.shape[1], g.shape[2], g.shape[3]};
  return grid_div(n);
}

template <typename T>
__device__ __forceinline__ T reg_value(const tg_cg_seg &g, const GridDiv &gd, gptr<const T> x,
                                       int64_t k) {
  return reg_at<T>(gd, g.axes, g.identity, x, (gptr<const T>)g.weight, k);
}

template <typename T, int OP>
__global__ __launch_bounds__(256) void k_cg_seg(const tg_cg_seg *__restrict__ segs, int nseg,
                                                double *partials) {
  __shared__ double lds[256];
  const int sidx = find_segment(segs, nseg, blockIdx.x);
  const tg_cg_seg &g = segs[sidx];
  const int64_t base = ((int64_t)blockIdx.x - g.block0) * kSegElems;
  // a solve's four vectors are distinct arrays: the loads of the unrolled
  // elements below may be issued before the stores of earlier ones
  gptr<T> __restrict__ x = (gptr<T>)g.x;
  gptr<T> __restrict__ r = (gptr<T>)g.r;
  gptr<T> __restrict__ p = (gptr<T>)g.p;
  gptr<T> __restrict__ q = (gptr<T>)g.q;
  gptr<const uint8_t> voxmask = (gptr<const uint8_t>)g.voxmask;
  GridDiv gd;
  if (OP == TG_CG_OP_REGULARIZE || OP == TG_CG_OP_REGULARIZE_DOT) gd = seg_grid(g);
  double acc = 0.0;
  T coef = 0;
  if (OP == TG_CG_OP_UPDATE) {
    const double den = g.scal[1];
    coef = (T)(den != 0.0 ? g.scal[0] / den : 0.0);
  } else if (OP == TG_CG_OP_DIRECTION) {
    const double den = g.scal[0];
    coef = (T)(den != 0.0 ? g.scal[2] / den : 0.0);
  }
#pragma unroll
  for (int it = 0; it < kSegElems / 256; it++) {
    const int64_t k = base + it * 256 + threadIdx.x;
    if (k >= g.n) continue;
    if (OP == TG_CG_OP_REGULARIZE) {
      if (!(voxmask && voxmask[k])) q[k] = q[k] + (T)g.lam * reg_value<T>(g, gd, p, k);
    } else if (OP == TG_CG_OP_REGULARIZE_DOT) {
      // REGULARIZE then DOT_PQ in one pass, with the same arithmetic
      T qk = q[k];
      if (!(voxmask && voxmask[k])) {
        qk = qk + (T)g.lam * reg_value<T>(g, gd, p, k);
        q[k] = qk;
      }
      acc += (double)p[k] * (double)qk;
    } else if (OP == TG_CG_OP_DOT_PQ) {
      acc += (double)p[k] * (double)q[k];
    } else if (OP == TG_CG_OP_UPDATE) {
      x[k] = x[k] + coef * p[k];
      const T rk = r[k] - coef * q[k];
      r[k] = rk;
      acc += (double)rk * (double)rk;
    } else if (OP == TG_CG_OP_DIRECTION) {
      p[k] = r[k] + coef * p[k];
    } else if (OP == TG_CG_OP_INIT) {
      acc += (double)r[k] * (double)r[k];
    } else if (OP == TG_CG_OP_ZERO_Q) {
      q[k] = (T)0;
    }
  }
  if (OP == TG_CG_OP_DOT_PQ || OP == TG_CG_OP_REGULARIZE_DOT || OP == TG_CG_OP_UPDATE ||
      OP == TG_CG_OP_INIT) {
    lds[threadIdx.x] = acc;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
      if ((int)threadIdx.x < off) lds[threadIdx.x] += lds[threadIdx.x + off];
      __syncthreads();
    }
    if (threadIdx.x == 0) partials[blockIdx.x] = lds[0];
  }
}

// One workgroup per segment: reduce its partials in a fixed order and update
// the segment's scalars.
constexpr int kFinalLanes = 1024;

template <int OP>
__global__ __launch_bounds__(kFinalLanes) void k_cg_seg_final(const tg_cg_seg *__restrict__ segs,
                                                              const double *partials) {
  __shared__ double lds[kFinalLanes];
  const tg_cg_seg &g = segs[blockIdx.x];
  const int64_t nb = (g.n + kSegElems - 1) / kSegElems;
  double s = 0.0;
  if (OP != TG_CG_OP_DIRECTION)
    for (int64_t k = threadIdx.x; k < nb; k += kFinalLanes) s += partials[g.block0 + k];
  lds[threadIdx.x] = s;
  __syncthreads();
  for (int off = kFinalLanes / 2; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) lds[threadIdx.x] += lds[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x != 0) return;
  double *sc = g.scal;
  if (OP == TG_CG_OP_DOT_PQ || OP == TG_CG_OP_REGULARIZE_DOT) {
    sc[1] = lds[0];
  } else if (OP == TG_CG_OP_UPDATE) {
    sc[3] = sc[1] != 0.0 ? sc[0] / sc[1] : 0.0;
    sc[2] = lds[0];
  } else if (OP == TG_CG_OP_DIRECTION) {
    sc[4] = sc[0] != 0.0 ? sc[2] / sc[0] : 0.0;
    sc[0] = sc[2];
  } else if (OP == TG_CG_OP_INIT) {
    sc[0] = lds[0];
    sc[5] = lds[0];
  }
}

template <typename T, int OP>
int launch_seg(const tg_cg_seg *segs, int nseg, int64_t nblocks, double *partials, hipStream_t s) {
  hipLaunchKernelGGL((k_cg_seg<T, OP>), dim3((uint32_t)nblocks), dim3(256), 0, s, segs, nseg,
                     partials);
  if (OP != TG_CG_OP_REGULARIZE && OP != TG_CG_OP_ZERO_Q)
    hipLaunchKernelGGL((k_cg_seg_final<OP>), dim3((uint32_t)nseg), dim3(kFinalLanes), 0, s, segs,
                       partials);
  return hip_check(hipGetLastError(), "segmented cg launch");
}

template <typename T>
int dispatch_seg(int op, const tg_cg_seg *segs, int nseg, int64_t nblocks, double *partials,
                 hipStream_t s) {
  switch (op) {
    case TG_CG_OP_REGULARIZE: return launch_seg<T, TG_CG_OP_REGULARIZE>(segs, nseg, nblocks, partials, s);
    case TG_CG_OP_DOT_PQ: return launch_seg<T, TG_CG_OP_DOT_PQ>(segs, nseg, nblocks, partials, s);
    case TG_CG_OP_REGULARIZE_DOT:
      return launch_seg<T, TG_CG_OP_REGULARIZE_DOT>(segs, nseg, nblocks, partials, s);
    case TG_CG_OP_UPDATE: return launch_seg<T, TG_CG_OP_UPDATE>(segs, nseg, nblocks, partials, s);
    case TG_CG_OP_DIRECTION: return launch_seg<T, TG_CG_OP_DIRECTION>(segs, nseg, nblocks, partials, s);
    case TG_CG_OP_INIT: return launch_seg<T, TG_CG_OP_INIT>(segs, nseg, nblocks, partials, s);
    case TG_CG_OP_ZERO_Q: return launch_seg<T, TG_CG_OP_ZERO_Q>(segs, nseg, nblocks, partials, s);
  }
  return set_error(TG_EINVAL, "bad segmented cg op %d", op);
}

inline uint32_t red_blocks(int64_t n) {
  int64_t b = (n + 255) / 256;
  if (b > kRedBlocks) b = kRedBlocks;
  if (b < 1) b = 1;
  return (uint32_t)b;
}

inline uint32_t ew_blocks(int64_t n) {
  int64_t b = (n + 255) / 256;
  if (b > 65536) b = 65536;
  if (b < 1) b = 1;
  return (uint32_t)b;
}

}  // namespace
}  // namespace tg

using namespace tg;

extern "C" {

int tg_regularize(int32_t dtype, const int32_t shape[4], uint32_t axes_mask, double lam,
                  int32_t identity, const void *x, void *out, const uint8_t *voxmask,
                  const void *weight, void *stream) {
  if (!shape || !x || !out) return set_error(TG_EINVAL, "null argument to tg_regularize");
  for (int k = 0; k < 4; k++)
    if (shape[k] <= 0) return set_error(TG_EINVAL, "shape[%d] = %d must be positive", k, shape[k]);
  hipStream_t s = (hipStream_t)stream;
  const int64_t n = (int64_t)shape[0] * shape[1] * shape[2] * shape[3];
  if (dtype == TG_F32)
    hipLaunchKernelGGL((k_regularize<float>), dim3(ew_blocks(n)), dim3(256), 0, s, shape[0],
                       shape[1], shape[2], shape[3], axes_mask, (float)lam, identity,
                       (const float *)x, (float *)out, voxmask, (const float *)weight);
  else if (dtype == TG_F64)
    hipLaunchKernelGGL((k_regularize<double>), dim3(ew_blocks(n)), dim3(256), 0, s, shape[0],
                       shape[1], shape[2], shape[3], axes_mask, lam, identity, (const double *)x,
                       (double *)out, voxmask, (const double *)weight);
  else
    return set_error(TG_EINVAL, "bad dtype %d", dtype);
  return hip_check(hipGetLastError(), "regularize launch");
}

int tg_smoothness(int32_t dtype, const int32_t shape[4], const double weights[4], const void *x,
                  const void *voxel_weight, void *grad, double *scal, int32_t slot,
                  void *stream) {
  if (!shape || !weights || !x || !scal || slot < 0 || slot >= 16)
    return set_error(TG_EINVAL, "bad tg_smoothness arguments");
  for (int k = 0; k < 4; k++)
    if (shape[k] <= 0) return set_error(TG_EINVAL, "shape[%d] = %d must be positive", k, shape[k]);
  hipStream_t s = (hipStream_t)stream;
  const int64_t n = (int64_t)shape[0] * shape[1] * shape[2] * shape[3];
  const uint32_t nb = red_blocks(n);
  if (dtype == TG_F32)
    hipLaunchKernelGGL((k_smooth<float>), dim3(nb), dim3(256), 0, s, shape[0], shape[1], shape[2],
                       shape[3], weights[0], weights[1], weights[2], weights[3], (const float *)x,
                       (float *)grad, scal, (const float *)voxel_weight);
  else if (dtype == TG_F64)
    hipLaunchKernelGGL((k_smooth<double>), dim3(nb), dim3(256), 0, s, shape[0], shape[1],
                       shape[2], shape[3], weights[0], weights[1], weights[2], weights[3],
                       (const double *)x, (double *)grad, scal, (const double *)voxel_weight);
  else
    return set_error(TG_EINVAL, "bad dtype %d", dtype);
  hipLaunchKernelGGL(k_dot_final, dim3(1), dim3(256), 0, s, scal, (int)nb, slot);
  return hip_check(hipGetLastError(), "smoothness launch");
}

int tg_cg_dot(int32_t dtype, int64_t n, const void *a, const void *b, double *scal, int32_t slot,
              void *stream) {
  if (!a || !b || !scal || slot < 0 || slot >= 16) return set_error(TG_EINVAL, "bad tg_cg_dot args");
  hipStream_t s = (hipStream_t)stream;
  const uint32_t nb = red_blocks(n);
  if (dtype == TG_F32)
    hipLaunchKernelGGL((k_dot_partial<float>), dim3(nb), dim3(256), 0, s, n, (const float *)a,
                       (const float *)b, scal);
  else if (dtype == TG_F64)
    hipLaunchKernelGGL((k_dot_partial<double>), dim3(nb), dim3(256), 0, s, n, (const double *)a,
                       (const double *)b, scal);
  else
    return set_error(TG_EINVAL, "bad dtype %d", dtype);
  hipLaunchKernelGGL(k_dot_final, dim3(1), dim3(256), 0, s, scal, (int)nb, slot);
  return hip_check(hipGetLastError(), "dot launch");
}

int tg_cg_update(int32_t dtype, int64_t n, void *x, void *r, const void *p, const void *q,
                 double *scal, void *stream) {
  if (!x || !r || !p || !q || !scal) return set_error(TG_EINVAL, "bad tg_cg_update args");
  hipStream_t s = (hipStream_t)stream;
  const uint32_t nb = red_blocks(n);
  if (dtype == TG_F32)
    hipLaunchKernelGGL((k_cg_update<float>), dim3(nb), dim3(256), 0, s, n, (float *)x, (float *)r,
                       (const float *)p, (const float *)q, scal);
  else if (dtype == TG_F64)
    hipLaunchKernelGGL((k_cg_update<double>), dim3(nb), dim3(256), 0, s, n, (double *)x,
                       (double *)r, (const double *)p, (const double *)q, scal);
  else
    return set_error(TG_EINVAL, "bad dtype %d", dtype);
  hipLaunchKernelGGL(k_dot_final, dim3(1), dim3(256), 0, s, scal, (int)nb, 2);
  return hip_check(hipGetLastError(), "cg update launch");
}

int tg_cg_direction(int32_t dtype, int64_t n, const void *r, void *p, double *scal, void *stream) {
  if (!r || !p || !scal) return set_error(TG_EINVAL, "bad tg_cg_direction args");
  hipStream_t s = (hipStream_t)stream;
  if (dtype == TG_F32)
    hipLaunchKernelGGL((k_cg_direction<float>), dim3(ew_blocks(n)), dim3(256), 0, s, n,
                       (const float *)r, (float *)p, scal);
  else if (dtype == TG_F64)
    hipLaunchKernelGGL((k_cg_direction<double>), dim3(ew_blocks(n)), dim3(256), 0, s, n,
                       (const double *)r, (double *)p, scal);
  else
    return set_error(TG_EINVAL, "bad dtype %d", dtype);
  hipLaunchKernelGGL(k_cg_shift, dim3(1), dim3(64), 0, s, scal);
  return hip_check(hipGetLastError(), "cg direction launch");
}

int64_t tg_cg_seg_blocks(int64_t n) { return (n + kSegElems - 1) / kSegElems; }

int tg_cg_batch(int32_t dtype, int32_t op, const tg_cg_seg *segs, int32_t nseg, int64_t nblocks,
                double *partials, void *stream) {
  if (!segs || nseg <= 0 || nblocks <= 0 || !partials)
    return set_error(TG_EINVAL, "bad tg_cg_batch args");
  hipStream_t s = (hipStream_t)stream;
  if (dtype == TG_F32) return dispatch_seg<float>(op, segs, nseg, nblocks, partials, s);
  if (dtype == TG_F64) return dispatch_seg<double>(op, segs, nseg, nblocks, partials, s);
  return set_error(TG_EINVAL, "bad dtype %d", dtype);
}

}  // extern "C"
