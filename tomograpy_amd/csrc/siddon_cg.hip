// Device-resident conjugate gradient pieces for the normal equations
// (P^T P + lam (I + sum_k D_k^T D_k)) x = P^T y that lo.acg solves in the
// reference (srt.py:225, models.py:85-94; lo is not vendored).
//
// D_k is the forward difference along axis k with the numpy.diff shape
// convention (output shorter by one), so D_k^T D_k x at index m along k is
//   [m >= 1] (x_m - x_{m-1}) - [m <= n-2] (x_{m+1} - x_m).
// Dot products are reduced in a fixed order (per-block partials, then one
// block), so a CG run is reproducible given reproducible operator outputs.
#include <hip/hip_runtime.h>

#include "siddon_common.h"
#include "siddon_host.h"

namespace tg {
namespace {

constexpr int kRedBlocks = 1024;  // max partials, lives at scal[16 .. 16 + kRedBlocks)
constexpr int kScalPartials = 16;

template <typename T>
__global__ __launch_bounds__(256) void k_regularize(int64_t n0, int64_t n1, int64_t n2,
                                                    int64_t n3, uint32_t axes, T lam,
                                                    int identity, const T *__restrict__ x,
                                                    T *__restrict__ out,
                                                    const uint8_t *__restrict__ mask) {
  const int64_t total = n0 * n1 * n2 * n3;
  const int64_t s3 = 1, s2 = n3, s1 = n2 * n3, s0 = n1 * n2 * n3;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < total;
       k += (int64_t)gridDim.x * blockDim.x) {
    if (mask && mask[k]) continue;
    const int64_t m3 = k % n3, m2 = (k / s2) % n2, m1 = (k / s1) % n1, m0 = k / s0;
    const T xc = x[k];
    T acc = identity ? xc : (T)0;
    const int64_t m[4] = {m0, m1, m2, m3};
    const int64_t n[4] = {n0, n1, n2, n3};
    const int64_t st[4] = {s0, s1, s2, s3};
#pragma unroll
    for (int ax = 0; ax < 4; ax++) {
      if (!(axes & (1u << ax))) continue;
      if (m[ax] >= 1) acc += xc - x[k - st[ax]];
      if (m[ax] <= n[ax] - 2) acc -= x[k + st[ax]] - xc;
    }
    out[k] = out[k] + lam * acc;
  }
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
                  void *stream) {
  if (!shape || !x || !out) return set_error(TG_EINVAL, "null argument to tg_regularize");
  for (int k = 0; k < 4; k++)
    if (shape[k] <= 0) return set_error(TG_EINVAL, "shape[%d] = %d must be positive", k, shape[k]);
  hipStream_t s = (hipStream_t)stream;
  const int64_t n = (int64_t)shape[0] * shape[1] * shape[2] * shape[3];
  if (dtype == TG_F32)
    hipLaunchKernelGGL((k_regularize<float>), dim3(ew_blocks(n)), dim3(256), 0, s, shape[0],
                       shape[1], shape[2], shape[3], axes_mask, (float)lam, identity,
                       (const float *)x, (float *)out, voxmask);
  else if (dtype == TG_F64)
    hipLaunchKernelGGL((k_regularize<double>), dim3(ew_blocks(n)), dim3(256), 0, s, shape[0],
                       shape[1], shape[2], shape[3], axes_mask, lam, identity, (const double *)x,
                       (double *)out, voxmask);
  else
    return set_error(TG_EINVAL, "bad dtype %d", dtype);
  return hip_check(hipGetLastError(), "regularize launch");
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

}  // extern "C"
