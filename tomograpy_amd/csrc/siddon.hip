// gfx950 Siddon conic projector / backprojector: kernels and the C ABI.
//
// One ray per lane.  A workgroup of 256 lanes covers a tile of BI x BJ pixels
// of one view (BJ = 1 << bj_log2 pixels along the image's second axis, which
// maps to the cube's fastest axis for near-equatorial observers), so the
// voxel reads of the 64 lanes of a wavefront fall on a few cache lines and the
// workgroups of one XCD sweep neighbouring rays of the same views.  The view
// header sits in scalar registers (its address is uniform per workgroup).
//
// Forward (pj, pjt): each lane owns its pixel, accumulates in a register in
// the reference's along-ray order starting from the pixel's current value and
// stores once, so the result is bit-identical to the reference's (C:915-920,
// C:931-936), whose per-pixel sums are also sequential.
// Backward (bpj, bpjt): an atomic scatter (C:923-928, C:939-944), or the
// fixed-order deterministic path in siddon_det.hip.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <vector>

#include "siddon_common.h"
#include "siddon_host.h"
#include "siddon_rays.h"

namespace tg {

// ---------------------------------------------------------------------------
// Ray-driven kernels
// ---------------------------------------------------------------------------

template <typename T, bool SUN, bool VOX, bool FAST>
__global__ __launch_bounds__(256) void k_project(KArgs<T> a) {
  project_tile<T, SUN, VOX, FAST>(a, (a.ray_i >= 0) ? 0u : xcd_remap(blockIdx.x, gridDim.x));
}

template <typename T, bool SUN, bool VOX, bool FAST>
__global__ __launch_bounds__(256) void k_backproject(KArgs<T> a) {
  backproject_tile<T, SUN, VOX, FAST>(a, (a.ray_i >= 0) ? 0u : xcd_remap(blockIdx.x, gridDim.x));
}

// Step count per ray (value independent).
template <typename T, bool SUN>
__global__ __launch_bounds__(256) void k_count(KArgs<T> a) {
  const uint32_t tile = xcd_remap(blockIdx.x, gridDim.x);
  int i, j, t;
  uint32_t c = 0;
  bool valid = lane_pixel(a, tile, i, j, t);
  if (valid && !(a.pixmask && a.pixmask[i * a.ms[0] + j * a.ms[1] + (int64_t)t * a.ms[2]] != 0)) {
    const ViewHdr<T> &h = a.views[t];
    T u0[3];
    pixel_unit_vector(a, h, t, i, j, u0);
    Ray<T> r;
    if (ray_setup(a, h.M, u0, r)) {
      if (!ray_state_ok(r)) {
        flag_status(a.status, TG_EGEOMETRY);
      } else {
        int guard = (int)(a.n[0] + a.n[1] + a.n[2]) + 2;
        while (in_map(a, r) && (!SUN || outside_sun(h.M, u0, r.ac))) {
          if (--guard < 0) {
            flag_status(a.status, TG_EGEOMETRY);
            break;
          }
          bool sx, sy, sz;
          step(r, sx, sy, sz);
          c++;
        }
      }
    }
  }
  if (valid && a.per_ray)
    a.per_ray[((size_t)(t - a.t0) * a.n1 + i) * a.n2 + j] = c;
  if (a.total) {
    // wavefront reduction, then one atomic per wavefront
    unsigned long long s = c;
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    if ((threadIdx.x & 63) == 0 && s) atomicAdd(a.total, s);
  }
}

// Unit vectors u[i][j][t][3] (C:618-648); skips pixels whose data is NaN.
template <typename T>
__global__ __launch_bounds__(256) void k_unit_vector(KArgs<T> a, T *u) {
  const uint32_t tile = blockIdx.x;
  int i, j, t;
  if (!lane_pixel(a, tile, i, j, t)) return;
  if (a.data) {
    const T x = a.data[i * a.ds[0] + j * a.ds[1] + (int64_t)t * a.ds[2]];
    if (x != x) return;
  }
  T u0[3];
  pixel_unit_vector(a, a.views[t], t, i, j, u0);
  const size_t o = (((size_t)i * a.n2 + j) * a.nv + t) * 3;
  u[o] = u0[0];
  u[o + 1] = u0[1];
  u[o + 2] = u0[2];
}

// a1/an per pixel from u (C:694-720).
template <typename T>
__global__ __launch_bounds__(256) void k_intersections(KArgs<T> a, const T *u, T *a1, T *an) {
  const uint32_t tile = blockIdx.x;
  int i, j, t;
  if (!lane_pixel(a, tile, i, j, t)) return;
  const size_t o = (((size_t)i * a.n2 + j) * a.nv + t) * 3;
  T u0[3] = {u[o], u[o + 1], u[o + 2]};
  T b1[3], bn[3];
  intersection_parameters(a, a.views[t].M, u0, b1, bn);
  for (int k = 0; k < 3; k++) {
    a1[o + k] = b1[k];
    an[o + k] = bn[k];
  }
}

// Elementwise cube *= weight (the voxel half of the Thomson adjoint).
template <typename T>
__global__ __launch_bounds__(256) void k_scale_cube(KArgs<T> a) {
  const int64_t nvox = (int64_t)a.n[0] * a.n[1] * a.n[2];
  const int64_t total = nvox * a.nt;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < total;
       k += (int64_t)gridDim.x * blockDim.x) {
    const int64_t b = k / nvox, v = k % nvox;
    const int64_t i0 = v / ((int64_t)a.n[1] * a.n[2]);
    const int64_t i1 = (v / a.n[2]) % a.n[1];
    const int64_t i2 = v % a.n[2];
    T *x = a.cube + i0 * a.cs[0] + i1 * a.cs[1] + i2 * a.cs[2] + b * a.cs[3];
    const T w = a.voxw[i0 * a.vs[0] + i1 * a.vs[1] + i2 * a.vs[2]];
    *x = w * *x;
  }
}

// cube[v] = vox_weight[v] * src[v], or 0 where voxmask[v] != 0, over every
// time bin: the voxel operand of a forward projection applied once per call
// instead of once per Siddon step.
template <typename T>
__global__ __launch_bounds__(256) void k_vox_apply(KArgs<T> a, const T *__restrict__ src,
                                                   int64_t s0, int64_t s1, int64_t s2, int64_t s3) {
  const int64_t nvox = (int64_t)a.n[0] * a.n[1] * a.n[2];
  const int64_t total = nvox * a.nt;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < total;
       k += (int64_t)gridDim.x * blockDim.x) {
    const int64_t b = k / nvox, v = k % nvox;
    const int64_t i0 = v / ((int64_t)a.n[1] * a.n[2]);
    const int64_t i1 = (v / a.n[2]) % a.n[1];
    const int64_t i2 = v % a.n[2];
    const int64_t m = i0 * a.vs[0] + i1 * a.vs[1] + i2 * a.vs[2];
    T x = src[i0 * s0 + i1 * s1 + i2 * s2 + b * s3];
    if (a.voxw) x = a.voxw[m] * x;
    if (a.voxmask && a.voxmask[m] != 0) x = (T)0;
    a.cube[i0 * a.cs[0] + i1 * a.cs[1] + i2 * a.cs[2] + b * a.cs[3]] = x;
  }
}

// k_vox_apply for a source with the time bins fastest (the (N0, N1, N2, T)
// C-order array) into a bin-major cube: a workgroup reads 256 consecutive
// voxels x nt bins in one coalesced sweep, transposes them through LDS, and
// writes each bin's 256 voxels as one coalesced run.  Voxel operands and the
// destination are dense (strides n1 n2, n2, 1 and nvox per bin).
constexpr int kVoxTile = 256;
constexpr int kVoxTileBins = 16;

template <typename T>
__global__ __launch_bounds__(kVoxTile) void k_vox_apply_bins(KArgs<T> a, const T *__restrict__ src) {
  __shared__ T tile[kVoxTileBins][kVoxTile + 1];
  const int nt = a.nt;
  const int64_t nvox = (int64_t)a.n[0] * a.n[1] * a.n[2];
  for (int64_t v0 = (int64_t)blockIdx.x * kVoxTile; v0 < nvox; v0 += (int64_t)gridDim.x * kVoxTile) {
    const int nv = (int)(nvox - v0 < kVoxTile ? nvox - v0 : kVoxTile);
    for (int i = threadIdx.x; i < nv * nt; i += kVoxTile) tile[i % nt][i / nt] = src[v0 * nt + i];
    __syncthreads();
    if ((int)threadIdx.x < nv) {
      const int64_t v = v0 + threadIdx.x;
      const T w = a.voxw ? a.voxw[v] : (T)1;
      const bool masked = a.voxmask && a.voxmask[v] != 0;
      for (int b = 0; b < nt; b++) {
        T x = tile[b][threadIdx.x];
        if (a.voxw) x = w * x;
        a.cube[(int64_t)b * nvox + v] = masked ? (T)0 : x;
      }
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------

static thread_local char g_err[1024] = "";

int set_error(int code, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return code;
}

int hip_check(hipError_t e, const char *what) {
  if (e == hipSuccess) return TG_OK;
  return set_error(TG_EHIP, "%s: %s", what, hipGetErrorString(e));
}

// Validates *c and fills the kernel arguments for views [t0, t1).
template <typename T>
int make_args(const tg_call *c, int t0, int t1, KArgs<T> &a, int tile_j = 0) {
  memset(&a, 0, sizeof(a));
  if (!c) return set_error(TG_EINVAL, "null call");
  if (c->n1 < 0 || c->n2 < 0 || c->nv < 0)
    return set_error(TG_EINVAL, "negative image stack shape (%d, %d, %d)", c->n1, c->n2, c->nv);
  for (int k = 0; k < 3; k++)
    if (c->map.n[k] <= 0)
      return set_error(TG_EINVAL, "map shape must be positive, got n[%d] = %d", k, c->map.n[k]);
  if (c->kind < TG_PJ || c->kind > TG_BPJT) return set_error(TG_EINVAL, "bad kind %d", c->kind);
  if (t0 < 0 || t1 > c->nv || t0 > t1)
    return set_error(TG_EINVAL, "view range [%d, %d) outside [0, %d)", t0, t1, c->nv);
  const bool time = (c->kind == TG_PJT || c->kind == TG_BPJT);
  const int nt = c->map.nt > 0 ? c->map.nt : 1;
  if (time && c->map.nt <= 0) return set_error(TG_EINVAL, "time kinds need map.nt >= 1");
  for (int k = 0; k < 3; k++) {
    a.mp[k] = (T)c->map.cdelt[k];
    a.md[k] = (T)c->map.pshape[k];
    a.mmin[k] = (T)c->map.mmin[k];
    a.mmax[k] = (T)c->map.mmax[k];
    a.n[k] = (uint32_t)c->map.n[k];
  }
  a.nt = time ? nt : 1;
  for (int k = 0; k < 4; k++) a.cs[k] = c->map.stride[k];
  a.n1 = c->n1;
  a.n2 = c->n2;
  a.nv = c->nv;
  a.t0 = t0;
  a.t1 = t1;
  a.ray_i = -1;
  a.ray_j = -1;
  // Default tile: one workgroup spans up to 256 pixels of an image column
  // (the image's second axis, rounded up to a power of two), so that its rays
  // share one direction across the first axis and walk the same voxel columns.
  int tj_default = 1;
  while (tj_default < c->n2 && tj_default < 256) tj_default <<= 1;
  int tj = tile_j > 0 ? tile_j : (c->tile_j > 0 ? c->tile_j : tj_default);
  if (tj > 256 || (tj & (tj - 1))) return set_error(TG_EINVAL, "tile_j must be a power of two <= 256");
  int lg = 0;
  while ((1 << lg) < tj) lg++;
  a.bj_log2 = lg;
  const int bi = 256 >> lg;
  a.tiles_i = (c->n1 + bi - 1) / bi;
  a.tiles_j = (c->n2 + tj - 1) / tj;
  a.kind = c->kind;
  a.flags = c->flags;
  for (int k = 0; k < 3; k++) {
    a.ds[k] = c->data_stride[k];
    a.ms[k] = c->mask_stride[k];
    a.vs[k] = c->vox_stride[k];
  }
  if (!c->views && c->nv > 0) return set_error(TG_EINVAL, "views descriptor pointer is null");
  a.views = (const ViewHdr<T> *)c->views;
  a.trig = (const double *)((const char *)c->views + views_hdr_bytes<T>(c->nv));
  a.cube = (T *)c->cube;
  a.data = (T *)c->data;
  a.pixmask = c->pixmask;
  a.tindex = c->tindex;
  a.pixw = (const T *)c->pix_weight;
  a.voxw = (const T *)c->vox_weight;
  a.voxmask = c->voxmask;
  a.status = c->status;
  // step counts are indexed in 32 bits; larger stacks walk each ray to its end
  a.ray_steps = ((c->kind == TG_PJ || c->kind == TG_PJT) &&
                 (uint64_t)c->nv * (uint64_t)c->n1 * (uint64_t)c->n2 <= 0xffffffffull)
                    ? c->ray_steps
                    : nullptr;
  a.cube_bytes = buffer_walk_bytes(a.n, a.cs, (int)sizeof(T));
  {
    const int64_t dn[3] = {c->n1, c->n2, c->nv};
    int64_t ext = 0;
    bool ok = true;
    for (int k = 0; k < 3; k++) {
      if (a.ds[k] < 0 || a.ds[k] >= (1ll << 31)) ok = false;
      ext += (dn[k] > 0 ? dn[k] - 1 : 0) * a.ds[k];
    }
    a.data32 = (ok && ext < (1ll << 31)) ? 1u : 0u;
  }
  a.acc = nullptr;
  if ((c->flags & TG_ACC_F64) && (c->kind == TG_BPJ || c->kind == TG_BPJT) &&
      !(c->flags & TG_DETERMINISTIC)) {
    if (!c->acc) return set_error(TG_EINVAL, "TG_ACC_F64 needs the acc buffer");
    const int64_t n0 = a.n[0], n1 = a.n[1], n2 = a.n[2];
    if (a.cs[2] != 1 || a.cs[1] != n2 || a.cs[0] != n1 * n2 ||
        (a.nt > 1 && a.cs[3] != n0 * n1 * n2))
      return set_error(TG_EINVAL, "TG_ACC_F64 needs a dense cube (time bin, axis 0, 1, 2 order)");
    a.acc = c->acc;
  }
  return TG_OK;
}

template int make_args<float>(const tg_call *, int, int, KArgs<float> &, int);
template int make_args<double>(const tg_call *, int, int, KArgs<double> &, int);


template <typename T, bool SUN, bool VOX, bool FAST>
void launch_siddon_f(const KArgs<T> &a, uint32_t blocks, hipStream_t s, bool bwd) {
  if (bwd)
    hipLaunchKernelGGL((k_backproject<T, SUN, VOX, FAST>), dim3(blocks), dim3(256), 0, s, a);
  else
    hipLaunchKernelGGL((k_project<T, SUN, VOX, FAST>), dim3(blocks), dim3(256), 0, s, a);
}

template <typename T, bool SUN, bool VOX>
void launch_siddon(const KArgs<T> &a, uint32_t blocks, hipStream_t s, bool bwd) {
  const bool fast = walk_fast_ok(a.n, a.cs) && (!VOX || walk_fast_ok(a.n, a.vs));
  if (fast) launch_siddon_f<T, SUN, VOX, true>(a, blocks, s, bwd);
  else launch_siddon_f<T, SUN, VOX, false>(a, blocks, s, bwd);
}

// TG_ACC_F64: cube += (T)acc, acc = 0, over the dense cube of every time bin.
template <typename T>
__global__ __launch_bounds__(256) void k_acc_finish(T *__restrict__ cube, double *__restrict__ acc,
                                                    int64_t total) {
  for (int64_t k = blockIdx.x * 256ll + threadIdx.x; k < total; k += (int64_t)gridDim.x * 256) {
    const double v = acc[k];
    if (v != 0.0) {
      cube[k] += (T)v;
      acc[k] = 0.0;
    }
  }
}

template <typename T>
int finish_acc(const KArgs<T> &a, hipStream_t s) {
  if (!a.acc) return TG_OK;
  const int64_t total = (int64_t)a.n[0] * a.n[1] * a.n[2] * a.nt;
  const int64_t want = (total + 255) / 256;
  const uint32_t blocks = (uint32_t)(want < 8192 ? want : 8192);
  if (blocks == 0) return TG_OK;
  hipLaunchKernelGGL((k_acc_finish<T>), dim3(blocks), dim3(256), 0, s, a.cube, a.acc, total);
  return hip_check(hipGetLastError(), "accumulator finish launch");
}

template int finish_acc<float>(const KArgs<float> &, hipStream_t);
template int finish_acc<double>(const KArgs<double> &, hipStream_t);

// k_acc_finish over the jobs of a device table: blockIdx.y is the job.
template <typename T>
__global__ __launch_bounds__(256) void k_acc_finish_table(const char *__restrict__ table,
                                                          size_t stride) {
  const KArgs<T> &a = *reinterpret_cast<const KArgs<T> *>(table + (size_t)blockIdx.y * stride);
  if (!a.acc) return;
  const int64_t total = (int64_t)a.n[0] * a.n[1] * a.n[2] * a.nt;
  for (int64_t k = blockIdx.x * 256ll + threadIdx.x; k < total; k += (int64_t)gridDim.x * 256) {
    const double v = a.acc[k];
    if (v != 0.0) {
      a.cube[k] += (T)v;
      a.acc[k] = 0.0;
    }
  }
}

template <typename T>
int finish_acc_table(const void *dev_table, size_t stride, const KArgs<T> *const *host_args,
                     int njobs, hipStream_t s) {
  int64_t most = 0;
  for (int k = 0; k < njobs; k++) {
    const KArgs<T> &a = *host_args[k];
    if (!a.acc) continue;
    const int64_t total = (int64_t)a.n[0] * a.n[1] * a.n[2] * a.nt;
    most = total > most ? total : most;
  }
  if (most == 0) return TG_OK;
  if (njobs > 65535) return set_error(TG_EINVAL, "accumulator finish of %d jobs", njobs);
  const int64_t want = (most + 255) / 256;
  const uint32_t bx = (uint32_t)(want < 1024 ? want : 1024);
  hipLaunchKernelGGL((k_acc_finish_table<T>), dim3(bx, (uint32_t)njobs), dim3(256), 0, s,
                     (const char *)dev_table, stride);
  return hip_check(hipGetLastError(), "accumulator finish launch");
}

template int finish_acc_table<float>(const void *, size_t, const KArgs<float> *const *, int,
                                     hipStream_t);
template int finish_acc_table<double>(const void *, size_t, const KArgs<double> *const *, int,
                                      hipStream_t);

template <typename T>
int run_siddon(const tg_call *c, int t0, int t1, int ray_i, int ray_j, hipStream_t s) {
  KArgs<T> a;
  int rc = make_args<T>(c, t0, t1, a);
  if (rc) return rc;
  if (!c->cube || !c->data) return set_error(TG_EINVAL, "cube and data pointers are required");
  const bool bwd = (c->kind == TG_BPJ || c->kind == TG_BPJT);
  if (bwd && (c->flags & TG_DETERMINISTIC)) {
    if (ray_i >= 0) return set_error(TG_EINVAL, "deterministic mode applies to full/image calls");
    if (t0 == 0 && t1 == c->nv) {
      // all views: through a canonical plan when it fits the workspace
      const int r = run_det_plan<T>(c, s);
      if (r != kDetPlanUnfit) return r;
    }
    return run_deterministic<T>(c, a, s);
  }
  const bool sun = (c->flags & TG_OBSTACLE_SUN) != 0;
  const bool vox = (c->voxmask != nullptr) || (c->vox_weight != nullptr && !bwd);
  auto launch = [&](const KArgs<T> &av, uint32_t blocks) {
    if (sun) {
      if (vox) launch_siddon<T, true, true>(av, blocks, s, bwd);
      else launch_siddon<T, true, false>(av, blocks, s, bwd);
    } else {
      if (vox) launch_siddon<T, false, true>(av, blocks, s, bwd);
      else launch_siddon<T, false, false>(av, blocks, s, bwd);
    }
  };
  if (ray_i >= 0) {
    if (ray_i >= c->n1 || ray_j >= c->n2 || ray_i < 0 || ray_j < 0)
      return set_error(TG_EINVAL, "pixel (%d, %d) outside the %d x %d image", ray_i, ray_j, c->n1,
                       c->n2);
    a.ray_i = ray_i;
    a.ray_j = ray_j;
    if (a.t1 <= a.t0) return TG_OK;
    launch(a, 1u);
  } else if ((rc = for_view_launches(a, launch))) {
    return rc;
  }
  if ((rc = hip_check(hipGetLastError(), "siddon kernel launch"))) return rc;
  return finish_acc<T>(a, s);
}

int dispatch_siddon(const tg_call *c, int t0, int t1, int ray_i, int ray_j, void *stream) {
  if (!c) return set_error(TG_EINVAL, "null call");
  hipStream_t s = (hipStream_t)stream;
  if (c->dtype == TG_F32) return run_siddon<float>(c, t0, t1, ray_i, ray_j, s);
  if (c->dtype == TG_F64) return run_siddon<double>(c, t0, t1, ray_i, ray_j, s);
  return set_error(TG_EINVAL, "bad dtype %d", c->dtype);
}

template <typename T>
int run_count(const tg_call *c, uint32_t *per_ray, unsigned long long *total, hipStream_t s) {
  KArgs<T> a;
  int rc = make_args<T>(c, 0, c->nv, a);
  if (rc) return rc;
  a.per_ray = per_ray;
  a.total = total;
  const bool sun = (c->flags & TG_OBSTACLE_SUN) != 0;
  rc = for_view_launches(a, [&](const KArgs<T> &av, uint32_t blocks) {
    KArgs<T> ac = av;
    // k_count writes per_ray relative to its launch's first view
    if (per_ray) ac.per_ray = per_ray + (size_t)(av.t0 - a.t0) * (size_t)a.n1 * (size_t)a.n2;
    if (sun) hipLaunchKernelGGL((k_count<T, true>), dim3(blocks), dim3(256), 0, s, ac);
    else hipLaunchKernelGGL((k_count<T, false>), dim3(blocks), dim3(256), 0, s, ac);
  });
  if (rc) return rc;
  return hip_check(hipGetLastError(), "count kernel launch");
}

template <typename T>
int run_unit(const tg_call *c, int t0, int t1, void *u, hipStream_t s) {
  KArgs<T> a;
  int rc = make_args<T>(c, t0, t1, a);
  if (rc) return rc;
  if (!u) return set_error(TG_EINVAL, "null u");
  rc = for_view_launches(a, [&](const KArgs<T> &av, uint32_t blocks) {
    hipLaunchKernelGGL((k_unit_vector<T>), dim3(blocks), dim3(256), 0, s, av, (T *)u);
  });
  if (rc) return rc;
  return hip_check(hipGetLastError(), "unit vector kernel launch");
}

template <typename T>
int run_inter(const tg_call *c, int t0, int t1, const void *u, void *a1, void *an, hipStream_t s) {
  KArgs<T> a;
  int rc = make_args<T>(c, t0, t1, a);
  if (rc) return rc;
  if (!u || !a1 || !an) return set_error(TG_EINVAL, "null u/a1/an");
  rc = for_view_launches(a, [&](const KArgs<T> &av, uint32_t blocks) {
    hipLaunchKernelGGL((k_intersections<T>), dim3(blocks), dim3(256), 0, s, av, (const T *)u,
                       (T *)a1, (T *)an);
  });
  if (rc) return rc;
  return hip_check(hipGetLastError(), "intersection kernel launch");
}

template <typename T>
int run_scale(const tg_call *c, hipStream_t s) {
  KArgs<T> a;
  tg_call cc = *c;
  if (cc.kind != TG_PJT && cc.kind != TG_BPJT) cc.map.nt = 1;
  int rc = make_args<T>(&cc, 0, cc.nv, a);
  if (rc) return rc;
  if (!c->cube || !c->vox_weight) return set_error(TG_EINVAL, "scale needs cube and vox_weight");
  const int64_t total = (int64_t)a.n[0] * a.n[1] * a.n[2] * a.nt;
  const uint32_t blocks = (uint32_t)((total + 255) / 256 < 65536 ? (total + 255) / 256 : 65536);
  hipLaunchKernelGGL((k_scale_cube<T>), dim3(blocks), dim3(256), 0, s, a);
  return hip_check(hipGetLastError(), "scale kernel launch");
}

template <typename T>
int run_vox_apply(const tg_call *c, const void *src, const int64_t src_stride[4], hipStream_t s) {
  KArgs<T> a;
  tg_call cc = *c;
  if (cc.kind != TG_PJT && cc.kind != TG_BPJT) cc.map.nt = 1;
  int rc = make_args<T>(&cc, 0, cc.nv, a);
  if (rc) return rc;
  if (!c->cube || !src || !src_stride) return set_error(TG_EINVAL, "vox_apply needs cube and src");
  if (src == (const void *)c->cube) return set_error(TG_EINVAL, "vox_apply: src and cube must differ");
  const int64_t total = (int64_t)a.n[0] * a.n[1] * a.n[2] * a.nt;
  if (total == 0) return TG_OK;
  const int64_t n1 = a.n[1], n2 = a.n[2], nt = a.nt, nvox = (int64_t)a.n[0] * n1 * n2;
  const bool bins_fastest = nt > 1 && nt <= kVoxTileBins && src_stride[3] == 1 &&
                            src_stride[2] == nt && src_stride[1] == n2 * nt &&
                            src_stride[0] == n1 * n2 * nt;
  const bool dense_dst = a.cs[0] == n1 * n2 && a.cs[1] == n2 && a.cs[2] == 1 && a.cs[3] == nvox;
  const bool dense_vox = (!a.voxw && !a.voxmask) ||
                         (a.vs[0] == n1 * n2 && a.vs[1] == n2 && a.vs[2] == 1);
  if (bins_fastest && dense_dst && dense_vox) {
    const int64_t want = (nvox + kVoxTile - 1) / kVoxTile;
    const uint32_t blocks = (uint32_t)(want < 65536 ? want : 65536);
    hipLaunchKernelGGL((k_vox_apply_bins<T>), dim3(blocks), dim3(kVoxTile), 0, s, a, (const T *)src);
    return hip_check(hipGetLastError(), "vox_apply kernel launch");
  }
  const uint32_t blocks = (uint32_t)((total + 255) / 256 < 65536 ? (total + 255) / 256 : 65536);
  hipLaunchKernelGGL((k_vox_apply<T>), dim3(blocks), dim3(256), 0, s, a, (const T *)src,
                     src_stride[0], src_stride[1], src_stride[2], src_stride[3]);
  return hip_check(hipGetLastError(), "vox_apply kernel launch");
}

// ---------------------------------------------------------------------------
// View packing (host): C:1043-1046 pixel -> angle in T, then libm cos/sin in
// double as C:1015-1017 calls them.
// ---------------------------------------------------------------------------
template <typename T>
int pack_views(int nv, int n1, int n2, const double *hdr, void *out, size_t bytes) {
  const size_t need = views_hdr_bytes<T>(nv) + views_trig_bytes(nv, n1, n2);
  if (bytes < need) return set_error(TG_EINVAL, "views buffer too small: %zu < %zu", bytes, need);
  if (nv > 0 && !hdr) return set_error(TG_EINVAL, "null header records");
  memset(out, 0, need);
  ViewHdr<T> *vh = (ViewHdr<T> *)out;
  double *trig = (double *)((char *)out + views_hdr_bytes<T>(nv));
  for (int t = 0; t < nv; t++) {
    const double *h = hdr + (size_t)t * TG_HDR;
    ViewHdr<T> &v = vh[t];
    v.s[0] = (T)h[TG_H_CRPIX1];
    v.s[1] = (T)h[TG_H_CRPIX2];
    v.p[0] = (T)h[TG_H_CDELT1];
    v.p[1] = (T)h[TG_H_CDELT2];
    v.v[0] = (T)h[TG_H_CRVAL1];
    v.v[1] = (T)h[TG_H_CRVAL2];
    for (int k = 0; k < 9; k++) v.R[k] = (T)h[TG_H_R11 + k];
    for (int k = 0; k < 3; k++) v.M[k] = (T)h[TG_H_M1 + k];
    double *tr = trig + (size_t)t * ((size_t)n1 + n2) * 2;
    for (unsigned int i = 0; i < (unsigned)n1; i++) {
      const T gamma = (i - v.s[0] + 1) * v.p[0] + v.v[0];
      tr[2 * i] = cos((double)gamma);
      tr[2 * i + 1] = sin((double)gamma);
    }
    for (unsigned int j = 0; j < (unsigned)n2; j++) {
      const T lambda = (j - v.s[1] + 1) * v.p[1] + v.v[1];
      tr[2 * ((size_t)n1 + j)] = cos((double)lambda);
      tr[2 * ((size_t)n1 + j) + 1] = sin((double)lambda);
    }
  }
  return TG_OK;
}

}  // namespace tg

using namespace tg;

extern "C" {

const char *tg_last_error(void) { return g_err; }

const char *tg_version(void) { return "tomograpy_amd 0.1.0 gfx950"; }

size_t tg_views_bytes(int32_t dtype, int32_t nv, int32_t n1, int32_t n2) {
  if (nv < 0 || n1 < 0 || n2 < 0) return 0;
  if (dtype == TG_F32) return views_hdr_bytes<float>(nv) + views_trig_bytes(nv, n1, n2);
  if (dtype == TG_F64) return views_hdr_bytes<double>(nv) + views_trig_bytes(nv, n1, n2);
  return 0;
}

int tg_views_pack(int32_t dtype, int32_t nv, int32_t n1, int32_t n2, const double *hdr, void *out,
                  size_t out_bytes) {
  if (nv < 0 || n1 < 0 || n2 < 0) return set_error(TG_EINVAL, "negative view/image sizes");
  if (!out) return set_error(TG_EINVAL, "null output buffer");
  if (dtype == TG_F32) return pack_views<float>(nv, n1, n2, hdr, out, out_bytes);
  if (dtype == TG_F64) return pack_views<double>(nv, n1, n2, hdr, out, out_bytes);
  return set_error(TG_EINVAL, "bad dtype %d", dtype);
}

int tg_conic_full_projector(const tg_call *c, void *stream) {
  if (!c) return set_error(TG_EINVAL, "null call");
  return dispatch_siddon(c, 0, c->nv, -1, -1, stream);
}

int tg_conic_image_projector(const tg_call *c, int32_t t, void *stream) {
  if (!c) return set_error(TG_EINVAL, "null call");
  if (t < 0 || t >= c->nv) return set_error(TG_EINVAL, "image index %d outside [0, %d)", t, c->nv);
  return dispatch_siddon(c, t, t + 1, -1, -1, stream);
}

int tg_ray_projector(const tg_call *c, int32_t i, int32_t j, int32_t t, void *stream) {
  if (!c) return set_error(TG_EINVAL, "null call");
  if (t < 0 || t >= c->nv) return set_error(TG_EINVAL, "image index %d outside [0, %d)", t, c->nv);
  return dispatch_siddon(c, t, t + 1, i, j, stream);
}

int tg_unit_vector(const tg_call *c, int32_t t0, int32_t t1, void *u, void *stream) {
  if (!c) return set_error(TG_EINVAL, "null call");
  hipStream_t s = (hipStream_t)stream;
  if (c->dtype == TG_F32) return run_unit<float>(c, t0, t1, u, s);
  if (c->dtype == TG_F64) return run_unit<double>(c, t0, t1, u, s);
  return set_error(TG_EINVAL, "bad dtype %d", c->dtype);
}

int tg_intersection_parameters(const tg_call *c, int32_t t0, int32_t t1, const void *u, void *a1,
                               void *an, void *stream) {
  if (!c) return set_error(TG_EINVAL, "null call");
  hipStream_t s = (hipStream_t)stream;
  if (c->dtype == TG_F32) return run_inter<float>(c, t0, t1, u, a1, an, s);
  if (c->dtype == TG_F64) return run_inter<double>(c, t0, t1, u, a1, an, s);
  return set_error(TG_EINVAL, "bad dtype %d", c->dtype);
}

int tg_count_steps(const tg_call *c, uint32_t *per_ray, unsigned long long *total, void *stream) {
  if (!c) return set_error(TG_EINVAL, "null call");
  hipStream_t s = (hipStream_t)stream;
  if (c->dtype == TG_F32) return run_count<float>(c, per_ray, total, s);
  if (c->dtype == TG_F64) return run_count<double>(c, per_ray, total, s);
  return set_error(TG_EINVAL, "bad dtype %d", c->dtype);
}

int tg_scale_cube(const tg_call *c, void *stream) {
  if (!c) return set_error(TG_EINVAL, "null call");
  hipStream_t s = (hipStream_t)stream;
  if (c->dtype == TG_F32) return run_scale<float>(c, s);
  if (c->dtype == TG_F64) return run_scale<double>(c, s);
  return set_error(TG_EINVAL, "bad dtype %d", c->dtype);
}

int tg_vox_apply(const tg_call *c, const void *src, const int64_t src_stride[4], void *stream) {
  if (!c) return set_error(TG_EINVAL, "null call");
  hipStream_t s = (hipStream_t)stream;
  if (c->dtype == TG_F32) return run_vox_apply<float>(c, src, src_stride, s);
  if (c->dtype == TG_F64) return run_vox_apply<double>(c, src, src_stride, s);
  return set_error(TG_EINVAL, "bad dtype %d", c->dtype);
}

}  // extern "C"
