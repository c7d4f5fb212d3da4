// Deterministic backprojection: each voxel is summed in the order the
// reference's single-threaded run adds into it (C:489-494 with nthread=1:
// view t, then pixel row i, then column j; C:923-928), so the result is
// bit-identical to that run and independent of scheduling.
//
// Per chunk of consecutive views:
//   1. count   contributions per voxel (atomic counters),
//   2. scan    counts -> segment offsets (exclusive, 64-bit),
//   3. fill    each contribution d * w with its ray key (t, i, j) into its
//              voxel's segment, one launch per view in view order; a voxel's
//              slots are handed out upwards, so its segment holds the views
//              in order and only the few rays of one view that cross the
//              voxel are in atomic order,
//   4. reduce  one lane per voxel insertion-sorts its segment by ray key
//              (linear work: only those per-view groups are out of order)
//              and adds the values into the voxel in that order.
// Chunks run in view order, so the per-voxel order across chunks is the
// reference's too.  The chunk size adapts to the workspace.
#include <hip/hip_runtime.h>
#include <string.h>

#include "siddon_common.h"
#include "siddon_host.h"

namespace tg {

namespace {

constexpr int kScanBlock = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanBlock * kScanItems;

template <typename T, bool SUN, bool VOX, bool FILL>
__global__ __launch_bounds__(256) void k_det_trace(KArgs<T> a, uint32_t *counts,
                                                   const unsigned long long *offsets,
                                                   uint32_t *keys, T *vals, int key_t0) {
  const uint32_t tile = xcd_remap(blockIdx.x, gridDim.x);
  int i, j, t;
  if (!lane_pixel(a, tile, i, j, t)) return;
  if (a.pixmask && a.pixmask[i * a.ms[0] + j * a.ms[1] + (int64_t)t * a.ms[2]] != 0) return;
  const int64_t did = i * a.ds[0] + j * a.ds[1] + (int64_t)t * a.ds[2];
  T w = a.data[did];
  if (a.pixw) w = a.pixw[did] * w;
  if (w == (T)0) return;
  const ViewHdr<T> &h = a.views[t];
  T u0[3];
  pixel_unit_vector(a, h, t, i, j, u0);
  Ray<T> r;
  if (!ray_setup(a, h.M, u0, r)) return;
  if (!ray_state_ok(r)) {
    flag_status(a.status, TG_EGEOMETRY);
    return;
  }
  int64_t toff;
  if (!time_offset(a, t, toff)) return;
  int64_t lin = toff + r.iv[0] * a.cs[0] + r.iv[1] * a.cs[1] + r.iv[2] * a.cs[2];
  const int64_t st0 = r.up[0] * a.cs[0], st1 = r.up[1] * a.cs[1], st2 = r.up[2] * a.cs[2];
  int64_t lm = 0, sm0 = 0, sm1 = 0, sm2 = 0;
  if (VOX) {
    lm = r.iv[0] * a.vs[0] + r.iv[1] * a.vs[1] + r.iv[2] * a.vs[2];
    sm0 = r.up[0] * a.vs[0];
    sm1 = r.up[1] * a.vs[1];
    sm2 = r.up[2] * a.vs[2];
  }
  const uint32_t key = ((uint32_t)(t - key_t0) * (uint32_t)a.n1 + (uint32_t)i) * (uint32_t)a.n2 +
                       (uint32_t)j;
  int guard = (int)(a.n[0] + a.n[1] + a.n[2]) + 2;
  while (in_map(a, r) && (!SUN || outside_sun(h.M, u0, r.ac))) {
    if (--guard < 0) {
      flag_status(a.status, TG_EGEOMETRY);
      break;
    }
    bool sx, sy, sz;
    const T d = step(r, sx, sy, sz);
    bool keep = true;
    if (VOX) {
      if (a.voxmask) keep = (a.voxmask[lm] == 0);
      lm += (sx ? sm0 : 0) + (sy ? sm1 : 0) + (sz ? sm2 : 0);
    }
    if (keep) {
      if (FILL) {
        // the counter runs down from the voxel's count: the first taker gets
        // the segment's first slot
        const uint32_t left = atomicSub(counts + lin, 1u);
        if (left == 0u) {
          // the fill met more contributions than the count pass: never write
          // outside the voxel's segment
          flag_status(a.status, TG_EGEOMETRY);
        } else {
          const unsigned long long pos = offsets[lin + 1] - left;
          keys[pos] = key;
          vals[pos] = d * w;
        }
      } else {
        atomicAdd(counts + lin, 1u);
      }
    }
    lin += (sx ? st0 : 0) + (sy ? st1 : 0) + (sz ? st2 : 0);
  }
}

__global__ __launch_bounds__(256) void k_fill_u32(uint32_t *p, int64_t n) {
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n;
       k += (int64_t)gridDim.x * blockDim.x)
    p[k] = 0u;
}

// Exclusive block-level scan of per-thread sums in LDS (Hillis-Steele).
__device__ unsigned long long block_exclusive_scan(unsigned long long v, unsigned long long *lds,
                                                   unsigned long long &block_total) {
  const int tid = threadIdx.x;
  lds[tid] = v;
  __syncthreads();
  for (int off = 1; off < kScanBlock; off <<= 1) {
    unsigned long long x = (tid >= off) ? lds[tid - off] : 0ull;
    __syncthreads();
    lds[tid] += x;
    __syncthreads();
  }
  block_total = lds[kScanBlock - 1];
  const unsigned long long incl = lds[tid];
  __syncthreads();
  return incl - v;
}

__global__ __launch_bounds__(kScanBlock) void k_scan_tiles(const uint32_t *cnt, int64_t n,
                                                           unsigned long long *tile_sums) {
  __shared__ unsigned long long lds[kScanBlock];
  const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
  unsigned long long s = 0;
  for (int q = 0; q < kScanItems; q++)
    if (base + q < n) s += cnt[base + q];
  unsigned long long tot;
  block_exclusive_scan(s, lds, tot);
  if (threadIdx.x == 0) tile_sums[blockIdx.x] = tot;
}

// Single workgroup: exclusive scan of the tile sums in place; tile_sums[ntiles]
// receives the grand total.
__global__ __launch_bounds__(kScanBlock) void k_scan_tile_sums(unsigned long long *tile_sums,
                                                               int64_t ntiles) {
  __shared__ unsigned long long lds[kScanBlock];
  unsigned long long carry = 0;
  for (int64_t b0 = 0; b0 < ntiles; b0 += kScanBlock) {
    const int64_t k = b0 + threadIdx.x;
    const unsigned long long v = (k < ntiles) ? tile_sums[k] : 0ull;
    unsigned long long tot;
    const unsigned long long ex = block_exclusive_scan(v, lds, tot);
    if (k < ntiles) tile_sums[k] = carry + ex;
    carry += tot;
  }
  if (threadIdx.x == 0) tile_sums[ntiles] = carry;
}

__global__ __launch_bounds__(kScanBlock) void k_scan_apply(const uint32_t *cnt, int64_t n,
                                                           const unsigned long long *tile_sums,
                                                           unsigned long long *off) {
  __shared__ unsigned long long lds[kScanBlock];
  const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
  unsigned long long s = 0;
  for (int q = 0; q < kScanItems; q++)
    if (base + q < n) s += cnt[base + q];
  unsigned long long tot;
  unsigned long long run = block_exclusive_scan(s, lds, tot) + tile_sums[blockIdx.x];
  for (int q = 0; q < kScanItems; q++) {
    if (base + q < n) {
      off[base + q] = run;
      run += cnt[base + q];
    }
  }
  if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) off[n] = tile_sums[gridDim.x];
}

// One lane per voxel: insertion-sort the segment by ray key, then add the
// values into the voxel in key order.
template <typename T>
__global__ __launch_bounds__(256) void k_det_reduce(int64_t n, const unsigned long long *off,
                                                    uint32_t *keys, T *vals, T *cube) {
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n;
       v += (int64_t)gridDim.x * blockDim.x) {
    const unsigned long long b = off[v], e = off[v + 1];
    if (b == e) continue;
    for (unsigned long long p = b + 1; p < e; p++) {
      const uint32_t k = keys[p];
      const T x = vals[p];
      unsigned long long q = p;
      while (q > b && keys[q - 1] > k) {
        keys[q] = keys[q - 1];
        vals[q] = vals[q - 1];
        q--;
      }
      keys[q] = k;
      vals[q] = x;
    }
    T acc = cube[v];
    for (unsigned long long p = b; p < e; p++) acc = acc + vals[p];
    cube[v] = acc;
  }
}

struct DetLayout {
  int64_t extent;
  int64_t ntiles;
  size_t off_counts, off_offsets, off_tiles, off_records;
  size_t fixed;
};

template <typename T>
DetLayout det_layout(const tg_call *c) {
  DetLayout L;
  L.extent = cube_extent(c);
  L.ntiles = (L.extent + kScanTile - 1) / kScanTile;
  L.off_counts = 0;
  L.off_offsets = align_up(L.off_counts + sizeof(uint32_t) * (size_t)L.extent, 256);
  L.off_tiles = align_up(L.off_offsets + sizeof(unsigned long long) * (size_t)(L.extent + 1), 256);
  L.off_records = align_up(L.off_tiles + sizeof(unsigned long long) * (size_t)(L.ntiles + 1), 256);
  L.fixed = L.off_records;
  return L;
}

template <typename T, bool FILL>
void launch_det_trace(const KArgs<T> &a, uint32_t blocks, hipStream_t s, bool sun, bool vox,
                      uint32_t *counts, const unsigned long long *off, uint32_t *keys, T *vals,
                      int key_t0) {
  if (sun) {
    if (vox)
      hipLaunchKernelGGL((k_det_trace<T, true, true, FILL>), dim3(blocks), dim3(256), 0, s, a,
                         counts, off, keys, vals, key_t0);
    else
      hipLaunchKernelGGL((k_det_trace<T, true, false, FILL>), dim3(blocks), dim3(256), 0, s, a,
                         counts, off, keys, vals, key_t0);
  } else {
    if (vox)
      hipLaunchKernelGGL((k_det_trace<T, false, true, FILL>), dim3(blocks), dim3(256), 0, s, a,
                         counts, off, keys, vals, key_t0);
    else
      hipLaunchKernelGGL((k_det_trace<T, false, false, FILL>), dim3(blocks), dim3(256), 0, s, a,
                         counts, off, keys, vals, key_t0);
  }
}

}  // namespace

template <typename T>
int run_deterministic(const tg_call *c, const KArgs<T> &a0, hipStream_t s) {
  for (int k = 0; k < 4; k++)
    if (c->map.stride[k] < 0) return set_error(TG_EINVAL, "deterministic mode needs non-negative cube strides");
  const DetLayout L = det_layout<T>(c);
  if (!c->workspace || c->workspace_bytes <= L.fixed)
    return set_error(TG_EWORKSPACE, "deterministic mode needs a workspace of more than %zu bytes",
                     L.fixed);
  char *ws = (char *)c->workspace;
  uint32_t *counts = (uint32_t *)(ws + L.off_counts);
  unsigned long long *off = (unsigned long long *)(ws + L.off_offsets);
  unsigned long long *tiles = (unsigned long long *)(ws + L.off_tiles);
  const size_t rec_bytes = sizeof(uint32_t) + sizeof(T);
  const unsigned long long cap = (c->workspace_bytes - L.fixed - 512) / rec_bytes;
  uint32_t *keys = (uint32_t *)(ws + L.off_records);
  T *vals = (T *)(ws + align_up(L.off_records + sizeof(uint32_t) * cap, 256));
  const bool sun = (c->flags & TG_OBSTACLE_SUN) != 0;
  const bool vox = c->voxmask != nullptr;
  const uint32_t fill_blocks = (uint32_t)((L.extent + 255) / 256 < 65536 ? (L.extent + 255) / 256 : 65536);
  const uint64_t rays_per_view = (uint64_t)a0.n1 * (uint64_t)a0.n2;
  int max_chunk = rays_per_view ? (int)((0xffffffffull / rays_per_view) < 4096 ? (0xffffffffull / rays_per_view) : 4096) : 4096;
  if (max_chunk < 1) return set_error(TG_EINVAL, "image too large for 32-bit ray keys");
  int chunk = a0.t1 - a0.t0 < max_chunk ? a0.t1 - a0.t0 : max_chunk;
  int t = a0.t0;
  int rc;
  while (t < a0.t1) {
    int k = chunk < a0.t1 - t ? chunk : a0.t1 - t;
    KArgs<T> a = a0;
    a.t0 = t;
    a.t1 = t + k;
    const uint32_t blocks = (uint32_t)k * (uint32_t)a.tiles_i * (uint32_t)a.tiles_j;
    hipLaunchKernelGGL(k_fill_u32, dim3(fill_blocks), dim3(256), 0, s, counts, L.extent);
    launch_det_trace<T, false>(a, blocks, s, sun, vox, counts, nullptr, nullptr, nullptr, t);
    hipLaunchKernelGGL(k_scan_tiles, dim3((uint32_t)L.ntiles), dim3(kScanBlock), 0, s, counts,
                       L.extent, tiles);
    hipLaunchKernelGGL(k_scan_tile_sums, dim3(1), dim3(kScanBlock), 0, s, tiles, L.ntiles);
    hipLaunchKernelGGL(k_scan_apply, dim3((uint32_t)L.ntiles), dim3(kScanBlock), 0, s, counts,
                       L.extent, tiles, off);
    if ((rc = hip_check(hipGetLastError(), "deterministic count/scan launch"))) return rc;
    unsigned long long total = 0;
    if ((rc = hip_check(hipMemcpyAsync(&total, off + L.extent, sizeof(total),
                                       hipMemcpyDeviceToHost, s),
                        "deterministic total copy")))
      return rc;
    if ((rc = hip_check(hipStreamSynchronize(s), "deterministic total sync"))) return rc;
    if (total > cap) {
      if (k == 1)
        return set_error(TG_EWORKSPACE,
                         "one view needs %llu records but the workspace holds %llu", total, cap);
      int nk = (int)((double)k * (double)cap / (double)total * 0.9);
      chunk = nk < 1 ? 1 : (nk < k ? nk : k - 1);
      continue;
    }
    if (total > 0) {
      // one launch per view, in view order (stream order), so each voxel's
      // segment receives the views in order
      const uint32_t view_blocks = (uint32_t)a.tiles_i * (uint32_t)a.tiles_j;
      for (int tv = t; tv < t + k; tv++) {
        KArgs<T> av = a;
        av.t0 = tv;
        av.t1 = tv + 1;
        launch_det_trace<T, true>(av, view_blocks, s, sun, vox, counts, off,
                                  keys, vals, t);
      }
      const uint32_t red_blocks =
          (uint32_t)((L.extent + 255) / 256 < 65536 ? (L.extent + 255) / 256 : 65536);
      hipLaunchKernelGGL((k_det_reduce<T>), dim3(red_blocks), dim3(256), 0, s, L.extent, off, keys,
                         vals, (T *)c->cube);
      if ((rc = hip_check(hipGetLastError(), "deterministic fill/reduce launch"))) return rc;
    }
    t += k;
    // grow the chunk when the last one left room
    if (total > 0 && total * 2 < cap) {
      const double per_view = (double)total / k;
      int nk = (int)((double)cap * 0.8 / per_view);
      if (nk > chunk) chunk = nk < max_chunk ? nk : max_chunk;
    }
  }
  return TG_OK;
}

template int run_deterministic<float>(const tg_call *, const KArgs<float> &, hipStream_t);
template int run_deterministic<double>(const tg_call *, const KArgs<double> &, hipStream_t);

}  // namespace tg

extern "C" size_t tg_det_workspace_min_bytes(const tg_call *c) {
  using namespace tg;
  if (!c) return 0;
  const size_t per_view_bound = (size_t)c->n1 * (size_t)c->n2 *
                                (size_t)(c->map.n[0] + c->map.n[1] + c->map.n[2] + 2);
  if (c->dtype == TG_F64) {
    DetLayout L = det_layout<double>(c);
    return L.fixed + 1024 + per_view_bound * (sizeof(uint32_t) + sizeof(double));
  }
  DetLayout L = det_layout<float>(c);
  return L.fixed + 1024 + per_view_bound * (sizeof(uint32_t) + sizeof(float));
}
