// Deterministic backprojection: each voxel is summed in the order the
// reference's single-threaded run adds into it (C:489-494 with nthread=1:
// view t, then pixel row i, then column j; C:923-928), so the result is
// bit-identical to that run and independent of scheduling.
//
// The rays are taken in chunks of consecutive pixel rows in (view, row)
// order.  Per chunk:
//   1. count   contributions per voxel (atomic counters),
//   2. group   voxels in runs of 64 consecutive linear indices (one wave
//              lane each); a group's segment holds max-count slots per
//              voxel, interleaved: slot p of the voxel in lane l sits at
//              (group offset + p) * 64 + l, so the reduce's loads coalesce;
//              the counters are reset to serve as slot cursors,
//   3. scan    group lengths -> group offsets (exclusive, 64-bit),
//   4. fill    each contribution d * w with its ray key (row, column
//              relative to the chunk) into the next slot of its voxel, one
//              launch per view in view order, so a voxel's slots hold the
//              views in order and only the few rays of one view that cross
//              the voxel are in atomic order,
//   5. reduce  one lane per voxel insertion-sorts its slots by ray key
//              (linear work: only those per-view runs are out of order) and
//              adds the values into the voxel in that order.
// Chunks run in ray order, so the per-voxel order across chunks is the
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
constexpr int kGroup = 64;  // voxels per interleaved group: one wave

// One contribution: its ray key and d * w, written with one store.
template <typename T>
struct DetRec;
template <>
struct alignas(8) DetRec<float> {
  uint32_t key;
  float val;
};
template <>
struct alignas(16) DetRec<double> {
  uint32_t key;
  uint32_t pad;
  double val;
};

template <typename T, bool SUN, bool VOX, bool FILL>
__global__ __launch_bounds__(256) void k_det_trace(KArgs<T> a, uint32_t *counts,
                                                   const unsigned long long *goff,
                                                   DetRec<T> *recs, int64_t row0, int64_t row1) {
  const uint32_t tile = xcd_remap(blockIdx.x, gridDim.x);
  int i, j, t;
  if (!lane_pixel(a, tile, i, j, t)) return;
  const int64_t row = (int64_t)t * a.n1 + i;
  if (row < row0 || row >= row1) return;
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
  const uint32_t key = (uint32_t)(row - row0) * (uint32_t)a.n2 + (uint32_t)j;
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
        const uint32_t slot = atomicAdd(counts + lin, 1u);
        const int64_t g = lin / kGroup;
        const unsigned long long base = goff[g];
        if (base + slot >= goff[g + 1]) {
          // the fill met more contributions than the count pass: never write
          // outside the group's segment
          flag_status(a.status, TG_EGEOMETRY);
        } else {
          DetRec<T> rec;
          rec.key = key;
          if constexpr (sizeof(DetRec<T>) == 16) rec.pad = 0u;
          rec.val = d * w;
          recs[(base + slot) * kGroup + (unsigned long long)(lin - g * kGroup)] = rec;
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

// One wave per group of 64 voxels: the group's length is the largest count
// among its voxels; the counters are cleared to serve as the fill's cursors.
__global__ __launch_bounds__(256) void k_det_group(int64_t ngroups, uint32_t *counts,
                                                   uint32_t *glen) {
  const int64_t nvox = ngroups * kGroup;
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < nvox;
       v += (int64_t)gridDim.x * blockDim.x) {
    uint32_t m = counts[v];
    counts[v] = 0u;
    for (int o = kGroup / 2; o > 0; o >>= 1) {
      const uint32_t x = (uint32_t)__shfl_xor((int)m, o, kGroup);
      m = x > m ? x : m;
    }
    if ((v & (kGroup - 1)) == 0) glen[v / kGroup] = m;
  }
}

// One lane per voxel (a wave per group): insertion-sort the voxel's slots by
// ray key, then add the values into the voxel in key order.
template <typename T>
__global__ __launch_bounds__(256) void k_det_reduce(int64_t n, const uint32_t *counts,
                                                    const unsigned long long *goff,
                                                    DetRec<T> *recs, T *cube) {
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n;
       v += (int64_t)gridDim.x * blockDim.x) {
    const int64_t g = v / kGroup;
    // a cursor past the group's length means the fill flagged an error; read
    // no further than the segment
    const uint32_t glen = (uint32_t)(goff[g + 1] - goff[g]);
    const uint32_t len = counts[v] < glen ? counts[v] : glen;
    if (len == 0u) continue;
    DetRec<T> *seg = recs + goff[g] * kGroup + (uint64_t)(v & (kGroup - 1));
    uint32_t prev = seg[0].key;
    for (uint32_t p = 1; p < len; p++) {
      const DetRec<T> x = seg[(uint64_t)p * kGroup];
      if (x.key > prev) {
        prev = x.key;
        continue;
      }
      uint32_t q = p;
      while (q > 0 && seg[(uint64_t)(q - 1) * kGroup].key > x.key) {
        seg[(uint64_t)q * kGroup] = seg[(uint64_t)(q - 1) * kGroup];
        q--;
      }
      seg[(uint64_t)q * kGroup] = x;
    }
    T acc = cube[v];
    for (uint32_t p = 0; p < len; p++) acc = acc + seg[(uint64_t)p * kGroup].val;
    cube[v] = acc;
  }
}

struct DetLayout {
  int64_t extent;
  int64_t ngroups;
  int64_t ntiles;
  size_t off_counts, off_glen, off_offsets, off_tiles, off_records;
  size_t fixed;
};

DetLayout det_layout(const tg_call *c) {
  DetLayout L;
  L.extent = cube_extent(c);
  L.ngroups = (L.extent + kGroup - 1) / kGroup;
  L.ntiles = (L.ngroups + kScanTile - 1) / kScanTile;
  L.off_counts = 0;
  L.off_glen = align_up(L.off_counts + sizeof(uint32_t) * (size_t)(L.ngroups * kGroup), 256);
  L.off_offsets = align_up(L.off_glen + sizeof(uint32_t) * (size_t)L.ngroups, 256);
  L.off_tiles = align_up(L.off_offsets + sizeof(unsigned long long) * (size_t)(L.ngroups + 1), 256);
  L.off_records = align_up(L.off_tiles + sizeof(unsigned long long) * (size_t)(L.ntiles + 1), 256);
  L.fixed = L.off_records;
  return L;
}

// Records one row of rays can need: a ray adds at most n0 + n1 + n2 + 2
// contributions and a group never holds more than 64 times its voxels'
// records.
size_t det_row_records(const tg_call *c) {
  return (size_t)kGroup * (size_t)c->n2 * (size_t)(c->map.n[0] + c->map.n[1] + c->map.n[2] + 2);
}

template <typename T, bool FILL>
void launch_det_trace(const KArgs<T> &a, uint32_t blocks, hipStream_t s, bool sun, bool vox,
                      uint32_t *counts, const unsigned long long *goff, DetRec<T> *recs,
                      int64_t row0, int64_t row1) {
  if (sun) {
    if (vox)
      hipLaunchKernelGGL((k_det_trace<T, true, true, FILL>), dim3(blocks), dim3(256), 0, s, a,
                         counts, goff, recs, row0, row1);
    else
      hipLaunchKernelGGL((k_det_trace<T, true, false, FILL>), dim3(blocks), dim3(256), 0, s, a,
                         counts, goff, recs, row0, row1);
  } else {
    if (vox)
      hipLaunchKernelGGL((k_det_trace<T, false, true, FILL>), dim3(blocks), dim3(256), 0, s, a,
                         counts, goff, recs, row0, row1);
    else
      hipLaunchKernelGGL((k_det_trace<T, false, false, FILL>), dim3(blocks), dim3(256), 0, s, a,
                         counts, goff, recs, row0, row1);
  }
}

uint32_t grid_for(int64_t n) {
  const int64_t b = (n + 255) / 256;
  return (uint32_t)(b < 65536 ? (b > 0 ? b : 1) : 65536);
}

}  // namespace

template <typename T>
int run_deterministic(const tg_call *c, const KArgs<T> &a0, hipStream_t s) {
  for (int k = 0; k < 4; k++)
    if (c->map.stride[k] < 0) return set_error(TG_EINVAL, "deterministic mode needs non-negative cube strides");
  const DetLayout L = det_layout(c);
  const size_t rec_bytes = sizeof(DetRec<T>);
  if (!c->workspace || c->workspace_bytes < L.fixed + 512 + rec_bytes * kGroup)
    return set_error(TG_EWORKSPACE, "deterministic mode needs a workspace of at least %zu bytes",
                     tg_det_workspace_min_bytes(c));
  char *ws = (char *)c->workspace;
  uint32_t *counts = (uint32_t *)(ws + L.off_counts);
  uint32_t *glen = (uint32_t *)(ws + L.off_glen);
  unsigned long long *goff = (unsigned long long *)(ws + L.off_offsets);
  unsigned long long *tiles = (unsigned long long *)(ws + L.off_tiles);
  DetRec<T> *recs = (DetRec<T> *)(ws + L.off_records);
  // capacity in records, a whole number of groups
  const unsigned long long cap =
      (c->workspace_bytes - L.fixed - 512) / rec_bytes / kGroup * kGroup;
  const bool sun = (c->flags & TG_OBSTACLE_SUN) != 0;
  const bool vox = c->voxmask != nullptr;
  const int64_t nvox = L.ngroups * kGroup;
  const int64_t row_end = (int64_t)a0.t1 * a0.n1;
  // ray keys are 32-bit: (row - row0) * n2 + j
  int64_t max_rows = (int64_t)(0xffffffffull / (uint64_t)(a0.n2 > 0 ? a0.n2 : 1));
  if (max_rows < 1) return set_error(TG_EINVAL, "image too wide for 32-bit ray keys");
  // the count pass launches every view a chunk touches in one grid; a chunk of
  // rows touches at most rows / n1 + 2 views
  const uint64_t per_view_groups = (uint64_t)a0.tiles_i * (uint64_t)a0.tiles_j;
  if (per_view_groups > kMaxGroups256 / 2)
    return set_error(TG_EINVAL, "a %d x %d image is too large for deterministic mode", a0.n1, a0.n2);
  const int64_t view_cap = (int64_t)(kMaxGroups256 / per_view_groups) - 2;
  if (view_cap < 1) max_rows = 1;
  else if (view_cap * (int64_t)a0.n1 < max_rows) max_rows = view_cap * (int64_t)a0.n1;
  // first chunk: what fits even if every ray crossed the whole cube
  const size_t per_row_bound = det_row_records(c) / kGroup;
  int64_t chunk = (int64_t)(cap / (per_row_bound > 0 ? per_row_bound : 1));
  if (chunk < 1) chunk = 1;
  if (chunk > max_rows) chunk = max_rows;
  const uint32_t view_blocks = (uint32_t)a0.tiles_i * (uint32_t)a0.tiles_j;
  int64_t row = (int64_t)a0.t0 * a0.n1;
  int rc;
  while (row < row_end) {
    const int64_t k = chunk < row_end - row ? chunk : row_end - row;
    const int64_t r1 = row + k;
    const int v0 = (int)(row / a0.n1), v1 = (int)((r1 - 1) / a0.n1) + 1;
    KArgs<T> a = a0;
    a.t0 = v0;
    a.t1 = v1;
    hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(nvox)), dim3(256), 0, s, counts, nvox);
    launch_det_trace<T, false>(a, (uint32_t)(v1 - v0) * view_blocks, s, sun, vox, counts,
                               nullptr, nullptr, row, r1);
    hipLaunchKernelGGL(k_det_group, dim3(grid_for(nvox)), dim3(256), 0, s, L.ngroups, counts,
                       glen);
    hipLaunchKernelGGL(k_scan_tiles, dim3((uint32_t)L.ntiles), dim3(kScanBlock), 0, s, glen,
                       L.ngroups, tiles);
    hipLaunchKernelGGL(k_scan_tile_sums, dim3(1), dim3(kScanBlock), 0, s, tiles, L.ntiles);
    hipLaunchKernelGGL(k_scan_apply, dim3((uint32_t)L.ntiles), dim3(kScanBlock), 0, s, glen,
                       L.ngroups, tiles, goff);
    if ((rc = hip_check(hipGetLastError(), "deterministic count/scan launch"))) return rc;
    unsigned long long total = 0;
    if ((rc = hip_check(hipMemcpyAsync(&total, goff + L.ngroups, sizeof(total),
                                       hipMemcpyDeviceToHost, s),
                        "deterministic total copy")))
      return rc;
    if ((rc = hip_check(hipStreamSynchronize(s), "deterministic total sync"))) return rc;
    const unsigned long long need = total * kGroup;
    if (need > cap) {
      if (k == 1)
        return set_error(TG_EWORKSPACE,
                         "one pixel row needs %llu records but the workspace holds %llu", need,
                         cap);
      const int64_t nk = (int64_t)((double)k * (double)cap / (double)need * 0.9);
      chunk = nk < 1 ? 1 : (nk < k ? nk : k - 1);
      continue;
    }
    if (total > 0) {
      // one launch per view, in view order (stream order), so each voxel's
      // slots receive the views in order
      for (int tv = v0; tv < v1; tv++) {
        KArgs<T> av = a;
        av.t0 = tv;
        av.t1 = tv + 1;
        launch_det_trace<T, true>(av, view_blocks, s, sun, vox, counts, goff, recs, row, r1);
      }
      hipLaunchKernelGGL((k_det_reduce<T>), dim3(grid_for(L.extent)), dim3(256), 0, s, L.extent,
                         counts, goff, recs, (T *)c->cube);
      if ((rc = hip_check(hipGetLastError(), "deterministic fill/reduce launch"))) return rc;
    }
    row = r1;
    // grow the chunk when the last one left room
    if (total > 0 && need * 2 < cap) {
      const double per_row = (double)need / (double)k;
      const int64_t nk = (int64_t)((double)cap * 0.8 / per_row);
      if (nk > chunk) chunk = nk < max_rows ? nk : max_rows;
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
  const DetLayout L = det_layout(c);
  const size_t rec_bytes = c->dtype == TG_F64 ? sizeof(DetRec<double>) : sizeof(DetRec<float>);
  return L.fixed + 512 + det_row_records(c) * rec_bytes;
}
