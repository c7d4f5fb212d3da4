// Per-ray bodies of the Siddon projector and backprojector, shared by the
// one-job kernels (siddon.hip) and the persistent multi-job kernel
// (siddon_batch.hip).  `tile` is the logical workgroup tile of the job.
#pragma once

#include "siddon_common.h"

namespace tg {

// Forward projection (pj / pjt), optionally with the voxel operands of
// O.P.M (weight, mask) and the pixel weight O.
template <typename T, bool SUN, bool VOX>
__device__ __forceinline__ void project_tile(const KArgs<T> &a, uint32_t tile) {
  int i, j, t;
  if (!lane_pixel(a, tile, i, j, t)) return;
  const int64_t did = i * a.ds[0] + j * a.ds[1] + (int64_t)t * a.ds[2];
  const bool zero_out = (a.flags & TG_ZERO_OUT) != 0;
  if (a.pixmask && a.pixmask[i * a.ms[0] + j * a.ms[1] + (int64_t)t * a.ms[2]] != 0) {
    if (zero_out) a.data[did] = (T)0;
    return;
  }

  const ViewHdr<T> &h = a.views[t];
  T u0[3];
  pixel_unit_vector(a, h, t, i, j, u0);
  Ray<T> r;
  const bool hit = ray_setup(a, h.M, u0, r);
  if (!hit || !ray_state_ok(r)) {
    if (zero_out) a.data[did] = (T)0;
    if (hit) flag_status(a.status, TG_EGEOMETRY);
    return;
  }
  int64_t toff;
  if (!time_offset(a, t, toff)) return;
  const T *cube = a.cube + toff;
  int64_t lin = r.iv[0] * a.cs[0] + r.iv[1] * a.cs[1] + r.iv[2] * a.cs[2];
  const int64_t st0 = r.up[0] * a.cs[0], st1 = r.up[1] * a.cs[1], st2 = r.up[2] * a.cs[2];
  int64_t lm = 0, sm0 = 0, sm1 = 0, sm2 = 0;
  if (VOX) {
    lm = r.iv[0] * a.vs[0] + r.iv[1] * a.vs[1] + r.iv[2] * a.vs[2];
    sm0 = r.up[0] * a.vs[0];
    sm1 = r.up[1] * a.vs[1];
    sm2 = r.up[2] * a.vs[2];
  }
  T acc = (a.pixw || zero_out) ? (T)0 : a.data[did];
  int guard = (int)(a.n[0] + a.n[1] + a.n[2]) + 2;
  while (in_map(a, r) && (!SUN || outside_sun(h.M, u0, r.ac))) {
    if (--guard < 0) {
      flag_status(a.status, TG_EGEOMETRY);
      break;
    }
    bool sx, sy, sz;
    const T d = step(r, sx, sy, sz);
    T v = cube[lin];
    bool keep = true;
    if (VOX) {
      if (a.voxw) v = a.voxw[lm] * v;
      if (a.voxmask) keep = (a.voxmask[lm] == 0);
      lm += (sx ? sm0 : 0) + (sy ? sm1 : 0) + (sz ? sm2 : 0);
    }
    if (keep) acc = acc + d * v;
    lin += (sx ? st0 : 0) + (sy ? st1 : 0) + (sz ? st2 : 0);
  }
  if (a.pixw)
    a.data[did] = (zero_out ? (T)0 : a.data[did]) + a.pixw[did] * acc;
  else
    a.data[did] = acc;
}

template <typename T>
__device__ __forceinline__ void atomic_add(T *p, T v) {
  unsafeAtomicAdd(p, v);
}

// Backprojection (bpj / bpjt) as an atomic scatter.
template <typename T, bool SUN, bool VOX>
__device__ __forceinline__ void backproject_tile(const KArgs<T> &a, uint32_t tile) {
  int i, j, t;
  if (!lane_pixel(a, tile, i, j, t)) return;
  if (a.pixmask && a.pixmask[i * a.ms[0] + j * a.ms[1] + (int64_t)t * a.ms[2]] != 0) return;
  const int64_t did = i * a.ds[0] + j * a.ds[1] + (int64_t)t * a.ds[2];
  T w = a.data[did];
  if (a.pixw) w = a.pixw[did] * w;
  // A zero pixel adds d * 0 to every voxel it crosses: nothing to do.
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
  T *cube = a.cube + toff;
  int64_t lin = r.iv[0] * a.cs[0] + r.iv[1] * a.cs[1] + r.iv[2] * a.cs[2];
  const int64_t st0 = r.up[0] * a.cs[0], st1 = r.up[1] * a.cs[1], st2 = r.up[2] * a.cs[2];
  int64_t lm = 0, sm0 = 0, sm1 = 0, sm2 = 0;
  if (VOX) {
    lm = r.iv[0] * a.vs[0] + r.iv[1] * a.vs[1] + r.iv[2] * a.vs[2];
    sm0 = r.up[0] * a.vs[0];
    sm1 = r.up[1] * a.vs[1];
    sm2 = r.up[2] * a.vs[2];
  }
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
    if (keep) atomic_add(cube + lin, d * w);
    lin += (sx ? st0 : 0) + (sy ? st1 : 0) + (sz ? st2 : 0);
  }
}

}  // namespace tg
