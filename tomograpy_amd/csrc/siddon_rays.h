// Per-ray bodies of the Siddon projector and backprojector, shared by the
// one-job kernels (siddon.hip) and the persistent multi-job kernel
// (siddon_batch.hip).  `tile` is the logical workgroup tile of the job.
// FAST selects the 32-bit counter walk (walk_fast_ok on the host).
#pragma once

#include "siddon_common.h"

namespace tg {

// Descriptor word 3 of a raw buffer resource on the gfx9 family (32-bit data
// format, no swizzle): loads past num_records return 0.
constexpr int kBufferWord3 = 0x00020000;

template <typename T>
__device__ __forceinline__ T buffer_load(__amdgpu_buffer_rsrc_t rs, uint32_t byte_off);

template <>
__device__ __forceinline__ float buffer_load<float>(__amdgpu_buffer_rsrc_t rs, uint32_t byte_off) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rs, (int)byte_off, 0, 0));
}

template <>
__device__ __forceinline__ double buffer_load<double>(__amdgpu_buffer_rsrc_t rs, uint32_t byte_off) {
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(rs, (int)byte_off, 0, 0));
}

// Forward projection (pj / pjt), optionally with the voxel operands of
// O.P.M (weight, mask) and the pixel weight O.
template <typename T, bool SUN, bool VOX, bool FAST>
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
  T acc = (a.pixw || zero_out) ? (T)0 : a.data[did];
  if (FAST && a.cube_bytes && a.ray_steps && in_map(a, r)) {
    // The ray's step count is known (value independent, counted with the same
    // traversal, obstacle included): the loop needs no per-axis counters, no
    // exit test on them and no obstacle test; the voxel offsets move by the
    // strides of the stepping axes.  Same steps and the same floating-point
    // operations as the counter walk below.
    const uint32_t nsteps =
        a.ray_steps[((uint32_t)t * (uint32_t)a.n1 + (uint32_t)i) * (uint32_t)a.n2 + (uint32_t)j];
    const __amdgpu_buffer_rsrc_t rs =
        __builtin_amdgcn_make_buffer_rsrc((void *)(a.cube + toff), (short)0, (int)a.cube_bytes,
                                          kBufferWord3);
    int dm[3], dv[3];
    uint32_t off = 0, offv = 0;
#pragma unroll
    for (int k = 0; k < 3; k++) {
      dm[k] = r.up[k] * (int)a.cs[k] * (int)sizeof(T);
      off += (uint32_t)r.iv[k] * (uint32_t)a.cs[k] * (uint32_t)sizeof(T);
      if (VOX) {
        dv[k] = r.up[k] * (int)a.vs[k];
        offv += (uint32_t)r.iv[k] * (uint32_t)a.vs[k];
      }
    }
    // One step: `cur` is this voxel's value, `next` receives the next voxel's
    // load, which is in flight during the following step.
    auto step = [&](const T cur_in, T &next) {
      T cur = cur_in;
      bool keep = true;
      if (VOX) {
        if (a.voxw) cur = a.voxw[offv] * cur;
        if (a.voxmask) keep = (a.voxmask[offv] == 0);
      }
      const T d = tmin(tmin(r.D[0], r.D[1]), r.D[2]);
      const bool sx = (r.D[0] == d), sy = (r.D[1] == d), sz = (r.D[2] == d);
      r.D[0] = sx ? r.pabs[0] : r.D[0] - d;
      r.D[1] = sy ? r.pabs[1] : r.D[1] - d;
      r.D[2] = sz ? r.pabs[2] : r.D[2] - d;
      off += (uint32_t)((sx ? dm[0] : 0) + (sy ? dm[1] : 0) + (sz ? dm[2] : 0));
      if (VOX) offv += (uint32_t)((sx ? dv[0] : 0) + (sy ? dv[1] : 0) + (sz ? dv[2] : 0));
      next = buffer_load<T>(rs, off);
      if (keep) acc = acc + d * cur;
    };
    // Unrolled by two so that the two values alternate registers: no copy of a
    // value whose load is still in flight.  The pairs of steps every lane of
    // the wavefront walks run on a scalar counter, the rest per lane.
    T v0 = buffer_load<T>(rs, off), v1 = (T)0;
    const uint32_t common = wave_floor(nsteps, __builtin_amdgcn_readfirstlane(nsteps)) & ~1u;
    uint32_t s = 0;
    for (; s < common; s += 2) {
      step(v0, v1);
      step(v1, v0);
    }
    while (true) {
      if (s >= nsteps) break;
      step(v0, v1);
      if (++s >= nsteps) break;
      step(v1, v0);
      ++s;
    }
  } else if (FAST && a.cube_bytes && in_map(a, r)) {
    // Byte walk over a buffer resource, one voxel ahead: the load of the next
    // voxel is in flight while this step's cascade runs.  After the last step
    // the walk may name a voxel outside the cube; the resource's range check
    // returns 0 for it and the value is never used.
    const __amdgpu_buffer_rsrc_t rs =
        __builtin_amdgcn_make_buffer_rsrc((void *)(a.cube + toff), (short)0, (int)a.cube_bytes,
                                          kBufferWord3);
    Walk<FAST> w, wm;
    walk_init<FAST>(r, a.n, a.cs, w, (int)sizeof(T));
    if (VOX) walk_init<FAST>(r, a.n, a.vs, wm);
    // Unrolled by two so that the two values alternate registers: no copy of a
    // value whose load is still in flight.  Every step lowers at least one
    // counter (the ray state is finite), so n0 + n1 + n2 steps end any ray;
    // the wave-uniform pair counter is the guard against a stalled walk.
    int pairs = (int)((a.n[0] + a.n[1] + a.n[2]) >> 1) + 2;
    T v0 = buffer_load<T>(rs, w.off()), v1 = (T)0;
    while (true) {
      if (!(walk_inside(w) && (!SUN || outside_sun(h.M, u0, r.ac)))) break;
      {
        T cur = v0;
        bool keep = true;
        if (VOX) {
          const auto lm = walk_off(wm, w);
          if (a.voxw) cur = a.voxw[lm] * cur;
          if (a.voxmask) keep = (a.voxmask[lm] == 0);
        }
        const T d = step_walk<T, SUN, FAST>(r, w);
        v1 = buffer_load<T>(rs, w.off());
        if (keep) acc = acc + d * cur;
      }
      if (!(walk_inside(w) && (!SUN || outside_sun(h.M, u0, r.ac)))) break;
      {
        T cur = v1;
        bool keep = true;
        if (VOX) {
          const auto lm = walk_off(wm, w);
          if (a.voxw) cur = a.voxw[lm] * cur;
          if (a.voxmask) keep = (a.voxmask[lm] == 0);
        }
        const T d = step_walk<T, SUN, FAST>(r, w);
        v0 = buffer_load<T>(rs, w.off());
        if (keep) acc = acc + d * cur;
      }
      if (--pairs <= 0) {
        if (walk_inside(w)) flag_status(a.status, TG_EGEOMETRY);
        break;
      }
    }
  } else if (in_map(a, r)) {
    const T *cube = a.cube + toff;
    Walk<FAST> w, wm;
    walk_init<FAST>(r, a.n, a.cs, w);
    if (VOX) walk_init<FAST>(r, a.n, a.vs, wm);
    int guard = (int)(a.n[0] + a.n[1] + a.n[2]) + 2;
    while (walk_inside(w) && (!SUN || outside_sun(h.M, u0, r.ac))) {
      if (--guard < 0) {
        flag_status(a.status, TG_EGEOMETRY);
        break;
      }
      const auto lin = w.off();
      T v = cube[lin];
      bool keep = true;
      if (VOX) {
        const auto lm = walk_off(wm, w);
        if (a.voxw) v = a.voxw[lm] * v;
        if (a.voxmask) keep = (a.voxmask[lm] == 0);
      }
      const T d = step_walk<T, SUN, FAST>(r, w);
      if (keep) acc = acc + d * v;
    }
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

// Walks one ray and adds the reference's per-step product d * w (rounded in T,
// C:923-928) to dst[voxel]: dst is the cube itself, or the float64
// accumulator of TG_ACC_F64 (same element offsets; the cube is then dense).
template <typename T, bool SUN, bool VOX, bool FAST, typename A>
__device__ __forceinline__ void scatter_ray(const KArgs<T> &a, const ViewHdr<T> &h, const T u0[3],
                                            Ray<T> &r, T w, A *dst) {
  Walk<FAST> wk, wm;
  walk_init<FAST>(r, a.n, a.cs, wk);
  if (VOX) walk_init<FAST>(r, a.n, a.vs, wm);
  int guard = (int)(a.n[0] + a.n[1] + a.n[2]) + 2;
  while (walk_inside(wk) && (!SUN || outside_sun(h.M, u0, r.ac))) {
    if (--guard < 0) {
      flag_status(a.status, TG_EGEOMETRY);
      break;
    }
    const auto lin = wk.off();
    bool keep = true;
    if (VOX) {
      if (a.voxmask) keep = (a.voxmask[walk_off(wm, wk)] == 0);
    }
    const T d = step_walk<T, SUN, FAST>(r, wk);
    if (keep) atomic_add(dst + lin, (A)(d * w));
  }
}

// Backprojection (bpj / bpjt) as an atomic scatter.
template <typename T, bool SUN, bool VOX, bool FAST>
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
  if (!in_map(a, r)) return;
  if (a.acc)
    scatter_ray<T, SUN, VOX, FAST>(a, h, u0, r, w, a.acc + toff);
  else
    scatter_ray<T, SUN, VOX, FAST>(a, h, u0, r, w, a.cube + toff);
}

}  // namespace tg
