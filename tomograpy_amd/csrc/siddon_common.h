// Shared device-side pieces of the gfx950 Siddon kernels: the per-view
// descriptor layout, the launch arguments and the per-ray traversal state.
//
// Arithmetic contract: every expression below is evaluated in the element
// type T in the order tomograpy/C_siddon.c.template (cited "C:") evaluates it,
// and the library is compiled with -ffp-contract=off, so that a*b+c is two
// roundings as in the reference's x86-64 build.  Divisions are IEEE correctly
// rounded (hipcc's default for float and double).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "tomograpy_amd.h"

namespace tg {

// C:17 -- INF stands in for "no crossing" along an axis.
constexpr int kInf = 100000;

// Per-view header in the element type: the 18 values C:1077-1101 reads,
// rounded to T as get_dict_ctype (C:1051-1056) does.
template <typename T>
struct ViewHdr {
  T s[2], p[2], v[2];
  T R[9];
  T M[3];
};

// Byte layout of the packed view buffer (tg_views_pack):
//   [0, hdr_bytes)             nv x ViewHdr<T>
//   [hdr_bytes, +trig_bytes)   per view: n1 x (cos gamma_i, sin gamma_i) then
//                              n2 x (cos lambda_j, sin lambda_j), as doubles.
inline size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

template <typename T>
inline size_t views_hdr_bytes(int nv) {
  return align_up(sizeof(ViewHdr<T>) * (size_t)(nv > 0 ? nv : 1), 256);
}

inline size_t views_trig_bytes(int nv, int n1, int n2) {
  return (size_t)nv * ((size_t)n1 + (size_t)n2) * 2 * sizeof(double);
}

// Kernel arguments (passed by value in the kernarg segment).
template <typename T>
struct KArgs {
  T mp[3], md[3], mmin[3], mmax[3];  // CDELT, PSHAPE, MMIN, MMAX in T (C:50-56)
  uint32_t n[3];
  int32_t nt;
  int64_t cs[4];  // cube element strides
  int32_t n1, n2, nv;
  int32_t t0, t1;  // view range of this launch
  int32_t ray_i, ray_j;  // single-ray mode when ray_i >= 0
  int32_t bj_log2;       // log2(pixels per block row)
  int32_t tiles_i, tiles_j;  // tiles per view
  int32_t kind;
  uint32_t flags;  // TG_ZERO_OUT
  int64_t ds[3];  // data strides
  int64_t ms[3];  // pixel-mask strides
  int64_t vs[3];  // voxel mask / voxel weight strides
  const ViewHdr<T> *views;
  const double *trig;
  T *cube;
  T *data;
  double *acc;  // TG_ACC_F64: dense float64 accumulator of a backprojection, else NULL
  // bytes of one time bin of the cube when the forward projector may read it
  // through a buffer resource (non-negative strides, byte offsets < 2^31), else 0
  uint32_t cube_bytes;
  // nonzero when every data (and pixel-weight) offset fits 31 bits with
  // non-negative strides: pixel offsets may use 32-bit arithmetic
  uint32_t data32;
  const uint8_t *pixmask;
  const int32_t *tindex;
  const T *pixw;
  const T *voxw;
  const uint8_t *voxmask;
  int32_t *status;
  uint32_t *per_ray;
  unsigned long long *total;
  const uint32_t *ray_steps;  // per-ray step counts (tg_call.ray_steps) or NULL
};

// Traversal state of one ray after the setup of C:561-573.
template <typename T>
struct Ray {
  T D[3];
  T pabs[3];
  T ac;
  int iv[3];
  int up[3];
};

// C:969-977
template <typename T>
__device__ __forceinline__ T min3(T x, T y, T z) {
  return ((x < y) && (x < z)) ? x : (((y <= x) && (y < z)) ? y : z);
}
// C:979-987
template <typename T>
__device__ __forceinline__ T max3(T x, T y, T z) {
  return ((x > y) && (x > z)) ? x : (((y >= x) && (y > z)) ? y : z);
}
// C:989-997
template <typename T>
__device__ __forceinline__ int sgn(T x) {
  return (x > 0) ? 1 : ((x < 0) ? -1 : 0);
}

// (unsigned int)x as x86-64 gcc generates it for C:785-786: truncate to a
// 64-bit integer, keep the low 32 bits.
template <typename T>
__device__ __forceinline__ uint32_t to_u32_x86(T x) {
  return (uint32_t)(long long)x;
}

// Unit vector of a pixel: C:1013-1032.  cg/sg/cl/sl are cos/sin of gamma and
// lambda in double (libm values from the packed tables); the double products
// are rounded to T on assignment, the rotation is evaluated in T.
template <typename T>
__device__ __forceinline__ void unit_vector(const ViewHdr<T> &h, double cg, double sg, double cl,
                                            double sl, T u0[3]) {
  T u2[3];
  u2[0] = (T)(cl * cg);
  u2[1] = (T)(cl * sg);
  u2[2] = (T)sl;
  u0[0] = h.R[0] * u2[0] + h.R[1] * u2[1] + h.R[2] * u2[2];
  u0[1] = h.R[3] * u2[0] + h.R[4] * u2[1] + h.R[5] * u2[2];
  u0[2] = h.R[6] * u2[0] + h.R[7] * u2[1] + h.R[8] * u2[2];
}

// C:723-735
template <typename T>
__device__ __forceinline__ void intersection_parameters(const KArgs<T> &a, const T M[3],
                                                        const T u0[3], T a1[3], T an[3]) {
#pragma unroll
  for (int k = 0; k < 3; k++) {
    if (u0[k] != 0) {
      a1[k] = (a.mmin[k] - M[k]) / u0[k];
      an[k] = (a.mmax[k] - M[k]) / u0[k];
    } else {
      a1[k] = (T)(-kInf);
      an[k] = (T)kInf;
    }
  }
}

// C:561-573 (with C:737-809).  Returns false when the ray misses the map
// (amin >= amax, C:564).
template <typename T>
__device__ __forceinline__ bool ray_setup(const KArgs<T> &a, const T M[3], const T u0[3],
                                          Ray<T> &r) {
  T a1[3], an[3], Imin[3], Imax[3], p[3];
  intersection_parameters(a, M, u0, a1, an);
#pragma unroll
  for (int k = 0; k < 3; k++) {  // C:999-1011
    const bool sw = a1[k] > an[k];
    Imin[k] = sw ? an[k] : a1[k];
    Imax[k] = sw ? a1[k] : an[k];
  }
  const T amin = max3(Imin[0], Imin[1], Imin[2]);
  const T amax = min3(Imax[0], Imax[1], Imax[2]);
  if (!(amin < amax)) return false;
#pragma unroll
  for (int k = 0; k < 3; k++) {  // C:750-774
    p[k] = (u0[k] != 0) ? a.mp[k] / u0[k] : (T)kInf;
    r.pabs[k] = fabs(p[k]);
    r.up[k] = sgn(u0[k]);
  }
#pragma unroll
  for (int k = 0; k < 3; k++) {  // C:776-788
    const T e = M[k] + amin * u0[k];
    const T ep = e - a.mmin[k];
    uint32_t iv = to_u32_x86(ep / a.mp[k]);
    iv -= to_u32_x86(ep / a.md[k]);
    r.iv[k] = (int)iv;
  }
#pragma unroll
  for (int k = 0; k < 3; k++) {  // C:790-809
    uint32_t next;
    if (r.up[k] == 1)
      next = (uint32_t)r.iv[k] + 1u;
    else if (r.up[k] == -1)
      next = (uint32_t)r.iv[k];
    else
      next = (uint32_t)kInf * a.n[k];
    r.D[k] = ((T)next) * p[k] + a1[k] - amin;
  }
  r.ac = amin;
  return true;
}

// True when the ray state can make progress: the reference's cascade
// (C:821-911) matches no branch when a distance is NaN and would loop forever.
template <typename T>
__device__ __forceinline__ bool ray_state_ok(const Ray<T> &r) {
  return (r.D[0] == r.D[0]) && (r.D[1] == r.D[1]) && (r.D[2] == r.D[2]) &&
         (r.pabs[0] == r.pabs[0]) && (r.pabs[1] == r.pabs[1]) && (r.pabs[2] == r.pabs[2]);
}

// C:948-953 (iv compared as unsigned against n, which also rejects iv < 0).
template <typename T>
__device__ __forceinline__ bool in_map(const KArgs<T> &a, const Ray<T> &r) {
  return ((uint32_t)r.iv[0] < a.n[0]) & ((uint32_t)r.iv[1] < a.n[1]) &
         ((uint32_t)r.iv[2] < a.n[2]);
}

// C:956-961 with C:1034-1041: true while the entry point of the current voxel
// is strictly outside the unit sphere.
template <typename T>
__device__ __forceinline__ bool outside_sun(const T M[3], const T u0[3], T ac) {
  T x0 = M[0] + ac * u0[0];
  T x1 = M[1] + ac * u0[1];
  T x2 = M[2] + ac * u0[2];
  T d = x0 * x0;
  d += x1 * x1;
  d += x2 * x2;
  return d > 1;
}

// One step of C:811-912, branch-free.  For non-NaN distances the cascade's
// seven branches are exactly "the axes whose D equals the minimum step
// together"; the length used is D of the first such axis in the order x, y, z
// (the value the cascade uses, which only matters for the sign of a zero).
// Returns the length; sx/sy/sz tell which axes stepped.
template <typename T>
__device__ __forceinline__ T step(Ray<T> &r, bool &sx, bool &sy, bool &sz) {
  const T m01 = r.D[0] < r.D[1] ? r.D[0] : r.D[1];
  const T dmin = m01 < r.D[2] ? m01 : r.D[2];
  sx = (r.D[0] == dmin);
  sy = (r.D[1] == dmin);
  sz = (r.D[2] == dmin);
  const T d = sx ? r.D[0] : (sy ? r.D[1] : r.D[2]);
  r.ac += d;
  r.D[0] = sx ? r.pabs[0] : r.D[0] - d;
  r.D[1] = sy ? r.pabs[1] : r.D[1] - d;
  r.D[2] = sz ? r.pabs[2] : r.D[2] - d;
  r.iv[0] += sx ? r.up[0] : 0;
  r.iv[1] += sy ? r.up[1] : 0;
  r.iv[2] += sz ? r.up[2] : 0;
  return d;
}

// ---------------------------------------------------------------------------
// Counter form of the traversal used by the hot loops.
//
// For a ray that starts inside the map, r_k counts the voxels left along axis
// k before the ray leaves [0, n_k): n_k - iv_k when stepping up, iv_k + 1 when
// stepping down, a large constant when the axis never steps (update 0, the
// cascade can still pick it; iv_k does not move then).  Each step along k
// decrements r_k, and the ray is inside exactly when every r_k > 0 -- the
// in_map test of C:948-953.  The element offset of the current voxel is
// base + sum_k r_k m_k with m_k = -update_k * stride_k, evaluated with 24-bit
// multiply-adds in 32-bit modular arithmetic (FAST) or in 64 bits.
// ---------------------------------------------------------------------------
constexpr int kNoStep = 1 << 22;

__device__ __forceinline__ float tmin(float a, float b) { return fminf(a, b); }
// The smaller of two doubles by one v_min_f64.  fmin would first quiet each
// operand (a v_max_f64 x, x, x apiece) to give a signalling NaN fmin's
// semantics; the traversal distances are never NaN (ray_state_ok rejects such
// rays before their walk), and for all other inputs the results are the same.
__device__ __forceinline__ double tmin(double a, double b) {
  double r;
  asm("v_min_f64 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
  return r;
}

// A value no larger than x on any active lane of the wavefront, found from the
// candidate c: while some active lane holds less, c moves to the value of the
// first such lane.  Every round lowers c, and it ends at the active lanes'
// minimum when c starts at an active lane's value (or at c, when c is already
// no larger).  c is wavefront-uniform, so loops bounded by the result run on
// a scalar counter with every lane active.
__device__ __forceinline__ uint32_t wave_floor(uint32_t x, uint32_t c) {
  for (;;) {
    const unsigned long long below = __ballot(x < c);
    if (below == 0ull) return c;
    c = (uint32_t)__builtin_amdgcn_readlane((int)x, (int)__builtin_ctzll(below));
  }
}

// step() with the minimum taken by one min3, as step_walk takes it: the same
// axes, distances and indices for the finite distances of a traversal (the
// value of the minimum is the length the cascade picks, up to the sign of a
// zero, and no distance is -0).
template <typename T>
__device__ __forceinline__ T step_min(Ray<T> &r, bool &sx, bool &sy, bool &sz) {
  const T d = tmin(tmin(r.D[0], r.D[1]), r.D[2]);
  sx = (r.D[0] == d);
  sy = (r.D[1] == d);
  sz = (r.D[2] == d);
  r.ac += d;
  r.D[0] = sx ? r.pabs[0] : r.D[0] - d;
  r.D[1] = sy ? r.pabs[1] : r.D[1] - d;
  r.D[2] = sz ? r.pabs[2] : r.D[2] - d;
  r.iv[0] += sx ? r.up[0] : 0;
  r.iv[1] += sy ? r.up[1] : 0;
  r.iv[2] += sz ? r.up[2] : 0;
  return d;
}

template <bool FAST>
struct Walk;

// 32-bit walk: r[k] = voxel crossings left on axis k (the ray is inside
// while every counter is positive); offset = base + sum r[k] m[k] with
// 24-bit multiplies.
template <>
struct Walk<true> {
  int r[3];
  int m[3];
  uint32_t base;
  __device__ __forceinline__ uint32_t off() const {
    return base + (uint32_t)__mul24(r[0], m[0]) + (uint32_t)__mul24(r[1], m[1]) +
           (uint32_t)__mul24(r[2], m[2]);
  }
};

template <>
struct Walk<false> {
  int r[3];
  int64_t m[3];
  int64_t base;
  __device__ __forceinline__ int64_t off() const {
    return base + r[0] * m[0] + r[1] * m[1] + r[2] * m[2];
  }
};

// Counters of a ray whose voxel iv is inside [0, n) (check in_map first), for
// an array with element strides st.
// Sets a 32-bit walk from the crossings left per axis (each >= 1), the byte or
// element multipliers and the offset of the current voxel.
__device__ __forceinline__ void walk_set(Walk<true> &w, const int cnt[3], const int m[3],
                                         uint32_t lin) {
#pragma unroll
  for (int k = 0; k < 3; k++) {
    w.r[k] = cnt[k];
    w.m[k] = m[k];
  }
  w.base = lin - ((uint32_t)(w.r[0] * w.m[0]) + (uint32_t)(w.r[1] * w.m[1]) +
                  (uint32_t)(w.r[2] * w.m[2]));
}

template <bool FAST, typename T, typename S>
__device__ __forceinline__ void walk_init(const Ray<T> &r, const uint32_t n[3], const S st[3],
                                          Walk<FAST> &w, int scale = 1) {
  if constexpr (FAST) {
    int cnt[3], m[3];
    uint32_t lin = 0;
#pragma unroll
    for (int k = 0; k < 3; k++) {
      cnt[k] = r.up[k] > 0 ? (int)n[k] - r.iv[k] : (r.up[k] < 0 ? r.iv[k] + 1 : kNoStep);
      m[k] = -(r.up[k] * (int)st[k] * scale);
      lin += (uint32_t)r.iv[k] * (uint32_t)(st[k] * scale);
    }
    walk_set(w, cnt, m, lin);
  } else {
#pragma unroll
    for (int k = 0; k < 3; k++) {
      w.r[k] = r.up[k] > 0 ? (int)n[k] - r.iv[k] : (r.up[k] < 0 ? r.iv[k] + 1 : kNoStep);
      w.m[k] = -(r.up[k] * st[k] * scale);
    }
    int64_t lin = 0;
#pragma unroll
    for (int k = 0; k < 3; k++) lin += (int64_t)r.iv[k] * (int64_t)st[k] * scale;
    w.base = lin - ((int64_t)w.r[0] * w.m[0] + (int64_t)w.r[1] * w.m[1] + (int64_t)w.r[2] * w.m[2]);
  }
}

template <bool FAST>
__device__ __forceinline__ bool walk_inside(const Walk<FAST> &w) {
  return min(min(w.r[0], w.r[1]), w.r[2]) > 0;
}

// Offset in the array of walk `g` (its multipliers and base) at the counters of
// walk `c`: a second array walked along the same ray (same shape, other
// strides) has the same counters, so only one set is stepped.
__device__ __forceinline__ uint32_t walk_off(const Walk<true> &g, const Walk<true> &c) {
  return g.base + (uint32_t)__mul24(c.r[0], g.m[0]) + (uint32_t)__mul24(c.r[1], g.m[1]) +
         (uint32_t)__mul24(c.r[2], g.m[2]);
}

__device__ __forceinline__ int64_t walk_off(const Walk<false> &g, const Walk<false> &c) {
  return g.base + c.r[0] * g.m[0] + c.r[1] * g.m[1] + c.r[2] * g.m[2];
}

// One step of C:811-912 on (D, ac) with the walk's counters instead of iv
// (a voxel-operand walk shares them, see walk_off).  Same floating-point operations as
// step(): the minimum is taken with min3 (exact: no NaN reaches here).  The
// reference's length is D of the first axis holding the minimum; the minimum
// itself is that value, up to the sign of a zero (-0 == +0), and a zero
// length adds a zero product that leaves any nonzero sum unchanged.
template <typename T, bool AC, bool FAST>
__device__ __forceinline__ T step_walk(Ray<T> &r, Walk<FAST> &w) {
  const T d = tmin(tmin(r.D[0], r.D[1]), r.D[2]);
  const bool sx = (r.D[0] == d);
  const bool sy = (r.D[1] == d);
  const bool sz = (r.D[2] == d);
  if (AC) r.ac += d;
  r.D[0] = sx ? r.pabs[0] : r.D[0] - d;
  r.D[1] = sy ? r.pabs[1] : r.D[1] - d;
  r.D[2] = sz ? r.pabs[2] : r.D[2] - d;
  w.r[0] -= sx;
  w.r[1] -= sy;
  w.r[2] -= sz;
  return d;
}

// Whether the forward projector may walk the cube in bytes through a buffer
// resource: non-negative strides, every stride in bytes < 2^23 (24-bit
// multiplies) and the 3-D extent in bytes < 2^31.  Returns the extent in
// bytes, or 0.
inline uint32_t buffer_walk_bytes(const uint32_t n[3], const int64_t st[3], int elem) {
  int64_t ext = 1;
  for (int k = 0; k < 3; k++) {
    if (st[k] < 0 || st[k] * elem >= (1 << 23)) return 0;
    ext += (int64_t)(n[k] - 1) * st[k];
  }
  ext *= elem;
  return ext < (1ll << 31) ? (uint32_t)ext : 0u;
}

// Whether the 32-bit walk applies: every |stride| < 2^23 (24-bit multiplies)
// and every offset of the 3-D part below 2^31.
inline bool walk_fast_ok(const uint32_t n[3], const int64_t st[3]) {
  int64_t ext = 1;
  for (int k = 0; k < 3; k++) {
    const int64_t a = st[k] < 0 ? -st[k] : st[k];
    if (a >= (1 << 23)) return false;
    ext += (int64_t)(n[k] - 1) * a;
  }
  return ext < (1ll << 31);
}

// Workgroup -> logical tile mapping that gives each XCD a contiguous range of
// logical tiles (workgroups are dispatched round-robin over the 8 XCDs), so
// one XCD's L2 serves neighbouring rays of the same views.
__device__ __forceinline__ uint32_t xcd_remap(uint32_t bid, uint32_t nb) {
  const uint32_t q = nb >> 3, rem = nb & 7u;
  const uint32_t xcd = bid & 7u, k = bid >> 3;
  return (xcd < rem) ? xcd * (q + 1) + k : rem * (q + 1) + (xcd - rem) * q + k;
}

// Resolves the pixel (i, j) and view t of this lane; false if out of range.
template <typename T>
__device__ __forceinline__ bool lane_pixel(const KArgs<T> &a, uint32_t tile, int &i, int &j,
                                           int &t) {
  if (a.ray_i >= 0) {
    i = a.ray_i;
    j = a.ray_j;
    t = a.t0;
    return threadIdx.x == 0;
  }
  const uint32_t tiles_per_view = (uint32_t)a.tiles_i * (uint32_t)a.tiles_j;
  t = a.t0 + (int)(tile / tiles_per_view);
  const uint32_t tv = tile % tiles_per_view;
  const int ti = (int)(tv / (uint32_t)a.tiles_j);
  const int tj = (int)(tv % (uint32_t)a.tiles_j);
  const int bi = (int)(blockDim.x >> a.bj_log2);
  i = ti * bi + (int)(threadIdx.x >> a.bj_log2);
  j = (tj << a.bj_log2) + (int)(threadIdx.x & ((1u << a.bj_log2) - 1u));
  return (i < a.n1) && (j < a.n2) && (t < a.t1);
}

template <typename T>
__device__ __forceinline__ void pixel_unit_vector(const KArgs<T> &a, const ViewHdr<T> &h, int t,
                                                  int i, int j, T u0[3]) {
  const double *tr = a.trig + (size_t)t * ((size_t)a.n1 + (size_t)a.n2) * 2;
  const double cg = tr[2 * i], sg = tr[2 * i + 1];
  const double cl = tr[2 * ((size_t)a.n1 + j)], sl = tr[2 * ((size_t)a.n1 + j) + 1];
  unit_vector(h, cg, sg, cl, sl, u0);
}

__device__ __forceinline__ void flag_status(int32_t *status, int32_t code) {
  if (status) *status = code;
}

// Offset of the time bin of view t in a 4-D map (C:931-944: the bin is the
// image index t, or tindex[t] when a view -> bin map is given).  A bin
// outside [0, nt) flags TG_EINVAL and the ray is skipped.
template <typename T>
__device__ __forceinline__ bool time_offset(const KArgs<T> &a, int t, int64_t &toff) {
  toff = 0;
  if (a.kind != TG_PJT && a.kind != TG_BPJT) return true;
  const int b = a.tindex ? a.tindex[t] : t;
  if (b < 0 || b >= a.nt) {
    flag_status(a.status, TG_EINVAL);
    return false;
  }
  toff = (int64_t)b * a.cs[3];
  return true;
}

}  // namespace tg
