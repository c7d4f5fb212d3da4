// Host-side helpers shared by the translation units of the library.
#pragma once

#include <hip/hip_runtime.h>

#include "siddon_common.h"

namespace tg {

// Records a thread-local error message and returns `code`.
int set_error(int code, const char *fmt, ...);
// Maps a HIP error to TG_EHIP with a message; TG_OK on success.
int hip_check(hipError_t e, const char *what);

// TG_ACC_F64: adds the float64 accumulator a.acc to the cube and clears it
// (no-op when a.acc is NULL); enqueued after a backprojection's kernels.
template <typename T>
int finish_acc(const KArgs<T> &a, hipStream_t s);

// finish_acc for every job of a batch in one launch: the jobs' KArgs<T> sit at
// the start of `stride`-byte entries of the device table `dev_table`; host_args
// are the same jobs' arguments on the host (for the launch size).
template <typename T>
int finish_acc_table(const void *dev_table, size_t stride, const KArgs<T> *const *host_args,
                     int njobs, hipStream_t s);

// The float64 accumulators of a batch's jobs are finished by one launch that
// runs every job's finish at once, each reading and clearing its accumulator
// and adding it to its cube without atomics.  Two jobs may therefore share
// neither an accumulator nor a cube.  Returns TG_EINVAL (with the job indices)
// when the ranges of two jobs' accumulators, or of their cubes, overlap.
template <typename T>
int check_acc_jobs(const KArgs<T> *const *args, int njobs) {
  for (int x = 0; x < njobs; x++) {
    const KArgs<T> &a = *args[x];
    if (!a.acc) continue;
    const int64_t na = (int64_t)a.n[0] * a.n[1] * a.n[2] * a.nt;
    const char *a0 = (const char *)a.acc, *a1 = a0 + na * (int64_t)sizeof(double);
    const char *c0 = (const char *)a.cube, *c1 = c0 + na * (int64_t)sizeof(T);
    for (int y = x + 1; y < njobs; y++) {
      const KArgs<T> &b = *args[y];
      if (!b.acc) continue;
      const int64_t nb = (int64_t)b.n[0] * b.n[1] * b.n[2] * b.nt;
      const char *b0 = (const char *)b.acc, *b1 = b0 + nb * (int64_t)sizeof(double);
      const char *d0 = (const char *)b.cube, *d1 = d0 + nb * (int64_t)sizeof(T);
      if (a0 < b1 && b0 < a1)
        return set_error(TG_EINVAL, "jobs %d and %d share a float64 accumulator", x, y);
      if (c0 < d1 && d0 < c1)
        return set_error(TG_EINVAL, "jobs %d and %d accumulate into the same cube", x, y);
    }
  }
  return TG_OK;
}

// Projects or backprojects views [t0, t1) of *c (ray_i >= 0: the one ray
// (ray_i, ray_j)); the body of tg_conic_full_projector and its relatives
// (siddon.hip).
int dispatch_siddon(const tg_call *c, int t0, int t1, int ray_i, int ray_j, void *stream);

// Deterministic backprojection through a canonical plan built in the call's
// workspace (siddon_plan.hip); kDetPlanUnfit when the geometry or workspace
// does not allow it, and the chunked path below is taken instead.
constexpr int kDetPlanUnfit = 1;
template <typename T>
int run_det_plan(const tg_call *c, hipStream_t s);
size_t det_plan_workspace_bytes(const tg_call *c);

// Deterministic backprojection (siddon_det.hip).
template <typename T>
int run_deterministic(const tg_call *c, const KArgs<T> &a, hipStream_t s);

// An HSA dispatch holds its grid size in work-items as 32 bits, so a launch
// of 256-lane workgroups has at most this many workgroups.
constexpr uint64_t kMaxGroups256 = 0xffffffffull / 256ull;

// Calls launch(view_args, groups) for consecutive view ranges of [a.t0, a.t1)
// whose ray-tile grids (256-lane workgroups, tiles_i x tiles_j per view) each
// fit one dispatch; view_args is `a` with t0/t1 narrowed to the range.
template <typename T, typename F>
int for_view_launches(const KArgs<T> &a, F &&launch) {
  const uint64_t per_view = (uint64_t)a.tiles_i * (uint64_t)a.tiles_j;
  if (per_view == 0 || a.t1 <= a.t0) return TG_OK;
  if (per_view > kMaxGroups256)
    return set_error(TG_EINVAL, "a %d x %d image needs %llu workgroups, more than one launch holds",
                     a.n1, a.n2, (unsigned long long)per_view);
  const uint64_t fit = kMaxGroups256 / per_view;
  const int step = fit < (uint64_t)(a.t1 - a.t0) ? (int)fit : a.t1 - a.t0;
  for (int v = a.t0; v < a.t1; v += step) {
    KArgs<T> av = a;
    av.t0 = v;
    av.t1 = a.t1 - v < step ? a.t1 : v + step;
    launch(av, (uint32_t)((uint64_t)(av.t1 - av.t0) * per_view));
    if (a.t1 - v <= step) break;
  }
  return TG_OK;
}

// Storage extent (max element offset + 1) a call can touch in the cube.
inline int64_t cube_extent(const tg_call *c) {
  const bool time = (c->kind == TG_PJT || c->kind == TG_BPJT);
  int64_t e = 1;
  for (int k = 0; k < 3; k++) e += (int64_t)(c->map.n[k] - 1) * c->map.stride[k];
  if (time) e += (int64_t)((c->map.nt > 0 ? c->map.nt : 1) - 1) * c->map.stride[3];
  return e;
}

}  // namespace tg
