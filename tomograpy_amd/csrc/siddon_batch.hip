// Many uneven projector (or backprojector) jobs in one persistent launch.
//
// The host lays the jobs out in queue order (the schedule) in a device job
// table; every job contributes its ray tiles (workgroup-sized pixel tiles of
// its views) to one global tile range.  Persistent workgroups take the next
// tile from an atomic counter until the range is drained, find the owning job
// by binary search over the jobs' first-tile offsets, and run the same
// per-ray body as the one-job kernels, so each job's results are those of a
// one-job launch (bit-identical for the forward projector).
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <string.h>

#include <algorithm>
#include <numeric>
#include <string>
#include <vector>

#include "siddon_common.h"
#include "siddon_host.h"
#include "siddon_rays.h"

namespace tg {

template <typename T>
int make_args(const tg_call *c, int t0, int t1, KArgs<T> &a, int tile_j = 0);

namespace {

constexpr size_t kQueueHeader = 256;

template <typename T>
struct JobDesc {
  KArgs<T> a;
  unsigned long long tile_begin;
  uint32_t sun;
  uint32_t vox;
  uint32_t fast;
  uint32_t pad;
};

size_t job_bytes() {
  return align_up(sizeof(JobDesc<double>) > sizeof(JobDesc<float>) ? sizeof(JobDesc<double>)
                                                                    : sizeof(JobDesc<float>),
                  16);
}

// One persistent kernel per variant (obstacle, voxel operand, walk kind): a
// single kernel dispatching all eight at run time would be register-allocated
// for the largest (172 VGPRs in fp32, 2 waves per SIMD), while each variant
// alone needs 26-62 and runs at full occupancy.  The host groups the jobs by
// variant and launches one kernel per non-empty group.
template <typename T, bool BWD, bool SUN, bool VOX, bool FAST>
__global__ __launch_bounds__(256) void k_batch(const JobDesc<T> *__restrict__ jobs, int njobs,
                                               unsigned long long total,
                                               unsigned long long *head) {
  __shared__ unsigned long long s_tile;
  while (true) {
    if (threadIdx.x == 0) s_tile = atomicAdd(head, 1ull);
    __syncthreads();
    const unsigned long long tile = s_tile;
    __syncthreads();
    if (tile >= total) break;  // every workgroup reaches this once the range is drained
    int lo = 0, hi = njobs - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (jobs[mid].tile_begin <= tile) lo = mid;
      else hi = mid - 1;
    }
    const JobDesc<T> &jd = jobs[lo];
    const uint32_t t = (uint32_t)(tile - jd.tile_begin);
    if (BWD) backproject_tile<T, SUN, VOX, FAST>(jd.a, t);
    else project_tile<T, SUN, VOX, FAST>(jd.a, t);
  }
}

// Header words of the device queue (64-bit): [1] tile total, [2] job count,
// [kHeadWord + v] the tile counter of variant group v.
constexpr int kHeadWord = 8;
constexpr int kVariants = 8;

__global__ void k_reset_heads(unsigned long long *hdr) {
  if (threadIdx.x < kVariants && blockIdx.x == 0) hdr[kHeadWord + threadIdx.x] = 0ull;
}

template <typename T, bool BWD, bool SUN, bool VOX>
void launch_group_f(bool fast, int blocks, hipStream_t s, const JobDesc<T> *jobs, int n,
                    unsigned long long total, unsigned long long *head) {
  if (fast)
    hipLaunchKernelGGL((k_batch<T, BWD, SUN, VOX, true>), dim3(blocks), dim3(256), 0, s, jobs, n,
                       total, head);
  else
    hipLaunchKernelGGL((k_batch<T, BWD, SUN, VOX, false>), dim3(blocks), dim3(256), 0, s, jobs, n,
                       total, head);
}

template <typename T, bool BWD>
void launch_group(int v, int blocks, hipStream_t s, const JobDesc<T> *jobs, int n,
                  unsigned long long total, unsigned long long *head) {
  const bool sun = v & 1, vox = v & 2, fast = v & 4;
  if (sun) {
    if (vox) launch_group_f<T, BWD, true, true>(fast, blocks, s, jobs, n, total, head);
    else launch_group_f<T, BWD, true, false>(fast, blocks, s, jobs, n, total, head);
  } else {
    if (vox) launch_group_f<T, BWD, false, true>(fast, blocks, s, jobs, n, total, head);
    else launch_group_f<T, BWD, false, false>(fast, blocks, s, jobs, n, total, head);
  }
}

template <typename T>
int run_batch(const tg_call *calls, int njobs, const tg_batch_plan *plan, void *queue,
              void *host_staging, hipStream_t s) {
  const int kind0 = calls[0].kind;
  const bool bwd = (kind0 == TG_BPJ || kind0 == TG_BPJT);
  std::vector<KArgs<T>> args(njobs);
  std::vector<double> cost(njobs);
  std::vector<unsigned long long> tiles(njobs);
  for (int k = 0; k < njobs; k++) {
    const tg_call &c = calls[k];
    if (c.dtype != calls[0].dtype) return set_error(TG_EINVAL, "batch jobs must share a dtype");
    const bool kb = (c.kind == TG_BPJ || c.kind == TG_BPJT);
    if (kb != bwd) return set_error(TG_EINVAL, "batch jobs must all project or all backproject");
    if (c.flags & TG_DETERMINISTIC)
      return set_error(TG_EINVAL, "deterministic backprojection is not batched");
    if (!c.cube || !c.data) return set_error(TG_EINVAL, "job %d: cube and data are required", k);
    tg_call cc = c;
    if (plan && plan->tile_j > 0) cc.tile_j = plan->tile_j;
    int rc = make_args<T>(&cc, 0, cc.nv, args[k]);
    if (rc) {
      const std::string msg = tg_last_error();
      return set_error(rc, "job %d: %s", k, msg.c_str());
    }
    tiles[k] = (unsigned long long)cc.nv * args[k].tiles_i * args[k].tiles_j;
    // longest rays first: a ray crosses at most n0 + n1 + n2 voxels
    cost[k] = (double)(cc.map.n[0] + cc.map.n[1] + cc.map.n[2]);
  }
  if (bwd) {
    std::vector<const KArgs<T> *> pa(njobs);
    for (int k = 0; k < njobs; k++) pa[k] = &args[k];
    int rc = check_acc_jobs<T>(pa.data(), njobs);
    if (rc) return rc;
  }
  // variant of each job: bit 0 obstacle, bit 1 voxel operand, bit 2 32-bit walk
  std::vector<int> variant(njobs);
  for (int k = 0; k < njobs; k++) {
    const uint32_t sun = (calls[k].flags & TG_OBSTACLE_SUN) ? 1u : 0u;
    const uint32_t vox =
        (calls[k].voxmask != nullptr || (!bwd && calls[k].vox_weight != nullptr)) ? 1u : 0u;
    const uint32_t fast =
        (walk_fast_ok(args[k].n, args[k].cs) && (!vox || walk_fast_ok(args[k].n, args[k].vs)))
            ? 1u : 0u;
    variant[k] = (int)(sun | (vox << 1) | (fast << 2));
  }
  // queue order: grouped by variant; inside a group, arrival order or longest rays first
  std::vector<int> order(njobs);
  std::iota(order.begin(), order.end(), 0);
  const bool longest_first = plan && plan->schedule == 1;
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
    if (variant[x] != variant[y]) return variant[x] < variant[y];
    return longest_first && cost[x] > cost[y];
  });
  char *stage = (char *)host_staging;
  memset(stage, 0, kQueueHeader);
  // The device reads the table as JobDesc<T>[njobs]: pack it with that stride.
  // tile_begin counts from the start of the job's variant group.
  JobDesc<T> *table = (JobDesc<T> *)(stage + kQueueHeader);
  int gfirst[kVariants], gcount[kVariants];
  unsigned long long gtiles[kVariants];
  for (int v = 0; v < kVariants; v++) {
    gfirst[v] = 0;
    gcount[v] = 0;
    gtiles[v] = 0;
  }
  unsigned long long total = 0;
  for (int q = 0; q < njobs; q++) {
    const int k = order[q];
    const int v = variant[k];
    if (gcount[v] == 0) gfirst[v] = q;
    JobDesc<T> jd;
    memset(&jd, 0, sizeof(jd));
    jd.a = args[k];
    jd.tile_begin = gtiles[v];
    jd.sun = (uint32_t)(v & 1);
    jd.vox = (uint32_t)((v >> 1) & 1);
    jd.fast = (uint32_t)((v >> 2) & 1);
    memcpy(&table[q], &jd, sizeof(jd));
    gcount[v]++;
    gtiles[v] += tiles[k];
    total += tiles[k];
  }
  const size_t bytes = kQueueHeader + (size_t)njobs * sizeof(JobDesc<T>);
  // the header records the tile total and the job count (read by tg_batch_plan_table)
  ((unsigned long long *)stage)[1] = total;
  ((unsigned long long *)stage)[2] = (unsigned long long)njobs;
  if (!queue) return TG_OK;  // plan only: the table is in host_staging
  int rc;
  if ((rc = hip_check(hipMemcpyAsync(queue, stage, bytes, hipMemcpyHostToDevice, s),
                      "batch job table copy")))
    return rc;
  if (total == 0) return TG_OK;
  unsigned long long *hdr = (unsigned long long *)queue;
  const JobDesc<T> *jobs = (const JobDesc<T> *)((char *)queue + kQueueHeader);
  int blocks = plan ? plan->blocks : 0;
  if (blocks <= 0) {
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        cus <= 0)
      cus = 256;
    blocks = cus * 8;
  }
  hipLaunchKernelGGL(k_reset_heads, dim3(1), dim3(64), 0, s, hdr);
  for (int v = 0; v < kVariants; v++) {
    if (gcount[v] == 0 || gtiles[v] == 0) continue;
    const int b = (unsigned long long)blocks > gtiles[v] ? (int)gtiles[v] : blocks;
    if (bwd)
      launch_group<T, true>(v, b, s, jobs + gfirst[v], gcount[v], gtiles[v], hdr + kHeadWord + v);
    else
      launch_group<T, false>(v, b, s, jobs + gfirst[v], gcount[v], gtiles[v], hdr + kHeadWord + v);
  }
  if ((rc = hip_check(hipGetLastError(), "batch kernel launch"))) return rc;
  if (bwd) {
    // the jobs' float64 accumulators, finished in one launch over the device table
    static_assert(offsetof(JobDesc<T>, a) == 0, "finish_acc_table reads KArgs at entry start");
    std::vector<const KArgs<T> *> hargs(njobs);
    for (int k = 0; k < njobs; k++) hargs[k] = &args[k];
    if ((rc = finish_acc_table<T>(jobs, sizeof(JobDesc<T>), hargs.data(), njobs, s))) return rc;
  }
  return TG_OK;
}

}  // namespace
}  // namespace tg

using namespace tg;

extern "C" {

size_t tg_batch_queue_bytes(int32_t njobs) {
  if (njobs < 0) return 0;
  return kQueueHeader + (size_t)njobs * job_bytes();
}

int tg_batch_run(const tg_call *calls, int32_t njobs, const tg_batch_plan *plan, void *queue,
                 void *host_staging, void *stream) {
  if (njobs <= 0 || !calls) return set_error(TG_EINVAL, "empty batch");
  if (!queue || !host_staging) return set_error(TG_EINVAL, "batch needs queue and staging buffers");
  hipStream_t s = (hipStream_t)stream;
  if (calls[0].dtype == TG_F32) return run_batch<float>(calls, njobs, plan, queue, host_staging, s);
  if (calls[0].dtype == TG_F64) return run_batch<double>(calls, njobs, plan, queue, host_staging, s);
  return set_error(TG_EINVAL, "bad dtype %d", calls[0].dtype);
}

int tg_batch_plan_table(const tg_call *calls, int32_t njobs, const tg_batch_plan *plan,
                        void *host_staging) {
  if (njobs <= 0 || !calls) return set_error(TG_EINVAL, "empty batch");
  if (!host_staging) return set_error(TG_EINVAL, "batch needs a staging buffer");
  if (calls[0].dtype == TG_F32) return run_batch<float>(calls, njobs, plan, nullptr, host_staging, 0);
  if (calls[0].dtype == TG_F64) return run_batch<double>(calls, njobs, plan, nullptr, host_staging, 0);
  return set_error(TG_EINVAL, "bad dtype %d", calls[0].dtype);
}

int tg_batch_job_layout(int32_t dtype, size_t *job_bytes_out, size_t *tile_begin_offset,
                        size_t *cube_offset) {
  if (dtype == TG_F32) {
    *job_bytes_out = sizeof(JobDesc<float>);
    *tile_begin_offset = offsetof(JobDesc<float>, tile_begin);
    *cube_offset = offsetof(KArgs<float>, cube);
    return TG_OK;
  }
  if (dtype == TG_F64) {
    *job_bytes_out = sizeof(JobDesc<double>);
    *tile_begin_offset = offsetof(JobDesc<double>, tile_begin);
    *cube_offset = offsetof(KArgs<double>, cube);
    return TG_OK;
  }
  return set_error(TG_EINVAL, "bad dtype %d", dtype);
}

}  // extern "C"
