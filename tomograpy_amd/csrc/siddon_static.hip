// tg_conic_full_projector_static: the full projector / backprojector with the
// float64 backprojection accumulator in a device array of the library itself,
// for callers that cannot allocate device memory (code under HIP graph
// capture, tight serving loops, harnesses that forbid hipMalloc).  It has no
// counterpart in the reference.
//
// The array is a file-scope __device__ object: every GPU the library runs on
// holds its own copy, of kStaticAccDoubles doubles, until the process ends.
// Host code never passes the array's name to a kernel (on the host that name
// is a stand-in, not a device address); it passes the address that
// hipGetSymbolAddress returns for the current device, looked up once per
// device on the entry point's first call there.
#include <hip/hip_runtime.h>

#include <mutex>

#include "siddon_common.h"
#include "siddon_host.h"

namespace tg {
namespace {

// 2^24 doubles (128 MiB): a 256^3 cube, or 128^3 with 8 time bins.
constexpr int64_t kStaticAccDoubles = 1ll << 24;
constexpr int kMaxDevices = 64;

__device__ __align__(256) double scratch[kStaticAccDoubles];

std::mutex g_scratch_mutex;
double *g_scratch_ptr[kMaxDevices] = {};

// The current device's copy of `scratch` (looked up once per device).
int scratch_for_current_device(double **out) {
  int dev = 0;
  int rc = hip_check(hipGetDevice(&dev), "hipGetDevice");
  if (rc) return rc;
  if (dev < 0 || dev >= kMaxDevices)
    return set_error(TG_EINVAL, "device %d beyond the %d the static scratch serves", dev,
                     kMaxDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    rc = hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "static scratch address");
    if (rc) return rc;
    if (!p) return set_error(TG_EHIP, "static scratch has no address on device %d", dev);
    g_scratch_ptr[dev] = (double *)p;
  }
  *out = g_scratch_ptr[dev];
  return TG_OK;
}

// acc[0 .. n) = 0 (a kernel rather than hipMemsetAsync, so that the call can be
// recorded into a HIP graph).
__global__ __launch_bounds__(256) void k_zero_f64(double *__restrict__ acc, int64_t n) {
  for (int64_t k = blockIdx.x * 256ll + threadIdx.x; k < n; k += (int64_t)gridDim.x * 256)
    acc[k] = 0.0;
}

}  // namespace
}  // namespace tg

using namespace tg;

extern "C" {

int64_t tg_static_acc_capacity(void) { return kStaticAccDoubles; }

int tg_conic_full_projector_static(const tg_call *c, void *stream) {
  if (!c) return set_error(TG_EINVAL, "null call");
  const bool bwd = (c->kind == TG_BPJ || c->kind == TG_BPJT);
  if (!bwd) return dispatch_siddon(c, 0, c->nv, -1, -1, stream);  // no scratch needed
  if (c->flags & TG_DETERMINISTIC)
    return set_error(TG_EINVAL, "the static entry point has no deterministic workspace");
  const bool time = (c->kind == TG_BPJT);
  int64_t n = 1;
  for (int k = 0; k < 3; k++) {
    if (c->map.n[k] <= 0) return set_error(TG_EINVAL, "map edge %d is %d", k, c->map.n[k]);
    n *= c->map.n[k];
  }
  if (time) n *= (c->map.nt > 0 ? c->map.nt : 1);
  if (n > kStaticAccDoubles)
    return set_error(TG_EINVAL,
                     "a map of %lld voxels exceeds the static accumulator's %lld (use "
                     "tg_conic_full_projector with a caller-owned acc)",
                     (long long)n, (long long)kStaticAccDoubles);
  double *acc = nullptr;
  int rc = scratch_for_current_device(&acc);
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;
  // the previous call on this device left its range zero unless it failed
  // part-way; clear what this call sums into
  const int64_t want = (n + 255) / 256;
  hipLaunchKernelGGL(k_zero_f64, dim3((uint32_t)(want < 4096 ? want : 4096)), dim3(256), 0, s,
                     acc, n);
  if ((rc = hip_check(hipGetLastError(), "static accumulator clear"))) return rc;
  tg_call cc = *c;
  cc.flags |= TG_ACC_F64;
  cc.acc = acc;
  return dispatch_siddon(&cc, 0, cc.nv, -1, -1, stream);
}

}  // extern "C"
