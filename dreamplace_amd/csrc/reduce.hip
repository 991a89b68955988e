// Error reporting and the final stage of the two-stage FP64 reductions.
#include <stdarg.h>
#include <string.h>

#include "dp_common.h"

namespace dp {

static thread_local char g_err[512] = "";
Probe g_probe;

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

// One workgroup: out[k] = sum (or max) of partials[k*kReduceMaxBlocks + 0..nblocks),
// each lane adding a fixed subset in a fixed order -> deterministic.
__global__ void __launch_bounds__(kReduceBlock) k_finish(const double* partials, int nblocks,
                                                         int nout, int max_mask, void* out,
                                                         int out_dtype) {
  __shared__ double smem[kReduceBlock / 64];
  for (int k = 0; k < nout; ++k) {
    const double* p = partials + (size_t)k * kReduceMaxBlocks;
    bool is_max = (max_mask >> k) & 1;
    double acc = is_max ? -HUGE_VAL : 0.0;
    // each lane takes i = lane, lane + 256, ... in that order; eight loads
    // are issued before their eight (ordered) additions, so the lanes do
    // not wait out one load latency per partial
    constexpr int U = 8;
    int i = threadIdx.x;
    for (; i + (U - 1) * kReduceBlock < nblocks; i += U * kReduceBlock) {
      double v[U];
#pragma unroll
      for (int u = 0; u < U; ++u) v[u] = p[i + u * kReduceBlock];
#pragma unroll
      for (int u = 0; u < U; ++u) acc = is_max ? (v[u] > acc ? v[u] : acc) : acc + v[u];
    }
    for (; i < nblocks; i += kReduceBlock) {
      const double v = p[i];
      acc = is_max ? (v > acc ? v : acc) : acc + v;
    }
    double r = is_max ? block_max<kReduceBlock>(acc, smem) : block_sum<kReduceBlock>(acc, smem);
    if (threadIdx.x == 0) {
      if (out_dtype == DP_FLOAT32)
        ((float*)out)[k] = (float)r;
      else
        ((double*)out)[k] = r;
    }
  }
}

int finish_reduction(const double* partials, int nblocks, int nout, const int* is_max_host,
                     void* out, int out_dtype, hipStream_t s) {
  int mask = 0;
  for (int k = 0; k < nout; ++k)
    if (is_max_host && is_max_host[k]) mask |= 1 << k;
  hipLaunchKernelGGL(k_finish, dim3(1), dim3(kReduceBlock), 0, s, partials, nblocks, nout, mask,
                     out, out_dtype);
  DP_LAUNCHED("k_finish");
  return DP_OK;
}

}  // namespace dp

extern "C" {

const char* dp_last_error(void) { return dp::g_err; }

int dp_version(void) { return 1; }

int dp_profile_kernel(int kernel_id, void* start_event, void* stop_event) {
  if (kernel_id >= DP_KERNEL_COUNT) {
    dp::set_error("dp_profile_kernel: unknown kernel id %d", kernel_id);
    return DP_ERR_INVALID_ARGUMENT;
  }
  if (kernel_id >= 0 && (!start_event || !stop_event)) {
    dp::set_error("dp_profile_kernel: null event");
    return DP_ERR_INVALID_ARGUMENT;
  }
  dp::g_probe.kernel_id = kernel_id < 0 ? -1 : kernel_id;
  dp::g_probe.start = (hipEvent_t)start_event;
  dp::g_probe.stop = (hipEvent_t)stop_event;
  return DP_OK;
}

}  // extern "C"
