// Internal helpers shared by the HIP translation units of libdreamplace_amd.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>

#include <cstdarg>
#include <type_traits>

#include "../../include/dreamplace_amd.h"

namespace dp {

// Per-thread message of the last failure (dp_last_error).
void set_error(const char* fmt, ...);

// Kernel probe state (dp_profile_kernel).  DP_PROBE(id, s, launch) records the
// probe's start event, launches, then records the stop event, all on `s`.
struct Probe {
  int kernel_id = -1;
  hipEvent_t start = nullptr, stop = nullptr;
};
extern Probe g_probe;
#define DP_PROBED(id, s, ...)                                                     \
  do {                                                                            \
    const bool probe_ = ::dp::g_probe.kernel_id == (id);                          \
    if (probe_) (void)hipEventRecord(::dp::g_probe.start, (s));                   \
    __VA_ARGS__;                                                                  \
    if (probe_) (void)hipEventRecord(::dp::g_probe.stop, (s));                    \
  } while (0)
// Whole-entry-point probe: records the start event where it is declared and
// the stop event when the entry point returns, after all of its launches.
struct ProbeScope {
  hipStream_t s;
  bool on;
  ProbeScope(int id, hipStream_t s_) : s(s_), on(g_probe.kernel_id == id) {
    if (on) (void)hipEventRecord(g_probe.start, s);
  }
  ~ProbeScope() {
    if (on) (void)hipEventRecord(g_probe.stop, s);
  }
};
#define DP_PROBE_OP(id, s) ::dp::ProbeScope probe_scope_((id), (s))

// Status helpers.  Every entry point returns through these.
#define DP_REQUIRE(cond, code, ...)        \
  do {                                     \
    if (!(cond)) {                         \
      ::dp::set_error(__VA_ARGS__);        \
      return (code);                       \
    }                                      \
  } while (0)

#define DP_ARG(cond, ...) DP_REQUIRE(cond, DP_ERR_INVALID_ARGUMENT, __VA_ARGS__)

#define DP_LAUNCHED(name)                                                          \
  do {                                                                             \
    hipError_t e_ = hipGetLastError();                                             \
    if (e_ != hipSuccess) {                                                        \
      ::dp::set_error("%s: launch failed: %s", name, hipGetErrorString(e_));       \
      return (int)e_;                                                              \
    }                                                                              \
  } while (0)

#define DP_CHECK_RC(expr)        \
  do {                           \
    int rc_ = (expr);            \
    if (rc_ != DP_OK) return rc_; \
  } while (0)

inline hipStream_t S(void* s) { return reinterpret_cast<hipStream_t>(s); }

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

// Launch-grid cap for grid-stride kernels: enough workgroups to fill 256 CUs
// several times over.
constexpr int kMaxGrid = 65536;
inline int grid_for(int64_t n, int block) {
  int64_t g = ceil_div(n, block);
  if (g < 1) g = 1;
  return (int)(g > kMaxGrid ? kMaxGrid : g);
}

// Workspace carving: 256-byte aligned sub-buffers.
inline size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

// dtype dispatch: DP_FLOAT32 -> float, DP_FLOAT64 -> double.
#define DP_DISPATCH(dtype, NAME, ...)                                           \
  [&]() -> int {                                                                \
    if ((dtype) == DP_FLOAT32) {                                                \
      using T = float;                                                          \
      return __VA_ARGS__();                                                     \
    } else if ((dtype) == DP_FLOAT64) {                                         \
      using T = double;                                                         \
      return __VA_ARGS__();                                                     \
    }                                                                           \
    ::dp::set_error("%s: dtype %d not supported (float32=0, float64=1)", NAME,  \
                    (int)(dtype));                                              \
    return DP_ERR_DTYPE;                                                        \
  }()

// ---------------------------------------------------------------------------
// Wavefront / block reductions (64-wide wavefronts).
// ---------------------------------------------------------------------------
template <typename V>
__device__ __forceinline__ V wave_sum(V v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
template <typename V>
__device__ __forceinline__ V wave_max(V v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    V w = __shfl_xor(v, o, 64);
    v = v > w ? v : w;
  }
  return v;
}

// Orders LDS accesses among the lanes of one wavefront without a workgroup
// barrier: a wavefront's LDS instructions execute in program order, so it is
// enough that the compiler moves no LDS access across this point.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Block sum of a double in a fixed order (deterministic): waves reduce with
// the xor butterfly, then wave 0 adds the per-wave results in wave order.
template <int BLOCK>
__device__ __forceinline__ double block_sum(double v, double* smem) {
  constexpr int W = BLOCK / 64;
  v = wave_sum(v);
  int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) smem[wid] = v;
  __syncthreads();
  double r = 0;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < W; ++i) r += smem[i];
  }
  __syncthreads();
  return r;  // valid in thread 0
}

template <int BLOCK>
__device__ __forceinline__ double block_max(double v, double* smem) {
  constexpr int W = BLOCK / 64;
  v = wave_max(v);
  int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) smem[wid] = v;
  __syncthreads();
  double r = -HUGE_VAL;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < W; ++i) r = r > smem[i] ? r : smem[i];
  }
  __syncthreads();
  return r;
}

// Two-component vector of T: (x, y) pairs stored as one 8- or 16-byte word.
template <typename T>
struct Vec2;
template <>
struct Vec2<float> {
  typedef float2 type;
};
template <>
struct Vec2<double> {
  typedef double2 type;
};

// Fill kernel (grid-stride).  Used instead of hipMemsetAsync so that the
// clears also behave under stream capture.
template <typename V>
__global__ void k_fill(V* p, int64_t n, V v) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

// Zero n 8-byte words at a 16-byte aligned p with 16-byte stores (the
// fixed-point maps: one 8-byte store per thread ran at 1.5 TB/s on 8 MB).
static __global__ void k_zero_words(unsigned long long* p, int64_t n) {
  const int64_t pairs = n >> 1;
  uint4* q = (uint4*)p;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < pairs;
       i += (int64_t)gridDim.x * blockDim.x)
    q[i] = make_uint4(0u, 0u, 0u, 0u);
  if ((n & 1) && blockIdx.x == 0 && threadIdx.x == 0) p[n - 1] = 0ull;
}

// Zero up to four regions in one launch: region r is n[r] 4-byte words at
// p[r] (16-byte aligned; null skips it), cleared with 16-byte stores and
// 4-byte stores for the rest.
struct ZeroRegions {
  uint32_t* p[4];
  int64_t n[4];
};
static __global__ void k_zero_regions(ZeroRegions z) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    if (!z.p[r]) continue;
    const int64_t q = z.n[r] >> 2;
    uint4* v = (uint4*)z.p[r];
    for (int64_t i = t; i < q; i += stride) v[i] = make_uint4(0u, 0u, 0u, 0u);
    for (int64_t i = 4 * q + t; i < z.n[r]; i += stride) z.p[r][i] = 0u;
  }
}
inline void zero_regions(const ZeroRegions& z, hipStream_t s) {
  int64_t m = 1;
  for (int r = 0; r < 4; ++r)
    if (z.p[r] && (z.n[r] >> 2) > m) m = z.n[r] >> 2;
  hipLaunchKernelGGL(k_zero_regions, dim3(grid_for(m, 256)), dim3(256), 0, s, z);
}

// Launch the zeroing of n words at p on s (p 16-byte aligned, else word stores).
inline void zero_words(unsigned long long* p, int64_t n, hipStream_t s) {
  if (((uintptr_t)p & 15) == 0)
    hipLaunchKernelGGL(k_zero_words, dim3(grid_for((n >> 1) > 0 ? (n >> 1) : 1, 256)), dim3(256), 0, s, p, n);
  else
    hipLaunchKernelGGL(k_fill<unsigned long long>, dim3(grid_for(n, 256)), dim3(256), 0, s, p, n, 0ull);
}

// Two-stage deterministic reduction helpers (reduce.hip).
constexpr int kReduceBlock = 256;
constexpr int kReduceMaxBlocks = 16384;
// Workgroups of the reductions that pair the x and y entries of a [2N]
// per-node vector (the step-size norms, fused or not): thread t of the
// grid-stride loop takes nodes t, t + grid, ...; every such reduction uses
// this grid, so their sums are bit-identical.
inline int node_reduce_blocks(int64_t num_nodes) {
  int64_t b = (num_nodes + kReduceBlock - 1) / kReduceBlock;
  if (b < 1) b = 1;
  return (int)(b > kReduceMaxBlocks ? kReduceMaxBlocks : b);
}
// Bytes of workspace used by one two-stage reduction slot.
inline size_t reduce_slot_bytes() { return align256(sizeof(double) * kReduceMaxBlocks * 2); }
// Final stage: out[k] (k < nout) = sum (or max when is_max[k]) of
// partials[k*kReduceMaxBlocks + 0..nblocks), then optionally cast to the
// dtype of out_t (DP_FLOAT32 / DP_FLOAT64 / -1 for double output).
int finish_reduction(const double* partials, int nblocks, int nout, const int* is_max_host,
                     void* out, int out_dtype, hipStream_t s);

}  // namespace dp
