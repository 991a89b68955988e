// Vector work of the Nesterov step and the preconditioner of the placement
// objective, on the device, with the step-size norms accumulated in FP64.
//
// Nesterov (dreamplace/NesterovAcceleratedGradientOptimizer.py:65-240):
//   a_{k+1} = (1 + sqrt(4 a_k^2 + 1)) / 2,  coef = (a_k - 1) / a_{k+1}
//   u_{k+1} = v_k - alpha_k * g_k
//   v_{k+1} = u_{k+1} + coef * (u_{k+1} - u_k)
// in the positions' type, in the reference's operation order; the
// Lipschitz / Barzilai-Borwein step sizes come from sum((v1-v0)^2),
// sum((g1-g0)^2) and sum((v1-v0)(g1-g0)), differences in the positions' type,
// squares and sums in FP64 (deterministic two-stage reduction).
//
// Preconditioner (dreamplace/PlaceObj.py:69-132, no fence regions):
//   precond = max(pws + (alpha * density_weight) * node_area, 1)
//   grad[x] /= precond, grad[y] /= precond.
#include "dp_common.h"

#pragma clang fp contract(off)

namespace dp {

template <typename T>
__global__ void k_nesterov_extrapolate(int64_t n, const T* __restrict__ v_k,
                                       const T* __restrict__ g_k, const T* __restrict__ u_k,
                                       const T* __restrict__ alpha_p, const T* __restrict__ a_p,
                                       T* u_kp1, T* v_kp1, T* a_kp1_out) {
  const T a = *a_p;
  const T a_kp1 = (T(1) + sqrt(T(4) * (a * a) + T(1))) / T(2);
  const T coef = (a - T(1)) / a_kp1;
  const T alpha = *alpha_p;
  const int64_t i0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i0 == 0 && a_kp1_out) *a_kp1_out = a_kp1;
  for (int64_t i = i0; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const T u = v_k[i] - alpha * g_k[i];
    const T uk = u_k[i];
    u_kp1[i] = u;
    v_kp1[i] = u + coef * (u - uk);
  }
}

// The die box move_boundary applies to movable and filler nodes
// (move_boundary_cuda_kernel.cu:10-40): x in [xl, xh - size_x],
// y in [yl, yh - size_y].  Uniform fillers take their size from
// filler_sx/sy instead of the per-node arrays.
template <typename T>
struct Box {
  const T* sx;
  const T* sy;
  T xl, yl, xh, yh, filler_sx, filler_sy;
  int64_t num_nodes, num_movable, num_filler;
  bool uniform_fillers;
};

// k_nesterov_extrapolate followed by move_boundary on v_kp1, in one pass:
// v_kp1 is formed as above, then clamped to the box of its node.  Entries
// of fixed nodes are left as formed, as move_boundary leaves them.  Thread
// t takes nodes t, t + grid, ... (x and y); with `partials` it also sums
// (v_kp1 - v_k)^2 in FP64 in the order k_step_norms does (the step-size
// numerator of step_nobb, NesterovAcceleratedGradientOptimizer.py:131).
template <typename T>
__device__ __forceinline__ T box_clamp(const Box<T>& box, T v, int64_t j, bool is_y) {
  const int64_t fill_begin = box.num_nodes - box.num_filler;
  if (j < box.num_movable || j >= fill_begin) {
    const T lo = is_y ? box.yl : box.xl;
    T size;
    if (box.uniform_fillers && j >= fill_begin)
      size = is_y ? box.filler_sy : box.filler_sx;
    else
      size = is_y ? box.sy[j] : box.sx[j];
    const T hi = (is_y ? box.yh : box.xh) - size;
    v = v < lo ? lo : v;
    v = hi < v ? hi : v;
  }
  return v;
}

template <typename T>
__global__ void __launch_bounds__(kReduceBlock) k_nesterov_extrapolate_box(
    const T* __restrict__ v_k, const T* __restrict__ g_k, const T* __restrict__ u_k,
    const T* __restrict__ alpha_p, const T* __restrict__ a_p, T* u_kp1, T* v_kp1, T* a_kp1_out,
    Box<T> box, double* partials) {
  __shared__ double smem[kReduceBlock / 64];
  const T a = *a_p;
  const T a_kp1 = (T(1) + sqrt(T(4) * (a * a) + T(1))) / T(2);
  const T coef = (a - T(1)) / a_kp1;
  const T alpha = *alpha_p;
  const int64_t N = box.num_nodes;
  const int64_t t0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t0 == 0 && a_kp1_out) *a_kp1_out = a_kp1;
  double svv = 0;
  for (int64_t j = t0; j < N; j += (int64_t)gridDim.x * blockDim.x) {
#pragma unroll
    for (int axis = 0; axis < 2; ++axis) {
      const int64_t i = axis ? N + j : j;
      const T vk = v_k[i];
      const T u = vk - alpha * g_k[i];
      const T uk = u_k[i];
      u_kp1[i] = u;
      const T v = box_clamp(box, u + coef * (u - uk), j, axis == 1);
      v_kp1[i] = v;
      if (partials) {
        const double dv = (double)(v - vk);
        svv += dv * dv;
      }
    }
  }
  if (partials) {
    const double r = block_sum<kReduceBlock>(svv, smem);
    if (threadIdx.x == 0) partials[blockIdx.x] = r;
  }
}

// partials of sum(da^2), sum(db^2), sum(da*db), da = a1 - a0, db = b1 - b0:
// the n entries taken as n/2 (x, y) pairs (i, n/2 + i) over the grid of
// node_reduce_blocks, thread t taking pairs t, t + grid, ... (the order the
// fused sums of the extrapolation and the node-gradient pass follow), and
// an odd last entry by thread 0 at the end.
template <typename T>
__global__ void __launch_bounds__(kReduceBlock) k_step_norms(int64_t n, const T* __restrict__ a1,
                                                             const T* __restrict__ a0,
                                                             const T* __restrict__ b1,
                                                             const T* __restrict__ b0,
                                                             double* partials) {
  __shared__ double smem[kReduceBlock / 64];
  double saa = 0, sbb = 0, sab = 0;
  const int64_t half = n / 2;
  const int64_t t0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  auto add = [&](int64_t i) {
    const double da = (double)(a1[i] - a0[i]);
    const double db = (double)(b1[i] - b0[i]);
    saa += da * da;
    sbb += db * db;
    sab += da * db;
  };
  for (int64_t j = t0; j < half; j += (int64_t)gridDim.x * blockDim.x) {
    add(j);
    add(half + j);
  }
  if ((n & 1) && t0 == 0) add(n - 1);
  double r = block_sum<kReduceBlock>(saa, smem);
  if (threadIdx.x == 0) partials[blockIdx.x] = r;
  r = block_sum<kReduceBlock>(sbb, smem);
  if (threadIdx.x == 0) partials[kReduceMaxBlocks + blockIdx.x] = r;
  r = block_sum<kReduceBlock>(sab, smem);
  if (threadIdx.x == 0) partials[2 * kReduceMaxBlocks + blockIdx.x] = r;
}

// partials of sum(|g|) over [0, n)
template <typename T>
__global__ void __launch_bounds__(kReduceBlock) k_l1_norm(int64_t n, const T* __restrict__ g,
                                                          double* partials) {
  __shared__ double smem[kReduceBlock / 64];
  double s = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    s += fabs((double)g[i]);
  const double r = block_sum<kReduceBlock>(s, smem);
  if (threadIdx.x == 0) partials[blockIdx.x] = r;
}

template <typename T>
__global__ void k_precondition(int64_t n, T* grad, const T* __restrict__ pws,
                               const T* __restrict__ area, const T* __restrict__ density_weight,
                               T alpha) {
  const T c = alpha * *density_weight;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    T p = pws[i] + c * area[i];
    p = p < T(1) ? T(1) : p;
    grad[i] = grad[i] / p;
    grad[n + i] = grad[n + i] / p;
  }
}

static int reduce_blocks(int64_t n) {
  int64_t b = ceil_div(n, kReduceBlock);
  if (b < 1) b = 1;
  return (int)(b > 1024 ? 1024 : b);
}

}  // namespace dp

using namespace dp;

extern "C" {

size_t dp_optimizer_workspace_bytes(void) {
  return align256(sizeof(double) * 3 * (size_t)kReduceMaxBlocks);
}

int dp_nesterov_extrapolate(int dtype, int64_t n, const void* v_k, const void* g_k,
                            const void* u_k, const void* alpha_k, const void* a_k, void* u_kp1,
                            void* v_kp1, void* a_kp1, void* stream) {
  DP_ARG(v_k && g_k && u_k && alpha_k && a_k && u_kp1 && v_kp1,
         "dp_nesterov_extrapolate: null pointer");
  DP_ARG(n >= 0, "dp_nesterov_extrapolate: bad size");
  DP_ARG(a_kp1 != a_k, "dp_nesterov_extrapolate: a_kp1 must not alias a_k");
  hipStream_t s = S(stream);
  return DP_DISPATCH(dtype, "dp_nesterov_extrapolate", [&] {
    hipLaunchKernelGGL(k_nesterov_extrapolate<T>, dim3(grid_for(n > 0 ? n : 1, 256)), dim3(256),
                       0, s, n, (const T*)v_k, (const T*)g_k, (const T*)u_k, (const T*)alpha_k,
                       (const T*)a_k, (T*)u_kp1, (T*)v_kp1, (T*)a_kp1);
    DP_LAUNCHED("k_nesterov_extrapolate");
    return DP_OK;
  });
}

int dp_nesterov_extrapolate_bounded(int dtype, int64_t num_nodes, const void* v_k, const void* g_k,
                                    const void* u_k, const void* alpha_k, const void* a_k,
                                    void* u_kp1, void* v_kp1, void* a_kp1,
                                    const void* node_size_x, const void* node_size_y, double xl,
                                    double yl, double xh, double yh, int64_t num_movable_nodes,
                                    int64_t num_filler_nodes, const double* filler_size,
                                    void* workspace, double* dv_norm2_out, void* stream) {
  DP_ARG(v_k && g_k && u_k && alpha_k && a_k && u_kp1 && v_kp1 && node_size_x && node_size_y,
         "dp_nesterov_extrapolate_bounded: null pointer");
  DP_ARG(!dv_norm2_out || workspace, "dp_nesterov_extrapolate_bounded: the norm needs a workspace");
  DP_ARG(num_nodes >= 0 && num_movable_nodes >= 0 && num_filler_nodes >= 0 &&
             num_movable_nodes + num_filler_nodes <= num_nodes,
         "dp_nesterov_extrapolate_bounded: bad node counts");
  DP_ARG(a_kp1 != a_k, "dp_nesterov_extrapolate_bounded: a_kp1 must not alias a_k");
  hipStream_t s = S(stream);
  const int nb = node_reduce_blocks(num_nodes);
  double* part = dv_norm2_out ? (double*)workspace : nullptr;
  return DP_DISPATCH(dtype, "dp_nesterov_extrapolate_bounded", [&] {
    Box<T> box;
    box.sx = (const T*)node_size_x;
    box.sy = (const T*)node_size_y;
    box.xl = (T)xl;
    box.yl = (T)yl;
    box.xh = (T)xh;
    box.yh = (T)yh;
    box.uniform_fillers = filler_size != nullptr;
    box.filler_sx = filler_size ? (T)filler_size[0] : T(0);
    box.filler_sy = filler_size ? (T)filler_size[1] : T(0);
    box.num_nodes = num_nodes;
    box.num_movable = num_movable_nodes;
    box.num_filler = num_filler_nodes;
    hipLaunchKernelGGL(k_nesterov_extrapolate_box<T>, dim3(nb), dim3(kReduceBlock), 0, s,
                       (const T*)v_k, (const T*)g_k, (const T*)u_k, (const T*)alpha_k,
                       (const T*)a_k, (T*)u_kp1, (T*)v_kp1, (T*)a_kp1, box, part);
    DP_LAUNCHED("k_nesterov_extrapolate_box");
    if (part) return finish_reduction(part, nb, 1, nullptr, dv_norm2_out, -1, s);
    return DP_OK;
  });
}

int dp_step_norms(int dtype, int64_t n, const void* a1, const void* a0, const void* b1,
                  const void* b0, void* workspace, double* out3, void* stream) {
  DP_ARG(a1 && a0 && b1 && b0 && workspace && out3, "dp_step_norms: null pointer");
  DP_ARG(n >= 0, "dp_step_norms: bad size");
  hipStream_t s = S(stream);
  double* part = (double*)workspace;
  const int nb = node_reduce_blocks(n / 2);
  return DP_DISPATCH(dtype, "dp_step_norms", [&] {
    hipLaunchKernelGGL(k_step_norms<T>, dim3(nb), dim3(kReduceBlock), 0, s, n, (const T*)a1,
                       (const T*)a0, (const T*)b1, (const T*)b0, part);
    DP_LAUNCHED("k_step_norms");
    return finish_reduction(part, nb, 3, nullptr, out3, -1, s);
  });
}

int dp_l1_norm(int dtype, int64_t n, const void* g, void* workspace, double* out, void* stream) {
  DP_ARG(g && workspace && out, "dp_l1_norm: null pointer");
  DP_ARG(n >= 0, "dp_l1_norm: bad size");
  hipStream_t s = S(stream);
  double* part = (double*)workspace;
  const int nb = reduce_blocks(n);
  return DP_DISPATCH(dtype, "dp_l1_norm", [&] {
    hipLaunchKernelGGL(k_l1_norm<T>, dim3(nb), dim3(kReduceBlock), 0, s, n, (const T*)g, part);
    DP_LAUNCHED("k_l1_norm");
    return finish_reduction(part, nb, 1, nullptr, out, -1, s);
  });
}

int dp_precondition(int dtype, void* grad, int64_t num_nodes, const void* pin_weight_sum,
                    const void* node_areas, const void* density_weight, double alpha,
                    void* stream) {
  DP_ARG(grad && pin_weight_sum && node_areas && density_weight, "dp_precondition: null pointer");
  DP_ARG(num_nodes >= 0, "dp_precondition: bad size");
  if (num_nodes == 0) return DP_OK;
  hipStream_t s = S(stream);
  return DP_DISPATCH(dtype, "dp_precondition", [&] {
    hipLaunchKernelGGL(k_precondition<T>, dim3(grid_for(num_nodes, 256)), dim3(256), 0, s,
                       num_nodes, (T*)grad, (const T*)pin_weight_sum, (const T*)node_areas,
                       (const T*)density_weight, (T)alpha);
    DP_LAUNCHED("k_precondition");
    return DP_OK;
  });
}

}  // extern "C"
