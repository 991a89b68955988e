// Small per-iteration ops of the loop: move_boundary and the pin-weight sum.
#include "dp_common.h"

namespace dp {

// x = min(max(x, xl), xh - w) for movables and fillers, in place
// (move_boundary_cuda_kernel.cu:10-24; both axes in one thread).
template <typename T>
__global__ void k_move_boundary(T* x, T* y, const T* __restrict__ wx, const T* __restrict__ wy,
                                T xl, T yl, T xh, T yh, int64_t n, int64_t nmov, int64_t nfill) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (i < nmov || i >= n - nfill) {
    T v = x[i];
    v = v < xl ? xl : v;
    T hx = xh - wx[i];
    x[i] = hx < v ? hx : v;
    T u = y[i];
    u = u < yl ? yl : u;
    T hy = yh - wy[i];
    y[i] = hy < u ? hy : u;
  }
}

// out[n] = sum_{pins p of node n} w[pin2net[p]]; fillers 0 (pws_cuda_kernel.cu:10-21).
template <typename T>
__global__ void k_pws(const T* __restrict__ w, const int* __restrict__ flat_nodepin,
                      const int* __restrict__ nodepin_start, const int* __restrict__ pin2net,
                      int64_t nphys, int64_t nnodes, T* out) {
  int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= nnodes) return;
  T s = 0;
  if (n < nphys)
    for (int j = nodepin_start[n]; j < nodepin_start[n + 1]; ++j) s += w[pin2net[flat_nodepin[j]]];
  out[n] = s;
}

}  // namespace dp

using namespace dp;

extern "C" {

int dp_move_boundary(int dtype, void* pos, int64_t num_nodes, const void* node_size_x,
                     const void* node_size_y, double xl, double yl, double xh, double yh,
                     int64_t num_movable_nodes, int64_t num_filler_nodes, void* stream) {
  DP_ARG(pos && node_size_x && node_size_y, "dp_move_boundary: null pointer");
  DP_ARG(num_movable_nodes + num_filler_nodes <= num_nodes, "dp_move_boundary: bad node counts");
  if (num_nodes == 0) return DP_OK;
  hipStream_t s = S(stream);
  return DP_DISPATCH(dtype, "dp_move_boundary", [&] {
    T* P = (T*)pos;
    hipLaunchKernelGGL(k_move_boundary<T>, dim3((unsigned)ceil_div(num_nodes, 256)), dim3(256), 0,
                       s, P, P + num_nodes, (const T*)node_size_x, (const T*)node_size_y, (T)xl,
                       (T)yl, (T)xh, (T)yh, num_nodes, num_movable_nodes, num_filler_nodes);
    DP_LAUNCHED("k_move_boundary");
    return DP_OK;
  });
}

int dp_pin_weight_sum(int dtype, const void* net_weights, const int32_t* flat_nodepin,
                      const int32_t* nodepin_start, const int32_t* pin2net_map,
                      int64_t num_physical_nodes, int64_t num_nodes, void* out, void* stream) {
  DP_ARG(net_weights && flat_nodepin && nodepin_start && pin2net_map && out,
         "dp_pin_weight_sum: null pointer");
  DP_ARG(num_physical_nodes <= num_nodes, "dp_pin_weight_sum: bad node counts");
  if (num_nodes == 0) return DP_OK;
  hipStream_t s = S(stream);
  return DP_DISPATCH(dtype, "dp_pin_weight_sum", [&] {
    hipLaunchKernelGGL(k_pws<T>, dim3((unsigned)ceil_div(num_nodes, 256)), dim3(256), 0, s,
                       (const T*)net_weights, (const int*)flat_nodepin, (const int*)nodepin_start,
                       (const int*)pin2net_map, num_physical_nodes, num_nodes, (T*)out);
    DP_LAUNCHED("k_pws");
    return DP_OK;
  });
}

}  // extern "C"
