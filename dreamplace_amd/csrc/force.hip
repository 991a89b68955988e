// Electric field gather: the density op's backward.
//
// For each movable or filler cell i:
//   g_x[i] = -(ratio_i * sum_{k,h} px_k*py_h*E_x[k,h]) * grad
// (dreamplace/ops/electric_potential/src/electric_force.cpp:116-172, the
// grad_out.mul_(grad_pos) at :96 and the negation of
// electric_potential.py:223).  Small cells (bin span <= 3x3) accumulate in the
// reference CPU order (k outer, h inner, in T, then *ratio); larger cells are
// queued to a wavefront each and reduced across lanes.  Terminals get 0.
#include "dp_density.h"

#pragma clang fp contract(off)

namespace dp {

template <typename T>
__global__ void __launch_bounds__(256) k_force_small(TriArgs<T> a, const T* __restrict__ fx,
                                                     const T* __restrict__ fy, const T* grad,
                                                     int64_t num_nodes, int64_t n_mov,
                                                     int64_t n_fill, int il, T* out, int* qcount,
                                                     int* qlist, int qcap) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  // index space: [0, n_mov) movables, then [n_mov, n_mov + n_fix) terminals
  // (written 0), then fillers; i.e. every node once.
  if (t >= num_nodes) return;
  int64_t i = t;
  if (i >= n_mov && i < num_nodes - n_fill) {
    put_grad(out, num_nodes, i, T(0), T(0), il);
    return;
  }
  CellBins<T> c = cell_bins(a, i);
  int nx = c.bxh - c.bxl, ny = c.byh - c.byl;
  T g = grad ? *grad : T(1);
  if (nx > 3 || ny > 3) {
    int slot = atomicAdd(qcount, 1);
    if (slot < qcap) {
      qlist[slot] = (int)i;
      return;
    }
    T gx = 0, gy = 0;
    const T dx = rel_origin(c.x, c.ex, a.xl, c.bxl, a.bsx);
    const T dy = rel_origin(c.y, c.ey, a.yl, c.byl, a.bsy);
    for (int k = c.bxl; k < c.bxh; ++k) {
      T px = overlap(a.acc, c.x, dx, c.sx, a.xl, c.bxl, k - c.bxl, a.bsx);
      for (int h = c.byl; h < c.byh; ++h) {
        T area = px * overlap(a.acc, c.y, dy, c.sy, a.yl, c.byl, h - c.byl, a.bsy);
        int64_t idx = (int64_t)k * a.nby + h;
        gx += area * fx[idx];
        gy += area * fy[idx];
      }
    }
    put_grad(out, num_nodes, i, -((gx * c.r) * g), -((gy * c.r) * g), il);
    return;
  }
  T px[3], py[3];
  const T dx = rel_origin(c.x, c.ex, a.xl, c.bxl, a.bsx);
  const T dy = rel_origin(c.y, c.ey, a.yl, c.byl, a.bsy);
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    px[j] = overlap(a.acc, c.x, dx, c.sx, a.xl, c.bxl, j, a.bsx);
    py[j] = overlap(a.acc, c.y, dy, c.sy, a.yl, c.byl, j, a.bsy);
  }
  T gx = 0, gy = 0;
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    if (j < nx) {
      const int64_t row = (int64_t)(c.bxl + j) * a.nby + c.byl;
#pragma unroll
      for (int l = 0; l < 3; ++l) {
        if (l < ny) {
          T area = px[j] * py[l];
          gx += area * fx[row + l];
          gy += area * fy[row + l];
        }
      }
    }
  }
  put_grad(out, num_nodes, i, -((gx * c.r) * g), -((gy * c.r) * g), il);
}

// One wavefront per queued cell: lanes stride over the cell's bins, then a
// fixed-order butterfly reduction.
template <typename T>
__global__ void __launch_bounds__(256) k_force_large(TriArgs<T> a, const T* __restrict__ fx,
                                                     const T* __restrict__ fy, const T* grad,
                                                     int64_t num_nodes, int il, T* out,
                                                     const int* qcount, const int* qlist,
                                                     int qcap) {
  int cnt = *qcount;
  if (cnt > qcap) cnt = qcap;
  const int lane = threadIdx.x & 63;
  const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int nwaves = (gridDim.x * blockDim.x) >> 6;
  T g = grad ? *grad : T(1);
  for (int q = wave; q < cnt; q += nwaves) {
    int64_t i = qlist[q];
    CellBins<T> c = cell_bins(a, i);
    int nx = c.bxh - c.bxl, ny = c.byh - c.byl;
    int total = nx * ny;
    T gx = 0, gy = 0;
    const T dx = rel_origin(c.x, c.ex, a.xl, c.bxl, a.bsx);
    const T dy = rel_origin(c.y, c.ey, a.yl, c.byl, a.bsy);
    for (int t = lane; t < total; t += 64) {
      int k = c.bxl + t / ny, h = c.byl + t % ny;
      T area = overlap(a.acc, c.x, dx, c.sx, a.xl, c.bxl, k - c.bxl, a.bsx) *
               overlap(a.acc, c.y, dy, c.sy, a.yl, c.byl, h - c.byl, a.bsy);
      int64_t idx = (int64_t)k * a.nby + h;
      gx += area * fx[idx];
      gy += area * fy[idx];
    }
    gx = wave_sum(gx);
    gy = wave_sum(gy);
    if (lane == 0) put_grad(out, num_nodes, i, -((gx * c.r) * g), -((gy * c.r) * g), il);
  }
}

// d/dx of tri(x, s, xl, k, b) from the right: [x+s < bin right] - [x >= bin
// left], the derivatives of the two sides of the overlap.
template <typename T>
__device__ __forceinline__ T dtri(T x, T s, T xl, int k, T b) {
  T bxl = xl + (T)k * b;
  T e = x + s, be = bxl + b;
  return ((e < be) ? T(1) : T(0)) - ((x >= bxl) ? T(1) : T(0));
}

// Exact adjoint of the energy: one thread per node over the cell's scatter
// bin range (k outer, h inner, accumulation in T, then * ratio * coef).
// Only the bins holding a cell edge have a non-zero overlap derivative, so
// a cell touches the potential map at the same bins as the scatter.
template <typename T>
__global__ void __launch_bounds__(256) k_energy_grad(TriArgs<T> a, const T* __restrict__ phi,
                                                     const T* grad, T coef, int64_t num_nodes,
                                                     int64_t n_mov, int64_t n_fill, int il,
                                                     T* out) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= num_nodes) return;
  if (i >= n_mov && i < num_nodes - n_fill) {
    put_grad(out, num_nodes, i, T(0), T(0), il);
    return;
  }
  CellBins<T> c = cell_bins(a, i);
  T g = grad ? *grad : T(1);
  T gx = 0, gy = 0;
  for (int k = c.bxl; k < c.bxh; ++k) {
    T px = tri(c.x, c.sx, a.xl, k, a.bsx);
    T dpx = dtri(c.x, c.sx, a.xl, k, a.bsx);
    const int64_t row = (int64_t)k * a.nby;
    for (int h = c.byl; h < c.byh; ++h) {
      T py = tri(c.y, c.sy, a.yl, h, a.bsy);
      T dpy = dtri(c.y, c.sy, a.yl, h, a.bsy);
      T p = phi[row + h];
      gx += (dpx * py) * p;
      gy += (px * dpy) * p;
    }
  }
  put_grad(out, num_nodes, i, ((gx * c.r) * coef) * g, ((gy * c.r) * coef) * g, il);
}

}  // namespace dp

using namespace dp;

extern "C" int dp_energy_gradient(int dtype, const void* grad_scale, const void* potential_map,
                                  const void* pos, int64_t num_nodes,
                                  const void* node_size_x_clamped, const void* node_size_y_clamped,
                                  const void* offset_x, const void* offset_y, const void* ratio,
                                  double xl, double yl, double bin_size_x, double bin_size_y,
                                  int64_t num_movable_nodes, int64_t num_filler_nodes,
                                  int num_bins_x, int num_bins_y, int grad_layout,
                                  const double* uniform_filler, void* grad_out, void* stream) {
  DP_ARG(potential_map && pos && node_size_x_clamped && node_size_y_clamped && offset_x &&
             offset_y && ratio && grad_out,
         "dp_energy_gradient: null pointer");
  DP_ARG(num_bins_x > 0 && num_bins_y > 0 && bin_size_x > 0 && bin_size_y > 0,
         "dp_energy_gradient: bad bins");
  DP_ARG(num_movable_nodes >= 0 && num_filler_nodes >= 0 &&
             num_movable_nodes + num_filler_nodes <= num_nodes && num_nodes < ((int64_t)1 << 31),
         "dp_energy_gradient: bad node counts");
  DP_ARG(grad_layout == DP_LAYOUT_SPLIT || grad_layout == DP_LAYOUT_INTERLEAVED,
         "dp_energy_gradient: bad grad_layout %d", grad_layout);
  if (num_nodes == 0) return DP_OK;
  hipStream_t s = S(stream);
  const int il = grad_layout == DP_LAYOUT_INTERLEAVED;
  return DP_DISPATCH(dtype, "dp_energy_gradient", [&] {
    TriArgs<T> a = make_tri_args<T>(pos, num_nodes, node_size_x_clamped, node_size_y_clamped,
                                    offset_x, offset_y, ratio, xl, yl, bin_size_x, bin_size_y,
                                    num_bins_x, num_bins_y, 1, uniform_filler,
                                    num_nodes - num_filler_nodes);
    const T coef = (T)(2.0 / (bin_size_x * bin_size_y));
    hipLaunchKernelGGL(k_energy_grad<T>, dim3((unsigned)ceil_div(num_nodes, 256)), dim3(256), 0,
                       s, a, (const T*)potential_map, (const T*)grad_scale, coef, num_nodes,
                       num_movable_nodes, num_filler_nodes, il, (T*)grad_out);
    DP_LAUNCHED("k_energy_grad");
    return DP_OK;
  });
}

// The field gather of dp_electric_force; with filler_grad_out (tiled only)
// the fillers' finished gradients go there (GradOut).
static int electric_force(int dtype, const void* grad_scale, const void* field_map_x,
                          const void* field_map_y, const void* pos, int64_t num_nodes,
                          const void* node_size_x_clamped, const void* node_size_y_clamped,
                          const void* offset_x, const void* offset_y, const void* ratio, double xl,
                          double yl, double bin_size_x, double bin_size_y,
                          int64_t num_movable_nodes, int64_t num_filler_nodes, int num_bins_x,
                          int num_bins_y, int algorithm, int64_t large_capacity, int grad_layout,
                          const double* uniform_filler, void* workspace, void* grad_out,
                          void* filler_grad_out, const void* density_weight, double alpha,
                          double filler_area, void* stream) {
  DP_ARG(field_map_x && field_map_y && pos && node_size_x_clamped && node_size_y_clamped &&
             offset_x && offset_y && ratio && grad_out && workspace,
         "dp_electric_force: null pointer");
  DP_ARG(num_bins_x > 0 && num_bins_y > 0, "dp_electric_force: bad bin counts");
  DP_ARG(num_movable_nodes >= 0 && num_filler_nodes >= 0 &&
             num_movable_nodes + num_filler_nodes <= num_nodes && num_nodes < ((int64_t)1 << 31),
         "dp_electric_force: bad node counts");
  DP_ARG(large_capacity >= 0 && large_capacity < ((int64_t)1 << 31),
         "dp_electric_force: bad large_capacity");
  hipStream_t s = S(stream);
  const int64_t nb = (int64_t)num_bins_x * num_bins_y;
  const int acc = (algorithm & DP_ARITH_BIN_RELATIVE) != 0;
  algorithm &= ~DP_ARITH_BIN_RELATIVE;
  DP_ARG(algorithm == DP_SCATTER_GLOBAL || algorithm == DP_SCATTER_TILED,
         "dp_electric_force: bad algorithm %d", algorithm);
  DP_ARG(grad_layout == DP_LAYOUT_SPLIT || grad_layout == DP_LAYOUT_INTERLEAVED,
         "dp_electric_force: bad grad_layout %d", grad_layout);
  const int il = grad_layout == DP_LAYOUT_INTERLEAVED;
  DP_PROBE_OP(DP_OP_ELECTRIC_FORCE, s);
  // Same workspace layout as dp_density_map: [int64 map][count][queue][tiled].
  char* p = (char*)workspace + density_queue_offset(nb);
  int* qcount = (int*)p;
  int* qlist = (int*)(p + 256);
  char* tiled_ws = (char*)workspace + density_tiled_offset(nb, large_capacity);
  if (num_nodes == 0) return DP_OK;
  return DP_DISPATCH(dtype, "dp_electric_force", [&] {
    TriArgs<T> a = make_tri_args<T>(pos, num_nodes, node_size_x_clamped, node_size_y_clamped,
                                    offset_x, offset_y, ratio, xl, yl, bin_size_x, bin_size_y,
                                    num_bins_x, num_bins_y, 1, uniform_filler,
                                    num_nodes - num_filler_nodes);
    a.acc = acc;
    const bool tiled = algorithm == DP_SCATTER_TILED && tiled_supported(num_bins_x, num_bins_y);
    DP_REQUIRE(!filler_grad_out || tiled, DP_ERR_UNSUPPORTED,
               "dp_electric_force_fillers_final: needs the tiled gather");
    if (tiled) {
      GradOut<T> go = {(T*)grad_out, num_nodes, il, (T*)filler_grad_out,
                       num_nodes - num_filler_nodes, (const T*)density_weight, (T)alpha,
                       (T)filler_area};
      return gather_tiled<T>(a, num_movable_nodes, num_filler_nodes, (const T*)field_map_x,
                             (const T*)field_map_y, (const T*)grad_scale, tiled_ws, go, s);
    }
    hipLaunchKernelGGL(k_fill<int>, dim3(1), dim3(64), 0, s, qcount, (int64_t)1, 0);
    DP_PROBED(DP_KERNEL_ELECTRIC_FORCE, s,
              hipLaunchKernelGGL(k_force_small<T>, dim3((unsigned)ceil_div(num_nodes, 256)),
                                 dim3(256), 0, s, a, (const T*)field_map_x,
                                 (const T*)field_map_y, (const T*)grad_scale, num_nodes,
                                 num_movable_nodes, num_filler_nodes, il, (T*)grad_out, qcount,
                                 qlist, (int)large_capacity));
    DP_LAUNCHED("k_force_small");
    hipLaunchKernelGGL(k_force_large<T>, dim3(256), dim3(256), 0, s, a, (const T*)field_map_x,
                       (const T*)field_map_y, (const T*)grad_scale, num_nodes, il, (T*)grad_out,
                       (const int*)qcount, (const int*)qlist, (int)large_capacity);
    DP_LAUNCHED("k_force_large");
    return DP_OK;
  });
}

extern "C" int dp_electric_force(int dtype, const void* grad_scale, const void* field_map_x,
                                 const void* field_map_y, const void* pos, int64_t num_nodes,
                                 const void* node_size_x_clamped, const void* node_size_y_clamped,
                                 const void* offset_x, const void* offset_y, const void* ratio,
                                 double xl, double yl, double bin_size_x, double bin_size_y,
                                 int64_t num_movable_nodes, int64_t num_filler_nodes,
                                 int num_bins_x, int num_bins_y, int algorithm,
                                 int64_t large_capacity, int grad_layout,
                                 const double* uniform_filler, void* workspace, void* grad_out,
                                 void* stream) {
  return electric_force(dtype, grad_scale, field_map_x, field_map_y, pos, num_nodes,
                        node_size_x_clamped, node_size_y_clamped, offset_x, offset_y, ratio, xl,
                        yl, bin_size_x, bin_size_y, num_movable_nodes, num_filler_nodes,
                        num_bins_x, num_bins_y, algorithm, large_capacity, grad_layout,
                        uniform_filler, workspace, grad_out, nullptr, nullptr, 0.0, -1.0, stream);
}

extern "C" int dp_electric_force_fillers_final(
    int dtype, const void* grad_scale, const void* field_map_x, const void* field_map_y,
    const void* pos, int64_t num_nodes, const void* node_size_x_clamped,
    const void* node_size_y_clamped, const void* offset_x, const void* offset_y,
    const void* ratio, double xl, double yl, double bin_size_x, double bin_size_y,
    int64_t num_movable_nodes, int64_t num_filler_nodes, int num_bins_x, int num_bins_y,
    int algorithm, int64_t large_capacity, int grad_layout, const double* uniform_filler,
    void* workspace, void* grad_out, void* grad_pos_out, const void* density_weight, double alpha,
    double filler_area, void* stream) {
  DP_ARG(grad_pos_out, "dp_electric_force_fillers_final: null grad_pos_out");
  DP_ARG(!density_weight || filler_area >= 0,
         "dp_electric_force_fillers_final: the preconditioner needs the fillers' area");
  return electric_force(dtype, grad_scale, field_map_x, field_map_y, pos, num_nodes,
                        node_size_x_clamped, node_size_y_clamped, offset_x, offset_y, ratio, xl,
                        yl, bin_size_x, bin_size_y, num_movable_nodes, num_filler_nodes,
                        num_bins_x, num_bins_y, algorithm, large_capacity, grad_layout,
                        uniform_filler, workspace, grad_out, grad_pos_out, density_weight, alpha,
                        filler_area, stream);
}
