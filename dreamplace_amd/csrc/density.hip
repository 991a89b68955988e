// Electrostatic density map: triangle (stretched-cell) scatter of movable and
// filler cells, exact-overlap map of fixed cells, fixed-point conversions and
// the overflow metric.
//
// All accumulation is in int64 fixed point: each (cell, bin) pair adds
// (int64)(((px*ratio)*py) * T(scale)), the arithmetic of
// dreamplace/ops/electric_potential/src/electric_density_map.cpp:254-309 with
// AtomicAdd<long> (dreamplace/ops/utility/src/atomic_ops.h:29-45).  Integer
// addition is associative, so the map is bit-identical for any launch shape
// and any cell order.  This file is compiled with -ffp-contract=off so that
// px, py and the bin indices round exactly as the reference's do.
#include "dp_common.h"
#include "dp_density.h"

#pragma clang fp contract(off)

namespace dp {

// ---------------------------------------------------------------------------
// Per-cell global-atomic scatter.  A thread owns one cell; cells whose bin
// span is at most 3x3 (every std cell and filler: their stretched size is
// sqrt(2) bins) are done inline with a fully unrolled 3x3 body; larger cells
// are pushed to a queue that one workgroup per cell drains, so a macro's
// thousands of bins do not serialise one lane of a wavefront.
// ---------------------------------------------------------------------------
// One launch covers both index ranges: t in [0, n_mov) are movables and
// t in [n_mov, n_mov + n_fill) map to the fillers at fill_begin + (t - n_mov).
template <typename T>
__global__ void __launch_bounds__(256) k_tri_scatter(TriArgs<T> a, int64_t n_mov,
                                                     int64_t fill_begin, int64_t total,
                                                     unsigned long long* map, int* qcount,
                                                     int* qlist, int qcap) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= total) return;
  const int64_t i = t < n_mov ? t : fill_begin + (t - n_mov);
  CellBins<T> c = cell_bins(a, i);
  int nx = c.bxh - c.bxl, ny = c.byh - c.byl;
  if (nx <= 0 || ny <= 0) return;
  if (nx > 3 || ny > 3) {
    int slot = atomicAdd(qcount, 1);
    if (slot < qcap) {
      qlist[slot] = (int)i;
      return;
    }
    // queue full: do the cell here (slow path, same arithmetic)
    const T dx = rel_origin(c.x, c.ex, a.xl, c.bxl, a.bsx);
    const T dy = rel_origin(c.y, c.ey, a.yl, c.byl, a.bsy);
    for (int k = c.bxl; k < c.bxh; ++k) {
      T pxr = overlap(a.acc, c.x, dx, c.sx, a.xl, c.bxl, k - c.bxl, a.bsx) * c.r;
      for (int h = c.byl; h < c.byh; ++h) {
        T area = pxr * overlap(a.acc, c.y, dy, c.sy, a.yl, c.byl, h - c.byl, a.bsy);
        long long v = (long long)(area * a.scale);
        atomicAdd(&map[(int64_t)k * a.nby + h], (unsigned long long)v);
      }
    }
    return;
  }
  T pxr[3], py[3];
  const T dx = rel_origin(c.x, c.ex, a.xl, c.bxl, a.bsx);
  const T dy = rel_origin(c.y, c.ey, a.yl, c.byl, a.bsy);
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    pxr[j] = overlap(a.acc, c.x, dx, c.sx, a.xl, c.bxl, j, a.bsx) * c.r;
    py[j] = overlap(a.acc, c.y, dy, c.sy, a.yl, c.byl, j, a.bsy);
  }
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    if (j < nx) {
#pragma unroll
      for (int l = 0; l < 3; ++l) {
        if (l < ny) {
          T area = pxr[j] * py[l];
          long long v = (long long)(area * a.scale);
          if (v != 0) atomicAdd(&map[(int64_t)(c.bxl + j) * a.nby + (c.byl + l)], (unsigned long long)v);
        }
      }
    }
  }
}

// One workgroup per queued (large) cell; lanes stride over its bins.
template <typename T>
__global__ void __launch_bounds__(256) k_tri_scatter_large(TriArgs<T> a, unsigned long long* map,
                                                           const int* qcount, const int* qlist,
                                                           int qcap) {
  int cnt = *qcount;
  if (cnt > qcap) cnt = qcap;
  for (int q = blockIdx.x; q < cnt; q += gridDim.x) {
    int64_t i = qlist[q];
    CellBins<T> c = cell_bins(a, i);
    int nx = c.bxh - c.bxl, ny = c.byh - c.byl;
    int total = nx * ny;
    const T dx = rel_origin(c.x, c.ex, a.xl, c.bxl, a.bsx);
    const T dy = rel_origin(c.y, c.ey, a.yl, c.byl, a.bsy);
    for (int t = threadIdx.x; t < total; t += blockDim.x) {
      int k = c.bxl + t / ny, h = c.byl + t % ny;
      T pxr = overlap(a.acc, c.x, dx, c.sx, a.xl, c.bxl, k - c.bxl, a.bsx) * c.r;
      T area = pxr * overlap(a.acc, c.y, dy, c.sy, a.yl, c.byl, h - c.byl, a.bsy);
      long long v = (long long)(area * a.scale);
      if (v != 0) atomicAdd(&map[(int64_t)k * a.nby + h], (unsigned long long)v);
    }
  }
}

// Exact overlap of fixed cells (electric_density_map.cpp:311-358,
// density_function.h:19-36 with fixed_node_flag = true: no boundary stretch).
template <typename T>
__global__ void __launch_bounds__(256) k_exact_scatter(const T* x, const T* y, const T* wx,
                                                       const T* wy, const T* bcx, const T* bcy,
                                                       int64_t n, int nbx, int nby, T xl, T yl,
                                                       T xh, T yh, T bsx, T bsy, T scale,
                                                       int rel, unsigned long long* map) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  T bxl = x[i], byl = y[i];
  T bxh = bxl + wx[i], byh = byl + wy[i];
  int kxl = max((int)((bxl - xl) / bsx), 0);
  int kxh = min((int)ceil((bxh - xl) / bsx) + 1, nbx);
  int kyl = max((int)((byl - yl) / bsy), 0);
  int kyh = min((int)ceil((byh - yl) / bsy) + 1, nby);
  for (int k = kxl; k < kxh; ++k) {
    T px = rel ? exact_overlap_rel(bxl, wx[i], bcx[k], bsx) : exact_overlap(bxl, bxh - bxl, bcx[k], bsx);
    for (int h = kyl; h < kyh; ++h) {
      T py = rel ? exact_overlap_rel(byl, wy[i], bcy[h], bsy)
                 : exact_overlap(byl, byh - byl, bcy[h], bsy);
      T area = px * py;
      long long v = (long long)(area * scale);
      if (v != 0) atomicAdd(&map[(int64_t)k * nby + h], (unsigned long long)v);
    }
  }
}

// CUDA-path quantisation of the initial map: u64(init * T(scale)).
template <typename T>
__global__ void k_quantize(const T* init, int64_t n, T scale, unsigned long long* buf) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    buf[i] = (unsigned long long)(init[i] * scale);
}

// CPU-path conversion: out = T(double(init) + double(sum) * (1.0/scale))
// (atomic_ops.h:48-54 scaleAdd on a clone of the initial map).
// buf2 (nullable): a second int64 map (the fillers') added to buf for `out`;
// out_part (nullable) receives the conversion of buf alone.
template <typename T>
__global__ void k_convert_cpu(const unsigned long long* buf, const unsigned long long* buf2,
                              const T* init, int64_t n, double inv_scale, T* out, T* out_part,
                              long long* raw) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const unsigned long long b = buf[i];
    long long s = (long long)(buf2 ? b + buf2[i] : b);
    double base = init ? (double)init[i] : 0.0;
    if (out) out[i] = (T)(base + (double)s * inv_scale);
    if (out_part) out_part[i] = (T)(base + (double)(long long)b * inv_scale);
    if (raw) raw[i] = s;
  }
}

// CUDA-path conversion: out = T(u64) * T(1.0/scale).  raw receives the sum
// without the quantised initial map.
template <typename T>
__global__ void k_convert_gpu(const unsigned long long* buf, const unsigned long long* buf2,
                              const T* init, int64_t n, T scale, T inv_scale, T* out,
                              T* out_part, long long* raw) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const unsigned long long b = buf[i];
    unsigned long long s = buf2 ? b + buf2[i] : b;
    if (out) out[i] = (T)s * inv_scale;
    if (out_part) out_part[i] = (T)b * inv_scale;
    if (raw) {
      unsigned long long q = init ? (unsigned long long)(init[i] * scale) : 0ull;
      raw[i] = (long long)(s - q);
    }
  }
}

// Fixed map finish (electric_density_map.cpp:205-209 clamp_max(bin area), then
// electric_overflow.py:258 mul_(target_density)), from the int64 sum.
template <typename T>
__global__ void k_fixed_finish(const unsigned long long* buf, int64_t n, int convention,
                               double inv_scale_d, T inv_scale_t, T bin_area, T target, T* out) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    T v;
    if (convention == DP_CONVERT_CPU)
      v = (T)(0.0 + (double)(long long)buf[i] * inv_scale_d);
    else
      v = (T)buf[i] * inv_scale_t;
    v = (bin_area < v) ? bin_area : v;
    out[i] = v * target;
  }
}

// Overflow metric partials: sum(max(d - target_area, 0)) and max(d), FP64.
template <typename T>
__global__ void __launch_bounds__(kReduceBlock) k_overflow_partial(const T* d, int64_t n,
                                                                   double target_area,
                                                                   double* partials) {
  __shared__ double smem[kReduceBlock / 64];
  double s = 0, m = -HUGE_VAL;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    double v = (double)d[i];
    double o = v - target_area;
    s += o > 0 ? o : 0;
    m = v > m ? v : m;
  }
  double bs = block_sum<kReduceBlock>(s, smem);
  double bm = block_max<kReduceBlock>(m, smem);
  if (threadIdx.x == 0) {
    partials[blockIdx.x] = bs;
    partials[kReduceMaxBlocks + blockIdx.x] = bm;
  }
}

// ---------------------------------------------------------------------------
// Host side.
// ---------------------------------------------------------------------------
struct DensityWorkspace {
  unsigned long long* map;
  unsigned long long* filler_map;
  int* qcount;
  int* qlist;
  char* tiled;
};

static DensityWorkspace carve(void* ws, int64_t nb, int64_t qcap) {
  char* p = (char*)ws;
  DensityWorkspace w;
  w.map = (unsigned long long*)p;
  w.filler_map = (unsigned long long*)(p + density_map_bytes(nb));
  w.qcount = (int*)(p + density_queue_offset(nb));
  w.qlist = (int*)(p + density_queue_offset(nb) + 256);
  w.tiled = p + density_tiled_offset(nb, qcap);
  return w;
}

template <typename T>
TriArgs<T> make_tri_args(const void* pos, int64_t num_nodes, const void* wxc, const void* wyc,
                         const void* offx, const void* offy, const void* ratio, double xl,
                         double yl, double bsx, double bsy, int nbx, int nby, int64_t scale,
                         const double* uniform_filler, int64_t fill_begin) {
  TriArgs<T> a;
  a.x = (const T*)pos;
  a.y = (const T*)pos + num_nodes;
  a.wxc = (const T*)wxc;
  a.wyc = (const T*)wyc;
  a.offx = (const T*)offx;
  a.offy = (const T*)offy;
  a.ratio = (const T*)ratio;
  a.xl = (T)xl;
  a.yl = (T)yl;
  a.bsx = (T)bsx;
  a.bsy = (T)bsy;
  // electric_density_map.cpp:265-266: T inv = 1.0 / bin_size (bin size already in T)
  a.inv_x = (T)(1.0 / (double)a.bsx);
  a.inv_y = (T)(1.0 / (double)a.bsy);
  a.nbx = nbx;
  a.nby = nby;
  a.scale = (T)scale;
  // the shared filler values come from arrays of T, so (T) is exact
  a.uni = uniform_filler != nullptr;
  a.fill_begin = fill_begin;
  a.uwx = a.uni ? (T)uniform_filler[0] : T(0);
  a.uwy = a.uni ? (T)uniform_filler[1] : T(0);
  a.uox = a.uni ? (T)uniform_filler[2] : T(0);
  a.uoy = a.uni ? (T)uniform_filler[3] : T(0);
  a.ur = a.uni ? (T)uniform_filler[4] : T(0);
  a.acc = 0;
  return a;
}

template TriArgs<float> make_tri_args<float>(const void*, int64_t, const void*, const void*,
                                             const void*, const void*, const void*, double,
                                             double, double, double, int, int, int64_t,
                                             const double*, int64_t);
template TriArgs<double> make_tri_args<double>(const void*, int64_t, const void*, const void*,
                                               const void*, const void*, const void*, double,
                                               double, double, double, int, int, int64_t,
                                               const double*, int64_t);

template <typename T>
static int scatter_global(const TriArgs<T>& a, int64_t n_mov, int64_t fill_begin, int64_t n_fill,
                          DensityWorkspace& w, int64_t qcap, hipStream_t s) {
  const int64_t total = n_mov + n_fill;
  if (total <= 0) return DP_OK;
  hipLaunchKernelGGL(k_fill<int>, dim3(1), dim3(64), 0, s, w.qcount, (int64_t)1, 0);
  DP_PROBED(DP_KERNEL_DENSITY_SCATTER, s,
            hipLaunchKernelGGL(k_tri_scatter<T>, dim3((unsigned)ceil_div(total, 256)), dim3(256),
                               0, s, a, n_mov, fill_begin, total, w.map, w.qcount, w.qlist,
                               (int)qcap));
  DP_LAUNCHED("k_tri_scatter");
  hipLaunchKernelGGL(k_tri_scatter_large<T>, dim3(1024), dim3(256), 0, s, a, w.map,
                     (const int*)w.qcount, (const int*)w.qlist, (int)qcap);
  DP_LAUNCHED("k_tri_scatter_large");
  return DP_OK;
}

}  // namespace dp

using namespace dp;

extern "C" {

int64_t dp_fixed_point_scale(double xl, double yl, double xh, double yh) {
  double diearea = (xh - xl) * (yh - yl);
  int integer_bits = (int)ceil(log2(diearea)) + 1;
  if (integer_bits < 32) integer_bits = 32;
  int fraction_bits = 64 - integer_bits;
  if (fraction_bits < 0) fraction_bits = 0;
  return (int64_t)1 << fraction_bits;
}

size_t dp_density_map_workspace_bytes(int64_t num_bins_x, int64_t num_bins_y, int64_t num_nodes,
                                      int64_t large_capacity, int algorithm, int dtype) {
  const int64_t nb = num_bins_x * num_bins_y;
  size_t b = density_tiled_offset(nb, large_capacity);
  const int alg = algorithm & ~DP_ARITH_BIN_RELATIVE;
  if (alg == DP_SCATTER_TILED && tiled_supported((int)num_bins_x, (int)num_bins_y))
    b += tiled_ws_bytes((int)num_bins_x, (int)num_bins_y, num_nodes, dtype,
                        (algorithm & DP_ARITH_BIN_RELATIVE) != 0);
  return b;
}

// dp_density_map, and with movable_map_out the map of the fixed and movable
// cells alone (the fillers scattered to their own int64 map) from the same
// scatter.
static int density_map_impl(const char* name, int dtype, const void* pos, int64_t num_nodes,
                            const void* node_size_x_clamped, const void* node_size_y_clamped,
                            const void* offset_x, const void* offset_y, const void* ratio,
                            const void* initial_density_map, double xl, double yl, double xh,
                            double yh, double bin_size_x, double bin_size_y,
                            int64_t num_movable_nodes, int64_t num_filler_nodes, int num_bins_x,
                            int num_bins_y, int convention, int algorithm,
                            int64_t large_capacity, const double* uniform_filler,
                            const int32_t* filler_tile_start, void* workspace,
                            void* density_map_out, int64_t* fixed_map_out,
                            void* movable_map_out, void* stream) {
  DP_ARG(pos && node_size_x_clamped && node_size_y_clamped && offset_x && offset_y && ratio,
         "dp_density_map: null input array");
  DP_ARG(workspace, "dp_density_map: null workspace");
  DP_ARG(num_bins_x > 0 && num_bins_y > 0, "dp_density_map: bad bin counts %d x %d", num_bins_x,
         num_bins_y);
  DP_ARG(num_movable_nodes >= 0 && num_filler_nodes >= 0 &&
             num_movable_nodes + num_filler_nodes <= num_nodes,
         "dp_density_map: bad node counts");
  DP_ARG(num_nodes < (int64_t)1 << 31, "dp_density_map: more than 2^31 nodes");
  DP_ARG(large_capacity >= 0 && large_capacity < (int64_t)1 << 31,
         "dp_density_map: bad large_capacity");
  DP_ARG(convention == DP_CONVERT_CPU || convention == DP_CONVERT_GPU,
         "dp_density_map: bad convention %d", convention);
  const int acc = (algorithm & DP_ARITH_BIN_RELATIVE) != 0;
  algorithm &= ~DP_ARITH_BIN_RELATIVE;
  DP_ARG(algorithm == DP_SCATTER_GLOBAL || algorithm == DP_SCATTER_TILED,
         "dp_density_map: bad algorithm %d", algorithm);
  DP_ARG(density_map_out || fixed_map_out || movable_map_out, "dp_density_map: no output");
  hipStream_t s = S(stream);
  DP_PROBE_OP(DP_OP_DENSITY_MAP, s);
  const int64_t nb = (int64_t)num_bins_x * num_bins_y;
  const int64_t scale = dp_fixed_point_scale(xl, yl, xh, yh);
  DensityWorkspace w = carve(workspace, nb, large_capacity);
  return DP_DISPATCH(dtype, "dp_density_map", [&] {
    TriArgs<T> a = make_tri_args<T>(pos, num_nodes, node_size_x_clamped, node_size_y_clamped,
                                    offset_x, offset_y, ratio, xl, yl, bin_size_x, bin_size_y,
                                    num_bins_x, num_bins_y, scale, uniform_filler,
                                    num_nodes - num_filler_nodes);
    a.acc = acc;
    const T* init = (const T*)initial_density_map;
    const bool tiled = algorithm == DP_SCATTER_TILED && tiled_supported(num_bins_x, num_bins_y);
    const bool quantize = convention == DP_CONVERT_GPU && init;
    if (quantize) {
      hipLaunchKernelGGL(k_quantize<T>, dim3(grid_for(nb, 256)), dim3(256), 0, s, init, nb,
                         (T)scale, w.map);
    } else if (!tiled) {
      zero_words(w.map, nb, s);  // the tiled scatter clears its maps itself
    }
    DP_LAUNCHED("density init");
    unsigned long long* fmap = nullptr;
    if (movable_map_out && num_filler_nodes > 0) {
      // the fillers' contributions in a map of their own: w.map then holds the
      // fixed and movable cells' alone, the map ElectricOverflow forms
      // (cleared by the tiled scatter, the only path that takes it)
      fmap = w.filler_map;
    }
    if (tiled) {
      DP_CHECK_RC(scatter_tiled<T>(a, num_movable_nodes, num_nodes - num_filler_nodes,
                                   num_filler_nodes, w.tiled, w.map, fmap,
                                   (const int*)filler_tile_start, !quantize, s));
    } else {
      DP_REQUIRE(!fmap, DP_ERR_UNSUPPORTED,
                 "%s: the movable map needs the tiled scatter", name);
      DP_CHECK_RC(scatter_global<T>(a, num_movable_nodes, num_nodes - num_filler_nodes,
                                    num_filler_nodes, w, large_capacity, s));
    }
    if (convention == DP_CONVERT_CPU) {
      DP_PROBED(DP_KERNEL_DENSITY_CONVERT, s,
                hipLaunchKernelGGL(k_convert_cpu<T>, dim3(grid_for(nb, 256)), dim3(256), 0, s,
                                   w.map, (const unsigned long long*)fmap, init, nb,
                                   1.0 / (double)scale, (T*)density_map_out,
                                   (T*)movable_map_out, (long long*)fixed_map_out));
    } else {
      DP_PROBED(DP_KERNEL_DENSITY_CONVERT, s,
                hipLaunchKernelGGL(k_convert_gpu<T>, dim3(grid_for(nb, 256)), dim3(256), 0, s,
                                   w.map, (const unsigned long long*)fmap, init, nb, (T)scale,
                                   (T)(1.0 / (double)scale), (T*)density_map_out,
                                   (T*)movable_map_out, (long long*)fixed_map_out));
    }
    DP_LAUNCHED("density convert");
    return DP_OK;
  });
}

int dp_density_map(int dtype, const void* pos, int64_t num_nodes, const void* node_size_x_clamped,
                   const void* node_size_y_clamped, const void* offset_x, const void* offset_y,
                   const void* ratio, const void* initial_density_map, double xl, double yl,
                   double xh, double yh, double bin_size_x, double bin_size_y,
                   int64_t num_movable_nodes, int64_t num_filler_nodes, int num_bins_x,
                   int num_bins_y, int convention, int algorithm, int64_t large_capacity,
                   const double* uniform_filler, void* workspace, void* density_map_out,
                   int64_t* fixed_map_out, void* stream) {
  return density_map_impl("dp_density_map", dtype, pos, num_nodes, node_size_x_clamped,
                          node_size_y_clamped, offset_x, offset_y, ratio, initial_density_map, xl,
                          yl, xh, yh, bin_size_x, bin_size_y, num_movable_nodes, num_filler_nodes,
                          num_bins_x, num_bins_y, convention, algorithm, large_capacity,
                          uniform_filler, nullptr, workspace, density_map_out, fixed_map_out,
                          nullptr, stream);
}

int dp_density_map_ex(int dtype, const void* pos, int64_t num_nodes,
                      const void* node_size_x_clamped, const void* node_size_y_clamped,
                      const void* offset_x, const void* offset_y, const void* ratio,
                      const void* initial_density_map, double xl, double yl, double xh,
                      double yh, double bin_size_x, double bin_size_y,
                      int64_t num_movable_nodes, int64_t num_filler_nodes, int num_bins_x,
                      int num_bins_y, int convention, int algorithm, int64_t large_capacity,
                      const double* uniform_filler, const int32_t* filler_tile_start,
                      void* workspace, void* density_map_out, void* movable_map_out,
                      void* stream) {
  return density_map_impl("dp_density_map_ex", dtype, pos, num_nodes, node_size_x_clamped,
                          node_size_y_clamped, offset_x, offset_y, ratio, initial_density_map, xl,
                          yl, xh, yh, bin_size_x, bin_size_y, num_movable_nodes, num_filler_nodes,
                          num_bins_x, num_bins_y, convention, algorithm, large_capacity,
                          uniform_filler,
                          (algorithm & ~DP_ARITH_BIN_RELATIVE) == DP_SCATTER_TILED
                              ? filler_tile_start
                              : nullptr,
                          workspace,
                          density_map_out, nullptr, movable_map_out, stream);
}

int dp_fixed_density_map(int dtype, const void* pos, int64_t num_nodes, const void* node_size_x,
                         const void* node_size_y, const void* bin_center_x,
                         const void* bin_center_y, double xl, double yl, double xh, double yh,
                         double bin_size_x, double bin_size_y, int64_t num_movable_nodes,
                         int64_t num_terminals, int num_bins_x, int num_bins_y,
                         double target_density, int convention, void* workspace, void* out,
                         void* stream) {
  DP_ARG(pos && node_size_x && node_size_y && bin_center_x && bin_center_y && out && workspace,
         "dp_fixed_density_map: null pointer");
  DP_ARG(num_bins_x > 0 && num_bins_y > 0, "dp_fixed_density_map: bad bin counts");
  DP_ARG(num_movable_nodes >= 0 && num_terminals >= 0 &&
             num_movable_nodes + num_terminals <= num_nodes,
         "dp_fixed_density_map: bad node counts");
  const int rel = (convention & DP_ARITH_BIN_RELATIVE) != 0;
  convention &= ~DP_ARITH_BIN_RELATIVE;
  DP_ARG(convention == DP_CONVERT_CPU || convention == DP_CONVERT_GPU,
         "dp_fixed_density_map: bad convention");
  hipStream_t s = S(stream);
  const int64_t nb = (int64_t)num_bins_x * num_bins_y;
  const int64_t scale = dp_fixed_point_scale(xl, yl, xh, yh);
  unsigned long long* map = (unsigned long long*)workspace;
  return DP_DISPATCH(dtype, "dp_fixed_density_map", [&] {
    const T* P = (const T*)pos;
    zero_words(map, nb, s);
    if (num_terminals > 0) {
      const int64_t b = num_movable_nodes;
      hipLaunchKernelGGL(k_exact_scatter<T>, dim3((unsigned)ceil_div(num_terminals, 256)),
                         dim3(256), 0, s, P + b, P + num_nodes + b, (const T*)node_size_x + b,
                         (const T*)node_size_y + b, (const T*)bin_center_x,
                         (const T*)bin_center_y, num_terminals, num_bins_x, num_bins_y, (T)xl,
                         (T)yl, (T)xh, (T)yh, (T)bin_size_x, (T)bin_size_y, (T)scale, rel, map);
      DP_LAUNCHED("k_exact_scatter");
    }
    T bin_area = (T)(bin_size_x * bin_size_y);
    if (num_terminals == 0) bin_area = (T)HUGE_VAL;  // reference clamps only when terminals exist
    hipLaunchKernelGGL(k_fixed_finish<T>, dim3(grid_for(nb, 256)), dim3(256), 0, s,
                       (const unsigned long long*)map, nb, convention, 1.0 / (double)scale,
                       (T)(1.0 / (double)scale), bin_area, (T)target_density, (T*)out);
    DP_LAUNCHED("k_fixed_finish");
    return DP_OK;
  });
}

size_t dp_reduce_workspace_bytes(int64_t n) {
  (void)n;
  return reduce_slot_bytes();
}

int dp_density_overflow(int dtype, const void* density_map, int64_t num_bins, double target_area,
                        void* workspace, double* out2, void* stream) {
  DP_ARG(density_map && workspace && out2 && num_bins > 0, "dp_density_overflow: bad argument");
  hipStream_t s = S(stream);
  double* partials = (double*)workspace;
  int nblk = (int)(ceil_div(num_bins, kReduceBlock) < kReduceMaxBlocks
                       ? ceil_div(num_bins, kReduceBlock)
                       : kReduceMaxBlocks);
  return DP_DISPATCH(dtype, "dp_density_overflow", [&] {
    hipLaunchKernelGGL(k_overflow_partial<T>, dim3(nblk), dim3(kReduceBlock), 0, s,
                       (const T*)density_map, num_bins, target_area, partials);
    DP_LAUNCHED("k_overflow_partial");
    int is_max[2] = {0, 1};
    return finish_reduction(partials, nblk, 2, is_max, out2, -1, s);
  });
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Allocation-free density map (dp_density_map_noalloc), for callers that
// cannot allocate (code under HIP graph capture, serving loops): the
// per-cell global-atomic scatter of dp_density_map with its int64 map and
// large-cell queue in a file-scope device array, `scratch`, instead of a
// caller workspace.  Each device holds its own copy (8.25 MiB) until the
// process ends; every call on a device shares it, so calls that overlap on
// one device race.  Nothing in it is assumed zero at entry: the map and the
// queue counter are cleared (or the map seeded) by kernels, not by
// hipMemsetAsync, so the call can be captured into a graph.
// ---------------------------------------------------------------------------
#include <mutex>

namespace dp {
constexpr int64_t kNoAllocMaxBins = (int64_t)1 << 20;  // 1024 x 1024 bins
constexpr int kNoAllocQueue = 1 << 16;                 // queued cells above 3x3 bins
constexpr size_t kNoAllocCountOffset = 8 * (size_t)kNoAllocMaxBins;  // after the int64 map
constexpr size_t kNoAllocQueueOffset = kNoAllocCountOffset + 256;
constexpr size_t kNoAllocBytes = kNoAllocQueueOffset + 4 * (size_t)kNoAllocQueue;
constexpr int kNoAllocMaxDevices = 64;
}  // namespace dp

static __device__ __align__(256) unsigned char scratch[dp::kNoAllocBytes];

namespace dp {
static std::mutex g_scratch_mutex;
static void* g_scratch_ptr[kNoAllocMaxDevices];

// The device address of `scratch` on the current device, looked up once per
// device (the array's name in host code is only a stand-in for the symbol).
static int scratch_address(void** out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) {
    set_error("dp_density_map_noalloc: hipGetDevice failed: %s", hipGetErrorString(e));
    return (int)e;
  }
  DP_REQUIRE(dev >= 0 && dev < kNoAllocMaxDevices, DP_ERR_UNSUPPORTED,
             "dp_density_map_noalloc: device %d beyond the %d supported", dev,
             kNoAllocMaxDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      set_error("dp_density_map_noalloc: hipGetSymbolAddress(scratch) failed on device %d: %s",
                dev, hipGetErrorString(e));
      return e != hipSuccess ? (int)e : DP_ERR_WORKSPACE;
    }
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return DP_OK;
}
}  // namespace dp

extern "C" int dp_density_map_noalloc(int dtype, const void* pos, int64_t num_nodes,
                                      const void* node_size_x_clamped,
                                      const void* node_size_y_clamped, const void* offset_x,
                                      const void* offset_y, const void* ratio,
                                      const void* initial_density_map, double xl, double yl,
                                      double xh, double yh, double bin_size_x, double bin_size_y,
                                      int64_t num_movable_nodes, int64_t num_filler_nodes,
                                      int num_bins_x, int num_bins_y, int convention,
                                      const double* uniform_filler, void* density_map_out,
                                      int64_t* fixed_map_out, void* stream) {
  DP_ARG(pos && node_size_x_clamped && node_size_y_clamped && offset_x && offset_y && ratio,
         "dp_density_map_noalloc: null input array");
  DP_ARG(num_bins_x > 0 && num_bins_y > 0, "dp_density_map_noalloc: bad bin counts %d x %d",
         num_bins_x, num_bins_y);
  DP_REQUIRE((int64_t)num_bins_x * num_bins_y <= kNoAllocMaxBins, DP_ERR_UNSUPPORTED,
             "dp_density_map_noalloc: %d x %d bins exceed its fixed scratch of %lld bins "
             "(1024 x 1024); use dp_density_map with a workspace",
             num_bins_x, num_bins_y, (long long)kNoAllocMaxBins);
  DP_ARG(num_movable_nodes >= 0 && num_filler_nodes >= 0 &&
             num_movable_nodes + num_filler_nodes <= num_nodes,
         "dp_density_map_noalloc: bad node counts");
  DP_ARG(num_nodes < (int64_t)1 << 31, "dp_density_map_noalloc: more than 2^31 nodes");
  DP_ARG(convention == DP_CONVERT_CPU || convention == DP_CONVERT_GPU,
         "dp_density_map_noalloc: bad convention %d", convention);
  DP_ARG(density_map_out || fixed_map_out, "dp_density_map_noalloc: no output");
  void* base = nullptr;
  DP_CHECK_RC(scratch_address(&base));
  hipStream_t s = S(stream);
  const int64_t nb = (int64_t)num_bins_x * num_bins_y;
  const int64_t scale = dp_fixed_point_scale(xl, yl, xh, yh);
  DensityWorkspace w;
  w.map = (unsigned long long*)base;
  w.filler_map = nullptr;
  w.qcount = (int*)((char*)base + kNoAllocCountOffset);
  w.qlist = (int*)((char*)base + kNoAllocQueueOffset);
  w.tiled = nullptr;
  return DP_DISPATCH(dtype, "dp_density_map_noalloc", [&] {
    TriArgs<T> a = make_tri_args<T>(pos, num_nodes, node_size_x_clamped, node_size_y_clamped,
                                    offset_x, offset_y, ratio, xl, yl, bin_size_x, bin_size_y,
                                    num_bins_x, num_bins_y, scale, uniform_filler,
                                    num_nodes - num_filler_nodes);
    const T* init = (const T*)initial_density_map;
    if (convention == DP_CONVERT_GPU && init)
      hipLaunchKernelGGL(k_quantize<T>, dim3(grid_for(nb, 256)), dim3(256), 0, s, init, nb,
                         (T)scale, w.map);
    else
      zero_words(w.map, nb, s);
    DP_LAUNCHED("density init (noalloc)");
    DP_CHECK_RC(scatter_global<T>(a, num_movable_nodes, num_nodes - num_filler_nodes,
                                  num_filler_nodes, w, kNoAllocQueue, s));
    if (convention == DP_CONVERT_CPU)
      hipLaunchKernelGGL(k_convert_cpu<T>, dim3(grid_for(nb, 256)), dim3(256), 0, s, w.map,
                         (const unsigned long long*)nullptr, init, nb, 1.0 / (double)scale,
                         (T*)density_map_out, (T*)nullptr, (long long*)fixed_map_out);
    else
      hipLaunchKernelGGL(k_convert_gpu<T>, dim3(grid_for(nb, 256)), dim3(256), 0, s, w.map,
                         (const unsigned long long*)nullptr, init, nb, (T)scale,
                         (T)(1.0 / (double)scale), (T*)density_map_out, (T*)nullptr,
                         (long long*)fixed_map_out);
    DP_LAUNCHED("density convert (noalloc)");
    return DP_OK;
  });
}
