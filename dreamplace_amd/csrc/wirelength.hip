// Pin positions, merged weighted-average (WA) wirelength, HPWL.
#include "dp_common.h"

namespace dp {

template <typename T>
__device__ __forceinline__ T tmax_value();
template <>
__device__ __forceinline__ float tmax_value<float>() { return 3.402823466e+38f; }
template <>
__device__ __forceinline__ double tmax_value<double>() { return 1.7976931348623157e+308; }

// pin_x[p] = off_x[p] + x[pin2node[p]]  (pin_pos_cuda_kernel.cu:11-26)
template <typename T, typename I>
__global__ void k_pin_pos_fwd(const T* __restrict__ x, const T* __restrict__ y,
                              const T* __restrict__ offx, const T* __restrict__ offy,
                              const I* __restrict__ pin2node, int64_t P, T* pin_x, T* pin_y) {
  int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P) return;
  int64_t nd = (int64_t)pin2node[p];
  pin_x[p] = offx[p] + x[nd];
  pin_y[p] = offy[p] + y[nd];
}

// grad[n] = sum over the CSR pins of node n (pin_pos_cuda_kernel.cu:50-70);
// nodes >= num_physical (fillers) get 0.
template <typename T>
__global__ void k_pin_pos_bwd(const T* __restrict__ gpx, const T* __restrict__ gpy,
                              const int* __restrict__ flat_node2pin,
                              const int* __restrict__ node2pin_start, int64_t nphys,
                              int64_t nnodes, T* gx, T* gy) {
  int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= nnodes) return;
  T sx = 0, sy = 0;
  if (n < nphys) {
    for (int j = node2pin_start[n]; j < node2pin_start[n + 1]; ++j) {
      int p = flat_node2pin[j];
      sx += gpx[p];
      sy += gpy[p];
    }
  }
  gx[n] = sx;
  gy[n] = sy;
}

// One net of the merged WA model, both axes: the extreme pin coordinates
// (pass 1) and the max-shifted exponential sums (pass 2),
// weighted_average_wirelength_merged.cpp:49-92.
template <typename T>
struct WaNet {
  T xmax, xmin, ymax, ymin;
  T xs, xns, es, ens, ys, yns, fs, fns;
  __device__ __forceinline__ T wl() const { return xs / es - xns / ens + ys / fs - yns / fns; }
};

template <typename T>
__device__ __forceinline__ WaNet<T> wa_net(const T* __restrict__ px, const T* __restrict__ py,
                                           const int* __restrict__ flat_netpin, int b, int en,
                                           T ig) {
  WaNet<T> n;
  n.xmax = n.ymax = -tmax_value<T>();
  n.xmin = n.ymin = tmax_value<T>();
  for (int j = b; j < en; ++j) {
    int p = flat_netpin[j];
    T xx = px[p], yy = py[p];
    n.xmax = xx > n.xmax ? xx : n.xmax;
    n.xmin = xx < n.xmin ? xx : n.xmin;
    n.ymax = yy > n.ymax ? yy : n.ymax;
    n.ymin = yy < n.ymin ? yy : n.ymin;
  }
  n.xs = n.xns = n.es = n.ens = n.ys = n.yns = n.fs = n.fns = 0;
  for (int j = b; j < en; ++j) {
    int p = flat_netpin[j];
    T xx = px[p], yy = py[p];
    T ex = exp((xx - n.xmax) * ig), enx = exp((n.xmin - xx) * ig);
    n.xs += xx * ex;
    n.xns += xx * enx;
    n.es += ex;
    n.ens += enx;
    T ey = exp((yy - n.ymax) * ig), eny = exp((n.ymin - yy) * ig);
    n.ys += yy * ey;
    n.yns += yy * eny;
    n.fs += ey;
    n.fns += eny;
  }
  return n;
}

// Pass 3: d(WL_n)/dx_p = (a+ + b+ x) e+ - (a- + b- x) e- per axis, with
// b+ = (1/gamma)/S+, a+ = (1 - b+ XS+)/S+, b- = -(1/gamma)/S-,
// a- = (1 - b- XS-)/S- (weighted_average_wirelength_merged.cpp:94-116).
template <typename T>
struct WaDeriv {
  T xmax, xmin, ymax, ymin, ig;
  T bx, ax, bnx, anx, by, ay, bny, any;
  __device__ __forceinline__ WaDeriv(const WaNet<T>& n, T ig_) {
    xmax = n.xmax; xmin = n.xmin; ymax = n.ymax; ymin = n.ymin; ig = ig_;
    bx = ig / n.es; ax = (T)((1.0 - bx * n.xs) / n.es);
    bnx = -ig / n.ens; anx = (T)((1.0 - bnx * n.xns) / n.ens);
    by = ig / n.fs; ay = (T)((1.0 - by * n.ys) / n.fs);
    bny = -ig / n.fns; any = (T)((1.0 - bny * n.yns) / n.fns);
  }
  __device__ __forceinline__ void at(T xx, T yy, T& gx, T& gy) const {
    T ex = exp((xx - xmax) * ig), enx = exp((xmin - xx) * ig);
    gx = (ax + bx * xx) * ex - (anx + bnx * xx) * enx;
    T ey = exp((yy - ymax) * ig), eny = exp((ymin - yy) * ig);
    gy = (ay + by * yy) * ey - (any + bny * yy) * eny;
  }
};

// Merged WA, one thread per net (both axes): passes 1-2, then (store form)
// pass 3 writes the per-pin derivative into gix/giy
// (weighted_average_wirelength_merged.cpp:31-116).  With gix == nullptr (the
// recompute form) nothing per pin is written.  Block partials of
// sum_n w_n*WL_n (FP64) feed the deterministic final reduction.
template <typename T>
__global__ void __launch_bounds__(256) k_wa_fwd(const T* __restrict__ px, const T* __restrict__ py,
                                                const int* __restrict__ flat_netpin,
                                                const int* __restrict__ netpin_start, int64_t E,
                                                const T* __restrict__ w,
                                                const uint8_t* __restrict__ mask,
                                                const T* __restrict__ inv_gamma_p, T* partial,
                                                T* gix, T* giy, double* partials) {
  __shared__ double smem[4];
  double contrib = 0;
  const T ig = *inv_gamma_p;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < E;
       e += (int64_t)gridDim.x * blockDim.x) {
    if (!mask[e]) {
      if (partial) partial[e] = 0;
      continue;
    }
    const int b = netpin_start[e], en = netpin_start[e + 1];
    const WaNet<T> n = wa_net(px, py, flat_netpin, b, en, ig);
    T wl = n.wl();
    if (partial) partial[e] = wl;
    contrib += (double)(w ? wl * w[e] : wl);
    if (gix) {
      const WaDeriv<T> d(n, ig);
      for (int j = b; j < en; ++j) {
        int p = flat_netpin[j];
        d.at(px[p], py[p], gix[p], giy[p]);
      }
    }
  }
  double bs = block_sum<256>(contrib, smem);
  if (threadIdx.x == 0) partials[blockIdx.x] = bs;
}

// Backward of the recompute form: passes 1-3 again from the pin positions,
// and the pin gradient written directly as ((dWL_n/dx_p) * grad) * w_n, the
// store form's k_wa_bwd arithmetic on the same derivative; pins under
// pin_mask get 0.  Pins of masked nets are not written (the caller's output
// holds zeros there, as the store form's derivative buffer does).
template <typename T>
__global__ void __launch_bounds__(256) k_wa_bwd_recompute(
    const T* __restrict__ px, const T* __restrict__ py, const int* __restrict__ flat_netpin,
    const int* __restrict__ netpin_start, int64_t E, const T* __restrict__ w,
    const uint8_t* __restrict__ mask, const uint8_t* __restrict__ pin_mask,
    const T* __restrict__ inv_gamma_p, const T* __restrict__ grad, T* gx_out, T* gy_out) {
  const T ig = *inv_gamma_p;
  const T g = *grad;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < E;
       e += (int64_t)gridDim.x * blockDim.x) {
    if (!mask[e]) continue;
    const int b = netpin_start[e], en = netpin_start[e + 1];
    const WaDeriv<T> d(wa_net(px, py, flat_netpin, b, en, ig), ig);
    for (int j = b; j < en; ++j) {
      int p = flat_netpin[j];
      T gx, gy;
      d.at(px[p], py[p], gx, gy);
      gx = gx * g;
      gy = gy * g;
      if (w) {
        T wn = w[e];
        gx *= wn;
        gy *= wn;
      }
      if (pin_mask && pin_mask[p]) {
        gx = 0;
        gy = 0;
      }
      gx_out[p] = gx;
      gy_out[p] = gy;
    }
  }
}

// Merged log-sum-exp wirelength, one thread per net (both axes)
// (logsumexp_wirelength_merged.cpp:14-85): pass 1 max/min, pass 2 the
// max-shifted exponential sums, pass 3 d/dx = e+/S+ - e-/S- per pin.
// WL = (log S+ + log S-)*gamma + max - min per axis, in the reference's order.
template <typename T>
__global__ void __launch_bounds__(256) k_lse_fwd(const T* __restrict__ px, const T* __restrict__ py,
                                                 const int* __restrict__ flat_netpin,
                                                 const int* __restrict__ netpin_start, int64_t E,
                                                 const T* __restrict__ w,
                                                 const uint8_t* __restrict__ mask,
                                                 const T* __restrict__ gamma_p, T* partial, T* gix,
                                                 T* giy, double* partials) {
  __shared__ double smem[4];
  double contrib = 0;
  const T gam = *gamma_p;
  const T ig = T(1) / gam;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < E;
       e += (int64_t)gridDim.x * blockDim.x) {
    if (!mask[e]) {
      if (partial) partial[e] = 0;
      continue;
    }
    const int b = netpin_start[e], en = netpin_start[e + 1];
    T xmax = -tmax_value<T>(), xmin = tmax_value<T>();
    T ymax = -tmax_value<T>(), ymin = tmax_value<T>();
    for (int j = b; j < en; ++j) {
      int p = flat_netpin[j];
      T xx = px[p], yy = py[p];
      xmax = xx > xmax ? xx : xmax;
      xmin = xx < xmin ? xx : xmin;
      ymax = yy > ymax ? yy : ymax;
      ymin = yy < ymin ? yy : ymin;
    }
    T es = 0, ens = 0, fs = 0, fns = 0;
    for (int j = b; j < en; ++j) {
      int p = flat_netpin[j];
      T xx = px[p], yy = py[p];
      es += exp((xx - xmax) * ig);
      ens += exp((xmin - xx) * ig);
      fs += exp((yy - ymax) * ig);
      fns += exp((ymin - yy) * ig);
    }
    T wl = (log(es) + log(ens)) * gam + xmax - xmin + (log(fs) + log(fns)) * gam + ymax - ymin;
    if (partial) partial[e] = wl;
    contrib += (double)(w ? wl * w[e] : wl);
    const T rx = (T)(1.0 / es), rnx = (T)(1.0 / ens), ry = (T)(1.0 / fs), rny = (T)(1.0 / fns);
    for (int j = b; j < en; ++j) {
      int p = flat_netpin[j];
      T xx = px[p], yy = py[p];
      gix[p] = exp((xx - xmax) * ig) * rx - exp((xmin - xx) * ig) * rnx;
      giy[p] = exp((yy - ymax) * ig) * ry - exp((ymin - yy) * ig) * rny;
    }
  }
  double bs = block_sum<256>(contrib, smem);
  if (threadIdx.x == 0) partials[blockIdx.x] = bs;
}

// g[p] = (gi[p]*grad)*w[net(p)] for unmasked nets; 0 where pin_mask
// (weighted_average_wirelength_cuda_merged.cpp:135, integrateNetWeights,
// weighted_average_wirelength.py:207-208).
template <typename T>
__global__ void k_wa_bwd(const T* __restrict__ grad, T* gi, int64_t P,
                         const int* __restrict__ pin2net, const T* __restrict__ w,
                         const uint8_t* __restrict__ net_mask, const uint8_t* __restrict__ pin_mask) {
  int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P) return;
  const T g = *grad;
  T gx = gi[p] * g, gy = gi[P + p] * g;
  if (w) {
    int n = pin2net[p];
    if (net_mask[n]) {
      T wn = w[n];
      gx *= wn;
      gy *= wn;
    }
  }
  if (pin_mask && pin_mask[p]) {
    gx = 0;
    gy = 0;
  }
  gi[p] = gx;
  gi[P + p] = gy;
}

template <typename T>
__global__ void k_wa_wgrad(const T* __restrict__ grad, const T* __restrict__ partial,
                           const uint8_t* __restrict__ mask, int64_t E, T* out) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= E) return;
  out[e] = mask[e] ? (*grad) * partial[e] : T(0);
}

// Per-net HPWL (hpwl_cuda_kernel.cu:10-39, both axes in one thread), times the
// net weight in T, summed in FP64.
template <typename T>
__global__ void __launch_bounds__(256) k_hpwl(const T* __restrict__ px, const T* __restrict__ py,
                                              const int* __restrict__ flat_netpin,
                                              const int* __restrict__ netpin_start, int64_t E,
                                              const T* __restrict__ w,
                                              const uint8_t* __restrict__ mask, double* partials) {
  __shared__ double smem[4];
  double acc = 0;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < E;
       e += (int64_t)gridDim.x * blockDim.x) {
    if (!mask[e]) continue;
    const int b = netpin_start[e], en = netpin_start[e + 1];
    if (en <= b) continue;
    T x0 = px[flat_netpin[b]], x1 = x0, y0 = py[flat_netpin[b]], y1 = y0;
    for (int j = b + 1; j < en; ++j) {
      int p = flat_netpin[j];
      T xx = px[p], yy = py[p];
      x0 = xx < x0 ? xx : x0;
      x1 = xx > x1 ? xx : x1;
      y0 = yy < y0 ? yy : y0;
      y1 = yy > y1 ? yy : y1;
    }
    T h = x1 - x0 + y1 - y0;
    acc += (double)(w ? h * w[e] : h);
  }
  double bs = block_sum<256>(acc, smem);
  if (threadIdx.x == 0) partials[blockIdx.x] = bs;
}

// Fused pin_pos + merged WA forward: the same per-net arithmetic as k_wa_fwd,
// with each pin position formed on the fly as offset[p] + pos[pin2node[p]].
template <typename T>
__global__ void __launch_bounds__(256) k_wa_fwd_nodes(
    const T* __restrict__ x, const T* __restrict__ y, const T* __restrict__ offx,
    const T* __restrict__ offy, const int* __restrict__ pin2node,
    const int* __restrict__ flat_netpin, const int* __restrict__ netpin_start, int64_t E,
    const T* __restrict__ w, const uint8_t* __restrict__ mask, const T* __restrict__ inv_gamma_p,
    T* partial, T* gix, T* giy, double* partials) {
  __shared__ double smem[4];
  double contrib = 0;
  const T ig = *inv_gamma_p;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < E;
       e += (int64_t)gridDim.x * blockDim.x) {
    if (!mask[e]) {
      if (partial) partial[e] = 0;
      continue;
    }
    const int b = netpin_start[e], en = netpin_start[e + 1];
    T xmax = -tmax_value<T>(), xmin = tmax_value<T>();
    T ymax = -tmax_value<T>(), ymin = tmax_value<T>();
    for (int j = b; j < en; ++j) {
      int p = flat_netpin[j];
      int nd = pin2node[p];
      T xx = offx[p] + x[nd], yy = offy[p] + y[nd];
      xmax = xx > xmax ? xx : xmax;
      xmin = xx < xmin ? xx : xmin;
      ymax = yy > ymax ? yy : ymax;
      ymin = yy < ymin ? yy : ymin;
    }
    T xs = 0, xns = 0, es = 0, ens = 0, ys = 0, yns = 0, fs = 0, fns = 0;
    for (int j = b; j < en; ++j) {
      int p = flat_netpin[j];
      int nd = pin2node[p];
      T xx = offx[p] + x[nd], yy = offy[p] + y[nd];
      T ex = exp((xx - xmax) * ig), enx = exp((xmin - xx) * ig);
      xs += xx * ex;
      xns += xx * enx;
      es += ex;
      ens += enx;
      T ey = exp((yy - ymax) * ig), eny = exp((ymin - yy) * ig);
      ys += yy * ey;
      yns += yy * eny;
      fs += ey;
      fns += eny;
    }
    T wl = xs / es - xns / ens + ys / fs - yns / fns;
    if (partial) partial[e] = wl;
    contrib += (double)(w ? wl * w[e] : wl);
    T bx = ig / es, ax = (T)((1.0 - bx * xs) / es);
    T bnx = -ig / ens, anx = (T)((1.0 - bnx * xns) / ens);
    T by = ig / fs, ay = (T)((1.0 - by * ys) / fs);
    T bny = -ig / fns, any = (T)((1.0 - bny * yns) / fns);
    for (int j = b; j < en; ++j) {
      int p = flat_netpin[j];
      int nd = pin2node[p];
      T xx = offx[p] + x[nd], yy = offy[p] + y[nd];
      T ex = exp((xx - xmax) * ig), enx = exp((xmin - xx) * ig);
      gix[p] = (ax + bx * xx) * ex - (anx + bnx * xx) * enx;
      T ey = exp((yy - ymax) * ig), eny = exp((ymin - yy) * ig);
      giy[p] = (ay + by * yy) * ey - (any + bny * yy) * eny;
    }
  }
  double bs = block_sum<256>(contrib, smem);
  if (threadIdx.x == 0) partials[blockIdx.x] = bs;
}

// Fused WA backward + pin_pos backward + gradient combine, node-parallel:
// g[n] = sum over the CSR pins p of n of (pin_mask[p] ? 0 : (gi[p]*grad)*w[net(p)])
//        + add[n]
// Per pin this is the arithmetic of k_wa_bwd, and the pins are summed in CSR
// order as k_pin_pos_bwd does, so the result equals the unfused sequence.
template <typename T>
__global__ void k_wa_pin_pos_bwd(const T* __restrict__ gix, const T* __restrict__ giy,
                                 const T* __restrict__ grad, const int* __restrict__ flat_node2pin,
                                 const int* __restrict__ node2pin_start,
                                 const int* __restrict__ pin2net, const T* __restrict__ w,
                                 const uint8_t* __restrict__ net_mask,
                                 const uint8_t* __restrict__ pin_mask, const T* __restrict__ add,
                                 int64_t nphys, int64_t nnodes, T* gx, T* gy) {
  int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= nnodes) return;
  const T g = *grad;
  T sx = 0, sy = 0;
  if (n < nphys) {
    for (int j = node2pin_start[n]; j < node2pin_start[n + 1]; ++j) {
      int p = flat_node2pin[j];
      if (pin_mask && pin_mask[p]) {
        sx += T(0);
        sy += T(0);
        continue;
      }
      T px = gix[p] * g, py = giy[p] * g;
      if (w) {
        int e = pin2net[p];
        if (net_mask[e]) {
          T wn = w[e];
          px *= wn;
          py *= wn;
        }
      }
      sx += px;
      sy += py;
    }
  }
  if (add) {
    sx += add[n];
    sy += add[nnodes + n];
  }
  gx[n] = sx;
  gy[n] = sy;
}

// k_hpwl's weighted FP64 sum over per-net HPWL values computed elsewhere
// (the pin-parallel WA pass): the same grid, loop order, mask test and
// empty-net test, so the total is bit-identical to k_hpwl's on the positions
// the values came from.
template <typename T>
__global__ void __launch_bounds__(256) k_hpwl_nets(const T* __restrict__ h_net,
                                                   const int* __restrict__ netpin_start, int64_t E,
                                                   const T* __restrict__ w,
                                                   const uint8_t* __restrict__ mask,
                                                   double* partials) {
  __shared__ double smem[4];
  double acc = 0;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < E;
       e += (int64_t)gridDim.x * blockDim.x) {
    if (!mask[e]) continue;
    if (netpin_start[e + 1] <= netpin_start[e]) continue;
    const T h = h_net[e];
    acc += (double)(w ? h * w[e] : h);
  }
  double bs = block_sum<256>(acc, smem);
  if (threadIdx.x == 0) partials[blockIdx.x] = bs;
}

static int reduce_grid(int64_t n) {
  int64_t g = ceil_div(n, 256);
  if (g < 1) g = 1;
  return (int)(g > kReduceMaxBlocks ? kReduceMaxBlocks : g);
}

}  // namespace dp

using namespace dp;

extern "C" {

int dp_pin_pos_forward(int dtype, const void* pos, int64_t num_nodes, const void* pin_offset_x,
                       const void* pin_offset_y, const void* pin2node_map, int index_type,
                       int64_t num_pins, void* pin_pos_out, void* stream) {
  DP_ARG(pos && pin_offset_x && pin_offset_y && pin2node_map && pin_pos_out,
         "dp_pin_pos_forward: null pointer");
  DP_ARG(index_type == DP_INDEX_INT32 || index_type == DP_INDEX_INT64,
         "dp_pin_pos_forward: bad index type");
  if (num_pins == 0) return DP_OK;
  hipStream_t s = S(stream);
  return DP_DISPATCH(dtype, "dp_pin_pos_forward", [&] {
    const T* P = (const T*)pos;
    T* out = (T*)pin_pos_out;
    dim3 g((unsigned)ceil_div(num_pins, 256)), b(256);
    if (index_type == DP_INDEX_INT32)
      DP_PROBED(DP_KERNEL_PIN_POS_FORWARD, s,
                hipLaunchKernelGGL((k_pin_pos_fwd<T, int>), g, b, 0, s, P, P + num_nodes,
                                   (const T*)pin_offset_x, (const T*)pin_offset_y,
                                   (const int*)pin2node_map, num_pins, out, out + num_pins));
    else
      DP_PROBED(DP_KERNEL_PIN_POS_FORWARD, s,
                hipLaunchKernelGGL((k_pin_pos_fwd<T, long long>), g, b, 0, s, P, P + num_nodes,
                                   (const T*)pin_offset_x, (const T*)pin_offset_y,
                                   (const long long*)pin2node_map, num_pins, out,
                                   out + num_pins));
    DP_LAUNCHED("k_pin_pos_fwd");
    return DP_OK;
  });
}

int dp_pin_pos_backward(int dtype, const void* grad_pin_pos, int64_t num_pins,
                        const int32_t* flat_node2pin_map, const int32_t* flat_node2pin_start_map,
                        int64_t num_physical_nodes, int64_t num_nodes, void* grad_pos_out,
                        void* stream) {
  DP_ARG(grad_pin_pos && flat_node2pin_map && flat_node2pin_start_map && grad_pos_out,
         "dp_pin_pos_backward: null pointer");
  DP_ARG(num_physical_nodes <= num_nodes, "dp_pin_pos_backward: num_physical_nodes > num_nodes");
  if (num_nodes == 0) return DP_OK;
  hipStream_t s = S(stream);
  return DP_DISPATCH(dtype, "dp_pin_pos_backward", [&] {
    const T* g = (const T*)grad_pin_pos;
    T* out = (T*)grad_pos_out;
    DP_PROBED(DP_KERNEL_PIN_POS_BACKWARD, s,
              hipLaunchKernelGGL(k_pin_pos_bwd<T>, dim3((unsigned)ceil_div(num_nodes, 256)),
                                 dim3(256), 0, s, g, g + num_pins, (const int*)flat_node2pin_map,
                                 (const int*)flat_node2pin_start_map, num_physical_nodes,
                                 num_nodes, out, out + num_nodes));
    DP_LAUNCHED("k_pin_pos_bwd");
    return DP_OK;
  });
}

size_t dp_wa_workspace_bytes(int64_t num_nets) {
  (void)num_nets;
  return reduce_slot_bytes();
}

int dp_wa_merged_forward(int dtype, const void* pin_pos, int64_t num_pins,
                         const int32_t* flat_netpin, const int32_t* netpin_start,
                         int64_t num_nets, const void* net_weights, const uint8_t* net_mask,
                         const void* inv_gamma, void* partial_wl, void* grad_intermediate,
                         void* wl_out, void* workspace, void* stream) {
  DP_ARG(pin_pos && flat_netpin && netpin_start && net_mask && inv_gamma && wl_out && workspace,
         "dp_wa_merged_forward: null pointer");
  hipStream_t s = S(stream);
  DP_PROBE_OP(DP_OP_WA_FORWARD, s);
  double* partials = (double*)workspace;
  int grid = reduce_grid(num_nets);
  return DP_DISPATCH(dtype, "dp_wa_merged_forward", [&] {
    const T* pp = (const T*)pin_pos;
    T* gi = (T*)grad_intermediate;
    DP_PROBED(DP_KERNEL_WA_FORWARD, s,
              hipLaunchKernelGGL(k_wa_fwd<T>, dim3(grid), dim3(256), 0, s, pp, pp + num_pins,
                                 (const int*)flat_netpin, (const int*)netpin_start, num_nets,
                                 (const T*)net_weights, net_mask, (const T*)inv_gamma,
                                 (T*)partial_wl, gi, gi ? gi + num_pins : nullptr, partials));
    DP_LAUNCHED("k_wa_fwd");
    return finish_reduction(partials, grid, 1, nullptr, wl_out, dtype, s);
  });
}

int dp_lse_merged_forward(int dtype, const void* pin_pos, int64_t num_pins,
                          const int32_t* flat_netpin, const int32_t* netpin_start,
                          int64_t num_nets, const void* net_weights, const uint8_t* net_mask,
                          const void* gamma, void* partial_wl, void* grad_intermediate,
                          void* wl_out, void* workspace, void* stream) {
  DP_ARG(pin_pos && flat_netpin && netpin_start && net_mask && gamma && grad_intermediate &&
             wl_out && workspace,
         "dp_lse_merged_forward: null pointer");
  hipStream_t s = S(stream);
  DP_PROBE_OP(DP_OP_WA_FORWARD, s);
  double* partials = (double*)workspace;
  int grid = reduce_grid(num_nets);
  return DP_DISPATCH(dtype, "dp_lse_merged_forward", [&] {
    const T* pp = (const T*)pin_pos;
    T* gi = (T*)grad_intermediate;
    DP_PROBED(DP_KERNEL_WA_FORWARD, s,
              hipLaunchKernelGGL(k_lse_fwd<T>, dim3(grid), dim3(256), 0, s, pp, pp + num_pins,
                                 (const int*)flat_netpin, (const int*)netpin_start, num_nets,
                                 (const T*)net_weights, net_mask, (const T*)gamma,
                                 (T*)partial_wl, gi, gi + num_pins, partials));
    DP_LAUNCHED("k_lse_fwd");
    return finish_reduction(partials, grid, 1, nullptr, wl_out, dtype, s);
  });
}

int dp_wa_merged_backward(int dtype, const void* grad, void* grad_intermediate, int64_t num_pins,
                          const int32_t* pin2net_map, const void* net_weights,
                          const uint8_t* net_mask, const uint8_t* pin_mask, void* stream) {
  DP_ARG(grad && grad_intermediate, "dp_wa_merged_backward: null pointer");
  DP_ARG(!net_weights || (pin2net_map && net_mask),
         "dp_wa_merged_backward: net weights need pin2net_map and net_mask");
  if (num_pins == 0) return DP_OK;
  hipStream_t s = S(stream);
  return DP_DISPATCH(dtype, "dp_wa_merged_backward", [&] {
    DP_PROBED(DP_KERNEL_WA_BACKWARD, s,
              hipLaunchKernelGGL(k_wa_bwd<T>, dim3((unsigned)ceil_div(num_pins, 256)), dim3(256),
                                 0, s, (const T*)grad, (T*)grad_intermediate, num_pins,
                                 (const int*)pin2net_map, (const T*)net_weights, net_mask,
                                 pin_mask));
    DP_LAUNCHED("k_wa_bwd");
    return DP_OK;
  });
}

int dp_wa_merged_backward_recompute(int dtype, const void* grad, const void* pin_pos,
                                    int64_t num_pins, const int32_t* flat_netpin,
                                    const int32_t* netpin_start, int64_t num_nets,
                                    const void* net_weights, const uint8_t* net_mask,
                                    const uint8_t* pin_mask, const void* inv_gamma,
                                    void* grad_pin_pos, void* stream) {
  DP_ARG(grad && pin_pos && flat_netpin && netpin_start && net_mask && inv_gamma && grad_pin_pos,
         "dp_wa_merged_backward_recompute: null pointer");
  if (num_nets == 0) return DP_OK;
  hipStream_t s = S(stream);
  int grid = reduce_grid(num_nets);
  return DP_DISPATCH(dtype, "dp_wa_merged_backward_recompute", [&] {
    const T* pp = (const T*)pin_pos;
    T* g = (T*)grad_pin_pos;
    DP_PROBED(DP_KERNEL_WA_BACKWARD, s,
              hipLaunchKernelGGL(k_wa_bwd_recompute<T>, dim3(grid), dim3(256), 0, s, pp,
                                 pp + num_pins, (const int*)flat_netpin,
                                 (const int*)netpin_start, num_nets, (const T*)net_weights,
                                 net_mask, pin_mask, (const T*)inv_gamma, (const T*)grad, g,
                                 g + num_pins));
    DP_LAUNCHED("k_wa_bwd_recompute");
    return DP_OK;
  });
}

int dp_wa_net_weight_grad(int dtype, const void* grad, const void* partial_wl,
                          const uint8_t* net_mask, int64_t num_nets, void* grad_net_weights,
                          void* stream) {
  DP_ARG(grad && partial_wl && net_mask && grad_net_weights, "dp_wa_net_weight_grad: null pointer");
  if (num_nets == 0) return DP_OK;
  hipStream_t s = S(stream);
  return DP_DISPATCH(dtype, "dp_wa_net_weight_grad", [&] {
    hipLaunchKernelGGL(k_wa_wgrad<T>, dim3((unsigned)ceil_div(num_nets, 256)), dim3(256), 0, s,
                       (const T*)grad, (const T*)partial_wl, net_mask, num_nets,
                       (T*)grad_net_weights);
    DP_LAUNCHED("k_wa_wgrad");
    return DP_OK;
  });
}

int dp_wa_forward_from_nodes(int dtype, const void* pos, int64_t num_nodes,
                             const void* pin_offset_x, const void* pin_offset_y,
                             const int32_t* pin2node_map, int64_t num_pins,
                             const int32_t* flat_netpin, const int32_t* netpin_start,
                             int64_t num_nets, const void* net_weights, const uint8_t* net_mask,
                             const void* inv_gamma, void* partial_wl, void* grad_intermediate,
                             void* wl_out, void* workspace, void* stream) {
  DP_ARG(pos && pin_offset_x && pin_offset_y && pin2node_map && flat_netpin && netpin_start &&
             net_mask && inv_gamma && grad_intermediate && wl_out && workspace,
         "dp_wa_forward_from_nodes: null pointer");
  hipStream_t s = S(stream);
  DP_PROBE_OP(DP_OP_WA_FORWARD, s);
  double* partials = (double*)workspace;
  int grid = reduce_grid(num_nets);
  return DP_DISPATCH(dtype, "dp_wa_forward_from_nodes", [&] {
    const T* P = (const T*)pos;
    T* gi = (T*)grad_intermediate;
    DP_PROBED(DP_KERNEL_WA_FORWARD, s,
              hipLaunchKernelGGL(k_wa_fwd_nodes<T>, dim3(grid), dim3(256), 0, s, P, P + num_nodes,
                                 (const T*)pin_offset_x, (const T*)pin_offset_y,
                                 (const int*)pin2node_map, (const int*)flat_netpin,
                                 (const int*)netpin_start, num_nets, (const T*)net_weights,
                                 net_mask, (const T*)inv_gamma, (T*)partial_wl, gi, gi + num_pins,
                                 partials));
    DP_LAUNCHED("k_wa_fwd_nodes");
    return finish_reduction(partials, grid, 1, nullptr, wl_out, dtype, s);
  });
}

int dp_wa_backward_to_nodes(int dtype, const void* grad, const void* grad_intermediate,
                            int64_t num_pins, const int32_t* flat_node2pin_map,
                            const int32_t* flat_node2pin_start_map, const int32_t* pin2net_map,
                            const void* net_weights, const uint8_t* net_mask,
                            const uint8_t* pin_mask, const void* add, int64_t num_physical_nodes,
                            int64_t num_nodes, void* grad_pos_out, void* stream) {
  DP_ARG(grad && grad_intermediate && flat_node2pin_map && flat_node2pin_start_map &&
             grad_pos_out,
         "dp_wa_backward_to_nodes: null pointer");
  DP_ARG(!net_weights || (pin2net_map && net_mask),
         "dp_wa_backward_to_nodes: net weights need pin2net_map and net_mask");
  DP_ARG(num_physical_nodes <= num_nodes, "dp_wa_backward_to_nodes: bad node counts");
  if (num_nodes == 0) return DP_OK;
  hipStream_t s = S(stream);
  DP_PROBE_OP(DP_OP_NODE_GRAD, s);
  return DP_DISPATCH(dtype, "dp_wa_backward_to_nodes", [&] {
    const T* gi = (const T*)grad_intermediate;
    T* out = (T*)grad_pos_out;
    DP_PROBED(DP_KERNEL_PIN_POS_BACKWARD, s,
              hipLaunchKernelGGL(k_wa_pin_pos_bwd<T>, dim3((unsigned)ceil_div(num_nodes, 256)),
                                 dim3(256), 0, s, gi, gi + num_pins, (const T*)grad,
                                 (const int*)flat_node2pin_map,
                                 (const int*)flat_node2pin_start_map, (const int*)pin2net_map,
                                 (const T*)net_weights, net_mask, pin_mask, (const T*)add,
                                 num_physical_nodes, num_nodes, out, out + num_nodes));
    DP_LAUNCHED("k_wa_pin_pos_bwd");
    return DP_OK;
  });
}

int dp_hpwl(int dtype, const void* pin_pos, int64_t num_pins, const int32_t* flat_netpin,
            const int32_t* netpin_start, int64_t num_nets, const void* net_weights,
            const uint8_t* net_mask, void* workspace, double* hpwl_out, void* stream) {
  DP_ARG(pin_pos && flat_netpin && netpin_start && net_mask && workspace && hpwl_out,
         "dp_hpwl: null pointer");
  hipStream_t s = S(stream);
  double* partials = (double*)workspace;
  int grid = reduce_grid(num_nets);
  return DP_DISPATCH(dtype, "dp_hpwl", [&] {
    const T* pp = (const T*)pin_pos;
    hipLaunchKernelGGL(k_hpwl<T>, dim3(grid), dim3(256), 0, s, pp, pp + num_pins,
                       (const int*)flat_netpin, (const int*)netpin_start, num_nets,
                       (const T*)net_weights, net_mask, partials);
    DP_LAUNCHED("k_hpwl");
    return finish_reduction(partials, grid, 1, nullptr, hpwl_out, -1, s);
  });
}

int dp_hpwl_from_nets(int dtype, const void* net_hpwl, const int32_t* netpin_start,
                      int64_t num_nets, const void* net_weights, const uint8_t* net_mask,
                      void* workspace, double* hpwl_out, void* stream) {
  DP_ARG(net_hpwl && netpin_start && net_mask && workspace && hpwl_out,
         "dp_hpwl_from_nets: null pointer");
  DP_ARG(num_nets >= 0, "dp_hpwl_from_nets: bad size");
  hipStream_t s = S(stream);
  double* partials = (double*)workspace;
  int grid = reduce_grid(num_nets);
  return DP_DISPATCH(dtype, "dp_hpwl_from_nets", [&] {
    hipLaunchKernelGGL(k_hpwl_nets<T>, dim3(grid), dim3(256), 0, s, (const T*)net_hpwl,
                       (const int*)netpin_start, num_nets, (const T*)net_weights, net_mask,
                       partials);
    DP_LAUNCHED("k_hpwl_nets");
    return finish_reduction(partials, grid, 1, nullptr, hpwl_out, -1, s);
  });
}

}  // extern "C"
