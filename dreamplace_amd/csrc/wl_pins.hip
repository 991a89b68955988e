// Pin-parallel weighted-average wirelength (the engine's WA forward) and the
// node-parallel backward that consumes its output.
//
// A static plan packs whole nets, in netlist order, into groups of at most
// 64 pins; one wavefront owns one group and one lane one pin.  Each lane
// gathers its pin position once and evaluates its four exponentials once;
// per-net max/min and exponential sums are reduced through LDS by the net's
// first lane; every lane then forms its own derivative.  Nets of more than
// 64 pins are listed apart and done by one wavefront each, in chunks.
// The derivative of each pin, already multiplied by its net weight (and 0 on
// masked pins), is written at the pin's position in node (CSR) order, so the
// backward is a contiguous per-node sum.
//
// Per net the arithmetic follows weighted_average_wirelength_merged.cpp:31-116;
// packed nets accumulate their exponential sums in pin order, as there.
#include "dp_common.h"

namespace dp {

template <typename T>
__device__ __forceinline__ T big_value();
template <>
__device__ __forceinline__ float big_value<float>() { return 3.402823466e+38f; }
template <>
__device__ __forceinline__ double big_value<double>() { return 1.7976931348623157e+308; }

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

// xy[n] = (x[n], y[n]) for the physical nodes: one 8- or 16-byte gather per pin.
template <typename T>
__global__ void k_interleave(const T* __restrict__ x, const T* __restrict__ y, int64_t n,
                             typename Vec2<T>::type* xy) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    typename Vec2<T>::type v;
    v.x = x[i];
    v.y = y[i];
    xy[i] = v;
  }
}

constexpr int kWaves = 4;           // wavefronts per workgroup
constexpr int kPackedGrid = 8192;   // cap of workgroups of the packed kernel
constexpr int kBigGrid = 1024;      // cap of workgroups of the big-net kernel

template <typename T>
__device__ __forceinline__ void store_grad(T* gnode, int q, T gx, T gy) {
  typename Vec2<T>::type gv;
  gv.x = gx;
  gv.y = gy;
  ((typename Vec2<T>::type*)gnode)[q] = gv;
}

template <typename T>
__global__ void __launch_bounds__(256) k_wa_pins(
    const typename Vec2<T>::type* __restrict__ xy, const T* __restrict__ offx,
    const T* __restrict__ offy, const int* __restrict__ pin2node,
    const int* __restrict__ flat_netpin, const int* __restrict__ netpin_start,
    const int* __restrict__ pin2net, const int2* __restrict__ groups, int ngroups,
    const T* __restrict__ w, const uint8_t* __restrict__ net_mask,
    const uint8_t* __restrict__ pin_mask, const int* __restrict__ node_order,
    const T* __restrict__ inv_gamma_p, T* partial, T* gnode, double* partials) {
  __shared__ T s_x[kWaves][64], s_y[kWaves][64];
  __shared__ T s_e[kWaves][4][64];
  __shared__ T s_c[kWaves][8][64];
  __shared__ double s_red[4];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const T ig = *inv_gamma_p;
  double contrib = 0;
  // the trip count depends on blockIdx only: every wave meets every barrier
  for (int base = blockIdx.x * kWaves; base < ngroups; base += gridDim.x * kWaves) {
    const int g = base + wv;
    int pb = 0, pe = 0;
    if (g < ngroups) {
      const int2 gr = groups[g];
      pb = netpin_start[gr.x];
      pe = netpin_start[gr.y];
    }
    const int j = pb + lane;
    const bool on = j < pe;
    int p = 0, e = 0, head = 0, deg = 0;
    bool live = false;
    T x = 0, y = 0;
    if (on) {
      p = flat_netpin[j];
      e = pin2net[p];
      head = netpin_start[e] - pb;
      deg = netpin_start[e + 1] - netpin_start[e];
      live = net_mask[e] != 0;
      const typename Vec2<T>::type v = xy[pin2node[p]];
      x = offx[p] + v.x;
      y = offy[p] + v.y;
    }
    s_x[wv][lane] = x;
    s_y[wv][lane] = y;
    __syncthreads();
    const bool is_head = on && lane == head;
    if (is_head && live) {
      T xmax = -big_value<T>(), xmin = big_value<T>(), ymax = -big_value<T>(), ymin = big_value<T>();
      for (int k = head; k < head + deg; ++k) {
        const T xx = s_x[wv][k], yy = s_y[wv][k];
        xmax = xx > xmax ? xx : xmax;
        xmin = xx < xmin ? xx : xmin;
        ymax = yy > ymax ? yy : ymax;
        ymin = yy < ymin ? yy : ymin;
      }
      s_c[wv][0][lane] = xmax;
      s_c[wv][1][lane] = xmin;
      s_c[wv][2][lane] = ymax;
      s_c[wv][3][lane] = ymin;
    }
    __syncthreads();
    T ex = 0, enx = 0, ey = 0, eny = 0;
    if (on && live) {
      ex = exp((x - s_c[wv][0][head]) * ig);
      enx = exp((s_c[wv][1][head] - x) * ig);
      ey = exp((y - s_c[wv][2][head]) * ig);
      eny = exp((s_c[wv][3][head] - y) * ig);
      s_e[wv][0][lane] = ex;
      s_e[wv][1][lane] = enx;
      s_e[wv][2][lane] = ey;
      s_e[wv][3][lane] = eny;
    }
    __syncthreads();
    if (is_head && live) {
      T xs = 0, xns = 0, es = 0, ens = 0, ys = 0, yns = 0, fs = 0, fns = 0;
      for (int k = head; k < head + deg; ++k) {
        const T xx = s_x[wv][k], yy = s_y[wv][k];
        const T a0 = s_e[wv][0][k], a1 = s_e[wv][1][k], a2 = s_e[wv][2][k], a3 = s_e[wv][3][k];
        xs += xx * a0;
        xns += xx * a1;
        es += a0;
        ens += a1;
        ys += yy * a2;
        yns += yy * a3;
        fs += a2;
        fns += a3;
      }
      const T wl = xs / es - xns / ens + ys / fs - yns / fns;
      if (partial) partial[e] = wl;
      contrib += (double)(w ? wl * w[e] : wl);
      const T bx = ig / es, bnx = -ig / ens, by = ig / fs, bny = -ig / fns;
      s_c[wv][0][lane] = (T)((1.0 - bx * xs) / es);
      s_c[wv][1][lane] = bx;
      s_c[wv][2][lane] = (T)((1.0 - bnx * xns) / ens);
      s_c[wv][3][lane] = bnx;
      s_c[wv][4][lane] = (T)((1.0 - by * ys) / fs);
      s_c[wv][5][lane] = by;
      s_c[wv][6][lane] = (T)((1.0 - bny * yns) / fns);
      s_c[wv][7][lane] = bny;
    }
    if (is_head && !live && partial) partial[e] = 0;
    __syncthreads();
    if (on) {
      T gx = 0, gy = 0;
      if (live && !(pin_mask && pin_mask[p])) {
        gx = (s_c[wv][0][head] + s_c[wv][1][head] * x) * ex -
             (s_c[wv][2][head] + s_c[wv][3][head] * x) * enx;
        gy = (s_c[wv][4][head] + s_c[wv][5][head] * y) * ey -
             (s_c[wv][6][head] + s_c[wv][7][head] * y) * eny;
        if (w) {
          const T wn = w[e];
          gx *= wn;
          gy *= wn;
        }
      }
      store_grad(gnode, node_order[p], gx, gy);
    }
    __syncthreads();  // LDS reuse by the next group set
  }
  const double bs = block_sum<256>(contrib, s_red);
  if (threadIdx.x == 0) partials[blockIdx.x] = bs;
}

// Nets of more than 64 pins: one wavefront per net.
template <typename T>
__global__ void __launch_bounds__(256) k_wa_big(
    const typename Vec2<T>::type* __restrict__ xy, const T* __restrict__ offx,
    const T* __restrict__ offy, const int* __restrict__ pin2node,
    const int* __restrict__ flat_netpin, const int* __restrict__ netpin_start,
    const int* __restrict__ big, int nbig, const T* __restrict__ w,
    const uint8_t* __restrict__ net_mask, const uint8_t* __restrict__ pin_mask,
    const int* __restrict__ node_order, const T* __restrict__ inv_gamma_p, T* partial, T* gnode,
    double* partials) {
  __shared__ double s_red[4];
  const int lane = threadIdx.x & 63;
  const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int nwaves = (gridDim.x * blockDim.x) >> 6;
  const T ig = *inv_gamma_p;
  double contrib = 0;
  for (int q = wave; q < nbig; q += nwaves) {
    const int e = big[q];
    const int pb = netpin_start[e], pe = netpin_start[e + 1];
    const bool live = net_mask[e] != 0;
    T xmax = -big_value<T>(), xmin = big_value<T>(), ymax = -big_value<T>(), ymin = big_value<T>();
    for (int j = pb + lane; j < pe; j += 64) {
      const int p = flat_netpin[j];
      const typename Vec2<T>::type v = xy[pin2node[p]];
      const T xx = offx[p] + v.x, yy = offy[p] + v.y;
      xmax = xx > xmax ? xx : xmax;
      xmin = xx < xmin ? xx : xmin;
      ymax = yy > ymax ? yy : ymax;
      ymin = yy < ymin ? yy : ymin;
    }
    xmax = wave_max(xmax);
    ymax = wave_max(ymax);
    xmin = -wave_max(-xmin);
    ymin = -wave_max(-ymin);
    T xs = 0, xns = 0, es = 0, ens = 0, ys = 0, yns = 0, fs = 0, fns = 0;
    for (int j = pb + lane; j < pe; j += 64) {
      const int p = flat_netpin[j];
      const typename Vec2<T>::type v = xy[pin2node[p]];
      const T xx = offx[p] + v.x, yy = offy[p] + v.y;
      const T a0 = exp((xx - xmax) * ig), a1 = exp((xmin - xx) * ig);
      const T a2 = exp((yy - ymax) * ig), a3 = exp((ymin - yy) * ig);
      xs += xx * a0;
      xns += xx * a1;
      es += a0;
      ens += a1;
      ys += yy * a2;
      yns += yy * a3;
      fs += a2;
      fns += a3;
    }
    xs = wave_sum(xs);
    xns = wave_sum(xns);
    es = wave_sum(es);
    ens = wave_sum(ens);
    ys = wave_sum(ys);
    yns = wave_sum(yns);
    fs = wave_sum(fs);
    fns = wave_sum(fns);
    const T wl = xs / es - xns / ens + ys / fs - yns / fns;
    if (lane == 0) {
      if (partial) partial[e] = live ? wl : T(0);
      if (live) contrib += (double)(w ? wl * w[e] : wl);
    }
    const T bx = ig / es, bnx = -ig / ens, by = ig / fs, bny = -ig / fns;
    const T ax = (T)((1.0 - bx * xs) / es), anx = (T)((1.0 - bnx * xns) / ens);
    const T ay = (T)((1.0 - by * ys) / fs), any = (T)((1.0 - bny * yns) / fns);
    for (int j = pb + lane; j < pe; j += 64) {
      const int p = flat_netpin[j];
      T gx = 0, gy = 0;
      if (live && !(pin_mask && pin_mask[p])) {
        const typename Vec2<T>::type v = xy[pin2node[p]];
        const T xx = offx[p] + v.x, yy = offy[p] + v.y;
        gx = (ax + bx * xx) * exp((xx - xmax) * ig) - (anx + bnx * xx) * exp((xmin - xx) * ig);
        gy = (ay + by * yy) * exp((yy - ymax) * ig) - (any + bny * yy) * exp((ymin - yy) * ig);
        if (w) {
          gx *= w[e];
          gy *= w[e];
        }
      }
      store_grad(gnode, node_order[p], gx, gy);
    }
  }
  const double bs = block_sum<256>(contrib, s_red);
  if (threadIdx.x == 0) partials[blockIdx.x] = bs;
}

// grad[n] = (*grad_scale) * sum of gnode over node n's CSR range + add[n]
// (add nullable).  Fillers (n >= nphys) get add[n] only.
template <typename T>
__global__ void k_node_grad(const typename Vec2<T>::type* __restrict__ gnode,
                            const int* __restrict__ node2pin_start, const T* __restrict__ grad,
                            const T* __restrict__ add, int64_t nphys, int64_t nnodes, T* gx_out,
                            T* gy_out) {
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= nnodes) return;
  T sx = 0, sy = 0;
  if (n < nphys) {
    const T g = *grad;
    for (int j = node2pin_start[n]; j < node2pin_start[n + 1]; ++j) {
      const typename Vec2<T>::type v = gnode[j];
      sx += v.x * g;
      sy += v.y * g;
    }
  }
  if (add) {
    sx += add[n];
    sy += add[nnodes + n];
  }
  gx_out[n] = sx;
  gy_out[n] = sy;
}

static int grid_cap(int64_t items, int cap) {
  int64_t g = ceil_div(items, kWaves);
  if (g < 1) g = 1;
  return (int)(g > cap ? cap : g);
}

}  // namespace dp

using namespace dp;

extern "C" {

int dp_wa_pack_nets(const int32_t* netpin_start, int64_t num_nets, int32_t* groups,
                    int64_t* num_groups, int32_t* big_nets, int64_t* num_big) {
  DP_ARG(netpin_start && groups && num_groups && big_nets && num_big,
         "dp_wa_pack_nets: null pointer");
  int64_t g = 0, b = 0, e = 0;
  while (e < num_nets) {
    const int64_t d0 = (int64_t)netpin_start[e + 1] - netpin_start[e];
    DP_ARG(d0 >= 0, "dp_wa_pack_nets: netpin_start not monotone at net %lld", (long long)e);
    if (d0 > 64) {
      big_nets[b++] = (int32_t)e++;
      continue;
    }
    const int64_t nb = e;
    int64_t pins = d0;
    ++e;
    while (e < num_nets) {
      const int64_t d = (int64_t)netpin_start[e + 1] - netpin_start[e];
      if (d > 64 || pins + d > 64) break;
      pins += d;
      ++e;
    }
    groups[2 * g] = (int32_t)nb;
    groups[2 * g + 1] = (int32_t)e;
    ++g;
  }
  *num_groups = g;
  *num_big = b;
  return DP_OK;
}

size_t dp_wa_pins_workspace_bytes(int64_t num_physical_nodes, int dtype) {
  const size_t eb = dtype == DP_FLOAT64 ? 16 : 8;
  return reduce_slot_bytes() + align256(eb * (size_t)(num_physical_nodes > 0 ? num_physical_nodes : 1));
}

int dp_wa_forward_pins(int dtype, const void* pos, int64_t num_nodes, int64_t num_physical_nodes,
                       const void* pin_offset_x, const void* pin_offset_y,
                       const int32_t* pin2node_map, const int32_t* flat_netpin,
                       const int32_t* netpin_start, const int32_t* pin2net_map,
                       const int32_t* groups, int64_t num_groups, const int32_t* big_nets,
                       int64_t num_big, const void* net_weights, const uint8_t* net_mask,
                       const uint8_t* pin_mask, const int32_t* pin_node_order,
                       const void* inv_gamma, void* partial_wl, void* grad_node_order,
                       void* wl_out, void* workspace, void* stream) {
  DP_ARG(pos && pin_offset_x && pin_offset_y && pin2node_map && flat_netpin && netpin_start &&
             pin2net_map && (groups || num_groups == 0) && (big_nets || num_big == 0) &&
             net_mask && pin_node_order && inv_gamma && grad_node_order && wl_out && workspace,
         "dp_wa_forward_pins: null pointer");
  DP_ARG(num_physical_nodes <= num_nodes && num_groups >= 0 && num_big >= 0,
         "dp_wa_forward_pins: bad counts");
  hipStream_t s = S(stream);
  double* partials = (double*)workspace;
  char* xy = (char*)workspace + reduce_slot_bytes();
  const int g1 = grid_cap(num_groups, kPackedGrid);
  const int g2 = grid_cap(num_big, kBigGrid);
  return DP_DISPATCH(dtype, "dp_wa_forward_pins", [&] {
    typedef typename Vec2<T>::type V;
    const T* P = (const T*)pos;
    hipLaunchKernelGGL(k_interleave<T>, dim3(grid_for(num_physical_nodes, 256)), dim3(256), 0, s,
                       P, P + num_nodes, num_physical_nodes, (V*)xy);
    DP_LAUNCHED("k_interleave");
    DP_PROBED(DP_KERNEL_WA_FORWARD, s,
              hipLaunchKernelGGL(k_wa_pins<T>, dim3(g1), dim3(256), 0, s, (const V*)xy,
                                 (const T*)pin_offset_x, (const T*)pin_offset_y,
                                 (const int*)pin2node_map, (const int*)flat_netpin,
                                 (const int*)netpin_start, (const int*)pin2net_map,
                                 (const int2*)groups, (int)num_groups, (const T*)net_weights,
                                 net_mask, pin_mask, (const int*)pin_node_order,
                                 (const T*)inv_gamma, (T*)partial_wl, (T*)grad_node_order,
                                 partials));
    DP_LAUNCHED("k_wa_pins");
    hipLaunchKernelGGL(k_wa_big<T>, dim3(g2), dim3(256), 0, s, (const V*)xy,
                       (const T*)pin_offset_x, (const T*)pin_offset_y, (const int*)pin2node_map,
                       (const int*)flat_netpin, (const int*)netpin_start, (const int*)big_nets,
                       (int)num_big, (const T*)net_weights, net_mask, pin_mask,
                       (const int*)pin_node_order, (const T*)inv_gamma, (T*)partial_wl,
                       (T*)grad_node_order, partials + g1);
    DP_LAUNCHED("k_wa_big");
    return finish_reduction(partials, g1 + g2, 1, nullptr, wl_out, dtype, s);
  });
}

int dp_node_grad_from_pins(int dtype, const void* grad, const void* grad_node_order,
                           const int32_t* flat_node2pin_start_map, const void* add,
                           int64_t num_physical_nodes, int64_t num_nodes, void* grad_pos_out,
                           void* stream) {
  DP_ARG(grad && grad_node_order && flat_node2pin_start_map && grad_pos_out,
         "dp_node_grad_from_pins: null pointer");
  DP_ARG(num_physical_nodes <= num_nodes, "dp_node_grad_from_pins: bad node counts");
  if (num_nodes == 0) return DP_OK;
  hipStream_t s = S(stream);
  return DP_DISPATCH(dtype, "dp_node_grad_from_pins", [&] {
    typedef typename Vec2<T>::type V;
    T* out = (T*)grad_pos_out;
    DP_PROBED(DP_KERNEL_PIN_POS_BACKWARD, s,
              hipLaunchKernelGGL(k_node_grad<T>, dim3((unsigned)ceil_div(num_nodes, 256)),
                                 dim3(256), 0, s, (const V*)grad_node_order,
                                 (const int*)flat_node2pin_start_map, (const T*)grad,
                                 (const T*)add, num_physical_nodes, num_nodes, out,
                                 out + num_nodes));
    DP_LAUNCHED("k_node_grad");
    return DP_OK;
  });
}

}  // extern "C"
