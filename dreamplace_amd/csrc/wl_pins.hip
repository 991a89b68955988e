// Pin-parallel weighted-average wirelength (the engine's WA forward) and the
// node-parallel backward that consumes its output.
//
// Data movement: the netlist connects nodes in no spatial or index order, so
// the forward has to permute between node order and net order twice (pin
// positions in, pin derivatives out).  Both permutations are done as
// scattered 8/16-byte stores of whole (x, y) words, which this memory system
// sustains at ~1.5x the rate of scattered loads:
//   k_pin_push   one thread per pin slot in node (CSR) order reads its node's
//                position and the pin offset (both contiguous) and stores
//                pos + offset at the pin's slot in "group order";
//   k_wa_pins    reads pin positions and per-pin metadata contiguously in
//                group order and stores each pin's derivative at the pin's
//                slot in node order;
//   k_node_grad  sums each node's contiguous CSR range.
//
// Group order is a static plan (dp_wa_pack_nets): nets of at most 64 pins are
// packed best-fit-decreasing by degree into groups of at most 64 pins, so a
// group's per-net serial loops are about as long as its nets' degree; one
// wavefront (one workgroup) owns one group and one lane one pin.  Each lane
// evaluates its four exponentials once; per-net max/min and exponential sums
// are reduced through LDS by the net's first lane; every lane then forms its
// own derivative.  Nets of more than 64 pins follow all groups in group order
// and are done by one wavefront each, in chunks.
//
// Per net the arithmetic follows weighted_average_wirelength_merged.cpp:31-116;
// packed nets accumulate their exponential sums in pin order, as there.
#include <vector>

#include "dp_common.h"

namespace dp {

template <typename T>
__device__ __forceinline__ T big_value();
template <>
__device__ __forceinline__ float big_value<float>() { return 3.402823466e+38f; }
template <>
__device__ __forceinline__ double big_value<double>() { return 1.7976931348623157e+308; }

constexpr int kWaves = 4;           // independent wavefronts per workgroup of the packed kernel
constexpr int kPackedGrid = 8192;   // cap of workgroups of the packed kernel
constexpr int kBigGrid = 1024;      // cap of workgroups of the big-net kernel (4 wavefronts each)
constexpr int kMaskBit = (int)0x80000000u;

// Per-pin metadata of the packed kernel.
typedef int4 PinMeta;

// ---- static preparation ----------------------------------------------------
// group-order slot s of pin r of net e (s = net_slot[e] + r):
//   meta[s] = {e | (pin masked ? sign bit : 0), node-order slot, r | deg << 8, 0}
// (deg only for nets of <= 64 pins; 16 bytes: one aligned dwordx4 load per
// lane, which measured faster here than a 12-byte record) and inv[pin] = s.
// (.y is filled by k_prep_node.)
__global__ void k_prep_net(const int* __restrict__ flat_netpin,
                           const int* __restrict__ netpin_start, const int* __restrict__ pin2net,
                           const uint8_t* __restrict__ pin_mask, const int* __restrict__ net_slot,
                           int64_t num_pins, PinMeta* meta, int* inv) {
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < num_pins;
       j += (int64_t)gridDim.x * blockDim.x) {
    const int p = flat_netpin[j];
    const int e = pin2net[p];
    const int b = netpin_start[e];
    const int r = (int)j - b, deg = netpin_start[e + 1] - b;
    const int s = net_slot[e] + r;
    PinMeta m;
    m.x = e | (pin_mask && pin_mask[p] ? kMaskBit : 0);
    m.y = 0;
    m.z = deg <= 64 ? (r | (deg << 8)) : 0;
    m.w = 0;
    meta[s] = m;
    inv[p] = s;
  }
}

// node-order slot i: dst[i] = group-order slot of the pin, meta[dst].y = i,
// off[i] = (offset_x, offset_y) of the pin
template <typename T>
__global__ void k_prep_node(const int* __restrict__ flat_node2pin, const int* __restrict__ inv,
                            const T* __restrict__ offx, const T* __restrict__ offy,
                            int64_t num_pins, int* dst, PinMeta* meta,
                            typename Vec2<T>::type* off) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < num_pins;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int p = flat_node2pin[i];
    const int s = inv[p];
    dst[i] = s;
    meta[s].y = (int)i;
    typename Vec2<T>::type v;
    v.x = offx[p];
    v.y = offy[p];
    off[i] = v;
  }
}

// Per group-order slot s of net e: meta[s].w = net_mask[e] != 0 and
// slot_w[s] = w[e] (1 without weights), for callers whose weights and mask
// stay fixed between calls (dp_wa_pins_slot_weights).
template <typename T>
__global__ void k_slot_weights(int64_t num_pins, PinMeta* meta, const T* __restrict__ w,
                               const uint8_t* __restrict__ net_mask, T* slot_w) {
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < num_pins;
       j += (int64_t)gridDim.x * blockDim.x) {
    const int e = meta[j].x & ~kMaskBit;
    meta[j].w = net_mask[e] != 0;
    slot_w[j] = w ? w[e] : T(1);
  }
}

// slot_node[i] = node owning node-order slot i
__global__ void k_prep_slot_node(const int* __restrict__ node2pin_start, int64_t nphys,
                                 int* slot_node) {
  for (int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; n < nphys;
       n += (int64_t)gridDim.x * blockDim.x)
    for (int i = node2pin_start[n]; i < node2pin_start[n + 1]; ++i) slot_node[i] = (int)n;
}

// ---- forward ---------------------------------------------------------------
// pin_xy[dst[i]] = (x[n] + off[i].x, y[n] + off[i].y), n = slot_node[i]
template <typename T>
__global__ void __launch_bounds__(256) k_pin_push(const T* __restrict__ x, const T* __restrict__ y,
                                                  int64_t num_pins,
                                                  const int* __restrict__ slot_node,
                                                  const typename Vec2<T>::type* __restrict__ off,
                                                  const int* __restrict__ dst,
                                                  typename Vec2<T>::type* pin_xy) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= num_pins) return;
  const int n = slot_node[i];
  const typename Vec2<T>::type o = off[i];
  typename Vec2<T>::type v;
  v.x = x[n] + o.x;
  v.y = y[n] + o.y;
  pin_xy[dst[i]] = v;
}

// Wirelength arithmetic of the pin-parallel kernels.
//   kWaRef: d/dx_p = (a+ + b+ x_p) e+_p - (a- + b- x_p) e-_p with
//     b+ = (1/g)/S+, a+ = (1 - b+ XS+)/S+ (XS+ = sum x e+), and the value
//     XS+/S+ - XS-/S- (weighted_average_wirelength_merged.cpp:85-111).
//   kWaCentred: the same function with x measured from the net's extremes,
//     u+ = x - x_max and u- = x - x_min (exact in floating point for a net's
//     pins, which lie within a factor of two of each other or straddle 0 by
//     less than a bin), C+ = sum u+ e+ / S+, C- = sum u- e- / S-:
//       value  = (x_max - x_min) + C+ - C-
//       d/dx_p = (e+_p/S+)(1 + (u+_p - C+)/g) - (e-_p/S-)(1 - (u-_p - C-)/g).
//     In exact arithmetic both are the same; the reference form cancels
//     b+ x_p (~ x/g, thousands at die coordinates) against a+, which costs
//     FP32 about 1e-5 relative in the gradient, while every term here is of
//     the order of the net's span.
constexpr int kWaRef = 0, kWaLse = 1, kWaCentred = 2;
// kWaCentred groups whose first (largest) net has more pins than this reduce
// their per-net extremes and sums across lanes in shuffle steps (as many as
// that net needs, at most six) instead of the head lane's serial loops over
// the net's pins (the order of the sums then differs from the reference's
// pin order, which only kWaRef keeps).
constexpr int kTreeDeg = 8;

// A pin's four exponentials as one aligned LDS word.
template <typename T>
struct alignas(4 * sizeof(T)) Exp4 {
  T a0, a1, a2, a3;
};

template <typename T>
__device__ __forceinline__ void store_grad(T* gnode, int q, T gx, T gy) {
  typename Vec2<T>::type gv;
  gv.x = gx;
  gv.y = gy;
  ((typename Vec2<T>::type*)gnode)[q] = gv;
}

// One lane's inputs of one group: the group's slot range and the lane's pin
// metadata, slot weight and position (zeros past the group's end).
template <typename T>
struct GroupLoad {
  int pb, pe;
  PinMeta m;
  T sw;
  typename Vec2<T>::type v;
};

template <typename T>
__device__ __forceinline__ void load_group(GroupLoad<T>& L, int g, int ngroups, int lane,
                                           const int* __restrict__ group_start,
                                           const PinMeta* __restrict__ meta,
                                           const T* __restrict__ slot_w,
                                           const typename Vec2<T>::type* __restrict__ pin_xy) {
  L.pb = 0;
  L.pe = 0;
  L.m = make_int4(0, 0, 0, 0);
  L.sw = T(1);
  L.v.x = 0;
  L.v.y = 0;
  if (g < ngroups) {
    L.pb = group_start[g];
    L.pe = group_start[g + 1];
    const int j = L.pb + lane;
    if (j < L.pe) {
      L.m = meta[j];
      if (slot_w) L.sw = slot_w[j];
      L.v = pin_xy[j];
    }
  }
}

// Four independent wavefronts per workgroup, each with its own LDS slice and
// its own grid-stride loop over groups; they exchange data only within the
// wavefront (wave_sync), so no wavefront waits for another until the final
// workgroup sum.  Each wavefront issues the loads of its next group before it
// evaluates the current one, so that their latency overlaps the evaluation
// (244 -> 206 us per call at 9.6M pins, the same bits).
// kMode: kWaRef, the reference's merged WA arithmetic; kWaLse, the
// log-sum-exp model (logsumexp_wirelength_merged.cpp:14-85) on the same
// skeleton, g_p then pointing at gamma instead of 1/gamma; kWaCentred, the
// same WA model with every sum and derivative taken relative to the net's
// extremes (see wa_centred below).
template <typename T, int kMode>
__global__ void __launch_bounds__(256) k_wa_pins(
    const typename Vec2<T>::type* __restrict__ pin_xy, const PinMeta* __restrict__ meta,
    const int* __restrict__ group_start, int ngroups, const T* __restrict__ w,
    const uint8_t* __restrict__ net_mask, const T* __restrict__ slot_w,
    const T* __restrict__ g_p, T* partial, T* net_hpwl, T* gnode, double* partials) {
  typedef typename Vec2<T>::type V;
  // per-lane words read by the net's first lane: one LDS read per pin for
  // (x, y) and one for its four exponentials
  __shared__ V sh_xy[kWaves][64];
  __shared__ Exp4<T> sh_e[kWaves][64];
  __shared__ T sh_c[kWaves][8][64];
  __shared__ double s_red[kWaves];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  V* s_xy = sh_xy[wv];
  Exp4<T>* s_e = sh_e[wv];
  T(*s_c)[64] = sh_c[wv];
  constexpr bool kLse = kMode == kWaLse;
  const T gam = kLse ? *g_p : T(0);
  const T ig = kLse ? T(1) / gam : *g_p;
  // weights (and the live flag) per pin slot when the caller prepared them
  // (contiguous), else gathered per net
  const bool weighted = slot_w != nullptr || w != nullptr;
  double contrib = 0;
  const int gstride = gridDim.x * kWaves;
  int g = blockIdx.x * kWaves + wv;
  GroupLoad<T> nxt;
  load_group(nxt, g, ngroups, lane, group_start, meta, slot_w, pin_xy);
  for (; g < ngroups; g += gstride) {
    const GroupLoad<T> cur = nxt;
    load_group(nxt, g + gstride, ngroups, lane, group_start, meta, slot_w, pin_xy);
    const int j = cur.pb + lane;
    const bool on = j < cur.pe;
    int e = 0, head = 0, deg = 0, dst = 0;
    bool live = false, pin_masked = false;
    T x = 0, y = 0, wn = T(1);
    if (on) {
      const PinMeta m = cur.m;
      e = m.x & ~kMaskBit;
      pin_masked = m.x < 0;
      dst = m.y;
      head = lane - (m.z & 0xff);
      deg = m.z >> 8;
      if (slot_w) {
        live = m.w != 0;
        wn = cur.sw;
      } else {
        live = net_mask[e] != 0;
        if (w) wn = w[e];
      }
      x = cur.v.x;
      y = cur.v.y;
    }
    {
      V v;
      v.x = x;
      v.y = y;
      s_xy[lane] = v;
    }
    wave_sync();
    const bool is_head = on && lane == head;
    // nets are packed largest first: lane 0 holds the group's largest degree.
    // A segment of d lanes is reduced into its head after the steps of
    // offsets below d; later steps add nothing to any head (lane + o is past
    // the segment), so the trees stop at the group's largest degree.
    const int maxdeg = __builtin_amdgcn_readfirstlane(deg);
    const bool tree = kMode == kWaCentred && maxdeg > kTreeDeg;
    const int seg_end = head + deg;  // one past this lane's net's last lane
    T xmax = -big_value<T>(), xmin = big_value<T>(), ymax = -big_value<T>(), ymin = big_value<T>();
    if (tree) {
      // segmented max / min: after the step of offset o, lane l holds the
      // extreme over lanes [l, min(l + 2o, seg_end)); the head ends with its net's
      T a = on ? x : -big_value<T>(), b = on ? x : big_value<T>();
      T c = on ? y : -big_value<T>(), dd = on ? y : big_value<T>();
#pragma unroll
      for (int o = 1; o < 64; o <<= 1) {
        if (o >= maxdeg) break;  // uniform
        const T a2 = __shfl_down(a, o, 64), b2 = __shfl_down(b, o, 64);
        const T c2 = __shfl_down(c, o, 64), d2 = __shfl_down(dd, o, 64);
        if (lane + o < seg_end) {
          a = a2 > a ? a2 : a;
          b = b2 < b ? b2 : b;
          c = c2 > c ? c2 : c;
          dd = d2 < dd ? d2 : dd;
        }
      }
      if (is_head && (live || net_hpwl)) {
        xmax = a;
        xmin = b;
        ymax = c;
        ymin = dd;
        s_c[0][lane] = xmax;
        s_c[1][lane] = xmin;
        s_c[2][lane] = ymax;
        s_c[3][lane] = ymin;
        if (net_hpwl) net_hpwl[e] = xmax - xmin + ymax - ymin;
      }
    } else if (is_head && (live || net_hpwl)) {
      for (int k = head; k < head + deg; ++k) {
        const V v = s_xy[k];
        const T xx = v.x, yy = v.y;
        xmax = xx > xmax ? xx : xmax;
        xmin = xx < xmin ? xx : xmin;
        ymax = yy > ymax ? yy : ymax;
        ymin = yy < ymin ? yy : ymin;
      }
      s_c[0][lane] = xmax;
      s_c[1][lane] = xmin;
      s_c[2][lane] = ymax;
      s_c[3][lane] = ymin;
      // the net's HPWL, in k_hpwl's expression order, for every net
      if (net_hpwl) net_hpwl[e] = xmax - xmin + ymax - ymin;
    }
    wave_sync();
    T ex = 0, enx = 0, ey = 0, eny = 0;
    T ux = 0, unx = 0, uy = 0, uny = 0;  // kWaCentred: x - x_max, x - x_min, ...
    if (on && live) {
      if (kMode == kWaCentred) {
        ux = x - s_c[0][head];
        unx = x - s_c[1][head];
        uy = y - s_c[2][head];
        uny = y - s_c[3][head];
        ex = exp(ux * ig);
        enx = exp(-unx * ig);
        ey = exp(uy * ig);
        eny = exp(-uny * ig);
      } else {
        ex = exp((x - s_c[0][head]) * ig);
        enx = exp((s_c[1][head] - x) * ig);
        ey = exp((y - s_c[2][head]) * ig);
        eny = exp((s_c[3][head] - y) * ig);
      }
      Exp4<T> q;
      q.a0 = ex;
      q.a1 = enx;
      q.a2 = ey;
      q.a3 = eny;
      s_e[lane] = q;
    }
    wave_sync();
    if (kLse && is_head && live) {
      T es = 0, ens = 0, fs = 0, fns = 0;
      for (int k = head; k < head + deg; ++k) {
        const Exp4<T> q = s_e[k];
        es += q.a0;
        ens += q.a1;
        fs += q.a2;
        fns += q.a3;
      }
      const T wl = (log(es) + log(ens)) * gam + xmax - xmin + (log(fs) + log(fns)) * gam + ymax - ymin;
      if (partial) partial[e] = wl;
      contrib += (double)(weighted ? wl * wn : wl);
      s_c[0][lane] = (T)(1.0 / es);
      s_c[1][lane] = (T)(1.0 / ens);
      s_c[2][lane] = (T)(1.0 / fs);
      s_c[3][lane] = (T)(1.0 / fns);
    }
    T sum[8] = {0, 0, 0, 0, 0, 0, 0, 0};  // kWaCentred, tree: this lane's partial sums
    if (kMode == kWaCentred && tree) {
      sum[0] = ux * ex;
      sum[1] = unx * enx;
      sum[2] = ex;
      sum[3] = enx;
      sum[4] = uy * ey;
      sum[5] = uny * eny;
      sum[6] = ey;
      sum[7] = eny;
#pragma unroll
      for (int o = 1; o < 64; o <<= 1) {
        if (o >= maxdeg) break;  // uniform
        const bool in = lane + o < seg_end;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
          const T t = __shfl_down(sum[k], o, 64);
          if (in) sum[k] += t;
        }
      }
    }
    if (kMode == kWaCentred && is_head && live) {
      T xs = 0, xns = 0, es = 0, ens = 0, ys = 0, yns = 0, fs = 0, fns = 0;
      if (tree) {
        xs = sum[0];
        xns = sum[1];
        es = sum[2];
        ens = sum[3];
        ys = sum[4];
        yns = sum[5];
        fs = sum[6];
        fns = sum[7];
      } else {
        for (int k = head; k < head + deg; ++k) {
          const V v = s_xy[k];
          const Exp4<T> q = s_e[k];
          const T a0 = q.a0, a1 = q.a1, a2 = q.a2, a3 = q.a3;
          xs += (v.x - xmax) * a0;
          xns += (v.x - xmin) * a1;
          es += a0;
          ens += a1;
          ys += (v.y - ymax) * a2;
          yns += (v.y - ymin) * a3;
          fs += a2;
          fns += a3;
        }
      }
      const T cx = xs / es, cnx = xns / ens, cy = ys / fs, cny = yns / fns;
      const T wl = ((xmax - xmin) + cx - cnx) + ((ymax - ymin) + cy - cny);
      if (partial) partial[e] = wl;
      contrib += (double)(weighted ? wl * wn : wl);
      s_c[0][lane] = T(1) / es;
      s_c[1][lane] = cx;
      s_c[2][lane] = T(1) / ens;
      s_c[3][lane] = cnx;
      s_c[4][lane] = T(1) / fs;
      s_c[5][lane] = cy;
      s_c[6][lane] = T(1) / fns;
      s_c[7][lane] = cny;
    }
    if (kMode == kWaRef && is_head && live) {
      T xs = 0, xns = 0, es = 0, ens = 0, ys = 0, yns = 0, fs = 0, fns = 0;
      for (int k = head; k < head + deg; ++k) {
        const V v = s_xy[k];
        const Exp4<T> q = s_e[k];
        const T xx = v.x, yy = v.y;
        const T a0 = q.a0, a1 = q.a1, a2 = q.a2, a3 = q.a3;
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
      contrib += (double)(weighted ? wl * wn : wl);
      const T bx = ig / es, bnx = -ig / ens, by = ig / fs, bny = -ig / fns;
      s_c[0][lane] = (T)((1.0 - bx * xs) / es);
      s_c[1][lane] = bx;
      s_c[2][lane] = (T)((1.0 - bnx * xns) / ens);
      s_c[3][lane] = bnx;
      s_c[4][lane] = (T)((1.0 - by * ys) / fs);
      s_c[5][lane] = by;
      s_c[6][lane] = (T)((1.0 - bny * yns) / fns);
      s_c[7][lane] = bny;
    }
    wave_sync();
    if (on) {
      T gx = 0, gy = 0;
      if (live && !pin_masked) {
        if (kLse) {
          gx = ex * s_c[0][head] - enx * s_c[1][head];
          gy = ey * s_c[2][head] - eny * s_c[3][head];
        } else if (kMode == kWaCentred) {
          gx = (ex * s_c[0][head]) * (T(1) + (ux - s_c[1][head]) * ig) -
               (enx * s_c[2][head]) * (T(1) - (unx - s_c[3][head]) * ig);
          gy = (ey * s_c[4][head]) * (T(1) + (uy - s_c[5][head]) * ig) -
               (eny * s_c[6][head]) * (T(1) - (uny - s_c[7][head]) * ig);
        } else {
          gx = (s_c[0][head] + s_c[1][head] * x) * ex - (s_c[2][head] + s_c[3][head] * x) * enx;
          gy = (s_c[4][head] + s_c[5][head] * y) * ey - (s_c[6][head] + s_c[7][head] * y) * eny;
        }
        if (weighted) {
          gx *= wn;
          gy *= wn;
        }
      }
      store_grad(gnode, dst, gx, gy);
    }
    wave_sync();  // LDS reuse by the next group
  }
  const double bs = block_sum<256>(contrib, s_red);
  if (threadIdx.x == 0) partials[blockIdx.x] = bs;
}

// Nets of more than 64 pins: one wavefront per net; big[q] = {net, first
// group-order slot, degree, 0}.
template <typename T, int kMode>
__global__ void __launch_bounds__(256) k_wa_big(
    const typename Vec2<T>::type* __restrict__ pin_xy, const PinMeta* __restrict__ meta,
    const int4* __restrict__ big, int nbig, const T* __restrict__ w,
    const uint8_t* __restrict__ net_mask, const T* __restrict__ g_p, T* partial, T* net_hpwl,
    T* gnode, double* partials) {
  __shared__ double s_red[4];
  const int lane = threadIdx.x & 63;
  const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int nwaves = (gridDim.x * blockDim.x) >> 6;
  constexpr bool kLse = kMode == kWaLse;
  constexpr bool kCen = kMode == kWaCentred;
  const T gam = kLse ? *g_p : T(0);
  const T ig = kLse ? T(1) / gam : *g_p;
  double contrib = 0;
  for (int q = wave; q < nbig; q += nwaves) {
    const int4 bq = big[q];
    const int e = bq.x, pb = bq.y, pe = bq.y + bq.z;
    const bool live = net_mask[e] != 0;
    T xmax = -big_value<T>(), xmin = big_value<T>(), ymax = -big_value<T>(), ymin = big_value<T>();
    for (int j = pb + lane; j < pe; j += 64) {
      const typename Vec2<T>::type v = pin_xy[j];
      xmax = v.x > xmax ? v.x : xmax;
      xmin = v.x < xmin ? v.x : xmin;
      ymax = v.y > ymax ? v.y : ymax;
      ymin = v.y < ymin ? v.y : ymin;
    }
    xmax = wave_max(xmax);
    ymax = wave_max(ymax);
    xmin = -wave_max(-xmin);
    ymin = -wave_max(-ymin);
    if (net_hpwl && lane == 0) net_hpwl[e] = xmax - xmin + ymax - ymin;
    T xs = 0, xns = 0, es = 0, ens = 0, ys = 0, yns = 0, fs = 0, fns = 0;
    for (int j = pb + lane; j < pe; j += 64) {
      const typename Vec2<T>::type v = pin_xy[j];
      const T xx = v.x, yy = v.y;
      const T a0 = exp((xx - xmax) * ig), a1 = exp((xmin - xx) * ig);
      const T a2 = exp((yy - ymax) * ig), a3 = exp((ymin - yy) * ig);
      xs += (kCen ? xx - xmax : xx) * a0;
      xns += (kCen ? xx - xmin : xx) * a1;
      es += a0;
      ens += a1;
      ys += (kCen ? yy - ymax : yy) * a2;
      yns += (kCen ? yy - ymin : yy) * a3;
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
    const T wl = kLse   ? (log(es) + log(ens)) * gam + xmax - xmin + (log(fs) + log(fns)) * gam + ymax - ymin
                 : kCen ? ((xmax - xmin) + xs / es - xns / ens) + ((ymax - ymin) + ys / fs - yns / fns)
                        : xs / es - xns / ens + ys / fs - yns / fns;
    if (lane == 0) {
      if (partial && live) partial[e] = wl;
      if (live) contrib += (double)(w ? wl * w[e] : wl);
    }
    // WA: d/dx = (a + b x) e+ - (a- + b- x) e-; LSE: a = 1/S, b = 0
    const T bx = kLse ? T(0) : ig / es, bnx = kLse ? T(0) : -ig / ens;
    const T by = kLse ? T(0) : ig / fs, bny = kLse ? T(0) : -ig / fns;
    const T ax = kLse ? (T)(1.0 / es) : (T)((1.0 - bx * xs) / es);
    const T anx = kLse ? (T)(1.0 / ens) : (T)((1.0 - bnx * xns) / ens);
    const T ay = kLse ? (T)(1.0 / fs) : (T)((1.0 - by * ys) / fs);
    const T any = kLse ? (T)(1.0 / fns) : (T)((1.0 - bny * yns) / fns);
    for (int j = pb + lane; j < pe; j += 64) {
      const PinMeta m = meta[j];
      T gx = 0, gy = 0;
      if (live && m.x >= 0) {
        const typename Vec2<T>::type v = pin_xy[j];
        const T xx = v.x, yy = v.y;
        if (kLse) {
          gx = exp((xx - xmax) * ig) * ax - exp((xmin - xx) * ig) * anx;
          gy = exp((yy - ymax) * ig) * ay - exp((ymin - yy) * ig) * any;
        } else if (kCen) {
          // the centred form of the packed kernel: 1/S and C = sum u e / S
          gx = (exp((xx - xmax) * ig) * (T(1) / es)) * (T(1) + ((xx - xmax) - xs / es) * ig) -
               (exp((xmin - xx) * ig) * (T(1) / ens)) * (T(1) - ((xx - xmin) - xns / ens) * ig);
          gy = (exp((yy - ymax) * ig) * (T(1) / fs)) * (T(1) + ((yy - ymax) - ys / fs) * ig) -
               (exp((ymin - yy) * ig) * (T(1) / fns)) * (T(1) - ((yy - ymin) - yns / fns) * ig);
        } else {
          gx = (ax + bx * xx) * exp((xx - xmax) * ig) - (anx + bnx * xx) * exp((xmin - xx) * ig);
          gy = (ay + by * yy) * exp((yy - ymax) * ig) - (any + bny * yy) * exp((ymin - yy) * ig);
        }
        if (w) {
          gx *= w[e];
          gy *= w[e];
        }
      }
      store_grad(gnode, m.y, gx, gy);
    }
  }
  const double bs = block_sum<256>(contrib, s_red);
  if (threadIdx.x == 0) partials[blockIdx.x] = bs;
}

// ---- backward --------------------------------------------------------------
// grad[n] = (*grad_scale) * sum of gnode over node n's CSR range + add[n]
// (add nullable, split or interleaved).  Fillers (n >= nphys) get add[n] only.
// kPre: then divided by max(pws[n] + (alpha * *dw) * area[n], 1), the
// preconditioner of PlaceObj.py:69-102 applied to the finished gradient;
// with farea >= 0 the fillers (n >= nphys: no pins, so pws[n] = 0, and one
// area) take farea without loading either array (0 + c * farea is the same
// value).
// Thread t takes nodes t, t + grid, ...; with `partials` it also sums
// (grad - g_prev)^2 over x and y in FP64, in k_step_norms's order (the
// step-size denominator of step_nobb, NesterovAcceleratedGradientOptimizer.py:131).
// fillers_done: the fillers' outputs are already final (the field gather
// wrote them); they are only read back for the sum, or not visited at all.
template <typename T, bool kPre>
__global__ void __launch_bounds__(256) k_node_grad(
    const typename Vec2<T>::type* __restrict__ gnode, const int* __restrict__ node2pin_start,
    const T* __restrict__ grad, const T* __restrict__ add, int add_interleaved, int64_t nphys,
    int64_t nnodes, T* gx_out, T* gy_out, const T* __restrict__ pws, const T* __restrict__ area,
    const T* __restrict__ dw, T alpha, T farea, const T* __restrict__ g_prev, double* partials,
    int fillers_done) {
  __shared__ double smem[4];
  double sgg = 0;
  const T g = *grad;
  const int64_t nend = fillers_done && !partials ? nphys : nnodes;
  for (int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; n < nend;
       n += (int64_t)gridDim.x * blockDim.x) {
    if (fillers_done && n >= nphys) {
      // only with partials: the same terms in the same order
      const double dx = (double)(gx_out[n] - g_prev[n]);
      sgg += dx * dx;
      const double dy = (double)(gy_out[n] - g_prev[nnodes + n]);
      sgg += dy * dy;
      continue;
    }
    T sx = 0, sy = 0;
    if (n < nphys) {
      for (int j = node2pin_start[n]; j < node2pin_start[n + 1]; ++j) {
        const typename Vec2<T>::type v = gnode[j];
        sx += v.x * g;
        sy += v.y * g;
      }
    }
    if (add) {
      if (add_interleaved) {
        const typename Vec2<T>::type v = reinterpret_cast<const typename Vec2<T>::type*>(add)[n];
        sx += v.x;
        sy += v.y;
      } else {
        sx += add[n];
        sy += add[nnodes + n];
      }
    }
    if (kPre) {
      const T c = alpha * *dw;
      T p = (farea >= T(0) && n >= nphys) ? c * farea : pws[n] + c * area[n];
      p = p < T(1) ? T(1) : p;
      sx = sx / p;
      sy = sy / p;
    }
    gx_out[n] = sx;
    gy_out[n] = sy;
    if (partials) {
      const double dx = (double)(sx - g_prev[n]);
      sgg += dx * dx;
      const double dy = (double)(sy - g_prev[nnodes + n]);
      sgg += dy * dy;
    }
  }
  if (partials) {
    const double r = block_sum<256>(sgg, smem);
    if (threadIdx.x == 0) partials[blockIdx.x] = r;
  }
}

static int grid_cap(int64_t items, int per_block, int cap) {
  int64_t g = ceil_div(items, per_block);
  if (g < 1) g = 1;
  return (int)(g > cap ? cap : g);
}

}  // namespace dp

using namespace dp;

extern "C" {

int dp_wa_pack_nets(const int32_t* netpin_start, int64_t num_nets, int32_t* group_start,
                    int64_t* num_groups, int32_t* big_nets, int64_t* num_big, int32_t* net_slot) {
  DP_ARG(netpin_start && group_start && num_groups && big_nets && num_big && net_slot,
         "dp_wa_pack_nets: null pointer");
  DP_ARG(num_nets >= 0 && num_nets < ((int64_t)1 << 31), "dp_wa_pack_nets: bad num_nets");
  // nets by degree (1..64), each bucket in netlist order
  std::vector<int64_t> count(65, 0);
  for (int64_t e = 0; e < num_nets; ++e) {
    const int64_t d = (int64_t)netpin_start[e + 1] - netpin_start[e];
    DP_ARG(d >= 0, "dp_wa_pack_nets: netpin_start not monotone at net %lld", (long long)e);
    if (d <= 64) ++count[d];
  }
  std::vector<int64_t> head(66, 0);
  for (int d = 1; d <= 64; ++d) head[d + 1] = head[d] + count[d];
  std::vector<int32_t> order((size_t)(head[65] > 0 ? head[65] : 1));
  {
    std::vector<int64_t> fill(head.begin(), head.end());
    for (int64_t e = 0; e < num_nets; ++e) {
      const int64_t d = (int64_t)netpin_start[e + 1] - netpin_start[e];
      if (d >= 1 && d <= 64) order[(size_t)fill[d]++] = (int32_t)e;
    }
  }
  std::vector<int64_t> cur(head.begin(), head.end());  // next unplaced net of each bucket
  uint64_t nonempty = 0;                                 // bit d-1: bucket d has nets left
  for (int d = 1; d <= 64; ++d)
    if (cur[d] < head[d + 1]) nonempty |= 1ull << (d - 1);
  // best-fit decreasing: open a group with the largest remaining net, then
  // keep adding the largest remaining net that still fits
  int64_t g = 0, slot = 0;
  group_start[0] = 0;
  while (nonempty) {
    int cap = 64;
    while (cap > 0) {
      const uint64_t fit = cap >= 64 ? nonempty : (nonempty & ((1ull << cap) - 1));
      if (!fit) break;
      const int d = 64 - __builtin_clzll(fit);  // largest degree <= cap with nets left
      const int32_t e = order[(size_t)cur[d]++];
      if (cur[d] == head[d + 1]) nonempty &= ~(1ull << (d - 1));
      net_slot[e] = (int32_t)slot;
      slot += d;
      cap -= d;
    }
    ++g;
    group_start[g] = (int32_t)slot;
  }
  // nets of more than 64 pins after all groups; nets without pins get the
  // end slot (they own no slot)
  int64_t b = 0;
  for (int64_t e = 0; e < num_nets; ++e) {
    const int64_t d = (int64_t)netpin_start[e + 1] - netpin_start[e];
    if (d > 64) {
      big_nets[4 * b] = (int32_t)e;
      big_nets[4 * b + 1] = (int32_t)slot;
      big_nets[4 * b + 2] = (int32_t)d;
      big_nets[4 * b + 3] = 0;
      net_slot[e] = (int32_t)slot;
      slot += d;
      ++b;
    } else if (d == 0) {
      net_slot[e] = (int32_t)slot;
    }
  }
  *num_groups = g;
  *num_big = b;
  return DP_OK;
}

size_t dp_wa_prepare_workspace_bytes(int64_t num_pins) {
  return align256(sizeof(int) * (size_t)(num_pins > 0 ? num_pins : 1));
}

int dp_wa_pins_prepare(int dtype, int64_t num_pins, int64_t num_physical_nodes,
                       const void* pin_offset_x, const void* pin_offset_y,
                       const int32_t* flat_netpin, const int32_t* netpin_start,
                       const int32_t* pin2net_map, const uint8_t* pin_mask,
                       const int32_t* flat_node2pin_map, const int32_t* flat_node2pin_start_map,
                       const int32_t* net_slot, void* workspace, void* node_offsets,
                       int32_t* node_dst, int32_t* slot_node, int32_t* pin_meta, void* stream) {
  DP_ARG(pin_offset_x && pin_offset_y && flat_netpin && netpin_start && pin2net_map &&
             flat_node2pin_map && flat_node2pin_start_map && net_slot && workspace &&
             node_offsets && node_dst && slot_node && pin_meta,
         "dp_wa_pins_prepare: null pointer");
  DP_ARG(num_pins >= 0 && num_pins < ((int64_t)1 << 31) && num_physical_nodes >= 0,
         "dp_wa_pins_prepare: bad counts");
  if (num_pins == 0) return DP_OK;
  hipStream_t s = S(stream);
  int* inv = (int*)workspace;
  hipLaunchKernelGGL(k_prep_net, dim3(grid_for(num_pins, 256)), dim3(256), 0, s,
                     (const int*)flat_netpin, (const int*)netpin_start, (const int*)pin2net_map,
                     pin_mask, (const int*)net_slot, num_pins, (PinMeta*)pin_meta, inv);
  DP_LAUNCHED("k_prep_net");
  hipLaunchKernelGGL(k_prep_slot_node, dim3(grid_for(num_physical_nodes, 256)), dim3(256), 0, s,
                     (const int*)flat_node2pin_start_map, num_physical_nodes, (int*)slot_node);
  DP_LAUNCHED("k_prep_slot_node");
  return DP_DISPATCH(dtype, "dp_wa_pins_prepare", [&] {
    hipLaunchKernelGGL(k_prep_node<T>, dim3(grid_for(num_pins, 256)), dim3(256), 0, s,
                       (const int*)flat_node2pin_map, (const int*)inv, (const T*)pin_offset_x,
                       (const T*)pin_offset_y, num_pins, (int*)node_dst, (PinMeta*)pin_meta,
                       (typename Vec2<T>::type*)node_offsets);
    DP_LAUNCHED("k_prep_node");
    return DP_OK;
  });
}

size_t dp_wa_pins_workspace_bytes(int64_t num_pins, int dtype) {
  const size_t eb = dtype == DP_FLOAT64 ? 16 : 8;
  return reduce_slot_bytes() + align256(eb * (size_t)(num_pins > 0 ? num_pins : 1));
}

static int forward_pins(int mode, const char* name, int dtype, const void* pos, int64_t num_nodes, int64_t num_pins,
                       const int32_t* slot_node, const void* node_offsets,
                       const int32_t* node_dst, const int32_t* pin_meta,
                       const int32_t* group_start, int64_t num_groups, const int32_t* big_nets,
                       int64_t num_big, int64_t num_nets, const void* net_weights,
                       const uint8_t* net_mask, const void* slot_weights, const void* g_p,
                       void* partial_wl, void* net_hpwl, void* grad_node_order, void* wl_out,
                       void* workspace, void* stream) {
  DP_ARG(pos && slot_node && node_offsets && node_dst && pin_meta &&
             (group_start || num_groups == 0) && (big_nets || num_big == 0) && net_mask &&
             g_p && grad_node_order && wl_out && workspace,
         "%s: null pointer", name);
  DP_ARG(num_pins >= 0 && num_pins < ((int64_t)1 << 31) && num_groups >= 0 && num_big >= 0 &&
             num_nets >= 0,
         "%s: bad counts", name);
  hipStream_t s = S(stream);
  DP_PROBE_OP(DP_OP_WA_FORWARD, s);
  double* partials = (double*)workspace;
  char* pin_xy = (char*)workspace + reduce_slot_bytes();
  const int g1 = grid_cap(num_groups, kWaves, kPackedGrid);
  const int g2 = grid_cap(num_big, 4, kBigGrid);
  return DP_DISPATCH(dtype, name, [&] {
    typedef typename Vec2<T>::type V;
    const T* P = (const T*)pos;
    if (partial_wl && num_nets > 0) {  // nets without a live group entry keep 0
      hipLaunchKernelGGL(k_fill<T>, dim3(grid_for(num_nets, 256)), dim3(256), 0, s,
                         (T*)partial_wl, num_nets, T(0));
      DP_LAUNCHED("partial_wl clear");
    }
    if (num_pins > 0) {
      DP_PROBED(DP_KERNEL_PIN_PUSH, s,
                hipLaunchKernelGGL(k_pin_push<T>, dim3((unsigned)ceil_div(num_pins, 256)),
                                   dim3(256), 0, s, P, P + num_nodes, num_pins,
                                   (const int*)slot_node, (const V*)node_offsets,
                                   (const int*)node_dst, (V*)pin_xy));
      DP_LAUNCHED("k_pin_push");
    }
    auto packed = mode == kWaLse       ? k_wa_pins<T, kWaLse>
                  : mode == kWaCentred ? k_wa_pins<T, kWaCentred>
                                       : k_wa_pins<T, kWaRef>;
    auto bignet = mode == kWaLse       ? k_wa_big<T, kWaLse>
                  : mode == kWaCentred ? k_wa_big<T, kWaCentred>
                                       : k_wa_big<T, kWaRef>;
    DP_PROBED(DP_KERNEL_WA_FORWARD, s,
              hipLaunchKernelGGL(packed, dim3(g1), dim3(256), 0, s, (const V*)pin_xy,
                                 (const PinMeta*)pin_meta, (const int*)group_start, (int)num_groups,
                                 (const T*)net_weights, net_mask, (const T*)slot_weights,
                                 (const T*)g_p, (T*)partial_wl, (T*)net_hpwl,
                                 (T*)grad_node_order, partials));
    DP_LAUNCHED("k_wa_pins");
    // without nets above 64 pins the big-net kernel is not launched
    if (num_big == 0) return finish_reduction(partials, g1, 1, nullptr, wl_out, dtype, s);
    hipLaunchKernelGGL(bignet, dim3(g2), dim3(256), 0, s, (const V*)pin_xy,
                       (const PinMeta*)pin_meta, (const int4*)big_nets, (int)num_big,
                       (const T*)net_weights, net_mask, (const T*)g_p, (T*)partial_wl,
                       (T*)net_hpwl, (T*)grad_node_order, partials + g1);
    DP_LAUNCHED("k_wa_big");
    return finish_reduction(partials, g1 + g2, 1, nullptr, wl_out, dtype, s);
  });
}

int dp_wa_forward_pins(int dtype, const void* pos, int64_t num_nodes, int64_t num_pins,
                       const int32_t* slot_node, const void* node_offsets,
                       const int32_t* node_dst, const int32_t* pin_meta,
                       const int32_t* group_start, int64_t num_groups, const int32_t* big_nets,
                       int64_t num_big, int64_t num_nets, const void* net_weights,
                       const uint8_t* net_mask, const void* slot_weights, const void* inv_gamma,
                       void* partial_wl, void* net_hpwl, void* grad_node_order, void* wl_out,
                       void* workspace, void* stream) {
  return forward_pins(kWaRef, "dp_wa_forward_pins", dtype, pos, num_nodes, num_pins, slot_node,
                      node_offsets, node_dst, pin_meta, group_start, num_groups, big_nets, num_big,
                      num_nets, net_weights, net_mask, slot_weights, inv_gamma, partial_wl,
                      net_hpwl, grad_node_order, wl_out, workspace, stream);
}

int dp_wa_forward_pins_centred(int dtype, const void* pos, int64_t num_nodes, int64_t num_pins,
                               const int32_t* slot_node, const void* node_offsets,
                               const int32_t* node_dst, const int32_t* pin_meta,
                               const int32_t* group_start, int64_t num_groups,
                               const int32_t* big_nets, int64_t num_big, int64_t num_nets,
                               const void* net_weights, const uint8_t* net_mask,
                               const void* slot_weights, const void* inv_gamma, void* partial_wl,
                               void* net_hpwl, void* grad_node_order, void* wl_out,
                               void* workspace, void* stream) {
  return forward_pins(kWaCentred, "dp_wa_forward_pins_centred", dtype, pos, num_nodes, num_pins,
                      slot_node, node_offsets, node_dst, pin_meta, group_start, num_groups,
                      big_nets, num_big, num_nets, net_weights, net_mask, slot_weights, inv_gamma,
                      partial_wl, net_hpwl, grad_node_order, wl_out, workspace, stream);
}

int dp_lse_forward_pins(int dtype, const void* pos, int64_t num_nodes, int64_t num_pins,
                        const int32_t* slot_node, const void* node_offsets,
                        const int32_t* node_dst, const int32_t* pin_meta,
                        const int32_t* group_start, int64_t num_groups, const int32_t* big_nets,
                        int64_t num_big, int64_t num_nets, const void* net_weights,
                        const uint8_t* net_mask, const void* slot_weights, const void* gamma,
                        void* partial_wl, void* net_hpwl, void* grad_node_order, void* wl_out,
                        void* workspace, void* stream) {
  return forward_pins(kWaLse, "dp_lse_forward_pins", dtype, pos, num_nodes, num_pins, slot_node,
                      node_offsets, node_dst, pin_meta, group_start, num_groups, big_nets, num_big,
                      num_nets, net_weights, net_mask, slot_weights, gamma, partial_wl,
                      net_hpwl, grad_node_order, wl_out, workspace, stream);
}

int dp_wa_pins_slot_weights(int dtype, int64_t num_pins, int32_t* pin_meta,
                            const void* net_weights, const uint8_t* net_mask, void* slot_weights,
                            void* stream) {
  DP_ARG(pin_meta && net_mask && slot_weights, "dp_wa_pins_slot_weights: null pointer");
  DP_ARG(num_pins >= 0 && num_pins < ((int64_t)1 << 31), "dp_wa_pins_slot_weights: bad count");
  if (num_pins == 0) return DP_OK;
  hipStream_t s = S(stream);
  return DP_DISPATCH(dtype, "dp_wa_pins_slot_weights", [&] {
    hipLaunchKernelGGL(k_slot_weights<T>, dim3(grid_for(num_pins, 256)), dim3(256), 0, s,
                       num_pins, (PinMeta*)pin_meta, (const T*)net_weights, net_mask,
                       (T*)slot_weights);
    DP_LAUNCHED("k_slot_weights");
    return DP_OK;
  });
}

static int node_grad(const char* name, int dtype, const void* grad, const void* grad_node_order,
                     const int32_t* flat_node2pin_start_map, const void* add, int add_layout,
                     int64_t num_physical_nodes, int64_t num_nodes, void* grad_pos_out,
                     const void* pws, const void* area, const void* dw, double alpha,
                     double filler_area, const void* grad_prev, void* workspace,
                     double* dg_norm2_out, void* stream) {
  DP_ARG(grad && grad_node_order && flat_node2pin_start_map && grad_pos_out,
         "%s: null pointer", name);
  DP_ARG(!dg_norm2_out || (grad_prev && workspace), "%s: the norm needs grad_prev and a workspace",
         name);
  DP_ARG(!dg_norm2_out || grad_prev != grad_pos_out, "%s: grad_prev must not be the output", name);
  DP_ARG(num_physical_nodes <= num_nodes, "%s: bad node counts", name);
  const int fillers_done = (add_layout & DP_LAYOUT_FILLERS_FINAL) != 0;
  add_layout &= ~DP_LAYOUT_FILLERS_FINAL;
  DP_ARG(add_layout == DP_LAYOUT_SPLIT || add_layout == DP_LAYOUT_INTERLEAVED,
         "%s: bad add_layout %d", name, add_layout);
  if (num_nodes == 0) return DP_OK;
  hipStream_t s = S(stream);
  DP_PROBE_OP(DP_OP_NODE_GRAD, s);
  return DP_DISPATCH(dtype, name, [&] {
    typedef typename Vec2<T>::type V;
    T* out = (T*)grad_pos_out;
    auto kern = pws ? k_node_grad<T, true> : k_node_grad<T, false>;
    double* part = dg_norm2_out ? (double*)workspace : nullptr;
    // the sum's order is fixed by the grid over all nodes
    const int nb = node_reduce_blocks(fillers_done && !part ? num_physical_nodes : num_nodes);
    DP_PROBED(DP_KERNEL_PIN_POS_BACKWARD, s,
              hipLaunchKernelGGL(kern, dim3(nb), dim3(256), 0, s, (const V*)grad_node_order,
                                 (const int*)flat_node2pin_start_map, (const T*)grad,
                                 (const T*)add, add_layout == DP_LAYOUT_INTERLEAVED,
                                 num_physical_nodes, num_nodes, out, out + num_nodes,
                                 (const T*)pws, (const T*)area, (const T*)dw, (T)alpha,
                                 (T)filler_area, (const T*)grad_prev, part, fillers_done));
    DP_LAUNCHED("k_node_grad");
    if (part) return finish_reduction(part, nb, 1, nullptr, dg_norm2_out, -1, s);
    return DP_OK;
  });
}

int dp_node_grad_from_pins(int dtype, const void* grad, const void* grad_node_order,
                           const int32_t* flat_node2pin_start_map, const void* add,
                           int add_layout, int64_t num_physical_nodes, int64_t num_nodes,
                           void* grad_pos_out, void* stream) {
  return node_grad("dp_node_grad_from_pins", dtype, grad, grad_node_order, flat_node2pin_start_map,
                   add, add_layout, num_physical_nodes, num_nodes, grad_pos_out, nullptr, nullptr,
                   nullptr, 0.0, -1.0, nullptr, nullptr, nullptr, stream);
}

int dp_node_grad_preconditioned(int dtype, const void* grad, const void* grad_node_order,
                                const int32_t* flat_node2pin_start_map, const void* add,
                                int add_layout, int64_t num_physical_nodes, int64_t num_nodes,
                                const void* pin_weight_sum, const void* node_areas,
                                const void* density_weight, double alpha, double filler_area,
                                void* grad_pos_out, const void* grad_prev, void* workspace,
                                double* dg_norm2_out, void* stream) {
  DP_ARG(pin_weight_sum && node_areas && density_weight,
         "dp_node_grad_preconditioned: null pointer");
  return node_grad("dp_node_grad_preconditioned", dtype, grad, grad_node_order,
                   flat_node2pin_start_map, add, add_layout, num_physical_nodes, num_nodes,
                   grad_pos_out, pin_weight_sum, node_areas, density_weight, alpha, filler_area,
                   grad_prev, workspace, dg_norm2_out, stream);
}

}  // extern "C"
