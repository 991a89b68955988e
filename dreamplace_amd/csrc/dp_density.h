// Shared device helpers of the density scatter (density.hip, density_tiled.hip)
// and the field gather (force.hip).
#pragma once

#include "dp_common.h"

#pragma clang fp contract(off)

namespace dp {

template <typename T>
struct TriArgs {
  const T *x, *y, *wxc, *wyc, *offx, *offy, *ratio;
  T xl, yl, bsx, bsy, inv_x, inv_y, scale;
  int nbx, nby;
  // uniform fillers: nodes >= fill_begin all have these clamped sizes,
  // offsets and ratio, which are then not loaded per cell
  int uni;
  int64_t fill_begin;
  T uwx, uwy, uox, uoy, ur;
  // bin-relative overlap arithmetic (DP_ARITH_BIN_RELATIVE; see overlap())
  int acc;
};

// Overlap of [x, x+s) with bin k along one axis:
// min(x+s, bin_xl+b) - max(x, bin_xl), bin_xl = xl + k*b
// (dreamplace/ops/electric_potential/src/density_function.h:12-17).
template <typename T>
__device__ __forceinline__ T tri(T x, T s, T xl, int k, T b) {
  T bxl = xl + (T)k * b;
  T e = x + s, be = bxl + b;
  T hi = (be < e) ? be : e;
  T lo = (x < bxl) ? bxl : x;
  return hi - lo;
}

// x + o == s + err exactly, for s = x + o rounded to T (Knuth's TwoSum; the
// file is compiled without contraction or reassociation).
template <typename T>
__device__ __forceinline__ T two_sum_err(T x, T o, T s) {
  const T bb = s - x;
  return (x - (s - bb)) + (o - bb);
}

// Bin-relative form of tri for bin first + q: the cell's stretched origin
// xs + err (err from two_sum_err) measured from the lower edge of bin
// `first`, d = (xs - (xl + first*b)) + err, then
// min(d + s, q*b + b) - max(d, q*b).  Every operand is of the order of a few
// bins, where tri's x + s rounds at the order of the die coordinates
// (2.4e-4 at 5000 units in FP32, ~2.5e-5 of a bin of 10 units).
template <typename T>
__device__ __forceinline__ T rel_origin(T xs, T err, T xl, int first, T b) {
  return (xs - (xl + (T)first * b)) + err;
}
template <typename T>
__device__ __forceinline__ T tri_rel(T d, T s, int q, T b) {
  const T lo_q = (T)q * b;
  const T e = d + s, be = lo_q + b;
  T hi = (be < e) ? be : e;
  T lo = (d < lo_q) ? lo_q : d;
  return hi - lo;
}
// Overlap of a cell's span with bin first + q along one axis, in the
// arithmetic `acc` selects: tri (the reference's) or tri_rel with d from
// rel_origin.
template <typename T>
__device__ __forceinline__ T overlap(int acc, T xs, T d, T s, T xl, int first, int q, T b) {
  return acc ? tri_rel(d, s, q, b) : tri(xs, s, xl, first + q, b);
}

// Exact overlap with a bin given its centre, no boundary stretch
// (density_function.h:19-36 with flag = true), clamped at 0.
template <typename T>
__device__ __forceinline__ T exact_overlap(T x, T s, T bc, T b) {
  T bxl = bc - b / 2, bxh = bc + b / 2;
  T e = x + s;
  T hi = (bxh < e) ? bxh : e;
  T lo = (x < bxl) ? bxl : x;
  T v = hi - lo;
  return (T(0) < v) ? v : T(0);
}
// exact_overlap in the bin-relative arithmetic: d = (x - bc) + b/2, then
// min(d + s, b) - max(d, 0), clamped at 0.
template <typename T>
__device__ __forceinline__ T exact_overlap_rel(T x, T s, T bc, T b) {
  T d = (x - bc) + b / 2;
  T e = d + s;
  T hi = (b < e) ? b : e;
  T lo = (d < T(0)) ? T(0) : d;
  T v = hi - lo;
  return (T(0) < v) ? v : T(0);
}

// A cell's stretched origin (x, y), its TwoSum residuals (ex, ey; 0 unless
// TriArgs::acc), clamped sizes, ratio and bin range.
template <typename T>
struct CellBins {
  T x, y, sx, sy, r, ex, ey;
  int bxl, bxh, byl, byh;
};

// The seven per-cell inputs of the scatter, as loaded.
template <typename T>
struct RawCell {
  T x, y, wx, wy, ox, oy, r;
};

template <typename T>
__device__ __forceinline__ RawCell<T> load_cell(const TriArgs<T>& a, int64_t i) {
  RawCell<T> c;
  c.x = a.x[i];
  c.y = a.y[i];
  if (a.uni && i >= a.fill_begin) {
    c.wx = a.uwx;
    c.wy = a.uwy;
    c.ox = a.uox;
    c.oy = a.uoy;
    c.r = a.ur;
  } else {
    c.wx = a.wxc[i];
    c.wy = a.wyc[i];
    c.ox = a.offx[i];
    c.oy = a.offy[i];
    c.r = a.ratio[i];
  }
  return c;
}

// Stretched cell origin and its clamped bin range
// (electric_density_map.cpp:274-293: int() truncation, exclusive upper end).
template <typename T>
__device__ __forceinline__ CellBins<T> bins_of(const TriArgs<T>& a, const RawCell<T>& v) {
  CellBins<T> c;
  c.sx = v.wx;
  c.sy = v.wy;
  c.x = v.x + v.ox;
  c.y = v.y + v.oy;
  c.ex = a.acc ? two_sum_err(v.x, v.ox, c.x) : T(0);
  c.ey = a.acc ? two_sum_err(v.y, v.oy, c.y) : T(0);
  c.r = v.r;
  c.bxl = max((int)((c.x - a.xl) * a.inv_x), 0);
  c.bxh = min((int)((c.x + c.sx - a.xl) * a.inv_x) + 1, a.nbx);
  c.byl = max((int)((c.y - a.yl) * a.inv_y), 0);
  c.byh = min((int)((c.y + c.sy - a.yl) * a.inv_y) + 1, a.nby);
  return c;
}

template <typename T>
__device__ __forceinline__ CellBins<T> cell_bins(const TriArgs<T>& a, int64_t i) {
  return bins_of(a, load_cell(a, i));
}

// Store a cell's gradient: interleaved (out[2i], out[2i+1]) as one 8/16-byte
// word, or split (out[i], out[num_nodes+i]) as the reference lays out [2N].
template <typename T>
__device__ __forceinline__ void put_grad(T* out, int64_t num_nodes, int64_t i, T gx, T gy,
                                         int interleaved) {
  if (interleaved) {
    typename Vec2<T>::type v;
    v.x = gx;
    v.y = gy;
    reinterpret_cast<typename Vec2<T>::type*>(out)[i] = v;
  } else {
    out[i] = gx;
    out[num_nodes + i] = gy;
  }
}

// Where the tiled gather puts a cell's gradient: put_grad into `out`, except
// that with `fin` the fillers (i >= f0) get their finished objective
// gradient in `fin` (split [2n]) instead: 0 + g, as the node-gradient pass
// adds a pinless node's density gradient to its zero wirelength sum, then,
// with `dw`, divided by max(alpha * (*dw) * farea, 1), that pass's
// preconditioner for a node of area farea and no pins.  Same operations in
// the same order, so the same bits as the two passes in sequence.
template <typename T>
struct GradOut {
  T* out;
  int64_t n;
  int interleaved;
  T* fin;
  int64_t f0;
  const T* dw;
  T alpha, farea;
};

// The preconditioner's density factor alpha * (*dw), read once per thread.
template <typename T>
__device__ __forceinline__ T grad_out_coef(const GradOut<T>& o) {
  return o.fin && o.dw ? o.alpha * *o.dw : T(0);
}

template <typename T>
__device__ __forceinline__ void put_cell_grad(const GradOut<T>& o, T c, int64_t i, T gx, T gy) {
  if (o.fin && i >= o.f0) {
    T vx = T(0), vy = T(0);
    vx += gx;
    vy += gy;
    if (o.dw) {
      T p = c * o.farea;
      p = p < T(1) ? T(1) : p;
      vx = vx / p;
      vy = vy / p;
    }
    o.fin[i] = vx;
    o.fin[o.n + i] = vy;
  } else {
    put_grad(o.out, o.n, i, gx, gy, o.interleaved);
  }
}

template <typename T>
TriArgs<T> make_tri_args(const void* pos, int64_t num_nodes, const void* wxc, const void* wyc,
                         const void* offx, const void* offy, const void* ratio, double xl,
                         double yl, double bsx, double bsy, int nbx, int nby, int64_t scale,
                         const double* uniform_filler, int64_t fill_begin);

// Tiled (LDS-binned) scatter and gather, density_tiled.hip.
bool tiled_supported(int nbx, int nby);
// acc: the records also hold the stretched origins' TwoSum residuals
size_t tiled_ws_bytes(int nbx, int nby, int64_t ncells, int dtype, bool acc);
// filler_map (nullable): uniform fillers' contributions go there instead of
// `map`, which then holds the other cells' only (requires uniform fillers).
// filler_ranges (nullable, device int[ntiles + 1]): fillers kept in tile
// order, scattered (and later gathered) from their positions by tile range
// instead of being binned.
// zero_map: clear `map` first (with the tiling counters, in one launch);
// filler_map, when given, is always cleared there.
template <typename T>
int scatter_tiled(const TriArgs<T>& a, int64_t n_mov, int64_t fill_begin, int64_t n_fill,
                  char* tiled_ws, unsigned long long* map, unsigned long long* filler_map,
                  const int* filler_ranges, bool zero_map, hipStream_t s);
template <typename T>
int gather_tiled(const TriArgs<T>& a, int64_t n_mov, int64_t n_fill, const T* fx, const T* fy,
                 const T* grad, char* tiled_ws, const GradOut<T>& go, hipStream_t s);

// Density workspace layout: [int64 map][int64 filler map][queue count (256 B)]
// [queue][tiled region].
inline size_t density_map_bytes(int64_t nb) { return align256(8 * (size_t)nb); }
inline size_t density_queue_offset(int64_t nb) { return 2 * density_map_bytes(nb); }
inline size_t density_tiled_offset(int64_t nb, int64_t qcap) {
  return density_queue_offset(nb) + 256 + align256(4 * (size_t)(qcap > 0 ? qcap : 1));
}

}  // namespace dp
