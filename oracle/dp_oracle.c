/*
 * dp_oracle.c -- CPU restatement of DREAMPlace's global-placement inner loop.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the oracle: tests/, the smoke check
 * and the CPU-baseline leg of bench.py may call it, as the checker.  The
 * product path (dreamplace_amd) never links, loads or calls it.
 *
 * Every function restates one reference routine; the citation names the
 * reference file:line whose arithmetic it follows.  Floating-point evaluation
 * order and conversions are kept as the reference writes them (the file is
 * compiled with -ffp-contract=off), because the fixed-point density map must
 * be bit-exact.
 *
 * Each routine exists for T = float (suffix _f32) and T = double (_f64).
 */
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#pragma STDC FP_CONTRACT OFF

static inline int imax(int a, int b) { return a > b ? a : b; }
static inline int imin(int a, int b) { return a < b ? a : b; }

/* Fixed-point scale: 2^(64 - max(ceil(log2(die_area)) + 1, 32)).
 * dreamplace/ops/electric_potential/src/electric_density_map.cpp:117-123. */
int64_t dpo_fixed_scale(double xl, double yl, double xh, double yh) {
  double diearea = (xh - xl) * (yh - yl);
  int integer_bits = imax((int)ceil(log2(diearea)) + 1, 32);
  int fraction_bits = imax(64 - integer_bits, 0);
  return (int64_t)1 << fraction_bits;
}

#define DPO_DEFINE(T, SFX)                                                                  \
  /* density_function.h:12-17 */                                                            \
  static inline T tri_##SFX(T x, T node_size, T xl, int k, T bin_size) {                   \
    T bin_xl = xl + k * bin_size;                                                            \
    T hi = (x + node_size < bin_xl + bin_size) ? x + node_size : bin_xl + bin_size;          \
    T lo = (x > bin_xl) ? x : bin_xl;                                                        \
    return hi - lo;                                                                          \
  }                                                                                          \
  /* density_function.h:19-36 */                                                            \
  static inline T exact_##SFX(T x, T node_size, T bin_center, T bin_size, T l, T h,          \
                              int flag) {                                                    \
    T bin_xl = bin_center - bin_size / 2;                                                    \
    T bin_xh = bin_center + bin_size / 2;                                                    \
    if (!flag) {                                                                             \
      if (bin_xl <= l) bin_xl = (bin_xl < x) ? bin_xl : x;                                   \
      if (bin_xh >= h) bin_xh = (bin_xh > x + node_size) ? bin_xh : x + node_size;           \
    }                                                                                        \
    T hi = (x + node_size < bin_xh) ? x + node_size : bin_xh;                                \
    T lo = (x > bin_xl) ? x : bin_xl;                                                        \
    T v = hi - lo;                                                                           \
    return (v > (T)0) ? v : (T)0;                                                            \
  }                                                                                          \
  /* Triangle (stretched-cell) scatter into an int64 fixed-point map.                     \
   * electric_density_map.cpp:254-309 with AtomicAdd<long> (atomic_ops.h:29-45):          \
   * each (cell, bin) adds (long)(((px*ratio)*py) * (T)scale). */                          \
  void dpo_triangle_density_##SFX(const T* xv, const T* yv, const T* wxc, const T* wyc,    \
                                  const T* offx, const T* offy, const T* ratio_v,            \
                                  int64_t n, int nbx, int nby, double xl_d, double yl_d,     \
                                  double bsx_d, double bsy_d, int64_t scale,                 \
                                  int64_t* buf) {                                            \
    const T xl = (T)xl_d, yl = (T)yl_d, bsx = (T)bsx_d, bsy = (T)bsy_d;                     \
    const T inv_x = (T)(1.0 / bsx), inv_y = (T)(1.0 / bsy);                                  \
    const T tscale = (T)scale;                                                               \
    _Pragma("omp parallel for schedule(static)")                                            \
    for (int64_t i = 0; i < n; ++i) {                                                        \
      T sx = wxc[i], sy = wyc[i];                                                            \
      T nx = xv[i] + offx[i], ny = yv[i] + offy[i];                                          \
      T r = ratio_v[i];                                                                      \
      int bxl = imax((int)((nx - xl) * inv_x), 0);                                           \
      int bxh = imin((int)((nx + sx - xl) * inv_x) + 1, nbx);                                \
      int byl = imax((int)((ny - yl) * inv_y), 0);                                           \
      int byh = imin((int)((ny + sy - yl) * inv_y) + 1, nby);                                \
      for (int k = bxl; k < bxh; ++k) {                                                      \
        T pxr = tri_##SFX(nx, sx, xl, k, bsx) * r;                                           \
        for (int h = byl; h < byh; ++h) {                                                    \
          T area = pxr * tri_##SFX(ny, sy, yl, h, bsy);                                      \
          int64_t sv = (int64_t)(area * tscale);                                             \
          _Pragma("omp atomic")                                                              \
          buf[(int64_t)k * nby + h] += sv;                                                   \
        }                                                                                    \
      }                                                                                      \
    }                                                                                        \
  }                                                                                          \
  /* Exact-overlap scatter of fixed cells into the int64 map.                              \
   * electric_density_map.cpp:311-358 (box2bin), fixed_node_flag as passed at :84-93. */   \
  void dpo_exact_density_##SFX(const T* xv, const T* yv, const T* wx, const T* wy,         \
                               const T* bcx, const T* bcy, int64_t n, int nbx, int nby,      \
                               double xl_d, double yl_d, double xh_d, double yh_d,           \
                               double bsx_d, double bsy_d, int fixed_flag, int64_t scale,    \
                               int64_t* buf) {                                               \
    const T xl = (T)xl_d, yl = (T)yl_d, xh = (T)xh_d, yh = (T)yh_d;                          \
    const T bsx = (T)bsx_d, bsy = (T)bsy_d;                                                  \
    const T tscale = (T)scale;                                                               \
    _Pragma("omp parallel for schedule(static)")                                            \
    for (int64_t i = 0; i < n; ++i) {                                                        \
      T bxl_ = xv[i], byl_ = yv[i];                                                          \
      T bxh_ = bxl_ + wx[i], byh_ = byl_ + wy[i];                                            \
      int kxl = imax((int)((bxl_ - xl) / bsx), 0);                                           \
      int kxh = imin((int)ceil((bxh_ - xl) / bsx) + 1, nbx);                                 \
      int kyl = imax((int)((byl_ - yl) / bsy), 0);                                           \
      int kyh = imin((int)ceil((byh_ - yl) / bsy) + 1, nby);                                 \
      for (int k = kxl; k < kxh; ++k) {                                                      \
        T px = exact_##SFX(bxl_, bxh_ - bxl_, bcx[k], bsx, xl, xh, fixed_flag);              \
        for (int h = kyl; h < kyh; ++h) {                                                    \
          T py = exact_##SFX(byl_, byh_ - byl_, bcy[h], bsy, yl, yh, fixed_flag);            \
          T area = px * py;                                                                  \
          int64_t sv = (int64_t)(area * tscale);                                             \
          _Pragma("omp atomic")                                                              \
          buf[(int64_t)k * nby + h] += sv;                                                   \
        }                                                                                    \
      }                                                                                      \
    }                                                                                        \
  }                                                                                          \
  /* CPU-path conversion: dst[i] += src[i] * (1.0/scale), in double (atomic_ops.h:48-54). */ \
  void dpo_fixed_add_cpu_##SFX(const int64_t* buf, int64_t nb, int64_t scale, T* dst) {    \
    double s = 1.0 / (double)scale;                                                          \
    for (int64_t i = 0; i < nb; ++i) dst[i] += (double)buf[i] * s;                           \
  }                                                                                          \
  /* CUDA-path conversions (atomic_ops.cuh:53-59 as used at                                 \
   * electric_density_map_cuda_kernel.cu:447-461): quantise the initial map with            \
   * u64(init*T(scale)), then map = T(u64) * T(1.0/scale). */                                \
  void dpo_fixed_quantize_gpu_##SFX(const T* init, int64_t nb, int64_t scale,              \
                                    int64_t* buf) {                                          \
    for (int64_t i = 0; i < nb; ++i)                                                         \
      buf[i] = (int64_t)(uint64_t)(init[i] * (T)scale);                                      \
  }                                                                                          \
  void dpo_fixed_to_map_gpu_##SFX(const int64_t* buf, int64_t nb, int64_t scale, T* dst) { \
    T s = (T)(1.0 / (double)scale);                                                          \
    for (int64_t i = 0; i < nb; ++i) dst[i] = (T)(uint64_t)buf[i] * s;                       \
  }                                                                                          \
  /* Electric force for a range of cells, times the incoming scalar gradient.               \
   * electric_force.cpp:116-172 (k-outer, h-inner accumulation in T, then *ratio),          \
   * :96 grad_out.mul_(grad_pos).  The caller negates (electric_potential.py:223). */       \
  void dpo_electric_force_##SFX(T grad, const T* fx, const T* fy, const T* xv, const T* yv, \
                                const T* wxc, const T* wyc, const T* offx, const T* offy,    \
                                const T* ratio_v, int64_t n, int nbx, int nby, double xl_d,  \
                                double yl_d, double bsx_d, double bsy_d, T* gxo, T* gyo) {   \
    const T xl = (T)xl_d, yl = (T)yl_d, bsx = (T)bsx_d, bsy = (T)bsy_d;                     \
    const T inv_x = (T)(1.0 / bsx), inv_y = (T)(1.0 / bsy);                                  \
    _Pragma("omp parallel for schedule(static)")                                            \
    for (int64_t i = 0; i < n; ++i) {                                                        \
      T sx = wxc[i], sy = wyc[i];                                                            \
      T nx = xv[i] + offx[i], ny = yv[i] + offy[i];                                          \
      T r = ratio_v[i];                                                                      \
      int bxl = imax((int)((nx - xl) * inv_x), 0);                                           \
      int bxh = imin((int)((nx + sx - xl) * inv_x) + 1, nbx);                                \
      int byl = imax((int)((ny - yl) * inv_y), 0);                                           \
      int byh = imin((int)((ny + sy - yl) * inv_y) + 1, nby);                                \
      T gx = 0, gy = 0;                                                                      \
      for (int k = bxl; k < bxh; ++k) {                                                      \
        T px = tri_##SFX(nx, sx, xl, k, bsx);                                                \
        for (int h = byl; h < byh; ++h) {                                                    \
          T py = tri_##SFX(ny, sy, yl, h, bsy);                                              \
          T area = px * py;                                                                  \
          int64_t idx = (int64_t)k * nby + h;                                                \
          gx += area * fx[idx];                                                              \
          gy += area * fy[idx];                                                              \
        }                                                                                    \
      }                                                                                      \
      gx *= r;                                                                               \
      gy *= r;                                                                               \
      gxo[i] = gx * grad;                                                                    \
      gyo[i] = gy * grad;                                                                    \
    }                                                                                        \
  }                                                                                          \
  /* d/dx of tri_(x, ...) from the right: [x + s < bin_xh] - [x >= bin_xl], the          \
   * derivatives of the two sides of density_function.h:12-17. */                          \
  static inline T dtri_##SFX(T x, T node_size, T xl, int k, T bin_size) {                  \
    T bin_xl = xl + k * bin_size;                                                            \
    T up = (x + node_size < bin_xl + bin_size) ? (T)1 : (T)0;                                \
    T dn = (x >= bin_xl) ? (T)1 : (T)0;                                                      \
    return up - dn;                                                                          \
  }                                                                                          \
  /* Exact gradient of the density energy E = sum_b phi_b*rho_b of                          \
   * electric_potential.py:137-186 (rho = D/(bsx*bsy); phi = L rho with L symmetric, so      \
   * dE/dD_b = 2*phi_b/(bsx*bsy)), through each cell's triangle overlaps                    \
   * (density_function.h:12-17) over the scatter's bin range                                \
   * (electric_density_map.cpp:274-293):                                                     \
   *   g_x[i] = grad * (2/(bsx*bsy)) * ratio_i * sum_{k,h} dpx_k * py_h * phi[k,h]           \
   * and g_y with px_k * dpy_h.  The reference's backward is the force instead              \
   * (electric_force.cpp); no reference routine computes this one.  Same loop order         \
   * as the force (k outer, h inner, accumulation in T, then * ratio). */                   \
  void dpo_energy_gradient_##SFX(T grad, const T* phi, const T* xv, const T* yv,           \
                                 const T* wxc, const T* wyc, const T* offx, const T* offy,   \
                                 const T* ratio_v, int64_t n, int nbx, int nby, double xl_d, \
                                 double yl_d, double bsx_d, double bsy_d, T* gxo, T* gyo) {  \
    const T xl = (T)xl_d, yl = (T)yl_d, bsx = (T)bsx_d, bsy = (T)bsy_d;                     \
    const T inv_x = (T)(1.0 / bsx), inv_y = (T)(1.0 / bsy);                                  \
    const T coef = (T)(2.0 / (bsx_d * bsy_d));                                               \
    _Pragma("omp parallel for schedule(static)")                                            \
    for (int64_t i = 0; i < n; ++i) {                                                        \
      T sx = wxc[i], sy = wyc[i];                                                            \
      T nx = xv[i] + offx[i], ny = yv[i] + offy[i];                                          \
      T r = ratio_v[i];                                                                      \
      int bxl = imax((int)((nx - xl) * inv_x), 0);                                           \
      int bxh = imin((int)((nx + sx - xl) * inv_x) + 1, nbx);                                \
      int byl = imax((int)((ny - yl) * inv_y), 0);                                           \
      int byh = imin((int)((ny + sy - yl) * inv_y) + 1, nby);                                \
      T gx = 0, gy = 0;                                                                      \
      for (int k = bxl; k < bxh; ++k) {                                                      \
        T px = tri_##SFX(nx, sx, xl, k, bsx);                                                \
        T dpx = dtri_##SFX(nx, sx, xl, k, bsx);                                              \
        for (int h = byl; h < byh; ++h) {                                                    \
          T py = tri_##SFX(ny, sy, yl, h, bsy);                                              \
          T dpy = dtri_##SFX(ny, sy, yl, h, bsy);                                            \
          T p = phi[(int64_t)k * nby + h];                                                   \
          gx += (dpx * py) * p;                                                              \
          gy += (px * dpy) * p;                                                              \
        }                                                                                    \
      }                                                                                      \
      gxo[i] = ((gx * r) * coef) * grad;                                                     \
      gyo[i] = ((gy * r) * coef) * grad;                                                     \
    }                                                                                        \
  }                                                                                          \
  /* Merged WA forward: weighted_average_wirelength_merged.cpp:31-116.                      \
   * partial_wl[i] is the pre-weight x+y wirelength of net i (0 when masked);               \
   * gi_x/gi_y receive d(WL_i)/d(pin) for every pin of an unmasked net. */                 \
  void dpo_wa_merged_fwd_##SFX(const T* px, const T* py, const int32_t* flat_netpin,       \
                               const int32_t* netpin_start, const uint8_t* net_mask,         \
                               int64_t num_nets, T inv_gamma, T* partial_wl, T* gi_x,        \
                               T* gi_y) {                                                    \
    _Pragma("omp parallel for schedule(dynamic, 256)")                                      \
    for (int64_t i = 0; i < num_nets; ++i) {                                                 \
      partial_wl[i] = 0;                                                                     \
      if (!net_mask[i]) continue;                                                            \
      int b = netpin_start[i], e = netpin_start[i + 1];                                      \
      const T tmax = (T)(sizeof(T) == 4 ? (double)FLT_MAX : DBL_MAX);                        \
      T x_max = -tmax, x_min = tmax, y_max = -tmax, y_min = tmax;                            \
      for (int j = b; j < e; ++j) {                                                          \
        T xx = px[flat_netpin[j]], yy = py[flat_netpin[j]];                                  \
        x_max = xx > x_max ? xx : x_max;                                                     \
        x_min = xx < x_min ? xx : x_min;                                                     \
        y_max = yy > y_max ? yy : y_max;                                                     \
        y_min = yy < y_min ? yy : y_min;                                                     \
      }                                                                                      \
      T xs = 0, xns = 0, es = 0, ens = 0, ys = 0, yns = 0, fs = 0, fns = 0;                  \
      for (int j = b; j < e; ++j) {                                                          \
        T xx = px[flat_netpin[j]], yy = py[flat_netpin[j]];                                  \
        T ex = exp((xx - x_max) * inv_gamma), enx = exp((x_min - xx) * inv_gamma);           \
        xs += xx * ex; xns += xx * enx; es += ex; ens += enx;                                \
        T ey = exp((yy - y_max) * inv_gamma), eny = exp((y_min - yy) * inv_gamma);           \
        ys += yy * ey; yns += yy * eny; fs += ey; fns += eny;                                \
      }                                                                                      \
      partial_wl[i] = xs / es - xns / ens + ys / fs - yns / fns;                             \
      T bx = inv_gamma / es, ax = (1.0 - bx * xs) / es;                                      \
      T bnx = -inv_gamma / ens, anx = (1.0 - bnx * xns) / ens;                               \
      T by = inv_gamma / fs, ay = (1.0 - by * ys) / fs;                                      \
      T bny = -inv_gamma / fns, any = (1.0 - bny * yns) / fns;                               \
      for (int j = b; j < e; ++j) {                                                          \
        int p = flat_netpin[j];                                                              \
        T xx = px[p], yy = py[p];                                                            \
        T ex = exp((xx - x_max) * inv_gamma), enx = exp((x_min - xx) * inv_gamma);           \
        gi_x[p] = (ax + bx * xx) * ex - (anx + bnx * xx) * enx;                              \
        T ey = exp((yy - y_max) * inv_gamma), eny = exp((y_min - yy) * inv_gamma);           \
        gi_y[p] = (ay + by * yy) * ey - (any + bny * yy) * eny;                              \
      }                                                                                      \
    }                                                                                        \
  }                                                                                          \
  /* pin_pos forward: pin_pos.cpp:133-146. */                                               \
  void dpo_pin_pos_fwd_##SFX(const T* xv, const T* yv, const T* offx, const T* offy,       \
                             const int64_t* pin2node, int64_t num_pins, T* pin_x,            \
                             T* pin_y) {                                                     \
    _Pragma("omp parallel for schedule(static)")                                            \
    for (int64_t i = 0; i < num_pins; ++i) {                                                 \
      int64_t nd = pin2node[i];                                                              \
      pin_x[i] = offx[i] + xv[nd];                                                           \
      pin_y[i] = offy[i] + yv[nd];                                                           \
    }                                                                                        \
  }                                                                                          \
  /* pin_pos backward: pin_pos.cpp:148-167 (CSR sum per physical node). */                  \
  void dpo_pin_pos_bwd_##SFX(const T* gpx, const T* gpy, const int32_t* flat_node2pin,     \
                             const int32_t* node2pin_start, int64_t num_physical_nodes,      \
                             T* gx, T* gy) {                                                 \
    _Pragma("omp parallel for schedule(static)")                                            \
    for (int64_t i = 0; i < num_physical_nodes; ++i) {                                       \
      T sx = 0, sy = 0;                                                                      \
      for (int j = node2pin_start[i]; j < node2pin_start[i + 1]; ++j) {                      \
        sx += gpx[flat_node2pin[j]];                                                         \
        sy += gpy[flat_node2pin[j]];                                                         \
      }                                                                                      \
      gx[i] = sx;                                                                            \
      gy[i] = sy;                                                                            \
    }                                                                                        \
  }                                                                                          \
  /* Per-net HPWL (unweighted) for unmasked nets: hpwl.cpp:60-84. */                        \
  void dpo_hpwl_##SFX(const T* px, const T* py, const int32_t* flat_netpin,                \
                      const int32_t* netpin_start, const uint8_t* net_mask, int64_t num_nets,\
                      T* hpwl) {                                                             \
    _Pragma("omp parallel for schedule(static)")                                            \
    for (int64_t i = 0; i < num_nets; ++i) {                                                 \
      hpwl[i] = 0;                                                                           \
      if (!net_mask[i] || netpin_start[i + 1] <= netpin_start[i]) continue;                  \
      int b = netpin_start[i];                                                               \
      T x0 = px[flat_netpin[b]], x1 = x0, y0 = py[flat_netpin[b]], y1 = y0;                  \
      for (int j = b; j < netpin_start[i + 1]; ++j) {                                        \
        T xx = px[flat_netpin[j]], yy = py[flat_netpin[j]];                                  \
        x0 = xx < x0 ? xx : x0; x1 = xx > x1 ? xx : x1;                                      \
        y0 = yy < y0 ? yy : y0; y1 = yy > y1 ? yy : y1;                                      \
      }                                                                                      \
      hpwl[i] = x1 - x0 + y1 - y0;                                                           \
    }                                                                                        \
  }                                                                                          \
  /* move_boundary: move_boundary.cpp:42-62 (movables and fillers, in place). */            \
  void dpo_move_boundary_##SFX(T* xv, T* yv, const T* wx, const T* wy, double xl_d,         \
                               double yl_d, double xh_d, double yh_d, int64_t num_nodes,     \
                               int64_t num_movable, int64_t num_filler) {                    \
    const T xl = (T)xl_d, yl = (T)yl_d, xh = (T)xh_d, yh = (T)yh_d;                          \
    for (int64_t i = 0; i < num_nodes; ++i) {                                                \
      if (i < num_movable || i >= num_nodes - num_filler) {                                  \
        T x = xv[i] > xl ? xv[i] : xl;                                                       \
        T xm = xh - wx[i];                                                                   \
        xv[i] = x < xm ? x : xm;                                                             \
        T y = yv[i] > yl ? yv[i] : yl;                                                       \
        T ym = yh - wy[i];                                                                   \
        yv[i] = y < ym ? y : ym;                                                             \
      }                                                                                      \
    }                                                                                        \
  }                                                                                          \
  /* pin weight sum: pws.cpp:37-48. */                                                      \
  void dpo_pws_##SFX(const T* net_weights, const int32_t* flat_nodepin,                    \
                     const int32_t* nodepin_start, const int32_t* pin2net,                   \
                     int64_t num_physical_nodes, T* node_weights) {                          \
    for (int64_t i = 0; i < num_physical_nodes; ++i) {                                       \
      T s = 0;                                                                               \
      for (int j = nodepin_start[i]; j < nodepin_start[i + 1]; ++j)                          \
        s += net_weights[pin2net[flat_nodepin[j]]];                                          \
      node_weights[i] += s;                                                                  \
    }                                                                                        \
  }

/* Bin-relative overlap arithmetic: this build's DP_ARITH_BIN_RELATIVE option,
 * which has no counterpart in the reference (parity of these functions is
 * pinned by mathematics in tests/test_oracle.py: their FP64 maps and forces
 * equal the reference arithmetic's to rounding, and their FP32 maps stay
 * within 1e-6 of FP64 where the reference arithmetic's do not).  Bins and the
 * fixed-point recipe are the reference's (electric_density_map.cpp:254-309);
 * only the overlap changes: the stretched origin xs = x + offset is carried
 * with its TwoSum residual err and measured from the lower edge of the
 * cell's first bin, d = (xs - (xl + bxl*b)) + err, and the overlap with bin
 * bxl + q is min(d + s, q*b + b) - max(d, q*b). */
#define DPO_DEFINE_REL(T, SFX)                                                              \
  static inline T twosum_err_##SFX(T a, T b, T s) {                                          \
    T bb = s - a;                                                                            \
    return (a - (s - bb)) + (b - bb);                                                        \
  }                                                                                          \
  static inline T tri_rel_##SFX(T d, T node_size, int q, T bin_size) {                       \
    T lo_q = (T)q * bin_size;                                                                \
    T e = d + node_size, be = lo_q + bin_size;                                               \
    T hi = (be < e) ? be : e;                                                                \
    T lo = (d < lo_q) ? lo_q : d;                                                            \
    return hi - lo;                                                                          \
  }                                                                                          \
  void dpo_triangle_density_rel_##SFX(const T* xv, const T* yv, const T* wxc, const T* wyc,\
                                      const T* offx, const T* offy, const T* ratio_v,        \
                                      int64_t n, int nbx, int nby, double xl_d, double yl_d, \
                                      double bsx_d, double bsy_d, int64_t scale,             \
                                      int64_t* buf) {                                        \
    const T xl = (T)xl_d, yl = (T)yl_d, bsx = (T)bsx_d, bsy = (T)bsy_d;                     \
    const T inv_x = (T)(1.0 / bsx), inv_y = (T)(1.0 / bsy);                                  \
    const T tscale = (T)scale;                                                               \
    _Pragma("omp parallel for schedule(static)")                                            \
    for (int64_t i = 0; i < n; ++i) {                                                        \
      T sx = wxc[i], sy = wyc[i];                                                            \
      T nx = xv[i] + offx[i], ny = yv[i] + offy[i];                                          \
      T ex = twosum_err_##SFX(xv[i], offx[i], nx), ey = twosum_err_##SFX(yv[i], offy[i], ny);\
      T r = ratio_v[i];                                                                      \
      int bxl = imax((int)((nx - xl) * inv_x), 0);                                           \
      int bxh = imin((int)((nx + sx - xl) * inv_x) + 1, nbx);                                \
      int byl = imax((int)((ny - yl) * inv_y), 0);                                           \
      int byh = imin((int)((ny + sy - yl) * inv_y) + 1, nby);                                \
      T dx = (nx - (xl + (T)bxl * bsx)) + ex, dy = (ny - (yl + (T)byl * bsy)) + ey;          \
      for (int k = bxl; k < bxh; ++k) {                                                      \
        T pxr = tri_rel_##SFX(dx, sx, k - bxl, bsx) * r;                                     \
        for (int h = byl; h < byh; ++h) {                                                    \
          T area = pxr * tri_rel_##SFX(dy, sy, h - byl, bsy);                                \
          int64_t sv = (int64_t)(area * tscale);                                             \
          _Pragma("omp atomic")                                                              \
          buf[(int64_t)k * nby + h] += sv;                                                   \
        }                                                                                    \
      }                                                                                      \
    }                                                                                        \
  }                                                                                          \
  /* exact overlap of a fixed cell with bin (centre bc): d = (x - bc) + b/2, then         \
   * min(d + s, b) - max(d, 0), clamped at 0 (density_function.h:19-36, flag true). */    \
  static inline T exact_rel_##SFX(T x, T node_size, T bin_center, T bin_size) {             \
    T d = (x - bin_center) + bin_size / 2;                                                   \
    T e = d + node_size;                                                                     \
    T hi = (bin_size < e) ? bin_size : e;                                                    \
    T lo = (d < (T)0) ? (T)0 : d;                                                            \
    T v = hi - lo;                                                                           \
    return ((T)0 < v) ? v : (T)0;                                                            \
  }                                                                                          \
  void dpo_exact_density_rel_##SFX(const T* xv, const T* yv, const T* wx, const T* wy,     \
                                   const T* bcx, const T* bcy, int64_t n, int nbx, int nby,  \
                                   double xl_d, double yl_d, double bsx_d, double bsy_d,     \
                                   int64_t scale, int64_t* buf) {                            \
    const T xl = (T)xl_d, yl = (T)yl_d, bsx = (T)bsx_d, bsy = (T)bsy_d;                     \
    const T tscale = (T)scale;                                                               \
    _Pragma("omp parallel for schedule(static)")                                            \
    for (int64_t i = 0; i < n; ++i) {                                                        \
      T bxl_ = xv[i], byl_ = yv[i];                                                          \
      T bxh_ = bxl_ + wx[i], byh_ = byl_ + wy[i];                                            \
      int kxl = imax((int)((bxl_ - xl) / bsx), 0);                                           \
      int kxh = imin((int)ceil((bxh_ - xl) / bsx) + 1, nbx);                                 \
      int kyl = imax((int)((byl_ - yl) / bsy), 0);                                           \
      int kyh = imin((int)ceil((byh_ - yl) / bsy) + 1, nby);                                 \
      for (int k = kxl; k < kxh; ++k) {                                                      \
        T px = exact_rel_##SFX(bxl_, wx[i], bcx[k], bsx);                                    \
        for (int h = kyl; h < kyh; ++h) {                                                    \
          T py = exact_rel_##SFX(byl_, wy[i], bcy[h], bsy);                                  \
          T area = px * py;                                                                  \
          int64_t sv = (int64_t)(area * tscale);                                             \
          _Pragma("omp atomic")                                                              \
          buf[(int64_t)k * nby + h] += sv;                                                   \
        }                                                                                    \
      }                                                                                      \
    }                                                                                        \
  }                                                                                          \
  void dpo_electric_force_rel_##SFX(T grad, const T* fx, const T* fy, const T* xv,         \
                                    const T* yv, const T* wxc, const T* wyc, const T* offx,  \
                                    const T* offy, const T* ratio_v, int64_t n, int nbx,     \
                                    int nby, double xl_d, double yl_d, double bsx_d,         \
                                    double bsy_d, T* gxo, T* gyo) {                          \
    const T xl = (T)xl_d, yl = (T)yl_d, bsx = (T)bsx_d, bsy = (T)bsy_d;                     \
    const T inv_x = (T)(1.0 / bsx), inv_y = (T)(1.0 / bsy);                                  \
    _Pragma("omp parallel for schedule(static)")                                            \
    for (int64_t i = 0; i < n; ++i) {                                                        \
      T sx = wxc[i], sy = wyc[i];                                                            \
      T nx = xv[i] + offx[i], ny = yv[i] + offy[i];                                          \
      T ex = twosum_err_##SFX(xv[i], offx[i], nx), ey = twosum_err_##SFX(yv[i], offy[i], ny);\
      T r = ratio_v[i];                                                                      \
      int bxl = imax((int)((nx - xl) * inv_x), 0);                                           \
      int bxh = imin((int)((nx + sx - xl) * inv_x) + 1, nbx);                                \
      int byl = imax((int)((ny - yl) * inv_y), 0);                                           \
      int byh = imin((int)((ny + sy - yl) * inv_y) + 1, nby);                                \
      T dx = (nx - (xl + (T)bxl * bsx)) + ex, dy = (ny - (yl + (T)byl * bsy)) + ey;          \
      T gx = 0, gy = 0;                                                                      \
      for (int k = bxl; k < bxh; ++k) {                                                      \
        T px = tri_rel_##SFX(dx, sx, k - bxl, bsx);                                          \
        for (int h = byl; h < byh; ++h) {                                                    \
          T py = tri_rel_##SFX(dy, sy, h - byl, bsy);                                        \
          T area = px * py;                                                                  \
          int64_t idx = (int64_t)k * nby + h;                                                \
          gx += area * fx[idx];                                                              \
          gy += area * fy[idx];                                                              \
        }                                                                                    \
      }                                                                                      \
      gx *= r;                                                                               \
      gy *= r;                                                                               \
      gxo[i] = gx * grad;                                                                    \
      gyo[i] = gy * grad;                                                                    \
    }                                                                                        \
  }

DPO_DEFINE(float, f32)
DPO_DEFINE(double, f64)
DPO_DEFINE_REL(float, f32)
DPO_DEFINE_REL(double, f64)
