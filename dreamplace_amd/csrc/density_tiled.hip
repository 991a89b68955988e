// LDS-binned density scatter and field gather (DP_SCATTER_TILED).
//
// Cells are binned by the 32x32-bin tile holding their lower-left bin:
//   1. k_bin_count  each workgroup histograms a fixed chunk of cells over the
//                   tiles in LDS, then reserves one range per (workgroup,
//                   tile) with a single global atomic;
//   2. k_bin_scan   one workgroup turns tile counts into tile offsets and a
//                   list of work items (a tile's cells, at most kChunk each);
//   3. k_bin_fill   the same chunks again: each cell gets its slot from an
//                   LDS cursor and writes a packed record (origin, stretched
//                   size, ratio, node index) in tile order;
//   4. k_tile_scatter  one workgroup per work item accumulates its cells'
//                   int64 contributions in a (32+2)^2 LDS tile (ds_add_u64),
//                   then adds the non-zero bins to the global map with one
//                   atomic each.
// The field gather (k_tile_gather) reuses the records of the same positions:
// a workgroup stages the tile's (32+2)^2 field values in LDS and each cell
// reads its <=3x3 bins from there.
// Cells spanning more than 3x3 bins get their records at the tail of the
// record array (slot total-1-q, q from a global counter) and are done one
// workgroup (scatter) or one wavefront (gather) per cell.  Small plus large
// cells never exceed the record count, so nothing can overflow.
// The int64 sum is order-free, so the map is bit-identical to the
// per-cell-atomic scatter.
#include "dp_density.h"

#pragma clang fp contract(off)

namespace dp {

constexpr int kTile = 32;                 // bins per tile side
constexpr int kTileH = kTile + 2;         // with the two-bin halo of a 3x3 span
constexpr int kMaxTiles = 8192;           // LDS histogram limit (32 KB)
constexpr int kBinGroups = 512;           // workgroups of the count/fill passes
constexpr int kChunk = 2048;              // cells per scatter/gather work item

template <typename T>
struct alignas(sizeof(T) == 4 ? 8 : 16) CellRec {
  T x, y, sx, sy, r;
  int idx;
};

struct TiledLayout {
  int ntx, nty, ntiles;
  int64_t max_work;
  size_t off_count, off_start, off_wgbase, off_work, off_nwork, off_rec, total;
};

static TiledLayout tiled_layout(int nbx, int nby, int64_t ncells, size_t rec_bytes) {
  TiledLayout L;
  L.ntx = (int)ceil_div(nbx, kTile);
  L.nty = (int)ceil_div(nby, kTile);
  L.ntiles = L.ntx * L.nty;
  L.max_work = ceil_div(ncells, kChunk) + L.ntiles + 1;
  size_t o = 0;
  L.off_count = o;
  o += align256(sizeof(int) * (size_t)L.ntiles);
  L.off_start = o;
  o += align256(sizeof(int) * (size_t)(L.ntiles + 1));
  L.off_wgbase = o;
  o += align256(sizeof(int) * (size_t)kBinGroups * L.ntiles);
  L.off_work = o;
  o += align256(sizeof(int4) * (size_t)L.max_work);
  L.off_nwork = o;
  o += 256;
  L.off_rec = o;
  o += align256(rec_bytes * (size_t)(ncells > 0 ? ncells : 1));
  L.total = o;
  return L;
}

bool tiled_supported(int nbx, int nby) {
  return (int64_t)ceil_div(nbx, kTile) * ceil_div(nby, kTile) <= kMaxTiles;
}

size_t tiled_ws_bytes(int nbx, int nby, int64_t ncells, int dtype) {
  size_t rb = dtype == DP_FLOAT64 ? sizeof(CellRec<double>) : sizeof(CellRec<float>);
  return tiled_layout(nbx, nby, ncells, rb).total;
}

// index t of the (movable, filler) space -> node index
__device__ __forceinline__ int64_t node_of(int64_t t, int64_t n_mov, int64_t fill_begin) {
  return t < n_mov ? t : fill_begin + (t - n_mov);
}

template <typename T>
__device__ __forceinline__ CellRec<T> make_rec(const CellBins<T>& c, int64_t i) {
  CellRec<T> r;
  r.x = c.x;
  r.y = c.y;
  r.sx = c.sx;
  r.sy = c.sy;
  r.r = c.r;
  r.idx = (int)i;
  return r;
}

template <typename T>
__global__ void __launch_bounds__(256) k_bin_count(TriArgs<T> a, int64_t n_mov, int64_t fill_begin,
                                                   int64_t total, int64_t chunk, int nty, int ntiles,
                                                   int* tile_count, int* wg_base, int* nlarge,
                                                   CellRec<T>* rec) {
  extern __shared__ int hist[];
  for (int i = threadIdx.x; i < ntiles; i += blockDim.x) hist[i] = 0;
  __syncthreads();
  const int64_t b = (int64_t)blockIdx.x * chunk;
  const int64_t e = b + chunk < total ? b + chunk : total;
  for (int64_t t = b + threadIdx.x; t < e; t += blockDim.x) {
    const int64_t i = node_of(t, n_mov, fill_begin);
    CellBins<T> c = cell_bins(a, i);
    const int nx = c.bxh - c.bxl, ny = c.byh - c.byl;
    if (nx <= 0 || ny <= 0) continue;
    if (nx > 3 || ny > 3) {
      const int q = atomicAdd(nlarge, 1);
      rec[total - 1 - q] = make_rec(c, i);
      continue;
    }
    atomicAdd(&hist[(c.bxl / kTile) * nty + c.byl / kTile], 1);
  }
  __syncthreads();
  int* my = wg_base + (size_t)blockIdx.x * ntiles;
  for (int k = threadIdx.x; k < ntiles; k += blockDim.x) {
    const int h = hist[k];
    my[k] = h ? atomicAdd(&tile_count[k], h) : 0;
  }
}

// One workgroup of 1024 threads: tile_start = exclusive scan of tile_count;
// work items (tile, begin, end) of at most kChunk cells per tile.
__global__ void __launch_bounds__(1024) k_bin_scan(const int* tile_count, int ntiles,
                                                   int* tile_start, int4* work, int* nwork) {
  __shared__ int s_cells[1024];
  __shared__ int s_items[1024];
  const int per = (ntiles + 1023) / 1024;
  const int k0 = threadIdx.x * per;
  int cells = 0, items = 0;
  for (int k = k0; k < k0 + per && k < ntiles; ++k) {
    const int c = tile_count[k];
    cells += c;
    items += (c + kChunk - 1) / kChunk;
  }
  s_cells[threadIdx.x] = cells;
  s_items[threadIdx.x] = items;
  __syncthreads();
  // Hillis-Steele inclusive scans
  for (int off = 1; off < 1024; off <<= 1) {
    int vc = threadIdx.x >= off ? s_cells[threadIdx.x - off] : 0;
    int vi = threadIdx.x >= off ? s_items[threadIdx.x - off] : 0;
    __syncthreads();
    s_cells[threadIdx.x] += vc;
    s_items[threadIdx.x] += vi;
    __syncthreads();
  }
  int cbase = s_cells[threadIdx.x] - cells;
  int ibase = s_items[threadIdx.x] - items;
  for (int k = k0; k < k0 + per && k < ntiles; ++k) {
    const int c = tile_count[k];
    tile_start[k] = cbase;
    for (int j = 0; j < c; j += kChunk) {
      const int end = j + kChunk < c ? j + kChunk : c;
      work[ibase++] = make_int4(k, cbase + j, cbase + end, 0);
    }
    cbase += c;
  }
  if (threadIdx.x == 1023) {
    tile_start[ntiles] = s_cells[1023];
    *nwork = s_items[1023];
  }
}

template <typename T>
__global__ void __launch_bounds__(256) k_bin_fill(TriArgs<T> a, int64_t n_mov, int64_t fill_begin,
                                                  int64_t total, int64_t chunk, int nty, int ntiles,
                                                  const int* tile_start, const int* wg_base,
                                                  CellRec<T>* rec) {
  extern __shared__ int cursor[];
  for (int i = threadIdx.x; i < ntiles; i += blockDim.x) cursor[i] = 0;
  __syncthreads();
  const int* my = wg_base + (size_t)blockIdx.x * ntiles;
  const int64_t b = (int64_t)blockIdx.x * chunk;
  const int64_t e = b + chunk < total ? b + chunk : total;
  for (int64_t t = b + threadIdx.x; t < e; t += blockDim.x) {
    const int64_t i = node_of(t, n_mov, fill_begin);
    CellBins<T> c = cell_bins(a, i);
    const int nx = c.bxh - c.bxl, ny = c.byh - c.byl;
    if (nx <= 0 || ny <= 0 || nx > 3 || ny > 3) continue;
    const int k = (c.bxl / kTile) * nty + c.byl / kTile;
    const int r = atomicAdd(&cursor[k], 1);
    rec[tile_start[k] + my[k] + r] = make_rec(c, i);
  }
}

// Bin range of a record (same arithmetic as cell_bins).
template <typename T>
__device__ __forceinline__ void rec_bins(const TriArgs<T>& a, const CellRec<T>& c, int& bxl,
                                         int& nx, int& byl, int& ny) {
  bxl = max((int)((c.x - a.xl) * a.inv_x), 0);
  int bxh = min((int)((c.x + c.sx - a.xl) * a.inv_x) + 1, a.nbx);
  byl = max((int)((c.y - a.yl) * a.inv_y), 0);
  int byh = min((int)((c.y + c.sy - a.yl) * a.inv_y) + 1, a.nby);
  nx = bxh - bxl;
  ny = byh - byl;
}

template <typename T>
__global__ void __launch_bounds__(256) k_tile_scatter(TriArgs<T> a, int nty, const int4* work,
                                                      const int* nwork, const CellRec<T>* rec,
                                                      unsigned long long* map) {
  __shared__ unsigned long long acc[kTileH * kTileH];
  const int W = *nwork;
  for (int w = blockIdx.x; w < W; w += gridDim.x) {
    const int4 it = work[w];
    const int tx0 = (it.x / nty) * kTile, ty0 = (it.x % nty) * kTile;
    for (int k = threadIdx.x; k < kTileH * kTileH; k += blockDim.x) acc[k] = 0ull;
    __syncthreads();
    for (int j = it.y + threadIdx.x; j < it.z; j += blockDim.x) {
      const CellRec<T> c = rec[j];
      int bxl, nx, byl, ny;
      rec_bins(a, c, bxl, nx, byl, ny);
      T pxr[3], py[3];
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        pxr[q] = tri(c.x, c.sx, a.xl, bxl + q, a.bsx) * c.r;
        py[q] = tri(c.y, c.sy, a.yl, byl + q, a.bsy);
      }
      const int lx = bxl - tx0, ly = byl - ty0;
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        if (q < nx) {
#pragma unroll
          for (int l = 0; l < 3; ++l) {
            if (l < ny) {
              T area = pxr[q] * py[l];
              long long v = (long long)(area * a.scale);
              if (v != 0) atomicAdd(&acc[(lx + q) * kTileH + (ly + l)], (unsigned long long)v);
            }
          }
        }
      }
    }
    __syncthreads();
    for (int k = threadIdx.x; k < kTileH * kTileH; k += blockDim.x) {
      const unsigned long long v = acc[k];
      if (v) {
        const int gx = tx0 + k / kTileH, gy = ty0 + k % kTileH;
        atomicAdd(&map[(int64_t)gx * a.nby + gy], v);
      }
    }
    __syncthreads();
  }
}

template <typename T>
__global__ void __launch_bounds__(256) k_tile_gather(TriArgs<T> a, int nty, const int4* work,
                                                     const int* nwork, const CellRec<T>* rec,
                                                     const T* __restrict__ fx,
                                                     const T* __restrict__ fy, const T* grad,
                                                     int64_t num_nodes, T* out) {
  __shared__ T sfx[kTileH * kTileH];
  __shared__ T sfy[kTileH * kTileH];
  const int W = *nwork;
  const T g = grad ? *grad : T(1);
  for (int w = blockIdx.x; w < W; w += gridDim.x) {
    const int4 it = work[w];
    const int tx0 = (it.x / nty) * kTile, ty0 = (it.x % nty) * kTile;
    for (int k = threadIdx.x; k < kTileH * kTileH; k += blockDim.x) {
      const int gx = tx0 + k / kTileH, gy = ty0 + k % kTileH;
      const bool in = gx < a.nbx && gy < a.nby;
      const int64_t idx = (int64_t)gx * a.nby + gy;
      sfx[k] = in ? fx[idx] : T(0);
      sfy[k] = in ? fy[idx] : T(0);
    }
    __syncthreads();
    for (int j = it.y + threadIdx.x; j < it.z; j += blockDim.x) {
      const CellRec<T> c = rec[j];
      int bxl, nx, byl, ny;
      rec_bins(a, c, bxl, nx, byl, ny);
      T px[3], py[3];
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        px[q] = tri(c.x, c.sx, a.xl, bxl + q, a.bsx);
        py[q] = tri(c.y, c.sy, a.yl, byl + q, a.bsy);
      }
      const int lx = bxl - tx0, ly = byl - ty0;
      T gx = 0, gy = 0;
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        if (q < nx) {
#pragma unroll
          for (int l = 0; l < 3; ++l) {
            if (l < ny) {
              T area = px[q] * py[l];
              const int k = (lx + q) * kTileH + (ly + l);
              gx += area * sfx[k];
              gy += area * sfy[k];
            }
          }
        }
      }
      out[c.idx] = -((gx * c.r) * g);
      out[num_nodes + c.idx] = -((gy * c.r) * g);
    }
    __syncthreads();
  }
}

// Large cells (tail records): one workgroup per cell, lanes over its bins.
template <typename T>
__global__ void __launch_bounds__(256) k_rec_scatter_large(TriArgs<T> a, const CellRec<T>* rec,
                                                           int64_t total, const int* nlarge,
                                                           unsigned long long* map) {
  const int cnt = *nlarge;
  for (int q = blockIdx.x; q < cnt; q += gridDim.x) {
    const CellRec<T> c = rec[total - 1 - q];
    int bxl, nx, byl, ny;
    rec_bins(a, c, bxl, nx, byl, ny);
    for (int t = threadIdx.x; t < nx * ny; t += blockDim.x) {
      const int k = bxl + t / ny, h = byl + t % ny;
      T pxr = tri(c.x, c.sx, a.xl, k, a.bsx) * c.r;
      T area = pxr * tri(c.y, c.sy, a.yl, h, a.bsy);
      long long v = (long long)(area * a.scale);
      if (v != 0) atomicAdd(&map[(int64_t)k * a.nby + h], (unsigned long long)v);
    }
  }
}

// Large cells (tail records): one wavefront per cell, fixed-order reduction.
template <typename T>
__global__ void __launch_bounds__(256) k_rec_gather_large(TriArgs<T> a, const CellRec<T>* rec,
                                                          int64_t total, const int* nlarge,
                                                          const T* __restrict__ fx,
                                                          const T* __restrict__ fy, const T* grad,
                                                          int64_t num_nodes, T* out) {
  const int cnt = *nlarge;
  const int lane = threadIdx.x & 63;
  const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int nwaves = (gridDim.x * blockDim.x) >> 6;
  const T g = grad ? *grad : T(1);
  for (int q = wave; q < cnt; q += nwaves) {
    const CellRec<T> c = rec[total - 1 - q];
    int bxl, nx, byl, ny;
    rec_bins(a, c, bxl, nx, byl, ny);
    T gx = 0, gy = 0;
    for (int t = lane; t < nx * ny; t += 64) {
      const int k = bxl + t / ny, h = byl + t % ny;
      T area = tri(c.x, c.sx, a.xl, k, a.bsx) * tri(c.y, c.sy, a.yl, h, a.bsy);
      const int64_t idx = (int64_t)k * a.nby + h;
      gx += area * fx[idx];
      gy += area * fy[idx];
    }
    gx = wave_sum(gx);
    gy = wave_sum(gy);
    if (lane == 0) {
      out[c.idx] = -((gx * c.r) * g);
      out[num_nodes + c.idx] = -((gy * c.r) * g);
    }
  }
}

// Zero the gradient of the nodes no record covers: terminals, and movable or
// filler cells entirely outside the bin grid (span <= 0).  Cells with a
// record or in the large queue are written by the gather kernels.
template <typename T>
__global__ void k_zero_uncovered(TriArgs<T> a, int64_t num_nodes, int64_t n_mov, int64_t n_fill,
                                 T* out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= num_nodes) return;
  bool zero = i >= n_mov && i < num_nodes - n_fill;
  if (!zero) {
    CellBins<T> c = cell_bins(a, i);
    zero = (c.bxh - c.bxl) <= 0 || (c.byh - c.byl) <= 0;
  }
  if (zero) {
    out[i] = 0;
    out[num_nodes + i] = 0;
  }
}

template <typename T>
int scatter_tiled(const TriArgs<T>& a, int64_t n_mov, int64_t fill_begin, int64_t n_fill,
                  char* tiled_ws, unsigned long long* map, hipStream_t s) {
  const int64_t total = n_mov + n_fill;
  TiledLayout L = tiled_layout(a.nbx, a.nby, total, sizeof(CellRec<T>));
  DP_REQUIRE(L.ntiles <= kMaxTiles, DP_ERR_UNSUPPORTED,
             "tiled scatter: %d tiles exceed the LDS histogram (%d)", L.ntiles, kMaxTiles);
  int* tile_count = (int*)(tiled_ws + L.off_count);
  int* tile_start = (int*)(tiled_ws + L.off_start);
  int* wg_base = (int*)(tiled_ws + L.off_wgbase);
  int4* work = (int4*)(tiled_ws + L.off_work);
  int* nwork = (int*)(tiled_ws + L.off_nwork);
  int* nlarge = nwork + 1;
  CellRec<T>* rec = (CellRec<T>*)(tiled_ws + L.off_rec);
  hipLaunchKernelGGL(k_fill<int>, dim3(grid_for(L.ntiles, 256)), dim3(256), 0, s, tile_count,
                     (int64_t)L.ntiles, 0);
  hipLaunchKernelGGL(k_fill<int>, dim3(1), dim3(64), 0, s, nwork, (int64_t)2, 0);
  DP_LAUNCHED("tiled init");
  if (total > 0) {
    const int64_t chunk = ceil_div(total, kBinGroups);
    const size_t lds = sizeof(int) * (size_t)L.ntiles;
    hipLaunchKernelGGL(k_bin_count<T>, dim3(kBinGroups), dim3(256), lds, s, a, n_mov, fill_begin,
                       total, chunk, L.nty, L.ntiles, tile_count, wg_base, nlarge, rec);
    DP_LAUNCHED("k_bin_count");
    hipLaunchKernelGGL(k_bin_scan, dim3(1), dim3(1024), 0, s, (const int*)tile_count, L.ntiles,
                       tile_start, work, nwork);
    DP_LAUNCHED("k_bin_scan");
    hipLaunchKernelGGL(k_bin_fill<T>, dim3(kBinGroups), dim3(256), lds, s, a, n_mov, fill_begin,
                       total, chunk, L.nty, L.ntiles, (const int*)tile_start,
                       (const int*)wg_base, rec);
    DP_LAUNCHED("k_bin_fill");
    DP_PROBED(DP_KERNEL_DENSITY_SCATTER, s,
              hipLaunchKernelGGL(k_tile_scatter<T>, dim3(2048), dim3(256), 0, s, a, L.nty,
                                 (const int4*)work, (const int*)nwork,
                                 (const CellRec<T>*)rec, map));
    DP_LAUNCHED("k_tile_scatter");
    hipLaunchKernelGGL(k_rec_scatter_large<T>, dim3(1024), dim3(256), 0, s, a,
                       (const CellRec<T>*)rec, total, (const int*)nlarge, map);
    DP_LAUNCHED("k_rec_scatter_large");
  }
  return DP_OK;
}

template <typename T>
int gather_tiled(const TriArgs<T>& a, int64_t num_nodes, int64_t n_mov, int64_t n_fill,
                 const T* fx, const T* fy, const T* grad, char* tiled_ws, T* out, hipStream_t s) {
  const int64_t total = n_mov + n_fill;
  TiledLayout L = tiled_layout(a.nbx, a.nby, total, sizeof(CellRec<T>));
  DP_REQUIRE(L.ntiles <= kMaxTiles, DP_ERR_UNSUPPORTED,
             "tiled gather: %d tiles exceed the LDS histogram (%d)", L.ntiles, kMaxTiles);
  const int4* work = (const int4*)(tiled_ws + L.off_work);
  const int* nwork = (const int*)(tiled_ws + L.off_nwork);
  const int* nlarge = nwork + 1;
  const CellRec<T>* rec = (const CellRec<T>*)(tiled_ws + L.off_rec);
  hipLaunchKernelGGL(k_zero_uncovered<T>, dim3((unsigned)ceil_div(num_nodes, 256)), dim3(256), 0,
                     s, a, num_nodes, n_mov, n_fill, out);
  DP_LAUNCHED("k_zero_uncovered");
  if (total > 0) {
    DP_PROBED(DP_KERNEL_ELECTRIC_FORCE, s,
              hipLaunchKernelGGL(k_tile_gather<T>, dim3(2048), dim3(256), 0, s, a, L.nty, work,
                                 nwork, rec, fx, fy, grad, num_nodes, out));
    DP_LAUNCHED("k_tile_gather");
    hipLaunchKernelGGL(k_rec_gather_large<T>, dim3(256), dim3(256), 0, s, a, rec, total, nlarge,
                       fx, fy, grad, num_nodes, out);
    DP_LAUNCHED("k_rec_gather_large");
  }
  return DP_OK;
}

#define DP_TILED_INST(T)                                                                     \
  template int scatter_tiled<T>(const TriArgs<T>&, int64_t, int64_t, int64_t, char*,        \
                                unsigned long long*, hipStream_t);                           \
  template int gather_tiled<T>(const TriArgs<T>&, int64_t, int64_t, int64_t, const T*,      \
                               const T*, const T*, char*, T*, hipStream_t);
DP_TILED_INST(float)
DP_TILED_INST(double)

}  // namespace dp
