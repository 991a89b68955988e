// LDS-binned density scatter and field gather (DP_SCATTER_TILED).
//
// Cells are binned by the 64x64-bin tile holding their lower-left bin:
//   1. k_bin_count  each workgroup histograms a fixed chunk of cells over the
//                   tiles in LDS, then reserves one range per (workgroup,
//                   tile) with a single global atomic;
//   2. k_bin_scan   one workgroup turns tile counts into tile offsets and a
//                   list of work items (a tile's cells, at most kChunk each);
//   3. k_bin_fill   the same chunks again, in sub-batches staged in LDS: each
//                   sub-batch is counting-sorted by tile in LDS and written
//                   out as one contiguous run per tile.  Records are a
//                   structure of arrays (stretched origin, stretched size,
//                   ratio, node index);
//   4. k_tile_scatter  one workgroup per work item accumulates its cells'
//                   int64 contributions in a (64+2)^2 LDS tile (ds_add_u64),
//                   then adds the non-zero bins to the global map with one
//                   atomic each.
// The field gather (k_tile_gather) reuses the records of the same positions:
// a workgroup stages the tile's (64+2)^2 (Ex, Ey) pairs in LDS and each cell
// reads its <=3x3 bins from there.
//
// Every permutation here is a push: scattered writes cost a 32-byte sector
// on this memory system while scattered reads cost a whole line, so records
// are written (in runs) to tile order and the per-cell gradients are written
// back to node order, one 8/16-byte store per cell in the interleaved layout.
//
// Uniform fillers (all of one size, TriArgs::uni) form a class of their own:
// the count and fill passes give them their own workgroups and their own
// record run per tile (after the other cells' run), and their records hold
// only the stretched origin and the node index.  Work items are per (tile,
// class), so a scatter or gather work item of fillers loads 8 (scatter) or
// 12 (gather) bytes per FP32 record instead of 20 or 24 and takes the size
// and ratio from the uniform constants.
//
// Cells spanning more than 3x3 bins, or no bin at all, get their records at
// the tail of the record arrays (slot total-1-q, q from a global counter) and
// are done one workgroup (scatter) or one wavefront (gather) per cell.  Tile
// plus tail records never exceed the record count, so nothing can overflow.
// The int64 sum is order-free, so the map is bit-identical to the per-cell
// global-atomic scatter.
#include "dp_density.h"

#pragma clang fp contract(off)

namespace dp {

constexpr int kTile = 64;                 // bins per tile side
// The LDS window of a tile: kHalo bins on every side plus the two above and
// right that a 3x3 span from the tile's last bin reaches.  Record cells start
// inside the tile; filler-range cells may have moved since they were sorted,
// and those whose span leaves the window go to the drift list, done by
// k_drift_scatter / k_drift_gather through global memory.
constexpr int kHalo = 2;
constexpr int kTileH = kTile + 2 + 2 * kHalo;
constexpr int kMaxTiles = 4096;           // LDS histogram limit (4096^2 bins)
constexpr int kBinGroups = 512;           // workgroups of the count/fill passes
constexpr int kFillThreads = 512;         // threads of a fill workgroup
constexpr int kScanThreads = 256;         // threads of the one scan workgroup
constexpr int kChunk = 8192;              // cells per scatter/gather work item
constexpr int kGatherSplit = 2;           // the gather does each work item in this many parts
// threads of a scatter/gather workgroup: its (64+2)^2-bin LDS tile (35 KB)
// allows 4 workgroups per CU, so 8 waves each keep 8 waves per SIMD in flight
constexpr int kTileThreads = 512;

// Cells per LDS-staged fill sub-batch: 26 (FP32) or 46 (FP64) bytes of LDS
// per cell, 8 more (FP32) or 16 (FP64) with the bin-relative arithmetic's
// residuals: 52 KB (FP32; 51 with the residuals) and 46 KB (FP64; 31), plus
// 12 bytes per tile (3 KB at 256 tiles, 48 KB at the 4096-tile limit).
// Workgroups of 512 threads with this much LDS leave room on a CU beside the
// wirelength kernels, which the step runs concurrently (0.733 against 0.750
// ms per config-1 step with 1024-thread workgroups and twice the sub-batch).
template <typename T, bool kAcc>
struct FillSub {
  static constexpr int value = sizeof(T) == 4 ? (kAcc ? 1536 : 2048) : (kAcc ? 512 : 1024);
};

// Records: one array per field, `total` entries each.  ex, ey (the
// stretched origins' TwoSum residuals) exist with the bin-relative
// arithmetic only, else they are null.
template <typename T>
struct Recs {
  T *x, *y, *sx, *sy, *r, *ex, *ey;
  int* idx;
  __device__ __forceinline__ void put(int64_t j, T x_, T y_, T sx_, T sy_, T r_, int i, T ex_,
                                      T ey_) const {
    x[j] = x_;
    y[j] = y_;
    sx[j] = sx_;
    sy[j] = sy_;
    r[j] = r_;
    idx[j] = i;
    if (ex) {
      ex[j] = ex_;
      ey[j] = ey_;
    }
  }
  // a uniform filler: size and ratio are the uniform constants
  __device__ __forceinline__ void put_xy(int64_t j, T x_, T y_, int i, T ex_, T ey_) const {
    x[j] = x_;
    y[j] = y_;
    idx[j] = i;
    if (ex) {
      ex[j] = ex_;
      ey[j] = ey_;
    }
  }
};

// Which cells of the (movable, filler) index space a count/fill workgroup
// takes: class 0 (all cells, or all but uniform fillers) in groups
// [0, groups0), class 1 (uniform fillers) in the rest.
struct ClassSplit {
  int64_t n0, total, chunk0, chunk1;
  int groups0;
};

__device__ __forceinline__ void chunk_of(const ClassSplit& cs, int wg, int64_t& b, int64_t& e,
                                         int& cls) {
  if (wg < cs.groups0) {
    b = (int64_t)wg * cs.chunk0;
    e = b + cs.chunk0 < cs.n0 ? b + cs.chunk0 : cs.n0;
    cls = 0;
  } else {
    b = cs.n0 + (int64_t)(wg - cs.groups0) * cs.chunk1;
    e = b + cs.chunk1 < cs.total ? b + cs.chunk1 : cs.total;
    cls = 1;
  }
}

static ClassSplit class_split(int64_t n0, int64_t total, int groups) {
  ClassSplit cs;
  cs.n0 = n0;
  cs.total = total;
  const int64_t n1 = total - n0;
  int g0 = groups;
  if (n1 > 0) {
    g0 = n0 == 0 ? 0 : (int)((double)groups * (double)n0 / (double)total + 0.5);
    if (n0 > 0 && g0 < 1) g0 = 1;
    if (g0 > groups - 1) g0 = groups - 1;
  }
  cs.groups0 = g0;
  cs.chunk0 = g0 > 0 ? ceil_div(n0, g0) : 0;
  cs.chunk1 = groups - g0 > 0 ? ceil_div(n1, groups - g0) : 0;
  return cs;
}

struct TiledLayout {
  int ntx, nty, ntiles;
  int64_t max_work;
  size_t off_count, off_start, off_wgbase, off_work, off_nwork, off_rec[5], off_err[2], off_idx,
      off_drift, total;
};

static TiledLayout tiled_layout(int nbx, int nby, int64_t ncells, size_t tsize, bool acc) {
  TiledLayout L;
  L.ntx = (int)ceil_div(nbx, kTile);
  L.nty = (int)ceil_div(nby, kTile);
  L.ntiles = L.ntx * L.nty;
  L.max_work = ceil_div(ncells, kChunk) + 2 * (int64_t)L.ntiles + 1;
  const size_t n = (size_t)(ncells > 0 ? ncells : 1);
  size_t o = 0;
  L.off_count = o;  // cells per (class, tile)
  o += align256(sizeof(int) * 2 * (size_t)L.ntiles);
  L.off_start = o;  // first record of (class, tile)
  o += align256(sizeof(int) * 2 * (size_t)L.ntiles);
  L.off_wgbase = o;
  o += align256(sizeof(int) * (size_t)kBinGroups * L.ntiles);
  L.off_work = o;
  o += align256(sizeof(int4) * (size_t)L.max_work);
  L.off_nwork = o;
  o += 256;
  for (int f = 0; f < 5; ++f) {
    L.off_rec[f] = o;
    o += align256(tsize * n);
  }
  L.off_idx = o;
  o += align256(sizeof(int) * n);
  L.off_drift = o;  // filler-range cells outside their tile's window
  o += align256(sizeof(int) * n);
  // the bin-relative arithmetic's residuals last, so that every other offset
  // is the same with and without them
  for (int f = 0; f < 2; ++f) {
    L.off_err[f] = acc ? o : 0;
    if (acc) o += align256(tsize * n);
  }
  L.total = o;
  return L;
}

template <typename T>
static Recs<T> recs_of(char* ws, const TiledLayout& L) {
  Recs<T> r;
  r.x = (T*)(ws + L.off_rec[0]);
  r.y = (T*)(ws + L.off_rec[1]);
  r.sx = (T*)(ws + L.off_rec[2]);
  r.sy = (T*)(ws + L.off_rec[3]);
  r.r = (T*)(ws + L.off_rec[4]);
  r.ex = L.off_err[0] ? (T*)(ws + L.off_err[0]) : nullptr;
  r.ey = L.off_err[1] ? (T*)(ws + L.off_err[1]) : nullptr;
  r.idx = (int*)(ws + L.off_idx);
  return r;
}

bool tiled_supported(int nbx, int nby) {
  return (int64_t)ceil_div(nbx, kTile) * ceil_div(nby, kTile) <= kMaxTiles;
}

size_t tiled_ws_bytes(int nbx, int nby, int64_t ncells, int dtype, bool acc) {
  return tiled_layout(nbx, nby, ncells, dtype == DP_FLOAT64 ? 8 : 4, acc).total;
}

// index t of the (movable, filler) space -> node index
__device__ __forceinline__ int64_t node_of(int64_t t, int64_t n_mov, int64_t fill_begin) {
  return t < n_mov ? t : fill_begin + (t - n_mov);
}

// counters[key] += 1 for every lane with key >= 0; returns the lane's rank
// (the counter before its own increment; -1 for key < 0).  The lanes whose
// key equals the first active lane's add with one atomic from their lowest
// lane, and rank by their position among those lanes: cells in tile order
// put most of a wavefront in one tile, where per-lane atomics on one LDS
// word would serialise.  Lanes that are not active take no part; every
// active lane counts once per call.
__device__ __forceinline__ int wave_counter_inc(int* counters, int key) {
  const int k0 = __builtin_amdgcn_readfirstlane(key);
  const unsigned long long same = __ballot(key >= 0 && key == k0);
  if (k0 >= 0 && key == k0) {
    const int lane = (int)__lane_id();
    const int leader = __ffsll((long long)same) - 1;  // uniform: from the ballot
    int base = 0;
    if (lane == leader) base = atomicAdd(&counters[k0], (int)__popcll(same));
    base = __builtin_amdgcn_readlane(base, leader);
    return base + (int)__popcll(same & ((1ull << lane) - 1ull));
  }
  return key >= 0 ? atomicAdd(&counters[key], 1) : -1;
}

// Tile of a cell whose span is 1..3 bins on both axes; -1 for a tail cell.
template <typename T>
__device__ __forceinline__ int tile_key(const CellBins<T>& c, int nty) {
  const int nx = c.bxh - c.bxl, ny = c.byh - c.byl;
  if (nx <= 0 || ny <= 0 || nx > 3 || ny > 3) return -1;
  return (c.bxl / kTile) * nty + c.byl / kTile;
}

template <typename T>
__global__ void __launch_bounds__(256) k_bin_count(TriArgs<T> a, int64_t n_mov, int64_t fill_begin,
                                                   int64_t total, ClassSplit cs, int nty, int ntiles,
                                                   int* tile_count, int* wg_base, int* ntail,
                                                   Recs<T> rec) {
  extern __shared__ int hist[];  // ntiles ints of dynamic LDS
  for (int i = threadIdx.x; i < ntiles; i += blockDim.x) hist[i] = 0;
  __syncthreads();
  int64_t b, e;
  int cls;
  chunk_of(cs, blockIdx.x, b, e, cls);
  constexpr int U = 4;  // cells per thread per pass, loaded before any is used
  for (int64_t t0 = b + threadIdx.x; t0 < e; t0 += U * blockDim.x) {
    RawCell<T> raw[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int64_t t = t0 + u * blockDim.x;
      if (t < e) raw[u] = load_cell(a, node_of(t, n_mov, fill_begin));
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int64_t t = t0 + u * blockDim.x;
      if (t >= e) continue;
      const int64_t i = node_of(t, n_mov, fill_begin);
      const CellBins<T> c = bins_of(a, raw[u]);
      const int k = tile_key(c, nty);
      if (k < 0) {
        const int q = atomicAdd(ntail, 1);
        rec.put(total - 1 - q, c.x, c.y, c.sx, c.sy, c.r, (int)i, c.ex, c.ey);
      } else {
        // per-lane atomics: this pass is bound by its loads, and wave
        // aggregation (as in k_bin_fill) measured slower here
        atomicAdd(&hist[k], 1);
      }
    }
  }
  __syncthreads();
  int* my = wg_base + (size_t)blockIdx.x * ntiles;
  int* cnt = tile_count + (size_t)cls * ntiles;
  for (int k = threadIdx.x; k < ntiles; k += blockDim.x) {
    const int h = hist[k];
    my[k] = h ? atomicAdd(&cnt[k], h) : 0;
  }
}

// One workgroup of kScanThreads threads: a tile's records are its class-0 run, then
// its class-1 run; tile_start[c * ntiles + k] = first record of (class c,
// tile k) (exclusive scan over tiles of both counts); work items (tile,
// begin, end, class) of at most kChunk cells.
// frange (nullable): filler-relative index ranges per tile, [frange[k],
// frange[k+1]) for tile k, of fillers kept in tile order; each range becomes
// class-2 work items.  Ranges that are not a monotone partition of [0, nf)
// are replaced by one range of all fillers on tile 0, which the scatter and
// gather still do exactly (their out-of-window path), only slower.
__global__ void __launch_bounds__(kScanThreads) k_bin_scan(const int* tile_count, int ntiles,
                                                   int* tile_start, int4* work, int* nwork,
                                                   const int* frange, int nf) {
  __shared__ int s_cells[kScanThreads];
  __shared__ int s_items[kScanThreads];
  __shared__ int s_bad;
  if (threadIdx.x == 0) s_bad = frange && (frange[0] != 0 || frange[ntiles] != nf) ? 1 : 0;
  __syncthreads();
  if (frange)
    for (int k = threadIdx.x; k < ntiles; k += kScanThreads)
      if (frange[k] > frange[k + 1]) atomicOr(&s_bad, 1);
  __syncthreads();
  const bool bad = s_bad != 0;
  auto f_lo = [&](int k) { return !frange ? 0 : bad ? (k == 0 ? 0 : nf) : frange[k]; };
  auto f_hi = [&](int k) { return !frange ? 0 : bad ? nf : frange[k + 1]; };
  const int per = (ntiles + kScanThreads - 1) / kScanThreads;
  const int k0 = threadIdx.x * per;
  int cells = 0, items = 0;
  for (int k = k0; k < k0 + per && k < ntiles; ++k) {
    const int c0 = tile_count[k], c1 = tile_count[ntiles + k];
    const int c2 = f_hi(k) - f_lo(k);
    cells += c0 + c1;
    items += (c0 + kChunk - 1) / kChunk + (c1 + kChunk - 1) / kChunk + (c2 + kChunk - 1) / kChunk;
  }
  s_cells[threadIdx.x] = cells;
  s_items[threadIdx.x] = items;
  __syncthreads();
  // Hillis-Steele inclusive scans
  for (int off = 1; off < kScanThreads; off <<= 1) {
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
    for (int cls = 0; cls < 2; ++cls) {
      const int c = tile_count[cls * ntiles + k];
      tile_start[cls * ntiles + k] = cbase;
      for (int j = 0; j < c; j += kChunk) {
        const int end = j + kChunk < c ? j + kChunk : c;
        work[ibase++] = make_int4(k, cbase + j, cbase + end, cls);
      }
      cbase += c;
    }
    const int lo = f_lo(k), hi = f_hi(k);
    for (int j = lo; j < hi; j += kChunk)
      work[ibase++] = make_int4(k, j, j + kChunk < hi ? j + kChunk : hi, 2);
  }
  if (threadIdx.x == kScanThreads - 1) *nwork = s_items[kScanThreads - 1];
}

// Exclusive scan of in[0, n) into out[0, n) by a workgroup of kFillThreads
// (n <= kMaxTiles: E = kMaxTiles / kFillThreads consecutive entries per
// thread).  wsum holds kFillThreads / 64 ints.  Ends with a barrier.
__device__ __forceinline__ void block_exclusive_scan(const int* in, int* out, int n, int* wsum) {
  constexpr int NW = kFillThreads / 64;
  constexpr int E = kMaxTiles / kFillThreads;
  static_assert(E * kFillThreads == kMaxTiles, "scan covers E entries per thread");
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int k0 = E * threadIdx.x;
  int v[E];
  int s = 0;
#pragma unroll
  for (int j = 0; j < E; ++j) {
    v[j] = k0 + j < n ? in[k0 + j] : 0;
    s += v[j];
  }
  int incl = s;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int u = __shfl_up(incl, o, 64);
    if (lane >= o) incl += u;
  }
  if (lane == 63) wsum[wid] = incl;
  __syncthreads();
  if (wid == 0) {
    const int w = lane < NW ? wsum[lane] : 0;
    int wi = w;
#pragma unroll
    for (int o = 1; o < NW; o <<= 1) {
      const int u = __shfl_up(wi, o, 64);
      if (lane >= o) wi += u;
    }
    if (lane < NW) wsum[lane] = wi - w;
  }
  __syncthreads();
  int base = wsum[wid] + incl - s;
#pragma unroll
  for (int j = 0; j < E; ++j) {
    if (k0 + j < n) out[k0 + j] = base;
    base += v[j];
  }
  __syncthreads();
}

// kAcc: the bin-relative arithmetic (records carry the TwoSum residuals).
template <typename T, bool kAcc>
__global__ void __launch_bounds__(kFillThreads) k_bin_fill(TriArgs<T> a, int64_t n_mov,
                                                           int64_t fill_begin, int64_t total,
                                                           ClassSplit cs, int nty, int ntiles,
                                                           const int* tile_start,
                                                           const int* wg_base, Recs<T> rec) {
  constexpr int S = FillSub<T, kAcc>::value;
  constexpr int PER = S / kFillThreads;
  static_assert(PER * kFillThreads == S, "whole sub-batches per thread");
  constexpr int SE = kAcc ? S : 1;
  // 3 * ntiles ints of dynamic LDS: per tile, the next record slot of (this
  // workgroup, tile), the sub-batch's cells, and their exclusive scan
  extern __shared__ int s_tiles[];
  int* s_base = s_tiles;
  int* s_hist = s_tiles + ntiles;
  int* s_off = s_tiles + 2 * ntiles;
  __shared__ int s_wsum[kFillThreads / 64];
  __shared__ T s_x[S], s_y[S], s_sx[S], s_sy[S], s_r[S];
  __shared__ T s_ex[SE], s_ey[SE];
  __shared__ int s_idx[S];
  __shared__ unsigned short s_key[S];
  int64_t b, e;
  int cls;
  chunk_of(cs, blockIdx.x, b, e, cls);
  const int* my = wg_base + (size_t)blockIdx.x * ntiles;
  const int* start = tile_start + (size_t)cls * ntiles;
  for (int k = threadIdx.x; k < ntiles; k += kFillThreads) s_base[k] = start[k] + my[k];
  // the next sub-batch's cells are loaded while this one is sorted in LDS
  RawCell<T> nxt[PER];
#pragma unroll
  for (int m = 0; m < PER; ++m) {
    const int64_t t = b + m * kFillThreads + threadIdx.x;
    if (t < e) nxt[m] = load_cell(a, node_of(t, n_mov, fill_begin));
  }
  // the trip count depends on blockIdx only: every thread meets every barrier
  for (int64_t sb = b; sb < e; sb += S) {
    RawCell<T> cur[PER];
#pragma unroll
    for (int m = 0; m < PER; ++m) cur[m] = nxt[m];
#pragma unroll
    for (int m = 0; m < PER; ++m) {
      const int64_t t = sb + S + m * kFillThreads + threadIdx.x;
      if (t < e) nxt[m] = load_cell(a, node_of(t, n_mov, fill_begin));
    }
    for (int k = threadIdx.x; k < ntiles; k += kFillThreads) s_hist[k] = 0;
    __syncthreads();
    CellBins<T> c[PER];
    int key[PER], rank[PER], node[PER];
#pragma unroll
    for (int m = 0; m < PER; ++m) {
      const int64_t t = sb + m * kFillThreads + threadIdx.x;
      key[m] = -1;
      if (t < e) {
        c[m] = bins_of(a, cur[m]);
        node[m] = (int)node_of(t, n_mov, fill_begin);
        key[m] = tile_key(c[m], nty);
      }
      rank[m] = wave_counter_inc(s_hist, key[m]);
    }
    __syncthreads();
    block_exclusive_scan(s_hist, s_off, ntiles, s_wsum);
#pragma unroll
    for (int m = 0; m < PER; ++m) {
      if (key[m] >= 0) {
        const int p = s_off[key[m]] + rank[m];
        s_x[p] = c[m].x;
        s_y[p] = c[m].y;
        s_sx[p] = c[m].sx;
        s_sy[p] = c[m].sy;
        s_r[p] = c[m].r;
        if (kAcc) {
          s_ex[p] = c[m].ex;
          s_ey[p] = c[m].ey;
        }
        s_idx[p] = node[m];
        s_key[p] = (unsigned short)key[m];
      }
    }
    __syncthreads();
    const int nsmall = s_off[ntiles - 1] + s_hist[ntiles - 1];
    for (int p = threadIdx.x; p < nsmall; p += kFillThreads) {
      const int k = s_key[p];
      const int64_t g = (int64_t)s_base[k] + (p - s_off[k]);
      const T ex = kAcc ? s_ex[p] : T(0), ey = kAcc ? s_ey[p] : T(0);
      if (cls)
        rec.put_xy(g, s_x[p], s_y[p], s_idx[p], ex, ey);
      else
        rec.put(g, s_x[p], s_y[p], s_sx[p], s_sy[p], s_r[p], s_idx[p], ex, ey);
    }
    __syncthreads();
    for (int k = threadIdx.x; k < ntiles; k += kFillThreads) s_base[k] += s_hist[k];
    // s_hist[k] is read above and cleared at the next sub-batch by the same thread
  }
}

// Bin range of a record (same arithmetic as cell_bins).
template <typename T>
__device__ __forceinline__ void rec_bins(const TriArgs<T>& a, T x, T y, T sx, T sy, int& bxl,
                                         int& nx, int& byl, int& ny) {
  bxl = max((int)((x - a.xl) * a.inv_x), 0);
  int bxh = min((int)((x + sx - a.xl) * a.inv_x) + 1, a.nbx);
  byl = max((int)((y - a.yl) * a.inv_y), 0);
  int byh = min((int)((y + sy - a.yl) * a.inv_y) + 1, a.nby);
  nx = bxh - bxl;
  ny = byh - byl;
}

// Cells per thread per pass of the scatter / gather loops: their record
// loads are issued together before any of them is used.
constexpr int kBatch = 4;

template <typename T>
struct RecBatch {
  T x[kBatch], y[kBatch], sx[kBatch], sy[kBatch], r[kBatch], ex[kBatch], ey[kBatch];
  int idx[kBatch];
  bool ok[kBatch];
};

// The three overlaps per axis of a record cell with bins bxl + q, byl + q
// (q = 0, 1, 2): the reference arithmetic (tri) or, with kAcc, the
// bin-relative one.  The tile kernels take the arithmetic as a template
// argument (a runtime choice costs them registers and occupancy); the
// drift kernels pass a.acc.
template <bool kAcc, typename T>
__device__ __forceinline__ void rec_overlaps(const TriArgs<T>& a, T x, T y, T sx, T sy, T ex,
                                             T ey, int bxl, int byl, T px[3], T py[3]) {
  if constexpr (kAcc) {
    const T dx = rel_origin(x, ex, a.xl, bxl, a.bsx);
    const T dy = rel_origin(y, ey, a.yl, byl, a.bsy);
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      px[q] = tri_rel(dx, sx, q, a.bsx);
      py[q] = tri_rel(dy, sy, q, a.bsy);
    }
  } else {
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      px[q] = tri(x, sx, a.xl, bxl + q, a.bsx);
      py[q] = tri(y, sy, a.yl, byl + q, a.bsy);
    }
  }
}
template <typename T>
__device__ __forceinline__ void rec_overlaps_rt(const TriArgs<T>& a, T x, T y, T sx, T sy, T ex,
                                                T ey, int bxl, int byl, T px[3], T py[3]) {
  if (a.acc)
    rec_overlaps<true>(a, x, y, sx, sy, ex, ey, bxl, byl, px, py);
  else
    rec_overlaps<false>(a, x, y, sx, sy, ex, ey, bxl, byl, px, py);
}

// uniform: a work item of uniform fillers (origin and index only stored).
template <bool kAcc, typename T>
__device__ __forceinline__ void load_batch(const TriArgs<T>& a, const Recs<T>& rec, int j0,
                                           int stride, int end, bool with_idx, bool uniform,
                                           RecBatch<T>& b) {
#pragma unroll
  for (int u = 0; u < kBatch; ++u) {
    const int j = j0 + u * stride;
    b.ok[u] = j < end;
    if (b.ok[u]) {
      b.x[u] = rec.x[j];
      b.y[u] = rec.y[j];
      b.ex[u] = kAcc ? rec.ex[j] : T(0);
      b.ey[u] = kAcc ? rec.ey[j] : T(0);
      if (uniform) {
        b.sx[u] = a.uwx;
        b.sy[u] = a.uwy;
        b.r[u] = a.ur;
      } else {
        b.sx[u] = rec.sx[j];
        b.sy[u] = rec.sy[j];
        b.r[u] = rec.r[j];
      }
      if (with_idx) b.idx[u] = rec.idx[j];
    }
  }
}

// list[next slots] = v for every lane with `put`: the lanes' slots come from
// one atomic on `count` per wavefront, ranked by lane (a list many cells of a
// wavefront append to would otherwise serialise on one counter).  Lanes that
// are not active take no part.
__device__ __forceinline__ void wave_append(int* list, int* count, bool put, int v) {
  const unsigned long long m = __ballot(put);
  if (m == 0ull) return;  // the same for every lane
  const int lane = (int)__lane_id();
  const int leader = __ffsll((long long)m) - 1;
  int base = 0;
  if (lane == leader) base = atomicAdd(count, (int)__popcll(m));
  base = __builtin_amdgcn_readlane(base, leader);
  if (put) list[base + (int)__popcll(m & ((1ull << lane) - 1ull))] = v;
}

// A work item of class 2: a range [j0, end) of uniform fillers (filler-
// relative), read from the positions directly.  The stretched origin is the
// same T sum the fill pass stores in a record.
template <bool kAcc, typename T>
__device__ __forceinline__ void load_range(const TriArgs<T>& a, int j0, int stride, int end,
                                           RecBatch<T>& b) {
#pragma unroll
  for (int u = 0; u < kBatch; ++u) {
    const int j = j0 + u * stride;
    b.ok[u] = j < end;
    if (b.ok[u]) {
      const int64_t i = a.fill_begin + j;
      const T x = a.x[i], y = a.y[i];
      b.x[u] = x + a.uox;
      b.y[u] = y + a.uoy;
      b.ex[u] = kAcc ? two_sum_err(x, a.uox, b.x[u]) : T(0);
      b.ey[u] = kAcc ? two_sum_err(y, a.uoy, b.y[u]) : T(0);
      b.sx[u] = a.uwx;
      b.sy[u] = a.uwy;
      b.r[u] = a.ur;
      b.idx[u] = (int)i;
    }
  }
}

// acc += (long long)f for an FP32 contribution f = area * scale >= 0 (or
// -0), with fewer instructions than the generic conversion: |f| < 1 (a
// zero contribution) adds nothing, and the truncation is split at 2^32
// exactly: hi = floor(f / 2^32); lo = f - hi * 2^32 is integral for
// f >= 2^32 (a float that large has no fraction, and its bits below 2^32
// fit 24) and is f itself below, where the conversion truncates it.
__device__ __forceinline__ void lds_add_fixed_nonneg(unsigned long long* p, float f) {
  const float m = fabsf(f);
  if (m >= 1.0f) {
    const float hi = floorf(m * 0x1p-32f);
    const float lo = fmaf(hi, -0x1p32f, m);
    atomicAdd(p, ((unsigned long long)(unsigned)hi << 32) | (unsigned long long)(unsigned)lo);
  }
}

// A cell's contributions to the LDS tile: bins (lx + q, ly + l) for q < nx,
// l < ny, each (long long)(pxr[q] * py[l] * scale) where not 0.  kNonNeg:
// every factor is known non-negative (lds_add_fixed_nonneg, FP32 only).
template <bool kNonNeg, typename T>
__device__ __forceinline__ void add_cell_bins(unsigned long long* acc, const T pxr[3],
                                              const T py[3], int nx, int ny, int lx, int ly,
                                              T scale) {
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    if (q < nx) {
#pragma unroll
      for (int l = 0; l < 3; ++l) {
        if (l < ny) {
          T area = pxr[q] * py[l];
          unsigned long long* word = &acc[(lx + q) * kTileH + (ly + l)];
          if constexpr (kNonNeg) {
            lds_add_fixed_nonneg(word, area * scale);
          } else {
            long long v = (long long)(area * scale);
            if (v != 0) atomicAdd(word, (unsigned long long)v);
          }
        }
      }
    }
  }
}

template <bool kAcc, typename T>
__device__ __forceinline__ void tile_scatter(const TriArgs<T>& a, int nty, const int4* work,
                                             const int* nwork, const Recs<T>& rec,
                                             unsigned long long* map,
                                             unsigned long long* filler_map, int* drift,
                                             int* ndrift) {
  __shared__ unsigned long long acc[kTileH * kTileH];
  const int W = *nwork;
  for (int w = blockIdx.x; w < W; w += gridDim.x) {
    const int4 it = work[w];
    const int wx0 = (it.x / nty) * kTile - kHalo, wy0 = (it.x % nty) * kTile - kHalo;
    // uniform fillers (classes 1, 2) go to the filler map when there is one
    unsigned long long* dst = (it.w != 0 && filler_map) ? filler_map : map;
    for (int k = threadIdx.x; k < kTileH * kTileH; k += blockDim.x) acc[k] = 0ull;
    __syncthreads();
    for (int j0 = it.y + threadIdx.x; j0 < it.z; j0 += kBatch * blockDim.x) {
      RecBatch<T> b;
      if (it.w == 2)
        load_range<kAcc>(a, j0, blockDim.x, it.z, b);
      else
        load_batch<kAcc>(a, rec, j0, blockDim.x, it.z, false, it.w != 0, b);
#pragma unroll
      for (int u = 0; u < kBatch; ++u) {
        const T x = b.x[u], y = b.y[u], sx = b.sx[u], sy = b.sy[u], r = b.r[u];
        int bxl, nx, byl, ny;
        rec_bins(a, x, y, sx, sy, bxl, nx, byl, ny);
        const int lx = bxl - wx0, ly = byl - wy0;
        // a range cell whose span left its tile's window goes to the drift
        // list (record cells start inside their tile: always in the window);
        // every lane reaches the append, which takes one atomic per wavefront
        const bool drifting = b.ok[u] && it.w == 2 && nx > 0 && ny > 0 &&
                              (lx < 0 || ly < 0 || lx + nx > kTileH || ly + ny > kTileH);
        wave_append(drift, ndrift, drifting, drifting ? b.idx[u] : 0);
        if (!b.ok[u] || drifting) continue;
        T pxr[3], py[3];
        rec_overlaps<kAcc>(a, x, y, sx, sy, b.ex[u], b.ey[u], bxl, byl, pxr, py);
#pragma unroll
        for (int q = 0; q < 3; ++q) pxr[q] = pxr[q] * r;
        if constexpr (sizeof(T) == 4) {
          // an overlap can round below zero at a span's edge: such a cell's
          // wavefront takes the generic conversion (a uniform branch)
          bool neg = false;
#pragma unroll
          for (int q = 0; q < 3; ++q) neg |= (q < nx && pxr[q] < T(0)) || (q < ny && py[q] < T(0));
          if (__ballot(neg) == 0ull)
            add_cell_bins<true>(acc, pxr, py, nx, ny, lx, ly, a.scale);
          else
            add_cell_bins<false>(acc, pxr, py, nx, ny, lx, ly, a.scale);
        } else {
          add_cell_bins<false>(acc, pxr, py, nx, ny, lx, ly, a.scale);
        }
      }
    }
    __syncthreads();
    for (int k = threadIdx.x; k < kTileH * kTileH; k += blockDim.x) {
      const unsigned long long v = acc[k];
      if (v) {  // only bins of the die receive contributions
        const int gx = wx0 + k / kTileH, gy = wy0 + k % kTileH;
        atomicAdd(&dst[(int64_t)gx * a.nby + gy], v);
      }
    }
    __syncthreads();
  }
}

// FP32 asks for 8 waves per SIMD explicitly: the extra arguments of the
// filler map and the drift list would otherwise cost a wave to SGPR pressure
// (a few SGPRs spill to VGPR lanes).  FP64 keeps the compiler's choice: there
// the same request would spill VGPRs to scratch.
__global__ void __launch_bounds__(kTileThreads) __attribute__((amdgpu_waves_per_eu(8, 8)))
k_tile_scatter_f32(TriArgs<float> a, int nty, const int4* work, const int* nwork,
                   Recs<float> rec, unsigned long long* map, unsigned long long* filler_map,
                   int* drift, int* ndrift) {
  tile_scatter<false>(a, nty, work, nwork, rec, map, filler_map, drift, ndrift);
}

// The bin-relative arithmetic's residuals take registers that 8 waves per
// SIMD would spill: the compiler's choice there.
__global__ void __launch_bounds__(kTileThreads)
k_tile_scatter_f32_acc(TriArgs<float> a, int nty, const int4* work, const int* nwork,
                       Recs<float> rec, unsigned long long* map, unsigned long long* filler_map,
                       int* drift, int* ndrift) {
  tile_scatter<true>(a, nty, work, nwork, rec, map, filler_map, drift, ndrift);
}

template <bool kAcc>
__global__ void __launch_bounds__(kTileThreads)
k_tile_scatter_f64(TriArgs<double> a, int nty, const int4* work, const int* nwork,
                   Recs<double> rec, unsigned long long* map, unsigned long long* filler_map,
                   int* drift, int* ndrift) {
  tile_scatter<kAcc>(a, nty, work, nwork, rec, map, filler_map, drift, ndrift);
}

template <typename T, bool kAcc>
struct TileScatterKernel;
template <>
struct TileScatterKernel<float, false> {
  static constexpr auto fn = k_tile_scatter_f32;
};
template <>
struct TileScatterKernel<float, true> {
  static constexpr auto fn = k_tile_scatter_f32_acc;
};
template <bool kAcc>
struct TileScatterKernel<double, kAcc> {
  static constexpr auto fn = k_tile_scatter_f64<kAcc>;
};

template <typename T, bool kAcc>
__global__ void __launch_bounds__(kTileThreads) k_tile_gather(TriArgs<T> a, int nty, const int4* work,
                                                     const int* nwork, Recs<T> rec,
                                                     const T* __restrict__ fx,
                                                     const T* __restrict__ fy, const T* grad,
                                                     GradOut<T> go) {
  typedef typename Vec2<T>::type V;
  __shared__ V sf[kTileH * kTileH];
  const int W = *nwork * kGatherSplit;
  const T g = grad ? *grad : T(1);
  const T c = grad_out_coef(go);
  // round-robin over the parts: a die-contiguous assignment (parts of a tile
  // sharing one L2) measured slower, as the tiles' loads are uneven
  for (int w = blockIdx.x; w < W; w += gridDim.x) {
    // part w % kGatherSplit of work item w / kGatherSplit
    int4 it = work[w / kGatherSplit];
    const int part = w % kGatherSplit, len = it.z - it.y;
    const int lo = it.y + (int)((int64_t)len * part / kGatherSplit);
    const int hi = it.y + (int)((int64_t)len * (part + 1) / kGatherSplit);
    it.y = lo;
    it.z = hi;
    const int wx0 = (it.x / nty) * kTile - kHalo, wy0 = (it.x % nty) * kTile - kHalo;
    for (int k = threadIdx.x; k < kTileH * kTileH; k += blockDim.x) {
      const int gx = wx0 + k / kTileH, gy = wy0 + k % kTileH;
      const bool in = gx >= 0 && gy >= 0 && gx < a.nbx && gy < a.nby;
      const int64_t idx = in ? (int64_t)gx * a.nby + gy : 0;
      V v;
      v.x = in ? fx[idx] : T(0);
      v.y = in ? fy[idx] : T(0);
      sf[k] = v;
    }
    __syncthreads();
    for (int j0 = it.y + threadIdx.x; j0 < it.z; j0 += kBatch * blockDim.x) {
      RecBatch<T> b;
      if (it.w == 2)
        load_range<kAcc>(a, j0, blockDim.x, it.z, b);
      else
        load_batch<kAcc>(a, rec, j0, blockDim.x, it.z, true, it.w != 0, b);
#pragma unroll
      for (int u = 0; u < kBatch; ++u) {
        if (!b.ok[u]) continue;
        const T x = b.x[u], y = b.y[u], sx = b.sx[u], sy = b.sy[u];
        int bxl, nx, byl, ny;
        rec_bins(a, x, y, sx, sy, bxl, nx, byl, ny);
        T px[3], py[3];
        rec_overlaps<kAcc>(a, x, y, sx, sy, b.ex[u], b.ey[u], bxl, byl, px, py);
        const int lx = bxl - wx0, ly = byl - wy0;
        // a range cell outside its tile's window with a non-empty span is on
        // the drift list (k_drift_gather); an empty span gives 0 here
        if (it.w == 2 && (lx < 0 || ly < 0 || lx + nx > kTileH || ly + ny > kTileH) && nx > 0 &&
            ny > 0)
          continue;
        T gx = 0, gy = 0;
#pragma unroll
        for (int q = 0; q < 3; ++q) {
          if (q < nx) {
#pragma unroll
            for (int l = 0; l < 3; ++l) {
              if (l < ny) {
                T area = px[q] * py[l];
                const V f = sf[(lx + q) * kTileH + (ly + l)];
                gx += area * f.x;
                gy += area * f.y;
              }
            }
          }
        }
        const T r = b.r[u];
        put_cell_grad(go, c, b.idx[u], -((gx * r) * g), -((gy * r) * g));
      }
    }
    __syncthreads();
  }
}

// Tail records: one workgroup per cell, lanes over its bins.  Records with an
// empty span contribute nothing.
// The drift list: filler-range cells whose span left their tile's window,
// done through global memory with the tile kernels' per-cell arithmetic, so
// the results are the bits the window path would give.
template <typename T>
__device__ __forceinline__ void drift_scatter(const TriArgs<T>& a, const int* drift,
                                              const int* ndrift, unsigned long long* dst) {
  const int cnt = *ndrift;
  for (int q0 = blockIdx.x * blockDim.x + threadIdx.x; q0 < cnt; q0 += gridDim.x * blockDim.x) {
    const int64_t i = drift[q0];
    const T x0 = a.x[i], y0 = a.y[i];
    const T x = x0 + a.uox, y = y0 + a.uoy, sx = a.uwx, sy = a.uwy, r = a.ur;
    const T ex = a.acc ? two_sum_err(x0, a.uox, x) : T(0);
    const T ey = a.acc ? two_sum_err(y0, a.uoy, y) : T(0);
    int bxl, nx, byl, ny;
    rec_bins(a, x, y, sx, sy, bxl, nx, byl, ny);
    T pxr[3], py[3];
    rec_overlaps_rt(a, x, y, sx, sy, ex, ey, bxl, byl, pxr, py);
#pragma unroll
    for (int q = 0; q < 3; ++q) pxr[q] = pxr[q] * r;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      if (q < nx) {
#pragma unroll
        for (int l = 0; l < 3; ++l) {
          if (l < ny) {
            T area = pxr[q] * py[l];
            long long v = (long long)(area * a.scale);
            if (v != 0)
              atomicAdd(&dst[(int64_t)(bxl + q) * a.nby + (byl + l)], (unsigned long long)v);
          }
        }
      }
    }
  }
}

template <typename T>
__device__ __forceinline__ void drift_gather(const TriArgs<T>& a, const int* drift,
                                             const int* ndrift, const T* __restrict__ fx,
                                             const T* __restrict__ fy, const T* grad,
                                             const GradOut<T>& go) {
  const int cnt = *ndrift;
  const T g = grad ? *grad : T(1);
  const T c = grad_out_coef(go);
  for (int q0 = blockIdx.x * blockDim.x + threadIdx.x; q0 < cnt; q0 += gridDim.x * blockDim.x) {
    const int64_t i = drift[q0];
    const T x0 = a.x[i], y0 = a.y[i];
    const T x = x0 + a.uox, y = y0 + a.uoy, sx = a.uwx, sy = a.uwy, r = a.ur;
    const T ex = a.acc ? two_sum_err(x0, a.uox, x) : T(0);
    const T ey = a.acc ? two_sum_err(y0, a.uoy, y) : T(0);
    int bxl, nx, byl, ny;
    rec_bins(a, x, y, sx, sy, bxl, nx, byl, ny);
    T px[3], py[3];
    rec_overlaps_rt(a, x, y, sx, sy, ex, ey, bxl, byl, px, py);
    T gx = 0, gy = 0;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      if (q < nx) {
#pragma unroll
        for (int l = 0; l < 3; ++l) {
          if (l < ny) {
            T area = px[q] * py[l];
            const int64_t k = (int64_t)(bxl + q) * a.nby + (byl + l);
            gx += area * fx[k];
            gy += area * fy[k];
          }
        }
      }
    }
    put_cell_grad(go, c, i, -((gx * r) * g), -((gy * r) * g));
  }
}

template <typename T>
__device__ __forceinline__ void rec_scatter_large(const TriArgs<T>& a, const Recs<T>& rec,
                                                  int64_t total, const int* ntail,
                                                  unsigned long long* map,
                                                  unsigned long long* filler_map,
                                                  int64_t fill_begin) {
  const int cnt = *ntail;
  for (int q = blockIdx.x; q < cnt; q += gridDim.x) {
    const int64_t j = total - 1 - q;
    const T x = rec.x[j], y = rec.y[j], sx = rec.sx[j], sy = rec.sy[j], r = rec.r[j];
    int bxl, nx, byl, ny;
    rec_bins(a, x, y, sx, sy, bxl, nx, byl, ny);
    if (nx <= 0 || ny <= 0) continue;
    unsigned long long* dst = (filler_map && rec.idx[j] >= fill_begin) ? filler_map : map;
    const T dx = rel_origin(x, rec.ex ? rec.ex[j] : T(0), a.xl, bxl, a.bsx);
    const T dy = rel_origin(y, rec.ey ? rec.ey[j] : T(0), a.yl, byl, a.bsy);
    for (int t = threadIdx.x; t < nx * ny; t += blockDim.x) {
      const int k = bxl + t / ny, h = byl + t % ny;
      T pxr = overlap(a.acc, x, dx, sx, a.xl, bxl, k - bxl, a.bsx) * r;
      T area = pxr * overlap(a.acc, y, dy, sy, a.yl, byl, h - byl, a.bsy);
      long long v = (long long)(area * a.scale);
      if (v != 0) atomicAdd(&dst[(int64_t)k * a.nby + h], (unsigned long long)v);
    }
  }
}

// Tail records: one wavefront per cell, fixed-order reduction; a cell with an
// empty span gets a zero gradient.
template <typename T>
__device__ __forceinline__ void rec_gather_large(const TriArgs<T>& a, const Recs<T>& rec,
                                                 int64_t total, const int* ntail,
                                                 const T* __restrict__ fx,
                                                 const T* __restrict__ fy, const T* grad,
                                                 const GradOut<T>& go) {
  const int cnt = *ntail;
  const int lane = threadIdx.x & 63;
  const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int nwaves = (gridDim.x * blockDim.x) >> 6;
  const T g = grad ? *grad : T(1);
  const T c = grad_out_coef(go);
  for (int q = wave; q < cnt; q += nwaves) {
    const int64_t j = total - 1 - q;
    const T x = rec.x[j], y = rec.y[j], sx = rec.sx[j], sy = rec.sy[j], r = rec.r[j];
    int bxl, nx, byl, ny;
    rec_bins(a, x, y, sx, sy, bxl, nx, byl, ny);
    const bool empty = nx <= 0 || ny <= 0;
    const int nbins = empty ? 0 : nx * ny;
    const T dx = rel_origin(x, rec.ex ? rec.ex[j] : T(0), a.xl, bxl, a.bsx);
    const T dy = rel_origin(y, rec.ey ? rec.ey[j] : T(0), a.yl, byl, a.bsy);
    T gx = 0, gy = 0;
    for (int t = lane; t < nbins; t += 64) {
      const int k = bxl + t / ny, h = byl + t % ny;
      T area = overlap(a.acc, x, dx, sx, a.xl, bxl, k - bxl, a.bsx) *
               overlap(a.acc, y, dy, sy, a.yl, byl, h - byl, a.bsy);
      const int64_t idx = (int64_t)k * a.nby + h;
      gx += area * fx[idx];
      gy += area * fy[idx];
    }
    gx = wave_sum(gx);
    gy = wave_sum(gy);
    if (lane == 0) {
      if (empty)
        put_cell_grad(go, c, rec.idx[j], T(0), T(0));
      else
        put_cell_grad(go, c, rec.idx[j], -((gx * r) * g), -((gy * r) * g));
    }
  }
}

// What the tile scatter leaves, in one launch: the drift list (a thread per
// cell; empty unless filler ranges are used), then the tail records (a
// workgroup per cell).  Every contribution is an order-free integer add.
template <typename T>
__global__ void __launch_bounds__(256) k_scatter_rest(TriArgs<T> a, const int* drift,
                                                      const int* ndrift,
                                                      unsigned long long* drift_dst, Recs<T> rec,
                                                      int64_t total, const int* ntail,
                                                      unsigned long long* map,
                                                      unsigned long long* filler_map,
                                                      int64_t fill_begin) {
  drift_scatter(a, drift, ndrift, drift_dst);
  rec_scatter_large(a, rec, total, ntail, map, filler_map, fill_begin);
}

// What the tile gather leaves, in one launch: the zero gradient of the
// terminals [t0, t1), the drift list and the tail records (a wavefront per
// cell).  Each node's gradient is written by exactly one of them.
template <typename T>
__global__ void __launch_bounds__(256) k_gather_rest(TriArgs<T> a, int64_t t0, int64_t t1,
                                                     const int* drift, const int* ndrift,
                                                     Recs<T> rec, int64_t total,
                                                     const int* ntail, const T* __restrict__ fx,
                                                     const T* __restrict__ fy, const T* grad,
                                                     GradOut<T> go) {
  for (int64_t i = t0 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < t1;
       i += (int64_t)gridDim.x * blockDim.x)
    put_grad(go.out, go.n, i, T(0), T(0), go.interleaved);
  drift_gather(a, drift, ndrift, fx, fy, grad, go);
  rec_gather_large(a, rec, total, ntail, fx, fy, grad, go);
}

template <typename T>
int scatter_tiled(const TriArgs<T>& a, int64_t n_mov, int64_t fill_begin, int64_t n_fill,
                  char* tiled_ws, unsigned long long* map, unsigned long long* filler_map,
                  const int* filler_ranges, bool zero_map, hipStream_t s) {
  const int64_t total = n_mov + n_fill;
  TiledLayout L = tiled_layout(a.nbx, a.nby, total, sizeof(T), a.acc != 0);
  DP_REQUIRE(L.ntiles <= kMaxTiles, DP_ERR_UNSUPPORTED,
             "tiled scatter: %d tiles exceed the LDS histogram (%d)", L.ntiles, kMaxTiles);
  DP_REQUIRE(!filler_map || n_fill == 0 || a.uni, DP_ERR_UNSUPPORTED,
             "tiled scatter: a separate filler map needs fillers of one size");
  // filler ranges: uniform fillers whose stretched size is at most 2 bins, so
  // every span is at most 3 bins (the tile kernels' loops)
  const bool ranges = filler_ranges != nullptr && n_fill > 0;
  DP_REQUIRE(!ranges || (a.uni && (double)a.uwx * (double)a.inv_x <= 2.0 &&
                         (double)a.uwy * (double)a.inv_y <= 2.0),
             DP_ERR_UNSUPPORTED,
             "tiled scatter: filler ranges need fillers of one size, at most 2 bins stretched");
  int* tile_count = (int*)(tiled_ws + L.off_count);
  int* tile_start = (int*)(tiled_ws + L.off_start);
  int* wg_base = (int*)(tiled_ws + L.off_wgbase);
  int4* work = (int4*)(tiled_ws + L.off_work);
  int* nwork = (int*)(tiled_ws + L.off_nwork);
  int* ntail = nwork + 1;
  Recs<T> rec = recs_of<T>(tiled_ws, L);
  int* ndrift = nwork + 2;
  int* drift = (int*)(tiled_ws + L.off_drift);
  // one launch clears the maps, the per-tile counts and the work / tail /
  // drift counters
  const int64_t nb = (int64_t)a.nbx * a.nby;
  ZeroRegions z = {};
  z.p[0] = (uint32_t*)tile_count;
  z.n[0] = 2 * (int64_t)L.ntiles;
  z.p[1] = (uint32_t*)nwork;
  z.n[1] = 3;
  z.p[2] = zero_map ? (uint32_t*)map : nullptr;
  z.n[2] = 2 * nb;
  z.p[3] = (uint32_t*)filler_map;
  z.n[3] = 2 * nb;
  zero_regions(z, s);
  DP_LAUNCHED("tiled init");
  if (total > 0) {
    // uniform fillers are class 1 (their own workgroups and record runs);
    // with filler ranges they are not binned at all (class 2 work items)
    const ClassSplit cs = ranges ? class_split(n_mov, n_mov, kBinGroups)
                                 : class_split(a.uni ? n_mov : total, total, kBinGroups);
    hipLaunchKernelGGL(k_bin_count<T>, dim3(kBinGroups), dim3(256),
                       sizeof(int) * (size_t)L.ntiles, s, a, n_mov, fill_begin,
                       total, cs, L.nty, L.ntiles, tile_count, wg_base, ntail, rec);
    DP_LAUNCHED("k_bin_count");
    hipLaunchKernelGGL(k_bin_scan, dim3(1), dim3(kScanThreads), 0, s, (const int*)tile_count, L.ntiles,
                       tile_start, work, nwork, ranges ? filler_ranges : nullptr, (int)n_fill);
    DP_LAUNCHED("k_bin_scan");
    auto fill = a.acc ? k_bin_fill<T, true> : k_bin_fill<T, false>;
    DP_PROBED(DP_KERNEL_DENSITY_BIN, s,
              hipLaunchKernelGGL(fill, dim3(kBinGroups), dim3(kFillThreads),
                                 3 * sizeof(int) * (size_t)L.ntiles, s, a, n_mov,
                                 fill_begin, total, cs, L.nty, L.ntiles, (const int*)tile_start,
                                 (const int*)wg_base, rec));
    DP_LAUNCHED("k_bin_fill");
    auto scat = a.acc ? TileScatterKernel<T, true>::fn : TileScatterKernel<T, false>::fn;
    DP_PROBED(DP_KERNEL_DENSITY_SCATTER, s,
              hipLaunchKernelGGL(scat, dim3(2048), dim3(kTileThreads), 0, s,
                                 a, L.nty, (const int4*)work, (const int*)nwork, rec, map,
                                 filler_map, drift, ndrift));
    DP_LAUNCHED("k_tile_scatter");
    hipLaunchKernelGGL(k_scatter_rest<T>, dim3(1024), dim3(256), 0, s, a, (const int*)drift,
                       (const int*)ndrift, filler_map ? filler_map : map, rec, total,
                       (const int*)ntail, map, filler_map, fill_begin);
    DP_LAUNCHED("k_scatter_rest");
  }
  return DP_OK;
}

template <typename T>
int gather_tiled(const TriArgs<T>& a, int64_t n_mov, int64_t n_fill, const T* fx, const T* fy,
                 const T* grad, char* tiled_ws, const GradOut<T>& go, hipStream_t s) {
  const int64_t num_nodes = go.n;
  const int64_t total = n_mov + n_fill;
  TiledLayout L = tiled_layout(a.nbx, a.nby, total, sizeof(T), a.acc != 0);
  DP_REQUIRE(L.ntiles <= kMaxTiles, DP_ERR_UNSUPPORTED,
             "tiled gather: %d tiles exceed the LDS histogram (%d)", L.ntiles, kMaxTiles);
  const int4* work = (const int4*)(tiled_ws + L.off_work);
  const int* nwork = (const int*)(tiled_ws + L.off_nwork);
  const int* ntail = nwork + 1;
  Recs<T> rec = recs_of<T>(tiled_ws, L);
  // terminals [n_mov, num_nodes - n_fill) get 0; every other node has a record
  const int64_t t0 = n_mov, t1 = num_nodes - n_fill;
  if (total > 0) {
    auto gath = a.acc ? k_tile_gather<T, true> : k_tile_gather<T, false>;
    DP_PROBED(DP_KERNEL_ELECTRIC_FORCE, s,
              hipLaunchKernelGGL(gath, dim3(2048), dim3(kTileThreads), 0, s, a, L.nty, work,
                                 nwork, rec, fx, fy, grad, go));
    DP_LAUNCHED("k_tile_gather");
  }
  // the terminals, the fillers taken by range whose span left their tile's
  // window (the list is empty unless the last density map used filler
  // ranges) and the tail records
  hipLaunchKernelGGL(k_gather_rest<T>, dim3(256), dim3(256), 0, s, a, t0, t1,
                     (const int*)(tiled_ws + L.off_drift), nwork + 2, rec, total, ntail, fx, fy,
                     grad, go);
  DP_LAUNCHED("k_gather_rest");
  return DP_OK;
}

#define DP_TILED_INST(T)                                                                     \
  template int scatter_tiled<T>(const TriArgs<T>&, int64_t, int64_t, int64_t, char*,        \
                                unsigned long long*, unsigned long long*, const int*, bool,  \
                                hipStream_t);                                                \
  template int gather_tiled<T>(const TriArgs<T>&, int64_t, int64_t, const T*, const T*,    \
                               const T*, char*, const GradOut<T>&, hipStream_t);
DP_TILED_INST(float)
DP_TILED_INST(double)

}  // namespace dp
