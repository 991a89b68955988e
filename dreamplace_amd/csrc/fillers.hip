// Filler relabelling support: the tile keys that order the fillers for the
// tiled density kernels, and the permutation of per-node vectors' filler
// entries.  Fillers are interchangeable (one size, no pins), so keeping them
// in tile order changes no result; these passes only have to be cheap, since
// a placement loop re-sorts every few tens of iterations.
#include "dp_common.h"

#pragma clang fp contract(off)

namespace dp {

constexpr int kKeyTile = 64;          // bins per tile side, as the tiled density kernels
constexpr int kKeyMaxTiles = 4096;    // their tile limit; keys fit int16
constexpr int kPermuteMaxVectors = 8;

__global__ void k_clear_ints(int* p, int n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0;
}

// key[j] = (bx / 64) * nty + by / 64 for filler j (node f0 + j), with
// bx = clamp(floor(((x + ox) - xl) * inv_x), 0, nbx - 1) evaluated in T in
// that order (the stretched origin the density kernels bin), by likewise;
// count[t] += 1 for every filler of tile t, through an LDS histogram.
template <typename T>
__global__ void __launch_bounds__(256) k_filler_keys(const T* __restrict__ x,
                                                     const T* __restrict__ y, int64_t nf, T ox,
                                                     T oy, T xl, T yl, T inv_x, T inv_y, int nbx,
                                                     int nby, int nty, int ntiles, int16_t* keys,
                                                     int* count) {
  __shared__ int hist[kKeyMaxTiles];
  if (count)
    for (int t = threadIdx.x; t < ntiles; t += blockDim.x) hist[t] = 0;
  __syncthreads();
  const T mx = (T)(nbx - 1), my = (T)(nby - 1);
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < nf;
       j += (int64_t)gridDim.x * blockDim.x) {
    T bx = floor(((x[j] + ox) - xl) * inv_x);
    T by = floor(((y[j] + oy) - yl) * inv_y);
    bx = bx > (T)0 ? bx : (T)0;  // a NaN position takes tile 0
    bx = bx < mx ? bx : mx;
    by = by > (T)0 ? by : (T)0;
    by = by < my ? by : my;
    const int k = ((int)bx / kKeyTile) * nty + (int)by / kKeyTile;
    keys[j] = (int16_t)k;
    if (count) atomicAdd(&hist[k], 1);
  }
  if (!count) return;  // uniform: the whole grid returns
  __syncthreads();
  for (int t = threadIdx.x; t < ntiles; t += blockDim.x)
    if (hist[t]) atomicAdd(&count[t], hist[t]);
}

struct PermuteVectors {
  void* v[kPermuteMaxVectors];
  int count;
};

// tmp[k][h][j] = v_k[h * n + f0 + perm[j]] (both halves h of each [2n] vector).
template <typename T>
__global__ void k_permute_gather(PermuteVectors pv, int64_t n, int64_t f0, int64_t nf,
                                 const int64_t* __restrict__ perm, T* tmp) {
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < nf;
       j += (int64_t)gridDim.x * blockDim.x) {
    int64_t q = perm[j];
    q = q < 0 ? 0 : (q >= nf ? nf - 1 : q);  // in bounds whatever the caller passed
    const int64_t src = f0 + q;
    for (int k = 0; k < pv.count; ++k) {
      const T* v = (const T*)pv.v[k];
      tmp[(2 * k) * nf + j] = v[src];
      tmp[(2 * k + 1) * nf + j] = v[n + src];
    }
  }
}

// v_k[h * n + f0 + j] = tmp[k][h][j]
template <typename T>
__global__ void k_permute_store(PermuteVectors pv, int64_t n, int64_t f0, int64_t nf,
                                const T* __restrict__ tmp) {
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < nf;
       j += (int64_t)gridDim.x * blockDim.x) {
    for (int k = 0; k < pv.count; ++k) {
      T* v = (T*)pv.v[k];
      v[f0 + j] = tmp[(2 * k) * nf + j];
      v[n + f0 + j] = tmp[(2 * k + 1) * nf + j];
    }
  }
}

}  // namespace dp

using namespace dp;

extern "C" {

int dp_filler_tile_keys(int dtype, const void* pos, int64_t num_nodes, int64_t num_fillers,
                        double xl, double yl, double bin_size_x, double bin_size_y,
                        int num_bins_x, int num_bins_y, double origin_x, double origin_y,
                        int16_t* keys, int32_t* tile_count, void* stream) {
  DP_ARG(pos && (keys || num_fillers == 0), "dp_filler_tile_keys: null pointer");
  DP_ARG(num_fillers >= 0 && num_fillers <= num_nodes, "dp_filler_tile_keys: bad node counts");
  DP_ARG(num_bins_x > 0 && num_bins_y > 0 && bin_size_x > 0 && bin_size_y > 0,
         "dp_filler_tile_keys: bad bins");
  const int ntx = (int)ceil_div(num_bins_x, kKeyTile), nty = (int)ceil_div(num_bins_y, kKeyTile);
  DP_REQUIRE((int64_t)ntx * nty <= kKeyMaxTiles, DP_ERR_UNSUPPORTED,
             "dp_filler_tile_keys: %d x %d tiles exceed %d", ntx, nty, kKeyMaxTiles);
  const int ntiles = ntx * nty;
  hipStream_t s = S(stream);
  if (tile_count) {
    hipLaunchKernelGGL(k_clear_ints, dim3(grid_for(ntiles, 256)), dim3(256), 0, s,
                       (int*)tile_count, ntiles);
    DP_LAUNCHED("k_clear_ints");
  }
  if (num_fillers == 0) return DP_OK;
  const int64_t f0 = num_nodes - num_fillers;
  return DP_DISPATCH(dtype, "dp_filler_tile_keys", [&] {
    const T* p = (const T*)pos;
    // few workgroups: each flushes its whole histogram once
    const int grid = (int)(ceil_div(num_fillers, 256) < 2048 ? ceil_div(num_fillers, 256) : 2048);
    hipLaunchKernelGGL(k_filler_keys<T>, dim3(grid), dim3(256), 0, s, p + f0, p + num_nodes + f0,
                       num_fillers, (T)origin_x, (T)origin_y, (T)xl, (T)yl,
                       (T)(1.0 / bin_size_x), (T)(1.0 / bin_size_y), num_bins_x, num_bins_y, nty,
                       ntiles, keys, (int*)tile_count);
    DP_LAUNCHED("k_filler_keys");
    return DP_OK;
  });
}

size_t dp_permute_fillers_workspace_bytes(int dtype, int64_t num_fillers, int count) {
  const size_t t = dtype == DP_FLOAT64 ? 8 : 4;
  return align256(t * 2 * (size_t)(num_fillers > 0 ? num_fillers : 1) *
                  (size_t)(count > 0 ? count : 1));
}

int dp_permute_fillers(int dtype, void* const* vectors, int count, int64_t num_nodes,
                       int64_t num_fillers, const int64_t* perm, void* workspace, void* stream) {
  DP_ARG(count >= 0 && count <= kPermuteMaxVectors, "dp_permute_fillers: count %d not in [0, %d]",
         count, kPermuteMaxVectors);
  DP_ARG(num_fillers >= 0 && num_fillers <= num_nodes, "dp_permute_fillers: bad node counts");
  if (count == 0 || num_fillers == 0) return DP_OK;
  DP_ARG(vectors && perm && workspace, "dp_permute_fillers: null pointer");
  PermuteVectors pv = {};
  pv.count = count;
  for (int k = 0; k < count; ++k) {
    DP_ARG(vectors[k], "dp_permute_fillers: null vector %d", k);
    pv.v[k] = vectors[k];
  }
  hipStream_t s = S(stream);
  const int64_t f0 = num_nodes - num_fillers;
  return DP_DISPATCH(dtype, "dp_permute_fillers", [&] {
    T* tmp = (T*)workspace;
    hipLaunchKernelGGL(k_permute_gather<T>, dim3(grid_for(num_fillers, 256)), dim3(256), 0, s, pv,
                       num_nodes, f0, num_fillers, perm, tmp);
    DP_LAUNCHED("k_permute_gather");
    hipLaunchKernelGGL(k_permute_store<T>, dim3(grid_for(num_fillers, 256)), dim3(256), 0, s, pv,
                       num_nodes, f0, num_fillers, (const T*)tmp);
    DP_LAUNCHED("k_permute_store");
    return DP_OK;
  });
}

}  // extern "C"
