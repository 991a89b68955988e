// Dense spectral transforms on the CDNA4 matrix cores, and the Poisson solve.
//
// Every 2-D transform of dreamplace/ops/dct/dct2_fft2.py is out = L * (x * R)
// with constant cosine / sine matrices (weights folded in):
//   DCT2       L = (2/M) C_M            R = (2/N) C_N^T
//   IDCT2      L = (c C_M)^T            R = c C_N
//   IDXST_IDCT L = 2 S_M^T              R = c C_N
//   IDCT_IDXST L = (c C_M)^T            R = 2 S_N
// with C_K[u,i] = cos(pi u (2i+1)/(2K)), S_K[u,i] = sin(...), c_0 = 1, c_k = 2.
// The GEMMs run on MFMA in three arithmetic forms:
//   FP64   v_mfma_f64_16x16x4_f64
//   FP32   v_mfma_f32_32x32x2_f32
//   BF16X3 v_mfma_f32_32x32x16_bf16 on a hi/lo bf16 split of both operands,
//          hi*hi + hi*lo + lo*hi accumulated in FP32.
// Grids whose sides are not multiples of 64 use a plain LDS-tiled FMA GEMM.
#include <math.h>

#include "dp_common.h"

namespace dp {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

enum MatId { L_DCT = 0, L_IDCT = 1, L_IDXST = 2, R_DCT = 3, R_IDCT = 4, R_IDXST = 5 };

static inline size_t elem_bytes(int precision) {
  return precision == DP_SPECTRAL_FP64 ? sizeof(double) : sizeof(float);
}

// Plan layout: the three M x M matrices, then the three N x N matrices.
static inline size_t plan_offset(int M, int N, int id, int precision) {
  size_t mm = (size_t)M * M, nn = (size_t)N * N;
  size_t off = id < 3 ? id * mm : 3 * mm + (id - 3) * nn;
  return off * elem_bytes(precision);
}

template <typename T>
__global__ void k_plan_fill(T* plan, int M, int N) {
  const size_t mm = (size_t)M * M, nn = (size_t)N * N;
  const size_t total = 3 * mm + 3 * nn;
  for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < total;
       t += (size_t)gridDim.x * blockDim.x) {
    int id, K;
    size_t e;
    if (t < 3 * mm) {
      id = (int)(t / mm);
      e = t % mm;
      K = M;
    } else {
      id = 3 + (int)((t - 3 * mm) / nn);
      e = (t - 3 * mm) % nn;
      K = N;
    }
    int r = (int)(e / K), c = (int)(e % K);
    // angle(u, i) = pi*u*(2i+1)/(2K), reduced mod 4K exactly in integers
    auto ang = [&](int u, int i) {
      long long q = ((long long)u * (2 * i + 1)) % (4LL * K);
      return M_PI * (double)q / (2.0 * K);
    };
    double v;
    switch (id) {
      case L_DCT: v = (2.0 / K) * cos(ang(r, c)); break;             // [u][i]
      case L_IDCT: v = (c == 0 ? 1.0 : 2.0) * cos(ang(c, r)); break;  // [m][u]
      case L_IDXST: v = 2.0 * sin(ang(c, r)); break;                  // [m][u]
      case R_DCT: v = (2.0 / K) * cos(ang(c, r)); break;              // [j][v]
      case R_IDCT: v = (r == 0 ? 1.0 : 2.0) * cos(ang(r, c)); break;  // [v][n]
      default: v = 2.0 * sin(ang(r, c)); break;                       // [v][n]
    }
    plan[t] = (T)v;
  }
}

// ---------------------------------------------------------------------------
// GEMM C[MxN] = alpha * A[MxK] * B[KxN] (row-major), a batch of up to 3.
// Optional epilogue: per-tile FP64 partial of sum(C .* R) (the energy).
// ---------------------------------------------------------------------------
struct GemmDesc {
  const void* A;
  const void* B;
  void* C;
  const void* R;      // nullable: energy reference map [M x N]
  double* partials;   // per-tile partials of sum(C*R) when R != null
};
struct GemmBatch {
  GemmDesc g[3];
  int M, N, K;
  double alpha;
  int TY, TX, count;  // tile grid of one GEMM, GEMMs in the batch
};

constexpr int BT = 64;          // block tile (M and N)
constexpr int kXcds = 8;        // accelerator dies; workgroup ids go round-robin over them
constexpr int kGroupRows = 4;   // tile rows per raster band
constexpr int kLargeGridTiles = 256;  // 128 x 128 tiles per launch from which the BF16x3 GEMM uses them

struct TileIdx {
  int g, ty, tx;
  bool ok;
};

// The MFMA kernels run on a 1-D grid of kXcds * ceil(tiles / kXcds)
// workgroups.  Workgroup b runs on die b % kXcds, so die x is given the
// contiguous range [x*per, (x+1)*per) of a banded raster over all tiles:
// its tiles form a compact block of rows and columns and the A and B slabs
// they read stay in that die's L2.
template <class Batch>
__device__ __forceinline__ TileIdx map_tile(const Batch& bt) {
  const int T = bt.TY * bt.TX, total = T * bt.count;
  const int per = (total + kXcds - 1) / kXcds;
  const int b = blockIdx.x;
  const int q = (b % kXcds) * per + b / kXcds;
  TileIdx r;
  r.ok = q < total;
  r.g = r.ty = r.tx = 0;
  if (!r.ok) return r;
  r.g = q / T;
  const int t = q - r.g * T;
  const int band = t / (kGroupRows * bt.TX);
  const int row0 = band * kGroupRows;
  const int rows = bt.TY - row0 < kGroupRows ? bt.TY - row0 : kGroupRows;
  const int w = t - band * kGroupRows * bt.TX;
  r.ty = row0 + w % rows;
  r.tx = w / rows;
  return r;
}

__device__ __forceinline__ const GemmDesc& desc_of(const GemmBatch& bt, int g) {
  return g == 0 ? bt.g[0] : (g == 1 ? bt.g[1] : bt.g[2]);
}

// ---- FP32: v_mfma_f32_32x32x2_f32, K tile 32 --------------------------------
// LDS rows hold 32 k values permuted as pos(k) = (k&1)*16 + (k>>1), so lane
// half h (= lane/32) reads its 16 k values (k = 2s + h) contiguously.
constexpr int F32_BK = 32;
constexpr int F32_LD = 36;  // padded row stride (floats)

__global__ void __launch_bounds__(256) k_gemm_f32(GemmBatch bt) {
  const TileIdx ti = map_tile(bt);
  if (!ti.ok) return;  // whole workgroup: before any barrier
  const GemmDesc& gd = desc_of(bt, ti.g);
  const float* A = (const float*)gd.A;
  const float* B = (const float*)gd.B;
  float* C = (float*)gd.C;
  const int N = bt.N, K = bt.K;
  const int m0 = ti.ty * BT, n0 = ti.tx * BT;
  __shared__ __attribute__((aligned(16))) float As[BT * F32_LD];
  __shared__ __attribute__((aligned(16))) float Bs[BT * F32_LD];
  const int t = threadIdx.x, lane = t & 63, wid = t >> 6;
  const int wm = wid >> 1, wn = wid & 1;
  const int h = lane >> 5, l32 = lane & 31;
  f32x16 acc = {};
  // global load assignment: A 8 consecutive k of one row, B 8 consecutive k
  // of one column (each B load coalesced across lanes); the next K tile is
  // loaded into registers while the current one feeds the MFMAs
  const int a_row = t >> 2, a_kc = (t & 3) * 8;
  const int b_col = t & 63, b_kc = (t >> 6) * 8;
  float4 a0, a1;
  float bv[8];
  auto load = [&](int k0) {
    const float4* ap = (const float4*)(A + (size_t)(m0 + a_row) * K + k0 + a_kc);
    a0 = ap[0];
    a1 = ap[1];
#pragma unroll
    for (int j = 0; j < 8; ++j) bv[j] = B[(size_t)(k0 + b_kc + j) * N + n0 + b_col];
  };
  load(0);
  for (int k0 = 0; k0 < K; k0 += F32_BK) {
    __syncthreads();
    {
      float* r = As + a_row * F32_LD;
      const int e = a_kc >> 1;  // even k -> e.., odd k -> 16 + e..
      *(float4*)(r + e) = make_float4(a0.x, a0.z, a1.x, a1.z);
      *(float4*)(r + 16 + e) = make_float4(a0.y, a0.w, a1.y, a1.w);
      float* q = Bs + b_col * F32_LD;
      const int f = b_kc >> 1;
      *(float4*)(q + f) = make_float4(bv[0], bv[2], bv[4], bv[6]);
      *(float4*)(q + 16 + f) = make_float4(bv[1], bv[3], bv[5], bv[7]);
    }
    __syncthreads();
    if (k0 + F32_BK < K) load(k0 + F32_BK);
    const float* ar = As + (wm * 32 + l32) * F32_LD + h * 16;
    const float* br = Bs + (wn * 32 + l32) * F32_LD + h * 16;
#pragma unroll
    for (int q4 = 0; q4 < 4; ++q4) {
      float4 av = *(const float4*)(ar + 4 * q4);
      float4 bv4 = *(const float4*)(br + 4 * q4);
      acc = __builtin_amdgcn_mfma_f32_32x32x2f32(av.x, bv4.x, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_32x32x2f32(av.y, bv4.y, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_32x32x2f32(av.z, bv4.z, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_32x32x2f32(av.w, bv4.w, acc, 0, 0, 0);
    }
  }
  // epilogue: acc[r] -> row 8*(r/4) + 4*h + r%4, col l32 of the wave tile
  const float alpha = (float)bt.alpha;
  const float* R = (const float*)gd.R;
  double e = 0;
  const int col = n0 + wn * 32 + l32;
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int row = m0 + wm * 32 + 8 * (r >> 2) + 4 * h + (r & 3);
    float v = alpha * acc[r];
    C[(size_t)row * N + col] = v;
    if (R) e += (double)(v * R[(size_t)row * N + col]);
  }
  if (R) {
    __shared__ double smem[4];
    double bs = block_sum<256>(e, smem);
    if (t == 0) gd.partials[ti.ty * bt.TX + ti.tx] = bs;
  }
}

// ---- BF16x3: v_mfma_f32_32x32x16_bf16, K tile 64 ---------------------------
constexpr int BF_BK = 64;
constexpr int BF_LD = 72;  // padded row stride (bf16 elements): 144 bytes

__device__ __forceinline__ void split8(const float* v, bf16x8& hi, bf16x8& lo) {
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    __bf16 hb = (__bf16)v[j];
    hi[j] = hb;
    lo[j] = (__bf16)(v[j] - (float)hb);
  }
}

// SPLIT = 2: a workgroup of 512 threads whose halves take the two halves of
// K (K a multiple of 2 * BF_BK), each with its own LDS tiles, and add their
// accumulators in a fixed order at the end.  At 1024^2 one GEMM has only 256
// tiles, so the serial chain of K steps, not the matrix cores, sets the time;
// halving it per workgroup and doubling the waves in flight shortens it.
template <int SPLIT>
__global__ void __launch_bounds__(256 * SPLIT) k_gemm_bf16x3(GemmBatch bt) {
  const TileIdx ti = map_tile(bt);
  if (!ti.ok) return;  // whole workgroup: before any barrier
  const GemmDesc& gd = desc_of(bt, ti.g);
  const float* A = (const float*)gd.A;
  const float* B = (const float*)gd.B;
  float* C = (float*)gd.C;
  const int N = bt.N, K = bt.K;
  const int m0 = ti.ty * BT, n0 = ti.tx * BT;
  __shared__ __attribute__((aligned(16))) __bf16 Ah_[SPLIT][BT * BF_LD];
  __shared__ __attribute__((aligned(16))) __bf16 Al_[SPLIT][BT * BF_LD];
  __shared__ __attribute__((aligned(16))) __bf16 Bh_[SPLIT][BT * BF_LD];
  __shared__ __attribute__((aligned(16))) __bf16 Bl_[SPLIT][BT * BF_LD];
  const int half = threadIdx.x >> 8;  // 0 when SPLIT == 1
  const int t = threadIdx.x & 255, lane = t & 63, wid = t >> 6;
  const int wm = wid >> 1, wn = wid & 1;
  const int h = lane >> 5, l32 = lane & 31;
  __bf16* Ah = Ah_[half];
  __bf16* Al = Al_[half];
  __bf16* Bh = Bh_[half];
  __bf16* Bl = Bl_[half];
  const int kspan = K / SPLIT, kbase = half * kspan;
  f32x16 acc = {};
  // A: 16 consecutive k of one row; B: 16 consecutive k of one column.  The
  // next K tile is loaded into registers while the current one is computed.
  const int a_row = t >> 2, a_kc = (t & 3) * 16;
  const int b_col = t & 63, b_kc = (t >> 6) * 16;
  float av[16], bv[16];
  auto load = [&](int k0) {
    const float4* ap = (const float4*)(A + (size_t)(m0 + a_row) * K + k0 + a_kc);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const float4 v = ap[j];
      av[4 * j] = v.x;
      av[4 * j + 1] = v.y;
      av[4 * j + 2] = v.z;
      av[4 * j + 3] = v.w;
    }
#pragma unroll
    for (int j = 0; j < 16; ++j) bv[j] = B[(size_t)(k0 + b_kc + j) * N + n0 + b_col];
  };
  load(kbase);
  // the same trip count in both halves: every thread meets every barrier
  for (int k0 = kbase; k0 < kbase + kspan; k0 += BF_BK) {
    bf16x8 ahi0, alo0, ahi1, alo1, bhi0, blo0, bhi1, blo1;
    split8(av, ahi0, alo0);
    split8(av + 8, ahi1, alo1);
    split8(bv, bhi0, blo0);
    split8(bv + 8, bhi1, blo1);
    __syncthreads();
    *(bf16x8*)(Ah + a_row * BF_LD + a_kc) = ahi0;
    *(bf16x8*)(Ah + a_row * BF_LD + a_kc + 8) = ahi1;
    *(bf16x8*)(Al + a_row * BF_LD + a_kc) = alo0;
    *(bf16x8*)(Al + a_row * BF_LD + a_kc + 8) = alo1;
    *(bf16x8*)(Bh + b_col * BF_LD + b_kc) = bhi0;
    *(bf16x8*)(Bh + b_col * BF_LD + b_kc + 8) = bhi1;
    *(bf16x8*)(Bl + b_col * BF_LD + b_kc) = blo0;
    *(bf16x8*)(Bl + b_col * BF_LD + b_kc + 8) = blo1;
    __syncthreads();
    if (k0 + BF_BK < kbase + kspan) load(k0 + BF_BK);
#pragma unroll
    for (int s = 0; s < BF_BK / 16; ++s) {
      const int ko = 16 * s + 8 * h;
      const int ra = (wm * 32 + l32) * BF_LD + ko;
      const int rb = (wn * 32 + l32) * BF_LD + ko;
      bf16x8 xh = *(const bf16x8*)(Ah + ra);
      bf16x8 xl = *(const bf16x8*)(Al + ra);
      bf16x8 yh = *(const bf16x8*)(Bh + rb);
      bf16x8 yl = *(const bf16x8*)(Bl + rb);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(xl, yh, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(xh, yl, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(xh, yh, acc, 0, 0, 0);
    }
  }
  if constexpr (SPLIT == 2) {
    // the upper half's accumulators (4 waves x 16 x 64 floats = 16 KB)
    // through the A-tile array of both halves (18 KB, no longer read), added
    // to the lower half's: (first K half) + (second K half)
    static_assert(sizeof(Ah_) >= 4 * 16 * 64 * sizeof(float), "reduction buffer");
    float* red = (float*)&Ah_[0][0];
    __syncthreads();
    if (half == 1) {
#pragma unroll
      for (int r = 0; r < 16; ++r) red[(r * 4 + wid) * 64 + lane] = acc[r];
    }
    __syncthreads();
    if (half == 0) {
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[r] = acc[r] + red[(r * 4 + wid) * 64 + lane];
    }
  }
  const float alpha = (float)bt.alpha;
  const float* R = (const float*)gd.R;
  double e = 0;
  if (half == 0) {
    const int col = n0 + wn * 32 + l32;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int row = m0 + wm * 32 + 8 * (r >> 2) + 4 * h + (r & 3);
      float v = alpha * acc[r];
      C[(size_t)row * N + col] = v;
      if (R) e += (double)(v * R[(size_t)row * N + col]);
    }
  }
  if (R) {
    __shared__ double smem[4 * SPLIT];
    double bs = block_sum<256 * SPLIT>(e, smem);
    if (threadIdx.x == 0) gd.partials[ti.ty * bt.TX + ti.tx] = bs;
  }
}

// BF16x3 on 128 x 128 block tiles, for large grids.  The 64 x 64 tiles
// above read each A row panel and B column panel once per tile, through L2:
// at 2048^2 that is 1 GB per GEMM and L2 bandwidth sets the time.  Here each
// of 4 wavefronts owns a 64 x 64 tile (2 x 2 MFMA tiles), which halves the
// L2 traffic and reads each LDS fragment for two MFMA tiles.  The two halves
// of the 512-thread workgroup take the two halves of K as above.  Every
// output element sees the same MFMA sequence (the same K order, lo*hi, hi*lo,
// hi*hi per K step, the same K-half addition) as in k_gemm_bf16x3<2>, so
// the two kernels give the same bits.
constexpr int BT2 = 128;

__global__ void __launch_bounds__(512) k_gemm_bf16x3_t128(GemmBatch bt) {
  const TileIdx ti = map_tile(bt);
  if (!ti.ok) return;  // whole workgroup: before any barrier
  const GemmDesc& gd = desc_of(bt, ti.g);
  const float* A = (const float*)gd.A;
  const float* B = (const float*)gd.B;
  float* C = (float*)gd.C;
  const int N = bt.N, K = bt.K;
  const int m0 = ti.ty * BT2, n0 = ti.tx * BT2;
  constexpr int kTileElems = BT2 * BF_LD;
  // [half][Ah, Al, Bh, Bl]; one array so the K-half reduction can reuse it
  __shared__ __attribute__((aligned(16))) __bf16 lds[2][4][kTileElems];
  const int half = threadIdx.x >> 8;
  const int t = threadIdx.x & 255, lane = t & 63, wid = t >> 6;
  const int wm = wid >> 1, wn = wid & 1;
  const int h = lane >> 5, l32 = lane & 31;
  __bf16* Ah = lds[half][0];
  __bf16* Al = lds[half][1];
  __bf16* Bh = lds[half][2];
  __bf16* Bl = lds[half][3];
  const int kspan = K / 2, kbase = half * kspan;
  f32x16 acc[2][2] = {};
  // A: 32 consecutive k of one row; B: 32 consecutive k of one column (each
  // B load coalesced across lanes).  The next K tile is loaded into
  // registers while the current one is computed.
  const int a_row = t >> 1, a_kc = (t & 1) * 32;
  const int b_col = t & 127, b_kc = (t >> 7) * 32;
  float av[32], bv[32];
  auto load = [&](int k0) {
    const float4* ap = (const float4*)(A + (size_t)(m0 + a_row) * K + k0 + a_kc);
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const float4 v = ap[j];
      av[4 * j] = v.x;
      av[4 * j + 1] = v.y;
      av[4 * j + 2] = v.z;
      av[4 * j + 3] = v.w;
    }
#pragma unroll
    for (int j = 0; j < 32; ++j) bv[j] = B[(size_t)(k0 + b_kc + j) * N + n0 + b_col];
  };
  load(kbase);
  for (int k0 = kbase; k0 < kbase + kspan; k0 += BF_BK) {
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      bf16x8 hi, lo;
      split8(av + 8 * q, hi, lo);
      *(bf16x8*)(Ah + a_row * BF_LD + a_kc + 8 * q) = hi;
      *(bf16x8*)(Al + a_row * BF_LD + a_kc + 8 * q) = lo;
      split8(bv + 8 * q, hi, lo);
      *(bf16x8*)(Bh + b_col * BF_LD + b_kc + 8 * q) = hi;
      *(bf16x8*)(Bl + b_col * BF_LD + b_kc + 8 * q) = lo;
    }
    __syncthreads();
    if (k0 + BF_BK < kbase + kspan) load(k0 + BF_BK);
#pragma unroll
    for (int s = 0; s < BF_BK / 16; ++s) {
      const int ko = 16 * s + 8 * h;
      bf16x8 xh[2], xl[2], yh[2], yl[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        const int ra = (wm * 64 + 32 * i + l32) * BF_LD + ko;
        const int rb = (wn * 64 + 32 * i + l32) * BF_LD + ko;
        xh[i] = *(const bf16x8*)(Ah + ra);
        xl[i] = *(const bf16x8*)(Al + ra);
        yh[i] = *(const bf16x8*)(Bh + rb);
        yl[i] = *(const bf16x8*)(Bl + rb);
      }
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(xl[i], yh[j], acc[i][j], 0, 0, 0);
          acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(xh[i], yl[j], acc[i][j], 0, 0, 0);
          acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(xh[i], yh[j], acc[i][j], 0, 0, 0);
        }
    }
  }
  // the upper half's accumulators (4 waves x 64 x 64 floats = 64 KB) through
  // the tile arrays (no longer read), added to the lower half's
  static_assert(sizeof(lds) >= 4 * 64 * 64 * sizeof(float), "reduction buffer");
  float* red = (float*)&lds[0][0][0];
  __syncthreads();
  if (half == 1) {
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int r = 0; r < 16; ++r) red[(((i * 2 + j) * 16 + r) * 4 + wid) * 64 + lane] = acc[i][j][r];
  }
  __syncthreads();
  const float alpha = (float)bt.alpha;
  const float* R = (const float*)gd.R;
  double e = 0;
  if (half == 0) {
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const int col = n0 + wn * 64 + 32 * j + l32;
#pragma unroll
        for (int r = 0; r < 16; ++r) {
          const float a = acc[i][j][r] + red[(((i * 2 + j) * 16 + r) * 4 + wid) * 64 + lane];
          const int row = m0 + wm * 64 + 32 * i + 8 * (r >> 2) + 4 * h + (r & 3);
          const float v = alpha * a;
          C[(size_t)row * N + col] = v;
          if (R) e += (double)(v * R[(size_t)row * N + col]);
        }
      }
  }
  if (R) {
    __shared__ double smem[8];
    double bs = block_sum<512>(e, smem);
    if (threadIdx.x == 0) gd.partials[ti.ty * bt.TX + ti.tx] = bs;
  }
}

// ---- BF16x3 on pre-split operands (the Poisson solve) -----------------------
// The operands arrive as bf16 hi/lo planes, both K-contiguous: A as [M][K]
// rows and B as [N][K] (the transpose of the K x N operand).  Their producers
// split each FP32 value v once, hi = bf16(v), lo = bf16(v - hi), exactly as
// split8 does in the kernels above: the centring (rho), the spectral weights,
// the first GEMM of each transform (which writes its product transposed, as
// the second GEMM's B) and the plan (the transform matrices).  A K tile is
// then a plain copy into LDS, with no conversion work in the GEMM.  The MFMA
// sequence per output element (K order, lo*hi, hi*lo, hi*hi, the two K
// halves added at the end) is the in-kernel-split kernels', so the products
// have their bits.
//   BM = 64: 64 x 64 tiles, four 32 x 32 wavefronts per K half, two LDS
//   stages: the next K tile is written while the current one feeds the
//   matrix cores, one barrier per K step.
//   BM = 128: 128 x 128 tiles (grids of >= 256 of them), four 64 x 64
//   wavefronts per K half, one LDS stage (two would exceed the LDS).
struct B3Desc {
  const __bf16* A;  // hi plane; lo plane at A + B3Batch::plane_a
  const __bf16* B;  // hi plane; lo plane at B + B3Batch::plane_b
  float* C;         // [M][N] FP32 (kOutT false)
  __bf16* CT;       // [N][M] hi plane, lo plane at CT + M*N (kOutT)
  const float* R;   // nullable: energy reference map [M][N] (kOutT false)
  double* partials;
};
struct B3Batch {
  B3Desc g[3];
  int M, N, K;
  int64_t plane_a, plane_b;
  float alpha;
  int TY, TX, count;
};

__device__ __forceinline__ const B3Desc& desc_of3(const B3Batch& bt, int g) {
  return g == 0 ? bt.g[0] : (g == 1 ? bt.g[1] : bt.g[2]);
}

__device__ __forceinline__ void split1(float v, __bf16& hi, __bf16& lo) {
  hi = (__bf16)v;
  lo = (__bf16)(v - (float)hi);
}

template <int BM, bool kOutT>
__global__ void __launch_bounds__(512) k_gemm_b3(B3Batch bt) {
  constexpr int STAGES = BM == 64 ? 2 : 1;
  constexpr int TE = BM * BF_LD;  // bf16 elements of one plane tile
  constexpr int WT = BM / 64;     // 32 x 32 MFMA tiles per wavefront side
  constexpr int RPT = BM / 64;    // rows per thread of each operand tile
  const TileIdx ti = map_tile(bt);
  if (!ti.ok) return;  // whole workgroup: before any barrier
  const B3Desc& gd = desc_of3(bt, ti.g);
  const int M = bt.M, N = bt.N, K = bt.K;
  const int m0 = ti.ty * BM, n0 = ti.tx * BM;
  // [K half][stage][Ah, Al, Bh, Bl]
  __shared__ __attribute__((aligned(16))) __bf16 lds[2][STAGES][4][TE];
  const int half = threadIdx.x >> 8;
  const int t = threadIdx.x & 255, lane = t & 63, wid = t >> 6;
  const int wm = wid >> 1, wn = wid & 1;
  const int h = lane >> 5, l32 = lane & 31;
  const int kspan = K / 2, kbase = half * kspan, nsteps = kspan / BF_BK;
  const __bf16* Ah = gd.A;
  const __bf16* Al = gd.A + bt.plane_a;
  const __bf16* Bh = gd.B;
  const __bf16* Bl = gd.B + bt.plane_b;
  // thread t copies rows (t >> 2) + 64 i of both operand tiles, 16-byte
  // chunks c and c + 4 (c = t & 3) of each 64-element K row; entry
  // u = 2 i + j of each register array holds row i, chunk j
  const int lr = t >> 2, lc = t & 3;
  constexpr int NR = 2 * RPT;
  // D register sets, each one K step's operand chunks: the loads of step
  // s + D are issued while step s is computed, so D steps of load latency
  // overlap (a K step's 12 MFMAs per wavefront are far shorter than one
  // load's latency).  Native vectors, compile-time indices: registers.  At
  // 2 waves per SIMD a wavefront has 256 registers: 64 x 64 tiles take two
  // sets (64 registers), 128 x 128 tiles one (its 64 accumulators and 2 x 2
  // operand fragments leave no room for a second).
  constexpr int D = BM == 64 ? 2 : 1;
  u32x4 rah[D][NR], ral[D][NR], rbh[D][NR], rbl[D][NR];
#define B3_LOAD(k0, d)                                                   \
  _Pragma("unroll") for (int u = 0; u < NR; ++u) {                       \
    const int r = lr + 64 * (u >> 1), q = lc + 4 * (u & 1);              \
    const size_t oa = (size_t)(m0 + r) * K + (k0) + 8 * q;               \
    const size_t ob = (size_t)(n0 + r) * K + (k0) + 8 * q;               \
    rah[d][u] = *(const u32x4*)(Ah + oa);                                \
    ral[d][u] = *(const u32x4*)(Al + oa);                                \
    rbh[d][u] = *(const u32x4*)(Bh + ob);                                \
    rbl[d][u] = *(const u32x4*)(Bl + ob);                                \
  }
#define B3_STORE(st, d)                                                  \
  _Pragma("unroll") for (int u = 0; u < NR; ++u) {                       \
    const int r = lr + 64 * (u >> 1), q = lc + 4 * (u & 1);              \
    const int o = r * BF_LD + 8 * q;                                     \
    *(u32x4*)&lds[half][st][0][o] = rah[d][u];                           \
    *(u32x4*)&lds[half][st][1][o] = ral[d][u];                           \
    *(u32x4*)&lds[half][st][2][o] = rbh[d][u];                           \
    *(u32x4*)&lds[half][st][3][o] = rbl[d][u];                           \
  }
  f32x16 acc[WT][WT] = {};
  auto compute = [&](int st) {
    const __bf16* Xh = lds[half][st][0];
    const __bf16* Xl = lds[half][st][1];
    const __bf16* Yh = lds[half][st][2];
    const __bf16* Yl = lds[half][st][3];
#pragma unroll
    for (int sub = 0; sub < BF_BK / 16; ++sub) {
      const int ko = 16 * sub + 8 * h;
      bf16x8 xh[WT], xl[WT], yh[WT], yl[WT];
#pragma unroll
      for (int i = 0; i < WT; ++i) {
        const int oa = (wm * 32 * WT + 32 * i + l32) * BF_LD + ko;
        const int ob = (wn * 32 * WT + 32 * i + l32) * BF_LD + ko;
        xh[i] = *(const bf16x8*)(Xh + oa);
        xl[i] = *(const bf16x8*)(Xl + oa);
        yh[i] = *(const bf16x8*)(Yh + ob);
        yl[i] = *(const bf16x8*)(Yl + ob);
      }
#pragma unroll
      for (int i = 0; i < WT; ++i)
#pragma unroll
        for (int j = 0; j < WT; ++j) {
          acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(xl[i], yh[j], acc[i][j], 0, 0, 0);
          acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(xh[i], yl[j], acc[i][j], 0, 0, 0);
          acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(xh[i], yh[j], acc[i][j], 0, 0, 0);
        }
    }
  };
  // The trip count is the same in both halves and every condition below is
  // uniform over the workgroup: every thread meets every barrier.  Register
  // set j of the ring holds K step s with s % D == j.  When the steps are a
  // whole number of rings, the loop issues its loads unconditionally and the
  // last D steps load nothing: with no branch around the loads the compiler
  // waits for the oldest set only, not for every load in flight.
  const bool rings = nsteps % D == 0;
  if (rings) {
#pragma unroll
    for (int d = 0; d < D; ++d) B3_LOAD(kbase + d * BF_BK, d);
  } else {
#pragma unroll
    for (int d = 0; d < D; ++d)
      if (d < nsteps) B3_LOAD(kbase + d * BF_BK, d);
  }
  if constexpr (STAGES == 2) {
    // step s is computed from LDS stage s & 1 while step s + 1 waits in
    // registers; it goes to the other stage after the computation
    static_assert(D % 2 == 0, "stage parity follows the ring index");  // 64 x 64: D = 2
    B3_STORE(0, 0);
    __syncthreads();
    if (rings) {
      for (int s0 = 0; s0 < nsteps - D; s0 += D) {
#pragma unroll
        for (int j = 0; j < D; ++j) {
          B3_LOAD(kbase + (s0 + j + D) * BF_BK, j);
          compute(j & 1);
          B3_STORE((j + 1) & 1, (j + 1) % D);
          __syncthreads();
        }
      }
#pragma unroll
      for (int j = 0; j < D; ++j) {  // the last D steps
        compute(j & 1);
        if (j + 1 < D) B3_STORE((j + 1) & 1, (j + 1) % D);
        __syncthreads();
      }
    } else {
      for (int s0 = 0; s0 < nsteps; s0 += D) {
#pragma unroll
        for (int j = 0; j < D; ++j) {
          const int s = s0 + j;
          if (s < nsteps) {
            if (s + D < nsteps) B3_LOAD(kbase + (s + D) * BF_BK, j);
            compute(s & 1);
            if (s + 1 < nsteps) B3_STORE((s + 1) & 1, (j + 1) % D);
            __syncthreads();
          }
        }
      }
    }
  } else {
    if (rings) {
      for (int s0 = 0; s0 < nsteps - D; s0 += D) {
#pragma unroll
        for (int j = 0; j < D; ++j) {
          __syncthreads();  // the previous step's reads of the stage are done
          B3_STORE(0, j);
          __syncthreads();
          B3_LOAD(kbase + (s0 + j + D) * BF_BK, j);
          compute(0);
        }
      }
#pragma unroll
      for (int j = 0; j < D; ++j) {  // the last D steps
        __syncthreads();
        B3_STORE(0, j);
        __syncthreads();
        compute(0);
      }
    } else {
      for (int s0 = 0; s0 < nsteps; s0 += D) {
#pragma unroll
        for (int j = 0; j < D; ++j) {
          const int s = s0 + j;
          if (s < nsteps) {
            __syncthreads();
            B3_STORE(0, j);
            __syncthreads();
            if (s + D < nsteps) B3_LOAD(kbase + (s + D) * BF_BK, j);
            compute(0);
          }
        }
      }
    }
  }
#undef B3_LOAD
#undef B3_STORE
  // the upper K half's accumulators through LDS (no longer read), added to
  // the lower half's: (first K half) + (second K half)
  static_assert(sizeof(lds) >= 4 * WT * WT * 16 * 64 * sizeof(float), "reduction buffer");
  float* red = (float*)&lds[0][0][0][0];
  __syncthreads();
  if (half == 1) {
#pragma unroll
    for (int i = 0; i < WT; ++i)
#pragma unroll
      for (int j = 0; j < WT; ++j)
#pragma unroll
        for (int r = 0; r < 16; ++r) red[(((i * WT + j) * 16 + r) * 4 + wid) * 64 + lane] = acc[i][j][r];
  }
  __syncthreads();
  const float alpha = bt.alpha;
  double e = 0;
  if (half == 0) {
#pragma unroll
    for (int i = 0; i < WT; ++i)
#pragma unroll
      for (int j = 0; j < WT; ++j) {
        const int col = n0 + wn * 32 * WT + 32 * j + l32;
        const int rbase = m0 + wm * 32 * WT + 32 * i + 4 * h;
#pragma unroll
        for (int g = 0; g < 4; ++g) {
          float v[4];
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const int r = 4 * g + u;
            const float a = acc[i][j][r] + red[(((i * WT + j) * 16 + r) * 4 + wid) * 64 + lane];
            v[u] = alpha * a;
          }
          const int row0 = rbase + 8 * g;
          if constexpr (kOutT) {
            // rows row0..row0+3 of column col: four consecutive entries of C^T
            union {
              __bf16 b[4];
              uint2 w;
            } hi, lo;
#pragma unroll
            for (int u = 0; u < 4; ++u) split1(v[u], hi.b[u], lo.b[u]);
            __bf16* ct = gd.CT + (size_t)col * M + row0;
            *(uint2*)ct = hi.w;
            *(uint2*)(ct + (size_t)M * N) = lo.w;
          } else {
#pragma unroll
            for (int u = 0; u < 4; ++u) {
              const size_t o = (size_t)(row0 + u) * N + col;
              gd.C[o] = v[u];
              if (gd.R) e += (double)(v[u] * gd.R[o]);
            }
          }
        }
      }
  }
  if (!kOutT && gd.R) {
    __shared__ double smem[8];
    double bs = block_sum<512>(e, smem);
    if (threadIdx.x == 0) gd.partials[ti.ty * bt.TX + ti.tx] = bs;
  }
}

// ---- FP64: v_mfma_f64_16x16x4_f64, K tile 16 --------------------------------
// Each wave owns a 32x32 tile = 2x2 MFMA tiles.  LDS rows hold 16 k values
// permuted as pos(k) = (k&3)*4 + (k>>2), so lane quarter q (= lane/16) reads
// its 4 k values (k = 4s + q) contiguously.
constexpr int F64_BK = 16;
constexpr int F64_LD = 18;  // padded row stride (doubles)

__global__ void __launch_bounds__(256) k_gemm_f64(GemmBatch bt) {
  const TileIdx ti = map_tile(bt);
  if (!ti.ok) return;  // whole workgroup: before any barrier
  const GemmDesc& gd = desc_of(bt, ti.g);
  const double* A = (const double*)gd.A;
  const double* B = (const double*)gd.B;
  double* C = (double*)gd.C;
  const int N = bt.N, K = bt.K;
  const int m0 = ti.ty * BT, n0 = ti.tx * BT;
  __shared__ __attribute__((aligned(16))) double As[BT * F64_LD];
  __shared__ __attribute__((aligned(16))) double Bs[BT * F64_LD];
  const int t = threadIdx.x, lane = t & 63, wid = t >> 6;
  const int wm = wid >> 1, wn = wid & 1;
  const int q = lane >> 4, l16 = lane & 15;
  f64x4 acc[2][2] = {};
  // 4 consecutive k of one row (A) / one column (B); the next K tile is
  // loaded into registers while the current one is computed
  const int a_row = t >> 2, a_kc = (t & 3) * 4;
  const int b_col = t & 63, b_kc = (t >> 6) * 4;
  double av[4], bv[4];
  auto load = [&](int k0) {
    const double2* ap = (const double2*)(A + (size_t)(m0 + a_row) * K + k0 + a_kc);
    const double2 p0 = ap[0], p1 = ap[1];
    av[0] = p0.x;
    av[1] = p0.y;
    av[2] = p1.x;
    av[3] = p1.y;
#pragma unroll
    for (int j = 0; j < 4; ++j) bv[j] = B[(size_t)(k0 + b_kc + j) * N + n0 + b_col];
  };
  load(0);
  for (int k0 = 0; k0 < K; k0 += F64_BK) {
    __syncthreads();
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int k = a_kc + j;
      As[a_row * F64_LD + (k & 3) * 4 + (k >> 2)] = av[j];
      const int kb = b_kc + j;
      Bs[b_col * F64_LD + (kb & 3) * 4 + (kb >> 2)] = bv[j];
    }
    __syncthreads();
    if (k0 + F64_BK < K) load(k0 + F64_BK);
#pragma unroll
    for (int s = 0; s < F64_BK / 4; ++s) {
      double a0 = As[(wm * 32 + l16) * F64_LD + q * 4 + s];
      double a1 = As[(wm * 32 + 16 + l16) * F64_LD + q * 4 + s];
      double b0 = Bs[(wn * 32 + l16) * F64_LD + q * 4 + s];
      double b1 = Bs[(wn * 32 + 16 + l16) * F64_LD + q * 4 + s];
      acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
      acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
      acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
      acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
    }
  }
  const double alpha = bt.alpha;
  const double* R = (const double*)gd.R;
  double e = 0;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        // v_mfma_f64_16x16x4_f64 accumulator layout: acc[r] of lane l holds
        // D[4*r + l/16][l%16] (measured with one-hot operands on gfx950)
        const int row = m0 + wm * 32 + 16 * i + 4 * r + q;
        const int col = n0 + wn * 32 + 16 * j + l16;
        double v = alpha * acc[i][j][r];
        C[(size_t)row * N + col] = v;
        if (R) e += v * R[(size_t)row * N + col];
      }
  if (R) {
    __shared__ double smem[4];
    double bs = block_sum<256>(e, smem);
    if (t == 0) gd.partials[ti.ty * bt.TX + ti.tx] = bs;
  }
}

// ---- Generic LDS-tiled FMA GEMM for any size (T = float or double) ---------
template <typename T>
__global__ void __launch_bounds__(256) k_gemm_generic(GemmBatch bt) {
  const GemmDesc& gd = bt.g[blockIdx.z];
  const T* A = (const T*)gd.A;
  const T* B = (const T*)gd.B;
  T* C = (T*)gd.C;
  const int M = bt.M, N = bt.N, K = bt.K;
  __shared__ T As[16][17];
  __shared__ T Bs[16][17];
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  const int row = blockIdx.y * 16 + ty, col = blockIdx.x * 16 + tx;
  T acc = 0;
  for (int k0 = 0; k0 < K; k0 += 16) {
    As[ty][tx] = (row < M && k0 + tx < K) ? A[(size_t)row * K + k0 + tx] : T(0);
    Bs[ty][tx] = (k0 + ty < K && col < N) ? B[(size_t)(k0 + ty) * N + col] : T(0);
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 16; ++k) acc += As[ty][k] * Bs[k][tx];
    __syncthreads();
  }
  double e = 0;
  if (row < M && col < N) {
    T v = (T)bt.alpha * acc;
    C[(size_t)row * N + col] = v;
    if (gd.R) e = (double)v * (double)((const T*)gd.R)[(size_t)row * N + col];
  }
  if (gd.R) {
    __shared__ double smem[4];
    double bs = block_sum<256>(e, smem);
    if (threadIdx.x == 0) gd.partials[blockIdx.y * gridDim.x + blockIdx.x] = bs;
  }
}

static bool mfma_shape(int M, int N) { return M % BT == 0 && N % BT == 0; }

// Launch a batch of `count` GEMMs of one shape; returns the number of
// per-GEMM energy partial blocks.
static int launch_gemm(int precision, GemmBatch& bt, int count, hipStream_t s, int* nblocks) {
  dim3 block(256);
  // 128 x 128 tiles where they still give every CU a workgroup
  if (precision == DP_SPECTRAL_BF16X3 && bt.M % BT2 == 0 && bt.N % BT2 == 0 &&
      bt.K % (2 * BF_BK) == 0 &&
      (int64_t)(bt.M / BT2) * (bt.N / BT2) * count >= kLargeGridTiles) {
    bt.TY = bt.M / BT2;
    bt.TX = bt.N / BT2;
    bt.count = count;
    const int64_t tiles = (int64_t)bt.TY * bt.TX * count;
    if (nblocks) *nblocks = bt.TY * bt.TX;
    hipLaunchKernelGGL(k_gemm_bf16x3_t128, dim3((unsigned)(kXcds * ceil_div(tiles, kXcds))),
                       dim3(512), 0, s, bt);
  } else if (mfma_shape(bt.M, bt.N) && bt.K % BT == 0) {
    bt.TY = bt.M / BT;
    bt.TX = bt.N / BT;
    bt.count = count;
    const int64_t tiles = (int64_t)bt.TY * bt.TX * count;
    dim3 grid((unsigned)(kXcds * ceil_div(tiles, kXcds)));
    if (nblocks) *nblocks = bt.TY * bt.TX;
    if (precision == DP_SPECTRAL_FP64)
      hipLaunchKernelGGL(k_gemm_f64, grid, block, 0, s, bt);
    else if (precision == DP_SPECTRAL_FP32)
      hipLaunchKernelGGL(k_gemm_f32, grid, block, 0, s, bt);
    else if (bt.K % (2 * BF_BK) == 0)
      hipLaunchKernelGGL(k_gemm_bf16x3<2>, grid, dim3(512), 0, s, bt);
    else
      hipLaunchKernelGGL(k_gemm_bf16x3<1>, grid, block, 0, s, bt);
  } else {
    dim3 grid((unsigned)ceil_div(bt.N, 16), (unsigned)ceil_div(bt.M, 16), count);
    bool energy = false;
    for (int i = 0; i < count; ++i) energy |= bt.g[i].R != nullptr;
    DP_REQUIRE(!energy || (int64_t)grid.x * grid.y <= kReduceMaxBlocks, DP_ERR_UNSUPPORTED,
               "spectral: %d x %d grid needs the matrix-core path (sides multiple of %d) for the "
               "fused energy",
               bt.M, bt.N, BT);
    if (nblocks) *nblocks = (int)(grid.x * grid.y);
    if (precision == DP_SPECTRAL_FP64)
      hipLaunchKernelGGL(k_gemm_generic<double>, grid, block, 0, s, bt);
    else
      hipLaunchKernelGGL(k_gemm_generic<float>, grid, block, 0, s, bt);
  }
  DP_LAUNCHED("spectral gemm");
  return DP_OK;
}

// ---------------------------------------------------------------------------
// Elementwise pieces of the solve.
// ---------------------------------------------------------------------------
// Partial sums of the map for its mean: each thread sums its few elements
// in T, the threads' sums are reduced in FP64 in a fixed order.
template <typename T>
__global__ void __launch_bounds__(kReduceBlock) k_sum_partial(const T* in, int64_t n,
                                                              double* partials) {
  __shared__ double smem[kReduceBlock / 64];
  T acc = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    acc += in[i];
  double bs = block_sum<kReduceBlock>((double)acc, smem);
  if (threadIdx.x == 0) partials[blockIdx.x] = bs;
}

// rho = map / (bx*by) (electric_potential.py:160), centred: in*s - mean as
// one fused multiply-add in T (the map's sum is a deterministic FP64
// reduction).  The DC coefficient of rho does not reach the fields or the
// potential (K[0,0] = 0, and sum(phi) = 0 so the energy sum(phi*rho) is
// unchanged), so the rounding of s and of the mean only moves that
// coefficient, while centring halves the rounding error of every other one.
template <typename T>
__device__ __forceinline__ T centre_scale(T s, const double* sum, int64_t n) {
  return (T)(*sum) * s / (T)n;
}
template <typename T>
__global__ void k_scale_centre(const T* in, int64_t n, T s, const double* sum, T* out) {
  const T mean = centre_scale(s, sum, n);
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    out[i] = fma(in[i], s, -mean);
}

// Spectral multipliers of electric_potential.py:504-519, evaluated per
// element in T from the two grid steps du = 2 pi/M and dv = 2 pi/N * (bx/by)
// (rounded to T once, on the host): wu = du*u, wv = dv*v,
// K = 1/(wu^2+wv^2), K[0,0] = 0.  In FP32 each multiplier is within a few
// ulp of its correctly rounded value.
template <typename T>
struct SpectralMul {
  T kx, ky, k, e;  // wu*K/2, wv*K/2, K, c_u c_v K (the energy's weight)
};
template <typename T>
__device__ __forceinline__ SpectralMul<T> spectral_mul(int u, int v, T du, T dv) {
  const T wu = du * (T)u, wv = dv * (T)v;
  const T w2 = wu * wu + wv * wv;
  const T inv = (u == 0 && v == 0) ? (T)0 : (T)1 / w2;
  const T c = (u ? (T)2 : (T)1) * (v ? (T)2 : (T)1);
  return {wu * inv * (T)0.5, wv * inv * (T)0.5, inv, c * inv};
}

// Writes a*wu*K/2, a*wv*K/2 and (optionally) a*K.  With `partials` it also
// forms the per-workgroup partials of the spectral energy
// (MN/4) sum c_u c_v K a^2 in the same pass (each thread sums its terms in T,
// the threads' sums are reduced in FP64).
template <typename T>
__global__ void __launch_bounds__(kReduceBlock) k_spectral_weights(
    const T* a, int M, int N, T du, T dv, T* ax, T* ay, T* ak, double* partials) {
  __shared__ double smem[kReduceBlock / 64];
  const int64_t n = (int64_t)M * N;
  T acc = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int u = (int)(i / N), v = (int)(i % N);
    const SpectralMul<T> m = spectral_mul<T>(u, v, du, dv);
    const T av = a[i];
    ax[i] = av * m.kx;
    ay[i] = av * m.ky;
    if (ak) ak[i] = av * m.k;
    if (partials) acc += m.e * av * av;
  }
  if (!partials) return;  // uniform: the whole grid returns
  double bs = block_sum<kReduceBlock>((double)acc, smem);
  if (threadIdx.x == 0) partials[blockIdx.x] = bs * ((double)M * N / 4.0);
}

// BF16x3 plans also hold the six matrices as bf16 hi/lo planes, after the
// FP32 matrices: the three M x M left matrices as A operands ([m][k], as
// stored) and the three N x N right matrices transposed as B operands
// ([n][k]); each matrix's hi plane, then its lo plane.
static inline size_t split_plane_bytes(int M, int N) {
  return 2 * sizeof(__bf16) * (3 * (size_t)M * M + 3 * (size_t)N * N);
}
// Elements from the planes' start to matrix `id`'s hi plane.
__host__ __device__ inline size_t split_plane_offset(int M, int N, int id) {
  const size_t mm = (size_t)M * M, nn = (size_t)N * N;
  return id < 3 ? 2 * id * mm : 6 * mm + 2 * (id - 3) * nn;
}

__global__ void k_plan_split(const float* plan, __bf16* planes, int M, int N) {
  const size_t mm = (size_t)M * M, nn = (size_t)N * N;
  const size_t total = 3 * mm + 3 * nn;
  for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < total;
       t += (size_t)gridDim.x * blockDim.x) {
    if (t < 3 * mm) {  // left matrix id, element (r, c) -> A plane [r][c]
      const int id = (int)(t / mm);
      const size_t e = t % mm;
      __bf16* hp = planes + split_plane_offset(M, N, id);
      split1(plan[t], hp[e], hp[mm + e]);
    } else {  // right matrix id, element (r, c) -> B plane [c][r]
      const size_t u = t - 3 * mm;
      const int id = 3 + (int)(u / nn);
      const size_t e = u % nn;
      const size_t r = e / N, c = e % N;
      __bf16* hp = planes + split_plane_offset(M, N, id);
      split1(plan[t], hp[c * N + r], hp[nn + c * N + r]);
    }
  }
}

// k_scale_centre with its output as the first GEMM's A planes (hi at P, lo
// at P + n) and, when r32 is given, as FP32 too (the energy reference).
__global__ void k_scale_centre_split(const float* in, int64_t n, float s, const double* sum,
                                     __bf16* P, float* r32) {
  const float mean = centre_scale(s, sum, n);
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const float v = fmaf(in[i], s, -mean);
    split1(v, P[i], P[n + i]);
    if (r32) r32[i] = v;
  }
}

// k_spectral_weights with the three products written as A planes (pk
// nullable) and the same spectral energy partials.
__global__ void __launch_bounds__(kReduceBlock) k_spectral_weights_split(
    const float* a, int M, int N, float du, float dv, __bf16* px, __bf16* py, __bf16* pk,
    double* partials) {
  __shared__ double smem[kReduceBlock / 64];
  const int64_t n = (int64_t)M * N;
  float acc = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int u = (int)(i / N), v = (int)(i % N);
    const SpectralMul<float> m = spectral_mul<float>(u, v, du, dv);
    const float av = a[i];
    split1(av * m.kx, px[i], px[n + i]);
    split1(av * m.ky, py[i], py[n + i]);
    if (pk) split1(av * m.k, pk[i], pk[n + i]);
    if (partials) acc += m.e * av * av;
  }
  if (!partials) return;  // uniform: the whole grid returns
  double bs = block_sum<kReduceBlock>((double)acc, smem);
  if (threadIdx.x == 0) partials[blockIdx.x] = bs * ((double)M * N / 4.0);
}

// A batch of `count` pre-split GEMMs of one shape (kOutT: products written as
// C^T planes).  The shape must satisfy presplit_shape.
static bool presplit_shape(int M, int N) {
  return M % (2 * BF_BK) == 0 && N % (2 * BF_BK) == 0 && M % BT == 0 && N % BT == 0;
}
static int launch_b3(B3Batch& bt, int count, bool out_t, hipStream_t s, int* nblocks) {
  const bool big = bt.M % BT2 == 0 && bt.N % BT2 == 0 &&
                   (int64_t)(bt.M / BT2) * (bt.N / BT2) * count >= kLargeGridTiles;
  const int bm = big ? BT2 : BT;
  bt.TY = bt.M / bm;
  bt.TX = bt.N / bm;
  bt.count = count;
  const int64_t tiles = (int64_t)bt.TY * bt.TX * count;
  dim3 grid((unsigned)(kXcds * ceil_div(tiles, kXcds)));
  if (nblocks) *nblocks = bt.TY * bt.TX;
  if (big) {
    if (out_t)
      hipLaunchKernelGGL((k_gemm_b3<BT2, true>), grid, dim3(512), 0, s, bt);
    else
      hipLaunchKernelGGL((k_gemm_b3<BT2, false>), grid, dim3(512), 0, s, bt);
  } else {
    if (out_t)
      hipLaunchKernelGGL((k_gemm_b3<BT, true>), grid, dim3(512), 0, s, bt);
    else
      hipLaunchKernelGGL((k_gemm_b3<BT, false>), grid, dim3(512), 0, s, bt);
  }
  DP_LAUNCHED("k_gemm_b3");
  return DP_OK;
}

// Workspace of the spectral ops: seven M*N slabs of the data type, then one
// reduction slot.  The single transforms use the first slab only.
static size_t ws_slab(int M, int N, int precision) {
  return align256((size_t)M * N * elem_bytes(precision));
}

// The BF16x3 solve on pre-split operands (k_gemm_b3): the same sequence of
// products as the in-kernel-split path and the same bits.  Workspace slabs:
// 0 rho (FP32, the energy reference of the potential), 1 rho planes, then
// Yphi^T; 2 the DCT2's first product (transposed planes), then Yx^T; 3 the
// DCT2 (FP32), then Yy^T; 4-6 the planes of a*wu*K/2, a*wv*K/2, a*K.
static int poisson_presplit(const void* plan, int M, int N, const float* map, double bsx,
                            double bsy, bool want_phi, float* field_x, float* field_y,
                            float* potential, float* energy, char* w, hipStream_t s) {
  const int precision = DP_SPECTRAL_BF16X3;
  const size_t slab = ws_slab(M, N, precision);
  const int64_t n = (int64_t)M * N;
  float* rho32 = (float*)w;
  __bf16* rhoP = (__bf16*)(w + slab);
  __bf16* tmpT = (__bf16*)(w + 2 * slab);
  float* a32 = (float*)(w + 3 * slab);
  __bf16* axP = (__bf16*)(w + 4 * slab);
  __bf16* ayP = (__bf16*)(w + 5 * slab);
  __bf16* akP = (__bf16*)(w + 6 * slab);
  __bf16* yphiT = rhoP;  // rho's planes are read by the first GEMM only
  __bf16* yxT = tmpT;    // read by the second GEMM only
  __bf16* yyT = (__bf16*)a32;  // the DCT2 is read by the weights only
  double* part = (double*)(w + 7 * slab);
  double* msum = part + kReduceMaxBlocks;
  const __bf16* planes = (const __bf16*)((const char*)plan + plan_offset(M, N, 6, precision));
  auto P = [&](int id) { return planes + split_plane_offset(M, N, id); };
  const bool energy_phi = energy && want_phi;
  // rho = map / (bx*by), centred, as planes
  const int nb = (int)(ceil_div(n, kReduceBlock) < 1024 ? ceil_div(n, kReduceBlock) : 1024);
  hipLaunchKernelGGL(k_sum_partial<float>, dim3(nb), dim3(kReduceBlock), 0, s, map, n, part);
  DP_LAUNCHED("k_sum_partial");
  DP_CHECK_RC(finish_reduction(part, nb, 1, nullptr, msum, -1, s));
  hipLaunchKernelGGL(k_scale_centre_split, dim3(grid_for(n, 256)), dim3(256), 0, s, map, n,
                     (float)(1.0 / (bsx * bsy)), (const double*)msum, rhoP,
                     energy_phi ? rho32 : nullptr);
  DP_LAUNCHED("k_scale_centre_split");
  B3Batch bt = {};
  bt.alpha = 1.0f;
  // DCT2: tmp = rho * R_DCT (written transposed), a = L_DCT * tmp
  bt.M = M;
  bt.N = N;
  bt.K = N;
  bt.plane_a = n;
  bt.plane_b = (int64_t)N * N;
  bt.g[0] = {rhoP, P(R_DCT), nullptr, tmpT, nullptr, nullptr};
  DP_CHECK_RC(launch_b3(bt, 1, true, s, nullptr));
  bt.K = M;
  bt.plane_a = (int64_t)M * M;
  bt.plane_b = n;
  bt.g[0] = {P(L_DCT), tmpT, a32, nullptr, nullptr, nullptr};
  DP_CHECK_RC(launch_b3(bt, 1, false, s, nullptr));
  // spectral weights (and the spectral energy when there is no potential)
  const bool spectral_energy = energy && !want_phi;
  const int wblk = (int)(ceil_div(n, kReduceBlock) < kReduceMaxBlocks ? ceil_div(n, kReduceBlock)
                                                                      : kReduceMaxBlocks);
  hipLaunchKernelGGL(k_spectral_weights_split, dim3(wblk), dim3(kReduceBlock), 0, s,
                     (const float*)a32, M, N, (float)(2.0 * M_PI / M),
                     (float)(2.0 * M_PI / N * (bsx / bsy)), axP, ayP, want_phi ? akP : nullptr,
                     spectral_energy ? part : nullptr);
  DP_LAUNCHED("k_spectral_weights_split");
  // inverse transforms, first GEMMs (one batch, written transposed):
  //   Yx = ax*R_IDCT, Yy = ay*R_IDXST, Yphi = ak*R_IDCT
  bt.K = N;
  bt.plane_a = n;
  bt.plane_b = (int64_t)N * N;
  bt.g[0] = {axP, P(R_IDCT), nullptr, yxT, nullptr, nullptr};
  bt.g[1] = {ayP, P(R_IDXST), nullptr, yyT, nullptr, nullptr};
  int count = 2;
  if (want_phi) {
    bt.g[2] = {akP, P(R_IDCT), nullptr, yphiT, nullptr, nullptr};
    count = 3;
  }
  DP_CHECK_RC(launch_b3(bt, count, true, s, nullptr));
  // second GEMMs: field_x = L_IDXST*Yx, field_y = L_IDCT*Yy,
  // potential = L_IDCT*Yphi (+ energy partials sum(potential*rho))
  bt.K = M;
  bt.plane_a = (int64_t)M * M;
  bt.plane_b = n;
  bt.g[0] = {P(L_IDXST), yxT, field_x, nullptr, nullptr, nullptr};
  bt.g[1] = {P(L_IDCT), yyT, field_y, nullptr, nullptr, nullptr};
  if (want_phi)
    bt.g[2] = {P(L_IDCT), yphiT, potential, nullptr, energy_phi ? rho32 : nullptr, part};
  int nblk = 0;
  DP_CHECK_RC(launch_b3(bt, count, false, s, &nblk));
  if (energy)
    DP_CHECK_RC(finish_reduction(part, want_phi ? nblk : wblk, 1, nullptr, energy, DP_FLOAT32, s));
  return DP_OK;
}

static int transform(int precision, const void* plan, int M, int N, const void* x, void* out,
                     void* workspace, int lid, int rid, hipStream_t s) {
  GemmBatch bt = {};
  const char* P = (const char*)plan;
  char* tmp = (char*)workspace;
  // stage 1: tmp = x * R   (M x N) = (M x N)(N x N)
  bt.M = M;
  bt.N = N;
  bt.K = N;
  bt.alpha = 1.0;
  bt.g[0].A = x;
  bt.g[0].B = P + plan_offset(M, N, rid, precision);
  bt.g[0].C = tmp;
  DP_CHECK_RC(launch_gemm(precision, bt, 1, s, nullptr));
  // stage 2: out = L * tmp  (M x N) = (M x M)(M x N)
  bt.K = M;
  bt.g[0].A = P + plan_offset(M, N, lid, precision);
  bt.g[0].B = tmp;
  bt.g[0].C = out;
  return launch_gemm(precision, bt, 1, s, nullptr);
}

static int check_common(int precision, const void* plan, int M, int N, const void* x, void* out,
                        void* ws, const char* name) {
  DP_ARG(precision == DP_SPECTRAL_FP64 || precision == DP_SPECTRAL_FP32 ||
             precision == DP_SPECTRAL_BF16X3,
         "%s: bad precision %d", name, precision);
  DP_ARG(plan && x && out && ws, "%s: null pointer", name);
  DP_ARG(M > 0 && N > 0 && M <= 16384 && N <= 16384, "%s: bad size %d x %d", name, M, N);
  return DP_OK;
}

}  // namespace dp

using namespace dp;

extern "C" {

size_t dp_spectral_plan_bytes(int M, int N, int precision) {
  size_t b = plan_offset(M, N, 6, precision);
  if (precision == DP_SPECTRAL_BF16X3) b += split_plane_bytes(M, N);
  return b;
}

int dp_spectral_plan_init(int M, int N, int precision, void* plan, void* stream) {
  DP_ARG(plan && M > 0 && N > 0, "dp_spectral_plan_init: bad argument");
  hipStream_t s = S(stream);
  size_t total = 3 * (size_t)M * M + 3 * (size_t)N * N;
  int grid = grid_for((int64_t)total, 256);
  if (precision == DP_SPECTRAL_FP64)
    hipLaunchKernelGGL(k_plan_fill<double>, dim3(grid), dim3(256), 0, s, (double*)plan, M, N);
  else if (precision == DP_SPECTRAL_FP32 || precision == DP_SPECTRAL_BF16X3)
    hipLaunchKernelGGL(k_plan_fill<float>, dim3(grid), dim3(256), 0, s, (float*)plan, M, N);
  else
    DP_ARG(false, "dp_spectral_plan_init: bad precision %d", precision);
  DP_LAUNCHED("k_plan_fill");
  if (precision == DP_SPECTRAL_BF16X3) {
    hipLaunchKernelGGL(k_plan_split, dim3(grid), dim3(256), 0, s, (const float*)plan,
                       (__bf16*)((char*)plan + plan_offset(M, N, 6, precision)), M, N);
    DP_LAUNCHED("k_plan_split");
  }
  return DP_OK;
}

size_t dp_spectral_workspace_bytes(int M, int N, int precision) {
  return 7 * ws_slab(M, N, precision) + reduce_slot_bytes();
}

#define DP_TRANSFORM_ENTRY(NAME, LID, RID)                                                   \
  int NAME(int precision, const void* plan, int M, int N, const void* x, void* out,          \
           void* workspace, void* stream) {                                                  \
    DP_CHECK_RC(check_common(precision, plan, M, N, x, out, workspace, #NAME));              \
    return transform(precision, plan, M, N, x, out, workspace, LID, RID, S(stream));         \
  }

DP_TRANSFORM_ENTRY(dp_dct2, L_DCT, R_DCT)
DP_TRANSFORM_ENTRY(dp_idct2, L_IDCT, R_IDCT)
DP_TRANSFORM_ENTRY(dp_idxst_idct, L_IDXST, R_IDCT)
DP_TRANSFORM_ENTRY(dp_idct_idxst, L_IDCT, R_IDXST)

int dp_poisson_solve(int precision, const void* plan, int M, int N, const void* density_map,
                     double bin_size_x, double bin_size_y, int compute_potential, void* field_x,
                     void* field_y, void* potential, void* energy, void* workspace,
                     void* stream) {
  DP_CHECK_RC(check_common(precision, plan, M, N, density_map, field_x, workspace,
                           "dp_poisson_solve"));
  DP_ARG(field_y, "dp_poisson_solve: null field_y");
  DP_ARG(!compute_potential || potential, "dp_poisson_solve: compute_potential needs potential");
  hipStream_t s = S(stream);
  DP_PROBE_OP(DP_OP_POISSON, s);
  if (precision == DP_SPECTRAL_BF16X3 && presplit_shape(M, N))
    return poisson_presplit(plan, M, N, (const float*)density_map, bin_size_x, bin_size_y,
                            compute_potential != 0, (float*)field_x, (float*)field_y,
                            (float*)potential, (float*)energy, (char*)workspace, s);
  const size_t slab = ws_slab(M, N, precision);
  char* w = (char*)workspace;
  void* rho = w;               // map/(bx*by), kept for the energy
  void* tmp = w + slab;        // stage-1 products
  void* auv = w + 2 * slab;    // DCT2(rho), then Yy
  void* ax = w + 3 * slab;     // a*wu*K/2
  void* ay = w + 4 * slab;     // a*wv*K/2
  void* ak = w + 5 * slab;     // a*K
  void* yphi = w + 6 * slab;   // ak * R_IDCT
  double* part = (double*)(w + 7 * slab);
  const char* P = (const char*)plan;
  const int64_t n = (int64_t)M * N;
  const int dt = precision == DP_SPECTRAL_FP64 ? DP_FLOAT64 : DP_FLOAT32;
  const double aspect = bin_size_x / bin_size_y;
  const bool want_phi = compute_potential != 0;
  // rho = map / (bx*by), centred; a = DCT2(rho); spectral multipliers
  double* msum = part + kReduceMaxBlocks;  // second half of the reduction slot
  DP_CHECK_RC(DP_DISPATCH(dt, "dp_poisson_solve", [&] {
    const int nb = (int)(ceil_div(n, kReduceBlock) < 1024 ? ceil_div(n, kReduceBlock) : 1024);
    hipLaunchKernelGGL(k_sum_partial<T>, dim3(nb), dim3(kReduceBlock), 0, s,
                       (const T*)density_map, n, part);
    DP_LAUNCHED("k_sum_partial");
    DP_CHECK_RC(finish_reduction(part, nb, 1, nullptr, msum, -1, s));
    hipLaunchKernelGGL(k_scale_centre<T>, dim3(grid_for(n, 256)), dim3(256), 0, s,
                       (const T*)density_map, n, (T)(1.0 / (bin_size_x * bin_size_y)),
                       (const double*)msum, (T*)rho);
    DP_LAUNCHED("k_scale_centre");
    return DP_OK;
  }));
  DP_CHECK_RC(transform(precision, plan, M, N, rho, auv, tmp, L_DCT, R_DCT, s));
  int wblk = 0;
  DP_CHECK_RC(DP_DISPATCH(dt, "dp_poisson_solve", [&] {
    // without the potential map the energy is spectral, formed in this pass
    const bool spectral_energy = energy && !want_phi;
    wblk = (int)(ceil_div(n, kReduceBlock) < kReduceMaxBlocks ? ceil_div(n, kReduceBlock)
                                                                        : kReduceMaxBlocks);
    hipLaunchKernelGGL(k_spectral_weights<T>, dim3(wblk), dim3(kReduceBlock), 0, s,
                       (const T*)auv, M, N, (T)(2.0 * M_PI / M), (T)(2.0 * M_PI / N * aspect),
                       (T*)ax, (T*)ay, want_phi ? (T*)ak : nullptr,
                       spectral_energy ? part : nullptr);
    DP_LAUNCHED("k_spectral_weights");
    return DP_OK;
  }));
  // stage 1 of the inverse transforms (one batched launch):
  //   Yx = ax*R_IDCT -> tmp, Yy = ay*R_IDXST -> auv, Yphi = ak*R_IDCT -> yphi
  GemmBatch bt = {};
  bt.M = M;
  bt.N = N;
  bt.K = N;
  bt.alpha = 1.0;
  bt.g[0] = {ax, P + plan_offset(M, N, R_IDCT, precision), tmp, nullptr, nullptr};
  bt.g[1] = {ay, P + plan_offset(M, N, R_IDXST, precision), auv, nullptr, nullptr};
  int count = 2;
  if (want_phi) {
    bt.g[2] = {ak, P + plan_offset(M, N, R_IDCT, precision), yphi, nullptr, nullptr};
    count = 3;
  }
  DP_CHECK_RC(launch_gemm(precision, bt, count, s, nullptr));
  // stage 2: field_x = L_IDXST*Yx, field_y = L_IDCT*Yy, potential = L_IDCT*Yphi
  // (+ energy partials sum(potential*rho) in the epilogue)
  bt.K = M;
  bt.g[0] = {P + plan_offset(M, N, L_IDXST, precision), tmp, field_x, nullptr, nullptr};
  bt.g[1] = {P + plan_offset(M, N, L_IDCT, precision), auv, field_y, nullptr, nullptr};
  if (want_phi)
    bt.g[2] = {P + plan_offset(M, N, L_IDCT, precision), yphi, potential,
               energy ? rho : nullptr, part};
  int nblk = 0;
  DP_CHECK_RC(launch_gemm(precision, bt, count, s, &nblk));
  if (energy) {
    DP_CHECK_RC(finish_reduction(part, want_phi ? nblk : wblk, 1, nullptr, energy, dt, s));
  }
  return DP_OK;
}

}  // extern "C"
