/*
 * dreamplace_amd.h -- C ABI of the MI355X (gfx950) global-placement inner loop.
 *
 * Every entry point takes plain device pointers, sizes and a hipStream_t
 * (passed as void*), enqueues its kernels on that stream and returns
 * DP_OK (0) or an error code; it never synchronises the stream and never
 * allocates device memory.  Scratch memory comes from the caller as a
 * `workspace` pointer sized by the matching *_workspace_bytes() query.
 * Negative codes are argument errors; positive codes are hipError_t values
 * from the launch.  dp_last_error() returns a message for the calling
 * thread's last failure.
 *
 * Floating-point arrays are typed by `dtype` (DP_FLOAT32 or DP_FLOAT64,
 * mirroring DREAMPLACE_DISPATCH_FLOATING_TYPES in
 * dreamplace/ops/utility/src/torch.h:83-92, which raises for other types).
 * Positions use the reference layout: pos[2*num_nodes] holds all x then all
 * y, nodes ordered movable, fixed (terminals), filler.  Bin maps are
 * [num_bins_x][num_bins_y] row-major (x-bin major).
 *
 * Each declaration names the reference native op it replaces (file:line).
 */
#ifndef DREAMPLACE_AMD_H
#define DREAMPLACE_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DP_OK 0
#define DP_ERR_INVALID_ARGUMENT (-1)
#define DP_ERR_DTYPE (-2)
#define DP_ERR_UNSUPPORTED (-3)
#define DP_ERR_WORKSPACE (-4)

enum { DP_FLOAT32 = 0, DP_FLOAT64 = 1 };
enum { DP_INDEX_INT32 = 0, DP_INDEX_INT64 = 1 };
/* Conversion of the int64 fixed-point map back to floating point:
 * DP_CONVERT_CPU follows electric_density_map.cpp:125-140 (init + sum*2^-f in
 * double, rounded to T); DP_CONVERT_GPU follows
 * electric_density_map_cuda_kernel.cu:433-463 (init quantised first, then
 * T(u64) * T(2^-f)). */
enum { DP_CONVERT_CPU = 0, DP_CONVERT_GPU = 1 };
/* Arithmetic of the dense spectral transforms. */
enum { DP_SPECTRAL_FP64 = 0, DP_SPECTRAL_FP32 = 1, DP_SPECTRAL_BF16X3 = 2 };
/* Density scatter algorithm: per-cell global int64 atomics, or cells binned
 * into LDS-resident bin tiles first. */
enum { DP_SCATTER_GLOBAL = 0, DP_SCATTER_TILED = 1 };
/* Overlap arithmetic (no counterpart in the reference), a flag OR'ed into
 * the `algorithm` argument of dp_density_map_workspace_bytes,
 * dp_density_map, dp_density_map_ex and dp_electric_force, and into the
 * `convention` argument of dp_fixed_density_map.  Without it every overlap
 * is the reference's min(x + s, bin_xh) - max(x, bin_xl) in T
 * (density_function.h:12-36), whose x + s rounds at the order of the die
 * coordinates.  With it each cell's stretched origin x + offset is carried
 * with its rounding residual (TwoSum) and measured from the lower edge of
 * the cell's first bin, so every overlap is formed from operands of the
 * order of a few bins: the same bins and the same fixed-point recipe, the
 * FP32 map and force within ~1e-7 of FP64 instead of ~2e-5.  Maps and
 * gradients are bit-identical across the global and tiled paths in either
 * arithmetic; only the reference arithmetic is bit-exact with the
 * reference. */
enum { DP_ARITH_BIN_RELATIVE = 0x100 };
/* Layout of a per-node gradient: DP_LAYOUT_SPLIT is the reference's [2N]
 * (all x, then all y); DP_LAYOUT_INTERLEAVED stores (x, y) of node i at
 * [2i, 2i+1], so one node is one 8- (FP32) or 16-byte (FP64) word. */
enum { DP_LAYOUT_SPLIT = 0, DP_LAYOUT_INTERLEAVED = 1 };
/* Flag OR-ed into the add_layout of dp_node_grad_from_pins and
 * dp_node_grad_preconditioned: the fillers' entries of grad_pos_out already
 * hold their finished gradient, as dp_electric_force_fillers_final wrote them
 * (with the same density_weight, alpha and filler_area when preconditioned);
 * only the other nodes are computed (and the fillers' entries read back for
 * dg_norm2_out), and add's filler entries are not read. */
enum { DP_LAYOUT_FILLERS_FINAL = 0x10 };

const char* dp_last_error(void);
int dp_version(void);

/* Kernel probe (no counterpart in the reference): when enabled, the library
 * records `start_event` immediately before and `stop_event` immediately after
 * every launch of the named kernel (DP_KERNEL_*), or around all launches of
 * one call of the named entry point (DP_OP_*), on the stream they are
 * launched on (a launch overwrites the previous pair).  kernel_id < 0
 * disables.  Events are hipEvent_t passed as void*.  Used by the benchmark to
 * time one kernel or op live. */
enum {
  DP_KERNEL_DENSITY_SCATTER = 0,  /* k_tile_scatter / k_tri_scatter */
  DP_KERNEL_ELECTRIC_FORCE = 1,   /* k_tile_gather / k_force_small */
  DP_KERNEL_WA_FORWARD = 2,       /* k_wa_pins / k_wa_fwd_nodes / k_wa_fwd */
  DP_KERNEL_WA_BACKWARD = 3,
  DP_KERNEL_PIN_POS_FORWARD = 4,
  DP_KERNEL_PIN_POS_BACKWARD = 5, /* k_node_grad / k_wa_pin_pos_bwd / k_pin_pos_bwd */
  DP_KERNEL_SPECTRAL_GEMM = 6,
  DP_KERNEL_DENSITY_CONVERT = 7,
  DP_KERNEL_DENSITY_BIN = 8,      /* k_bin_fill */
  DP_KERNEL_PIN_PUSH = 9,         /* k_pin_push */
  DP_OP_DENSITY_MAP = 16,         /* dp_density_map */
  DP_OP_ELECTRIC_FORCE = 17,      /* dp_electric_force */
  DP_OP_WA_FORWARD = 18,          /* dp_wa_forward_pins / dp_wa_forward_from_nodes / dp_wa_merged_forward */
  DP_OP_NODE_GRAD = 19,           /* dp_node_grad_from_pins / dp_wa_backward_to_nodes */
  DP_OP_POISSON = 20,             /* dp_poisson_solve */
  DP_KERNEL_COUNT = 32
};
int dp_profile_kernel(int kernel_id, void* start_event, void* stop_event);

/* ---------------------------------------------------------------------------
 * Density map (fixed-point int64 accumulation; always deterministic).
 * ------------------------------------------------------------------------- */

/* Fixed-point scale 2^(64 - max(ceil(log2(area))+1, 32)) for a die
 * (electric_density_map.cpp:117-123). */
int64_t dp_fixed_point_scale(double xl, double yl, double xh, double yh);

/* Bytes of the workspace shared by dp_density_map, dp_fixed_density_map and
 * dp_electric_force.  With DP_SCATTER_TILED it includes one packed record per
 * movable/filler cell (the binning of the last dp_density_map call). */
size_t dp_density_map_workspace_bytes(int64_t num_bins_x, int64_t num_bins_y, int64_t num_nodes,
                                      int64_t large_capacity, int algorithm, int dtype);

/* Replaces electric_potential_cuda.density_map
 * (dreamplace/ops/electric_potential/src/electric_density_map_cuda.cpp:71-136;
 * CPU twin electric_density_map.cpp:100-152).  Scatters movable cells
 * [0, num_movable_nodes) and fillers [num_nodes-num_filler_nodes, num_nodes)
 * with the triangle (stretched-cell) model into an int64 fixed-point map and
 * writes density_map_out = initial_density_map + map (per `convention`).
 * `fixed_map_out` (nullable) receives the raw int64 sum (no initial map).
 * `sorted_node_map` is accepted for signature parity and not needed: the
 * integer sum is order-free.  large_capacity bounds the queue of cells
 * spanning more than 3x3 bins (any count is correct; a larger queue only
 * moves more of them off the per-thread path).  uniform_filler (nullable,
 * host memory): {clamped size x, clamped size y, offset x, offset y, ratio}
 * shared by every filler node, which PlaceDB's fillers are
 * (dreamplace/PlaceDB.py:917-936); the filler entries of the five arrays are
 * then not read.  The result is the same either way. */
int dp_density_map(int dtype, const void* pos, int64_t num_nodes, const void* node_size_x_clamped,
                   const void* node_size_y_clamped, const void* offset_x, const void* offset_y,
                   const void* ratio, const void* initial_density_map, double xl, double yl,
                   double xh, double yh, double bin_size_x, double bin_size_y,
                   int64_t num_movable_nodes, int64_t num_filler_nodes, int num_bins_x,
                   int num_bins_y, int convention, int algorithm, int64_t large_capacity,
                   const double* uniform_filler, void* workspace, void* density_map_out,
                   int64_t* fixed_map_out, void* stream);

/* New: dp_density_map with two options for a placement loop, both nullable.
 * movable_map_out[num_bins] (`dtype`): the map of the fixed and movable
 * cells alone, converted as density_map_out is: the map ElectricOverflow's
 * density_map forms with num_filler_nodes = 0 (electric_overflow.py:260-278,
 * the second scatter NonLinearPlace.py:414 runs for the overflow metric),
 * bit-identical, from this one scatter (the fillers go to an int64 map of
 * their own and are added for density_map_out).
 * filler_tile_start (device int32[ntiles + 1], ntiles = ceil(M/64) *
 * ceil(N/64), tile t = (bx/64) * ceil(N/64) + by/64 of a cell's lower-left
 * bin): the fillers are kept in tile order and [start[t], start[t+1]) are
 * the (filler-relative) fillers of tile t.  They are then scattered, and
 * gathered by the next dp_electric_force on this workspace, from their
 * positions by tile instead of being binned.  The maps and gradients are
 * bit-identical to the binned path for any ranges: a filler outside its
 * tile's LDS window (moved since it was sorted) is done through global
 * memory, and ranges that are not a monotone partition of the fillers are
 * replaced by one range; only speed depends on how well they fit.
 * Both need DP_SCATTER_TILED (filler_tile_start is ignored otherwise) and
 * fillers of one size (uniform_filler) or none; filler_tile_start also needs
 * their stretched size at most 2 bins.  Else DP_ERR_UNSUPPORTED. */
int dp_density_map_ex(int dtype, const void* pos, int64_t num_nodes,
                      const void* node_size_x_clamped, const void* node_size_y_clamped,
                      const void* offset_x, const void* offset_y, const void* ratio,
                      const void* initial_density_map, double xl, double yl, double xh,
                      double yh, double bin_size_x, double bin_size_y,
                      int64_t num_movable_nodes, int64_t num_filler_nodes, int num_bins_x,
                      int num_bins_y, int convention, int algorithm, int64_t large_capacity,
                      const double* uniform_filler, const int32_t* filler_tile_start,
                      void* workspace, void* density_map_out, void* movable_map_out,
                      void* stream);

/* No counterpart in the reference.  For callers that cannot allocate (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc): dp_density_map with algorithm DP_SCATTER_GLOBAL and no
 * workspace argument.  The int64 map and the large-cell queue live in a
 * file-scope device array of the library (8.25 MiB, one copy per device,
 * held until the process ends); the call allocates nothing.  Limit: at most
 * 1024 x 1024 bins (DP_ERR_UNSUPPORTED beyond).  Every call on a device
 * shares that array, so calls that overlap on one device (two streams, two
 * threads) race.  Same arguments and results as dp_density_map otherwise
 * (bit-identical map and fixed-point sum). */
int dp_density_map_noalloc(int dtype, const void* pos, int64_t num_nodes,
                           const void* node_size_x_clamped, const void* node_size_y_clamped,
                           const void* offset_x, const void* offset_y, const void* ratio,
                           const void* initial_density_map, double xl, double yl, double xh,
                           double yh, double bin_size_x, double bin_size_y,
                           int64_t num_movable_nodes, int64_t num_filler_nodes, int num_bins_x,
                           int num_bins_y, int convention, const double* uniform_filler,
                           void* density_map_out, int64_t* fixed_map_out, void* stream);

/* Replaces electric_potential_cuda.fixed_density_map
 * (electric_density_map_cuda.cpp:139-182, electric_density_map.cpp:155-213)
 * followed by the Python-side scaling of electric_overflow.py:258:
 * out = min(exact overlap map of terminals, bin area) * target_density.
 * Terminals are nodes [num_movable_nodes, num_movable_nodes+num_terminals). */
int dp_fixed_density_map(int dtype, const void* pos, int64_t num_nodes, const void* node_size_x,
                         const void* node_size_y, const void* bin_center_x,
                         const void* bin_center_y, double xl, double yl, double xh, double yh,
                         double bin_size_x, double bin_size_y, int64_t num_movable_nodes,
                         int64_t num_terminals, int num_bins_x, int num_bins_y,
                         double target_density, int convention, void* workspace, void* out,
                         void* stream);

/* Replaces -electric_potential_cuda.electric_force(...) as used by
 * ElectricPotentialFunction.backward (electric_potential.py:219-245;
 * electric_force_cuda.cpp:56-126, electric_force.cpp:76-172).
 * grad_out[2N] = -(ratio_i * sum_bins px*py*E) * (*grad_scale) for movables and
 * fillers, 0 for terminals.  grad_scale is a device scalar of `dtype`
 * (nullable: 1).  workspace as for dp_density_map (same large_capacity).
 * With DP_SCATTER_TILED the gather reuses the cell binning that the last
 * dp_density_map call (DP_SCATTER_TILED, same workspace) made: `pos` must be
 * the positions of that call.  grad_layout: DP_LAYOUT_SPLIT (the reference's
 * layout) or DP_LAYOUT_INTERLEAVED (one word per node, as
 * dp_node_grad_from_pins reads it).  uniform_filler as for dp_density_map. */
int dp_electric_force(int dtype, const void* grad_scale, const void* field_map_x,
                      const void* field_map_y, const void* pos, int64_t num_nodes,
                      const void* node_size_x_clamped, const void* node_size_y_clamped,
                      const void* offset_x, const void* offset_y, const void* ratio, double xl,
                      double yl, double bin_size_x, double bin_size_y, int64_t num_movable_nodes,
                      int64_t num_filler_nodes, int num_bins_x, int num_bins_y, int algorithm,
                      int64_t large_capacity, int grad_layout, const double* uniform_filler,
                      void* workspace, void* grad_out, void* stream);

/* No counterpart in the reference: dp_electric_force with the gradient
 * combine of the fillers folded in, for a placement loop that combines the
 * two gradients on the device (PlaceObj.obj_and_grad, PlaceObj.py:394-411,
 * then PreconditionOp, PlaceObj.py:69-102).  Fillers have no pins, so their
 * objective gradient is their density gradient alone: for every filler i
 * (i >= num_nodes - num_filler_nodes) grad_pos_out (split [2N], `dtype`)
 * receives 0 + g_i, with g_i what dp_electric_force would write to grad_out,
 * divided, when density_weight (device scalar) is given, by
 * max(alpha * (*density_weight) * filler_area, 1) (filler_area >= 0, the
 * fillers' one area).  These are the bits dp_node_grad_from_pins /
 * dp_node_grad_preconditioned give a filler; pass DP_LAYOUT_FILLERS_FINAL to
 * them so that they skip the fillers.  Every other node's density gradient
 * goes to grad_out as dp_electric_force writes it; grad_out's filler entries
 * are left unwritten.  Tiled gather only (DP_ERR_UNSUPPORTED otherwise). */
int dp_electric_force_fillers_final(
    int dtype, const void* grad_scale, const void* field_map_x, const void* field_map_y,
    const void* pos, int64_t num_nodes, const void* node_size_x_clamped,
    const void* node_size_y_clamped, const void* offset_x, const void* offset_y,
    const void* ratio, double xl, double yl, double bin_size_x, double bin_size_y,
    int64_t num_movable_nodes, int64_t num_filler_nodes, int num_bins_x, int num_bins_y,
    int algorithm, int64_t large_capacity, int grad_layout, const double* uniform_filler,
    void* workspace, void* grad_out, void* grad_pos_out, const void* density_weight, double alpha,
    double filler_area, void* stream);

/* The exact gradient of the density energy E = sum(phi * rho) that
 * ElectricPotentialFunction.forward returns (electric_potential.py:137-186;
 * rho = map / (bin_size_x * bin_size_y), phi = IDCT2(DCT2(rho) * K)).  No
 * counterpart in the reference, whose backward is the field gather above
 * (electric_potential.py:219-245): this is its exact adjoint, offered beside it.
 * The solve is symmetric in rho, so dE/dmap_b = 2 * phi_b / (bin_size_x *
 * bin_size_y); through each cell's triangle overlaps over the scatter's bin
 * range (density_function.h:12-17, electric_density_map.cpp:274-293):
 *   grad_out_x[i] = (*grad_scale) * 2/(bin_size_x*bin_size_y) * ratio_i *
 *                   sum_{k,h} dpx_k * py_h * phi[k,h]
 * and y with px_k * dpy_h, where dpx_k = [x+sx < bin_k right] - [x >= bin_k
 * left] is the overlap's derivative from the right.  0 for terminals.
 * potential_map is phi ([num_bins_x * num_bins_y], `dtype`, as dp_poisson_solve
 * writes it); grad_scale nullable (1); grad_layout and uniform_filler as for
 * dp_electric_force.  Needs no workspace. */
int dp_energy_gradient(int dtype, const void* grad_scale, const void* potential_map,
                       const void* pos, int64_t num_nodes, const void* node_size_x_clamped,
                       const void* node_size_y_clamped, const void* offset_x, const void* offset_y,
                       const void* ratio, double xl, double yl, double bin_size_x,
                       double bin_size_y, int64_t num_movable_nodes, int64_t num_filler_nodes,
                       int num_bins_x, int num_bins_y, int grad_layout,
                       const double* uniform_filler, void* grad_out, void* stream);

/* ElectricOverflow metric (electric_overflow.py:260-278): out2[0] =
 * sum(max(map - target_area, 0)), out2[1] = max(map); FP64 accumulation,
 * deterministic.  out2 is a device double[2]. */
size_t dp_reduce_workspace_bytes(int64_t n);
int dp_density_overflow(int dtype, const void* density_map, int64_t num_bins, double target_area,
                        void* workspace, double* out2, void* stream);

/* ---------------------------------------------------------------------------
 * Spectral transforms (dense, on the matrix cores) and the Poisson solve.
 * Data arrays are double for DP_SPECTRAL_FP64 and float otherwise.
 * ------------------------------------------------------------------------- */

/* A plan holds the cosine/sine transform matrices for an M x N grid. */
size_t dp_spectral_plan_bytes(int M, int N, int precision);
int dp_spectral_plan_init(int M, int N, int precision, void* plan, void* stream);
size_t dp_spectral_workspace_bytes(int M, int N, int precision);

/* Replace dct2_fft2_cuda.{dct2_fft2, idct2_fft2, idct_idxst, idxst_idct}
 * (dreamplace/ops/dct/src/dct2_fft2_cuda.cpp:14-141); conventions of
 * dreamplace/ops/dct/dct2_fft2.py:21-150:
 *   DCT2       out[u,v] = 4/(MN) sum x[i,j] cos(pi u(2i+1)/2M) cos(pi v(2j+1)/2N)
 *   IDCT2      out[m,n] = sum c_u c_v x[u,v] cos(.. m ..) cos(.. n ..), c_0=1, c_k=2
 *   IDXST_IDCT out[m,n] = 2 sum c_v x[u,v] sin(pi u(2m+1)/2M) cos(pi v(2n+1)/2N)
 *   IDCT_IDXST out[m,n] = 2 sum c_u x[u,v] cos(pi u(2m+1)/2M) sin(pi v(2n+1)/2N) */
int dp_dct2(int precision, const void* plan, int M, int N, const void* x, void* out,
            void* workspace, void* stream);
int dp_idct2(int precision, const void* plan, int M, int N, const void* x, void* out,
             void* workspace, void* stream);
int dp_idxst_idct(int precision, const void* plan, int M, int N, const void* x, void* out,
                  void* workspace, void* stream);
int dp_idct_idxst(int precision, const void* plan, int M, int N, const void* x, void* out,
                  void* workspace, void* stream);

/* The solve of ElectricPotentialFunction.forward (electric_potential.py:137-186):
 * rho = density_map/(bx*by); a = DCT2(rho); field_x = IDXST_IDCT(a*wu/(2(wu^2+wv^2)));
 * field_y = IDCT_IDXST(a*wv/(2(wu^2+wv^2))); optionally potential = IDCT2(a/(wu^2+wv^2))
 * and *energy = sum(potential*rho) (device scalar of the data type; nullable).
 * With compute_potential = 0 and energy non-null, the energy is evaluated in
 * the spectral domain, (MN/4) sum c_u c_v a^2/(wu^2+wv^2), which equals
 * sum(potential*rho) exactly in exact arithmetic. */
int dp_poisson_solve(int precision, const void* plan, int M, int N, const void* density_map,
                     double bin_size_x, double bin_size_y, int compute_potential, void* field_x,
                     void* field_y, void* potential, void* energy, void* workspace, void* stream);

/* ---------------------------------------------------------------------------
 * Pin positions and weighted-average wirelength.
 * ------------------------------------------------------------------------- */

/* Replaces pin_pos_cuda.forward (dreamplace/ops/pin_pos/src/pin_pos_cuda.cpp:27-54,
 * kernel pin_pos_cuda_kernel.cu:11-26): pin_pos[p] = offset[p] + pos[pin2node[p]]. */
int dp_pin_pos_forward(int dtype, const void* pos, int64_t num_nodes, const void* pin_offset_x,
                       const void* pin_offset_y, const void* pin2node_map, int index_type,
                       int64_t num_pins, void* pin_pos_out, void* stream);

/* Replaces pin_pos_cuda.backward (pin_pos_cuda.cpp:57-93, kernel
 * pin_pos_cuda_kernel.cu:50-100): grad_pos[n] = sum of grad_pin_pos over the
 * CSR pins of physical node n; fillers get 0. */
int dp_pin_pos_backward(int dtype, const void* grad_pin_pos, int64_t num_pins,
                        const int32_t* flat_node2pin_map, const int32_t* flat_node2pin_start_map,
                        int64_t num_physical_nodes, int64_t num_nodes, void* grad_pos_out,
                        void* stream);

size_t dp_wa_workspace_bytes(int64_t num_nets);

/* Replaces weighted_average_wirelength_cuda_merged.forward
 * (weighted_average_wirelength_cuda_merged.cpp:63-110, kernel
 * weighted_average_wirelength_cuda_merged_kernel.cu:16-85).  Writes the
 * pre-weight per-net WL (partial_wl, nullable), the per-pin derivative of the
 * net WL (grad_intermediate[2P], entries of masked nets untouched) and
 * *wl_out = sum_n w_n*WL_n (device scalar of dtype, FP64 accumulation,
 * deterministic).  net_weights nullable (all 1).  inv_gamma is a device scalar.
 * grad_intermediate may be null: then nothing per pin is stored, the forward
 * of the recompute form (dp_wa_merged_backward_recompute). */
int dp_wa_merged_forward(int dtype, const void* pin_pos, int64_t num_pins,
                         const int32_t* flat_netpin, const int32_t* netpin_start,
                         int64_t num_nets, const void* net_weights, const uint8_t* net_mask,
                         const void* inv_gamma, void* partial_wl, void* grad_intermediate,
                         void* wl_out, void* workspace, void* stream);

/* Replaces logsumexp_wirelength_cuda_merged.forward
 * (logsumexp_wirelength_cuda_merged.cpp; CPU twin
 * logsumexp_wirelength_merged.cpp:99-146): per unmasked net and axis
 * WL = (log S+ + log S-)*gamma + max - min, S+- the max-shifted exponential
 * sums; grad_intermediate[2P] = e+/S+ - e-/S- per pin (untouched for masked
 * nets: the caller zeroes it, as the reference allocates it with zeros).
 * `gamma` is a device scalar of `dtype`.  The backward
 * (logsumexp_wirelength_merged.cpp:160-205) is dp_wa_merged_backward: the
 * same grad, net-weight and pin-mask step on grad_intermediate. */
int dp_lse_merged_forward(int dtype, const void* pin_pos, int64_t num_pins,
                          const int32_t* flat_netpin, const int32_t* netpin_start,
                          int64_t num_nets, const void* net_weights, const uint8_t* net_mask,
                          const void* gamma, void* partial_wl, void* grad_intermediate,
                          void* wl_out, void* workspace, void* stream);

/* Replaces weighted_average_wirelength_cuda_merged.backward
 * (weighted_average_wirelength_cuda_merged.cpp:121-160,
 * integrate_net_weights_cuda_kernel.cu:17-48) plus the pin_mask fill of
 * weighted_average_wirelength.py:207-208, in place on grad_intermediate:
 * g[p] = gi[p]*(*grad)*w[net(p)] (weight applied for unmasked nets), 0 where
 * pin_mask[p] != 0. */
int dp_wa_merged_backward(int dtype, const void* grad, void* grad_intermediate, int64_t num_pins,
                          const int32_t* pin2net_map, const void* net_weights,
                          const uint8_t* net_mask, const uint8_t* pin_mask, void* stream);

/* The backward of the recompute form: the reference keeps the per-pin
 * derivative from the forward (weighted_average_wirelength.py:189, the store
 * form above; its net-by-net function keeps the exponential terms themselves,
 * :59-64); this entry point recomputes it from pin_pos instead,
 * one thread per net (weighted_average_wirelength_merged.cpp:49-116), and
 * writes grad_pin_pos[2P] = (dWL_n/dx_p * (*grad)) * w_n, 0 where
 * pin_mask[p] != 0: the values dp_wa_merged_backward leaves in
 * grad_intermediate.  Pins of masked nets are not written (the caller zeroes
 * grad_pin_pos).  Saves the [2P] derivative held between forward and
 * backward, for one more pass over the pins. */
int dp_wa_merged_backward_recompute(int dtype, const void* grad, const void* pin_pos,
                                    int64_t num_pins, const int32_t* flat_netpin,
                                    const int32_t* netpin_start, int64_t num_nets,
                                    const void* net_weights, const uint8_t* net_mask,
                                    const uint8_t* pin_mask, const void* inv_gamma,
                                    void* grad_pin_pos, void* stream);

/* New (no counterpart in the reference, which returns None for the weights,
 * weighted_average_wirelength.py:213): d(sum_n w_n WL_n)/d(w_n) times *grad,
 * = (*grad)*partial_wl[n] for unmasked nets, 0 otherwise. */
int dp_wa_net_weight_grad(int dtype, const void* grad, const void* partial_wl,
                          const uint8_t* net_mask, int64_t num_nets, void* grad_net_weights,
                          void* stream);

/* Timing-driven net weighting: replaces the weighting step of
 * timing_cpp.update_net_weights (timing.py:182-206 ->
 * net_weighting_scheme.h:106-190) on the device, given what a static timer
 * reports (the timer itself, OpenTimer / HeteroSTA in the reference, is not
 * part of this library).  scheme DP_NET_WEIGHTING_ADAMS (the reference's 0):
 * net_critical[e] != 0 marks nets on the reported critical paths;
 * DP_NET_WEIGHTING_LILITH (1): net_slack[e] is the net's worst sink slack and
 * wns the design's worst negative slack.  net_criticality and net_weights
 * (`dtype`, per net) are updated in place; nets of degree
 * (netpin_start[e+1] - netpin_start[e]) above ignore_net_degree keep their
 * weight.  max_net_weight caps LILITH's weights (+inf: no cap); decay is the
 * momentum decay factor (params.json: 0.5). */
enum { DP_NET_WEIGHTING_ADAMS = 0, DP_NET_WEIGHTING_LILITH = 1 };
int dp_net_weighting(int dtype, int scheme, int64_t num_nets, const void* net_slack,
                     const uint8_t* net_critical, double wns, const int32_t* netpin_start,
                     int ignore_net_degree, double decay, double max_net_weight,
                     void* net_criticality, void* net_weights, void* stream);

/* Fused forward of the wirelength op chain WA(PinPos(pos))
 * (dreamplace/PlaceObj.py:478-479): dp_pin_pos_forward followed by
 * dp_wa_merged_forward, with each pin position formed on the fly from
 * pos[pin2node[p]] + offset[p] instead of being written to memory.  Outputs as
 * for dp_wa_merged_forward. */
int dp_wa_forward_from_nodes(int dtype, const void* pos, int64_t num_nodes,
                             const void* pin_offset_x, const void* pin_offset_y,
                             const int32_t* pin2node_map, int64_t num_pins,
                             const int32_t* flat_netpin, const int32_t* netpin_start,
                             int64_t num_nets, const void* net_weights, const uint8_t* net_mask,
                             const void* inv_gamma, void* partial_wl, void* grad_intermediate,
                             void* wl_out, void* workspace, void* stream);

/* Fused backward of that chain plus the gradient combine of
 * PlaceObj.obj_and_grad_fn (dreamplace/PlaceObj.py:394-411): per pin the
 * arithmetic of dp_wa_merged_backward (not in place), summed over each
 * physical node's CSR pins as dp_pin_pos_backward does, plus add[2N]
 * (nullable; e.g. the density gradient).  grad_intermediate is read only. */
int dp_wa_backward_to_nodes(int dtype, const void* grad, const void* grad_intermediate,
                            int64_t num_pins, const int32_t* flat_node2pin_map,
                            const int32_t* flat_node2pin_start_map, const int32_t* pin2net_map,
                            const void* net_weights, const uint8_t* net_mask,
                            const uint8_t* pin_mask, const void* add, int64_t num_physical_nodes,
                            int64_t num_nodes, void* grad_pos_out, void* stream);

/* Pin-parallel form of the fused wirelength chain (no reference counterpart:
 * the reference runs one thread per net, weighted_average_wirelength_cuda_
 * merged_kernel.cu:16-85).  Host routine, plain host memory, no GPU work:
 * packs the nets of 1..64 pins best-fit-decreasing by degree into groups of
 * at most 64 pins and lays their pins out in "group order": group g owns
 * slots [group_start[g], group_start[g+1]), a net's pins are contiguous in
 * it, in flat_netpin order.  Nets of more than 64 pins follow, one per
 * big_nets entry {net, first slot, degree, 0}.  net_slot[e] = first slot of
 * net e.  Capacities: group_start num_nets+1, big_nets 4*num_nets, net_slot
 * num_nets. */
int dp_wa_pack_nets(const int32_t* netpin_start, int64_t num_nets, int32_t* group_start,
                    int64_t* num_groups, int32_t* big_nets, int64_t* num_big, int32_t* net_slot);

/* Static preparation of the pin-parallel form for a netlist (run once; the
 * pin offsets, maps and pin mask are constant during placement).  Writes
 *   node_offsets[2P]  (offset_x, offset_y) of the pin in node-order slot i
 *                     (slot i of flat_node2pin_map), interleaved;
 *   node_dst[P]       the pin's slot in group order;
 *   slot_node[P]      the physical node owning node-order slot i;
 *   pin_meta[4P]      per group-order slot s: {net | sign bit when pin_mask
 *                     marks the pin (pin_mask nullable), node-order slot,
 *                     rank in the net | degree << 8 (nets of <= 64 pins), 0}.
 * flat_netpin and flat_node2pin_map must both be permutations of
 * [0, num_pins) and flat_node2pin_start_map must cover them over the physical
 * nodes, as PlaceDB builds them.  workspace:
 * dp_wa_prepare_workspace_bytes(num_pins). */
size_t dp_wa_prepare_workspace_bytes(int64_t num_pins);
int dp_wa_pins_prepare(int dtype, int64_t num_pins, int64_t num_physical_nodes,
                       const void* pin_offset_x, const void* pin_offset_y,
                       const int32_t* flat_netpin, const int32_t* netpin_start,
                       const int32_t* pin2net_map, const uint8_t* pin_mask,
                       const int32_t* flat_node2pin_map, const int32_t* flat_node2pin_start_map,
                       const int32_t* net_slot, void* workspace, void* node_offsets,
                       int32_t* node_dst, int32_t* slot_node, int32_t* pin_meta, void* stream);

size_t dp_wa_pins_workspace_bytes(int64_t num_pins, int dtype);

/* The chain WA(PinPos(pos)) with one wavefront per net group and one lane per
 * pin.  Each pin slot in node order first pushes pos + offset to the pin's
 * group-order slot; the WA kernel then reads pins contiguously.  Outputs
 * partial_wl[num_nets] (nullable; 0 for masked nets) and *wl_out as
 * dp_wa_merged_forward, and grad_node_order[2P]: each pin's derivative
 * (x, y interleaved), times its net weight, 0 for masked nets and pins,
 * stored at the pin's node-order slot.  net_hpwl[num_nets] (nullable, new):
 * every net's HPWL (max - min per axis, unweighted, in hpwl_cuda_kernel.cu's
 * expression) from the same pass, masked or not, for dp_hpwl_from_nets. */
int dp_wa_forward_pins(int dtype, const void* pos, int64_t num_nodes, int64_t num_pins,
                       const int32_t* slot_node, const void* node_offsets,
                       const int32_t* node_dst, const int32_t* pin_meta,
                       const int32_t* group_start, int64_t num_groups, const int32_t* big_nets,
                       int64_t num_big, int64_t num_nets, const void* net_weights,
                       const uint8_t* net_mask, const void* slot_weights, const void* inv_gamma,
                       void* partial_wl, void* net_hpwl, void* grad_node_order, void* wl_out,
                       void* workspace, void* stream);

/* dp_wa_forward_pins with the same model evaluated relative to each net's
 * extremes (no counterpart in the reference, which evaluates
 * weighted_average_wirelength_merged.cpp:85-111's (a + b x) e form): with
 * u+ = x - x_max, u- = x - x_min, S = sum e, C = sum u e / S per net and
 * axis, WL = (x_max - x_min) + C+ - C- and
 * dWL/dx_p = (e+_p/S+)(1 + (u+_p - C+)/g) - (e-_p/S-)(1 - (u-_p - C-)/g).
 * The exponentials are the reference's bits; the value and the derivatives
 * are the same function without the reference form's cancellation of
 * x/gamma-sized terms, so FP32 keeps about 1e-7 relative instead of ~2e-5.
 * Same arguments, plan and outputs as dp_wa_forward_pins. */
int dp_wa_forward_pins_centred(int dtype, const void* pos, int64_t num_nodes, int64_t num_pins,
                               const int32_t* slot_node, const void* node_offsets,
                               const int32_t* node_dst, const int32_t* pin_meta,
                               const int32_t* group_start, int64_t num_groups,
                               const int32_t* big_nets, int64_t num_big, int64_t num_nets,
                               const void* net_weights, const uint8_t* net_mask,
                               const void* slot_weights, const void* inv_gamma, void* partial_wl,
                               void* net_hpwl, void* grad_node_order, void* wl_out,
                               void* workspace, void* stream);

/* The log-sum-exp model on the same plan and skeleton: the chain
 * LSE(PinPos(pos)) of logsumexp_wirelength_merged.cpp:14-85 (CUDA twin
 * logsumexp_wirelength_cuda_merged_kernel.cu), with `gamma` (device scalar of
 * `dtype`, not its inverse) as LogSumExpWirelength passes it.  Outputs as
 * dp_wa_forward_pins; dp_node_grad_from_pins finishes the backward. */
int dp_lse_forward_pins(int dtype, const void* pos, int64_t num_nodes, int64_t num_pins,
                        const int32_t* slot_node, const void* node_offsets,
                        const int32_t* node_dst, const int32_t* pin_meta,
                        const int32_t* group_start, int64_t num_groups, const int32_t* big_nets,
                        int64_t num_big, int64_t num_nets, const void* net_weights,
                        const uint8_t* net_mask, const void* slot_weights, const void* gamma,
                        void* partial_wl, void* net_hpwl, void* grad_node_order, void* wl_out,
                        void* workspace, void* stream);

/* Optional companion of the two entry points above, with no counterpart in
 * the reference: for callers whose net weights and net mask stay fixed
 * between calls, writes each group-order pin slot's net weight to
 * slot_weights[num_pins] (`dtype`; 1 where net_weights is null) and its
 * net's mask into pin_meta.  Passing slot_weights to dp_wa_forward_pins /
 * dp_lse_forward_pins then replaces the per-net gathers of net_weights and
 * net_mask by one contiguous read; run it again after either changes. */
int dp_wa_pins_slot_weights(int dtype, int64_t num_pins, int32_t* pin_meta,
                            const void* net_weights, const uint8_t* net_mask, void* slot_weights,
                            void* stream);

/* Backward of that chain plus the gradient combine: grad_pos[n] =
 * (*grad) * (sum of grad_node_order over node n's CSR range) + add[n]
 * (add nullable, in add_layout: DP_LAYOUT_SPLIT or DP_LAYOUT_INTERLEAVED,
 * optionally with DP_LAYOUT_FILLERS_FINAL), read contiguously; fillers get
 * add only.  grad_pos_out is split. */
int dp_node_grad_from_pins(int dtype, const void* grad, const void* grad_node_order,
                           const int32_t* flat_node2pin_start_map, const void* add,
                           int add_layout, int64_t num_physical_nodes, int64_t num_nodes,
                           void* grad_pos_out, void* stream);

/* dp_node_grad_from_pins followed by PreconditionOp.__call__ without fence
 * regions (PlaceObj.py:69-102) in the same pass: each node's finished
 * gradient is divided by max(pin_weight_sum[n] + (alpha * *density_weight) *
 * node_areas[n], 1), as dp_precondition does; density_weight is a device
 * scalar of `dtype`.  dg_norm2_out (device double, nullable; then grad_prev,
 * a [2N] gradient not aliasing the output, and a dp_optimizer_workspace_bytes
 * workspace are needed): sum((grad_pos_out - grad_prev)^2), the Lipschitz
 * step's denominator (NesterovAcceleratedGradientOptimizer.py:131), formed in
 * the same pass with the bits dp_step_norms gives.  filler_area >= 0: the
 * fillers (nodes num_physical_nodes .. num_nodes - 1, which have no pins, so
 * their pin_weight_sum is 0) all have this area, and neither array is read
 * for them (same results); negative: every node's entries are read. */
int dp_node_grad_preconditioned(int dtype, const void* grad, const void* grad_node_order,
                                const int32_t* flat_node2pin_start_map, const void* add,
                                int add_layout, int64_t num_physical_nodes, int64_t num_nodes,
                                const void* pin_weight_sum, const void* node_areas,
                                const void* density_weight, double alpha, double filler_area,
                                void* grad_pos_out, const void* grad_prev, void* workspace,
                                double* dg_norm2_out, void* stream);

/* Replaces hpwl_cuda.forward (dreamplace/ops/hpwl/src/hpwl_cuda.cpp:27-64):
 * *hpwl_out = sum_n w_n*(max-min)_x + (max-min)_y over unmasked nets, FP64
 * accumulation into a device double, deterministic. */
int dp_hpwl(int dtype, const void* pin_pos, int64_t num_pins, const int32_t* flat_netpin,
            const int32_t* netpin_start, int64_t num_nets, const void* net_weights,
            const uint8_t* net_mask, void* workspace, double* hpwl_out, void* stream);

/* New: the HPWL metric from per-net values another pass already formed
 * (dp_wa_forward_pins' net_hpwl), for a loop that evaluates the objective
 * and the metric at the same positions (NonLinearPlace.py:414 evaluates both
 * there): *hpwl_out = sum over nets with net_mask and at least one pin of
 * net_hpwl[n] * net_weights[n], with dp_hpwl's grid and summation order, so
 * the result is bit-identical to dp_hpwl on those positions.  workspace:
 * dp_wa_workspace_bytes(num_nets). */
int dp_hpwl_from_nets(int dtype, const void* net_hpwl, const int32_t* netpin_start,
                      int64_t num_nets, const void* net_weights, const uint8_t* net_mask,
                      void* workspace, double* hpwl_out, void* stream);

/* Replaces move_boundary_cuda.forward (dreamplace/ops/move_boundary/src/
 * move_boundary_cuda_kernel.cu:10-75): in-place clamp of movables and fillers. */
int dp_move_boundary(int dtype, void* pos, int64_t num_nodes, const void* node_size_x,
                     const void* node_size_y, double xl, double yl, double xh, double yh,
                     int64_t num_movable_nodes, int64_t num_filler_nodes, void* stream);

/* Replaces pws_cuda.forward (dreamplace/ops/pin_weight_sum/src/pws_cuda_kernel.cu:10-21):
 * out[n] = sum of net_weights over the nets of physical node n's pins; out has
 * num_nodes entries, fillers 0. */
int dp_pin_weight_sum(int dtype, const void* net_weights, const int32_t* flat_nodepin,
                      const int32_t* nodepin_start, const int32_t* pin2net_map,
                      int64_t num_physical_nodes, int64_t num_nodes, void* out, void* stream);

/* ---------------------------------------------------------------------------
 * Optimizer (dreamplace/NesterovAcceleratedGradientOptimizer.py:65-240) and
 * preconditioner (dreamplace/PlaceObj.py:69-132).  Scalars live on the device
 * (one element of `dtype`) so that a step needs no host round trip.
 * ------------------------------------------------------------------------- */

size_t dp_optimizer_workspace_bytes(void);

/* One Nesterov extrapolation (step_nobb/step_bb inner update, :120-124,211-212):
 * a_kp1 = (1 + sqrt(4 a_k^2 + 1)) / 2, coef = (a_k - 1) / a_kp1,
 * u_kp1 = v_k - alpha_k * g_k, v_kp1 = u_kp1 + coef * (u_kp1 - u_k), in
 * `dtype` and in the reference's operation order.  a_kp1 (nullable) must not
 * alias a_k. */
int dp_nesterov_extrapolate(int dtype, int64_t n, const void* v_k, const void* g_k,
                            const void* u_k, const void* alpha_k, const void* a_k, void* u_kp1,
                            void* v_kp1, void* a_kp1, void* stream);

/* dp_nesterov_extrapolate over [2*num_nodes] positions followed by the
 * optimizer's constraint, move_boundary on v_kp1
 * (NesterovAcceleratedGradientOptimizer.py:123,212 with PlaceObj's
 * constraint_fn; move_boundary_cuda_kernel.cu:10-40), in one pass: movable
 * and filler nodes are clamped to x in [xl, xh - node_size_x],
 * y in [yl, yh - node_size_y].  The same bits as the two calls.
 * filler_size (host double[2], nullable): the one size all fillers share,
 * read instead of their node_size entries.  dv_norm2_out (device double,
 * nullable; then a dp_optimizer_workspace_bytes workspace is needed):
 * sum((v_kp1 - v_k)^2), the Lipschitz step's numerator (:131), formed in the
 * same pass with the bits dp_step_norms gives. */
int dp_nesterov_extrapolate_bounded(int dtype, int64_t num_nodes, const void* v_k, const void* g_k,
                                    const void* u_k, const void* alpha_k, const void* a_k,
                                    void* u_kp1, void* v_kp1, void* a_kp1,
                                    const void* node_size_x, const void* node_size_y, double xl,
                                    double yl, double xh, double yh, int64_t num_movable_nodes,
                                    int64_t num_filler_nodes, const double* filler_size,
                                    void* workspace, double* dv_norm2_out, void* stream);

/* out3 (device double[3]) = {sum((a1-a0)^2), sum((b1-b0)^2), sum((a1-a0)(b1-b0))}:
 * differences in `dtype`, squares and sums in FP64, deterministic (the n
 * entries are summed as n/2 (x, y) pairs, in the order the fused sums of
 * dp_nesterov_extrapolate_bounded and dp_node_grad_preconditioned follow).  With
 * a = positions and b = gradients these give the Lipschitz step
 * sqrt(out3[0]/out3[1]) (:131) and the Barzilai-Borwein steps (:200-205). */
int dp_step_norms(int dtype, int64_t n, const void* a1, const void* a0, const void* b1,
                  const void* b0, void* workspace, double* out3, void* stream);

/* *out (device double) = sum |g_i| in FP64 (the norms of PlaceObj.
 * initialize_density_weight, dreamplace/PlaceObj.py:777-815). */
int dp_l1_norm(int dtype, int64_t n, const void* g, void* workspace, double* out, void* stream);

/* Replaces PreconditionOp.__call__ without fence regions (PlaceObj.py:69-82,
 * 100-102): precond = max(pin_weight_sum + (alpha * density_weight) *
 * node_areas, 1); grad[0:N] and grad[N:2N] are divided by it in place.
 * density_weight is a device scalar of `dtype`. */
int dp_precondition(int dtype, void* grad, int64_t num_nodes, const void* pin_weight_sum,
                    const void* node_areas, const void* density_weight, double alpha,
                    void* stream);

/* ---- Filler relabelling ----------------------------------------------------
 * No counterpart in the reference: fillers are interchangeable (one size, no
 * pins, PlaceDB.py:917-936), and this library keeps them in the tiled density
 * kernels' tile order so that those kernels take them by tile range.  A
 * placement loop re-sorts them every few tens of iterations; these two calls
 * are the sort's device passes. */

/* key[j] (int16) = tile of filler j (node num_nodes - num_fillers + j):
 * (bx / 64) * ceil(num_bins_y / 64) + by / 64, the tile numbering of the
 * tiled density kernels, for bx = clamp(floor(((x + origin_x) - xl) *
 * (1 / bin_size_x)), 0, num_bins_x - 1) evaluated in `dtype` in that order
 * (origin: the density kernels' stretch offset of a filler's origin), by
 * likewise.  tile_count (nullable, int32 [ntiles]): cleared, then the
 * fillers per tile.  At most 4096 tiles (DP_ERR_UNSUPPORTED beyond).  A
 * stable sort of the keys gives the tile order; the prefix sums of the
 * counts give dp_density_map_ex's filler_tile_start. */
int dp_filler_tile_keys(int dtype, const void* pos, int64_t num_nodes, int64_t num_fillers,
                        double xl, double yl, double bin_size_x, double bin_size_y,
                        int num_bins_x, int num_bins_y, double origin_x, double origin_y,
                        int16_t* keys, int32_t* tile_count, void* stream);

/* In place, for `count` (at most 8) [2 * num_nodes] vectors of `dtype`
 * (vectors: a host array of device pointers): the filler entries of both
 * halves take the order perm (int64 [num_fillers], a permutation of
 * [0, num_fillers), new slot -> old slot): v[h*N + f0 + j] <- old
 * v[h*N + f0 + perm[j]], f0 = num_nodes - num_fillers.  One gather into the
 * workspace (dp_permute_fillers_workspace_bytes(dtype, num_fillers, count))
 * and one store back.  Entries of perm outside [0, num_fillers) are clamped
 * into it (no access leaves the vectors; the result is then not a
 * relabelling). */
size_t dp_permute_fillers_workspace_bytes(int dtype, int64_t num_fillers, int count);
int dp_permute_fillers(int dtype, void* const* vectors, int count, int64_t num_nodes,
                       int64_t num_fillers, const int64_t* perm, void* workspace, void* stream);

/* ---- Bookshelf input / output (host only) ------------------------------- */
/* Node orientations and pin directions as PlaceDB names them. */
enum {
  DP_ORIENT_N = 0, DP_ORIENT_S = 1, DP_ORIENT_W = 2, DP_ORIENT_E = 3,
  DP_ORIENT_FN = 4, DP_ORIENT_FS = 5, DP_ORIENT_FW = 6, DP_ORIENT_FE = 7,
  DP_ORIENT_UNKNOWN = 8
};
enum { DP_PIN_INPUT = 0, DP_PIN_OUTPUT = 1, DP_PIN_INOUT = 2 };
/* Node status after reading: unplaced, placed, dummy-fixed (movable node
 * taller than 5 rows), fixed. */
enum { DP_NODE_UNPLACED = 0, DP_NODE_PLACED = 1, DP_NODE_DUMMY_FIXED = 2, DP_NODE_FIXED = 3 };

/* Replaces PlaceIOFunction.read + PyPlaceDB construction for Bookshelf input
 * (dreamplace/ops/place_io/place_io.py:17-30; PlaceDB.cpp:904-1290,
 * 1726-1734, 1854-2090; PyPlaceDB.cpp:161-440): parses the .nodes, .nets,
 * .wts, .scl and .pl files an .aux file names into a host-side handle, with
 * the reference's node order (movable, fixed, IO pins), net order (by degree
 * when sort_nets_by_degree) and pin order (by net).  Errors: DP_ERR_* with
 * dp_last_error(). */
int dp_bookshelf_read(const char* aux_path, int sort_nets_by_degree, void** handle);
/* sizes[9] = {num_nodes, num_terminals (fixed, not IO), num_terminal_NIs,
 * num_nets, num_pins, num_rows, bytes of the NUL-separated node names, bytes
 * of the net names, num_movable_pins}. */
int dp_bookshelf_sizes(void* handle, int64_t* sizes);
/* Copies the design out (every pointer nullable; host memory):
 * node_x/y (lower-left), node_size_x/y, node_orient (DP_ORIENT_*),
 * node_status (DP_NODE_*), node_names; pin_offset_x/y (from the lower-left
 * corner), pin_direct (DP_PIN_*), pin2node_map, pin2net_map;
 * flat_net2pin_map + flat_net2pin_start_map[num_nets+1],
 * flat_node2pin_map + flat_node2pin_start_map[num_nodes+1]; net_weights,
 * net_ignore (fewer than 2 distinct nodes), net_names; rows[4*num_rows]
 * (xl, yl, xh, yh); die[6] = {xl, yl, xh, yh, row_height, site_width}. */
int dp_bookshelf_export(void* handle, double* node_x, double* node_y, double* node_size_x,
                        double* node_size_y, int32_t* node_orient, int32_t* node_status,
                        char* node_names, double* pin_offset_x, double* pin_offset_y,
                        int32_t* pin_direct, int32_t* pin2node_map, int32_t* pin2net_map,
                        int32_t* flat_net2pin_map, int32_t* flat_net2pin_start_map,
                        int32_t* flat_node2pin_map, int32_t* flat_node2pin_start_map,
                        double* net_weights, uint8_t* net_ignore, char* net_names, double* rows,
                        double* die);
void dp_bookshelf_free(void* handle);
/* Area of the union of n rectangles clipped to clip[4] = {xl, yl, xh, yh}
 * (nullable): the fixed-cell area of PyPlaceDB::computeAreaStatistics
 * (PyPlaceDB.cpp:741-770), which the space area and filler count use. */
double dp_rect_union_area(int64_t n, const double* xl, const double* yl, const double* xh,
                          const double* yh, const double* clip);

#ifdef __cplusplus
}
#endif

#endif /* DREAMPLACE_AMD_H */
