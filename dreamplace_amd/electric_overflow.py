"""Electric density map and overflow: drop-in for
dreamplace/ops/electric_potential/electric_overflow.py.

``ElectricOverflow`` keeps the reference constructor
(dreamplace/ops/electric_potential/electric_overflow.py:117-159) and
``reset`` / ``compute_initial_density_map`` / ``forward`` semantics
(:161-278).  The density map is always accumulated in int64 fixed point on the
GPU (bit-exact with the reference's deterministic mode for either
``deterministic_flag``), by libdreamplace_amd's ``dp_density_map``.
"""
import math

import torch
from torch import nn

from . import _native as N


def _f(v):
    return float(v.item()) if torch.is_tensor(v) else float(v)


class ElectricDensityMapFunction:
    """Density map of movables + fillers over the initial (fixed-cell) map.

    Same argument list as the reference's static forward
    (electric_overflow.py:57-114); returns the [num_bins_x, num_bins_y] map.
    Extra keyword arguments: ``convention`` ("cpu" follows
    electric_density_map.cpp:125-140, "gpu" follows
    electric_density_map_cuda_kernel.cu:433-463), ``algorithm`` (global |
    tiled scatter), ``large_capacity``, ``workspace`` (a uint8 tensor),
    ``fixed_map_out`` (int64 tensor receiving the raw fixed-point sum) and
    ``arithmetic`` ("reference": the reference's overlaps, bit-exact with its
    map; "bin_relative": overlaps measured from each cell's first bin, see
    DP_ARITH_BIN_RELATIVE).
    """

    @staticmethod
    def forward(pos, node_size_x_clamped, node_size_y_clamped, offset_x, offset_y, ratio,
                bin_center_x, bin_center_y, initial_density_map, target_density, xl, yl, xh, yh,
                bin_size_x, bin_size_y, num_movable_nodes, num_filler_nodes, padding,
                padding_mask, num_bins_x, num_bins_y, num_movable_impacted_bins_x,
                num_movable_impacted_bins_y, num_filler_impacted_bins_x,
                num_filler_impacted_bins_y, deterministic_flag, sorted_node_map,
                convention="cpu", algorithm="global", large_capacity=None, workspace=None,
                fixed_map_out=None, uniform_filler=None, arithmetic="reference"):
        name = "ElectricDensityMapFunction"
        N.require_gpu(pos, node_size_x_clamped, initial_density_map, name=name)
        pos = pos.view(pos.numel())
        N.require_contiguous(pos, node_size_x_clamped, node_size_y_clamped, offset_x, offset_y,
                             ratio, name=name)
        if pos.numel() % 2:
            raise RuntimeError("%s: pos must have an even number of elements" % name)
        num_nodes = pos.numel() // 2
        nbx, nby = int(num_bins_x), int(num_bins_y)
        if large_capacity is None:
            large_capacity = num_nodes
        alg = algorithm_code(algorithm, arithmetic)
        nbytes = N.query("dp_density_map_workspace_bytes", nbx, nby, num_nodes, int(large_capacity),
                         alg, N.dtype_code(pos))
        if workspace is None or workspace.numel() < nbytes:
            workspace = N.workspace.get("density_map", nbytes, pos.device)
        init = initial_density_map
        if init is not None:
            init = init.to(pos.dtype).contiguous()
        out = torch.empty(nbx, nby, dtype=pos.dtype, device=pos.device)
        N.call("dp_density_map", N.dtype_code(pos), N.ptr(pos), num_nodes,
               N.ptr(node_size_x_clamped), N.ptr(node_size_y_clamped), N.ptr(offset_x),
               N.ptr(offset_y), N.ptr(ratio), N.ptr(init), _f(xl), _f(yl), _f(xh), _f(yh),
               _f(bin_size_x), _f(bin_size_y), int(num_movable_nodes), int(num_filler_nodes), nbx,
               nby, N.DP_CONVERT_CPU if convention == "cpu" else N.DP_CONVERT_GPU, alg,
               int(large_capacity), N.host_doubles(uniform_filler), N.ptr(workspace), N.ptr(out),
               N.ptr(fixed_map_out), N.stream_ptr(pos.device))
        if padding > 0:
            out.masked_fill_(padding_mask.bool(), _f(target_density) * _f(bin_size_x) * _f(bin_size_y))
        return out


def algorithm_code(algorithm, arithmetic="reference"):
    """The `algorithm` argument of dp_density_map / dp_electric_force."""
    if arithmetic not in ("reference", "bin_relative"):
        raise ValueError("arithmetic must be 'reference' or 'bin_relative', got %r" % (arithmetic,))
    code = N.DP_SCATTER_TILED if algorithm == "tiled" else N.DP_SCATTER_GLOBAL
    return code | (N.DP_ARITH_BIN_RELATIVE if arithmetic == "bin_relative" else 0)


class ElectricOverflow(nn.Module):
    """electric_overflow.py:117-278 on the HIP path.  Extra keyword arguments
    beyond the reference's: ``convention``, ``algorithm`` and ``arithmetic``
    (see ElectricDensityMapFunction)."""

    def __init__(self, node_size_x, node_size_y, bin_center_x, bin_center_y, target_density, xl,
                 yl, xh, yh, bin_size_x, bin_size_y, num_movable_nodes, num_terminals,
                 num_filler_nodes, padding, deterministic_flag, sorted_node_map,
                 movable_macro_mask=None, convention="cpu", algorithm="global",
                 arithmetic="reference"):
        super(ElectricOverflow, self).__init__()
        algorithm_code(algorithm, arithmetic)  # validates
        self.arithmetic = arithmetic
        self.node_size_x = node_size_x
        self.node_size_y = node_size_y
        self.bin_center_x = bin_center_x
        self.bin_center_y = bin_center_y
        self.target_density = target_density
        self.xl = xl
        self.yl = yl
        self.xh = xh
        self.yh = yh
        self.bin_size_x = bin_size_x
        self.bin_size_y = bin_size_y
        self.num_movable_nodes = num_movable_nodes
        self.num_terminals = num_terminals
        self.num_filler_nodes = num_filler_nodes
        self.padding = padding
        self.sorted_node_map = sorted_node_map
        self.movable_macro_mask = movable_macro_mask
        self.deterministic_flag = deterministic_flag
        self.convention = convention
        self.algorithm = algorithm
        # pass the fillers' shared size to the kernels when reset() finds one
        # (set False to make the kernels read every filler's arrays)
        self.use_uniform_filler = True
        # the tiled gather reuses the cell binning of the last full density
        # call; this counts those calls so a stale binning can be detected
        self.binning_generation = 0
        self._ws = None
        self._ws_overflow = None
        self._red_ws = None
        self.reset()

    def reset(self):
        """electric_overflow.py:162-227: stretched sizes, offsets, ratios, bins."""
        sqrt2 = math.sqrt(2)
        self.node_size_x_clamped = self.node_size_x.clamp(min=self.bin_size_x * sqrt2)
        self.offset_x = (self.node_size_x - self.node_size_x_clamped).mul(0.5)
        self.node_size_y_clamped = self.node_size_y.clamp(min=self.bin_size_y * sqrt2)
        self.offset_y = (self.node_size_y - self.node_size_y_clamped).mul(0.5)
        node_areas = self.node_size_x * self.node_size_y
        self.ratio = node_areas / (self.node_size_x_clamped * self.node_size_y_clamped)

        self.num_movable_macros = 0
        if _f(self.target_density) < 1 and self.movable_macro_mask is not None:
            self.num_movable_macros = int(self.movable_macro_mask.sum().item())
            r = self.ratio[:self.num_movable_nodes]
            r[self.movable_macro_mask.to(r.device).bool()] = _f(self.target_density)

        self.num_bins_x = int(round((self.xh - self.xl) / self.bin_size_x))
        self.num_bins_y = int(round((self.yh - self.yl) / self.bin_size_y))
        # impacted-bin counts are kept for interface parity; the kernels derive
        # each cell's span themselves
        nm, nf = self.num_movable_nodes, self.num_filler_nodes
        if nm:
            self.num_movable_impacted_bins_x = int(math.ceil(min(
                (_f(self.node_size_x[:nm].max()) + 2 * sqrt2 * self.bin_size_x) / self.bin_size_x,
                self.num_bins_x)))
            self.num_movable_impacted_bins_y = int(math.ceil(min(
                (_f(self.node_size_y[:nm].max()) + 2 * sqrt2 * self.bin_size_y) / self.bin_size_y,
                self.num_bins_y)))
        else:
            self.num_movable_impacted_bins_x = self.num_movable_impacted_bins_y = 0
        if nf:
            self.num_filler_impacted_bins_x = int(math.ceil(min(
                (_f(self.node_size_x[-nf:].max()) + 2 * sqrt2 * self.bin_size_x) / self.bin_size_x,
                self.num_bins_x)))
            self.num_filler_impacted_bins_y = int(math.ceil(min(
                (_f(self.node_size_y[-nf:].max()) + 2 * sqrt2 * self.bin_size_y) / self.bin_size_y,
                self.num_bins_y)))
        else:
            self.num_filler_impacted_bins_x = self.num_filler_impacted_bins_y = 0
        self.padding_mask = torch.zeros(self.num_bins_x, self.num_bins_y, dtype=torch.uint8,
                                        device=self.node_size_x.device)
        if self.padding > 0:
            self.padding_mask.fill_(1)
            self.padding_mask[self.padding:self.num_bins_x - self.padding,
                              self.padding:self.num_bins_y - self.padding].fill_(0)
        # cells whose stretched size can span more than 3 bins go through the
        # per-cell workgroup path; the queue is sized for them
        big = (self.node_size_x_clamped > 1.9 * self.bin_size_x) | \
              (self.node_size_y_clamped > 1.9 * self.bin_size_y)
        self.large_capacity = int(big.sum().item()) + 64
        # PlaceDB gives every filler one size (dreamplace/PlaceDB.py:917-936):
        # then the kernels take the fillers' clamped size, offset and ratio as
        # five scalars instead of reading them per cell
        self.uniform_filler = None
        if nf > 0:
            vals = []
            for t in (self.node_size_x_clamped, self.node_size_y_clamped, self.offset_x,
                      self.offset_y, self.ratio):
                f = t[-nf:]
                if not bool((f == f[0]).all().item()):
                    vals = None
                    break
                vals.append(float(f[0].item()))
            self.uniform_filler = vals
        self.initial_density_map = None

    @property
    def algorithm_code(self):
        return algorithm_code(self.algorithm, self.arithmetic)

    def uniform_filler_arg(self):
        """The uniform_filler argument of dp_density_map / dp_electric_force."""
        return N.host_doubles(self.uniform_filler if self.use_uniform_filler else None)

    def workspace(self, device, overflow=False):
        """The density workspace (its tiled region holds the binning the field
        gather reuses); `overflow` gives a second one for the movables-only
        overflow map, so that map never replaces that binning."""
        nbytes = N.query("dp_density_map_workspace_bytes", self.num_bins_x, self.num_bins_y,
                         self.node_size_x.numel(), self.large_capacity, self.algorithm_code,
                         N.dtype_code(self.node_size_x))
        if overflow:
            if self._ws_overflow is None or self._ws_overflow.numel() < nbytes:
                self._ws_overflow = torch.empty(nbytes, dtype=torch.uint8, device=device)
            return self._ws_overflow
        if self._ws is None or self._ws.numel() < nbytes:
            self._ws = torch.empty(nbytes, dtype=torch.uint8, device=device)
        return self._ws

    def reduce_workspace(self, device):
        if self._red_ws is None:
            self._red_ws = torch.empty(N.query("dp_reduce_workspace_bytes", 0), dtype=torch.uint8,
                                       device=device)
        return self._red_ws

    def compute_initial_density_map(self, pos):
        """electric_overflow.py:229-258 (fixed_density_map, then * target)."""
        N.require_gpu(pos, name="ElectricOverflow")
        pos = pos.view(-1).contiguous()
        out = torch.empty(self.num_bins_x, self.num_bins_y, dtype=pos.dtype, device=pos.device)
        N.call("dp_fixed_density_map", N.dtype_code(pos), N.ptr(pos), pos.numel() // 2,
               N.ptr(self.node_size_x), N.ptr(self.node_size_y), N.ptr(self.bin_center_x),
               N.ptr(self.bin_center_y), float(self.xl), float(self.yl), float(self.xh),
               float(self.yh), float(self.bin_size_x), float(self.bin_size_y),
               int(self.num_movable_nodes), int(self.num_terminals), self.num_bins_x,
               self.num_bins_y, _f(self.target_density),
               (N.DP_CONVERT_CPU if self.convention == "cpu" else N.DP_CONVERT_GPU)
               | (N.DP_ARITH_BIN_RELATIVE if self.arithmetic == "bin_relative" else 0),
               N.ptr(self.workspace(pos.device)), N.ptr(out), N.stream_ptr(pos.device))
        self.initial_density_map = out

    def density_map(self, pos, num_filler_nodes=None, fixed_map_out=None):
        nf = self.num_filler_nodes if num_filler_nodes is None else num_filler_nodes
        full = nf == self.num_filler_nodes
        if full:
            self.binning_generation += 1
        return ElectricDensityMapFunction.forward(
            pos, self.node_size_x_clamped, self.node_size_y_clamped, self.offset_x, self.offset_y,
            self.ratio, self.bin_center_x, self.bin_center_y, self.initial_density_map,
            self.target_density, self.xl, self.yl, self.xh, self.yh, self.bin_size_x,
            self.bin_size_y, self.num_movable_nodes, nf, self.padding, self.padding_mask,
            self.num_bins_x, self.num_bins_y, self.num_movable_impacted_bins_x,
            self.num_movable_impacted_bins_y, self.num_filler_impacted_bins_x,
            self.num_filler_impacted_bins_y, self.deterministic_flag, self.sorted_node_map,
            convention=self.convention, algorithm=self.algorithm, arithmetic=self.arithmetic,
            large_capacity=self.large_capacity,
            workspace=self.workspace(pos.device, overflow=not full), fixed_map_out=fixed_map_out,
            uniform_filler=self.uniform_filler if self.use_uniform_filler else None)

    def overflow_from_map(self, density_map):
        """(sum max(D - t*A_bin, 0), max(D)/A_bin) with FP64 accumulation, as
        tensors of the map's dtype, shape [1] (electric_overflow.py:271-278)."""
        bin_area = self.bin_size_x * self.bin_size_y
        out2 = torch.empty(2, dtype=torch.float64, device=density_map.device)
        N.call("dp_density_overflow", N.dtype_code(density_map), N.ptr(density_map),
               density_map.numel(), _f(self.target_density) * bin_area,
               N.ptr(self.reduce_workspace(density_map.device)), N.ptr(out2),
               N.stream_ptr(density_map.device))
        cost = out2[0:1].to(density_map.dtype)
        mx = (out2[1:2] / bin_area).to(density_map.dtype)
        return cost, mx

    def forward(self, pos):
        if self.initial_density_map is None:
            self.compute_initial_density_map(pos)
        density_map = self.density_map(pos)
        return self.overflow_from_map(density_map)
