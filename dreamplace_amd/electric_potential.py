"""Electrostatic density penalty: drop-in for
dreamplace/ops/electric_potential/electric_potential.py.

``ElectricPotential`` keeps the reference constructor
(electric_potential.py:282-307) and ``forward(pos, mode="density"|"overflow")``
(:433-553).  In "density" mode it returns the energy sum(phi*rho) as a 0-dim
tensor whose backward is the reference's electric force
(-ratio * sum overlap * E, electric_potential.py:219-245).  The chain is:
int64 fixed-point scatter -> rho -> DCT2 -> three inverse transforms on the
matrix cores -> energy; backward is the fused field gather.

Additions (keyword-only, defaults reproduce the reference):
``precision`` ("fp64" | "fp32" | "bf16x3") for the spectral solve,
``energy_mode`` ("potential": materialise phi = IDCT2(...) as the reference
does; "spectral": evaluate the same energy from the DCT coefficients without
the fourth transform), ``convention``, ``algorithm`` and ``arithmetic``
("reference" | "bin_relative", DP_ARITH_BIN_RELATIVE) of the density map and
the field gather, and ``backward``: "force" (the reference's field gather) or "adjoint" (the
exact derivative of the returned energy, dp_energy_gradient; the forward
then also materialises phi).
Fence regions (``region_id`` / ``fence_regions``) are not supported and raise.
"""
import logging

import torch
from torch.autograd import Function

from . import _native as N
from .electric_overflow import ElectricOverflow, _f
from .spectral import default_precision, get_plan

logger = logging.getLogger(__name__)


class ElectricPotentialFunction(Function):
    """Forward: energy of the density map of ``pos``; backward: field gather."""

    @staticmethod
    def forward(ctx, pos, op):
        density_map = op.density_map(pos)
        fx, fy, energy = op.solve(density_map)
        ctx.op = op
        ctx.field_map_x = fx
        ctx.field_map_y = fy
        ctx.potential_map = op.potential_map
        ctx.binning = op.binning_generation
        ctx.save_for_backward(pos)
        op.last_density_map = density_map
        return energy

    @staticmethod
    def backward(ctx, grad_pos):
        (pos,) = ctx.saved_tensors
        op = ctx.op
        if op.algorithm == "tiled" and op.binning_generation != ctx.binning:
            # another density call has re-binned the cells since this forward:
            # bin these positions again for the gather
            op.density_map(pos)
        if op.backward == "adjoint":
            return op.energy_gradient(pos, ctx.potential_map, grad_pos), None
        grad = op.electric_force(pos, ctx.field_map_x, ctx.field_map_y, grad_pos)
        return grad, None


class ElectricPotential(ElectricOverflow):
    """electric_potential.py:278-553 on the HIP path."""

    def __init__(self, node_size_x, node_size_y, bin_center_x, bin_center_y, target_density, xl,
                 yl, xh, yh, bin_size_x, bin_size_y, num_movable_nodes, num_terminals,
                 num_filler_nodes, padding, deterministic_flag, sorted_node_map,
                 movable_macro_mask=None, fast_mode=False, region_id=None, fence_regions=None,
                 node2fence_region_map=None, placedb=None, *, precision=None,
                 energy_mode="potential", convention="cpu", algorithm="global",
                 backward="force", arithmetic="reference"):
        if region_id is not None or fence_regions is not None:
            raise NotImplementedError("dreamplace_amd.ElectricPotential: fence regions are not supported")
        super(ElectricPotential, self).__init__(
            node_size_x=node_size_x, node_size_y=node_size_y, bin_center_x=bin_center_x,
            bin_center_y=bin_center_y, target_density=target_density, xl=xl, yl=yl, xh=xh, yh=yh,
            bin_size_x=bin_size_x, bin_size_y=bin_size_y, num_movable_nodes=num_movable_nodes,
            num_terminals=num_terminals, num_filler_nodes=num_filler_nodes, padding=padding,
            deterministic_flag=deterministic_flag, sorted_node_map=sorted_node_map,
            movable_macro_mask=movable_macro_mask, convention=convention, algorithm=algorithm,
            arithmetic=arithmetic)
        self.fast_mode = fast_mode
        self.region_id = region_id
        self.fence_regions = fence_regions
        self.placedb = placedb
        if energy_mode not in ("potential", "spectral"):
            raise ValueError("energy_mode must be 'potential' or 'spectral'")
        if backward not in ("force", "adjoint"):
            raise ValueError("backward must be 'force' or 'adjoint'")
        self.precision = precision
        self.energy_mode = energy_mode
        self.backward = backward
        self.last_density_map = None
        self.potential_map = None

    def reset(self):
        super(ElectricPotential, self).reset()
        logger.info("%d movable cells are treated as macros by the density model" % self.num_movable_macros)

    def plan(self, dtype, device):
        prec = self.precision or default_precision(dtype)
        return get_plan(self.num_bins_x, self.num_bins_y, prec, device)

    def solve(self, density_map):
        """(field_x, field_y, energy) of a density map, in the map's dtype."""
        plan = self.plan(density_map.dtype, density_map.device)
        d = density_map if density_map.dtype == plan.dtype else density_map.to(plan.dtype)
        want_energy = not self.fast_mode
        want_phi = (want_energy and self.energy_mode == "potential") or self.backward == "adjoint"
        fx, fy, phi, energy = plan.poisson(d, self.bin_size_x, self.bin_size_y,
                                           compute_potential=want_phi, compute_energy=want_energy)
        if phi is not None and phi.dtype != density_map.dtype:
            phi = phi.to(density_map.dtype)
        self.potential_map = phi
        if fx.dtype != density_map.dtype:
            fx, fy = fx.to(density_map.dtype), fy.to(density_map.dtype)
            energy = energy.to(density_map.dtype) if energy is not None else None
        if energy is None:  # fast mode: a dummy that still carries the backward
            energy = torch.zeros((), dtype=density_map.dtype, device=density_map.device)
        return fx, fy, energy

    def electric_force(self, pos, field_x, field_y, grad_pos, grad_layout="split"):
        """-(ratio * sum overlap*E) * grad for movables and fillers; 0 for
        terminals (electric_potential.py:219-245).  grad_layout "split" is the
        reference's [2N] (all x, then all y); "interleaved" holds (x, y) of
        node i at [2i, 2i+1]."""
        layout = {"split": N.DP_LAYOUT_SPLIT, "interleaved": N.DP_LAYOUT_INTERLEAVED}[grad_layout]
        pos = pos.view(-1)
        grad_scale = grad_pos.reshape(()).to(pos.dtype).contiguous()
        out = torch.empty_like(pos)
        N.call("dp_electric_force", N.dtype_code(pos), N.ptr(grad_scale), N.ptr(field_x),
               N.ptr(field_y), N.ptr(pos), pos.numel() // 2, N.ptr(self.node_size_x_clamped),
               N.ptr(self.node_size_y_clamped), N.ptr(self.offset_x), N.ptr(self.offset_y),
               N.ptr(self.ratio), float(self.xl), float(self.yl), float(self.bin_size_x),
               float(self.bin_size_y), int(self.num_movable_nodes), int(self.num_filler_nodes),
               self.num_bins_x, self.num_bins_y, self.algorithm_code, self.large_capacity,
               layout, self.uniform_filler_arg(), N.ptr(self.workspace(pos.device)), N.ptr(out),
               N.stream_ptr(pos.device))
        return out

    def energy_gradient(self, pos, potential_map, grad_pos, grad_layout="split"):
        """The exact derivative of the energy sum(phi*rho) with respect to
        the positions, times grad_pos (dp_energy_gradient): 2/(bx*by) *
        ratio * sum (overlap derivative) * phi for movables and fillers, 0
        for terminals.  grad_layout as for electric_force."""
        layout = {"split": N.DP_LAYOUT_SPLIT, "interleaved": N.DP_LAYOUT_INTERLEAVED}[grad_layout]
        pos = pos.view(-1)
        grad_scale = grad_pos.reshape(()).to(pos.dtype).contiguous()
        phi = potential_map.to(pos.dtype).contiguous()
        out = torch.empty_like(pos)
        N.call("dp_energy_gradient", N.dtype_code(pos), N.ptr(grad_scale), N.ptr(phi), N.ptr(pos),
               pos.numel() // 2, N.ptr(self.node_size_x_clamped), N.ptr(self.node_size_y_clamped),
               N.ptr(self.offset_x), N.ptr(self.offset_y), N.ptr(self.ratio), float(self.xl),
               float(self.yl), float(self.bin_size_x), float(self.bin_size_y),
               int(self.num_movable_nodes), int(self.num_filler_nodes), self.num_bins_x,
               self.num_bins_y, layout, self.uniform_filler_arg(), N.ptr(out),
               N.stream_ptr(pos.device))
        return out

    def forward(self, pos, mode="density"):
        assert mode in {"density", "overflow"}, "forward mode must be 'density' or 'overflow'"
        N.require_gpu(pos, name="ElectricPotential")
        if self.initial_density_map is None:
            self.compute_initial_density_map(pos)
            logger.info("fixed density map: average %g, max %g, bin area %g" %
                        (_f(self.initial_density_map.mean()), _f(self.initial_density_map.max()),
                         self.bin_size_x * self.bin_size_y))
        if mode == "density":
            return ElectricPotentialFunction.apply(pos, self)
        density_map = self.density_map(pos, num_filler_nodes=0)
        cost, mx = self.overflow_from_map(density_map)
        return cost[0], mx[0]
