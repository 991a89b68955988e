"""Weighted-average wirelength: drop-in for
dreamplace/ops/weighted_average_wirelength/weighted_average_wirelength.py.

``WeightedAverageWirelength(flat_netpin, netpin_start, pin2net_map,
net_weights, net_mask, pin_mask, gamma, algorithm)`` keeps the reference
constructor (weighted_average_wirelength.py:216-264) and ``forward(pos)`` on
pin positions.  Every algorithm name runs the merged net-parallel kernel
(the one PlaceObj selects, dreamplace/PlaceObj.py:467-475).  ``gamma`` is a
0-dim device tensor read on every call (PlaceObj mutates it in place).

New: when ``net_weights.requires_grad`` the backward also returns
d(wl)/d(net_weights) = grad * WL_n for unmasked nets (the reference returns
None there, weighted_average_wirelength.py:213).

New: ``terms`` chooses what the forward keeps for the backward.  "store"
keeps the per-pin derivative [2P], as the reference's merged function does
(weighted_average_wirelength.py:189).  "recompute" keeps nothing per pin and
recomputes the exp/max terms in the backward (one more pass over the pins),
writing the pin gradient directly; the result is the same.  "auto" (the
default) stores when the derivative's 2*P*itemsize bytes fit in
``store_limit_bytes`` (default: 1/32 of the device's memory) and recomputes
otherwise.
"""
import torch
from torch import nn
from torch.autograd import Function

from . import _native as N


class WeightedAverageWirelengthMergedFunction(Function):
    @staticmethod
    def forward(ctx, pos, flat_netpin, netpin_start, pin2net_map, net_weights, net_mask, pin_mask,
                inv_gamma, recompute=False):
        """Same inputs as weighted_average_wirelength.py:165-195 (plus the
        store/recompute choice); returns the 0-dim weighted wirelength."""
        name = "WeightedAverageWirelength"
        N.require_gpu(pos, flat_netpin, netpin_start, net_weights, net_mask, name=name)
        pos_in = pos
        pos = pos.view(pos.numel())
        N.require_contiguous(pos, flat_netpin, netpin_start, name=name)
        if pos.numel() % 2:
            raise RuntimeError("%s: pos must have an even number of elements" % name)
        num_pins = pos.numel() // 2
        num_nets = netpin_start.numel() - 1
        dt = pos.dtype
        inv_gamma = inv_gamma.reshape(()).to(dt).contiguous()
        w = net_weights.detach().to(dt).contiguous() if net_weights is not None and net_weights.numel() else None
        mask = N.as_u8(net_mask).contiguous()
        partial = torch.empty(num_nets, dtype=dt, device=pos.device)
        grad_intermediate = None if recompute else torch.zeros_like(pos)
        wl = torch.empty((), dtype=dt, device=pos.device)
        ws = N.workspace.get("wa", N.query("dp_wa_workspace_bytes", num_nets), pos.device)
        fnp, nps = N.as_i32(flat_netpin), N.as_i32(netpin_start)
        N.call("dp_wa_merged_forward", N.dtype_code(pos), N.ptr(pos), num_pins, N.ptr(fnp),
               N.ptr(nps), num_nets, N.ptr(w), N.ptr(mask), N.ptr(inv_gamma), N.ptr(partial),
               N.ptr(grad_intermediate), N.ptr(wl), N.ptr(ws), N.stream_ptr(pos.device))
        ctx.pin2net_map = N.as_i32(pin2net_map).contiguous()
        ctx.net_weights = w
        ctx.net_mask = mask
        ctx.pin_mask = N.as_u8(pin_mask).contiguous() if pin_mask is not None else None
        ctx.grad_intermediate = grad_intermediate
        if recompute:
            # what the backward recomputes from: the pin positions (saved
            # without a copy; autograd rejects a backward after they were
            # changed in place), the net CSR and 1/gamma
            ctx.save_for_backward(pos_in)
            ctx.flat_netpin, ctx.netpin_start = fnp, nps
            ctx.inv_gamma, ctx.num_nets = inv_gamma, num_nets
        ctx.partial_wl = partial
        ctx.num_pins = num_pins
        ctx.weights_need_grad = net_weights is not None and net_weights.requires_grad
        ctx.weights_dtype = net_weights.dtype if net_weights is not None else dt
        return wl

    @staticmethod
    def backward(ctx, grad_pos):
        gi = ctx.grad_intermediate
        # drop the context's reference, so autograd can take the buffer as
        # the gradient instead of copying it
        ctx.grad_intermediate = None
        if gi is None:
            pos = ctx.saved_tensors[0].reshape(-1)
            grad = grad_pos.reshape(()).to(pos.dtype).contiguous()
            gi = torch.zeros_like(pos)
            N.call("dp_wa_merged_backward_recompute", N.dtype_code(pos), N.ptr(grad), N.ptr(pos),
                   ctx.num_pins, N.ptr(ctx.flat_netpin), N.ptr(ctx.netpin_start), ctx.num_nets,
                   N.ptr(ctx.net_weights), N.ptr(ctx.net_mask), N.ptr(ctx.pin_mask),
                   N.ptr(ctx.inv_gamma), N.ptr(gi), N.stream_ptr(pos.device))
        else:
            grad = grad_pos.reshape(()).to(gi.dtype).contiguous()
            # in place on grad_intermediate, as the reference does
            N.call("dp_wa_merged_backward", N.dtype_code(gi), N.ptr(grad), N.ptr(gi),
                   ctx.num_pins, N.ptr(ctx.pin2net_map), N.ptr(ctx.net_weights),
                   N.ptr(ctx.net_mask), N.ptr(ctx.pin_mask), N.stream_ptr(gi.device))
        grad_w = None
        if ctx.weights_need_grad:
            grad_w = torch.empty_like(ctx.partial_wl)
            N.call("dp_wa_net_weight_grad", N.dtype_code(gi), N.ptr(grad), N.ptr(ctx.partial_wl),
                   N.ptr(ctx.net_mask), ctx.partial_wl.numel(), N.ptr(grad_w),
                   N.stream_ptr(gi.device))
            grad_w = grad_w.to(ctx.weights_dtype)
        return gi, None, None, None, grad_w, None, None, None, None


class WeightedAverageWirelength(nn.Module):
    def __init__(self, flat_netpin=None, netpin_start=None, pin2net_map=None, net_weights=None,
                 net_mask=None, pin_mask=None, gamma=None, algorithm='atomic', terms="auto",
                 store_limit_bytes=None):
        super(WeightedAverageWirelength, self).__init__()
        assert net_weights is not None and net_mask is not None and pin_mask is not None \
            and gamma is not None, \
            "weighted-average wirelength: net_weights, net_mask, pin_mask and gamma must all be given"
        assert flat_netpin is not None and netpin_start is not None and pin2net_map is not None, \
            "the merged weighted-average kernels need the net CSR and pin2net_map"
        if algorithm not in ("net-by-net", "atomic", "merged"):
            raise ValueError("unknown algorithm %s" % algorithm)
        self.flat_netpin = flat_netpin
        self.netpin_start = netpin_start
        self.pin2net_map = pin2net_map
        self.net_weights = net_weights
        self.net_mask = net_mask
        self.pin_mask = pin_mask
        self.gamma = gamma
        self.algorithm = algorithm
        if terms not in ("auto", "store", "recompute"):
            raise ValueError("terms must be auto, store or recompute, got %r" % (terms,))
        self.terms = terms
        self.store_limit_bytes = store_limit_bytes

    def recomputes(self, pos):
        """Whether a forward on `pos` keeps nothing per pin for the backward."""
        if self.terms != "auto":
            return self.terms == "recompute"
        limit = self.store_limit_bytes
        if limit is None:
            limit = torch.cuda.get_device_properties(pos.device).total_memory // 32
        return pos.numel() * pos.element_size() > limit

    def forward(self, pos):
        return WeightedAverageWirelengthMergedFunction.apply(
            pos, self.flat_netpin, self.netpin_start, self.pin2net_map, self.net_weights,
            self.net_mask, self.pin_mask, 1.0 / self.gamma, self.recomputes(pos))
