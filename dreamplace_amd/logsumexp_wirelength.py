"""Log-sum-exp wirelength: drop-in for
dreamplace/ops/logsumexp_wirelength/logsumexp_wirelength.py.

``LogSumExpWirelength(flat_netpin, netpin_start, pin2net_map, net_weights,
net_mask, pin_mask, gamma, algorithm)`` keeps the reference constructor
(logsumexp_wirelength.py:124-165) and ``forward(pos)`` on pin positions.
Both algorithm names ('merged', 'atomic') run the merged net-parallel
kernel: the two compute the same per-net sums, and 'atomic' only differs in
the reference's CUDA scheduling.  ``gamma`` is a 0-dim device tensor (the
model's gamma itself, not its inverse) read on every call.

As for the WA op, when ``net_weights.requires_grad`` the backward also
returns d(wl)/d(net_weights) = grad * WL_n for unmasked nets (the reference
returns None, logsumexp_wirelength.py:121).
"""
import torch
from torch import nn
from torch.autograd import Function

from . import _native as N


class LogSumExpWirelengthMergedFunction(Function):
    @staticmethod
    def forward(ctx, pos, flat_netpin, netpin_start, pin2net_map, net_weights, net_mask, pin_mask,
                gamma):
        """Same inputs as logsumexp_wirelength.py:73-104; returns the 0-dim
        weighted wirelength."""
        name = "LogSumExpWirelength"
        N.require_gpu(pos, flat_netpin, netpin_start, net_weights, net_mask, name=name)
        pos = pos.view(pos.numel())
        N.require_contiguous(pos, flat_netpin, netpin_start, name=name)
        if pos.numel() % 2:
            raise RuntimeError("%s: pos must have an even number of elements" % name)
        num_pins = pos.numel() // 2
        num_nets = netpin_start.numel() - 1
        dt = pos.dtype
        gamma = gamma.reshape(()).to(dt).contiguous()
        w = net_weights.detach().to(dt).contiguous() if net_weights is not None and net_weights.numel() else None
        mask = N.as_u8(net_mask).contiguous()
        partial = torch.empty(num_nets, dtype=dt, device=pos.device)
        grad_intermediate = torch.zeros_like(pos)
        wl = torch.empty((), dtype=dt, device=pos.device)
        ws = N.workspace.get("wa", N.query("dp_wa_workspace_bytes", num_nets), pos.device)
        fnp, nps = N.as_i32(flat_netpin), N.as_i32(netpin_start)
        N.call("dp_lse_merged_forward", N.dtype_code(pos), N.ptr(pos), num_pins, N.ptr(fnp),
               N.ptr(nps), num_nets, N.ptr(w), N.ptr(mask), N.ptr(gamma), N.ptr(partial),
               N.ptr(grad_intermediate), N.ptr(wl), N.ptr(ws), N.stream_ptr(pos.device))
        ctx.pin2net_map = N.as_i32(pin2net_map).contiguous()
        ctx.net_weights = w
        ctx.net_mask = mask
        ctx.pin_mask = N.as_u8(pin_mask).contiguous() if pin_mask is not None else None
        ctx.grad_intermediate = grad_intermediate
        ctx.partial_wl = partial
        ctx.num_pins = num_pins
        ctx.weights_need_grad = net_weights is not None and net_weights.requires_grad
        ctx.weights_dtype = net_weights.dtype if net_weights is not None else dt
        return wl

    @staticmethod
    def backward(ctx, grad_pos):
        """logsumexp_wirelength.py:106-121: grad_intermediate * grad, times
        the net weight, pin_mask pins zeroed, in place."""
        gi = ctx.grad_intermediate
        grad = grad_pos.reshape(()).to(gi.dtype).contiguous()
        N.call("dp_wa_merged_backward", N.dtype_code(gi), N.ptr(grad), N.ptr(gi), ctx.num_pins,
               N.ptr(ctx.pin2net_map), N.ptr(ctx.net_weights), N.ptr(ctx.net_mask),
               N.ptr(ctx.pin_mask), N.stream_ptr(gi.device))
        grad_w = None
        if ctx.weights_need_grad:
            grad_w = torch.empty_like(ctx.partial_wl)
            N.call("dp_wa_net_weight_grad", N.dtype_code(gi), N.ptr(grad), N.ptr(ctx.partial_wl),
                   N.ptr(ctx.net_mask), ctx.partial_wl.numel(), N.ptr(grad_w),
                   N.stream_ptr(gi.device))
            grad_w = grad_w.to(ctx.weights_dtype)
        return gi, None, None, None, grad_w, None, None, None


class LogSumExpWirelength(nn.Module):
    def __init__(self, flat_netpin=None, netpin_start=None, pin2net_map=None, net_weights=None,
                 net_mask=None, pin_mask=None, gamma=None, algorithm='merged'):
        super(LogSumExpWirelength, self).__init__()
        if net_weights is None or net_mask is None or pin_mask is None or gamma is None:
            raise AssertionError("log-sum-exp wirelength: net_weights, net_mask, pin_mask and gamma must all be given")
        if algorithm not in ("merged", "atomic"):
            raise AssertionError("unknown wirelength algorithm %r (expected 'merged' or 'atomic')" % algorithm)
        if flat_netpin is None or netpin_start is None or pin2net_map is None:
            raise AssertionError("log-sum-exp wirelength: the net CSR (flat_netpin, netpin_start) and pin2net_map must all be given")
        self.flat_netpin = flat_netpin
        self.netpin_start = netpin_start
        self.pin2net_map = pin2net_map
        self.net_weights = net_weights
        self.net_mask = net_mask
        self.pin_mask = pin_mask
        self.gamma = gamma
        self.algorithm = algorithm

    def forward(self, pos):
        return LogSumExpWirelengthMergedFunction.apply(
            pos, self.flat_netpin, self.netpin_start, self.pin2net_map, self.net_weights,
            self.net_mask, self.pin_mask, self.gamma)
