"""Half-perimeter wirelength: drop-in for dreamplace/ops/hpwl/hpwl.py
(``HPWL``, hpwl.py:59-107).  Both algorithm names run the net-parallel
kernel; the weighted sum is accumulated in FP64 on the GPU.  ``forward``
returns a 0-dim tensor of the position dtype; ``forward_fp64`` returns the
FP64 value itself.
"""
import torch
from torch import nn

from . import _native as N


def hpwl_fp64(pos, flat_netpin, netpin_start, net_weights, net_mask):
    N.require_gpu(pos, name="HPWL")
    pos = pos.view(-1).contiguous()
    num_pins = pos.numel() // 2
    num_nets = netpin_start.numel() - 1
    w = net_weights.to(pos.dtype).contiguous() if net_weights is not None and net_weights.numel() else None
    out = torch.empty((), dtype=torch.float64, device=pos.device)
    ws = N.workspace.get("hpwl", N.query("dp_reduce_workspace_bytes", num_nets), pos.device)
    N.call("dp_hpwl", N.dtype_code(pos), N.ptr(pos), num_pins, N.ptr(N.as_i32(flat_netpin)),
           N.ptr(N.as_i32(netpin_start)), num_nets, N.ptr(w), N.ptr(N.as_u8(net_mask)), N.ptr(ws),
           N.ptr(out), N.stream_ptr(pos.device))
    return out


class HPWL(nn.Module):
    def __init__(self, flat_netpin=None, netpin_start=None, pin2net_map=None, net_weights=None,
                 net_mask=None, algorithm='atomic'):
        super(HPWL, self).__init__()
        assert net_mask is not None, "HPWL needs a net mask (net_mask=None given)"
        assert flat_netpin is not None and netpin_start is not None, \
            "flat_netpin, netpin_start are required by the net-parallel kernel"
        self.flat_netpin = N.as_i32(flat_netpin)
        self.netpin_start = N.as_i32(netpin_start)
        self.pin2net_map = pin2net_map
        self.net_weights = net_weights
        self.net_mask = N.as_u8(net_mask)
        self.algorithm = algorithm

    def forward_fp64(self, pos):
        return hpwl_fp64(pos, self.flat_netpin, self.netpin_start, self.net_weights, self.net_mask)

    def forward(self, pos):
        return self.forward_fp64(pos).to(pos.dtype)
