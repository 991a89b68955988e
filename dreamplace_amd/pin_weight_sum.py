"""Per-node sum of net weights over the node's pins: drop-in for
dreamplace/ops/pin_weight_sum/pin_weight_sum.py (``PinWeightSum``,
pin_weight_sum.py:30-65); used by the placer's preconditioner."""
import torch
from torch import nn

from . import _native as N


class PinWeightSum(nn.Module):
    def __init__(self, flat_nodepin=None, nodepin_start=None, pin2net_map=None, num_nodes=None,
                 algorithm='node-by-node'):
        super(PinWeightSum, self).__init__()
        assert flat_nodepin is not None and nodepin_start is not None, \
            "the node-by-node sum needs the node-to-pin CSR (flat_nodepin and nodepin_start)"
        self.flat_nodepin = N.as_i32(flat_nodepin).contiguous()
        self.nodepin_start = N.as_i32(nodepin_start).contiguous()
        self.pin2net_map = N.as_i32(pin2net_map).contiguous()
        self.num_nodes = num_nodes
        self.algorithm = algorithm

    def forward(self, net_weights):
        N.require_gpu(net_weights, name="PinWeightSum")
        w = net_weights.contiguous()
        out = torch.empty(self.num_nodes, dtype=w.dtype, device=w.device)
        N.call("dp_pin_weight_sum", N.dtype_code(w), N.ptr(w), N.ptr(self.flat_nodepin),
               N.ptr(self.nodepin_start), N.ptr(self.pin2net_map), self.nodepin_start.numel() - 1,
               int(self.num_nodes), N.ptr(out), N.stream_ptr(w.device))
        return out
