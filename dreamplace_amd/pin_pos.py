"""Pin positions from cell positions: drop-in for
dreamplace/ops/pin_pos/pin_pos.py (``PinPos``, pin_pos.py:103-156).

Forward: pin_pos[p] = offset[p] + pos[pin2node[p]] (x block then y block).
Backward: node gradient = sum of pin gradients over the node's CSR pins
(flat_node2pin_map / flat_node2pin_start_map); fillers receive 0.
Both algorithm names ('segment', 'node-by-node') run the same kernels.
"""
import torch
from torch import nn
from torch.autograd import Function

from . import _native as N


class PinPosFunction(Function):
    @staticmethod
    def forward(ctx, pos, pin_offset_x, pin_offset_y, pin2node_map, flat_node2pin_map,
                flat_node2pin_start_map, num_physical_nodes):
        N.require_gpu(pos, pin_offset_x, pin2node_map, name="PinPos")
        pos = pos.view(pos.numel()).contiguous()
        num_nodes = pos.numel() // 2
        num_pins = pin2node_map.numel()
        out = torch.empty(2 * num_pins, dtype=pos.dtype, device=pos.device)
        idx_type = N.DP_INDEX_INT64 if pin2node_map.dtype == torch.int64 else N.DP_INDEX_INT32
        p2n = pin2node_map if pin2node_map.dtype in (torch.int32, torch.int64) else pin2node_map.long()
        N.call("dp_pin_pos_forward", N.dtype_code(pos), N.ptr(pos), num_nodes, N.ptr(pin_offset_x),
               N.ptr(pin_offset_y), N.ptr(p2n.contiguous()), idx_type, num_pins, N.ptr(out),
               N.stream_ptr(pos.device))
        ctx.num_nodes = num_nodes
        ctx.num_pins = num_pins
        ctx.num_physical_nodes = int(num_physical_nodes)
        ctx.flat_node2pin_map = N.as_i32(flat_node2pin_map).contiguous()
        ctx.flat_node2pin_start_map = N.as_i32(flat_node2pin_start_map).contiguous()
        return out

    @staticmethod
    def backward(ctx, grad_pin_pos):
        g = grad_pin_pos.contiguous()
        out = torch.empty(2 * ctx.num_nodes, dtype=g.dtype, device=g.device)
        N.call("dp_pin_pos_backward", N.dtype_code(g), N.ptr(g), ctx.num_pins,
               N.ptr(ctx.flat_node2pin_map), N.ptr(ctx.flat_node2pin_start_map),
               ctx.num_physical_nodes, ctx.num_nodes, N.ptr(out), N.stream_ptr(g.device))
        return out, None, None, None, None, None, None


class PinPos(nn.Module):
    def __init__(self, pin_offset_x, pin_offset_y, pin2node_map, flat_node2pin_map,
                 flat_node2pin_start_map, num_physical_nodes, algorithm='segment'):
        super(PinPos, self).__init__()
        self.pin_offset_x = pin_offset_x
        self.pin_offset_y = pin_offset_y
        self.pin2node_map = pin2node_map.long()
        self.flat_node2pin_map = flat_node2pin_map
        self.flat_node2pin_start_map = flat_node2pin_start_map
        self.num_physical_nodes = num_physical_nodes
        self.algorithm = algorithm

    def forward(self, pos):
        assert pos.numel() % 2 == 0
        return PinPosFunction.apply(pos, self.pin_offset_x, self.pin_offset_y, self.pin2node_map,
                                    self.flat_node2pin_map, self.flat_node2pin_start_map,
                                    self.num_physical_nodes)
