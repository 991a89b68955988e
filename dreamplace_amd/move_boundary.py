"""Clamp movable and filler cells into the die, in place: drop-in for
dreamplace/ops/move_boundary/move_boundary.py (``MoveBoundary``,
move_boundary.py:34-63)."""
from . import _native as N


class MoveBoundaryFunction:
    @staticmethod
    def forward(pos, node_size_x, node_size_y, xl, yl, xh, yh, num_movable_nodes, num_filler_nodes):
        N.require_gpu(pos, node_size_x, name="MoveBoundary")
        flat = pos.data.view(-1)
        N.require_contiguous(flat, name="MoveBoundary")
        N.call("dp_move_boundary", N.dtype_code(flat), N.ptr(flat), flat.numel() // 2,
               N.ptr(node_size_x), N.ptr(node_size_y), float(xl), float(yl), float(xh), float(yh),
               int(num_movable_nodes), int(num_filler_nodes), N.stream_ptr(flat.device))
        return pos


class MoveBoundary(object):
    def __init__(self, node_size_x, node_size_y, xl, yl, xh, yh, num_movable_nodes, num_filler_nodes):
        super(MoveBoundary, self).__init__()
        self.node_size_x = node_size_x
        self.node_size_y = node_size_y
        self.xl = xl
        self.yl = yl
        self.xh = xh
        self.yh = yh
        self.num_movable_nodes = num_movable_nodes
        self.num_filler_nodes = num_filler_nodes

    def forward(self, pos):
        return MoveBoundaryFunction.forward(pos, node_size_x=self.node_size_x,
                                            node_size_y=self.node_size_y, xl=self.xl, yl=self.yl,
                                            xh=self.xh, yh=self.yh,
                                            num_movable_nodes=self.num_movable_nodes,
                                            num_filler_nodes=self.num_filler_nodes)

    def __call__(self, pos):
        return self.forward(pos)
