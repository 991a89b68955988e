"""Shared helpers of the tests: designs as device tensors, oracle density params."""
import numpy as np
import torch

from dreamplace_amd import synth


def design(config=0, seed=0, dtype=np.float64, **kw):
    return synth.make_design(config, seed=seed, dtype=dtype, **kw)


def to_dev(d, device):
    """Torch tensors of a design's arrays on `device`."""
    t = {}
    for k, v in d.items():
        if isinstance(v, np.ndarray):
            t[k] = torch.from_numpy(np.ascontiguousarray(v)).to(device)
    return t


def density_params(oracle, d):
    return oracle.DensityParams(d.node_size_x, d.node_size_y, d.xl, d.yl, d.xh, d.yh, d.bin_size_x,
                                d.bin_size_y, d.num_movable_nodes, d.num_terminals,
                                d.num_filler_nodes, d.target_density, d.movable_macro_mask)


def off_grid(d, seed=3):
    """A copy of design `d` with some movable and filler cells moved past each
    die edge: far outside (no bin at all), and just outside or straddling an
    edge, where the truncated bin index still names an edge bin (a partial or
    negative overlap, as the reference computes it)."""
    rng = np.random.default_rng(seed)
    d2 = synth.Design(d)
    n = d.num_nodes
    pos = d.pos.copy()
    cells = np.concatenate([np.arange(d.num_movable_nodes), np.arange(n - d.num_filler_nodes, n)])
    pick = rng.choice(cells, size=min(96, len(cells)), replace=False)
    bx, by = d.bin_size_x, d.bin_size_y
    for j, i in enumerate(pick):
        w, h = d.node_size_x[i], d.node_size_y[i]
        kind = j % 8
        if kind == 0:
            pos[i] = d.xl - 10 * bx - w                   # far left
        elif kind == 1:
            pos[i] = d.xl - w - 0.4 * bx                  # ends just left of the die
        elif kind == 2:
            pos[i] = d.xh + 5 * bx                        # far right
        elif kind == 3:
            pos[i] = d.xh - 0.5 * w                       # straddles the right edge
        elif kind == 4:
            pos[n + i] = d.yl - 10 * by - h               # far below
        elif kind == 5:
            pos[n + i] = d.yl - h - 0.3 * by              # ends just below the die
        elif kind == 6:
            pos[n + i] = d.yh + 7 * by                    # far above
        else:
            pos[n + i] = d.yh - 0.5 * h                   # straddles the top edge
    d2.pos = pos.astype(d.pos.dtype)
    return d2


def relabel_pins(d, seed=4):
    """A copy of design `d` whose pins carry new, random ids.  Each net's pin
    sequence (flat_net2pin_map) and each node's (flat_node2pin_map) is kept,
    only renamed, so every per-net and per-node sum runs in the same order."""
    rng = np.random.default_rng(seed)
    P = d.num_pins
    perm = rng.permutation(P).astype(np.int32)  # new id of old pin p
    d2 = synth.Design(d)
    for k in ("pin_offset_x", "pin_offset_y", "pin2node_map", "pin2net_map",
              "pin_mask_ignore_fixed_macros"):
        v = np.empty_like(d[k])
        v[perm] = d[k]
        d2[k] = v
    d2.flat_net2pin_map = perm[d.flat_net2pin_map]
    d2.flat_node2pin_map = perm[d.flat_node2pin_map]
    return d2


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def shifted(d, dx, dy):
    """A copy of design `d` translated by (dx, dy): positions, die and bin
    centres move together, so the placement problem is the same up to
    rounding, at coordinates the size of a large die's."""
    d2 = synth.Design(d)
    n = d.num_nodes
    pos = d.pos.astype(np.float64).copy()
    pos[:n] += dx
    pos[n:] += dy
    d2.pos = pos.astype(d.pos.dtype)
    d2.xl, d2.xh = d.xl + dx, d.xh + dx
    d2.yl, d2.yh = d.yl + dy, d.yh + dy
    d2.bin_center_x = (d.bin_center_x.astype(np.float64) + dx).astype(d.bin_center_x.dtype)
    d2.bin_center_y = (d.bin_center_y.astype(np.float64) + dy).astype(d.bin_center_y.dtype)
    return d2


def with_dtype(d, dtype):
    """A copy of design `d` with every floating-point array cast to `dtype`
    (FP32 -> FP64 keeps the FP32 values exactly)."""
    d2 = synth.Design(d)
    for k, v in d.items():
        if isinstance(v, np.ndarray) and v.dtype.kind == "f":
            d2[k] = v.astype(dtype)
    d2.dtype = np.dtype(dtype).name
    return d2


def fine_grid(d, nb, shrink):
    """A copy of design `d` on an nb x nb bin grid with every node (and pin
    offset) shrunk by `shrink`, so that most cells still span at most 3 x 3
    of the finer bins: the tiled density kernels' tile path at nb / 64 tiles
    per side."""
    d2 = synth.Design(d)
    dt = d.pos.dtype
    for k in ("node_size_x", "node_size_y", "pin_offset_x", "pin_offset_y"):
        d2[k] = (d[k].astype(np.float64) * shrink).astype(d[k].dtype)
    bsx, bsy = (d.xh - d.xl) / nb, (d.yh - d.yl) / nb
    d2.update(num_bins_x=nb, num_bins_y=nb, bin_size_x=bsx, bin_size_y=bsy, wa_gamma=4.0 * bsx)
    d2.bin_center_x = synth._bin_centers(d.xl, d.xh, bsx, nb).astype(dt)
    d2.bin_center_y = synth._bin_centers(d.yl, d.yh, bsy, nb).astype(dt)
    return d2
