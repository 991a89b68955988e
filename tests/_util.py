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


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))
