"""dp_density_map_noalloc (the allocation-free density map, scratch in a
file-scope device array) against dp_density_map with a caller workspace:
bit-identical maps and fixed-point sums on two calls in a row, the second
with a smaller design and grid, in both conversions and both dtypes; sizes
beyond its fixed scratch are refused before anything is launched."""
import numpy as np
import pytest
import torch

from _util import design, to_dev

pytestmark = pytest.mark.gpu


def _op(d, t, convention):
    from dreamplace_amd.electric_potential import ElectricPotential
    return ElectricPotential(
        t["node_size_x"], t["node_size_y"], t["bin_center_x"], t["bin_center_y"],
        d.target_density, d.xl, d.yl, d.xh, d.yh, d.bin_size_x, d.bin_size_y,
        d.num_movable_nodes, d.num_terminals, d.num_filler_nodes, 0, True,
        t["sorted_node_map"], movable_macro_mask=t["movable_macro_mask"],
        convention=convention, algorithm="global")


def _noalloc(op, pos, out, raw):
    from dreamplace_amd import _native as N
    N.call("dp_density_map_noalloc", N.dtype_code(pos), N.ptr(pos), pos.numel() // 2,
           N.ptr(op.node_size_x_clamped), N.ptr(op.node_size_y_clamped), N.ptr(op.offset_x),
           N.ptr(op.offset_y), N.ptr(op.ratio), N.ptr(op.initial_density_map), float(op.xl),
           float(op.yl), float(op.xh), float(op.yh), float(op.bin_size_x), float(op.bin_size_y),
           int(op.num_movable_nodes), int(op.num_filler_nodes), op.num_bins_x, op.num_bins_y,
           N.DP_CONVERT_CPU if op.convention == "cpu" else N.DP_CONVERT_GPU,
           op.uniform_filler_arg(), N.ptr(out), N.ptr(raw), N.stream_ptr(pos.device))


@pytest.mark.parametrize("dtype", [np.float64, np.float32], ids=["f64", "f32"])
@pytest.mark.parametrize("convention", ["cpu", "gpu"])
def test_noalloc_density_map_matches_workspace_form(gpu, dtype, convention):
    # the largest call first (config 0 on 1024 x 1024 bins, the limit: most
    # cells span more than 3x3 bins and go through the queue), then smaller
    # ones (config 4 at 0.4% with movable macros, config 0 on 45 x 45 bins)
    # whose maps must not see an earlier call's leftovers
    cases = [(0, {"bins": 1024}, "config0-1024x1024bins"),
             (4, {"scale": 0.004, "max_fillers": 200000}, "config4-small-macros"),
             (0, {"bins": 45}, "config0-45x45bins")]
    for cfg, kw, name in cases:
        print("case noalloc %s %s %s" % (name, np.dtype(dtype).name, convention), flush=True)
        d = design(cfg, dtype=dtype, **kw)
        t = to_dev(d, gpu)
        op = _op(d, t, convention)
        op.compute_initial_density_map(t["pos"])
        nb = d.num_bins_x * d.num_bins_y
        raw_ref = torch.empty(nb, dtype=torch.int64, device=gpu)
        ref = op.density_map(t["pos"], fixed_map_out=raw_ref)
        out = torch.full((d.num_bins_x, d.num_bins_y), float("nan"), dtype=t["pos"].dtype,
                         device=gpu)
        raw = torch.full((nb,), -7, dtype=torch.int64, device=gpu)
        _noalloc(op, t["pos"], out, raw)
        torch.cuda.synchronize()
        assert torch.equal(raw, raw_ref), name
        assert torch.equal(out, ref), name


def test_noalloc_refuses_grids_beyond_its_scratch(gpu):
    """2048 x 2048 bins exceed the fixed scratch (1024 x 1024): the call
    returns DP_ERR_UNSUPPORTED with a message, launching nothing."""
    from dreamplace_amd import _native as N
    print("case noalloc oversize refused", flush=True)
    d = design(0, dtype=np.float64)
    t = to_dev(d, gpu)
    op = _op(d, t, "cpu")
    op.compute_initial_density_map(t["pos"])
    pos = t["pos"]
    out = torch.empty(8, dtype=pos.dtype, device=gpu)
    rc = N.lib().dp_density_map_noalloc(
        N.dtype_code(pos), N.ptr(pos), pos.numel() // 2, N.ptr(op.node_size_x_clamped),
        N.ptr(op.node_size_y_clamped), N.ptr(op.offset_x), N.ptr(op.offset_y), N.ptr(op.ratio),
        None, float(op.xl), float(op.yl), float(op.xh), float(op.yh), float(op.bin_size_x),
        float(op.bin_size_y), int(op.num_movable_nodes), int(op.num_filler_nodes), 2048, 2048,
        N.DP_CONVERT_CPU, None, N.ptr(out), None, N.stream_ptr(gpu))
    assert rc == -3  # DP_ERR_UNSUPPORTED
    assert b"1024 x 1024" in N.lib().dp_last_error()
