"""GPU parity of the density scatter, fixed-cell map, overflow and field gather
against the oracle.  Integer work is bit-exact; the gather is compared within
a floating-point tolerance stated per test."""
import numpy as np
import pytest
import torch

from _util import density_params, design, fine_grid, off_grid, to_dev

pytestmark = pytest.mark.gpu

CASES = [
    (0, {}, "config0"),
    # an odd bin count: the map clear's 16-byte stores leave one 8-byte tail
    (0, {"bins": 45}, "config0-45x45bins"),
    (1, {"scale": 0.002}, "config1-small"),
    (4, {"scale": 0.004, "max_fillers": 200000}, "config4-small-macros"),
]


def _op(d, t, dtype, convention="cpu", algorithm="global"):
    from dreamplace_amd.electric_potential import ElectricPotential
    return ElectricPotential(
        t["node_size_x"], t["node_size_y"], t["bin_center_x"], t["bin_center_y"],
        d.target_density, d.xl, d.yl, d.xh, d.yh, d.bin_size_x, d.bin_size_y,
        d.num_movable_nodes, d.num_terminals, d.num_filler_nodes, 0, True,
        t["sorted_node_map"], movable_macro_mask=t["movable_macro_mask"],
        convention=convention, algorithm=algorithm)


@pytest.mark.parametrize("algorithm", ["global", "tiled"])
@pytest.mark.parametrize("dtype", [np.float64, np.float32], ids=["f64", "f32"])
@pytest.mark.parametrize("convention", ["cpu", "gpu"])
@pytest.mark.parametrize("case", CASES, ids=[c[2] for c in CASES])
def test_density_map_bit_exact(oracle, gpu, case, dtype, convention, algorithm):
    cfg, kw, name = case
    print("case density_map %s %s %s %s" % (name, np.dtype(dtype).name, convention, algorithm),
          flush=True)
    d = design(cfg, dtype=dtype, **kw)
    t = to_dev(d, gpu)
    op = _op(d, t, t["pos"].dtype, convention, algorithm)
    op.compute_initial_density_map(t["pos"])
    raw = torch.empty(d.num_bins_x * d.num_bins_y, dtype=torch.int64, device=gpu)
    dmap = op.density_map(t["pos"], fixed_map_out=raw)
    torch.cuda.synchronize()
    dp = density_params(oracle, d)
    init = oracle.fixed_density_map(dp, d.pos, d.bin_center_x, d.bin_center_y)
    assert np.array_equal(op.initial_density_map.cpu().numpy(), init), "fixed map differs"
    ref_raw = oracle.density_fixed_sum(dp, d.pos)
    assert np.array_equal(raw.cpu().numpy(), ref_raw), "int64 fixed-point sum differs"
    ref_map = oracle.density_map(dp, d.pos, init, convention=convention)
    assert np.array_equal(dmap.cpu().numpy(), ref_map), "converted map differs"


def test_density_map_order_free(oracle, gpu):
    """Permuting movables gives a bit-identical int64 map (any launch order)."""
    print("case density_map permutation", flush=True)
    d = design(0)
    t = to_dev(d, gpu)
    op = _op(d, t, torch.float64)
    op.compute_initial_density_map(t["pos"])
    raw0 = torch.empty(d.num_bins_x * d.num_bins_y, dtype=torch.int64, device=gpu)
    op.density_map(t["pos"], fixed_map_out=raw0)
    perm = np.random.default_rng(1).permutation(d.num_movable_nodes)
    n = d.num_nodes
    idx = np.concatenate([perm, np.arange(d.num_movable_nodes, n)])
    d2 = dict(d)
    for k in ("node_size_x", "node_size_y"):
        d2[k] = d[k][idx]
    pos2 = np.concatenate([d.pos[:n][idx], d.pos[n:][idx]])
    t2 = {k: torch.from_numpy(np.ascontiguousarray(v)).to(gpu) for k, v in
          [("node_size_x", d2["node_size_x"]), ("node_size_y", d2["node_size_y"]), ("pos", pos2)]}
    t3 = dict(t)
    t3.update(t2)
    op2 = _op(d, t3, torch.float64)
    op2.initial_density_map = op.initial_density_map
    raw1 = torch.empty_like(raw0)
    op2.density_map(t3["pos"], fixed_map_out=raw1)
    assert torch.equal(raw0, raw1)


@pytest.mark.parametrize("dtype", [np.float64, np.float32], ids=["f64", "f32"])
def test_overflow_metric(oracle, gpu, dtype):
    print("case overflow %s" % np.dtype(dtype).name, flush=True)
    d = design(0, dtype=dtype)
    t = to_dev(d, gpu)
    op = _op(d, t, t["pos"].dtype)
    cost, mx = op(t["pos"], mode="overflow")
    dp = density_params(oracle, d)
    init = oracle.fixed_density_map(dp, d.pos, d.bin_center_x, d.bin_center_y)
    m = oracle.density_map(dp, d.pos, init, num_filler_nodes=0)
    a = d.bin_size_x * d.bin_size_y
    ref_cost = np.sum(np.maximum(m.astype(np.float64) - d.target_density * a, 0))
    assert abs(cost.item() - ref_cost) <= 1e-6 * ref_cost
    assert abs(mx.item() - m.max() / a) <= 1e-6 * (m.max() / a)


@pytest.mark.parametrize("algorithm", ["global", "tiled"])
@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-12), (np.float32, 2e-5)], ids=["f64", "f32"])
@pytest.mark.parametrize("case", CASES, ids=[c[2] for c in CASES])
def test_electric_force(oracle, gpu, case, dtype, tol, algorithm):
    cfg, kw, name = case
    print("case electric_force %s %s %s" % (name, np.dtype(dtype).name, algorithm), flush=True)
    d = design(cfg, dtype=dtype, **kw)
    t = to_dev(d, gpu)
    op = _op(d, t, t["pos"].dtype, algorithm=algorithm)
    # the tiled gather reuses the binning of the density call at the same positions
    op.compute_initial_density_map(t["pos"])
    op.density_map(t["pos"])
    rng = np.random.default_rng(7)
    fx = rng.standard_normal((d.num_bins_x, d.num_bins_y)).astype(dtype)
    fy = rng.standard_normal((d.num_bins_x, d.num_bins_y)).astype(dtype)
    g = torch.tensor(0.75, dtype=t["pos"].dtype, device=gpu)
    out = op.electric_force(t["pos"], torch.from_numpy(fx).to(gpu), torch.from_numpy(fy).to(gpu), g)
    dp = density_params(oracle, d)
    ref = oracle.electric_force(dp, d.pos, fx, fy, grad=0.75)
    got = out.cpu().numpy()
    scale = np.abs(ref).max()
    assert np.max(np.abs(got - ref)) <= tol * scale
    nm, nt = d.num_movable_nodes, d.num_terminals
    n = d.num_nodes
    assert np.all(got[nm:nm + nt] == 0) and np.all(got[n + nm:n + nm + nt] == 0)


@pytest.mark.parametrize("algorithm", ["global", "tiled"])
@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-12), (np.float32, 2e-5)], ids=["f64", "f32"])
def test_off_grid_cells(oracle, gpu, dtype, tol, algorithm):
    """Cells past the die edges (no bin, or a truncated index naming an edge
    bin): bit-exact map, force as the oracle, zero force where no bin."""
    print("case off_grid %s %s" % (np.dtype(dtype).name, algorithm), flush=True)
    d = off_grid(design(0, dtype=dtype))
    t = to_dev(d, gpu)
    op = _op(d, t, t["pos"].dtype, algorithm=algorithm)
    op.compute_initial_density_map(t["pos"])
    raw = torch.empty(d.num_bins_x * d.num_bins_y, dtype=torch.int64, device=gpu)
    dmap = op.density_map(t["pos"], fixed_map_out=raw)
    dp = density_params(oracle, d)
    init = oracle.fixed_density_map(dp, d.pos, d.bin_center_x, d.bin_center_y)
    assert np.array_equal(raw.cpu().numpy(), oracle.density_fixed_sum(dp, d.pos))
    assert np.array_equal(dmap.cpu().numpy(), oracle.density_map(dp, d.pos, init))
    rng = np.random.default_rng(5)
    fx = rng.standard_normal((d.num_bins_x, d.num_bins_y)).astype(dtype)
    fy = rng.standard_normal((d.num_bins_x, d.num_bins_y)).astype(dtype)
    g = torch.tensor(1.5, dtype=t["pos"].dtype, device=gpu)
    out = op.electric_force(t["pos"], torch.from_numpy(fx).to(gpu), torch.from_numpy(fy).to(gpu), g)
    ref = oracle.electric_force(dp, d.pos, fx, fy, grad=1.5)
    got = out.cpu().numpy()
    assert np.max(np.abs(got - ref)) <= tol * np.abs(ref).max()
    assert np.all(got[ref == 0] == 0)


@pytest.mark.parametrize("algorithm", ["global", "tiled"])
@pytest.mark.parametrize("dtype", [np.float64, np.float32], ids=["f64", "f32"])
@pytest.mark.parametrize("case", CASES, ids=[c[2] for c in CASES])
def test_electric_force_interleaved_layout(gpu, case, dtype, algorithm):
    """The interleaved gradient is the split one, word for word."""
    cfg, kw, name = case
    print("case force layout %s %s %s" % (name, np.dtype(dtype).name, algorithm), flush=True)
    d = design(cfg, dtype=dtype, **kw)
    t = to_dev(d, gpu)
    op = _op(d, t, t["pos"].dtype, algorithm=algorithm)
    op.compute_initial_density_map(t["pos"])
    op.density_map(t["pos"])
    rng = np.random.default_rng(9)
    fx = torch.from_numpy(rng.standard_normal((d.num_bins_x, d.num_bins_y)).astype(dtype)).to(gpu)
    fy = torch.from_numpy(rng.standard_normal((d.num_bins_x, d.num_bins_y)).astype(dtype)).to(gpu)
    g = torch.tensor(0.5, dtype=t["pos"].dtype, device=gpu)
    split = op.electric_force(t["pos"], fx, fy, g)
    inter = op.electric_force(t["pos"], fx, fy, g, grad_layout="interleaved")
    assert torch.equal(inter.view(-1, 2).t().reshape(-1), split)


@pytest.mark.parametrize("algorithm", ["global", "tiled"])
@pytest.mark.parametrize("dtype", [np.float64, np.float32], ids=["f64", "f32"])
def test_uniform_filler_path_identical(gpu, dtype, algorithm):
    """Taking the fillers' shared size as scalars gives the identical int64 map
    and force as reading every filler's arrays."""
    print("case uniform filler %s %s" % (np.dtype(dtype).name, algorithm), flush=True)
    d = off_grid(design(1, dtype=dtype, scale=0.002))
    t = to_dev(d, gpu)
    op = _op(d, t, t["pos"].dtype, algorithm=algorithm)
    assert op.uniform_filler is not None
    op.compute_initial_density_map(t["pos"])
    rng = np.random.default_rng(3)
    fx = torch.from_numpy(rng.standard_normal((d.num_bins_x, d.num_bins_y)).astype(dtype)).to(gpu)
    fy = torch.from_numpy(rng.standard_normal((d.num_bins_x, d.num_bins_y)).astype(dtype)).to(gpu)
    g = torch.tensor(0.5, dtype=t["pos"].dtype, device=gpu)
    out = []
    for uni in (True, False):
        op.use_uniform_filler = uni
        raw = torch.empty(d.num_bins_x * d.num_bins_y, dtype=torch.int64, device=gpu)
        m = op.density_map(t["pos"], fixed_map_out=raw)
        out.append((raw, m, op.electric_force(t["pos"], fx, fy, g)))
    for a, b in zip(*out):
        assert torch.equal(a, b)


def test_backward_after_rebinning(gpu):
    """The tiled gather reuses the forward's cell binning; a density call at
    other positions between forward and backward must not change the
    gradient, nor must an overflow-mode call."""
    print("case backward after rebinning", flush=True)
    d = design(0)
    t = to_dev(d, gpu)
    op = _op(d, t, torch.float64, algorithm="tiled")
    pos = t["pos"].clone().requires_grad_(True)
    op(pos).backward()
    ref = pos.grad.clone()
    other = t["pos"] + 0.37
    for interleave in ("density", "overflow"):
        pos.grad = None
        e = op(pos)
        if interleave == "density":
            op(other)
        else:
            op(other, mode="overflow")
        e.backward()
        assert torch.equal(pos.grad, ref), interleave


@pytest.mark.parametrize("arith", ["reference", "bin_relative"])
@pytest.mark.parametrize("dtype", [np.float64, np.float32], ids=["f64", "f32"])
def test_tiled_at_the_tile_limit(oracle, gpu, dtype, arith):
    """4096 x 4096 bins, 4096 tiles (the tiled kernels' limit), 3.3M cells
    shrunk to span at most 3 x 3 bins: the tiled scatter's int64 sum and
    converted map are the oracle's bits and the tiled gather is the oracle's
    force within rounding."""
    from dreamplace_amd.electric_potential import ElectricPotential
    print("case tiled density at 4096 tiles %s %s" % (np.dtype(dtype).name, arith), flush=True)
    d = fine_grid(design(1, dtype=dtype, scale=0.25), 4096, 0.2)
    t = to_dev(d, gpu)
    op = ElectricPotential(
        t["node_size_x"], t["node_size_y"], t["bin_center_x"], t["bin_center_y"],
        d.target_density, d.xl, d.yl, d.xh, d.yh, d.bin_size_x, d.bin_size_y,
        d.num_movable_nodes, d.num_terminals, d.num_filler_nodes, 0, True,
        t["sorted_node_map"], movable_macro_mask=t["movable_macro_mask"],
        algorithm="tiled", arithmetic=arith)
    op.compute_initial_density_map(t["pos"])
    nb = d.num_bins_x * d.num_bins_y
    raw = torch.empty(nb, dtype=torch.int64, device=gpu)
    dmap = op.density_map(t["pos"], fixed_map_out=raw)
    torch.cuda.synchronize()
    dp = density_params(oracle, d)
    kw = dict(arithmetic=arith)
    assert np.array_equal(raw.cpu().numpy(), oracle.density_fixed_sum(dp, d.pos, **kw))
    init = oracle.fixed_density_map(dp, d.pos, d.bin_center_x, d.bin_center_y, **kw)
    assert np.array_equal(dmap.cpu().numpy(), oracle.density_map(dp, d.pos, init, **kw))
    rng = np.random.default_rng(11)
    fx = rng.standard_normal((d.num_bins_x, d.num_bins_y)).astype(dtype)
    fy = rng.standard_normal((d.num_bins_x, d.num_bins_y)).astype(dtype)
    g = torch.tensor(0.75, dtype=t["pos"].dtype, device=gpu)
    out = op.electric_force(t["pos"], torch.from_numpy(fx).to(gpu), torch.from_numpy(fy).to(gpu), g)
    ref = oracle.electric_force(dp, d.pos, fx, fy, grad=0.75, **kw)
    tol = 1e-12 if dtype == np.float64 else 1e-6
    assert np.max(np.abs(out.cpu().numpy() - ref)) <= tol * np.abs(ref).max()


@pytest.mark.parametrize("arith", ["reference", "bin_relative"])
def test_tiled_negative_contributions(gpu, arith):
    """Negative contributions (here from negated ratios of one movable cell in
    seven) take the FP32 tile scatter's generic conversion, the others its
    two-conversion truncation: the int64 map equals the global scatter's bit
    for bit, which converts every contribution generically."""
    from dreamplace_amd.electric_potential import ElectricPotential
    print("case tiled negative contributions %s" % arith, flush=True)
    d = design(1, dtype=np.float32, scale=0.002)
    t = to_dev(d, gpu)
    raws = []
    for algorithm, negate in (("tiled", False), ("global", True), ("tiled", True)):
        op = ElectricPotential(
            t["node_size_x"], t["node_size_y"], t["bin_center_x"], t["bin_center_y"],
            d.target_density, d.xl, d.yl, d.xh, d.yh, d.bin_size_x, d.bin_size_y,
            d.num_movable_nodes, d.num_terminals, d.num_filler_nodes, 0, True,
            t["sorted_node_map"], movable_macro_mask=t["movable_macro_mask"],
            algorithm=algorithm, arithmetic=arith)
        if negate:
            op.ratio[:d.num_movable_nodes:7] *= -1
        op.compute_initial_density_map(t["pos"])
        raw = torch.empty(d.num_bins_x * d.num_bins_y, dtype=torch.int64, device=gpu)
        op.density_map(t["pos"], fixed_map_out=raw)
        raws.append(raw)
    torch.cuda.synchronize()
    assert not torch.equal(raws[0], raws[1])  # the negated cells count
    assert torch.equal(raws[1], raws[2])
