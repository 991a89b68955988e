"""Host-side logic without a GPU: the synthetic designs follow the reference's
formulas, the op modules prepare the reference's derived arrays, and every op
refuses CPU tensors instead of falling back."""
import math

import numpy as np
import pytest
import torch

from dreamplace_amd import synth


@pytest.mark.parametrize("cfg,kw", [(0, {}), (1, {"scale": 0.003}), (3, {"scale": 0.001}),
                                    (4, {"scale": 0.003, "max_fillers": 20000})])
def test_design_invariants(cfg, kw):
    d = synth.make_design(cfg, **kw)
    n, P, E = d.num_nodes, d.num_pins, d.num_nets
    assert d.pos.shape == (2 * n,) and d.node_size_x.shape == (n,)
    assert d.num_movable_nodes + d.num_terminals + d.num_filler_nodes == n
    # CSR maps are consistent
    assert d.flat_net2pin_start_map[-1] == P and np.array_equal(np.sort(d.flat_net2pin_map), np.arange(P))
    for e in range(0, E, max(1, E // 50)):
        pins = d.flat_net2pin_map[d.flat_net2pin_start_map[e]:d.flat_net2pin_start_map[e + 1]]
        assert np.all(d.pin2net_map[pins] == e)
        assert len(set(d.pin2node_map[pins])) == len(pins)  # distinct nodes per net
    assert d.flat_node2pin_start_map[-1] == P
    nodes = np.repeat(np.arange(d.num_physical_nodes), np.diff(d.flat_node2pin_start_map))
    assert np.array_equal(d.pin2node_map[d.flat_node2pin_map], nodes)
    # pins inside their nodes
    assert np.all(d.pin_offset_x >= 0) and np.all(d.pin_offset_x <= d.node_size_x[d.pin2node_map])
    # masks (BasicPlace.py:159-183)
    deg = np.diff(d.flat_net2pin_start_map)
    assert np.array_equal(d.net_mask_all, (deg >= 2).astype(np.uint8))
    assert np.array_equal(d.net_mask_ignore_large_degrees, ((deg >= 2) & (deg < 100)).astype(np.uint8))
    assert np.array_equal(d.pin_mask_ignore_fixed_macros, d.pin2node_map >= d.num_movable_nodes)
    # fixed blocks do not overlap each other
    m0, m1 = d.num_movable_nodes, d.num_movable_nodes + d.num_terminals
    fx, fy = d.pos[m0:m1], d.pos[n + m0:n + m1]
    fw, fh = d.node_size_x[m0:m1], d.node_size_y[m0:m1]
    if cfg != 4 and m1 > m0:
        k = min(300, m1 - m0)
        ox = np.minimum(fx[:k, None] + fw[:k, None], fx[None, :k] + fw[None, :k]) - np.maximum(fx[:k, None], fx[None, :k])
        oy = np.minimum(fy[:k, None] + fh[:k, None], fy[None, :k] + fh[None, :k]) - np.maximum(fy[:k, None], fy[None, :k])
        over = (ox > 1e-9) & (oy > 1e-9)
        np.fill_diagonal(over, False)
        assert not over.any()


def test_filler_formula():
    """PlaceDB.py:917-936: filler width = mean of the 5%..95% width band,
    count = round(max(space*t - movable cell area, 0) / filler area)."""
    d = synth.make_design(0)
    w = d.node_size_x[:d.num_movable_nodes]
    order = np.argsort(w, kind="stable")
    lb, ub = int(len(w) * 0.05), int(len(w) * 0.95)
    fw = np.mean(w[order[lb:ub]])
    assert d.filler_size_x == pytest.approx(fw)
    space = (d.xh - d.xl) * (d.yh - d.yl) - d.total_fixed_node_area
    expect = int(round(max(space * d.target_density - (w * d.node_size_y[:d.num_movable_nodes]).sum(), 0) / fw))
    assert d.num_filler_nodes == expect


def test_degree_distribution():
    d = synth.make_design(0, seed=3)
    deg = np.diff(d.flat_net2pin_start_map)
    assert abs(np.mean(deg == 2) - 0.60) < 0.02
    assert abs(np.mean(deg == 3) - 0.20) < 0.02
    assert abs(np.mean((deg >= 4) & (deg <= 10)) - 0.15) < 0.02
    assert abs(np.mean(deg >= 11) - 0.05) < 0.01
    assert deg.max() <= 50


def test_overflow_module_reset_matches_oracle(oracle):
    """ElectricOverflow.reset derives the reference's stretched sizes, offsets
    and ratios (electric_overflow.py:162-227) -- host tensors are enough."""
    from dreamplace_amd.electric_overflow import ElectricOverflow
    d = synth.make_design(4, scale=0.003, max_fillers=20000)
    T = lambda a: torch.from_numpy(a)
    op = ElectricOverflow(T(d.node_size_x), T(d.node_size_y), T(d.bin_center_x), T(d.bin_center_y),
                          d.target_density, d.xl, d.yl, d.xh, d.yh, d.bin_size_x, d.bin_size_y,
                          d.num_movable_nodes, d.num_terminals, d.num_filler_nodes, 0, True,
                          T(d.sorted_node_map), movable_macro_mask=T(d.movable_macro_mask))
    dp = oracle.DensityParams(d.node_size_x, d.node_size_y, d.xl, d.yl, d.xh, d.yh, d.bin_size_x,
                              d.bin_size_y, d.num_movable_nodes, d.num_terminals, d.num_filler_nodes,
                              d.target_density, d.movable_macro_mask)
    for a, b in [(op.node_size_x_clamped, dp.wxc), (op.offset_x, dp.offx), (op.ratio, dp.ratio)]:
        assert np.array_equal(a.numpy(), b)
    assert op.num_bins_x == d.num_bins_x
    big = (dp.wxc > 1.9 * d.bin_size_x) | (dp.wyc > 1.9 * d.bin_size_y)
    assert op.large_capacity == int(big.sum()) + 64
    with pytest.raises(RuntimeError, match="GPU"):
        op.compute_initial_density_map(T(d.pos))


def test_uniform_filler_detection():
    """reset() hands the fillers' shared clamped size, offset and ratio to the
    kernels only when every filler has them."""
    from dreamplace_amd.electric_overflow import ElectricOverflow
    d = synth.make_design(0)
    T = lambda a: torch.from_numpy(np.ascontiguousarray(a))

    def build(sx):
        return ElectricOverflow(T(sx), T(d.node_size_y), T(d.bin_center_x), T(d.bin_center_y),
                                d.target_density, d.xl, d.yl, d.xh, d.yh, d.bin_size_x,
                                d.bin_size_y, d.num_movable_nodes, d.num_terminals,
                                d.num_filler_nodes, 0, True, T(d.sorted_node_map))
    op = build(d.node_size_x)
    nf = d.num_filler_nodes
    assert op.uniform_filler == [float(op.node_size_x_clamped[-1]), float(op.node_size_y_clamped[-1]),
                                 float(op.offset_x[-1]), float(op.offset_y[-1]), float(op.ratio[-1])]
    sx = d.node_size_x.copy()
    sx[-nf // 2] *= 1.5
    assert build(sx).uniform_filler is None


def test_ops_refuse_cpu_tensors():
    from dreamplace_amd.hpwl import hpwl_fp64
    from dreamplace_amd.pin_pos import PinPos
    d = synth.make_design(0)
    T = lambda a: torch.from_numpy(a)
    pp = PinPos(T(d.pin_offset_x), T(d.pin_offset_y), T(d.pin2node_map), T(d.flat_node2pin_map),
                T(d.flat_node2pin_start_map), d.num_physical_nodes)
    with pytest.raises(RuntimeError, match="GPU"):
        pp(T(d.pos))
    with pytest.raises(RuntimeError, match="GPU"):
        hpwl_fp64(T(d.pos), T(d.flat_net2pin_map), T(d.flat_net2pin_start_map), T(d.net_weights),
                  T(d.net_mask_all))


def test_fence_regions_not_supported():
    from dreamplace_amd.electric_potential import ElectricPotential
    d = synth.make_design(0)
    T = lambda a: torch.from_numpy(a)
    with pytest.raises(NotImplementedError):
        ElectricPotential(T(d.node_size_x), T(d.node_size_y), T(d.bin_center_x), T(d.bin_center_y),
                          1.0, d.xl, d.yl, d.xh, d.yh, d.bin_size_x, d.bin_size_y,
                          d.num_movable_nodes, d.num_terminals, d.num_filler_nodes, 0, True,
                          T(d.sorted_node_map), region_id=0)


def _pack(start):
    import ctypes
    from dreamplace_amd import _native as N
    E = len(start) - 1
    start = np.ascontiguousarray(start, dtype=np.int32)
    gs = np.zeros(E + 1, dtype=np.int32)
    big = np.zeros(4 * max(E, 1), dtype=np.int32)
    slot = np.zeros(max(E, 1), dtype=np.int32)
    ng, nb = ctypes.c_int64(0), ctypes.c_int64(0)
    vp = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    N.call("dp_wa_pack_nets", vp(start), E, vp(gs), ctypes.byref(ng), vp(big), ctypes.byref(nb),
           vp(slot))
    return gs[:ng.value + 1], big[:4 * nb.value].reshape(-1, 4), slot[:E]


@pytest.mark.parametrize("seed", [0, 1])
def test_wa_pack_nets_layout(seed):
    """Host packing (no GPU work): groups of <= 64 pins made of whole nets,
    every net's pins contiguous and every slot owned exactly once, nets of
    more than 64 pins after the groups, and the packing nearly full."""
    rng = np.random.default_rng(seed)
    deg = synth._degrees(rng, 20000, 200)
    deg[rng.choice(len(deg), 40, replace=False)] = rng.choice([0, 1, 64, 65, 130], 40)
    start = np.concatenate([[0], np.cumsum(deg)]).astype(np.int32)
    gs, big, slot = _pack(start)
    P = int(start[-1])
    owner = np.full(P, -1)
    for e, d in enumerate(deg):
        if d > 0:
            assert np.all(owner[slot[e]:slot[e] + d] == -1)
            owner[slot[e]:slot[e] + d] = e
    assert np.all(owner >= 0)                              # every slot owned exactly once
    small = int(deg[(deg >= 1) & (deg <= 64)].sum())
    assert gs[0] == 0 and gs[-1] == small and np.all(np.diff(gs) >= 1) and np.all(np.diff(gs) <= 64)
    for g in range(len(gs) - 1):                           # groups hold whole nets
        nets = set(owner[gs[g]:gs[g + 1]])
        assert all(slot[e] >= gs[g] and slot[e] + deg[e] <= gs[g + 1] for e in nets)
    assert sorted(big[:, 0]) == sorted(np.nonzero(deg > 64)[0])
    assert np.all(big[:, 2] == deg[big[:, 0]]) and np.all(big[:, 1] == slot[big[:, 0]])
    assert np.all(big[:, 1] >= small)
    assert len(gs) - 1 <= math.ceil(small / 64 * 1.02) + 1    # best-fit decreasing packs tightly


def test_netlist_check_rejects_bad_maps():
    from dreamplace_amd.engine import check_netlist
    d = synth.make_design(0)
    args = [d.flat_net2pin_map, d.flat_net2pin_start_map, d.pin2net_map, d.flat_node2pin_map,
            d.flat_node2pin_start_map, d.num_pins, d.num_physical_nodes]
    check_netlist(*args)
    bad = list(args)
    bad[0] = d.flat_net2pin_map.copy()
    bad[0][0] = bad[0][1]
    with pytest.raises(ValueError):
        check_netlist(*bad)
    bad = list(args)
    bad[2] = d.pin2net_map.copy()
    bad[2][d.flat_net2pin_map[0]] += 1
    with pytest.raises(ValueError):
        check_netlist(*bad)
    bad = list(args)
    bad[4] = d.flat_node2pin_start_map.copy()
    bad[4][-1] -= 1
    with pytest.raises(ValueError):
        check_netlist(*bad)


def test_placer_control_logic_matches_oracle():
    """The placer's plateau / divergence tests and density-weight factor agree
    with the oracle's restatement on random histories."""
    from dreamplace_amd import placer as P
    from oracle import placer as OP
    rng = np.random.default_rng(0)
    for k in range(200):
        n = int(rng.integers(10, 80))
        base = rng.uniform(0.1, 0.5)
        ov = list(base + rng.uniform(-0.05, 0.05, n) * rng.choice([0.0, 0.01, 1.0]))
        wl = list(rng.uniform(90, 200, n))
        pairs = list(zip(wl, ov))
        best = dict(hpwl=float(rng.uniform(80, 150)), overflow=float(rng.uniform(0.1, 0.5)))
        thr = float(rng.uniform(0.02, 0.4))
        assert P.check_plateau([1] + ov, 15, 0.001) == OP.plateau([1] + ov, 15, 0.001)
        assert bool(P.check_divergence(pairs, best, 0.1, 50, thr)) == OP.diverging(pairs, best, 0.1, 50, thr)
    p = dict(P.DEFAULTS)
    for cur, prev, it in [(90.0, 100.0, 3), (100.0, 90.0, 7), (1e6, 1.0, 2), (1.0, 1.0, 500)]:
        mu = P.density_weight_factor(cur, prev, it, p)
        assert p["upper_pcof"] * p["lower_pcof"] <= mu <= p["upper_pcof"] ** 2
    assert P.density_weight_factor(1e9, 0.0, 1, p) == pytest.approx(p["upper_pcof"] * p["lower_pcof"])
    assert P.density_weight_factor(0.0, 1.0, 1, p) == pytest.approx(p["upper_pcof"] * 0.9999)


def test_filler_tile_order_on_host():
    """DeviceDesign.reorder_fillers on CPU tensors: a permutation of the
    fillers only, tile keys non-decreasing afterwards, and the generation
    order recovered through any number of relabellings."""
    import torch
    from dreamplace_amd.engine import DeviceDesign
    d = synth.make_design(0)
    dd = DeviceDesign(d, "cpu")
    n, nf = d.num_nodes, d.num_filler_nodes
    pos0 = dd.pos.clone()
    perm = dd.reorder_fillers()
    assert sorted(perm.tolist()) == list(range(nf))
    f0 = n - nf
    assert torch.equal(dd.pos[:f0], pos0[:f0]) and torch.equal(dd.pos[n:n + f0], pos0[n:n + f0])
    bx = torch.floor((dd.pos[f0:n] - d.xl) / d.bin_size_x).clamp(0, d.num_bins_x - 1).long() // 64
    by = torch.floor((dd.pos[n + f0:] - d.yl) / d.bin_size_y).clamp(0, d.num_bins_y - 1).long() // 64
    key = bx * ((d.num_bins_y + 63) // 64) + by
    assert bool((key[1:] >= key[:-1]).all())
    assert torch.equal(dd.generation_order(dd.pos), pos0)
    assert torch.equal(dd.current_order(pos0), dd.pos)
    dd.reorder_fillers(torch.randperm(nf, generator=torch.Generator().manual_seed(3)))
    assert torch.equal(dd.generation_order(dd.pos), pos0)
    assert torch.equal(dd.current_order(pos0), dd.pos)


def test_bench_forms_share_one_placement(oracle):
    """bench.py's set-up: the timed FP32 design is relabelled (movable cells
    in netlist order, fillers in tile order) and the FP64 form and the CPU
    baseline are built from its host_design.  All three must be one
    placement: the same positions, sizes and netlist in the same order, so
    the objective is the same number and each cell's gradient moves with it.
    (Relabelling the positions alone, or relabelling twice, fails here.)"""
    import bench
    from dreamplace_amd.engine import DeviceDesign
    d = synth.make_design(0, dtype=np.float32)
    dd = DeviceDesign(d, "cpu")
    bench.relabel(dd, "netlist")
    bench.relabel(dd, None, (0.3, -0.2))
    h = dd.host_design(d)
    dd64 = DeviceDesign(h, "cpu", dtype=torch.float64)
    assert torch.equal(dd64.pos, dd.pos.double())
    for k in ("node_size_x", "node_size_y", "pin_offset_x", "pin_offset_y"):
        assert torch.equal(dd64.t[k], dd.t[k].double()), k
    for k in ("pin2node_map", "flat_node2pin_map", "flat_node2pin_start_map", "flat_net2pin_map",
              "pin2net_map", "sorted_node_map"):
        assert torch.equal(dd64.t[k], dd.t[k]), k
        assert h[k].dtype == d[k].dtype, k
    # the relabelling is real: positions, pin owners and node CSR all moved
    assert not np.array_equal(h["pos"], d["pos"])
    assert not np.array_equal(h["pin2node_map"], d["pin2node_map"])
    f0, g0, p0 = oracle.objective_and_gradient(d, 0.5, d.wa_gamma)
    f1, g1, p1 = oracle.objective_and_gradient(h, 0.5, d.wa_gamma)
    assert f1 == f0
    assert np.array_equal(p1["density_map"], p0["density_map"])
    assert np.array_equal(g1, dd.current_order(torch.from_numpy(g0)).numpy())
    assert np.array_equal(dd.generation_order(torch.from_numpy(g1)).numpy(), g0)


def test_reorder_movables_without_movable_cells():
    """A design with no movable cells: relabelling them is a no-op."""
    from dreamplace_amd.engine import DeviceDesign
    d = synth.make_design(0)
    dd = DeviceDesign(d, "cpu")
    before = {k: v.clone() for k, v in dd.t.items()}
    dd.scalars["num_movable_nodes"] = 0
    dd.reorder_movables(torch.empty(0, dtype=torch.int64))
    for k, v in before.items():
        assert torch.equal(dd.t[k], v), k
    assert dd.__dict__.get("movable_perm") is None


def test_metric_tracking_gate():
    """The placer tracks its metrics inside the objective only where the
    objective can (pins schedule, tiled density within its tile limit, no
    padding, fillers of one size); otherwise the separate passes run."""
    from dreamplace_amd.engine import MAX_TILES, can_track_metrics
    assert can_track_metrics("pins", "tiled", 2048, 2048, 0, True)  # 32 x 32 tiles
    assert can_track_metrics("pins", "tiled", 64 * MAX_TILES, 64, 0, True)  # the limit itself
    assert not can_track_metrics("pins", "tiled", 64 * MAX_TILES + 1, 64, 0, True)  # one more
    assert not can_track_metrics("pins", "global", 1024, 1024, 0, True)
    assert not can_track_metrics("fused", "tiled", 1024, 1024, 0, True)
    assert not can_track_metrics("pins", "tiled", 1024, 1024, 2, True)
    assert not can_track_metrics("pins", "tiled", 1024, 1024, 0, False)
