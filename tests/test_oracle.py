"""Pinning the oracle (CPU, no GPU):

1. against the golden vectors of the reference's own unit tests (tests/golden);
2. against the reference's own CPU ops compiled into oracle/_ref (skipped when
   the reference sources were not available to build them);
3. against what the mathematics fixes: brute-force overlaps, mass
   conservation, order independence, closed-form single-mode solutions of the
   Poisson solve, the DCT inversion identity, finite differences, the
   gamma -> 0 limit of WA, adjoint identities.
"""
import json
import math
import os

import numpy as np
import pytest
import torch

from _util import density_params, design, off_grid, rel_l2, relabel_pins
from dreamplace_amd import synth

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _gold():
    with open(os.path.join(GOLD, "ops_golden.json")) as f:
        return json.load(f)


def _csr(groups, n_items):
    flat = np.array([i for g in groups for i in g], dtype=np.int32)
    start = np.zeros(len(groups) + 1, dtype=np.int32)
    start[1:] = np.cumsum([len(g) for g in groups])
    owner = np.zeros(n_items, dtype=np.int32)
    for k, g in enumerate(groups):
        owner[g] = k
    return flat, start, owner


# ---------------------------------------------------------------------------
# 1. golden vectors of the reference's unit tests
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("dtype,atol", [(np.float64, 1e-12), (np.float32, 1e-5)])
def test_wa_golden(oracle, dtype, atol):
    g = _gold()["weighted_average_wirelength"]
    pin = np.array(g["pin_pos"], dtype=dtype)
    flat, start, _ = _csr(g["net2pin"], len(pin))
    mask = np.array([len(n) < g["ignore_net_degree"] for n in g["net2pin"]], dtype=np.uint8)
    wl, _, _ = oracle.wa_merged_forward(np.concatenate([pin[:, 0], pin[:, 1]]), flat, start,
                                        np.array(g["net_weights"], dtype=dtype), mask, g["gamma"])
    assert abs(wl - g["expected"]) <= atol


def test_hpwl_golden(oracle):
    g = _gold()["hpwl"]
    pin = np.array(g["pin_pos"])
    flat, start, _ = _csr(g["net2pin"], len(pin))
    v = oracle.hpwl(np.concatenate([pin[:, 0], pin[:, 1]]), flat, start, np.array(g["net_weights"]),
                    np.ones(2, dtype=np.uint8))
    assert v == pytest.approx(g["expected"], rel=1e-15)


def test_pin_pos_golden(oracle):
    g = _gold()["pin_pos"]
    pos = np.array(g["pos"])
    pos = np.concatenate([pos[:, 0], pos[:, 1]])
    out = oracle.pin_pos_forward(pos, g["pin_offset_x"], g["pin_offset_y"], g["pin2node"])
    assert np.allclose(out, g["expected_pin_pos"], rtol=0, atol=1e-15)
    flat, start, _ = _csr(g["node2pin"], 5)
    grad = oracle.pin_pos_backward(np.array(g["grad_pin_pos"]), flat, start, g["num_physical_nodes"], 3)
    assert np.array_equal(grad, np.array(g["expected_grad_pos"]))


# ---------------------------------------------------------------------------
# 2. against oracle/_ref (the reference's own CPU ops)
# ---------------------------------------------------------------------------
def _ref_density_inputs(d, dp, dtype):
    T = lambda a: torch.from_numpy(np.ascontiguousarray(a, dtype=dtype))
    return dict(pos=T(d.pos), wxc=T(dp.wxc), wyc=T(dp.wyc), offx=T(dp.offx), offy=T(dp.offy),
                ratio=T(dp.ratio), bcx=T(d.bin_center_x), bcy=T(d.bin_center_y))


@pytest.mark.parametrize("dtype", [np.float64, np.float32], ids=["f64", "f32"])
@pytest.mark.parametrize("cfg,kw", [(0, {}), (1, {"scale": 0.002}), (4, {"scale": 0.002, "max_fillers": 50000})],
                         ids=["config0", "config1-small", "config4-small"])
def test_density_maps_match_reference(oracle, ref, cfg, kw, dtype):
    d = design(cfg, dtype=dtype, **kw)
    dp = density_params(oracle, d)
    t = _ref_density_inputs(d, dp, dtype)
    ep = ref["electric_potential_cpp"]
    nm, nt = d.num_movable_nodes, d.num_terminals
    init_ref = ep.fixed_density_map(t["pos"], torch.from_numpy(d.node_size_x), torch.from_numpy(d.node_size_y),
                                    t["bcx"], t["bcy"], d.xl, d.yl, d.xh, d.yh, d.bin_size_x, d.bin_size_y,
                                    nm, nt, d.num_bins_x, d.num_bins_y, 3, 3, 1)
    init_ref.mul_(d.target_density)
    init = oracle.fixed_density_map(dp, d.pos, d.bin_center_x, d.bin_center_y)
    assert np.array_equal(init, init_ref.numpy())
    m_ref = ep.density_map(t["pos"], t["wxc"], t["wyc"], t["offx"], t["offy"], t["ratio"], t["bcx"], t["bcy"],
                           torch.from_numpy(init), d.target_density, d.xl, d.yl, d.xh, d.yh, d.bin_size_x,
                           d.bin_size_y, nm, d.num_filler_nodes, 0, d.num_bins_x, d.num_bins_y, 3, 3, 3, 3, 1)
    m = oracle.density_map(dp, d.pos, init)
    assert np.array_equal(m.reshape(-1), m_ref.numpy().reshape(-1))


def test_density_84_node_fixture_matches_reference(oracle, ref):
    """The reference unit test's 84-node fixture (1 movable, 83 fixed, 1024^2 bins)."""
    f = np.load(os.path.join(GOLD, "electric_potential_84.npz"))
    nm = int(f["num_movable_nodes"])
    n = len(f["x"])
    xl, yl, xh, yh = (float(f[k]) for k in ("xl", "yl", "xh", "yh"))
    nb = int(f["num_bins_x"])
    bsx, bsy = (xh - xl) / nb, (yh - yl) / nb
    bcx = oracle.bin_centers(xl, xh, bsx)
    bcy = oracle.bin_centers(yl, yh, bsy)
    dp = oracle.DensityParams(f["node_size_x"], f["node_size_y"], xl, yl, xh, yh, bsx, bsy, nm, n - nm, 0,
                              float(f["target_density"]))
    pos = np.concatenate([f["x"], f["y"]])
    init = oracle.fixed_density_map(dp, pos, bcx, bcy)
    ep = ref["electric_potential_cpp"]
    init_ref = ep.fixed_density_map(torch.from_numpy(pos), torch.from_numpy(f["node_size_x"]),
                                    torch.from_numpy(f["node_size_y"]), torch.from_numpy(bcx),
                                    torch.from_numpy(bcy), xl, yl, xh, yh, bsx, bsy, nm, n - nm, nb, nb, 3, 3, 1)
    init_ref.mul_(float(f["target_density"]))
    assert np.array_equal(init, init_ref.numpy())
    t = {k: torch.from_numpy(np.ascontiguousarray(v)) for k, v in
         dict(wxc=dp.wxc, wyc=dp.wyc, offx=dp.offx, offy=dp.offy, ratio=dp.ratio).items()}
    m_ref = ep.density_map(torch.from_numpy(pos), t["wxc"], t["wyc"], t["offx"], t["offy"], t["ratio"],
                           torch.from_numpy(bcx), torch.from_numpy(bcy), torch.from_numpy(init),
                           float(f["target_density"]), xl, yl, xh, yh, bsx, bsy, nm, 0, 0, nb, nb, 3, 3, 3, 3, 1)
    m = oracle.density_map(dp, pos, init, num_filler_nodes=0)
    assert np.array_equal(m.reshape(-1), m_ref.numpy().reshape(-1))


@pytest.mark.parametrize("dtype", [np.float64, np.float32], ids=["f64", "f32"])
def test_electric_force_matches_reference(oracle, ref, dtype):
    d = design(0, dtype=dtype)
    dp = density_params(oracle, d)
    t = _ref_density_inputs(d, dp, dtype)
    rng = np.random.default_rng(11)
    fx = rng.standard_normal((d.num_bins_x, d.num_bins_y)).astype(dtype)
    fy = rng.standard_normal((d.num_bins_x, d.num_bins_y)).astype(dtype)
    g = torch.tensor([0.5], dtype=torch.from_numpy(fx).dtype)
    out = ref["electric_potential_cpp"].electric_force(
        g, d.num_bins_x, d.num_bins_y, 3, 3, 3, 3, torch.from_numpy(fx).view(-1), torch.from_numpy(fy).view(-1),
        t["pos"], t["wxc"], t["wyc"], t["offx"], t["offy"], t["ratio"], t["bcx"], t["bcy"], d.xl, d.yl, d.xh,
        d.yh, d.bin_size_x, d.bin_size_y, d.num_movable_nodes, d.num_filler_nodes)
    mine = oracle.electric_force(dp, d.pos, fx, fy, grad=0.5)
    assert np.array_equal(mine, -out.numpy())


@pytest.mark.parametrize("dtype", [np.float64, np.float32], ids=["f64", "f32"])
def test_off_grid_cells_match_reference(oracle, ref, dtype):
    """Cells past the die edges: no bin at all, or a truncated index naming an
    edge bin (partial or negative overlap); map and force as the reference."""
    d = off_grid(design(0, dtype=dtype))
    dp = density_params(oracle, d)
    t = _ref_density_inputs(d, dp, dtype)
    ep = ref["electric_potential_cpp"]
    init = oracle.fixed_density_map(dp, d.pos, d.bin_center_x, d.bin_center_y)
    m_ref = ep.density_map(t["pos"], t["wxc"], t["wyc"], t["offx"], t["offy"], t["ratio"], t["bcx"], t["bcy"],
                           torch.from_numpy(init), d.target_density, d.xl, d.yl, d.xh, d.yh, d.bin_size_x,
                           d.bin_size_y, d.num_movable_nodes, d.num_filler_nodes, 0, d.num_bins_x,
                           d.num_bins_y, 3, 3, 3, 3, 1)
    m = oracle.density_map(dp, d.pos, init)
    assert np.array_equal(m.reshape(-1), m_ref.numpy().reshape(-1))
    rng = np.random.default_rng(12)
    fx = rng.standard_normal((d.num_bins_x, d.num_bins_y)).astype(dtype)
    fy = rng.standard_normal((d.num_bins_x, d.num_bins_y)).astype(dtype)
    g = torch.tensor([1.25], dtype=torch.from_numpy(fx).dtype)
    out = ep.electric_force(
        g, d.num_bins_x, d.num_bins_y, 3, 3, 3, 3, torch.from_numpy(fx).view(-1), torch.from_numpy(fy).view(-1),
        t["pos"], t["wxc"], t["wyc"], t["offx"], t["offy"], t["ratio"], t["bcx"], t["bcy"], d.xl, d.yl, d.xh,
        d.yh, d.bin_size_x, d.bin_size_y, d.num_movable_nodes, d.num_filler_nodes)
    mine = oracle.electric_force(dp, d.pos, fx, fy, grad=1.25)
    assert np.array_equal(mine, -out.numpy())


@pytest.mark.parametrize("shape", [(4, 8), (4, 4), (64, 32), (30, 18)])
@pytest.mark.parametrize("name,fn", [("dct2", "dct2_fft2"), ("idct2", "idct2_fft2"),
                                     ("idxst_idct", "idxst_idct"), ("idct_idxst", "idct_idxst")])
def test_transforms_match_reference(oracle, ref, name, fn, shape):
    from oracle.ref_iteration import exact_expk
    M, N = shape
    x = np.random.default_rng(5).standard_normal((M, N))
    out = torch.empty(M, N, dtype=torch.float64)
    buf = torch.empty(M, N // 2 + 1, 2, dtype=torch.float64)
    getattr(ref["dct2_fft2_cpp"], fn)(torch.from_numpy(x), exact_expk(M, torch.float64),
                                      exact_expk(N, torch.float64), out, buf, 1)
    assert rel_l2(getattr(oracle, name)(x), out.numpy()) <= 1e-13


@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-13), (np.float32, 1e-5)], ids=["f64", "f32"])
def test_wa_matches_reference(oracle, ref, dtype, tol):
    d = design(0, dtype=dtype)
    pin = oracle.pin_pos_forward(d.pos, d.pin_offset_x, d.pin_offset_y, d.pin2node_map)
    m = ref["weighted_average_wirelength_cpp_merged"]
    T = lambda a: torch.from_numpy(np.ascontiguousarray(a))
    ig = torch.tensor(1.0 / d.wa_gamma, dtype=T(pin).dtype)
    wl_ref, gi_ref = m.forward(T(pin), T(d.flat_net2pin_map), T(d.flat_net2pin_start_map), T(d.pin2net_map),
                               T(d.net_weights), T(d.net_mask_ignore_large_degrees), ig)
    wl, partial, gi = oracle.wa_merged_forward(pin, d.flat_net2pin_map, d.flat_net2pin_start_map,
                                               d.net_weights, d.net_mask_ignore_large_degrees, d.wa_gamma)
    assert abs(wl - wl_ref.item()) <= tol * abs(wl)
    assert np.max(np.abs(gi - gi_ref.numpy())) <= tol * np.abs(gi).max()
    g_ref = m.backward(torch.tensor(1.0, dtype=ig.dtype), T(pin), gi_ref.clone(), T(d.flat_net2pin_map),
                       T(d.flat_net2pin_start_map), T(d.pin2net_map), T(d.net_weights),
                       T(d.net_mask_ignore_large_degrees), ig)
    P = d.num_pins
    g_ref[:P].masked_fill_(T(d.pin_mask_ignore_fixed_macros), 0)
    g_ref[P:].masked_fill_(T(d.pin_mask_ignore_fixed_macros), 0)
    g = oracle.wa_merged_backward(1.0, gi, d.pin2net_map, d.net_weights, d.net_mask_ignore_large_degrees,
                                  d.pin_mask_ignore_fixed_macros)
    assert np.max(np.abs(g - g_ref.numpy())) <= tol * np.abs(g).max()


def test_small_ops_match_reference(oracle, ref):
    d = design(0)
    T = lambda a: torch.from_numpy(np.ascontiguousarray(a))
    pin_ref = ref["pin_pos_cpp"].forward(T(d.pos), T(d.pin_offset_x), T(d.pin_offset_y),
                                         T(d.pin2node_map.astype(np.int64)), T(d.flat_node2pin_map),
                                         T(d.flat_node2pin_start_map))
    pin = oracle.pin_pos_forward(d.pos, d.pin_offset_x, d.pin_offset_y, d.pin2node_map)
    assert np.array_equal(pin, pin_ref.numpy())
    g = np.random.default_rng(1).standard_normal(2 * d.num_pins)
    gpos_ref = ref["pin_pos_cpp"].backward(T(g), T(d.pos), T(d.pin_offset_x), T(d.pin_offset_y),
                                           T(d.pin2node_map.astype(np.int64)), T(d.flat_node2pin_map),
                                           T(d.flat_node2pin_start_map), d.num_physical_nodes)
    gpos = oracle.pin_pos_backward(g, d.flat_node2pin_map, d.flat_node2pin_start_map, d.num_physical_nodes,
                                   d.num_nodes)
    assert np.array_equal(gpos, gpos_ref.numpy())
    h_ref = ref["hpwl_cpp"].forward(T(pin), T(d.flat_net2pin_map), T(d.flat_net2pin_start_map),
                                    T(d.net_weights), T(d.net_mask_all))
    h = oracle.hpwl(pin, d.flat_net2pin_map, d.flat_net2pin_start_map, d.net_weights, d.net_mask_all)
    assert h == pytest.approx(h_ref.item(), rel=1e-13)
    pos = d.pos + np.random.default_rng(2).normal(0, 30, d.pos.shape)
    pr = T(pos.copy())
    ref["move_boundary_cpp"].forward(pr, T(d.node_size_x), T(d.node_size_y), d.xl, d.yl, d.xh, d.yh,
                                     d.num_movable_nodes, d.num_filler_nodes)
    mine = oracle.move_boundary(pos.copy(), d.node_size_x, d.node_size_y, d.xl, d.yl, d.xh, d.yh,
                                d.num_movable_nodes, d.num_filler_nodes)
    assert np.array_equal(mine, pr.numpy())
    w = np.exp(np.random.default_rng(3).uniform(-0.7, 1.4, d.num_nets))
    pws_ref = ref["pws_cpp"].forward(T(w), T(d.flat_node2pin_map), T(d.flat_node2pin_start_map),
                                     T(d.pin2net_map), d.num_nodes)
    pws = oracle.pin_weight_sum(w, d.flat_node2pin_map, d.flat_node2pin_start_map, d.pin2net_map, d.num_nodes)
    assert np.array_equal(pws, pws_ref.numpy())


def test_full_iteration_matches_reference(oracle, ref):
    from oracle import ref_iteration as R
    d = design(0)
    it = R.RefIteration(d, dtype=torch.float64, density_weight=0.25)
    rf, rg, det = it.iterate()
    f, g, parts = oracle.objective_and_gradient(d, 0.25, d.wa_gamma)
    assert f == pytest.approx(rf.item(), rel=1e-13)
    assert rel_l2(g, rg.numpy()) <= 1e-12


# ---------------------------------------------------------------------------
# 3. mathematics
# ---------------------------------------------------------------------------
def _tiny_design(rng, n=40, nb=8, die=16.0):
    bs = die / nb
    w = rng.uniform(0.3, 3.5, n)
    h = rng.uniform(0.3, 3.5, n)
    x = rng.uniform(0, die - w)
    y = rng.uniform(0, die - h)
    dp = type("DP", (), {})()
    dp.wxc = np.maximum(w, bs * math.sqrt(2))
    dp.wyc = np.maximum(h, bs * math.sqrt(2))
    dp.offx = (w - dp.wxc) * 0.5
    dp.offy = (h - dp.wyc) * 0.5
    dp.ratio = w * h / (dp.wxc * dp.wyc)
    return x, y, w, h, dp, bs, nb, die


def test_triangle_scatter_brute_force(oracle):
    """Each (cell, bin) adds the overlap of the stretched cell with the bin,
    times the area ratio (within the 2^-32 fixed-point step per pair)."""
    rng = np.random.default_rng(0)
    x, y, w, h, dp, bs, nb, die = _tiny_design(rng)
    scale = oracle.fixed_scale(0, 0, die, die)
    buf = oracle.triangle_density_fixed(x, y, dp.wxc, dp.wyc, dp.offx, dp.offy, dp.ratio, nb, nb, 0.0, 0.0, bs,
                                        bs, scale)
    brute = np.zeros((nb, nb))
    for i in range(len(x)):
        x0, y0 = x[i] + dp.offx[i], y[i] + dp.offy[i]
        for k in range(nb):
            ox = max(0.0, min(x0 + dp.wxc[i], (k + 1) * bs) - max(x0, k * bs))
            for l in range(nb):
                oy = max(0.0, min(y0 + dp.wyc[i], (l + 1) * bs) - max(y0, l * bs))
                brute[k, l] += ox * oy * dp.ratio[i]
    got = buf.reshape(nb, nb) / scale
    assert np.max(np.abs(got - brute)) <= len(x) * 9 * 2.0 ** -32 + 1e-12


def test_triangle_scatter_mass_and_order(oracle):
    """Cells inside the die: the map sums to the cells' true area; any
    permutation of the cells gives the identical int64 map."""
    rng = np.random.default_rng(1)
    die, nb = 64.0, 32
    bs = die / nb
    n = 500
    w, h = rng.uniform(0.5, 6, n), rng.uniform(0.5, 6, n)
    wxc, wyc = np.maximum(w, bs * math.sqrt(2)), np.maximum(h, bs * math.sqrt(2))
    offx, offy = (w - wxc) / 2, (h - wyc) / 2
    ratio = w * h / (wxc * wyc)
    x = rng.uniform(-offx, die - wxc - offx)
    y = rng.uniform(-offy, die - wyc - offy)
    scale = oracle.fixed_scale(0, 0, die, die)
    buf = oracle.triangle_density_fixed(x, y, wxc, wyc, offx, offy, ratio, nb, nb, 0.0, 0.0, bs, bs, scale)
    assert abs(buf.sum() / scale - (w * h).sum()) <= 1e-6
    p = rng.permutation(n)
    buf2 = oracle.triangle_density_fixed(x[p], y[p], wxc[p], wyc[p], offx[p], offy[p], ratio[p], nb, nb, 0.0,
                                         0.0, bs, bs, scale)
    assert np.array_equal(buf, buf2)


def test_fixed_point_scale(oracle):
    assert oracle.fixed_scale(0, 0, 512, 512) == 2 ** 32
    assert oracle.fixed_scale(0, 0, 1e4, 1e4) == 2 ** 32
    assert oracle.fixed_scale(0, 0, 2e4, 2e4) == 2 ** 32
    # area 2^40 -> ceil(log2)+1 = 41 integer bits -> 23 fraction bits
    assert oracle.fixed_scale(0, 0, 2.0 ** 20, 2.0 ** 20) == 2 ** 23


def test_gpu_conversion_known_answers(oracle):
    """The CUDA deterministic path's conversion (atomic_ops.cuh:53-59 as
    electric_density_map_cuda_kernel.cu:447-461 uses it): the initial map is
    quantised with u64(init * T(scale)) (truncation), the integer scatter sum
    is added, and the map is T(u64) * T(1/scale).  Expected values are
    derived by hand (scale 2^32) and chosen so that rounding instead of
    truncating, adding init after the conversion, or the CPU path's double
    sum each give another answer (noted per entry)."""
    s32 = 2 ** 32
    init = np.array([3 * 2.0 ** -33,   # init*2^32 = 1.5 -> 1 (round: 2; init after / cpu: 3*2^-33)
                     0.25,             # 2^30 + 3 -> T rounds to 2^30 (ulp 128)
                     2.0 ** -9,        # 2^23 + (2^23 + 3) = 2^24 + 3 -> tie to even: 2^24 + 4
                     1 - 2.0 ** -24,   # 2^32 - 256 + 5 -> nearest float is 2^32 - 256
                     5 * 2.0 ** -35],  # 0.625 -> 0 (round: 1), + 7
                    dtype=np.float32)
    sums = np.array([0, 3, 2 ** 23 + 3, 5, 7], dtype=np.int64)
    expect = np.array([2.0 ** -32, 0.25, 2.0 ** -8 + 2.0 ** -30, 1 - 2.0 ** -24, 7 * 2.0 ** -32],
                      dtype=np.float32)
    q = oracle.fixed_quantize_gpu(init, s32)
    assert q.tolist() == [1, 2 ** 30, 2 ** 23, 2 ** 32 - 256, 0]
    got = oracle.fixed_to_map_gpu(q + sums, s32, np.float32)
    assert np.array_equal(got, expect)
    # the planted mistakes give other maps
    rounded = np.floor(init.astype(np.float64) * s32 + 0.5).astype(np.int64)
    assert not np.array_equal(oracle.fixed_to_map_gpu(rounded + sums, s32, np.float32), expect)
    after = (oracle.fixed_to_map_gpu(sums, s32, np.float32) + init).astype(np.float32)
    assert not np.array_equal(after, expect)
    cpu = oracle.fixed_add_cpu(sums, s32, init.copy())
    assert not np.array_equal(cpu, expect)
    # FP64: 0.1*2^32 = 429496729.6... -> 429496729 (round: ...730; cpu: 0.1);
    # 0.75*2^32 + 2^33 + 1 = 11*2^30 + 1, exact in double
    init64 = np.array([0.1, 0.75])
    q64 = oracle.fixed_quantize_gpu(init64, s32)
    assert q64.tolist() == [429496729, 3 * 2 ** 30]
    got64 = oracle.fixed_to_map_gpu(q64 + np.array([0, 2 ** 33 + 1]), s32, np.float64)
    assert got64.tolist() == [429496729 / 2.0 ** 32, 2.75 + 2.0 ** -32]


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_gpu_and_cpu_conversions_agree_within_a_quantum(oracle, dtype):
    """On a whole config-0 map the two conventions differ only by what
    quantising the initial map drops (< 2^-32 per bin) plus each side's final
    rounding to T; and T(u64)*T(1/scale) equals rounding the exact double
    product, since the scale is a power of two and every sum is below 2^53.
    With target density 0.9 the FP64 initial map has fractions below 2^-32,
    so there the two maps differ."""
    d = design(0, dtype=dtype)
    d.target_density = 0.9
    dp = density_params(oracle, d)
    assert dp.scale & (dp.scale - 1) == 0
    init = oracle.fixed_density_map(dp, d.pos, d.bin_center_x, d.bin_center_y)
    cpu = oracle.density_map(dp, d.pos, init, convention="cpu").reshape(-1)
    gpu = oracle.density_map(dp, d.pos, init, convention="gpu").reshape(-1)
    ulp = np.spacing(np.maximum(np.abs(cpu), np.abs(gpu)))
    assert np.all(np.abs(gpu.astype(np.float64) - cpu) <= 1.0 / dp.scale + ulp)
    if dtype == np.float64:
        assert not np.array_equal(gpu, cpu)
    buf = oracle.fixed_quantize_gpu(init.reshape(-1), dp.scale) + oracle.density_fixed_sum(dp, d.pos)
    assert buf.max() < 2 ** 53
    assert np.array_equal(gpu, (buf.astype(np.float64) / dp.scale).astype(dtype))


def test_dct_brute_force(oracle):
    """The four transforms against their defining double sums (4x8)."""
    M, N = 4, 8
    x = np.random.default_rng(2).standard_normal((M, N))
    cu = lambda k: 1.0 if k == 0 else 2.0
    C = lambda u, i, K: math.cos(math.pi * u * (2 * i + 1) / (2 * K))
    S = lambda u, i, K: math.sin(math.pi * u * (2 * i + 1) / (2 * K))
    d2 = np.zeros((M, N))
    i2 = np.zeros((M, N))
    xs = np.zeros((M, N))
    cs = np.zeros((M, N))
    for a in range(M):
        for b in range(N):
            d2[a, b] = 4 / (M * N) * sum(x[i, j] * C(a, i, M) * C(b, j, N) for i in range(M) for j in range(N))
            i2[a, b] = sum(cu(u) * cu(v) * x[u, v] * C(u, a, M) * C(v, b, N) for u in range(M) for v in range(N))
            xs[a, b] = 2 * sum(cu(v) * x[u, v] * S(u, a, M) * C(v, b, N) for u in range(M) for v in range(N))
            cs[a, b] = 2 * sum(cu(u) * x[u, v] * C(u, a, M) * S(v, b, N) for u in range(M) for v in range(N))
    assert np.allclose(oracle.dct2(x), d2, atol=1e-13)
    assert np.allclose(oracle.idct2(x), i2, atol=1e-13)
    assert np.allclose(oracle.idxst_idct(x), xs, atol=1e-13)
    assert np.allclose(oracle.idct_idxst(x), cs, atol=1e-13)


def test_dct_inversion(oracle):
    """IDCT2(DCT2(x)) = 4x for these normalisations (c_0=1, c_k=2 per axis)."""
    x = np.random.default_rng(3).standard_normal((16, 12))
    assert np.allclose(oracle.idct2(oracle.dct2(x)), 4 * x, atol=1e-12)


@pytest.mark.parametrize("u,v", [(1, 0), (3, 2), (0, 5), (7, 7)])
def test_poisson_single_mode(oracle, u, v):
    """rho = C_M[u,m] C_N[v,n].  From sum_m C[u,m]C[u',m] = delta*(M if u=0
    else M/2) and the transform definitions:
      a = DCT2(rho) = 4/(c_u c_v) e_uv,   phi = 4 K rho,
      E_x = (4/c_u) wu K S_M[u,m] C_N[v,n],   E_y = (4/c_v) wv K C_M[u,m] S_N[v,n],
      energy = sum(phi*rho) = 4 K sum(rho^2),
    with K = 1/(wu^2+wv^2); the bin aspect 2 exercises the wv scaling."""
    M, N, bx, by = 16, 32, 2.0, 1.0
    m = np.arange(M)[:, None]
    n = np.arange(N)[None, :]
    Cu = np.cos(np.pi * u * (2 * m + 1) / (2 * M))
    Cv = np.cos(np.pi * v * (2 * n + 1) / (2 * N))
    Su = np.sin(np.pi * u * (2 * m + 1) / (2 * M))
    Sv = np.sin(np.pi * v * (2 * n + 1) / (2 * N))
    rho = Cu * Cv
    r = oracle.poisson_solve(rho * (bx * by), bx, by)
    wu = 2 * np.pi * u / M
    wv = 2 * np.pi * v / N * (bx / by)
    K = 1.0 / (wu ** 2 + wv ** 2)
    cu = 1.0 if u == 0 else 2.0
    cv = 1.0 if v == 0 else 2.0
    a_expect = np.zeros((M, N))
    a_expect[u, v] = 4.0 / (cu * cv)
    assert np.allclose(r["auv"], a_expect, atol=1e-12)
    assert np.allclose(r["phi"], 4.0 * K * rho, atol=1e-12)
    assert np.allclose(r["field_x"], (4.0 / cu) * wu * K * Su * Cv, atol=1e-12)
    assert np.allclose(r["field_y"], (4.0 / cv) * wv * K * Cu * Sv, atol=1e-12)
    assert r["energy"] == pytest.approx(4.0 * K * float((rho * rho).sum()), rel=1e-12)


def test_poisson_constant_map(oracle):
    """A uniform density has no field and no energy."""
    r = oracle.poisson_solve(np.full((16, 16), 3.7), 1.5, 1.5)
    assert np.abs(r["field_x"]).max() < 1e-12 and np.abs(r["field_y"]).max() < 1e-12
    assert abs(r["energy"]) < 1e-9


def test_poisson_field_is_potential_gradient(oracle):
    """E_x = -d(phi)/dm of phi's cosine interpolant, on a random map.
    phi = sum c_u c_v b[u,v] C_M[u,m] C_N[v,n] with b = DCT2(phi)/4, so
    d(phi)/dm = sum c_u c_v b (-(pi u/M) S_M[u,m]) C_N[v,n]."""
    rng = np.random.default_rng(4)
    M, N = 12, 10
    r = oracle.poisson_solve(rng.standard_normal((M, N)), 1.0, 1.0)
    b = oracle.dct2(r["phi"]) / 4.0
    u = np.arange(M)
    v = np.arange(N)
    c_u = np.where(u == 0, 1.0, 2.0)
    c_v = np.where(v == 0, 1.0, 2.0)
    mm = np.arange(M)
    nn = np.arange(N)
    dS = -(np.pi * u[:, None] / M) * np.sin(np.pi * u[:, None] * (2 * mm[None, :] + 1) / (2 * M))  # [u, m]
    Cn = np.cos(np.pi * v[:, None] * (2 * nn[None, :] + 1) / (2 * N))  # [v, n]
    dphi = dS.T @ (c_u[:, None] * c_v[None, :] * b) @ Cn
    assert np.allclose(r["field_x"], -dphi, atol=1e-10 * np.abs(r["field_x"]).max())


def _wa_fixture(rng, n_nets=30, dtype=np.float64):
    deg = rng.integers(2, 7, n_nets)
    P = int(deg.sum())
    flat = np.arange(P, dtype=np.int32)
    start = np.zeros(n_nets + 1, dtype=np.int32)
    start[1:] = np.cumsum(deg)
    pin2net = np.repeat(np.arange(n_nets, dtype=np.int32), deg)
    pos = rng.uniform(0, 50, 2 * P).astype(dtype)
    w = rng.uniform(0.5, 4, n_nets).astype(dtype)
    mask = np.ones(n_nets, dtype=np.uint8)
    mask[::7] = 0
    return pos, flat, start, pin2net, w, mask


def _edge_clear_coordinates(rng, d, dp, count, margin):
    """`count` (cell, axis) pairs of movable or filler cells inside the die
    whose stretched edges are at least `margin` from every bin edge."""
    n = d.num_nodes
    cells = np.concatenate([np.arange(d.num_movable_nodes),
                            np.arange(n - d.num_filler_nodes, n)])
    out = []
    while len(out) < count:
        i = int(rng.choice(cells))
        ax = len(out) % 2
        b = d.bin_size_x if ax == 0 else d.bin_size_y
        lo = d.pos[ax * n + i] + (dp.offx if ax == 0 else dp.offy)[i]
        hi = lo + (dp.wxc if ax == 0 else dp.wyc)[i]
        l0, h0 = (d.xl, d.xh) if ax == 0 else (d.yl, d.yh)
        if lo < l0 + b or hi > h0 - b:
            continue
        if min(abs(v / b - round(v / b)) * b for v in (lo - l0, hi - l0)) < margin:
            continue
        out.append((i, ax))
    return out


def test_energy_gradient_finite_differences(oracle):
    """The exact energy gradient against central differences of the
    oracle's own energy (FP64, config 0).  Between bin edges the map is
    linear in a cell's position and the energy quadratic, so central
    differences are exact up to the map's fixed-point quantum (2^-32 per
    contribution): h = 1e-2*b must agree to 1e-6 of the largest derivative;
    h = 1e-4*b (the quantum divided by a 100x smaller step) to 1e-3."""
    d = design(0)
    dp = density_params(oracle, d)
    n, b = d.num_nodes, d.bin_size_x
    init = oracle.fixed_density_map(dp, d.pos, d.bin_center_x, d.bin_center_y)

    def energy(pos):
        return oracle.poisson_solve(oracle.density_map(dp, pos, init), b, d.bin_size_y)

    sol = energy(d.pos)
    g = oracle.energy_gradient(dp, d.pos, sol["phi"])
    rng = np.random.default_rng(5)
    for hf, tol in ((1e-2, 1e-6), (1e-4, 1e-3)):
        h = hf * b
        fd, an = [], []
        for i, ax in _edge_clear_coordinates(rng, d, dp, 12, 3 * h):
            p = d.pos.copy()
            p[ax * n + i] += h
            ep = energy(p)["energy"]
            p[ax * n + i] -= 2 * h
            em = energy(p)["energy"]
            fd.append((ep - em) / (2 * h))
            an.append(g[ax * n + i])
        fd, an = np.array(fd), np.array(an)
        assert np.abs(fd - an).max() <= tol * np.abs(fd).max(), hf
    # fixed cells do not move the energy
    assert not g[d.num_movable_nodes:n - d.num_filler_nodes].any()


def test_energy_gradient_linear_potential_closed_form(oracle):
    """For a linear potential phi[k,h] = a*(k+1/2) + c*(h+1/2) with its exact
    field (E_x = -a, E_y = -c), both gathers have closed forms per cell inside
    the die (stretched sizes sx, sy; edge bins kL, kR and hB, hT):
      exact gradient  g_x = 2/(bx*by) * ratio * a * (kR - kL) * sy,
      reference force f_x = ratio * a * sx * sy,
    so f_x = (bx^2*by/2) * (sx/bx)/(kR - kL) * g_x: the force is the exact
    gradient times bx^2*by/2 when a cell's edges lie its own width apart in
    bins, as they do on average over positions (the relation of the
    ePlace force to the energy's derivative)."""
    d = design(0)
    dp = density_params(oracle, d)
    n, M, Nb = d.num_nodes, d.num_bins_x, d.num_bins_y
    bx, by = d.bin_size_x, d.bin_size_y
    a, c = 0.37, -1.3
    phi = a * (np.arange(M)[:, None] + 0.5) + c * (np.arange(Nb)[None, :] + 0.5)
    g = oracle.energy_gradient(dp, d.pos, phi)
    f = oracle.electric_force(dp, d.pos, np.full((M, Nb), -a), np.full((M, Nb), -c))
    cells = np.concatenate([np.arange(d.num_movable_nodes), np.arange(n - d.num_filler_nodes, n)])
    x = d.pos[cells] + dp.offx[cells]
    y = d.pos[n + cells] + dp.offy[cells]
    sx, sy, r = dp.wxc[cells], dp.wyc[cells], dp.ratio[cells]
    inside = (x > d.xl) & (x + sx < d.xh) & (y > d.yl) & (y + sy < d.yh)
    assert inside.sum() > 0.9 * len(cells)
    cells, x, y, sx, sy, r = (v[inside] for v in (cells, x, y, sx, sy, r))
    kspan = np.floor((x + sx - d.xl) / bx) - np.floor((x - d.xl) / bx)
    hspan = np.floor((y + sy - d.yl) / by) - np.floor((y - d.yl) / by)
    coef = 2.0 / (bx * by)
    assert np.allclose(g[cells], coef * r * a * kspan * sy, rtol=1e-12, atol=0)
    assert np.allclose(g[n + cells], coef * r * c * hspan * sx, rtol=1e-12, atol=0)
    assert np.allclose(f[cells], r * a * sx * sy, rtol=1e-12, atol=0)
    assert np.allclose(f[n + cells], r * c * sx * sy, rtol=1e-12, atol=0)
    const = bx * bx * by / 2
    assert np.allclose(f[cells], const * (sx / bx) / kspan * g[cells], rtol=1e-12, atol=0)


def test_wa_gradient_finite_difference(oracle):
    rng = np.random.default_rng(5)
    pos, flat, start, pin2net, w, mask = _wa_fixture(rng)
    gamma = 4.0
    wl, partial, gi = oracle.wa_merged_forward(pos, flat, start, w, mask, gamma)
    g = oracle.wa_merged_backward(1.0, gi, pin2net, w, mask, np.zeros(len(pos) // 2, dtype=bool))
    h = 1e-6
    for k in rng.choice(len(pos), 40, replace=False):
        pp, pm = pos.copy(), pos.copy()
        pp[k] += h
        pm[k] -= h
        fd = (oracle.wa_merged_forward(pp, flat, start, w, mask, gamma)[0] -
              oracle.wa_merged_forward(pm, flat, start, w, mask, gamma)[0]) / (2 * h)
        assert fd == pytest.approx(g[k], abs=1e-6)


def test_wa_net_weight_adjoint_finite_difference(oracle):
    rng = np.random.default_rng(6)
    pos, flat, start, pin2net, w, mask = _wa_fixture(rng)
    wl, partial, gi = oracle.wa_merged_forward(pos, flat, start, w, mask, 2.0)
    gw = oracle.wa_net_weight_grad(1.0, partial, mask)
    h = 1e-5
    for e in range(len(w)):
        wp, wm = w.copy(), w.copy()
        wp[e] += h
        wm[e] -= h
        fd = (oracle.wa_merged_forward(pos, flat, start, wp, mask, 2.0)[0] -
              oracle.wa_merged_forward(pos, flat, start, wm, mask, 2.0)[0]) / (2 * h)
        assert fd == pytest.approx(gw[e], abs=1e-6)


def test_wa_limits_and_invariance(oracle):
    """gamma -> 0 gives HPWL; shifting all pins leaves WA and its gradient unchanged."""
    rng = np.random.default_rng(7)
    pos, flat, start, pin2net, w, mask = _wa_fixture(rng)
    wl = oracle.wa_merged_forward(pos, flat, start, w, mask, 1e-3)[0]
    hp = oracle.hpwl(pos, flat, start, w, mask)
    assert wl == pytest.approx(hp, rel=1e-9)
    P = len(pos) // 2
    shifted = pos.copy()
    shifted[:P] += 123.0
    shifted[P:] -= 77.0
    a = oracle.wa_merged_forward(pos, flat, start, w, mask, 3.0)
    b = oracle.wa_merged_forward(shifted, flat, start, w, mask, 3.0)
    assert a[0] == pytest.approx(b[0], rel=1e-10)
    assert np.allclose(a[2], b[2], atol=1e-10)


def test_pin_pos_adjoint(oracle):
    d = design(0)
    rng = np.random.default_rng(8)
    dx = rng.standard_normal(2 * d.num_nodes)
    dx[d.num_physical_nodes:d.num_nodes] = 0
    dx[d.num_nodes + d.num_physical_nodes:] = 0
    zero = np.zeros(d.num_pins)
    fwd = oracle.pin_pos_forward(dx, zero, zero, d.pin2node_map)
    g = rng.standard_normal(2 * d.num_pins)
    bwd = oracle.pin_pos_backward(g, d.flat_node2pin_map, d.flat_node2pin_start_map, d.num_physical_nodes,
                                  d.num_nodes)
    assert float(fwd @ g) == pytest.approx(float(dx @ bwd), rel=1e-12)


def test_pin_weight_sum_closed_form(oracle):
    d = design(0)
    w = np.random.default_rng(9).uniform(0.5, 4, d.num_nets)
    expect = np.zeros(d.num_nodes)
    np.add.at(expect, d.pin2node_map, w[d.pin2net_map])
    got = oracle.pin_weight_sum(w, d.flat_node2pin_map, d.flat_node2pin_start_map, d.pin2net_map, d.num_nodes)
    assert np.allclose(got, expect, rtol=1e-13)


def test_objective_invariant_under_pin_relabelling(oracle):
    """Pin ids carry no meaning beyond the net and node pin sequences: renaming
    them gives the identical objective and gradient (the GPU schedules that
    permute pins between net and node order are held to the same)."""
    d = design(0, big_nets=5)
    d2 = relabel_pins(d)
    assert not np.array_equal(d2.flat_net2pin_map, d.flat_net2pin_map)
    f1, g1, _ = oracle.objective_and_gradient(d, 0.3, d.wa_gamma)
    f2, g2, _ = oracle.objective_and_gradient(d2, 0.3, d2.wa_gamma)
    assert f1 == f2 and np.array_equal(g1, g2)


def test_planted_sign_error_is_caught(oracle, monkeypatch):
    """A sign flip in the oracle's field would be caught by the single-mode
    closed form (this plants the mistake in the oracle, on the CPU)."""
    orig = oracle.idxst_idct
    monkeypatch.setattr(oracle, "idxst_idct", lambda x: -orig(x))
    with pytest.raises(AssertionError):
        test_poisson_single_mode(oracle, 3, 2)


# ---------------------------------------------------------------------------
# the placement driver (oracle/placer.py)
# ---------------------------------------------------------------------------
def _quadratic(seed=0, n=64):
    a = np.random.default_rng(seed).uniform(0.5, 20.0, n)

    def f(x):
        return 0.5 * float(np.sum(a * x * x)), a * x
    return a, f


@pytest.mark.parametrize("use_bb", [False, True], ids=["nobb", "bb"])
def test_nesterov_converges_on_quadratic(use_bb):
    """f = 1/2 sum a_i x_i^2 (minimiser 0, Lipschitz constant L = max a): the
    restated optimizer meets the accelerated-gradient bound
    f(x_k) <= 2 L |x_0|^2 / (k+1)^2 and keeps descending."""
    from oracle import placer as OP
    a, f = _quadratic()
    x = np.random.default_rng(1).standard_normal(len(a))
    f0, r0 = f(x)[0], float(np.dot(x, x))
    opt = OP.Nesterov(x, 0.01, f, lambda v: None, use_bb=use_bb)
    for k in range(1, 61):
        opt.step()
        assert f(x)[0] <= 2 * a.max() * r0 / (k + 1) ** 2
    assert f(x)[0] < 1e-6 * f0


def test_nesterov_first_step_closed_form():
    """First step_nobb: a_1 = (1+sqrt 5)/2, coef = 0, alpha from the initial
    secant; on a quadratic with all a_i = c the secant step is exactly 1/c,
    so v_1 = u_1 = x0 - (1/c) * c * x0 = 0 and the line search accepts it."""
    from oracle import placer as OP
    c = 4.0
    x = np.linspace(-1, 1, 16)

    def f(v):
        return 0.5 * c * float(np.dot(v, v)), c * v
    opt = OP.Nesterov(x, 0.05, f, lambda v: None)
    opt.step()
    assert np.allclose(x, 0.0, atol=1e-15)
    assert opt.alpha_k == pytest.approx(1.0 / c)
    assert opt.a_k == pytest.approx((1 + 5 ** 0.5) / 2)


def test_placement_loop_decreases_hpwl(oracle):
    """Six iterations of the restated loop on config 0: HPWL falls from the
    spread-out start, the density weight starts at the gradient-norm ratio and
    gamma follows the overflow schedule."""
    from oracle import placer as OP
    d = design(0)
    pos, hist, _ = OP.place(d, 6)
    assert hist[-1]["hpwl"] < 0.5 * hist[0]["hpwl"]
    base = OP.DEFAULTS["gamma"] * (d.bin_size_x + d.bin_size_y)
    assert hist[0]["gamma"] == pytest.approx(10 * base)
    for h0, h1 in zip(hist, hist[1:]):
        assert h1["gamma"] == pytest.approx(base * 10 ** ((h0["overflow"] - 0.1) * 20 / 9 - 1))


def test_plateau_and_divergence_known_answers():
    """check_plateau / check_divergence (NonLinearPlace.py:527-570) on
    histories whose answer follows from the definitions."""
    from oracle import placer as OP
    assert OP.plateau([1.0] * 15, 15, 0.001)
    assert not OP.plateau([1.0] * 14, 15, 0.001)          # window not full
    assert not OP.plateau(list(np.linspace(1.0, 0.9, 15)), 15, 0.001)
    best = dict(hpwl=100.0, overflow=0.2)
    flat_wl = [(100.0, 0.3)] * 50
    assert not OP.diverging(flat_wl, best, 0.1, 50, 0.05)  # wirelength has not grown
    grown = [(150.0, 0.3)] * 50                            # overflow mean 50% above best
    assert OP.diverging(grown, best, 0.1, 50, 0.05)
    assert not OP.diverging(grown[:49], best, 0.1, 50, 0.05)
    assert not OP.diverging(grown, None, 0.1, 50, 0.05)
    level = [(150.0, 0.2)] * 50                            # overflow flat at the best
    assert OP.diverging(level, best, 0.1, 50, 0.05)
    # overflow at the best on average, spread 13% of the mean: only the
    # fraction of rising steps decides (40 of 49 rise; 25 of 49 rise)
    best3 = dict(hpwl=100.0, overflow=0.3)
    saw = [(150.0, 0.28 + 0.01 * (i % 5)) for i in range(50)]
    assert OP.diverging(saw, best3, 0.1, 50, 0.05)
    alt = [(150.0, 0.28 + 0.04 * (i % 2)) for i in range(50)]
    assert not OP.diverging(alt, best3, 0.1, 50, 0.05)
    falling = [(150.0, 0.2 - 0.002 * i) for i in range(50)]  # steady descent, ratio < 0
    assert not OP.diverging(falling, dict(hpwl=100.0, overflow=0.2), 0.1, 50, 0.05)


def test_entropy_injection_invariants():
    """Shrink about the movables' centroid keeps that centroid and scales
    every offset by the factor; fixed nodes stay put; the noise is added as
    given."""
    from oracle import placer as OP
    d = design(0)
    n, m0 = d.num_nodes, d.num_movable_nodes
    m1 = n - d.num_filler_nodes
    pos = d.pos.copy()
    xc = pos[:m0].mean()
    OP.entropy_injection(d, pos, 0.5, 0.0, None)
    assert pos[:m0].mean() == pytest.approx(xc, rel=1e-12)
    assert np.allclose(pos[:m0] - xc, 0.5 * (d.pos[:m0] - xc), rtol=0, atol=1e-9)
    assert np.array_equal(pos[m0:m1], d.pos[m0:m1]) and np.array_equal(pos[n + m0:n + m1], d.pos[n + m0:n + m1])
    pos2 = d.pos.copy()
    z = OP.injection_noise(n, 0, 0)
    OP.entropy_injection(d, pos2, 1, 2.0, z)
    mov = np.r_[0:m0, m1:n + m0, n + m1:2 * n]
    assert np.allclose(pos2[mov] - d.pos[mov], 2.0 * z[mov], rtol=0, atol=1e-12)
    assert np.array_equal(pos2[m0:m1], d.pos[m0:m1])


def test_quadratic_penalty_derivative(oracle):
    """With the quadratic penalty the density part of the objective is
    Q(E) = f*lambda*E*(1 + c*E): along a direction u, the ratio of the
    density gradients (penalised / plain, lambda = 1) equals the ratio of
    finite differences of Q and E along u, f*lambda*(1 + c*(E(+t) + E(-t)))."""
    d = design(0)
    gamma, lam, f = d.wa_gamma, 0.7, 2.0
    _, _, p0 = oracle.objective_and_gradient(d, 1.0, gamma)
    e0 = p0["energy"]
    c = 1000.0 / e0
    _, _, pq = oracle.objective_and_gradient(d, lam, gamma, density_factor=f, quad_coeff=c)
    u = np.zeros(2 * d.num_nodes)
    u[:d.num_movable_nodes] = np.random.default_rng(5).standard_normal(d.num_movable_nodes)
    t = 1e-3
    es = []
    for s in (t, -t):
        d2 = synth.Design(d)
        d2.pos = d.pos + s * u
        es.append(oracle.objective_and_gradient(d2, 1.0, gamma)[2]["energy"])
    q = lambda e: f * lam * e * (1 + c * e)
    fd_ratio = (q(es[0]) - q(es[1])) / (es[0] - es[1])
    grad_ratio = float(np.dot(pq["density_grad"], u)) / float(np.dot(p0["density_grad"], u))
    assert grad_ratio == pytest.approx(fd_ratio, rel=1e-6)
    assert grad_ratio == pytest.approx(f * lam * (1 + 2 * c * e0), rel=1e-6)


# ---------------------------------------------------------------------------
# log-sum-exp wirelength (oracle.lse_merged_forward)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-13), (np.float32, 2e-5)], ids=["f64", "f32"])
def test_lse_matches_reference(oracle, ref, dtype, tol):
    d = design(0, dtype=dtype, big_nets=5)
    pin = oracle.pin_pos_forward(d.pos, d.pin_offset_x, d.pin_offset_y, d.pin2node_map)
    m = ref["logsumexp_wirelength_cpp_merged"]
    T = lambda a: torch.from_numpy(np.ascontiguousarray(a))
    gam = torch.tensor(d.wa_gamma, dtype=T(pin).dtype)
    wl_ref, gi_ref = m.forward(T(pin), T(d.flat_net2pin_map), T(d.flat_net2pin_start_map),
                               T(d.pin2net_map), T(d.net_weights), T(d.net_mask_ignore_large_degrees), gam)
    wl, partial, gi = oracle.lse_merged_forward(pin, d.flat_net2pin_map, d.flat_net2pin_start_map,
                                                d.net_weights, d.net_mask_ignore_large_degrees, d.wa_gamma)
    assert abs(wl - wl_ref.item()) <= tol * abs(wl)
    assert np.max(np.abs(gi - gi_ref.numpy())) <= tol * np.abs(gi).max()
    g_ref = m.backward(torch.tensor(1.0, dtype=gam.dtype), T(pin), gi_ref.clone(), T(d.flat_net2pin_map),
                       T(d.flat_net2pin_start_map), T(d.pin2net_map), T(d.net_weights),
                       T(d.net_mask_ignore_large_degrees), gam)
    P = d.num_pins
    g_ref[:P].masked_fill_(T(d.pin_mask_ignore_fixed_macros), 0)
    g_ref[P:].masked_fill_(T(d.pin_mask_ignore_fixed_macros), 0)
    # the LSE backward is the WA backward's weight-and-mask step on its own derivative
    g = oracle.wa_merged_backward(1.0, gi, d.pin2net_map, d.net_weights, d.net_mask_ignore_large_degrees,
                                  d.pin_mask_ignore_fixed_macros)
    assert np.max(np.abs(g - g_ref.numpy())) <= tol * np.abs(g).max()


def test_lse_two_pin_closed_form(oracle):
    """One two-pin net per row: WL = |dx| + |dy| + 2*gamma*(log(1+exp(-|dx|/gamma))
    + log(1+exp(-|dy|/gamma))), d/dx1 = tanh((x1-x2)/(2 gamma))."""
    rng = np.random.default_rng(3)
    E = 50
    pos = rng.uniform(-20, 20, 4 * E)            # pins 2e, 2e+1; x then y
    P = 2 * E
    starts = np.arange(0, P + 1, 2)
    gamma = 3.0
    wl, partial, gi = oracle.lse_merged_forward(pos, np.arange(P), starts, np.ones(E), np.ones(E), gamma)
    dx = pos[0:P:2] - pos[1:P:2]
    dy = pos[P::2] - pos[P + 1::2]
    f = lambda d: np.abs(d) + 2 * gamma * np.log1p(np.exp(-np.abs(d) / gamma))
    assert np.allclose(partial, f(dx) + f(dy), rtol=1e-13)
    assert np.allclose(gi[0:P:2], np.tanh(dx / (2 * gamma)), atol=1e-14)
    assert np.allclose(gi[P::2], np.tanh(dy / (2 * gamma)), atol=1e-14)
    assert np.allclose(gi[0:P:2], -gi[1:P:2], atol=1e-15)


def test_lse_limits_and_gradient(oracle):
    """gamma -> 0 gives HPWL; HPWL <= LSE <= HPWL + 2*gamma*log(deg) per axis;
    the derivative matches central finite differences."""
    d = design(0)
    pin = oracle.pin_pos_forward(d.pos, d.pin_offset_x, d.pin_offset_y, d.pin2node_map)
    ones = np.ones(d.num_nets, dtype=np.uint8)
    h = oracle.hpwl(pin, d.flat_net2pin_map, d.flat_net2pin_start_map, d.net_weights, ones)
    wl0, _, _ = oracle.lse_merged_forward(pin, d.flat_net2pin_map, d.flat_net2pin_start_map, d.net_weights,
                                          ones, 1e-6)
    assert wl0 == pytest.approx(h, rel=1e-9)
    gamma = 2.0
    wl, partial, gi = oracle.lse_merged_forward(pin, d.flat_net2pin_map, d.flat_net2pin_start_map,
                                                d.net_weights, ones, gamma)
    deg = np.diff(d.flat_net2pin_start_map)
    assert h <= wl <= h + float(np.sum(d.net_weights * 4 * gamma * np.log(deg))) + 1e-6
    rng = np.random.default_rng(4)
    idx = rng.choice(2 * d.num_pins, 40, replace=False)
    hh = 1e-4
    for i in idx:
        e = d.pin2net_map[i % d.num_pins]   # difference the owning net's own wirelength
        a, b = pin.copy(), pin.copy()
        a[i] += hh
        b[i] -= hh
        fa = oracle.lse_merged_forward(a, d.flat_net2pin_map, d.flat_net2pin_start_map, np.ones(d.num_nets),
                                       ones, gamma)[1][e]
        fb = oracle.lse_merged_forward(b, d.flat_net2pin_map, d.flat_net2pin_start_map, np.ones(d.num_nets),
                                       ones, gamma)[1][e]
        assert (fa - fb) / (2 * hh) == pytest.approx(gi[i], abs=1e-7)


def test_reference_ops_drive_the_same_loop(oracle, ref):
    """The restated loop driven by the reference's own CPU ops
    (ref_iteration.reference_objective) follows the loop driven by the
    oracle's objective on config 0 (FP64): the ops agree to ~1e-13, so eight
    iterations agree to 1e-8."""
    import torch
    from oracle import placer as OP
    from oracle import ref_iteration as R
    d = design(0)
    fn = R.reference_objective(d, torch.float64, num_threads=4)
    xr, hr, _ = OP.place(d, 8, objective=fn)
    xo, ho, _ = OP.place(d, 8)
    for a, b in zip(hr, ho):
        for k in ("hpwl", "overflow", "density_weight", "gamma"):
            assert a[k] == pytest.approx(b[k], rel=1e-8), k
    assert rel_l2(xr, xo) <= 1e-8


def test_net_weighting_known_answers(oracle):
    """ADAMS and LILITH updates (net_weighting_scheme.h:106-188) derived by
    hand: nets above ignore_net_degree keep their weights; LILITH with decay
    0.5 gives c' = sqrt((1+c)(1+nc)) - 1, nc = max(0, slack/wns) for negative
    slacks; weights are capped at max_net_weight; without a negative WNS the
    criticality is kept and the weights still grow by 1 + c."""
    start = np.array([0, 2, 5, 205, 207])  # degrees 2, 3, 200, 2
    w, c = oracle.net_weighting("adams", np.array([1.0, 2.0, 3.0, 4.0]),
                                np.array([0.0, 0.5, 1.0, 0.25]), start, 100,
                                net_critical=np.array([1, 0, 1, 0], dtype=np.uint8))
    assert w.tolist() == [1.5, 2.5, 3.0, 4.5] and c.tolist() == [0.5, 0.25, 1.0, 0.125]
    w, c = oracle.net_weighting("lilith", np.array([1.0, 2.0, 3.0, 1.0]),
                                np.array([0.0, 0.0, 0.44, 0.0]), start, 100,
                                net_slack=np.array([-1.0, 0.5, -2.0, -4.0]), wns=-2.0,
                                decay=0.5, max_net_weight=1.5)
    ce = [math.sqrt(1.5) - 1, 0.0, math.sqrt(1.44 * 2.0) - 1, math.sqrt(3.0) - 1]
    assert np.allclose(c, ce, rtol=1e-15, atol=1e-15)
    # net 1 keeps c = 0 but its weight 2 is above the cap; net 2 is above the degree limit
    assert np.allclose(w, [1 + ce[0], 1.5, 3.0, 1.5], rtol=1e-15)
    w, c = oracle.net_weighting("lilith", np.array([1.0, 2.0]), np.array([0.5, 0.0]),
                                np.array([0, 2, 4]), 100, net_slack=np.array([-1.0, 1.0]), wns=0.0)
    assert c.tolist() == [0.5, 0.0] and w.tolist() == [1.5, 2.0]


# ---------------------------------------------------------------------------
# The bin-relative overlap arithmetic (DP_ARITH_BIN_RELATIVE; no reference
# counterpart): the same function as the reference arithmetic, evaluated from
# operands of the order of a bin.  Pinned by agreement with the reference
# arithmetic in FP64 and by FP32 accuracy against FP64 at large coordinates.
# ---------------------------------------------------------------------------
def _shifted_small(dtype):
    """Config 1's cells and bins (bin 10000/1024 units) on a 625-unit die
    moved to (5000, 3000), where FP32 coordinates round at 2.4e-4 units;
    every bin edge stays exact in FP32, as at configs 1 and 4."""
    from _util import shifted, with_dtype
    d = shifted(design(1, scale=0.00390625), 5000.0, 3000.0)
    assert d.num_bins_x == 64 and d.bin_size_x == 10000.0 / 1024
    return with_dtype(d, dtype)


def _maps(oracle, d, arithmetic):
    """(fixed map, movable + filler int64 sum as float64, params)."""
    dp = density_params(oracle, d)
    init = oracle.fixed_density_map(dp, d.pos, d.bin_center_x, d.bin_center_y,
                                    arithmetic=arithmetic)
    s = oracle.density_fixed_sum(dp, d.pos, arithmetic=arithmetic).astype(np.float64) / dp.scale
    return init, s, dp


def _smooth_field(d, dtype):
    u = np.linspace(0, 3 * np.pi, d.num_bins_x)[:, None]
    v = np.linspace(0, 2 * np.pi, d.num_bins_y)[None, :]
    return (np.sin(u) * np.cos(v) + 0.3).astype(dtype), (np.cos(2 * u) * np.sin(v) - 0.1).astype(dtype)


def test_bin_relative_fp64_equals_reference_arithmetic(oracle):
    """FP64: fixed map, density map and force of the bin-relative arithmetic
    equal the reference arithmetic's to rounding (the same function)."""
    d = _shifted_small(np.float64)
    init_r, m_r, dp = _maps(oracle, d, "reference")
    init_b, m_b, _ = _maps(oracle, d, "bin_relative")
    assert rel_l2(init_b, init_r) <= 1e-12
    assert rel_l2(m_b, m_r) <= 1e-12
    fx, fy = _smooth_field(d, np.float64)
    g_r = oracle.electric_force(dp, d.pos, fx, fy)
    g_b = oracle.electric_force(dp, d.pos, fx, fy, arithmetic="bin_relative")
    assert rel_l2(g_b, g_r) <= 1e-12


def test_bin_relative_fp32_accuracy(oracle):
    """FP32 at die coordinates ~5000 against FP64 on the same FP32 inputs:
    the reference arithmetic's x + s rounds at ~2.4e-4 units (~2.5e-5 of a
    bin); the bin-relative arithmetic keeps the maps and the force within
    1e-6.  A form that drops the TwoSum residual (or measures from the die
    origin) stays at the reference arithmetic's error and fails."""
    from _util import with_dtype
    d32 = _shifted_small(np.float32)
    d64 = with_dtype(d32, np.float64)
    init64, m64, dp64 = _maps(oracle, d64, "reference")
    fx, fy = _smooth_field(d32, np.float32)
    g64 = oracle.electric_force(dp64, d64.pos, fx.astype(np.float64), fy.astype(np.float64))
    err = {}
    for arith in ("reference", "bin_relative"):
        init, m, dp = _maps(oracle, d32, arith)
        g = oracle.electric_force(dp, d32.pos, fx, fy, arithmetic=arith)
        err[arith] = (rel_l2(init, init64), rel_l2(m, m64), rel_l2(g, g64))
    print("relative errors (fixed map, scatter, force):", err)
    assert max(err["bin_relative"]) <= 1e-6
    # the stretched cells' x + s rounds in the reference arithmetic; fixed
    # blocks of integer size do not (x + w stays exact within a binade), so
    # there both are at the map conversion's rounding
    for k in (1, 2):
        assert err["reference"][k] > 5 * err["bin_relative"][k]
    assert err["bin_relative"][0] <= 1.01 * err["reference"][0]


def test_bin_relative_mass_and_order(oracle):
    """Cells inside the die: the bin-relative map sums to the cells' area, and
    any permutation of the cells gives the identical int64 map."""
    rng = np.random.default_rng(5)
    die, nb = 64.0, 32
    bs = die / nb
    n = 500
    w, h = rng.uniform(0.5, 6, n), rng.uniform(0.5, 6, n)
    wxc, wyc = np.maximum(w, bs * math.sqrt(2)), np.maximum(h, bs * math.sqrt(2))
    offx, offy = (w - wxc) / 2, (h - wyc) / 2
    ratio = w * h / (wxc * wyc)
    x = rng.uniform(-offx, die - wxc - offx)
    y = rng.uniform(-offy, die - wyc - offy)
    scale = oracle.fixed_scale(0, 0, die, die)
    kw = dict(arithmetic="bin_relative")
    buf = oracle.triangle_density_fixed(x, y, wxc, wyc, offx, offy, ratio, nb, nb, 0.0, 0.0, bs, bs,
                                        scale, **kw)
    assert abs(buf.sum() / scale - (w * h).sum()) <= 1e-6
    p = rng.permutation(n)
    buf2 = oracle.triangle_density_fixed(x[p], y[p], wxc[p], wyc[p], offx[p], offy[p], ratio[p], nb, nb,
                                         0.0, 0.0, bs, bs, scale, **kw)
    assert np.array_equal(buf, buf2)
