"""FP32 forms against the FP64 oracle at the same (FP32) inputs, at die
coordinates of a large design: the reference arithmetic and the better
conditioned arithmetic this build offers beside it."""
import numpy as np
import pytest
import torch

from _util import design, rel_l2, shifted, with_dtype

pytestmark = pytest.mark.gpu

CASES = [(0, {"big_nets": 40}), (4, {"scale": 0.003, "max_fillers": 100000, "big_nets": 10})]
CASE_IDS = ["config0-bignets", "config4-small"]


def _wa_oracle(oracle, d):
    """FP64 WA value, per-net WA and node gradient (pin_pos -> WA merged ->
    WA backward -> pin_pos backward) of design `d`."""
    pp = oracle.pin_pos_forward(d.pos, d.pin_offset_x, d.pin_offset_y, d.pin2node_map)
    wl, part, gi = oracle.wa_merged_forward(pp, d.flat_net2pin_map, d.flat_net2pin_start_map,
                                            d.net_weights, d.net_mask_ignore_large_degrees,
                                            d.wa_gamma)
    gpin = oracle.wa_merged_backward(1.0, gi, d.pin2net_map, d.net_weights,
                                     d.net_mask_ignore_large_degrees,
                                     d.pin_mask_ignore_fixed_macros)
    g = oracle.pin_pos_backward(gpin, d.flat_node2pin_map, d.flat_node2pin_start_map,
                                d.num_physical_nodes, d.num_nodes)
    return wl, part, g


def _wa_engine(d, dev, wa_arith):
    """(wl, partial_wl, gradient) of the pins schedule with a zero density
    weight, so the gradient is the wirelength's."""
    from dreamplace_amd.engine import DeviceDesign, PlacementObjective
    dd = DeviceDesign(d, dev)
    obj = PlacementObjective(dd, d.wa_gamma, density_weight=0.0, precision="fp64",
                             algorithm="tiled", schedule="pins", keep_partial_wl=True,
                             wa_arith=wa_arith)
    f, g = obj.evaluate(dd.pos)
    return (float(obj.wl.double().item()), obj.partial_wl.double().cpu().numpy(),
            g.double().cpu().numpy())


@pytest.mark.parametrize("cfg,kw", CASES, ids=CASE_IDS)
def test_wa_centred_fp64_matches_oracle(oracle, gpu, cfg, kw):
    """The centred evaluation is the reference's function: FP64 agrees with
    the FP64 oracle (reference arithmetic) to rounding."""
    print("case wa centred fp64 cfg%d" % cfg, flush=True)
    d = shifted(design(cfg, **kw), 5000.0, 3000.0)
    wl, part, g = _wa_oracle(oracle, d)
    wl_c, part_c, g_c = _wa_engine(d, gpu, "centred")
    assert abs(wl_c - wl) <= 1e-12 * abs(wl)
    assert rel_l2(part_c, part) <= 1e-12
    assert rel_l2(g_c, g) <= 1e-11


@pytest.mark.parametrize("cfg,kw", CASES, ids=CASE_IDS)
def test_wa_centred_fp32_accuracy(oracle, gpu, cfg, kw):
    """FP32 at die coordinates ~5000 against FP64 at the same FP32 inputs:
    the reference arithmetic loses ~1.7e-5 of the gradient to the
    cancellation of x/gamma-sized terms; the centred form keeps ~2.4e-6, the
    level the FP32 pin positions and the node sums of pin derivatives of
    both signs leave."""
    print("case wa centred fp32 accuracy cfg%d" % cfg, flush=True)
    d32 = with_dtype(shifted(design(cfg, **kw), 5000.0, 3000.0), np.float32)
    d64 = with_dtype(d32, np.float64)
    wl, part, g = _wa_oracle(oracle, d64)
    errs = {}
    for arith in ("reference", "centred"):
        wl_a, part_a, g_a = _wa_engine(d32, gpu, arith)
        errs[arith] = (abs(wl_a - wl) / abs(wl), rel_l2(part_a, part), rel_l2(g_a, g))
    print("  relative errors (wl, per-net wl, gradient):", errs, flush=True)
    # the per-net values keep the FP32 rounding of pos + offset (2.4e-4 units
    # at these coordinates, against net spans of a few units)
    assert errs["centred"][0] <= 1e-6
    assert errs["centred"][1] <= 5e-6
    assert errs["centred"][2] <= 5e-6
    # the reference arithmetic at these coordinates is what the centred form fixes
    assert errs["reference"][2] > 2 * errs["centred"][2]


# ---------------------------------------------------------------------------
# Bin-relative overlap arithmetic of the density map, fixed map and gather
# ---------------------------------------------------------------------------
def _grid_design(dtype):
    """Config 1's cells and bin size (10000/1024) on a 625-unit die moved to
    (5000, 3000): FP32 coordinates round at 2.4e-4 units, bin edges are exact."""
    d = shifted(design(1, scale=0.00390625), 5000.0, 3000.0)
    return with_dtype(d, dtype)


DENSITY_CASES = [
    (lambda dt: design(0, dtype=dt), "config0"),
    (_grid_design, "config1-grid-shifted"),
    (lambda dt: design(4, dtype=dt, scale=0.004, max_fillers=200000), "config4-small-macros"),
]


def _density_op(d, t, algorithm):
    from dreamplace_amd.electric_potential import ElectricPotential
    return ElectricPotential(
        t["node_size_x"], t["node_size_y"], t["bin_center_x"], t["bin_center_y"],
        d.target_density, d.xl, d.yl, d.xh, d.yh, d.bin_size_x, d.bin_size_y,
        d.num_movable_nodes, d.num_terminals, d.num_filler_nodes, 0, True,
        t["sorted_node_map"], movable_macro_mask=t["movable_macro_mask"],
        algorithm=algorithm, arithmetic="bin_relative")


@pytest.mark.parametrize("algorithm", ["global", "tiled"])
@pytest.mark.parametrize("dtype", [np.float64, np.float32], ids=["f64", "f32"])
@pytest.mark.parametrize("case", DENSITY_CASES, ids=[c[1] for c in DENSITY_CASES])
def test_bin_relative_density_bit_exact(oracle, gpu, case, dtype, algorithm):
    """The fixed map, the int64 sum and the converted map of the
    bin-relative arithmetic are the oracle's bits on either path; the force
    is the oracle's within rounding (its large cells reduce across lanes)."""
    from _util import density_params, to_dev
    make, name = case
    print("case bin-relative density %s %s %s" % (name, np.dtype(dtype).name, algorithm),
          flush=True)
    d = make(dtype)
    t = to_dev(d, gpu)
    op = _density_op(d, t, algorithm)
    op.compute_initial_density_map(t["pos"])
    raw = torch.empty(d.num_bins_x * d.num_bins_y, dtype=torch.int64, device=gpu)
    dmap = op.density_map(t["pos"], fixed_map_out=raw)
    torch.cuda.synchronize()
    dp = density_params(oracle, d)
    kw = dict(arithmetic="bin_relative")
    init = oracle.fixed_density_map(dp, d.pos, d.bin_center_x, d.bin_center_y, **kw)
    assert np.array_equal(op.initial_density_map.cpu().numpy(), init), "fixed map differs"
    assert np.array_equal(raw.cpu().numpy(), oracle.density_fixed_sum(dp, d.pos, **kw)), \
        "int64 fixed-point sum differs"
    assert np.array_equal(dmap.cpu().numpy(), oracle.density_map(dp, d.pos, init, **kw)), \
        "converted map differs"
    rng = np.random.default_rng(7)
    fx = rng.standard_normal((d.num_bins_x, d.num_bins_y)).astype(dtype)
    fy = rng.standard_normal((d.num_bins_x, d.num_bins_y)).astype(dtype)
    g = torch.tensor(0.75, dtype=t["pos"].dtype, device=gpu)
    out = op.electric_force(t["pos"], torch.from_numpy(fx).to(gpu), torch.from_numpy(fy).to(gpu), g)
    ref = oracle.electric_force(dp, d.pos, fx, fy, grad=0.75, **kw)
    tol = 1e-12 if dtype == np.float64 else 1e-6
    assert np.max(np.abs(out.cpu().numpy() - ref)) <= tol * np.abs(ref).max()


@pytest.mark.parametrize("precision,tol", [("fp64", 5e-6), ("bf16x3", 1e-5)])
def test_engine_accurate_fp32_whole_step(oracle, gpu, precision, tol):
    """One whole step in FP32 with the centred WA and the bin-relative
    density, against the FP64 oracle (reference arithmetic) at the same FP32
    inputs and large coordinates: density gradient and total gradient within
    the stated relative L2 (the FP64 solve leaves the FP32 rounding of the
    chain; BF16x3 adds its own ~6e-6); the reference arithmetic does not."""
    from dreamplace_amd.engine import DeviceDesign, PlacementObjective
    print("case accurate fp32 whole step %s" % precision, flush=True)
    d32 = _grid_design(np.float32)
    d64 = with_dtype(d32, np.float64)
    lam = 0.25
    rf, rg, parts = oracle.objective_and_gradient(d64, lam, d64.wa_gamma)
    errs = {}
    for arith in ("reference", "accurate"):
        acc = arith == "accurate"
        dd = DeviceDesign(d32, gpu)
        ob = PlacementObjective(dd, d32.wa_gamma, density_weight=lam, precision=precision,
                                algorithm="tiled", schedule="pins",
                                wa_arith="centred" if acc else "reference",
                                density_arith="bin_relative" if acc else "reference")
        f, g = ob.evaluate(dd.pos)
        dg = ob.density_grad.double().cpu().numpy()
        errs[arith] = (abs(f.item() - rf) / abs(rf), rel_l2(dg, parts["density_grad"]),
                       rel_l2(g.double().cpu().numpy(), rg))
    print("  relative errors (objective, density gradient, gradient):", errs, flush=True)
    assert max(errs["accurate"]) <= tol
    if precision == "fp64":  # with BF16x3 its own ~3e-6 dominates at this size
        assert errs["reference"][1] > 10 * errs["accurate"][1]
