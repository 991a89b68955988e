"""GPU parity of one full objective+gradient evaluation (the bench step)
against the oracle and against the reference's own CPU ops (oracle/_ref)."""
import numpy as np
import pytest
import torch

from _util import design, rel_l2, relabel_pins

pytestmark = pytest.mark.gpu


def _objective(d, dev, lam, fused, precision, dtype=None, algorithm="global", schedule=None):
    from dreamplace_amd.engine import DeviceDesign, PlacementObjective
    dd = DeviceDesign(d, dev, dtype=dtype)
    return dd, PlacementObjective(dd, d.wa_gamma, density_weight=lam, precision=precision, fused=fused,
                                  algorithm=algorithm, schedule=schedule, keep_partial_wl=True)


@pytest.mark.parametrize("algorithm", ["global", "tiled"])
@pytest.mark.parametrize("fused", [True, False], ids=["fused", "unfused"])
@pytest.mark.parametrize("cfg,kw", [(0, {}), (4, {"scale": 0.003, "max_fillers": 100000})],
                         ids=["config0", "config4-small"])
def test_engine_fp64_matches_oracle(oracle, gpu, cfg, kw, fused, algorithm):
    print("case engine fp64 cfg%d fused=%s %s" % (cfg, fused, algorithm), flush=True)
    d = design(cfg, **kw)
    lam = 0.25
    dd, obj = _objective(d, gpu, lam, fused, "fp64", algorithm=algorithm)
    f, g = obj.evaluate(dd.pos)
    rf, rg, parts = oracle.objective_and_gradient(d, lam, d.wa_gamma)
    assert abs(f.item() - rf) <= 1e-10 * abs(rf)
    assert rel_l2(g.cpu().numpy(), rg) <= 1e-10
    assert np.array_equal(obj.density_map.cpu().numpy(), parts["density_map"])


def test_engine_fused_equals_unfused(gpu):
    print("case engine fused == unfused", flush=True)
    d = design(0)
    dd1, o1 = _objective(d, gpu, 0.3, True, "fp64")
    dd2, o2 = _objective(d, gpu, 0.3, False, "fp64")
    f1, g1 = o1.evaluate(dd1.pos)
    f2, g2 = o2.evaluate(dd2.pos)
    assert torch.equal(g1, g2)
    assert f1.item() == f2.item()


@pytest.mark.parametrize("precision,tol", [("fp32", 2e-5), ("bf16x3", 4e-5)])
def test_engine_fp32_forms(oracle, gpu, precision, tol):
    """FP32 positions with an FP32 or BF16x3 spectral solve, against the FP64
    oracle: relative L2 of the density gradient and of the total gradient."""
    print("case engine fp32/%s" % precision, flush=True)
    d = design(0, dtype=np.float32)
    lam = 0.25
    dd, obj = _objective(d, gpu, lam, True, precision)
    f, g = obj.evaluate(dd.pos)
    d64 = design(0)
    rf, rg, parts = oracle.objective_and_gradient(d64, lam, d64.wa_gamma)
    assert rel_l2(obj.density_grad.double().cpu().numpy(), parts["density_grad"]) <= tol
    assert rel_l2(g.double().cpu().numpy(), rg) <= tol
    assert abs(f.item() - rf) <= tol * abs(rf)


def test_engine_matches_reference_cpu_ops(gpu):
    """Against the reference's own CPU ops (oracle/_ref), config 0, FP64."""
    from oracle import ref_iteration as R
    if not R.available():
        pytest.skip("oracle/_ref not built")
    print("case engine vs oracle/_ref", flush=True)
    d = design(0)
    lam = 0.25
    dd, obj = _objective(d, gpu, lam, True, "fp64")
    f, g = obj.evaluate(dd.pos)
    it = R.RefIteration(d, dtype=torch.float64, density_weight=lam)
    rf, rg, det = it.iterate()
    # the reference scales the map in place by 1/(bx*by) before the DCT
    assert np.array_equal(obj.density_map.cpu().numpy() * (1.0 / (d.bin_size_x * d.bin_size_y)),
                          det["density_map"].numpy())
    assert abs(f.item() - rf.item()) <= 1e-10 * abs(rf.item())
    assert rel_l2(g.cpu().numpy(), rg.numpy()) <= 1e-10


@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-10), (np.float32, 2e-5)], ids=["f64", "f32"])
@pytest.mark.parametrize("cfg,kw", [(0, {"big_nets": 40}), (1, {"scale": 0.003}),
                                    (4, {"scale": 0.003, "max_fillers": 100000, "big_nets": 10})],
                         ids=["config0-bignets", "config1-small", "config4-small"])
def test_engine_pins_schedule(oracle, gpu, cfg, kw, dtype, tol):
    """Pin-parallel WA + node-order backward + tiled density against the FP64
    oracle (FP64 solve in both forms, so only the data type differs)."""
    print("case engine pins cfg%d %s" % (cfg, np.dtype(dtype).name), flush=True)
    d = design(cfg, dtype=dtype, **kw)
    lam = 0.25
    dd, obj = _objective(d, gpu, lam, True, "fp64", algorithm="tiled", schedule="pins")
    f, g = obj.evaluate(dd.pos)
    d64 = design(cfg, **kw)
    rf, rg, parts = oracle.objective_and_gradient(d64, lam, d64.wa_gamma)
    assert abs(f.item() - rf) <= tol * abs(rf)
    assert rel_l2(obj.partial_wl.double().cpu().numpy(),
                  oracle.wa_merged_forward(oracle.pin_pos_forward(d64.pos, d64.pin_offset_x, d64.pin_offset_y,
                                                                  d64.pin2node_map),
                                           d64.flat_net2pin_map, d64.flat_net2pin_start_map, d64.net_weights,
                                           d64.net_mask_ignore_large_degrees, d64.wa_gamma)[1]) <= tol
    assert rel_l2(g.double().cpu().numpy(), rg) <= tol


@pytest.mark.parametrize("schedule", ["pins", "fused", "unfused"])
def test_engine_pin_relabelling(gpu, schedule):
    """Pin ids in no particular order (flat_net2pin_map not the identity):
    renaming the pins while keeping each net's and node's pin sequence gives
    the identical objective and gradient."""
    print("case engine pin relabelling %s" % schedule, flush=True)
    d = design(0, big_nets=20)
    d2 = relabel_pins(d)
    assert not np.array_equal(d2.flat_net2pin_map, np.arange(d.num_pins))
    dd1, o1 = _objective(d, gpu, 0.3, schedule != "unfused", "fp64", algorithm="tiled",
                         schedule=schedule)
    dd2, o2 = _objective(d2, gpu, 0.3, schedule != "unfused", "fp64", algorithm="tiled",
                         schedule=schedule)
    f1, g1 = o1.evaluate(dd1.pos)
    f2, g2 = o2.evaluate(dd2.pos)
    assert torch.equal(g1, g2)
    assert f1.item() == f2.item()


def test_engine_pins_equals_unfused_fp64(gpu):
    """The pin-parallel schedule keeps the per-pin arithmetic and the CSR
    summation order: FP64 gradients equal the reference-order schedule."""
    print("case engine pins == unfused", flush=True)
    d = design(0, big_nets=20)
    dd1, o1 = _objective(d, gpu, 0.3, True, "fp64", algorithm="tiled", schedule="pins")
    dd2, o2 = _objective(d, gpu, 0.3, False, "fp64")
    f1, g1 = o1.evaluate(dd1.pos)
    f2, g2 = o2.evaluate(dd2.pos)
    assert rel_l2(g1.cpu().numpy(), g2.cpu().numpy()) <= 1e-14
    assert abs(f1.item() - f2.item()) <= 1e-13 * abs(f2.item())


def test_engine_pins_net_weights(gpu):
    """Non-unit net weights on the pins schedule (read per pin slot) equal the
    reference-order schedule, also after the weights change."""
    print("case engine pins net weights", flush=True)
    d = design(0, big_nets=20)
    rng = np.random.default_rng(8)
    d.net_weights = np.exp(rng.uniform(np.log(0.5), np.log(4.0), d.num_nets))
    dd1, o1 = _objective(d, gpu, 0.3, True, "fp64", algorithm="tiled", schedule="pins")
    dd2, o2 = _objective(d, gpu, 0.3, False, "fp64")
    for k in range(2):
        if k:
            w2 = torch.from_numpy(np.exp(rng.uniform(-1.0, 1.5, d.num_nets))).to(gpu)
            o1.set_net_weights(w2)
            o2.set_net_weights(w2)
        f1, g1 = o1.evaluate(dd1.pos)
        f2, g2 = o2.evaluate(dd2.pos)
        assert rel_l2(g1.cpu().numpy(), g2.cpu().numpy()) <= 1e-14
        assert abs(f1.item() - f2.item()) <= 1e-13 * abs(f2.item())


@pytest.mark.parametrize("dtype", [np.float64, np.float32], ids=["f64", "f32"])
def test_engine_overlap_is_deterministic(gpu, dtype):
    """The wirelength forward on its own stream gives the same bits as the
    single-stream schedule, and repeated evaluations give the same bits."""
    from dreamplace_amd.engine import DeviceDesign, PlacementObjective
    print("case engine overlap determinism %s" % np.dtype(dtype).name, flush=True)
    d = design(0, dtype=dtype, big_nets=10)
    out = []
    for overlap in (True, False):
        dd = DeviceDesign(d, gpu)
        ob = PlacementObjective(dd, d.wa_gamma, density_weight=0.3, precision="bf16x3" if dtype == np.float32
                                else "fp64", algorithm="tiled", schedule="pins", overlap=overlap)
        for _ in range(3):
            f, g = ob.evaluate(dd.pos)
            out.append((f.item(), g.clone()))
    for f, g in out[1:]:
        assert f == out[0][0]
        assert torch.equal(g, out[0][1])


@pytest.mark.parametrize("dtype", [np.float64, np.float32], ids=["f64", "f32"])
def test_engine_filler_tile_order(gpu, dtype):
    """Fillers relabelled into tile order: the same objective bits, and the
    same gradient bits once put back in the generation order."""
    from dreamplace_amd.engine import DeviceDesign, PlacementObjective
    print("case engine filler tile order %s" % np.dtype(dtype).name, flush=True)
    d = design(1, dtype=dtype, scale=0.01)  # 102 x 102 bins: 2 x 2 tiles
    out = []
    for sort in (False, True):
        dd = DeviceDesign(d, gpu)
        if sort:
            perm = dd.reorder_fillers()
            assert not torch.equal(perm.cpu(), torch.arange(d.num_filler_nodes))
        ob = PlacementObjective(dd, d.wa_gamma, density_weight=0.3, precision="bf16x3" if dtype == np.float32
                                else "fp64", algorithm="tiled", schedule="pins")
        f, g = ob.evaluate(dd.pos)
        out.append((f.item(), dd.generation_order(g)))
    assert out[0][0] == out[1][0]
    assert torch.equal(out[0][1], out[1][1])


@pytest.mark.parametrize("origin", [(0.0, 0.0), (0.37, -0.21)], ids=["origin0", "origin-shifted"])
@pytest.mark.parametrize("dtype", [np.float64, np.float32], ids=["f64", "f32"])
@pytest.mark.parametrize("cfg,kw", [(0, {}), (1, {"scale": 0.01}), (4, {"scale": 0.004, "max_fillers": 200000})],
                         ids=["config0", "config1-small", "config4-small"])
def test_filler_tile_keys_native(gpu, cfg, kw, dtype, origin):
    """dp_filler_tile_keys gives the keys of the indexing form (including
    fillers on and beyond the die's edges, which clamp), per-tile counts
    equal to a bincount, and the same stable order."""
    from dreamplace_amd.engine import filler_tile_keys, filler_tile_keys_device
    print("case filler tile keys cfg%d %s %s" % (cfg, np.dtype(dtype).name, origin), flush=True)
    d = design(cfg, dtype=dtype, **kw)
    n, nf = d.num_nodes, d.num_filler_nodes
    pos = torch.from_numpy(np.ascontiguousarray(d.pos)).to(gpu)
    # a few fillers pushed past each edge of the die
    f0 = n - nf
    pos[f0:f0 + 3] = torch.tensor([d.xl - 50.0, d.xh + 50.0, d.xh], dtype=pos.dtype)
    pos[n + f0 + 3:n + f0 + 6] = torch.tensor([d.yl - 7.0, d.yh + 7.0, d.yh], dtype=pos.dtype)
    args = (pos, n, nf, float(d.xl), float(d.yl), float(d.bin_size_x), float(d.bin_size_y),
            d.num_bins_x, d.num_bins_y, origin)
    ref = filler_tile_keys(*args)
    keys, counts = filler_tile_keys_device(*args)
    torch.cuda.synchronize()
    ntiles = ((d.num_bins_x + 63) // 64) * ((d.num_bins_y + 63) // 64)
    assert keys.dtype == torch.int16 and keys.numel() == nf
    assert torch.equal(keys.long(), ref)
    assert torch.equal(counts.long(), torch.bincount(ref, minlength=ntiles))
    assert torch.equal(torch.argsort(keys, stable=True), torch.argsort(ref, stable=True))


@pytest.mark.parametrize("count", [1, 3, 6])
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32], ids=["f64", "f32"])
def test_permute_fillers_native(gpu, dtype, count):
    """dp_permute_fillers on several [2N] vectors (in groups of four) gives
    the indexing form's bits; the movable and fixed entries stay."""
    from dreamplace_amd.engine import permute_fillers_many_
    print("case permute fillers %s x%d" % (dtype, count), flush=True)
    n, nf = 100003, 77777
    g = torch.Generator(device="cpu").manual_seed(11)
    vecs = [torch.randn(2 * n, generator=g, dtype=torch.float64).to(gpu, dtype) for _ in range(count)]
    perm = torch.randperm(nf, generator=g).to(gpu)
    want = []
    for v in vecs:
        w = v.clone().view(2, n)
        w[:, n - nf:] = w[:, n - nf:][:, perm]
        want.append(w.view(-1))
    permute_fillers_many_(vecs, n, nf, perm)
    torch.cuda.synchronize()
    for v, w in zip(vecs, want):
        assert torch.equal(v, w)


@pytest.mark.parametrize("arith", ["reference", "bin_relative"])
@pytest.mark.parametrize("dtype", [np.float64, np.float32], ids=["f64", "f32"])
def test_tracked_metrics_equal_the_separate_passes(gpu, dtype, arith):
    """HPWL from the wirelength pass's per-net extremes and overflow from the
    objective's own scatter give the same bits as dp_hpwl on pin positions
    and ElectricOverflow's movables-only scatter, with nets above 64 pins and
    nets the WA mask drops (HPWL counts them)."""
    from dreamplace_amd import _native as N
    from dreamplace_amd.engine import DeviceDesign, PlacementObjective
    print("case engine tracked metrics %s %s" % (np.dtype(dtype).name, arith), flush=True)
    d = design(0, dtype=dtype, big_nets=20)
    assert (np.diff(d.flat_net2pin_start_map) >= 100).any()
    dd = DeviceDesign(d, gpu)
    dd.reorder_fillers()
    ob = PlacementObjective(dd, d.wa_gamma, density_weight=0.3, precision="bf16x3" if dtype == np.float32
                            else "fp64", algorithm="tiled", schedule="pins", track_metrics=True,
                            density_arith=arith)
    ob.evaluate(dd.pos)
    assert ob.metric_ptr == dd.pos.data_ptr()
    got = ob.metric_out.clone()
    code = N.dtype_code(dd.pos)
    pin_pos = torch.empty(2 * d.num_pins, dtype=dd.dtype, device=gpu)
    N.call("dp_pin_pos_forward", code, N.ptr(dd.pos), d.num_nodes, N.ptr(dd.pin_offset_x),
           N.ptr(dd.pin_offset_y), N.ptr(dd.pin2node_map), N.DP_INDEX_INT32, d.num_pins, N.ptr(pin_pos),
           N.stream_ptr(gpu))
    ref = torch.empty(3, dtype=torch.float64, device=gpu)
    ws = torch.empty(N.query("dp_wa_workspace_bytes", d.num_nets), dtype=torch.uint8, device=gpu)
    N.call("dp_hpwl", code, N.ptr(pin_pos), d.num_pins, N.ptr(dd.flat_net2pin_map),
           N.ptr(dd.flat_net2pin_start_map), d.num_nets, N.ptr(dd.net_weights), N.ptr(dd.net_mask_all),
           N.ptr(ws), N.ptr(ref[0:1]), N.stream_ptr(gpu))
    ep = ob.density
    m = ep.density_map(dd.pos, num_filler_nodes=0)
    assert torch.equal(m, ob.movable_map)
    N.call("dp_density_overflow", N.dtype_code(m), N.ptr(m), m.numel(),
           float(ep.target_density) * ep.bin_size_x * ep.bin_size_y, N.ptr(ep.reduce_workspace(gpu)),
           N.ptr(ref[1:3]), N.stream_ptr(gpu))
    assert torch.equal(got, ref), (got, ref)
    # the objective's full map is unchanged by the split scatter
    ob2 = PlacementObjective(dd, d.wa_gamma, density_weight=0.3, precision="bf16x3" if dtype == np.float32
                             else "fp64", algorithm="tiled", schedule="pins", density_arith=arith)
    f2, g2 = ob2.evaluate(dd.pos)
    f1, g1 = ob.evaluate(dd.pos)
    assert torch.equal(ob.density_map, ob2.density_map)
    assert f1.item() == f2.item() and torch.equal(g1, g2)


def _check_filler_ranges(d, gpu, dtype, arith):
    """Fillers taken by tile range instead of binned give the density map,
    objective and gradient bits of the binned path: freshly sorted, after
    the fillers have moved (up to 3 bins: some leave their tile's window and
    take the global path), and with ranges that are not a partition
    (replaced by one range); with the movable-only map of the tracked
    metrics too."""
    from dreamplace_amd.engine import DeviceDesign, PlacementObjective
    prec = "bf16x3" if dtype == np.float32 else "fp64"
    dd = DeviceDesign(d, gpu)
    kw = dict(precision=prec, algorithm="tiled", schedule="pins", track_metrics=True,
              density_arith=arith)
    ref = PlacementObjective(dd, d.wa_gamma, density_weight=0.3, **kw)
    rng = PlacementObjective(dd, d.wa_gamma, density_weight=0.3, **kw)
    dd.reorder_fillers(origin_offset=rng.filler_origin_offset())
    assert rng.update_filler_ranges(dd.pos)
    assert ref.filler_ranges is None

    def same():
        f0, g0 = ref.evaluate(dd.pos)
        f0, g0, m0, mm0 = f0.item(), g0.clone(), ref.density_map.clone(), ref.movable_map.clone()
        f1, g1 = rng.evaluate(dd.pos)
        assert torch.equal(m0, rng.density_map)
        assert torch.equal(mm0, rng.movable_map)
        assert f0 == f1.item()
        assert torch.equal(g0, g1)

    same()
    n, nf = d.num_nodes, d.num_filler_nodes
    b = float(d.bin_size_x)
    gen = torch.Generator(device="cpu").manual_seed(7)
    step = ((torch.rand(2, nf, generator=gen, dtype=torch.float64) - 0.5) * 6 * b).to(gpu, dd.dtype)
    with torch.no_grad():
        dd.pos.view(2, n)[:, n - nf:] += step
    same()
    bad = rng.filler_ranges.clone()
    rng.filler_ranges[1] = bad[-1] + 5  # not monotone, past the end
    same()
    rng.filler_ranges.copy_(bad)


@pytest.mark.parametrize("arith", ["reference", "bin_relative"])
@pytest.mark.parametrize("dtype", [np.float64, np.float32], ids=["f64", "f32"])
def test_filler_ranges_are_exact(gpu, dtype, arith):
    """_check_filler_ranges on 102 x 102 bins (2 x 2 tiles)."""
    print("case engine filler ranges %s %s" % (np.dtype(dtype).name, arith), flush=True)
    _check_filler_ranges(design(1, dtype=dtype, scale=0.01), gpu, dtype, arith)


def test_filler_ranges_at_the_tile_limit(gpu):
    """_check_filler_ranges on 4096 x 4096 bins, 4096 tiles (the tiled
    kernels' limit), with 3.3M cells shrunk so that every one spans at most
    3 x 3 bins (FP32, bin-relative arithmetic, BF16x3 solve)."""
    from _util import fine_grid
    print("case engine filler ranges at 4096 tiles", flush=True)
    d = fine_grid(design(1, dtype=np.float32, scale=0.25), 4096, 0.2)
    _check_filler_ranges(d, gpu, np.float32, "bin_relative")


@pytest.mark.parametrize("dtype", [np.float64, np.float32], ids=["f64", "f32"])
def test_engine_movable_relabelling(gpu, dtype):
    """Movable cells relabelled in the wirelength pass's netlist order
    (reorder_movables) give the same objective bits, and each cell's
    gradient moves with it; generation_order / current_order undo and redo
    the relabelling (with fillers reordered too)."""
    from dreamplace_amd.engine import DeviceDesign, PlacementObjective, movable_netlist_order
    print("case movable relabelling %s" % np.dtype(dtype).name, flush=True)
    d = design(0, dtype=dtype, big_nets=20)
    res = []
    for relabel in (False, True):
        dd = DeviceDesign(d, gpu)
        if relabel:
            order = movable_netlist_order(dd)
            assert not torch.equal(order.cpu(), torch.arange(d.num_movable_nodes))
            dd.reorder_movables(order)
        ob = PlacementObjective(dd, d.wa_gamma, density_weight=0.3, precision="fp64",
                                algorithm="tiled", schedule="pins")
        perm = dd.reorder_fillers(origin_offset=ob.filler_origin_offset())
        ob.update_filler_ranges(dd.pos)
        f, g = ob.evaluate(dd.pos)
        res.append((f.item(), dd.generation_order(g), dd.generation_order(dd.pos), dd))
    assert res[0][0] == res[1][0]
    assert torch.equal(res[0][1], res[1][1])
    assert torch.equal(res[0][2], res[1][2])
    assert torch.equal(res[0][2].cpu(), torch.from_numpy(d.pos).to(res[0][2].dtype))
    dd = res[1][3]
    assert torch.equal(dd.current_order(res[1][2]), dd.pos)
    with pytest.raises(ValueError):
        dd.reorder_movables(torch.zeros(d.num_movable_nodes, dtype=torch.int64))


@pytest.mark.parametrize("ranges", [False, True], ids=["binned", "ranges"])
@pytest.mark.parametrize("dtype", [np.float64, np.float32], ids=["f64", "f32"])
def test_fillers_final_is_bit_identical(gpu, dtype, ranges):
    """The field gather finishing the fillers' gradients itself
    (dp_electric_force_fillers_final with DP_LAYOUT_FILLERS_FINAL) gives the
    bits of the node-gradient pass doing it, zero signs included:
    unpreconditioned, preconditioned with the fused step norm, with the
    fillers binned or taken by tile range after some have left their tile's
    window (the drift list)."""
    from dreamplace_amd import _native as N
    from dreamplace_amd.engine import DeviceDesign, PlacementObjective
    print("case fillers final %s %s" % (np.dtype(dtype).name, ranges), flush=True)
    prec = "bf16x3" if dtype == np.float32 else "fp64"
    d = design(1, dtype=dtype, scale=0.01)
    dd = DeviceDesign(d, gpu)
    ob = PlacementObjective(dd, d.wa_gamma, density_weight=0.3, precision=prec, algorithm="tiled",
                            schedule="pins", density_arith="bin_relative")
    assert ob.fillers_final
    n, nf = dd.num_nodes, dd.num_filler_nodes
    if ranges:
        dd.reorder_fillers(origin_offset=ob.filler_origin_offset())
        assert ob.update_filler_ranges(dd.pos)
        b = float(d.bin_size_x)
        gen = torch.Generator(device="cpu").manual_seed(3)
        step = ((torch.rand(2, nf, generator=gen, dtype=torch.float64) - 0.5) * 6 * b)
        with torch.no_grad():
            dd.pos.view(2, n)[:, n - nf:] += step.to(gpu, dd.dtype)
    code = N.dtype_code(dd.pos)
    pws = torch.empty(n, dtype=dd.dtype, device=gpu)
    N.call("dp_pin_weight_sum", code, N.ptr(dd.net_weights), N.ptr(dd.flat_node2pin_map),
           N.ptr(dd.flat_node2pin_start_map), N.ptr(dd.pin2net_map), dd.num_physical_nodes, n,
           N.ptr(pws), N.stream_ptr(gpu))
    areas = (dd.node_size_x * dd.node_size_y).contiguous()
    farea = float(areas[n - 1].item())
    gen = torch.Generator(device="cpu").manual_seed(4)
    gref = torch.randn(2 * n, generator=gen, dtype=torch.float64).to(gpu, dd.dtype)
    bits = torch.int64 if dtype == np.float64 else torch.int32
    outs = []
    for on in (False, True):
        ob.fillers_final = on
        f, g = ob.evaluate(dd.pos)
        res = [f.item(), g.clone().view(bits), ob.density_grad.clone()]
        _, gp = ob.evaluate(dd.pos, precondition=(pws, areas, 0.7, farea), grad_ref=gref)
        res += [gp.clone().view(bits), ob.last_dg2.item()]
        if on:
            with pytest.raises(RuntimeError):
                ob.density_grad  # noqa: B018 -- preconditioned: the fillers' were not kept
        outs.append(res)
    a, b = outs
    assert a[0] == b[0]
    assert torch.equal(a[1], b[1])
    assert torch.equal(a[2], b[2])
    assert torch.equal(a[3], b[3])
    assert a[4] == b[4]
