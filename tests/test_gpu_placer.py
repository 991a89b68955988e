"""GPU parity of the optimizer and of the global-placement loop against their
CPU restatement (oracle/placer.py)."""
import numpy as np
import pytest
import torch

from _util import design, rel_l2

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("use_bb", [False, True], ids=["nobb", "bb"])
def test_nesterov_matches_oracle_on_quadratic(gpu, use_bb):
    """Same step sizes, extrapolation and line search as the restated
    optimizer on f = 1/2 sum a_i x_i^2 (FP64)."""
    from dreamplace_amd.nesterov import NesterovAcceleratedGradientOptimizer
    from oracle import placer as OP
    print("case nesterov quadratic bb=%s" % use_bb, flush=True)
    n = 5000
    a = np.random.default_rng(0).uniform(0.5, 20.0, n)
    x0 = np.random.default_rng(1).standard_normal(n)
    at = torch.from_numpy(a).to(gpu)

    def f_dev(x):
        return 0.5 * torch.sum(at * x * x), at * x

    x = torch.from_numpy(x0.copy()).to(gpu)
    opt = NesterovAcceleratedGradientOptimizer([x], lr=0.01, obj_and_grad_fn=f_dev, use_bb=use_bb)
    xr = x0.copy()
    ref = OP.Nesterov(xr, 0.01, lambda v: (0.5 * float(np.sum(a * v * v)), a * v), lambda v: None,
                      use_bb=use_bb)
    for _ in range(25):
        opt.step()
        ref.step()
    assert rel_l2(x.cpu().numpy(), xr) <= 1e-10
    if not use_bb:
        assert opt.backtracks == ref.backtracks


def test_nesterov_gradient_into_caller_buffer(gpu):
    """grad_into: the line search evaluates into g_k_1's storage and keeps
    it as g_k; the iterates are the same bits as with the copy, and the
    gradient lands in buffers the optimizer owns."""
    from dreamplace_amd.nesterov import NesterovAcceleratedGradientOptimizer
    print("case nesterov grad_into", flush=True)
    n = 5000
    a = torch.from_numpy(np.random.default_rng(0).uniform(0.5, 20.0, n)).to(gpu)
    x0 = torch.from_numpy(np.random.default_rng(1).standard_normal(n)).to(gpu)
    own = torch.empty(n, dtype=torch.float64, device=gpu)
    targets = []

    def f_copy(x):
        torch.mul(a, x, out=own)
        return 0.5 * torch.sum(a * x * x), own

    def f_into(x, grad_out=None):
        g = own if grad_out is None else grad_out
        if grad_out is not None:
            targets.append(grad_out.data_ptr())
        torch.mul(a, x, out=g)
        return 0.5 * torch.sum(a * x * x), g

    xa, xb = x0.clone(), x0.clone()
    oa = NesterovAcceleratedGradientOptimizer([xa], lr=0.01, obj_and_grad_fn=f_copy, use_bb=False)
    ob = NesterovAcceleratedGradientOptimizer([xb], lr=0.01, obj_and_grad_fn=f_into, use_bb=False,
                                              grad_into=True)
    for _ in range(25):
        oa.step()
        ob.step()
        assert torch.equal(xa, xb)
        assert torch.equal(oa.param_groups[0]["g_k"][0], ob.param_groups[0]["g_k"][0])
    assert oa.backtracks == ob.backtracks
    assert targets and own.data_ptr() not in targets
    assert set(targets) <= {ob.param_groups[0]["g_k"][0].data_ptr(),
                            ob.param_groups[0]["g_k_1"][0].data_ptr()}


def test_best_positions_without_copies(gpu):
    """BestPositions with deferred copies (storage taken from the
    optimizer's v_k_1) holds the same best positions as copying at every
    improvement, through relabellings, roll-backs and injected noise, and
    leaves the optimizer's iterates unchanged."""
    from dreamplace_amd.nesterov import NesterovAcceleratedGradientOptimizer
    from dreamplace_amd.placer import BestPositions
    print("case best positions deferred", flush=True)
    n = 4000
    a = torch.from_numpy(np.random.default_rng(0).uniform(0.5, 20.0, n)).to(gpu)
    x0 = torch.from_numpy(np.random.default_rng(1).standard_normal(n)).to(gpu)

    def f(x, grad_out=None):
        g = torch.empty_like(x) if grad_out is None else grad_out
        torch.mul(a, x, out=g)
        return 0.5 * torch.sum(a * x * x), g

    xa, xb = x0.clone(), x0.clone()
    oa = NesterovAcceleratedGradientOptimizer([xa], lr=0.01, obj_and_grad_fn=f, use_bb=False,
                                              grad_into=True)
    ob = NesterovAcceleratedGradientOptimizer([xb], lr=0.01, obj_and_grad_fn=f, use_bb=False,
                                              grad_into=True)
    bp = BestPositions(oa, defer=True)
    ref = None
    rng = np.random.default_rng(3)
    gen = torch.Generator().manual_seed(4)
    events = {"restore": 0, "permute": 0, "inject": 0, "taken": 0}
    for it in range(80):
        if ref is not None and rng.random() < 0.1:
            bp.restore(xa)
            xb.copy_(ref)
            assert torch.equal(xa, ref)
            events["restore"] += 1
        if rng.random() < 0.1:
            perm = torch.randperm(n, generator=gen).to(gpu)
            fn = lambda v: v.copy_(v[perm])
            for x, o in ((xa, oa), (xb, ob)):
                fn(x.data)
                o.permute_state_(fn, include_param=False)
            bp.permute_(fn)
            if ref is not None:
                ref = ref[perm]
            events["permute"] += 1
        oa.step()
        ob.step()
        assert torch.equal(xa, xb)
        if it == 0 or rng.random() < 0.6:
            bp.improved(xa)
            ref = xa.detach().clone()
        else:
            events["taken"] += int(bp.pending)
            bp.not_improved()
        if rng.random() < 0.1:
            bp.materialize(xa)
            noise = torch.randn(n, generator=gen, dtype=torch.float64).to(gpu) * 1e-3
            xa.data.add_(noise)
            xb.data.add_(noise)
            events["inject"] += 1
        assert torch.equal(bp.value(xa), ref), it
    print("  events", events, flush=True)
    assert all(v > 0 for v in events.values())


def test_placer_matches_oracle(gpu):
    """Eight iterations of the placement loop on config 0 (FP64, FP64 solve):
    the same metrics, density weights, gammas, line-search counts and
    positions as the restated loop."""
    from dreamplace_amd.placer import GlobalPlacer
    from oracle import placer as OP
    print("case placer config0 fp64", flush=True)
    d = design(0)
    gp = GlobalPlacer(d, gpu, dtype=torch.float64, precision="fp64", algorithm="tiled",
                      schedule="pins")
    x, hist = gp.run(8)
    assert gp.filler_sorts == 1  # at set-up: the loop keeps the fillers in tile order
    xr, hr, _ = OP.place(d, 8)
    assert len(hist) == len(hr)
    for h, r in zip(hist, hr):
        for k in ("hpwl", "overflow", "max_density", "density_weight", "gamma"):
            assert h[k] == pytest.approx(r[k], rel=1e-9), k
    assert rel_l2(gp.positions().cpu().numpy(), xr) <= 1e-9


def test_placer_fp32_tracks_fp64(gpu):
    """The FP32 + BF16x3 form of the loop follows the FP64 form on config 0
    over 30 iterations (same HPWL within 1%)."""
    from dreamplace_amd.placer import GlobalPlacer
    print("case placer config0 fp32 vs fp64", flush=True)
    d64 = design(0)
    d32 = design(0, dtype=np.float32)
    h64 = GlobalPlacer(d64, gpu, dtype=torch.float64, precision="fp64", algorithm="tiled",
                       schedule="pins").run(30)[1]
    h32 = GlobalPlacer(d32, gpu, precision="bf16x3", algorithm="tiled", schedule="pins").run(30)[1]
    assert h32[-1]["hpwl"] == pytest.approx(h64[-1]["hpwl"], rel=1e-2)
    assert h32[-1]["hpwl"] < 0.5 * h32[0]["hpwl"]


def test_placer_quadratic_penalty_matches_oracle(gpu):
    """The loop with the quadratic density penalty and a doubled density
    factor in force from the start (the plateau response) follows the
    restated loop (FP64)."""
    from dreamplace_amd.placer import GlobalPlacer
    from oracle import placer as OP
    print("case placer config0 fp64 quadratic penalty", flush=True)
    d = design(0)
    prm = dict(quad_penalty=True, density_factor=2.0)
    gp = GlobalPlacer(d, gpu, params=prm, dtype=torch.float64, precision="fp64", algorithm="tiled",
                      schedule="pins")
    x, hist = gp.run(6)
    xr, hr, info = OP.place(d, 6, params=prm)
    for h, r in zip(hist, hr):
        for k in ("hpwl", "overflow", "density_weight"):
            assert h[k] == pytest.approx(r[k], rel=1e-9), k
    assert rel_l2(gp.positions().cpu().numpy(), xr) <= 1e-9
    assert gp.best["overflow"] == pytest.approx(info["best"]["overflow"], rel=1e-9)


def test_entropy_injection_matches_oracle(gpu):
    """Shrink + Gaussian noise + fixed-node restore on the device equals the
    restated injection with the same draws."""
    from dreamplace_amd.placer import GlobalPlacer
    from oracle import placer as OP
    print("case entropy injection", flush=True)
    d = design(0)
    gp = GlobalPlacer(d, gpu, dtype=torch.float64, precision="fp64", algorithm="tiled",
                      schedule="pins")
    x = gp.pos.clone()  # fillers in tile order (sorted at set-up)
    ref = gp.dd.generation_order(x).cpu().numpy()
    gp.entropy_injection(x, 0.996, 47.5)
    OP.entropy_injection(gp.d, ref, 0.996, 47.5, OP.injection_noise(d.num_nodes, 0, 0))
    # each filler keeps the draw of its generated index
    assert np.allclose(gp.dd.generation_order(x).cpu().numpy(), ref, rtol=1e-13, atol=1e-9)


@pytest.mark.parametrize("dtype", [np.float64, np.float32], ids=["f64", "f32"])
def test_fused_preconditioner_equals_separate(gpu, dtype):
    """The preconditioner folded into the node-gradient pass gives the same
    bits as the separate dp_precondition pass after the evaluation."""
    from dreamplace_amd import _native as N
    from dreamplace_amd.engine import DeviceDesign, PlacementObjective
    print("case fused preconditioner %s" % np.dtype(dtype).name, flush=True)
    d = design(0, dtype=dtype)
    dd = DeviceDesign(d, gpu)
    ob = PlacementObjective(dd, d.wa_gamma, density_weight=3e-5, precision="fp32", algorithm="tiled",
                            schedule="pins")
    n = dd.num_nodes
    code = N.dtype_code(dd.pos)
    pws = torch.empty(n, dtype=dd.dtype, device=gpu)
    N.call("dp_pin_weight_sum", code, N.ptr(dd.net_weights), N.ptr(dd.flat_node2pin_map),
           N.ptr(dd.flat_node2pin_start_map), N.ptr(dd.pin2net_map), dd.num_physical_nodes, n,
           N.ptr(pws), N.stream_ptr(gpu))
    areas = (dd.node_size_x * dd.node_size_y).contiguous()
    _, g = ob.evaluate(dd.pos)
    sep = g.clone()
    N.call("dp_precondition", code, N.ptr(sep), n, N.ptr(pws), N.ptr(areas), N.ptr(ob.lam), 1.0,
           N.stream_ptr(gpu))
    _, g2 = ob.evaluate(dd.pos, precondition=(pws, areas, 1.0))
    assert torch.equal(g2, sep)
    # the fillers' one area given instead of read per node: the same bits
    nf = dd.num_filler_nodes
    assert nf > 0 and dd.uniform_fillers()
    _, g3 = ob.evaluate(dd.pos, precondition=(pws, areas, 1.0, float(areas[n - nf].item())))
    assert torch.equal(g3, sep)


def test_filler_sorting_keeps_the_loop(gpu):
    """Relabelling the fillers into tile order every 3 iterations (with the
    optimizer state and the best positions) leaves the FP64 loop on config 0
    where it is without relabelling: only FP64 summation orders change."""
    from dreamplace_amd.placer import GlobalPlacer
    print("case placer filler sorting", flush=True)
    d = design(0)
    runs = []
    for interval in (0, 3):
        gp = GlobalPlacer(d, gpu, params=dict(filler_sort_interval=interval, early_stop=False),
                          dtype=torch.float64, precision="fp64", algorithm="tiled", schedule="pins")
        _, hist = gp.run(10)
        runs.append((gp, hist))
    assert runs[0][0].filler_sorts == 0 and runs[1][0].filler_sorts == 4
    for h0, h1 in zip(runs[0][1], runs[1][1]):
        for k in ("hpwl", "overflow", "density_weight"):
            assert h1[k] == pytest.approx(h0[k], rel=1e-9), k
    p0, p1 = (gp.positions().cpu().numpy() for gp, _ in runs)
    assert rel_l2(p1, p0) <= 1e-9
    # the relabelled run's fillers are in tile order (up to the moves since the last sort)
    assert not torch.equal(runs[1][0].pos, runs[1][0].positions())


@pytest.mark.parametrize("dtype", [np.float64, np.float32], ids=["f64", "f32"])
def test_fused_move_boundary_is_bit_identical(gpu, dtype):
    """The optimizer's extrapolation with the die clamp fused in gives the
    same bits as extrapolation followed by move_boundary: identical metric
    histories and positions over 12 iterations on config 0."""
    from dreamplace_amd.placer import GlobalPlacer
    print("case placer fused move_boundary %s" % np.dtype(dtype).name, flush=True)
    d = design(0, dtype=dtype)
    runs = []
    for fuse in (False, True):
        gp = GlobalPlacer(d, gpu, params=dict(fuse_move_boundary=fuse, early_stop=False),
                          precision="fp64" if dtype == np.float64 else "bf16x3", algorithm="tiled",
                          schedule="pins")
        _, hist = gp.run(12)
        runs.append((gp.pos.clone(), hist))
    assert torch.equal(runs[0][0], runs[1][0])
    for h0, h1 in zip(runs[0][1], runs[1][1]):
        assert h0 == h1


@pytest.mark.parametrize("dtype", [np.float64, np.float32], ids=["f64", "f32"])
def test_fused_metrics_keep_the_loop(gpu, dtype):
    """Metrics taken from the objective's last evaluation (when it was at the
    same positions) give the same loop, bit for bit, as the separate passes."""
    from dreamplace_amd.placer import GlobalPlacer
    print("case placer fused metrics %s" % np.dtype(dtype).name, flush=True)
    d = design(0, dtype=dtype)
    runs = []
    for fused in (False, True):
        gp = GlobalPlacer(d, gpu, params=dict(fused_metrics=fused, early_stop=False),
                          precision="fp64" if dtype == np.float64 else "bf16x3", algorithm="tiled",
                          schedule="pins")
        assert gp.objective.track_metrics == fused
        _, hist = gp.run(12)
        runs.append((gp.pos.clone(), hist))
    assert torch.equal(runs[0][0], runs[1][0])
    for h0, h1 in zip(runs[0][1], runs[1][1]):
        assert h0 == h1


def test_bookshelf_design_through_the_loop(gpu, tmp_path):
    """The reference's `simple` Bookshelf fixture, read and converted
    (bookshelf.to_design), placed for 6 iterations (FP64): the same metrics
    and positions as the restated loop.  run() returns the positions in the
    design's node order, so the .pl written from them puts every movable
    node where the restated loop puts it."""
    import os

    from dreamplace_amd import bookshelf as B
    from dreamplace_amd.placer import GlobalPlacer
    from oracle import placer as OP
    print("case bookshelf simple through the loop", flush=True)
    here = os.path.dirname(os.path.abspath(__file__))
    db = B.read(os.path.join(here, "golden", "bookshelf_simple", "simple.aux"))
    d = B.to_design(db, target_density=1.0)
    gp = GlobalPlacer(d, gpu, dtype=torch.float64, precision="fp64", algorithm="tiled",
                      schedule="pins")
    pos, hist = gp.run(6)
    xr, hr, _ = OP.place(d, 6)
    assert len(hist) == len(hr)
    for h, r in zip(hist, hr):
        for k in ("hpwl", "overflow", "density_weight", "gamma"):
            assert h[k] == pytest.approx(r[k], rel=1e-9), k
    assert torch.equal(pos, gp.dd.generation_order(gp.pos))
    assert rel_l2(pos.cpu().numpy(), xr) <= 1e-9
    x, y = B.positions_in_file_units(d, pos.cpu().numpy())
    out = tmp_path / "simple.gp.pl"
    B.write_pl(str(out), db, x, y)
    rows = {}
    for line in out.read_text().splitlines()[1:]:
        t = line.split()
        if t:
            rows[t[0]] = (float(t[1]), float(t[2]))
    xe, ye = B.positions_in_file_units(d, xr)
    for i in range(db.num_movable_nodes):
        got = rows[db.node_names[i]]
        assert got[0] == pytest.approx(xe[i], rel=1e-5) and got[1] == pytest.approx(ye[i], rel=1e-5)


@pytest.mark.parametrize("precision", ["fp64", "bf16x3"])
def test_fused_step_norms_bit_identical(gpu, precision):
    """The step sizes' FP64 norms formed inside the extrapolation and the
    node-gradient passes (fused_norms, the default) give the loop the same
    bits as the separate dp_step_norms pass: 40 iterations of config 0 with
    identical metrics, line-search counts and positions."""
    from dreamplace_amd.placer import GlobalPlacer
    print("case fused step norms %s" % precision, flush=True)
    dt = np.float64 if precision == "fp64" else np.float32
    d = design(0, dtype=dt)
    runs = []
    for fused in (True, False):
        gp = GlobalPlacer(d, gpu, params=dict(fused_norms=fused, early_stop=False),
                          dtype=torch.float64 if dt == np.float64 else torch.float32,
                          precision=precision, algorithm="tiled", schedule="pins")
        pos, hist = gp.run(40)
        runs.append((pos.cpu(), hist, list(gp.optimizer.backtracks)))
    (p0, h0, b0), (p1, h1, b1) = runs
    assert b0 == b1
    assert h0 == h1
    assert torch.equal(p0, p1)
