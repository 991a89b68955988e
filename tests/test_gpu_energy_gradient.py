"""GPU checks of the exact density-energy gradient (dp_energy_gradient,
ElectricPotential(backward="adjoint"), PlacementObjective(density_backward=
"adjoint")): parity with the oracle, central finite differences of the GPU
energy itself, the linear-potential closed form through both HIP gathers, and
the relation of the reference's force to it on the config-0 potential."""
import numpy as np
import pytest
import torch

from _util import density_params, design, to_dev

pytestmark = pytest.mark.gpu

CASES = [
    (0, {}, "config0"),
    (0, {"bins": 45}, "config0-45x45bins"),
    (1, {"scale": 0.002}, "config1-small"),
    (4, {"scale": 0.004, "max_fillers": 200000}, "config4-small-macros"),
]


def _op(d, t, algorithm="tiled", precision=None, backward="adjoint"):
    from dreamplace_amd.electric_potential import ElectricPotential
    return ElectricPotential(
        t["node_size_x"], t["node_size_y"], t["bin_center_x"], t["bin_center_y"],
        d.target_density, d.xl, d.yl, d.xh, d.yh, d.bin_size_x, d.bin_size_y,
        d.num_movable_nodes, d.num_terminals, d.num_filler_nodes, 0, True,
        t["sorted_node_map"], movable_macro_mask=t["movable_macro_mask"],
        algorithm=algorithm, precision=precision, backward=backward)


@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-12), (np.float32, 2e-5)], ids=["f64", "f32"])
@pytest.mark.parametrize("case", CASES, ids=[c[2] for c in CASES])
def test_energy_gradient_matches_oracle(oracle, gpu, case, dtype, tol):
    cfg, kw, name = case
    print("case energy_gradient %s %s" % (name, np.dtype(dtype).name), flush=True)
    d = design(cfg, dtype=dtype, **kw)
    t = to_dev(d, gpu)
    op = _op(d, t)
    phi = np.random.default_rng(11).standard_normal((d.num_bins_x, d.num_bins_y)).astype(dtype)
    g = torch.tensor(0.75, dtype=t["pos"].dtype, device=gpu)
    phi_t = torch.from_numpy(phi).to(gpu)
    out = op.energy_gradient(t["pos"], phi_t, g)
    inter = op.energy_gradient(t["pos"], phi_t, g, grad_layout="interleaved")
    dp = density_params(oracle, d)
    ref = oracle.energy_gradient(dp, d.pos, phi, grad=0.75)
    got = out.cpu().numpy()
    assert np.max(np.abs(got - ref)) <= tol * np.abs(ref).max()
    nm, nt, n = d.num_movable_nodes, d.num_terminals, d.num_nodes
    assert np.all(got[nm:nm + nt] == 0) and np.all(got[n + nm:n + nm + nt] == 0)
    assert torch.equal(inter.view(-1, 2).t().reshape(-1), out)


def _clear_coordinates(rng, d, dp, count, margin):
    n = d.num_nodes
    cells = np.concatenate([np.arange(d.num_movable_nodes), np.arange(n - d.num_filler_nodes, n)])
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


@pytest.mark.parametrize("case", CASES[:3], ids=[c[2] for c in CASES[:3]])
def test_energy_gradient_central_differences(oracle, gpu, case):
    """The module's backward="adjoint" gradient against central differences
    of its own energy (FP64, FP64 solve), at 100 coordinates of movable and
    filler cells whose edges keep clear of bin edges.  Between bin edges the
    energy is quadratic in a coordinate, so the differences are exact up to
    the map's 2^-32 fixed-point quantum divided by the step: h = 1e-2*b
    agrees to 1e-6 of the largest derivative, h = 1e-4*b to 1e-3."""
    cfg, kw, name = case
    print("case energy_gradient FD %s" % name, flush=True)
    d = design(cfg, dtype=np.float64, **kw)
    t = to_dev(d, gpu)
    op = _op(d, t, precision="fp64")
    pos = t["pos"].clone().requires_grad_(True)
    op(pos).backward()
    g = pos.grad.cpu().numpy()
    dp = density_params(oracle, d)
    rng = np.random.default_rng(3)
    n = d.num_nodes
    with torch.no_grad():
        for hf, tol in ((1e-2, 1e-6), (1e-4, 1e-3)):
            h = hf * d.bin_size_x
            fd, an = [], []
            for i, ax in _clear_coordinates(rng, d, dp, 100, 3 * h):
                p = t["pos"].clone()
                p[ax * n + i] += h
                ep = op(p).item()
                p[ax * n + i] -= 2 * h
                em = op(p).item()
                fd.append((ep - em) / (2 * h))
                an.append(g[ax * n + i])
            fd, an = np.array(fd), np.array(an)
            err = np.abs(fd - an).max() / np.abs(fd).max()
            print("  h = %g b: max error %.2e of the largest derivative" % (hf, err), flush=True)
            assert err <= tol, hf


@pytest.mark.parametrize("dtype", [np.float64, np.float32], ids=["f64", "f32"])
def test_linear_potential_closed_form(oracle, gpu, dtype):
    """Both HIP gathers on a linear potential phi = a*(k+1/2) + c*(h+1/2)
    with its exact field (-a, -c), config 0: per cell inside the die the
    exact gradient is 2/(bx*by)*ratio*a*(kR - kL)*sy and the reference
    force ratio*a*sx*sy, so force = (bx^2*by/2) * (sx/bx)/(kR - kL) * exact."""
    print("case energy_gradient linear potential %s" % np.dtype(dtype).name, flush=True)
    d = design(0, dtype=dtype)
    t = to_dev(d, gpu)
    op = _op(d, t, algorithm="global")
    M, Nb, n = d.num_bins_x, d.num_bins_y, d.num_nodes
    bx, by = d.bin_size_x, d.bin_size_y
    a, c = 0.37, -1.3
    phi = (a * (np.arange(M)[:, None] + 0.5) + c * (np.arange(Nb)[None, :] + 0.5)).astype(dtype)
    one = torch.ones((), dtype=t["pos"].dtype, device=gpu)
    g = op.energy_gradient(t["pos"], torch.from_numpy(phi).to(gpu), one).double().cpu().numpy()
    f = op.electric_force(t["pos"], torch.full((M, Nb), -a, dtype=t["pos"].dtype, device=gpu),
                          torch.full((M, Nb), -c, dtype=t["pos"].dtype, device=gpu),
                          one).double().cpu().numpy()
    dp = density_params(oracle, d)
    cells = np.concatenate([np.arange(d.num_movable_nodes), np.arange(n - d.num_filler_nodes, n)])
    x = (d.pos[cells] + dp.offx[cells]).astype(np.float64)
    y = (d.pos[n + cells] + dp.offy[cells]).astype(np.float64)
    sx, sy = dp.wxc[cells].astype(np.float64), dp.wyc[cells].astype(np.float64)
    r = dp.ratio[cells].astype(np.float64)
    inside = (x > d.xl + bx) & (x + sx < d.xh - bx) & (y > d.yl + by) & (y + sy < d.yh - by)
    # cells whose edges sit within rounding of a bin edge may take either bin
    eps = 1e-3 if dtype == np.float32 else 1e-9
    for v, b in ((x, bx), (x + sx, bx), (y, by), (y + sy, by)):
        inside &= np.abs(v / b - np.round(v / b)) > eps
    cells, x, y, sx, sy, r = (v[inside] for v in (cells, x, y, sx, sy, r))
    kspan = np.floor((x + sx - d.xl) / bx) - np.floor((x - d.xl) / bx)
    hspan = np.floor((y + sy - d.yl) / by) - np.floor((y - d.yl) / by)
    # FP32: the exact gradient is phi(kR) - phi(kL) summed from terms of
    # |phi| up to ~330, a cancellation that costs ~5e-5 of the difference
    rt = 1e-12 if dtype == np.float64 else 5e-4
    coef = 2.0 / (bx * by)
    assert np.allclose(g[cells], coef * r * a * kspan * sy, rtol=rt, atol=0)
    assert np.allclose(g[n + cells], coef * r * c * hspan * sx, rtol=rt, atol=0)
    assert np.allclose(f[cells], r * a * sx * sy, rtol=rt, atol=0)
    assert np.allclose(f[cells], (bx * bx * by / 2) * (sx / bx) / kspan * g[cells], rtol=rt, atol=0)


def test_force_against_exact_gradient_at_config0(gpu):
    """On the config-0 potential the reference's force (x part) is the
    energy's exact x-derivative times ~bx^2*by/2 = 4: the least-squares factor
    over the cells lies within 15% of 4 and the two agree in direction
    (cosine >= 0.9).  (Row-height cells, stretched to 1.41 bins, sample phi
    at edge bins 1 or 2 apart, so the y parts agree only on average.)"""
    print("case force vs exact gradient config0", flush=True)
    d = design(0)
    t = to_dev(d, gpu)
    force_op = _op(d, t, precision="fp64", backward="force")
    adj_op = _op(d, t, precision="fp64", backward="adjoint")
    grads = []
    for op in (force_op, adj_op):
        pos = t["pos"].clone().requires_grad_(True)
        op(pos).backward()
        grads.append(pos.grad.cpu().numpy())
    f, g = grads
    n, m0, m1 = d.num_nodes, d.num_movable_nodes, d.num_nodes - d.num_filler_nodes
    sel = np.r_[0:m0, m1:n]
    fx, gx = f[sel], g[sel]
    factor = float(np.dot(fx, gx) / np.dot(gx, gx))
    cos = float(np.dot(fx, gx) / np.linalg.norm(fx) / np.linalg.norm(gx))
    print("  x: least-squares factor %.3f (bx^2*by/2 = %.3f), cosine %.3f"
          % (factor, d.bin_size_x ** 2 * d.bin_size_y / 2, cos), flush=True)
    assert abs(factor / (d.bin_size_x ** 2 * d.bin_size_y / 2) - 1) <= 0.15
    assert cos >= 0.9


def test_engine_adjoint_matches_oracle(oracle, gpu):
    """PlacementObjective(density_backward="adjoint"), FP64: the gradient is
    the wirelength gradient plus lambda times the oracle's exact energy
    gradient at the oracle's potential."""
    from dreamplace_amd.engine import DeviceDesign, PlacementObjective
    print("case engine adjoint config0 fp64", flush=True)
    d = design(0)
    lam, gamma = 0.5, d.wa_gamma
    dd = DeviceDesign(d, gpu)
    obj = PlacementObjective(dd, gamma, density_weight=lam, precision="fp64", algorithm="tiled",
                             schedule="pins", energy_mode="spectral", density_backward="adjoint")
    f, g = obj.evaluate(dd.pos)
    rf, _, parts = oracle.objective_and_gradient(d, lam, gamma)
    dp = density_params(oracle, d)
    ref = parts["wl_grad"] + oracle.energy_gradient(dp, d.pos, parts["phi"], grad=lam)
    assert abs(f.item() - rf) <= 1e-10 * abs(rf)
    got = g.cpu().numpy()
    assert np.linalg.norm(got - ref) <= 1e-10 * np.linalg.norm(ref)
