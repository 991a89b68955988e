"""The WA module's two ways of keeping the exp/max terms for the backward
(the stored per-pin derivative, or recomputing it from the pin positions),
and the differentiation study's finite-difference checks at the config-3
shape (log-uniform net weights in [0.5, 4], gamma in {1, 4, 16} x bin width,
h = 1e-4 x bin width in FP64, 1000 random coordinates and 1000 random net
weights)."""
import numpy as np
import pytest
import torch

from _util import design, to_dev

pytestmark = pytest.mark.gpu


def _module(t, w, gamma, terms, **kw):
    from dreamplace_amd.weighted_average_wirelength import WeightedAverageWirelength
    return WeightedAverageWirelength(t["flat_net2pin_map"], t["flat_net2pin_start_map"], t["pin2net_map"], w,
                                     t["net_mask_ignore_large_degrees"], t["pin_mask_ignore_fixed_macros"],
                                     gamma, algorithm="merged", terms=terms, **kw)


def _grads(d, t, gpu, terms, gamma_value, grad_out=1.5):
    gamma = torch.tensor(gamma_value, dtype=t["pos"].dtype, device=gpu)
    w = t["net_weights"].clone().requires_grad_(True)
    op = _module(t, w, gamma, terms)
    from oracle import oracle as o
    x = torch.from_numpy(o.pin_pos_forward(d.pos, d.pin_offset_x, d.pin_offset_y, d.pin2node_map)).to(gpu)
    x.requires_grad_(True)
    wl = op(x)
    wl.backward(torch.tensor(grad_out, dtype=x.dtype, device=gpu))
    return wl.item(), x.grad.clone(), w.grad.clone()


@pytest.mark.parametrize("dtype", [np.float64, np.float32], ids=["f64", "f32"])
@pytest.mark.parametrize("cfg,kw", [(0, {"big_nets": 20}), (3, {"scale": 0.002})],
                         ids=["config0-bignets", "config3-small"])
def test_recompute_equals_store(gpu, cfg, kw, dtype):
    print("case wa terms store == recompute cfg%d %s" % (cfg, np.dtype(dtype).name), flush=True)
    d = design(cfg, dtype=dtype, **kw)
    t = to_dev(d, gpu)
    f_s, g_s, w_s = _grads(d, t, gpu, "store", d.wa_gamma)
    f_r, g_r, w_r = _grads(d, t, gpu, "recompute", d.wa_gamma)
    assert f_s == f_r
    assert torch.equal(w_s, w_r)
    # the same device functions compute the derivative in both forms
    assert torch.equal(g_s, g_r)
    # pins of fixed macros (pin_mask) and of masked nets get exactly 0
    pm = torch.from_numpy(d.pin_mask_ignore_fixed_macros.astype(bool)).to(gpu)
    P = d.num_pins
    assert not g_r[:P][pm].any() and not g_r[P:][pm].any()


def test_auto_terms_follow_the_memory_limit(gpu):
    print("case wa terms auto", flush=True)
    d = design(0)
    t = to_dev(d, gpu)
    gamma = torch.tensor(d.wa_gamma, dtype=torch.float64, device=gpu)
    x = torch.zeros(2 * d.num_pins, dtype=torch.float64, device=gpu)
    assert not _module(t, t["net_weights"], gamma, "auto").recomputes(x)
    assert _module(t, t["net_weights"], gamma, "auto", store_limit_bytes=x.numel() * 8 - 1).recomputes(x)
    assert not _module(t, t["net_weights"], gamma, "auto", store_limit_bytes=x.numel() * 8).recomputes(x)
    assert _module(t, t["net_weights"], gamma, "recompute").recomputes(x)
    with pytest.raises(ValueError):
        _module(t, t["net_weights"], gamma, "sometimes")


def test_recompute_rejects_positions_changed_before_backward(gpu):
    print("case wa terms in-place change", flush=True)
    d = design(0)
    t = to_dev(d, gpu)
    gamma = torch.tensor(d.wa_gamma, dtype=torch.float64, device=gpu)
    op = _module(t, t["net_weights"], gamma, "recompute")
    x = torch.rand(2 * d.num_pins, dtype=torch.float64, device=gpu) * 500
    x.requires_grad_(True)
    y = x * 1.0
    wl = op(y)
    with torch.no_grad():
        y.add_(1.0)
    with pytest.raises(RuntimeError):
        wl.backward()


@pytest.fixture(scope="module")
def config3_shape(gpu):
    """Config 3 at 5% scale: ~500K pins on ~125K nets, weights log-uniform."""
    d = design(3, scale=0.05)
    return d, to_dev(d, gpu)


@pytest.mark.parametrize("gamma_bins", [1, 4, 16])
def test_finite_differences_config3(gpu, config3_shape, gamma_bins):
    from dreamplace_amd import gradcheck
    from oracle import oracle as o
    d, t = config3_shape
    print("case wa FD config3 gamma %dxbin" % gamma_bins, flush=True)
    b = d.bin_size_x
    gamma = gamma_bins * b
    h = 1e-4 * b
    pin64 = o.pin_pos_forward(d.pos, d.pin_offset_x, d.pin_offset_y, d.pin2node_map)
    # FP32 study: FP32 weights and positions; the FD runs in FP64 at the same point
    pin32 = torch.from_numpy(pin64.astype(np.float32)).to(gpu)
    w32 = torch.from_numpy(d.net_weights.astype(np.float32)).to(gpu)
    x64 = pin32.double()
    wv64 = w32.double()
    fwd = gradcheck.WaForward64(t["flat_net2pin_map"], t["flat_net2pin_start_map"],
                                t["net_mask_ignore_large_degrees"], gamma)
    rng = np.random.default_rng(11 + gamma_bins)
    nets, slots, axes = gradcheck.pick_pin_coordinates(rng, d.flat_net2pin_start_map, 1000)
    pins = d.flat_net2pin_map[slots]
    fd = gradcheck.pin_fd(fwd, x64, wv64, nets, pins, axes, h)
    fd = np.where(d.pin_mask_ignore_fixed_macros[pins], 0.0, fd)
    wnets = rng.choice(d.num_nets, size=1000, replace=False)
    fdw = gradcheck.weight_fd(fwd, x64, wv64, wnets, h)
    P = d.num_pins
    coord = axes * P + pins
    # A weight's difference is taken on the weighted total, whose FP64
    # rounding (about eps*|total|) over 2h bounds the FD itself: ~1e-8 of the
    # largest weight derivative here.  Coordinates difference one net's
    # partial, ~1e-11 of it.
    total = fwd(x64, wv64).item()
    fd_floor_w = 8 * np.finfo(np.float64).eps * abs(total) / (2 * h) / np.abs(fdw).max()
    print("  FD rounding floor of the weight derivatives %.1e" % fd_floor_w, flush=True)
    for dt, x0, w0, tol_x, tol_w in ((torch.float64, x64, wv64, 1e-8, max(1e-9, fd_floor_w)),
                                     (torch.float32, pin32, w32, 2e-4, 1e-5)):
        for terms in ("store", "recompute"):
            x = x0.clone().requires_grad_(True)
            w = w0.clone().requires_grad_(True)
            op = _module(t, w, torch.tensor(gamma, dtype=dt, device=gpu), terms)
            op(x).backward()
            gx = x.grad.double().cpu().numpy()[coord]
            gw = w.grad.double().cpu().numpy()[wnets]
            ex = np.abs(gx - fd).max() / max(1.0, np.abs(fd).max())
            ew = np.abs(gw - fdw).max() / np.abs(fdw).max()
            print("  %s %s: coordinates %.2e, weights %.2e" % (dt, terms, ex, ew), flush=True)
            assert ex <= tol_x, (dt, terms, ex)
            assert ew <= tol_w, (dt, terms, ew)
