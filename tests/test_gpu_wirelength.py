"""GPU parity of pin_pos, the merged WA wirelength (forward, backward, the new
net-weight adjoint), HPWL, move_boundary and the pin-weight sum against the
oracle.  Integer-indexed gathers are exact; FP results within the tolerance
stated per test."""
import numpy as np
import pytest
import torch

from _util import design, to_dev

pytestmark = pytest.mark.gpu

TOL = {np.float64: 1e-11, np.float32: 2e-5}


def _setup(d, gpu):
    t = to_dev(d, gpu)
    return t


@pytest.mark.parametrize("dtype", [np.float64, np.float32], ids=["f64", "f32"])
def test_pin_pos(oracle, gpu, dtype):
    from dreamplace_amd.pin_pos import PinPos
    print("case pin_pos %s" % np.dtype(dtype).name, flush=True)
    d = design(0, dtype=dtype)
    t = _setup(d, gpu)
    op = PinPos(t["pin_offset_x"], t["pin_offset_y"], t["pin2node_map"], t["flat_node2pin_map"],
                t["flat_node2pin_start_map"], d.num_physical_nodes)
    pos = t["pos"].clone().requires_grad_(True)
    pp = op(pos)
    ref = oracle.pin_pos_forward(d.pos, d.pin_offset_x, d.pin_offset_y, d.pin2node_map)
    assert np.array_equal(pp.detach().cpu().numpy(), ref)
    g = np.random.default_rng(2).standard_normal(2 * d.num_pins).astype(dtype)
    pp.backward(torch.from_numpy(g).to(gpu))
    refg = oracle.pin_pos_backward(g, d.flat_node2pin_map, d.flat_node2pin_start_map,
                                   d.num_physical_nodes, d.num_nodes)
    assert np.array_equal(pos.grad.cpu().numpy(), refg)


@pytest.mark.parametrize("dtype", [np.float64, np.float32], ids=["f64", "f32"])
@pytest.mark.parametrize("cfg,kw", [(0, {}), (3, {"scale": 0.001})], ids=["config0", "config3-small"])
def test_wa_forward_backward(oracle, gpu, cfg, kw, dtype):
    from dreamplace_amd.weighted_average_wirelength import WeightedAverageWirelength
    print("case wa cfg%d %s" % (cfg, np.dtype(dtype).name), flush=True)
    d = design(cfg, dtype=dtype, **kw)
    t = _setup(d, gpu)
    pin_pos = oracle.pin_pos_forward(d.pos, d.pin_offset_x, d.pin_offset_y, d.pin2node_map)
    gamma = torch.tensor(d.wa_gamma, dtype=t["pos"].dtype, device=gpu)
    w = t["net_weights"].clone().requires_grad_(True)
    op = WeightedAverageWirelength(t["flat_net2pin_map"], t["flat_net2pin_start_map"],
                                   t["pin2net_map"], w, t["net_mask_ignore_large_degrees"],
                                   t["pin_mask_ignore_fixed_macros"], gamma, algorithm="merged")
    x = torch.from_numpy(pin_pos).to(gpu).requires_grad_(True)
    wl = op(x)
    rwl, partial, gi = oracle.wa_merged_forward(pin_pos, d.flat_net2pin_map, d.flat_net2pin_start_map,
                                                d.net_weights, d.net_mask_ignore_large_degrees,
                                                d.wa_gamma)
    tol = TOL[dtype]
    assert abs(wl.item() - rwl) <= tol * abs(rwl)
    wl.backward(torch.tensor(1.5, dtype=x.dtype, device=gpu))
    refg = oracle.wa_merged_backward(1.5, gi, d.pin2net_map, d.net_weights,
                                     d.net_mask_ignore_large_degrees, d.pin_mask_ignore_fixed_macros)
    got = x.grad.cpu().numpy()
    assert np.max(np.abs(got - refg)) <= tol * np.abs(refg).max()
    refw = oracle.wa_net_weight_grad(1.5, partial, d.net_mask_ignore_large_degrees)
    assert np.max(np.abs(w.grad.cpu().numpy() - refw)) <= tol * np.abs(refw).max()


def test_wa_net_weight_grad_fd(gpu):
    """Central finite differences in FP64 on random net weights."""
    from dreamplace_amd.weighted_average_wirelength import WeightedAverageWirelength
    print("case wa net-weight FD", flush=True)
    d = design(3, scale=0.0005)
    t = _setup(d, gpu)
    from oracle import oracle as o
    pin_pos = torch.from_numpy(o.pin_pos_forward(d.pos, d.pin_offset_x, d.pin_offset_y,
                                                 d.pin2node_map)).to(gpu)
    gamma = torch.tensor(d.wa_gamma, dtype=torch.float64, device=gpu)
    w = t["net_weights"].clone().requires_grad_(True)
    op = WeightedAverageWirelength(t["flat_net2pin_map"], t["flat_net2pin_start_map"],
                                   t["pin2net_map"], w, t["net_mask_ignore_large_degrees"],
                                   t["pin_mask_ignore_fixed_macros"], gamma, algorithm="merged")
    op(pin_pos).backward()
    idx = np.random.default_rng(5).choice(d.num_nets, size=50, replace=False)
    h = 1e-4
    for e in idx:
        wp = w.detach().clone()
        wp[e] += h
        wm = w.detach().clone()
        wm[e] -= h
        op.net_weights = wp
        fp = op(pin_pos).item()
        op.net_weights = wm
        fm = op(pin_pos).item()
        fd = (fp - fm) / (2 * h)
        assert abs(fd - w.grad[e].item()) <= 1e-6 * max(1.0, abs(fd))


@pytest.mark.parametrize("dtype", [np.float64, np.float32], ids=["f64", "f32"])
def test_hpwl(oracle, gpu, dtype):
    from dreamplace_amd.hpwl import HPWL
    print("case hpwl %s" % np.dtype(dtype).name, flush=True)
    d = design(0, dtype=dtype)
    t = _setup(d, gpu)
    pin_pos = oracle.pin_pos_forward(d.pos, d.pin_offset_x, d.pin_offset_y, d.pin2node_map)
    op = HPWL(t["flat_net2pin_map"], t["flat_net2pin_start_map"], t["pin2net_map"], t["net_weights"],
              t["net_mask_all"], algorithm="net-by-net")
    got = op.forward_fp64(torch.from_numpy(pin_pos).to(gpu)).item()
    ref = oracle.hpwl(pin_pos, d.flat_net2pin_map, d.flat_net2pin_start_map, d.net_weights, d.net_mask_all)
    assert abs(got - ref) <= 1e-12 * ref


@pytest.mark.parametrize("dtype", [np.float64, np.float32], ids=["f64", "f32"])
def test_move_boundary_and_pws(oracle, gpu, dtype):
    from dreamplace_amd.move_boundary import MoveBoundary
    from dreamplace_amd.pin_weight_sum import PinWeightSum
    print("case move_boundary/pws %s" % np.dtype(dtype).name, flush=True)
    d = design(0, dtype=dtype)
    t = _setup(d, gpu)
    pos = d.pos.copy()
    rng = np.random.default_rng(4)
    pos += rng.normal(0, 40, size=pos.shape).astype(dtype)
    p = torch.from_numpy(pos).to(gpu)
    MoveBoundary(t["node_size_x"], t["node_size_y"], d.xl, d.yl, d.xh, d.yh, d.num_movable_nodes,
                 d.num_filler_nodes)(p)
    ref = oracle.move_boundary(pos.copy(), d.node_size_x, d.node_size_y, d.xl, d.yl, d.xh, d.yh,
                               d.num_movable_nodes, d.num_filler_nodes)
    assert np.array_equal(p.cpu().numpy(), ref)
    w = np.exp(rng.uniform(np.log(0.5), np.log(4), size=d.num_nets)).astype(dtype)
    pws = PinWeightSum(t["flat_node2pin_map"], t["flat_node2pin_start_map"], t["pin2net_map"],
                       d.num_nodes)(torch.from_numpy(w).to(gpu))
    refp = oracle.pin_weight_sum(w, d.flat_node2pin_map, d.flat_node2pin_start_map, d.pin2net_map,
                                 d.num_nodes)
    assert np.array_equal(pws.cpu().numpy(), refp)
