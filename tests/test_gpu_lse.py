"""GPU parity of the log-sum-exp wirelength against the oracle: the drop-in
module (net-parallel, forward, backward, net-weight adjoint) and the
engine's pin-parallel and reference-order schedules."""
import numpy as np
import pytest
import torch

from _util import design, rel_l2, to_dev

pytestmark = pytest.mark.gpu

TOL = {np.float64: 1e-11, np.float32: 2e-5}


@pytest.mark.parametrize("dtype", [np.float64, np.float32], ids=["f64", "f32"])
@pytest.mark.parametrize("cfg,kw", [(0, {"big_nets": 6}), (3, {"scale": 0.001})],
                         ids=["config0-bignets", "config3-small"])
def test_lse_module_forward_backward(oracle, gpu, cfg, kw, dtype):
    from dreamplace_amd.logsumexp_wirelength import LogSumExpWirelength
    print("case lse module cfg%d %s" % (cfg, np.dtype(dtype).name), flush=True)
    d = design(cfg, dtype=dtype, **kw)
    t = to_dev(d, gpu)
    pin_pos = oracle.pin_pos_forward(d.pos, d.pin_offset_x, d.pin_offset_y, d.pin2node_map)
    gamma = torch.tensor(d.wa_gamma, dtype=t["pos"].dtype, device=gpu)
    w = t["net_weights"].clone().requires_grad_(True)
    op = LogSumExpWirelength(t["flat_net2pin_map"], t["flat_net2pin_start_map"], t["pin2net_map"], w,
                             t["net_mask_ignore_large_degrees"], t["pin_mask_ignore_fixed_macros"], gamma,
                             algorithm="merged")
    x = torch.from_numpy(pin_pos).to(gpu).requires_grad_(True)
    wl = op(x)
    rwl, partial, gi = oracle.lse_merged_forward(pin_pos, d.flat_net2pin_map, d.flat_net2pin_start_map,
                                                 d.net_weights, d.net_mask_ignore_large_degrees, d.wa_gamma)
    tol = TOL[dtype]
    assert abs(wl.item() - rwl) <= tol * abs(rwl)
    wl.backward(torch.tensor(1.5, dtype=x.dtype, device=gpu))
    refg = oracle.wa_merged_backward(1.5, gi, d.pin2net_map, d.net_weights, d.net_mask_ignore_large_degrees,
                                     d.pin_mask_ignore_fixed_macros)
    assert np.max(np.abs(x.grad.cpu().numpy() - refg)) <= tol * np.abs(refg).max()
    refw = oracle.wa_net_weight_grad(1.5, partial, d.net_mask_ignore_large_degrees)
    assert np.max(np.abs(w.grad.cpu().numpy() - refw)) <= tol * np.abs(refw).max()


def _lse_chain(oracle, d, gamma):
    """Oracle LSE(PinPos(pos)) and its node gradient (no density term)."""
    pin = oracle.pin_pos_forward(d.pos, d.pin_offset_x, d.pin_offset_y, d.pin2node_map)
    wl, _, gi = oracle.lse_merged_forward(pin, d.flat_net2pin_map, d.flat_net2pin_start_map, d.net_weights,
                                          d.net_mask_ignore_large_degrees, gamma)
    gp = oracle.wa_merged_backward(1.0, gi, d.pin2net_map, d.net_weights, d.net_mask_ignore_large_degrees,
                                   d.pin_mask_ignore_fixed_macros)
    g = oracle.pin_pos_backward(gp, d.flat_node2pin_map, d.flat_node2pin_start_map, d.num_physical_nodes,
                                d.num_nodes)
    return wl, g


@pytest.mark.parametrize("schedule", ["pins", "unfused"])
def test_engine_lse_matches_oracle(oracle, gpu, schedule):
    """Objective with density weight 0: the wirelength term alone."""
    from dreamplace_amd.engine import DeviceDesign, PlacementObjective
    print("case engine lse %s" % schedule, flush=True)
    d = design(0, big_nets=6)
    dd = DeviceDesign(d, gpu)
    ob = PlacementObjective(dd, d.wa_gamma, density_weight=0.0, precision="fp64", algorithm="tiled",
                            schedule=schedule, fused=schedule != "unfused", wirelength="logsumexp")
    f, g = ob.evaluate(dd.pos)
    rwl, rg = _lse_chain(oracle, d, d.wa_gamma)
    assert abs(f.item() - rwl) <= 1e-11 * abs(rwl)
    assert rel_l2(g.cpu().numpy(), rg) <= 1e-11


def test_engine_lse_pins_equals_unfused_with_density(gpu):
    """Full objective (both terms), pin-parallel vs reference order, FP32."""
    from dreamplace_amd.engine import DeviceDesign, PlacementObjective
    print("case engine lse pins == unfused fp32", flush=True)
    d = design(0, dtype=np.float32)
    out = []
    for schedule in ("pins", "unfused"):
        dd = DeviceDesign(d, gpu)
        ob = PlacementObjective(dd, d.wa_gamma, density_weight=0.3, precision="fp32", algorithm="tiled",
                                schedule=schedule, fused=schedule != "unfused", wirelength="logsumexp")
        f, g = ob.evaluate(dd.pos)
        out.append((f.item(), g.cpu().numpy().astype(np.float64)))
    assert out[0][0] == pytest.approx(out[1][0], rel=1e-6)
    assert rel_l2(out[0][1], out[1][1]) <= 1e-6
