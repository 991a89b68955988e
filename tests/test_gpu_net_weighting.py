"""GPU checks of the timing-driven net weighting (dp_net_weighting,
net_weighting.NetWeighting / WirelengthTimer, GlobalPlacer(timing_opt=True))
against the oracle's restatement of the ADAMS and LILITH schemes."""
import math

import numpy as np
import pytest
import torch

from _util import design

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-14), (np.float32, 2e-6)], ids=["f64", "f32"])
@pytest.mark.parametrize("scheme", ["adams", "lilith"])
def test_net_weighting_matches_oracle(oracle, gpu, scheme, dtype, tol):
    """Three updates in a row on 50k nets (degrees 2..150, so some are above
    ignore_net_degree), random slacks with a negative WNS, LILITH capped."""
    from dreamplace_amd.net_weighting import NetWeighting
    print("case net weighting %s %s" % (scheme, np.dtype(dtype).name), flush=True)
    rng = np.random.default_rng(2)
    E = 50000
    deg = rng.integers(2, 151, size=E)
    start = np.zeros(E + 1, dtype=np.int32)
    start[1:] = np.cumsum(deg)
    w0 = rng.uniform(0.5, 4.0, size=E).astype(dtype)
    nw = NetWeighting(torch.from_numpy(start).to(gpu), scheme=scheme, momentum_decay_factor=0.5,
                      max_net_weight=30.0, ignore_net_degree=100,
                      dtype=torch.float64 if dtype == np.float64 else torch.float32)
    w_dev = torch.from_numpy(w0.copy()).to(gpu)
    w_ref, c_ref = w0.copy(), np.zeros(E, dtype=dtype)
    for k in range(3):
        slack = rng.normal(0.0, 1.0, size=E).astype(dtype)
        wns = float(slack.min())
        flags = (slack < -1.0).astype(np.uint8)
        nw.update(w_dev, net_slack=torch.from_numpy(slack).to(gpu), wns=wns,
                  net_critical=torch.from_numpy(flags).to(gpu))
        w_ref, c_ref = oracle.net_weighting(scheme, w_ref, c_ref, start, 100, net_slack=slack,
                                            wns=wns, net_critical=flags, decay=0.5,
                                            max_net_weight=30.0)
    got_w = w_dev.cpu().numpy()
    got_c = nw.net_criticality.cpu().numpy()
    assert np.allclose(got_w, w_ref, rtol=tol, atol=0)
    assert np.allclose(got_c, c_ref, rtol=tol, atol=tol)
    assert np.array_equal(got_w[deg > 100], w0[deg > 100])


@pytest.mark.parametrize("scheme", ["lilith", "adams"])
def test_placer_timing_driven_weights(oracle, gpu, scheme):
    """GlobalPlacer(timing_opt=True) on config 0 (FP64) with updates after
    iterations 3, 6 and 9: the reports of its wirelength-delay timer (whose
    delays are the net-weight adjoint), replayed through the oracle's
    weighting, give the device weights; the preconditioner's pin weights and
    the objective then use them (the objective at the final positions equals
    the oracle's with the new weights)."""
    from dreamplace_amd.net_weighting import WirelengthTimer
    from dreamplace_amd.placer import GlobalPlacer
    print("case placer timing-driven %s" % scheme, flush=True)
    d = design(0)
    reports = []
    timer = WirelengthTimer(quantile=0.9)

    def recording(placer):
        r = timer(placer)
        reports.append({k: (v.cpu().numpy() if torch.is_tensor(v) else v) for k, v in r.items()})
        return r

    prm = dict(timing_opt=True, net_weighting_scheme=scheme, timing_start=2, timing_interval=3,
               max_net_weight=math.inf, early_stop=False)
    gp = GlobalPlacer(d, gpu, params=prm, dtype=torch.float64, precision="fp64",
                      algorithm="tiled", schedule="pins", timer=recording)
    gp.objective.keep_partial_wl = True
    gp.run(10)
    assert gp.net_weight_updates == 3 and len(reports) == 3
    w, c = d["net_weights"].copy(), np.zeros(d.num_nets)
    start = d["flat_net2pin_start_map"]
    for r in reports:
        assert r["wns"] < 0
        w, c = oracle.net_weighting(scheme, w, c, start, 100, net_slack=r["net_slack"], wns=r["wns"],
                                    net_critical=r["net_critical"], decay=0.5)
    got = gp.dd.net_weights.cpu().numpy()
    assert np.allclose(got, w, rtol=1e-13, atol=0)
    assert not np.allclose(got, d["net_weights"])
    dd = gp.dd
    pws = oracle.pin_weight_sum(got, dd.flat_node2pin_map.cpu().numpy(),
                                dd.flat_node2pin_start_map.cpu().numpy(),
                                dd.pin2net_map.cpu().numpy(), d.num_nodes)
    assert np.allclose(gp.pws.cpu().numpy(), pws, rtol=1e-14, atol=0)
    # the objective at the loop's positions, with the new weights, in the design's order
    f, _ = gp.objective.evaluate(gp.pos)
    d2 = dict(d)
    d2["pos"] = gp.positions().cpu().numpy()
    d2["net_weights"] = got  # nets keep their ids under the node relabelling
    rf, _, _ = oracle.objective_and_gradient(d2, gp.objective.lam_host, gp.gamma)
    assert f.item() == pytest.approx(rf, rel=1e-10)
