"""GPU parity of the dense spectral transforms and the Poisson solve against
the FP64 closed forms of the oracle.  Tolerances are relative L2 errors:
FP64 1e-12, FP32 2e-6, BF16x3 2e-5 per single transform; the accuracy gate of
the solve (<= 1e-5 for the fields and energy in the FP32 / BF16x3 forms) is
tested at the config-0 map."""
import numpy as np
import pytest
import torch

from _util import density_params, design, rel_l2

pytestmark = pytest.mark.gpu

TOL = {"fp64": 1e-12, "fp32": 2e-6, "bf16x3": 2e-5}
SHAPES = [(64, 128), (128, 64), (256, 256), (20, 36)]


@pytest.mark.parametrize("shape", SHAPES, ids=["%dx%d" % s for s in SHAPES])
@pytest.mark.parametrize("precision", ["fp64", "fp32", "bf16x3"])
@pytest.mark.parametrize("name", ["dct2", "idct2", "idxst_idct", "idct_idxst"])
def test_transform(oracle, gpu, name, precision, shape):
    from dreamplace_amd.spectral import data_dtype, get_plan
    print("case transform %s %s %s" % (name, precision, shape), flush=True)
    M, N = shape
    x = np.random.default_rng(3).standard_normal((M, N))
    plan = get_plan(M, N, precision, gpu)
    xt = torch.from_numpy(x).to(gpu, data_dtype(precision))
    got = getattr(plan, name)(xt).double().cpu().numpy()
    ref = getattr(oracle, name)(xt.double().cpu().numpy())
    assert rel_l2(got, ref) <= TOL[precision]


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-11), ("fp32", 1e-5), ("bf16x3", 1e-5)])
def test_poisson_solve(oracle, gpu, precision, tol):
    """Config-0 density map: fields and energy within `tol` relative L2 of FP64."""
    from dreamplace_amd.spectral import data_dtype, get_plan
    print("case poisson %s" % precision, flush=True)
    d = design(0)
    dp = density_params(oracle, d)
    init = oracle.fixed_density_map(dp, d.pos, d.bin_center_x, d.bin_center_y)
    dmap = oracle.density_map(dp, d.pos, init)
    ref = oracle.poisson_solve(dmap, d.bin_size_x, d.bin_size_y)
    plan = get_plan(d.num_bins_x, d.num_bins_y, precision, gpu)
    dt = torch.from_numpy(dmap).to(gpu, data_dtype(precision))
    fx, fy, phi, energy = plan.poisson(dt, d.bin_size_x, d.bin_size_y, True, True)
    assert rel_l2(fx.double().cpu(), ref["field_x"]) <= tol
    assert rel_l2(fy.double().cpu(), ref["field_y"]) <= tol
    assert rel_l2(phi.double().cpu(), ref["phi"]) <= tol
    assert abs(energy.item() - ref["energy"]) <= tol * abs(ref["energy"])
    # spectral-domain energy (no fourth transform) gives the same value
    _, _, _, e2 = plan.poisson(dt, d.bin_size_x, d.bin_size_y, False, True)
    assert abs(e2.item() - ref["energy"]) <= tol * abs(ref["energy"])
