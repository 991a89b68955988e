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


@pytest.mark.parametrize("precision", ["bf16x3", "fp32"])
@pytest.mark.parametrize("shape", [(2048, 2048), (2048, 4096)], ids=["2048x2048", "2048x4096"])
def test_large_grid_against_fp64(gpu, shape, precision):
    """Large grids (the BF16x3 form takes its 128-tile GEMM from 256 tiles of
    128 x 128 on): the DCT2 and the Poisson solve (fields, potential and the
    energy from the GEMM epilogue) of the FP32 and BF16x3 forms stay within
    the accuracy gate of the FP64 form on the device (whose kernels the
    oracle pins at the sizes above)."""
    from dreamplace_amd.spectral import get_plan
    print("case large grid %s %s" % (precision, shape), flush=True)
    M, N = shape
    g = torch.Generator().manual_seed(7)
    x64 = (torch.rand(M, N, generator=g, dtype=torch.float64) * 2.0).to(gpu)
    p64 = get_plan(M, N, "fp64", gpu)
    p16 = get_plan(M, N, precision, gpu)
    x32 = x64.float()
    assert rel_l2(p16.dct2(x32).double().cpu().numpy(), p64.dct2(x64).cpu().numpy()) <= 2e-5
    if M != N:
        return
    r64 = p64.poisson(x64, 1.5, 2.0, True, True)
    r16 = p16.poisson(x32, 1.5, 2.0, True, True)
    for a, b in zip(r16[:3], r64[:3]):
        assert rel_l2(a.double().cpu().numpy(), b.cpu().numpy()) <= 1e-5
    assert abs(r16[3].item() - r64[3].item()) <= 1e-5 * abs(r64[3].item())
