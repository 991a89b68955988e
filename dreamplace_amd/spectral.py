"""Dense spectral transforms and the Poisson solve on the matrix cores.

``SpectralPlan`` holds the cosine / sine matrices of an M x N grid for one
arithmetic form ("fp64", "fp32" or "bf16x3") and the scratch the transforms
need; it wraps dp_dct2 / dp_idct2 / dp_idxst_idct / dp_idct_idxst and
dp_poisson_solve of libdreamplace_amd.
"""
import torch

from . import _native as N


def default_precision(dtype):
    return "fp64" if dtype == torch.float64 else "fp32"


def data_dtype(precision):
    return torch.float64 if precision == "fp64" else torch.float32


class SpectralPlan:
    def __init__(self, M, N_, precision, device):
        if precision not in N.PRECISIONS:
            raise ValueError("precision must be one of %s" % sorted(N.PRECISIONS))
        self.M, self.N = int(M), int(N_)
        self.precision = precision
        self.code = N.PRECISIONS[precision]
        self.device = torch.device(device)
        self.dtype = data_dtype(precision)
        nbytes = N.query("dp_spectral_plan_bytes", self.M, self.N, self.code)
        self.plan = torch.empty(nbytes, dtype=torch.uint8, device=self.device)
        N.call("dp_spectral_plan_init", self.M, self.N, self.code, N.ptr(self.plan),
               N.stream_ptr(self.device))
        self.ws = torch.empty(N.query("dp_spectral_workspace_bytes", self.M, self.N, self.code),
                              dtype=torch.uint8, device=self.device)

    def _check(self, x):
        N.require_gpu(x, name="SpectralPlan")
        if x.dtype != self.dtype:
            raise RuntimeError("SpectralPlan(%s) expects %s data, got %s" % (self.precision, self.dtype, x.dtype))
        if x.shape[-2:] != (self.M, self.N):
            raise RuntimeError("SpectralPlan is for %dx%d grids, got %s" % (self.M, self.N, tuple(x.shape)))
        return x.contiguous()

    def transform(self, name, x, out=None):
        x = self._check(x)
        if out is None:
            out = torch.empty(self.M, self.N, dtype=self.dtype, device=self.device)
        N.call(name, self.code, N.ptr(self.plan), self.M, self.N, N.ptr(x), N.ptr(out),
               N.ptr(self.ws), N.stream_ptr(self.device))
        return out

    def dct2(self, x, out=None):
        return self.transform("dp_dct2", x, out)

    def idct2(self, x, out=None):
        return self.transform("dp_idct2", x, out)

    def idxst_idct(self, x, out=None):
        return self.transform("dp_idxst_idct", x, out)

    def idct_idxst(self, x, out=None):
        return self.transform("dp_idct_idxst", x, out)

    def poisson(self, density_map, bin_size_x, bin_size_y, compute_potential=True,
                compute_energy=True, out=None):
        """(field_x, field_y, potential or None, energy 0-dim tensor or None)."""
        d = self._check(density_map)
        if out is None:
            fx = torch.empty(self.M, self.N, dtype=self.dtype, device=self.device)
            fy = torch.empty_like(fx)
            phi = torch.empty_like(fx) if compute_potential else None
            energy = torch.empty((), dtype=self.dtype, device=self.device) if compute_energy else None
        else:
            fx, fy, phi, energy = out
        N.call("dp_poisson_solve", self.code, N.ptr(self.plan), self.M, self.N, N.ptr(d),
               float(bin_size_x), float(bin_size_y), 1 if compute_potential else 0, N.ptr(fx),
               N.ptr(fy), N.ptr(phi), N.ptr(energy), N.ptr(self.ws), N.stream_ptr(self.device))
        return fx, fy, phi, energy


_plans = {}


def get_plan(M, N_, precision, device):
    key = (int(M), int(N_), precision, str(torch.device(device)))
    p = _plans.get(key)
    if p is None:
        p = SpectralPlan(M, N_, precision, device)
        _plans[key] = p
    return p
