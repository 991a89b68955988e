"""Drop-in for dreamplace/ops/dct/dct2_fft2.py: DCT2, IDCT2, IDCT_IDXST and
IDXST_IDCT modules with the reference constructor ``(expkM=None, expkN=None)``
and ``forward(x)`` (dct2_fft2.py:31-150).

The transforms are dense products with precomputed cosine / sine matrices on
the matrix cores (see spectral.py), so the twiddle tables ``expkM`` / ``expkN``
are accepted for interface parity and not used.  ``precision`` selects
"fp64", "fp32" or "bf16x3" (default: fp64 for float64 input, fp32 for float32).
"""
from torch import nn

from .spectral import default_precision, get_plan


class _Transform(nn.Module):
    _method = None

    def __init__(self, expkM=None, expkN=None, precision=None):
        super(_Transform, self).__init__()
        self.expkM = expkM
        self.expkN = expkN
        self.precision = precision

    def forward(self, x):
        prec = self.precision or default_precision(x.dtype)
        plan = get_plan(x.size(-2), x.size(-1), prec, x.device)
        return getattr(plan, self._method)(x)


class DCT2(_Transform):
    """out[u,v] = 4/(MN) sum_ij x[i,j] cos(pi u(2i+1)/2M) cos(pi v(2j+1)/2N)."""
    _method = "dct2"


class IDCT2(_Transform):
    """out[m,n] = sum_uv c_u c_v x[u,v] cos(pi u(2m+1)/2M) cos(pi v(2n+1)/2N)."""
    _method = "idct2"


class IDXST_IDCT(_Transform):
    """out[m,n] = 2 sum_uv c_v x[u,v] sin(pi u(2m+1)/2M) cos(pi v(2n+1)/2N)."""
    _method = "idxst_idct"


class IDCT_IDXST(_Transform):
    """out[m,n] = 2 sum_uv c_u x[u,v] cos(pi u(2m+1)/2M) sin(pi v(2n+1)/2N)."""
    _method = "idct_idxst"
