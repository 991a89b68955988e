"""ctypes binding of libdreamplace_amd.so (the C ABI declared in
include/dreamplace_amd.h).

PyTorch supplies device memory and the current HIP stream; every call passes
raw pointers.  There is no fallback: if the library is missing, or a tensor
is not on the GPU, the call raises.
"""
import ctypes
import os
import threading

import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libdreamplace_amd.so")

DP_FLOAT32, DP_FLOAT64 = 0, 1
DP_INDEX_INT32, DP_INDEX_INT64 = 0, 1
DP_CONVERT_CPU, DP_CONVERT_GPU = 0, 1
DP_SPECTRAL_FP64, DP_SPECTRAL_FP32, DP_SPECTRAL_BF16X3 = 0, 1, 2
DP_SCATTER_GLOBAL, DP_SCATTER_TILED = 0, 1
DP_LAYOUT_SPLIT, DP_LAYOUT_INTERLEAVED = 0, 1
DP_LAYOUT_FILLERS_FINAL = 0x10
DP_ARITH_BIN_RELATIVE = 0x100
DP_NET_WEIGHTING_ADAMS, DP_NET_WEIGHTING_LILITH = 0, 1

PRECISIONS = {"fp64": DP_SPECTRAL_FP64, "fp32": DP_SPECTRAL_FP32, "bf16x3": DP_SPECTRAL_BF16X3}

_P = ctypes.c_void_p
_I = ctypes.c_int
_L = ctypes.c_int64
_D = ctypes.c_double
_S = ctypes.c_size_t

# name -> (restype, argtypes)
_SIGNATURES = {
    "dp_last_error": (ctypes.c_char_p, []),
    "dp_version": (_I, []),
    "dp_fixed_point_scale": (_L, [_D, _D, _D, _D]),
    "dp_density_map_workspace_bytes": (_S, [_L, _L, _L, _L, _I, _I]),
    "dp_density_map": (_I, [_I, _P, _L, _P, _P, _P, _P, _P, _P, _D, _D, _D, _D, _D, _D, _L, _L,
                            _I, _I, _I, _I, _L, _P, _P, _P, _P, _P]),
    "dp_density_map_ex": (_I, [_I, _P, _L, _P, _P, _P, _P, _P, _P, _D, _D, _D, _D, _D, _D,
                               _L, _L, _I, _I, _I, _I, _L, _P, _P, _P, _P, _P, _P]),
    "dp_density_map_noalloc": (_I, [_I, _P, _L, _P, _P, _P, _P, _P, _P, _D, _D, _D, _D, _D, _D,
                                    _L, _L, _I, _I, _I, _P, _P, _P, _P]),
    "dp_fixed_density_map": (_I, [_I, _P, _L, _P, _P, _P, _P, _D, _D, _D, _D, _D, _D, _L, _L, _I,
                                  _I, _D, _I, _P, _P, _P]),
    "dp_electric_force": (_I, [_I, _P, _P, _P, _P, _L, _P, _P, _P, _P, _P, _D, _D, _D, _D, _L, _L,
                               _I, _I, _I, _L, _I, _P, _P, _P, _P]),
    "dp_electric_force_fillers_final": (_I, [_I, _P, _P, _P, _P, _L, _P, _P, _P, _P, _P, _D, _D,
                                             _D, _D, _L, _L, _I, _I, _I, _L, _I, _P, _P, _P, _P,
                                             _P, _D, _D, _P]),
    "dp_energy_gradient": (_I, [_I, _P, _P, _P, _L, _P, _P, _P, _P, _P, _D, _D, _D, _D, _L, _L,
                                _I, _I, _I, _P, _P, _P]),
    "dp_reduce_workspace_bytes": (_S, [_L]),
    "dp_density_overflow": (_I, [_I, _P, _L, _D, _P, _P, _P]),
    "dp_spectral_plan_bytes": (_S, [_I, _I, _I]),
    "dp_spectral_plan_init": (_I, [_I, _I, _I, _P, _P]),
    "dp_spectral_workspace_bytes": (_S, [_I, _I, _I]),
    "dp_dct2": (_I, [_I, _P, _I, _I, _P, _P, _P, _P]),
    "dp_idct2": (_I, [_I, _P, _I, _I, _P, _P, _P, _P]),
    "dp_idxst_idct": (_I, [_I, _P, _I, _I, _P, _P, _P, _P]),
    "dp_idct_idxst": (_I, [_I, _P, _I, _I, _P, _P, _P, _P]),
    "dp_poisson_solve": (_I, [_I, _P, _I, _I, _P, _D, _D, _I, _P, _P, _P, _P, _P, _P]),
    "dp_pin_pos_forward": (_I, [_I, _P, _L, _P, _P, _P, _I, _L, _P, _P]),
    "dp_pin_pos_backward": (_I, [_I, _P, _L, _P, _P, _L, _L, _P, _P]),
    "dp_wa_workspace_bytes": (_S, [_L]),
    "dp_wa_merged_forward": (_I, [_I, _P, _L, _P, _P, _L, _P, _P, _P, _P, _P, _P, _P, _P]),
    "dp_lse_merged_forward": (_I, [_I, _P, _L, _P, _P, _L, _P, _P, _P, _P, _P, _P, _P, _P]),
    "dp_wa_merged_backward": (_I, [_I, _P, _P, _L, _P, _P, _P, _P, _P]),
    "dp_wa_merged_backward_recompute": (_I, [_I, _P, _P, _L, _P, _P, _L, _P, _P, _P, _P, _P,
                                             _P]),
    "dp_wa_net_weight_grad": (_I, [_I, _P, _P, _P, _L, _P, _P]),
    "dp_net_weighting": (_I, [_I, _I, _L, _P, _P, _D, _P, _I, _D, _D, _P, _P, _P]),
    "dp_wa_forward_from_nodes": (_I, [_I, _P, _L, _P, _P, _P, _L, _P, _P, _L, _P, _P, _P, _P,
                                      _P, _P, _P, _P]),
    "dp_wa_backward_to_nodes": (_I, [_I, _P, _P, _L, _P, _P, _P, _P, _P, _P, _P, _L, _L, _P, _P]),
    "dp_profile_kernel": (_I, [_I, _P, _P]),
    "dp_wa_pack_nets": (_I, [_P, _L, _P, _P, _P, _P, _P]),
    "dp_wa_prepare_workspace_bytes": (_S, [_L]),
    "dp_wa_pins_prepare": (_I, [_I, _L, _L, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P,
                                _P, _P]),
    "dp_wa_pins_workspace_bytes": (_S, [_L, _I]),
    "dp_wa_forward_pins": (_I, [_I, _P, _L, _L, _P, _P, _P, _P, _P, _L, _P, _L, _L, _P, _P, _P,
                                _P, _P, _P, _P, _P, _P, _P]),
    "dp_wa_forward_pins_centred": (_I, [_I, _P, _L, _L, _P, _P, _P, _P, _P, _L, _P, _L, _L, _P, _P, _P,
                                _P, _P, _P, _P, _P, _P, _P]),
    "dp_lse_forward_pins": (_I, [_I, _P, _L, _L, _P, _P, _P, _P, _P, _L, _P, _L, _L, _P, _P, _P,
                                 _P, _P, _P, _P, _P, _P, _P]),
    "dp_wa_pins_slot_weights": (_I, [_I, _L, _P, _P, _P, _P, _P]),
    "dp_node_grad_from_pins": (_I, [_I, _P, _P, _P, _P, _I, _L, _L, _P, _P]),
    "dp_node_grad_preconditioned": (_I, [_I, _P, _P, _P, _P, _I, _L, _L, _P, _P, _P, _D, _D, _P,
                                         _P, _P, _P, _P]),
    "dp_hpwl": (_I, [_I, _P, _L, _P, _P, _L, _P, _P, _P, _P, _P]),
    "dp_hpwl_from_nets": (_I, [_I, _P, _P, _L, _P, _P, _P, _P, _P]),
    "dp_move_boundary": (_I, [_I, _P, _L, _P, _P, _D, _D, _D, _D, _L, _L, _P]),
    "dp_filler_tile_keys": (_I, [_I, _P, _L, _L, _D, _D, _D, _D, _I, _I, _D, _D, _P, _P, _P]),
    "dp_permute_fillers_workspace_bytes": (_S, [_I, _L, _I]),
    "dp_permute_fillers": (_I, [_I, _P, _I, _L, _L, _P, _P, _P]),
    "dp_pin_weight_sum": (_I, [_I, _P, _P, _P, _P, _L, _L, _P, _P]),
    "dp_optimizer_workspace_bytes": (_S, []),
    "dp_nesterov_extrapolate": (_I, [_I, _L, _P, _P, _P, _P, _P, _P, _P, _P, _P]),
    "dp_nesterov_extrapolate_bounded": (_I, [_I, _L, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _D,
                                             _D, _D, _D, _L, _L, _P, _P, _P, _P]),
    "dp_step_norms": (_I, [_I, _L, _P, _P, _P, _P, _P, _P, _P]),
    "dp_l1_norm": (_I, [_I, _L, _P, _P, _P, _P]),
    "dp_precondition": (_I, [_I, _P, _L, _P, _P, _P, _D, _P]),
    # host-only Bookshelf reader
    "dp_bookshelf_read": (_I, [ctypes.c_char_p, _I, ctypes.POINTER(_P)]),
    "dp_bookshelf_sizes": (_I, [_P, _P]),
    "dp_bookshelf_export": (_I, [_P] + [_P] * 21),
    "dp_bookshelf_free": (None, [_P]),
    "dp_rect_union_area": (_D, [_L, _P, _P, _P, _P, _P]),
}

_lib = None
_lock = threading.Lock()


def lib():
    """The loaded native library; raises RuntimeError if it is not built."""
    global _lib
    if _lib is None:
        with _lock:
            if _lib is None:
                if not os.path.exists(LIB_PATH):
                    raise RuntimeError(
                        "dreamplace_amd: native library %s is not built; run "
                        "`python -m dreamplace_amd.build` (hipcc, gfx950)" % LIB_PATH)
                L = ctypes.CDLL(LIB_PATH)
                for name, (res, args) in _SIGNATURES.items():
                    fn = getattr(L, name)
                    fn.restype = res
                    fn.argtypes = args
                _lib = L
    return _lib


def exported_symbols():
    return list(_SIGNATURES)


def call(name, *args):
    """Call a status-returning entry point and raise on a non-zero status."""
    rc = getattr(lib(), name)(*args)
    if rc != 0:
        msg = lib().dp_last_error().decode(errors="replace")
        raise RuntimeError("%s failed (status %d): %s" % (name, rc, msg))


def query(name, *args):
    return getattr(lib(), name)(*args)


def ptr(t):
    """Device pointer of a tensor (None for None)."""
    if t is None:
        return None
    return ctypes.c_void_p(t.data_ptr())


def ptr_np(a):
    """Host pointer of a C-contiguous numpy array (None for None)."""
    if a is None:
        return None
    if not a.flags["C_CONTIGUOUS"]:
        raise ValueError("host array must be C-contiguous")
    return ctypes.c_void_p(a.ctypes.data)


def host_doubles(values):
    """A host double array for an argument read during the call (None for None)."""
    if values is None:
        return None
    return (ctypes.c_double * len(values))(*[float(v) for v in values])


def require_gpu(*tensors, name="dreamplace_amd"):
    for t in tensors:
        if t is not None and not t.is_cuda:
            raise RuntimeError("%s: tensors must be on the GPU (got a %s tensor); the "
                               "dreamplace_amd ops run only on the HIP path" % (name, t.device))


def require_contiguous(*tensors, name="dreamplace_amd"):
    for t in tensors:
        if t is not None and not t.is_contiguous():
            raise RuntimeError("%s: tensor must be contiguous" % name)


def dtype_code(t):
    if t.dtype == torch.float32:
        return DP_FLOAT32
    if t.dtype == torch.float64:
        return DP_FLOAT64
    raise RuntimeError("dreamplace_amd: dtype %s not implemented (float32/float64 only)" % t.dtype)


def stream_ptr(device=None):
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def as_i32(t):
    return t if t.dtype == torch.int32 else t.to(torch.int32)


def as_u8(t):
    if t is None:
        return None
    return t if t.dtype == torch.uint8 else t.to(torch.uint8)


class Workspace:
    """Grow-only scratch buffers, one per (key, device)."""

    def __init__(self):
        self._bufs = {}

    def get(self, key, nbytes, device):
        k = (key, str(device))
        b = self._bufs.get(k)
        if b is None or b.numel() < nbytes:
            b = torch.empty(max(int(nbytes), 256), dtype=torch.uint8, device=device)
            self._bufs[k] = b
        return b


workspace = Workspace()
