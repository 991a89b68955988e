"""The C-ABI library without a GPU: it loads, exports every entry point that
include/dreamplace_amd.h declares, answers its host-side size queries, and
rejects bad arguments with the documented codes before touching the device."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "dreamplace_amd.h")


def _declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(dp_[a-z0-9_]+)\s*\(", text)))


@pytest.fixture(scope="module")
def lib():
    from dreamplace_amd import _native as N
    return N.lib()


def test_every_declared_symbol_is_exported(lib):
    names = _declared()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_the_header():
    from dreamplace_amd import _native as N
    assert sorted(N.exported_symbols()) == _declared()


def test_fixed_point_scale_matches_oracle(lib, oracle):
    for die in [(0, 0, 512, 512), (0, 6, 16728, 11430), (0, 0, 1e4, 1e4), (0, 0, 2e4, 2e4),
                (0, 0, 2.0 ** 20, 2.0 ** 20), (0, 0, 3e5, 3e5)]:
        assert lib.dp_fixed_point_scale(*[float(v) for v in die]) == oracle.fixed_scale(*die)


def test_workspace_queries(lib):
    from dreamplace_amd import _native as N
    nb = 1024 * 1024
    b0 = N.query("dp_density_map_workspace_bytes", 1024, 1024, 1000, 0, 0, 0)
    b1 = N.query("dp_density_map_workspace_bytes", 1024, 1024, 1000, 1000, 0, 0)
    assert b0 >= 8 * nb and b1 - b0 >= 4 * 1000 - 256
    assert N.query("dp_spectral_plan_bytes", 1024, 512, N.DP_SPECTRAL_FP32) == 4 * 3 * (1024 ** 2 + 512 ** 2)
    assert N.query("dp_spectral_plan_bytes", 64, 64, N.DP_SPECTRAL_FP64) == 8 * 6 * 64 * 64
    assert N.query("dp_spectral_workspace_bytes", 1024, 1024, N.DP_SPECTRAL_FP32) >= 7 * 4 * nb
    assert N.query("dp_wa_workspace_bytes", 10) > 0
    assert N.query("dp_reduce_workspace_bytes", 10) > 0


def test_argument_errors_return_codes(lib):
    from dreamplace_amd import _native as N
    fake = ctypes.c_void_p(0x1000)
    # null input arrays
    rc = lib.dp_density_map(0, None, 10, None, None, None, None, None, None, 0., 0., 1., 1., .1, .1,
                            5, 0, 10, 10, 0, 0, 0, None, fake, fake, None, None)
    assert rc == -1 and b"null" in lib.dp_last_error()
    # unsupported dtype is reported before any launch
    rc = lib.dp_move_boundary(7, fake, 4, fake, fake, 0., 0., 1., 1., 2, 0, None)
    assert rc == -2 and b"dtype" in lib.dp_last_error()
    rc = lib.dp_pin_pos_forward(0, fake, 4, fake, fake, fake, 9, 4, fake, None)
    assert rc == -1
    rc = lib.dp_dct2(9, fake, 8, 8, fake, fake, fake, None)
    assert rc == -1 and b"precision" in lib.dp_last_error()
    assert lib.dp_profile_kernel(99, None, None) == -1
    assert lib.dp_profile_kernel(-1, None, None) == 0
    with pytest.raises(RuntimeError, match="dp_move_boundary failed"):
        N.call("dp_move_boundary", 7, fake, 4, fake, fake, 0., 0., 1., 1., 2, 0, None)


def test_fillers_final_argument_errors(lib):
    """dp_electric_force_fillers_final checks its extra arguments before any
    launch: the finished-gradient output, and the fillers' area wherever the
    preconditioner is asked for."""
    from dreamplace_amd import _native as N
    fake = ctypes.c_void_p(0x1000)
    base = [0, fake, fake, fake, fake, 10, fake, fake, fake, fake, fake, 0., 0., 1., 1., 5, 3,
            10, 10, N.DP_SCATTER_TILED, 10, N.DP_LAYOUT_INTERLEAVED, None, fake, fake]
    with pytest.raises(RuntimeError, match="null grad_pos_out"):
        N.call("dp_electric_force_fillers_final", *base, None, None, 0.0, -1.0, None)
    with pytest.raises(RuntimeError, match="fillers' area"):
        N.call("dp_electric_force_fillers_final", *base, fake, fake, 1.0, -1.0, None)
    # the node-gradient passes accept the flag and reject other bits
    with pytest.raises(RuntimeError, match="bad add_layout"):
        N.call("dp_node_grad_from_pins", 0, fake, fake, fake, fake,
               N.DP_LAYOUT_INTERLEAVED | 0x20, 5, 10, fake, None)
