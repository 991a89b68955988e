import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for _p in (ROOT, os.path.join(ROOT, "tests")):
    if _p not in sys.path:
        sys.path.insert(0, _p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")
    config.addinivalue_line("markers", "slow: long-running")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as o
    o.build()
    return o


def _ref_dir():
    return os.path.join(ROOT, "oracle", "_ref")


@pytest.fixture(scope="session")
def ref():
    """The reference's own CPU ops (oracle/_ref), or skip when not built."""
    d = _ref_dir()
    if not os.path.isdir(d) or not any(f.endswith(".so") for f in os.listdir(d)):
        pytest.skip("oracle/_ref not built (needs the reference sources)")
    import importlib
    import torch  # noqa: F401  (the modules link against libtorch)
    if d not in sys.path:
        sys.path.insert(0, d)
    mods = {}
    for name in ("electric_potential_cpp", "dct2_fft2_cpp", "weighted_average_wirelength_cpp_merged",
                 "logsumexp_wirelength_cpp_merged", "pin_pos_cpp", "hpwl_cpp", "move_boundary_cpp",
                 "pws_cpp"):
        try:
            mods[name] = importlib.import_module(name)
        except ImportError as e:  # pragma: no cover
            pytest.skip("oracle/_ref module %s unavailable: %s" % (name, e))
    return mods


@pytest.fixture(scope="session")
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda:0")
