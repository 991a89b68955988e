"""Build recipe for ``oracle/_ref``: the reference's own CPU ops, compiled from
their sources where they lie under /root/reference, as test infrastructure.

Only the checker uses what this builds (tests, smoke, the CPU-baseline leg of
bench.py); the product path never loads it.

Each module is a pybind11 extension named exactly as the reference names it
(``dreamplace/ops/*/CMakeLists.txt`` ``add_pytorch_extension(<name>_cpp ...)``)
and is compiled from the reference's own source files plus
``dreamplace/ops/utility/src/msg.cpp``, with the compile definitions the
reference's ``cmake/TorchExtension.cmake:128-136`` passes
(TORCH_EXTENSION_NAME, TORCH_VERSION_MAJOR/MINOR, ENABLE_CUDA).  Nothing is
copied into this repository and no stand-in header is written: torch and
pybind11 come from the installed PyTorch.  Outputs go to ``oracle/_ref/``
only.

Usage:  python oracle/build_ref.py [--jobs N] [--force]
"""
import argparse
import concurrent.futures as cf
import os
import subprocess
import sys
import sysconfig

REF = os.environ.get("DREAMPLACE_REFERENCE_ROOT", "/root/reference")
OPS = os.path.join(REF, "dreamplace", "ops")
HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "_ref")

# module name -> reference sources (relative to dreamplace/ops)
MODULES = {
    "electric_potential_cpp": [
        "electric_potential/src/electric_density_map.cpp",
        "electric_potential/src/electric_force.cpp",
    ],
    "dct2_fft2_cpp": ["dct/src/dct2_fft2.cpp"],
    "weighted_average_wirelength_cpp_merged": [
        "weighted_average_wirelength/src/weighted_average_wirelength_merged.cpp"
    ],
    "logsumexp_wirelength_cpp_merged": [
        "logsumexp_wirelength/src/logsumexp_wirelength_merged.cpp"
    ],
    "pin_pos_cpp": ["pin_pos/src/pin_pos.cpp"],
    "hpwl_cpp": ["hpwl/src/hpwl.cpp"],
    "move_boundary_cpp": ["move_boundary/src/move_boundary.cpp"],
    "pws_cpp": ["pin_weight_sum/src/pws.cpp"],
}


def _torch_flags():
    import torch
    from torch.utils.cpp_extension import include_paths, library_paths

    major, minor = torch.__version__.split(".")[:2]
    inc = ["-I" + p for p in include_paths()]
    inc.append("-I" + sysconfig.get_paths()["include"])
    libs = ["-L" + p for p in library_paths()]
    libs += ["-Wl,-rpath," + p for p in library_paths()]
    libs += ["-ltorch", "-ltorch_cpu", "-lc10", "-ltorch_python"]
    abi = int(torch._C._GLIBCXX_USE_CXX11_ABI)
    return major, minor, inc, libs, abi


def build_one(name, srcs, flags, force=False):
    major, minor, inc, libs, abi = flags
    ext = sysconfig.get_config_var("EXT_SUFFIX")
    target = os.path.join(OUT, name + ext)
    full = [os.path.join(OPS, s) for s in srcs] + [os.path.join(OPS, "utility/src/msg.cpp")]
    if not force and os.path.exists(target):
        newest = max(os.path.getmtime(s) for s in full)
        if os.path.getmtime(target) >= newest:
            return name, True, "up to date"
    cmd = (["g++", "-O3", "-fopenmp", "-ffp-contract=off", "-shared", "-fPIC", "-std=c++17",
            "-w", "-DTORCH_EXTENSION_NAME=" + name, "-DTORCH_VERSION_MAJOR=" + major,
            "-DTORCH_VERSION_MINOR=" + minor, "-DENABLE_CUDA=FALSE",
            "-D_GLIBCXX_USE_CXX11_ABI=%d" % abi, "-I" + OPS] + inc + full + libs
           + ["-o", target + ".tmp"])
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        return name, False, r.stderr[-2000:]
    os.replace(target + ".tmp", target)
    return name, True, "built"


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=4)
    ap.add_argument("--force", action="store_true")
    args = ap.parse_args(argv)
    if not os.path.isdir(OPS):
        print("reference sources not present at %s; oracle/_ref not built" % OPS)
        return 0
    os.makedirs(OUT, exist_ok=True)
    flags = _torch_flags()
    ok = True
    with cf.ThreadPoolExecutor(max_workers=args.jobs) as ex:
        futs = [ex.submit(build_one, n, s, flags, args.force) for n, s in MODULES.items()]
        for f in cf.as_completed(futs):
            name, good, msg = f.result()
            print("[oracle/_ref] %-45s %s" % (name, msg if good else "FAILED\n" + msg))
            ok &= good
    return 0 if ok else 1


if __name__ == "__main__":
    sys.exit(main())
