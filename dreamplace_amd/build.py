"""Build libdreamplace_amd.so for gfx950 with hipcc, in-tree.

    python -m dreamplace_amd.build [--force] [--jobs N]

Each csrc/*.hip translation unit (and each host-only csrc/*.cpp, with g++) is
compiled to an object under
dreamplace_amd/_build/ and the objects are linked into
dreamplace_amd/libdreamplace_amd.so.  Flags:

* ``--offload-arch=gfx950``: CDNA4 only.
* ``-ffp-contract=off``: no a*b+c -> FMA contraction, so the density scatter's
  bin indices and overlaps round exactly as the reference's.
* ``-Xclang -target-feature -Xclang -packed-fp32-ops``: no packed FP32 VALU ops.
"""
import argparse
import concurrent.futures as cf
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libdreamplace_amd.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")
CXXFLAGS = ["-O2", "-fPIC", "-std=c++17", "-Wall"]

CFLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result", "-Wno-unknown-pragmas",
]


def _deps_newer(obj, src):
    if not os.path.exists(obj):
        return True
    t = os.path.getmtime(obj)
    deps = [src] + glob.glob(os.path.join(CSRC, "*.h")) + [os.path.join(HERE, "..", "include", "dreamplace_amd.h")]
    return any(os.path.getmtime(d) > t for d in deps if os.path.exists(d))


def _compile(src, force):
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    if not force and not _deps_newer(obj, src):
        return obj, True, ""
    if src.endswith(".cpp"):  # host-only sources
        cmd = [CXX] + CXXFLAGS + ["-c", src, "-o", obj]
    else:
        cmd = [HIPCC] + CFLAGS + ["-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    return obj, r.returncode == 0, (r.stdout + r.stderr)


def build(force=False, jobs=8, verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")) + glob.glob(os.path.join(CSRC, "*.cpp")))
    objs, ok, logs = [], True, []
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        for obj, good, log in ex.map(lambda s: _compile(s, force), srcs):
            objs.append(obj)
            ok &= good
            if log.strip() and (verbose or not good):
                logs.append(log)
    if not ok:
        raise RuntimeError("hipcc failed:\n" + "\n".join(logs))
    if logs:
        print("\n".join(logs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stdout + r.stderr)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--force", action="store_true")
    ap.add_argument("--jobs", type=int, default=8)
    ap.add_argument("-v", "--verbose", action="store_true")
    a = ap.parse_args()
    print(build(a.force, a.jobs, a.verbose))
    sys.exit(0)
