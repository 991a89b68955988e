#!/usr/bin/env python
"""Benchmark: ms per density + WA forward+backward iteration at 2.2M cells.

One step is one evaluation of the placement objective and its gradient
(PlaceObj.obj_and_grad_fn minus the preconditioner,
dreamplace/PlaceObj.py:394-411) on the config-1 workload: 2.2M movable std
cells, 25k fixed blocks, ~10.9M fillers, 2.2M nets / ~9.6M pins, 1024x1024
bins, synthetic, all inputs resident in HBM.  The timed form is FP32 with the
spectral solve in BF16x3 (split bf16 on the matrix cores, FP32 accumulate);
the same steps are also timed in FP64 for the speedup, and the FP32/BF16x3
fields and gradient are checked against the FP64 run (relative L2).

Prints one JSON line.  --gpus above 1 is not supported (one GPU per box).
"""
import argparse
import json
import math
import os
import statistics
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "ms per density+WA fwd+bwd iter at 2.2M cells, FP32/BF16x3 vs FP64 GPU, at L2err<=1e-5"
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def algorithmic_bytes(d, kernel, schedule="pins", algorithm="tiled"):
    """Algorithmic HBM bytes of one launch of `kernel` (FP32 data, int32
    indices): every array element the kernel must read or write once."""
    Ns = d.num_movable_nodes + d.num_filler_nodes
    B = d.num_bins_x * d.num_bins_y
    P, E, n, nphys = d.num_pins, d.num_nets, d.num_nodes, d.num_physical_nodes
    if kernel == "density_scatter":
        if algorithm == "tiled":  # one 24-byte record per cell in, int64 map in + out
            return 24 * Ns + 16 * B
        return 28 * Ns + 16 * B  # 7 cell arrays in, int64 map in + out
    if kernel == "electric_force":
        if algorithm == "tiled":  # record in, 2 gradients out, both FP32 field maps in
            return 24 * Ns + 8 * Ns + 8 * B
        return 28 * Ns + 8 * n + 8 * B
    if kernel == "wa_forward":
        if schedule == "pins":
            # per pin: netpin index, pin2net, pin2node, 2 offsets, gathered (x,y),
            # pin mask, node-order index, (gx,gy) out; per net: start, mask, weight, WL
            return 41 * P + 13 * E
        return 32 * P + 13 * E
    if kernel == "wa_pin_pos_backward":
        if schedule == "pins":  # (gx,gy) per pin in node order; start per node; add in, grad out
            return 8 * P + 4 * nphys + 16 * n
        return 21 * P + 4 * nphys + 16 * n
    raise KeyError(kernel)


KERNELS = {  # name -> probe id (include/dreamplace_amd.h)
    "density_scatter": 0,
    "electric_force": 1,
    "wa_forward": 2,
    "wa_pin_pos_backward": 5,
}


def time_steps(torch, obj, pos, steps, warmup):
    for _ in range(warmup):
        obj.evaluate(pos)
    torch.cuda.synchronize()
    starts = [torch.cuda.Event(enable_timing=True) for _ in range(steps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(steps)]
    t0 = time.perf_counter()
    for i in range(steps):
        starts[i].record()
        obj.evaluate(pos)
        ends[i].record()
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) * 1e3
    per = [s.elapsed_time(e) for s, e in zip(starts, ends)]
    return wall / steps, per


def probe_kernel(torch, N, obj, pos, kid, reps):
    """Average duration (ms) of one kernel over `reps` evaluations, timed with
    HIP events the library records around that launch on its stream."""
    evs = []
    for _ in range(reps):
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record()  # creates the underlying hipEvent_t; the library re-records it
        b.record()
        N.call("dp_profile_kernel", kid, a.cuda_event, b.cuda_event)
        obj.evaluate(pos)
        evs.append((a, b))
    N.call("dp_profile_kernel", -1, None, None)
    torch.cuda.synchronize()
    return statistics.mean(a.elapsed_time(b) for a, b in evs)


def cpu_baseline(d, budget_s):
    """The reference's own CPU ops (oracle/_ref) on the same design, one
    warm-up then timed iterations until the budget is used (at least one)."""
    import torch
    from oracle import ref_iteration as R
    if not R.available():
        return None
    threads = min(os.cpu_count() or 1, 16)
    it = R.RefIteration(d, dtype=torch.float32, num_threads=threads, density_weight=1e-3)
    t0 = time.perf_counter()
    it.iterate()
    first = time.perf_counter() - t0
    times = []
    t_end = time.perf_counter() + max(budget_s - first, 0)
    while True:
        t1 = time.perf_counter()
        it.iterate()
        times.append(time.perf_counter() - t1)
        if time.perf_counter() > t_end or len(times) >= 10:
            break
    return dict(value=round(statistics.median(times) * 1e3, 3), unit="ms", cores=threads,
                kind="reference",
                sample="full config-1 iteration through oracle/_ref (the reference's CPU ops, FP32, "
                       "%d threads): 1 warm-up + %d timed, median" % (threads, len(times)))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--config", type=int, default=1)
    ap.add_argument("--scale", type=float, default=1.0)
    ap.add_argument("--precision", default="bf16x3", choices=["bf16x3", "fp32"])
    ap.add_argument("--schedule", default="pins", choices=["pins", "fused", "unfused"],
                    help="pins: pin-parallel WA; fused: net-parallel fused; unfused: reference op order")
    ap.add_argument("--algorithm", default="tiled", choices=["tiled", "global"],
                    help="density scatter / field gather: LDS-binned tiles or per-cell global atomics")
    ap.add_argument("--skip-fp64", action="store_true")
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: only --gpus 1 is supported; this build runs on one MI355X "
              "(multi-GPU sharding is a design note, not a measured path)", file=sys.stderr)
        sys.exit(2)

    import torch
    from dreamplace_amd import _native as N
    from dreamplace_amd import synth
    from dreamplace_amd.engine import DeviceDesign, PlacementObjective
    dev = torch.device("cuda:0")
    N.lib()

    t0 = time.perf_counter()
    d = synth.make_design(args.config, seed=0, dtype=np.float32, scale=args.scale)
    log("[bench] design built in %.1fs: %s" % (time.perf_counter() - t0, synth.summary(d)))
    gamma = d.wa_gamma
    lam = 1e-3
    dd32 = DeviceDesign(d, dev)
    obj32 = PlacementObjective(dd32, gamma, density_weight=lam, precision=args.precision,
                               algorithm=args.algorithm, schedule=args.schedule)
    pos32 = dd32.pos
    ms32, per32 = time_steps(torch, obj32, pos32, args.steps, args.warmup)
    log("[bench] fp32/%s: %.3f ms/step (event median %.3f)" % (args.precision, ms32, statistics.median(per32)))

    # kernel probes: find the dominant kernel, then time it live
    kms = {k: probe_kernel(torch, N, obj32, pos32, kid, 3) for k, kid in KERNELS.items()}
    dom = max(kms, key=kms.get)
    dom_ms = probe_kernel(torch, N, obj32, pos32, KERNELS[dom], max(args.steps, 10))
    dom_bytes = algorithmic_bytes(d, dom, args.schedule, args.algorithm)
    achieved = dom_bytes / (dom_ms * 1e-3) / 1e9
    log("[bench] kernel ms: " + ", ".join("%s %.3f" % kv for kv in kms.items()))

    result = {}
    if not args.skip_fp64:
        obj32.evaluate(pos32)
        g32 = obj32.grad.double().clone()
        fx32 = obj32.field_x.double().clone()
        dd64 = DeviceDesign(d, dev, dtype=torch.float64)
        obj64 = PlacementObjective(dd64, gamma, density_weight=lam, precision="fp64",
                                   algorithm=args.algorithm, schedule=args.schedule)
        ms64, per64 = time_steps(torch, obj64, dd64.pos, max(args.steps // 2, 5), 2)
        log("[bench] fp64: %.3f ms/step" % ms64)
        obj64.evaluate(dd64.pos)
        # density gradient alone (the WA part is identical up to FP32 rounding)
        obj32.evaluate(pos32)
        dg32 = obj32.density_grad.double()
        dg64 = obj64.density_grad
        rel = lambda a, b: float(torch.linalg.norm(a - b) / torch.linalg.norm(b))
        result.update(fp64_ms_per_step=round(ms64, 4), speedup_vs_fp64=round(ms64 / ms32, 3),
                      l2err_field_x=rel(fx32, obj64.field_x), l2err_density_grad=rel(dg32, dg64),
                      l2err_grad=rel(g32, obj64.grad))
        del obj64, dd64
        torch.cuda.empty_cache()

    cpu = None if args.skip_cpu else cpu_baseline(d, args.cpu_budget)

    line = {
        "metric": METRIC,
        "value": round(ms32, 4),
        "unit": "ms",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms32, 4),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {
            "workload": "config%d%s: %d movable cells, %d fixed, %d fillers, %d nets, %d pins, %dx%d bins"
                        % (args.config, "" if args.scale == 1.0 else " scale %g" % args.scale,
                           d.num_movable_nodes, d.num_terminals, d.num_filler_nodes, d.num_nets,
                           d.num_pins, d.num_bins_x, d.num_bins_y),
            "form": "fp32 + %s spectral solve, %s schedule, %s density" % (
                args.precision, args.schedule, args.algorithm),
            "kernel_ms": {k: round(v, 4) for k, v in kms.items()},
            **result,
        },
        "roofline": {
            "kernel": dom,
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": None,
            "kernel_ms": round(dom_ms, 4),
            "algorithmic_bytes": dom_bytes,
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
