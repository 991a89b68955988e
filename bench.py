#!/usr/bin/env python
"""Benchmark: ms per density + WA forward+backward iteration at 2.2M cells.

One step is one evaluation of the placement objective and its gradient
(PlaceObj.obj_and_grad_fn minus the preconditioner,
dreamplace/PlaceObj.py:394-411) on the config-1 workload: 2.2M movable std
cells, 25k fixed blocks, ~10.9M fillers, 2.2M nets / ~9.6M pins, 1024x1024
bins, synthetic, all inputs resident in HBM.  The timed form is FP32 with the
spectral solve in BF16x3 (split bf16 on the matrix cores, FP32 accumulate);
the same steps are also timed in FP64 for the speedup.

Arithmetic: the timed FP32 form evaluates the wirelength relative to each
net's extremes (dp_wa_forward_pins_centred) and every density overlap
relative to the cell's first bin (DP_ARITH_BIN_RELATIVE): the same functions
as the reference's, without its cancellation at die coordinates.  The form
with the reference's own arithmetic (bit-exact with its FP32 map) is timed
too (ms_per_step_reference_arith).

Accuracy (the gate, l2err_gate): the whole step of the timed FP32 form
against the FP64 form at the same positions -- potential, field, energy,
per-cell density gradient and total gradient (l2err_step_*).  The solve
alone on one map (l2err_solve_*) and the reference-arithmetic form's whole
step (l2err_step_*_reference_arith) are reported beside it.

The placement loop also re-sorts the fillers into the density kernels' tile
order every 25 steps; that pass is timed on the same design and reported
with the step it amortizes over (ms_per_filler_resort,
ms_per_step_with_resort_every_25), and so is the step at positions that
have drifted from the sorted order by half a bin
(ms_per_step_drifted_half_bin).

Roofline: each op of the step (one C-ABI entry point, all of its kernels) is
timed live with HIP events recorded by the library around its launches on
its stream; achieved GB/s = the op's algorithmic bytes / its duration.

Prints one JSON line.  --gpus above 1 is not supported (one GPU per box).
"""
import argparse
import contextlib
import ctypes
import json
import os
import statistics
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "ms per density+WA fwd+bwd iter at 2.2M cells, FP32/BF16x3 vs FP64 GPU, at L2err<=1e-5"
HBM_PEAK_GBS = 8000.0
MFMA_PEAK_TFLOPS = {"bf16x3": 2500.0, "fp32": 157.3, "fp64": 78.6}
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_config1.json")

# probe ids (include/dreamplace_amd.h)
DP_OP_DENSITY_MAP, DP_OP_ELECTRIC_FORCE, DP_OP_WA_FORWARD, DP_OP_NODE_GRAD, DP_OP_POISSON = \
    16, 17, 18, 19, 20

# kernels of each op in the pins schedule with tiled density (rocprof names)
OP_KERNELS = {
    "wa_forward": ["k_pin_push", "k_wa_pins", "k_finish (no k_wa_big: no net above 64 pins)"],
    "density_map": ["k_zero_regions", "k_bin_count", "k_bin_scan", "k_bin_fill",
                    "k_tile_scatter_f32(_acc)", "k_scatter_rest", "k_convert_cpu"],
    "poisson": ["k_sum_partial", "k_scale_centre_split", "k_gemm_b3", "k_spectral_weights_split",
                "k_finish"],
    "electric_force": ["k_tile_gather", "k_gather_rest"],
    "node_grad": ["k_node_grad"],
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def op_bytes(d, schedule, fillers_final=False):
    """Algorithmic HBM bytes of one call of each HBM-bound op, from the
    per-unit figures of the hot-path analysis (FP32 data, int32 indices):
    pin_pos fwd 28 B/pin; WA fwd (stored derivative) 20 B/pin + 13 B/net;
    scatter 28 B/cell + 20 B/bin; force 36 B/cell + 8 B/bin; WA bwd fused
    with pin_pos bwd 20 B/pin + 8 B/physical node, plus 16 B/node for adding
    the density gradient into the node gradient (physical nodes only when the
    force writes the fillers' finished gradients, fillers_final)."""
    Ns = d.num_movable_nodes + d.num_filler_nodes
    B = d.num_bins_x * d.num_bins_y
    P, E, n, nphys = d.num_pins, d.num_nets, d.num_nodes, d.num_physical_nodes
    wa = 20 * P + 13 * E
    if schedule in ("pins", "fused"):
        wa += 28 * P  # pin positions are formed inside the op
    ops = {"wa_forward": wa, "density_map": 28 * Ns + 20 * B, "electric_force": 36 * Ns + 8 * B}
    if schedule in ("pins", "fused"):
        ops["node_grad"] = 20 * P + 8 * nphys + 16 * (nphys if fillers_final else n)
    return ops


def poisson_flops(M, N, transforms=4):
    """Dense-transform flops of one solve: `transforms` 2-D transforms of 2
    GEMMs each (DCT2 + two field transforms, + the potential's IDCT2 when the
    energy is taken from the potential map)."""
    return transforms * 2 * (M * M * N + M * N * N)


PROBES = {"wa_forward": DP_OP_WA_FORWARD, "density_map": DP_OP_DENSITY_MAP,
          "poisson": DP_OP_POISSON, "electric_force": DP_OP_ELECTRIC_FORCE,
          "node_grad": DP_OP_NODE_GRAD}


def relabel(dd, movable_order, filler_origin=None):
    """The timed form's node order: movable cells in the wirelength pass's
    netlist order ("netlist") and, when `filler_origin` (the density kernels'
    filler origin offset) is given, fillers in tile order.  Neither changes
    the objective's bits.  Every other form of the same placement (FP64, the
    CPU baseline) is made from dd.host_design(...) afterwards, so that all of
    them hold the same positions, sizes and netlist in the same order."""
    from dreamplace_amd.engine import movable_netlist_order
    if movable_order == "netlist":
        dd.reorder_movables(movable_netlist_order(dd))
    if filler_origin is not None:
        dd.reorder_fillers(origin_offset=filler_origin)


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


def resort_ms(torch, dd, obj, reps):
    """Median ms of the placer's filler re-sort (GlobalPlacer.sort_fillers)
    on this design: dd.reorder_fillers with five more [2N] vectors (the
    optimizer's four and the best positions) and the objective's filler
    ranges.  At the bench's fixed positions the order found is the current
    one, which the passes do not exploit."""
    extra = [dd.pos.clone() for _ in range(5)]
    times = []
    for _ in range(reps + 1):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        dd.reorder_fillers(origin_offset=obj.filler_origin_offset(), extra=extra)
        obj.update_filler_ranges(dd.pos)
        torch.cuda.synchronize()
        times.append((time.perf_counter() - t0) * 1e3)
    return statistics.median(times[1:])


def probe(torch, N, obj, pos, pid, reps):
    """Average duration (ms) of one kernel or op over `reps` evaluations,
    timed with HIP events the library records around its launches on the
    stream they go to."""
    evs = []
    for _ in range(reps):
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record()  # creates the underlying hipEvent_t; the library re-records it
        b.record()
        N.call("dp_profile_kernel", pid, a.cuda_event, b.cuda_event)
        obj.evaluate(pos)
        evs.append((a, b))
    N.call("dp_profile_kernel", -1, None, None)
    torch.cuda.synchronize()
    return statistics.mean(a.elapsed_time(b) for a, b in evs)


def measured_traffic(op):
    """HBM bytes per call of `op` from the committed counter summary of the
    same workload (FETCH_SIZE and WRITE_SIZE per dispatch, each scaled by
    what the counter reads for a plain copy of known size), or None."""
    try:
        with open(TRAFFIC_FILE) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    ops = t.get("ops", {})
    return ops.get(op)


def rel(torch, a, b):
    a = a.double().reshape(-1)
    b = b.double().reshape(-1)
    return float(torch.linalg.norm(a - b) / torch.linalg.norm(b))


def accuracy(torch, obj32, obj64, dd64, suffix=""):
    """Solve-only errors on the same density map, and whole-step FP32-form vs
    FP64-form differences (the gate).  Both objectives were last evaluated at
    the same positions.  `suffix` names a second FP32 form's fields."""
    ep = obj64.density
    m = obj32.density_map.double()
    fx64, fy64, phi64, e64 = obj64.plan.poisson(m, ep.bin_size_x, ep.bin_size_y)
    # the reduced form's potential map (not needed by its step when the energy
    # is spectral), solved here outside the timed region for the gate
    _, _, phi32, _ = obj32.plan.poisson(obj32.density_map.to(obj32.plan.dtype), ep.bin_size_x,
                                        ep.bin_size_y, compute_energy=False)
    one = torch.ones((), dtype=torch.float64, device=m.device)
    pos = dd64.pos
    g64 = ep.electric_force(pos, fx64, fy64, one)
    g32 = ep.electric_force(pos, obj32.field_x.double(), obj32.field_y.double(), one)
    both = lambda a, b: torch.cat([a.reshape(-1), b.reshape(-1)])
    # the FP64 form's potential map (its step takes the energy spectrally)
    _, _, phi64s, _ = obj64.plan.poisson(obj64.density_map, ep.bin_size_x, ep.bin_size_y,
                                         compute_energy=False)
    e64s = float(obj64.energy.double().item())
    step = {
        "l2err_step_potential": rel(torch, phi32, phi64s),
        "l2err_step_field": rel(torch, both(obj32.field_x, obj32.field_y),
                                both(obj64.field_x, obj64.field_y)),
        "l2err_step_energy": abs(float(obj32.energy.double().item()) - e64s) / abs(e64s),
        "l2err_step_density_grad": rel(torch, obj32.density_grad, obj64.density_grad),
        "l2err_step_grad": rel(torch, obj32.grad, obj64.grad),
    }
    if suffix:
        return {k + suffix: v for k, v in step.items()}
    out = {
        "l2err_solve_potential": rel(torch, phi32, phi64),
        "l2err_solve_energy": abs(float(obj32.energy.double().item()) - float(e64.item()))
                              / abs(float(e64.item())),
        "l2err_solve_field": rel(torch, both(obj32.field_x, obj32.field_y), both(fx64, fy64)),
        "l2err_solve_density_grad": rel(torch, g32, g64),
    }
    out.update(step)
    out["l2err_gate"] = max(step.values())
    return out


@contextlib.contextmanager
def stdout_to_stderr():
    """Send file descriptor 1 to stderr, so that native code printing to
    stdout (the reference's CPU ops log through printf) cannot add a line
    to the JSON output; C stdio is flushed before stdout comes back."""
    sys.stdout.flush()
    libc = ctypes.CDLL(None)
    saved = os.dup(1)
    os.dup2(2, 1)
    try:
        yield
    finally:
        libc.fflush(None)
        os.dup2(saved, 1)
        os.close(saved)


def cpu_baseline(d, budget_s):
    with stdout_to_stderr():
        return _cpu_baseline(d, budget_s)


def host_cpu():
    """The host CPU: model name, logical CPUs, whether SMT is on, and the
    CPUs this process may run on (its affinity set)."""
    model, smt = None, None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        with open("/sys/devices/system/cpu/smt/active") as f:
            smt = f.read().strip() == "1"
    except OSError:
        pass
    try:
        affinity = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        affinity = os.cpu_count() or 1
    return dict(model=model, logical_cpus=os.cpu_count(), smt=smt, affinity_cpus=affinity)


def baseline_threads(cpu):
    """Threads for the CPU baseline: the CPU share this process is given
    (OMP_NUM_THREADS where the host sets it, as it does to apportion a shared
    machine's cores per GPU), else every CPU of its affinity set."""
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        return min(int(env), cpu["affinity_cpus"])
    return cpu["affinity_cpus"]


def _cpu_baseline(d, budget_s, warmup=3):
    """The reference's own CPU ops (oracle/_ref) on the same design: `warmup`
    untimed iterations, then timed ones until the budget is used (at least
    one, at most 10); the median."""
    import torch
    from oracle import ref_iteration as R
    if not R.available():
        return None
    cpu = host_cpu()
    threads = baseline_threads(cpu)
    it = R.RefIteration(d, dtype=torch.float32, num_threads=threads, density_weight=1e-3)
    t0 = time.perf_counter()
    for _ in range(warmup):
        it.iterate()
    first = (time.perf_counter() - t0) / max(warmup, 1)
    times = []
    t_end = time.perf_counter() + max(budget_s - first * warmup, 0)
    while True:
        t1 = time.perf_counter()
        it.iterate()
        times.append(time.perf_counter() - t1)
        if time.perf_counter() > t_end or len(times) >= 10:
            break
    return dict(value=round(statistics.median(times) * 1e3, 3), unit="ms", cores=threads,
                kind="reference", cpu_model=cpu["model"], logical_cpus=cpu["logical_cpus"],
                smt=cpu["smt"], affinity_cpus=cpu["affinity_cpus"],
                sample="full config-1 iteration through oracle/_ref (the reference's CPU ops, FP32, "
                       "%d threads = this process's CPU share): %d warm-up + %d timed, median"
                       % (threads, warmup, len(times)))


def summarize_trace(trace_dir, warmup, steps):
    """Text summary of `rocprofv3 --kernel-trace --stats --output-format csv`
    over `bench.py --steps K --warmup W --skip-cpu`: per library kernel, the
    calls and average over the whole run (the --stats table), and its
    average in three phases of the FP32 run: the timed steps (wirelength
    forward on its own stream), the timed one-stream steps, and the per-op
    probe evaluations whose HIP-event durations make `config.ops` and the
    roofline.  A dispatch belongs to evaluation i when i k_node_grad<float>
    dispatches precede it (one per evaluation, its last kernel)."""
    import csv
    import glob
    import re

    def short(name):
        m = re.match(r"(?:void )?([\w:]+(?:<[^>(]*>)?)", name)
        return m.group(1) if m else name[:60]

    stats = list(csv.DictReader(open(glob.glob(os.path.join(trace_dir, "*kernel_stats.csv"))[0])))
    trace = sorted(csv.DictReader(open(glob.glob(os.path.join(trace_dir, "*kernel_trace.csv"))[0])),
                   key=lambda r: int(r["Dispatch_Id"]))
    K, W = steps, warmup
    phases = {"overlap": (W, W + K), "one_stream": (W + K + 2, W + 2 * K + 2),
              "probes": (W + 2 * K + 2, W + 2 * K + 2 + 5 * 3 + max(K, 10))}
    acc = {}
    ev = 0
    for r in trace:
        name = short(r["Kernel_Name"])
        dur = (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3
        for ph, (a, b) in phases.items():
            if a <= ev < b and name.startswith("dp::"):
                acc.setdefault(name, {}).setdefault(ph, []).append(dur)
        if name.startswith("dp::k_node_grad<float"):
            ev += 1
    lines = ["# %-44s %6s %10s %10s | %10s %10s %10s" % ("kernel", "calls", "total", "average",
                                                         "overlap", "one-stream", "probes")]
    for r in stats:
        name = short(r["Name"])
        if not name.startswith("dp::"):
            continue
        cols = []
        for ph in ("overlap", "one_stream", "probes"):
            v = acc.get(name, {}).get(ph)
            cols.append("%10.1f" % statistics.mean(v) if v else "%10s" % "-")
        lines.append("%-46s %6d %10.1f %10.1f | %s" % (name[:46], int(r["Calls"]),
                                                       float(r["TotalDurationNs"]) / 1e3,
                                                       float(r["AverageNs"]) / 1e3, " ".join(cols)))
    lines.append("# evaluations counted: %d; phases (evaluation index ranges): %s" % (ev, phases))
    return "\n".join(lines)


# kernels whose HBM bytes make each op's measured traffic (per-dispatch
# medians; the map clear belongs to the scatter; the wirelength total's
# k_finish reads a few KB and is left out)
TRAFFIC_KERNELS = {
    "wa_forward": ["dp::k_pin_push<", "dp::k_wa_pins<", "dp::k_wa_big<"],
    "density_map": ["dp::k_zero_regions", "dp::k_bin_count<", "dp::k_bin_scan",
                    "dp::k_bin_fill<", "dp::k_tile_scatter", "dp::k_scatter_rest<",
                    "dp::k_convert_cpu<"],
    "electric_force": ["dp::k_tile_gather<", "dp::k_gather_rest<"],
    "node_grad": ["dp::k_node_grad<"],
}


def summarize_traffic(cal_fetch, cal_write, pmc_fetch, pmc_write, workload):
    """HBM bytes per dispatch of each library kernel, and per op, from four
    `rocprofv3 --pmc ... --output-format csv` runs: FETCH_SIZE and WRITE_SIZE
    of `bench.py --calibrate-copy` (1 GiB read and written by one elementwise
    kernel) and of the step.  bytes = median counter (KiB) x 1024 x scale,
    scale = 1 GiB / what the counter reads for the copy."""
    import collections
    import csv
    import glob
    import re

    def short(name):
        m = re.match(r"(?:void )?([\w:]+(?:<[^>(]*>)?)", name)
        return m.group(1) if m else name[:60]

    def counters(d):
        out = collections.defaultdict(list)
        for r in csv.DictReader(open(glob.glob(os.path.join(d, "*counter_collection.csv"))[0])):
            out[short(r["Kernel_Name"])].append(float(r["Counter_Value"]))
        return out

    def copy_kib(d):
        for k, v in counters(d).items():
            if "elementwise" in k:
                return statistics.median(v)
        raise SystemExit("no elementwise copy kernel in " + d)

    gib_kib = (1 << 30) / 1024.0
    fscale, wscale = gib_kib / copy_kib(cal_fetch), gib_kib / copy_kib(cal_write)
    f, w = counters(pmc_fetch), counters(pmc_write)
    kernels = {}
    for k in f:
        if k.startswith("dp::"):
            kernels[k] = {"fetch_bytes": round(statistics.median(f[k]) * 1024 * fscale),
                          "write_bytes": round(statistics.median(w.get(k, [0])) * 1024 * wscale),
                          "dispatches": len(f[k])}
    ops = {}
    for op, prefixes in TRAFFIC_KERNELS.items():
        ops[op] = sum(v["fetch_bytes"] + v["write_bytes"] for k, v in kernels.items()
                      if any(k.startswith(p) for p in prefixes))
    return {"workload": workload,
            "scaling": "bytes = median per-dispatch counter (KiB) x 1024 x scale; scale = 1 GiB / "
                       "what the counter reads for the 1 GiB elementwise copy of --calibrate-copy",
            "fetch_scale": fscale, "write_scale": wscale, "ops": ops, "kernels": kernels}


def wa_terms_study(torch, args):
    """Differentiation study on the config-3 netlist (10M pins, 2.5M nets,
    net weights log-uniform in [0.5, 4]), in FP32 and FP64, gamma in
    {1, 4, 16} x bin width:
    * the WA module's two ways of keeping the exp/max terms for the backward
      (the stored per-pin derivative, or recomputing it from the positions):
      forward and backward ms (events, median), the HBM the forward leaves
      held for the backward and the peak over forward + backward (torch
      allocator statistics above what was allocated before the forward);
    * the gradients against FP64 central differences (h = 1e-4 x bin width)
      at 1000 random pin coordinates and 1000 random net weights; FP32
      gradients are compared at the FP32-rounded point.
    Prints one JSON line per (dtype, gamma, terms)."""
    from dreamplace_amd import _native as N
    from dreamplace_amd import gradcheck, synth
    from dreamplace_amd.pin_pos import PinPos
    from dreamplace_amd.weighted_average_wirelength import WeightedAverageWirelength
    dev = torch.device("cuda:0")
    N.lib()
    t0 = time.perf_counter()
    d = synth.make_design(3, seed=0, dtype=np.float64, scale=args.scale)
    log("[study] design built in %.1fs: %s" % (time.perf_counter() - t0, synth.summary(d)))

    def T(a, dtype=None):
        t = torch.from_numpy(np.ascontiguousarray(a)).to(dev)
        return t if dtype is None else t.to(dtype)

    fnp, nps = T(d.flat_net2pin_map), T(d.flat_net2pin_start_map)
    p2n, net_mask = T(d.pin2net_map), T(d.net_mask_ignore_large_degrees)
    pin_mask = T(d.pin_mask_ignore_fixed_macros)
    pp = PinPos(T(d.pin_offset_x), T(d.pin_offset_y), T(d.pin2node_map), T(d.flat_node2pin_map),
                T(d.flat_node2pin_start_map), d.num_physical_nodes)
    pin32 = pp(T(d.pos)).detach().float()
    w32 = T(d.net_weights, torch.float32)
    x64, w64 = pin32.double(), w32.double()
    P, b = d.num_pins, d.bin_size_x
    pmask = d.pin_mask_ignore_fixed_macros
    rng = np.random.default_rng(0)
    reps = 10
    for gamma_bins in (1, 4, 16):
        gamma = gamma_bins * b
        h = 1e-4 * b
        fwd = gradcheck.WaForward64(fnp, nps, net_mask, gamma)
        nets, slots, axes = gradcheck.pick_pin_coordinates(rng, d.flat_net2pin_start_map, 1000)
        pins = d.flat_net2pin_map[slots]
        t1 = time.perf_counter()
        fd = gradcheck.pin_fd(fwd, x64, w64, nets, pins, axes, h)
        fd = np.where(pmask[pins], 0.0, fd)
        wnets = rng.choice(d.num_nets, size=1000, replace=False)
        fdw = gradcheck.weight_fd(fwd, x64, w64, wnets, h)
        log("[study] gamma %dxbin: FD in %.1fs" % (gamma_bins, time.perf_counter() - t1))
        coord = axes * P + pins
        for dt, x0, w0 in ((torch.float32, pin32, w32), (torch.float64, x64, w64)):
            auto = WeightedAverageWirelength(fnp, nps, p2n, w0, net_mask, pin_mask,
                                             torch.tensor(gamma, dtype=dt, device=dev),
                                             algorithm="merged").recomputes(x0)
            for terms in ("store", "recompute"):
                w = w0.clone().requires_grad_(True)
                op = WeightedAverageWirelength(fnp, nps, p2n, w, net_mask, pin_mask,
                                               torch.tensor(gamma, dtype=dt, device=dev),
                                               algorithm="merged", terms=terms)
                x = x0.detach().clone().requires_grad_(True)
                for _ in range(2):
                    x.grad, w.grad = None, None
                    op(x).backward()
                fms, bms = [], []
                for _ in range(reps):
                    x.grad, w.grad = None, None
                    e0, e1, e2 = (torch.cuda.Event(enable_timing=True) for _ in range(3))
                    e0.record()
                    wl = op(x)
                    e1.record()
                    wl.backward()
                    e2.record()
                    torch.cuda.synchronize()
                    fms.append(e0.elapsed_time(e1))
                    bms.append(e1.elapsed_time(e2))
                x.grad, w.grad = None, None
                del wl
                torch.cuda.synchronize()
                torch.cuda.reset_peak_memory_stats(dev)
                base = torch.cuda.memory_allocated(dev)
                wl = op(x)
                torch.cuda.synchronize()
                held = torch.cuda.memory_allocated(dev) - base
                wl.backward()
                torch.cuda.synchronize()
                peak = torch.cuda.max_memory_allocated(dev) - base
                gx = x.grad.double().cpu().numpy()[coord]
                gw = w.grad.double().cpu().numpy()[wnets]
                rec = dict(study="wa_terms", config=3, dtype=str(dt).replace("torch.", ""),
                           gamma_bins=gamma_bins, terms=terms, num_pins=P, num_nets=d.num_nets,
                           ms_forward=round(statistics.median(fms), 4),
                           ms_backward=round(statistics.median(bms), 4),
                           mb_held_forward_to_backward=round(held / 1e6, 2),
                           mb_peak_forward_backward=round(peak / 1e6, 2),
                           auto_choice="recompute" if auto else "store",
                           fd_max_err_coordinates=float(np.abs(gx - fd).max() / max(1.0, np.abs(fd).max())),
                           fd_max_err_weights=float(np.abs(gw - fdw).max() / np.abs(fdw).max()),
                           fd_points=[len(fd), len(fdw)], fd_h=h)
                print(json.dumps(rec), flush=True)
                del wl, op, x, w


def with_bins(d, nb):
    """A copy of design `d` on an nb x nb bin grid (bin sizes, centres and
    the WA gamma of 4 bin widths follow; the cells are the same)."""
    from dreamplace_amd import synth
    d2 = synth.Design(d)
    bsx, bsy = (d.xh - d.xl) / nb, (d.yh - d.yl) / nb
    d2.update(num_bins_x=nb, num_bins_y=nb, bin_size_x=bsx, bin_size_y=bsy, wa_gamma=4.0 * bsx)
    dt = d.pos.dtype
    d2.bin_center_x = synth._bin_centers(d.xl, d.xh, bsx, nb).astype(dt)
    d2.bin_center_y = synth._bin_centers(d.yl, d.yh, bsy, nb).astype(dt)
    return d2


def precision_study(torch, args):
    """The precision study on config-1 shapes: the Poisson solve at 1024^2
    and 2048^2 bins in three forms -- FP64, FP32 on the matrix cores, and
    BF16x3 (bf16 hi/lo splits, FP32 accumulation) -- all on the same FP32
    density map, the one that is bit-exact with the reference's
    deterministic FP32 map (tests/test_gpu_density.py holds that).  Against
    the FP64 form: relative L2 of the potential, of the field (x and y
    together), of the energy, and of the per-cell density gradient, which the
    FP32 field gather forms from each form's field (so the FP32 gather is in
    the chain; the FP64 row isolates the gather).  The reference gradient is
    the FP64 gather of the FP64 field at the same positions.  ms per solve:
    the step's solve (fields + energy, no potential map), median of HIP-event
    timings.  Prints one JSON line per (grid, form)."""
    from dreamplace_amd import _native as N
    from dreamplace_amd import synth
    from dreamplace_amd.electric_potential import ElectricPotential
    from dreamplace_amd.spectral import get_plan
    dev = torch.device("cuda:0")
    N.lib()
    t0 = time.perf_counter()
    base = synth.make_design(1, seed=0, dtype=np.float32, scale=args.scale)
    log("[study] design built in %.1fs: %s" % (time.perf_counter() - t0, synth.summary(base)))
    reps = 20

    def module(d, t, dtype, algorithm, arithmetic="reference"):
        return ElectricPotential(
            t["node_size_x"].to(dtype), t["node_size_y"].to(dtype), t["bin_center_x"].to(dtype),
            t["bin_center_y"].to(dtype), d.target_density, d.xl, d.yl, d.xh, d.yh, d.bin_size_x,
            d.bin_size_y, d.num_movable_nodes, d.num_terminals, d.num_filler_nodes, 0, True,
            t["sorted_node_map"], movable_macro_mask=t["movable_macro_mask"], algorithm=algorithm,
            arithmetic=arithmetic)

    for nb in ([1024, 2048] if args.scale == 1.0 else [base.num_bins_x, 2 * base.num_bins_x]):
        d = with_bins(base, nb)
        t = {k: torch.from_numpy(np.ascontiguousarray(v)).to(dev) for k, v in d.items()
             if isinstance(v, np.ndarray)}
        pos32, pos64 = t["pos"], t["pos"].double()
        ep32 = module(d, t, torch.float32, "tiled")
        ep32.compute_initial_density_map(pos32)
        m32 = ep32.density_map(pos32)  # also bins the cells for the FP32 gather
        ep32b = module(d, t, torch.float32, "tiled", "bin_relative")
        ep32b.compute_initial_density_map(pos32)
        ep32b.density_map(pos32)  # bins the cells for its gather
        ep64 = module(d, t, torch.float64, "global")
        one32 = torch.ones((), dtype=torch.float32, device=dev)
        one64 = torch.ones((), dtype=torch.float64, device=dev)
        bx, by = d.bin_size_x, d.bin_size_y
        ref = get_plan(nb, nb, "fp64", dev).poisson(m32.double(), bx, by, True, True)
        g_ref = ep64.electric_force(pos64, ref[0], ref[1], one64)
        both = lambda a, b: torch.cat([a.reshape(-1), b.reshape(-1)])
        for prec in ("fp64", "fp32", "bf16x3"):
            plan = get_plan(nb, nb, prec, dev)
            m = m32.double() if prec == "fp64" else m32
            fx, fy, phi, e = plan.poisson(m, bx, by, True, True)
            g = ep32.electric_force(pos32, fx.float(), fy.float(), one32)
            gb = ep32b.electric_force(pos32, fx.float(), fy.float(), one32)
            ms = []
            for _ in range(3):
                plan.poisson(m, bx, by, False, True)
            for _ in range(reps):
                a_, b_ = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                a_.record()
                plan.poisson(m, bx, by, False, True)
                b_.record()
                torch.cuda.synchronize()
                ms.append(a_.elapsed_time(b_))
            med = statistics.median(ms)
            fl = poisson_flops(nb, nb, 3)
            tf = fl / (med * 1e-3) / 1e12
            rec = dict(study="precision", workload="config1 cells (%d cells, %d fillers) on %dx%d bins"
                       % (d.num_movable_nodes, d.num_filler_nodes, nb, nb),
                       form=prec, ms_per_solve=round(med, 4), flops_per_solve=fl,
                       achieved_tflops=round(tf, 1),
                       l2err_potential=rel(torch, phi, ref[2]),
                       l2err_field=rel(torch, both(fx, fy), both(ref[0], ref[1])),
                       l2err_energy=abs(float(e.double().item()) - float(ref[3].item()))
                       / abs(float(ref[3].item())),
                       l2err_density_grad_fp32_gather=rel(torch, g, g_ref),
                       l2err_density_grad_fp32_gather_bin_relative=rel(torch, gb, g_ref))
            if prec != "fp64":
                rec["frac_of_fp32_peak"] = round(tf / MFMA_PEAK_TFLOPS["fp32"], 4)
            if prec == "bf16x3":
                rec["frac_of_bf16_peak"] = round(3 * tf / MFMA_PEAK_TFLOPS["bf16x3"], 4)
            if prec == "fp64":
                rec["frac_of_fp64_peak"] = round(tf / MFMA_PEAK_TFLOPS["fp64"], 4)
            print(json.dumps(rec), flush=True)
        del ep32, ep32b, ep64, t
        torch.cuda.empty_cache()


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
    ap.add_argument("--energy-mode", default="spectral", choices=["spectral", "potential"],
                    help="energy from the DCT coefficients (Parseval; no potential map) or "
                         "from the materialised potential map as the reference does")
    ap.add_argument("--movable-order", default="netlist", choices=["netlist", "generation"],
                    help="movable cells relabelled in the wirelength pass's netlist order "
                         "(same objective bits) or kept in the generated order")
    ap.add_argument("--filler-order", default="tile", choices=["tile", "generation"],
                    help="tile: fillers relabelled into the density kernels' tile order before the "
                         "timed steps and taken by tile range (as the placer keeps them); "
                         "generation: as generated, binned every step")
    ap.add_argument("--arith", default="accurate", choices=["accurate", "reference"],
                    help="arithmetic of the timed FP32 form: accurate (centred wirelength, "
                         "bin-relative density overlaps) or the reference's own")
    ap.add_argument("--skip-fp64", action="store_true")
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--study", choices=["wa-terms", "precision"],
                    help="wa-terms: the config-3 differentiation study (store vs recompute of the WA "
                         "exp/max terms, FD checks); precision: the config-2 precision study (FP64, "
                         "FP32 and BF16x3 solves at 1024^2 and 2048^2 bins); instead of the step "
                         "benchmark")
    ap.add_argument("--summarize-trace", metavar="DIR",
                    help="print the per-phase kernel summary of a rocprofv3 csv kernel trace of "
                         "`bench.py --steps K --warmup W --skip-cpu` (pass the same K and W) and exit")
    ap.add_argument("--summarize-traffic", nargs=4,
                    metavar=("CAL_FETCH", "CAL_WRITE", "PMC_FETCH", "PMC_WRITE"),
                    help="print the HBM-traffic JSON from four rocprofv3 --pmc csv output folders "
                         "(FETCH_SIZE and WRITE_SIZE of --calibrate-copy, then of the step) and exit")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a plain 1 GiB -> 1 GiB device copy kernel 5 times (the known-size "
                         "reference for FETCH_SIZE / WRITE_SIZE counter scaling) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: only --gpus 1 is supported; this build runs on one MI355X "
              "(multi-GPU sharding is a design note, not a measured path)", file=sys.stderr)
        sys.exit(2)

    if args.summarize_traffic:
        wl = ("config1: 2200000 movable cells, 25000 fixed, 10924750 fillers, 2200000 nets, "
              "9623921 pins, 1024x1024 bins; FP32 + bf16x3 solve, pins schedule, tiled density; "
              "fillers in tile order")
        print(json.dumps(summarize_traffic(*args.summarize_traffic, workload=wl), indent=1))
        return
    if args.summarize_trace:
        print(summarize_trace(args.summarize_trace, args.warmup, args.steps))
        return
    import torch
    if args.study == "wa-terms":
        wa_terms_study(torch, args)
        return
    if args.study == "precision":
        precision_study(torch, args)
        return
    if args.calibrate_copy:
        n = (1 << 30) // 4
        a = torch.rand(n, device="cuda:0")
        b = torch.empty_like(a)
        for _ in range(5):
            torch.mul(a, 1.0, out=b)  # one elementwise kernel: reads 1 GiB, writes 1 GiB
        torch.cuda.synchronize()
        print(json.dumps({"calibrate_copy": {"read_bytes": 4 * n, "write_bytes": 4 * n, "launches": 5}}))
        return

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
    relabel(dd32, args.movable_order)
    arith = {"accurate": dict(wa_arith="centred", density_arith="bin_relative"),
             "reference": dict(wa_arith="reference", density_arith="reference")}
    obj32 = PlacementObjective(dd32, gamma, density_weight=lam, precision=args.precision,
                               algorithm=args.algorithm, schedule=args.schedule,
                               energy_mode=args.energy_mode, **arith[args.arith])
    if args.filler_order == "tile":
        # fillers are interchangeable (one size, no pins): the engine keeps
        # them in the density kernels' tile order and takes them by tile range
        relabel(dd32, None, obj32.filler_origin_offset())
        obj32.update_filler_ranges(dd32.pos)
    # the FP64 form and the CPU baseline: the same arrays in the same order
    d = dd32.host_design(d)
    pos32 = dd32.pos
    ms32, per32 = time_steps(torch, obj32, pos32, args.steps, args.warmup)
    log("[bench] fp32/%s: %.3f ms/step (event median %.3f)" % (args.precision, ms32,
                                                                statistics.median(per32)))

    # The step above runs the wirelength forward on a second stream beside the
    # density chain.  The ops are timed each alone (all on one stream), so each
    # op's duration, and its roofline, is its own and not shared with a
    # concurrent op; the one-stream step time is reported beside the value.
    ms_serial = None
    if obj32.overlap:
        obj32.set_overlap(False)
        ms_serial, _ = time_steps(torch, obj32, pos32, args.steps, 2)
        log("[bench] fp32/%s one stream: %.3f ms/step" % (args.precision, ms_serial))

    # per-op live timing
    hbm_bytes = op_bytes(d, args.schedule, obj32.fillers_final)
    names = [k for k in PROBES if k in hbm_bytes or k == "poisson"]
    op_ms = {k: probe(torch, N, obj32, pos32, PROBES[k], 3) for k in names}
    dom = max(hbm_bytes, key=lambda k: op_ms[k])
    dom_ms = probe(torch, N, obj32, pos32, PROBES[dom], max(args.steps, 10))
    op_ms[dom] = dom_ms
    log("[bench] op ms: " + ", ".join("%s %.3f" % kv for kv in op_ms.items()))
    ops = {}
    for k in names:
        if k == "poisson":
            fl = poisson_flops(d.num_bins_x, d.num_bins_y, 3 if args.energy_mode == "spectral" else 4)
            tf = fl / (op_ms[k] * 1e-3) / 1e12
            ops[k] = {"ms": round(op_ms[k], 4), "bound": "mfma", "flops": fl,
                      "achieved_tflops": round(tf, 1),
                      "frac_of_fp32_peak": round(tf / MFMA_PEAK_TFLOPS["fp32"], 4)}
            if args.precision == "bf16x3":  # three bf16 products per FP32 product
                ops[k]["frac_of_bf16_peak"] = round(3 * tf / MFMA_PEAK_TFLOPS["bf16x3"], 4)
        else:
            gbs = hbm_bytes[k] / (op_ms[k] * 1e-3) / 1e9
            ops[k] = {"ms": round(op_ms[k], 4), "bound": "hbm", "algorithmic_bytes": hbm_bytes[k],
                      "achieved_gbs": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4)}
    dom_bytes = hbm_bytes[dom]
    achieved = dom_bytes / (dom_ms * 1e-3) / 1e9

    result = {}
    if ms_serial is not None:
        result.update(ms_per_step_one_stream=round(ms_serial, 4),
                      ops_timing="each op alone, every kernel on one stream")
    # the same step with the net-weight adjoint: d obj / d w_n = WL_n, the
    # per-net pre-weight wirelength, which the wirelength pass then also stores
    if args.schedule in ("pins", "fused"):
        obj32.set_overlap(True)
        obj32.keep_partial_wl = True
        ms_nw, _ = time_steps(torch, obj32, pos32, max(args.steps // 2, 5), 2)
        obj32.keep_partial_wl = False
        log("[bench] fp32/%s with the net-weight gradient: %.3f ms/step" % (args.precision, ms_nw))
        result.update(ms_per_step_with_net_weight_grad=round(ms_nw, 4))
    if args.filler_order == "tile":
        # what the placement loop adds per step: the fillers' re-sort into
        # tile order every 25 steps (keys, sort, the positions and the four
        # optimizer vectors and the best positions permuted, new filler
        # ranges), amortized over those 25 steps
        ms_rs = resort_ms(torch, dd32, obj32, 3)
        result.update(ms_per_filler_resort=round(ms_rs, 4),
                      ms_per_step_with_resort_every_25=round(ms32 + ms_rs / 25.0, 4))
        log("[bench] filler re-sort: %.3f ms (%.4f ms per step every 25 steps)" % (ms_rs, ms_rs / 25))
        # between re-sorts the cells drift from the tile order: the same step
        # at positions moved by a normal step of 0.5 bin per coordinate (some
        # fillers leave their tile's window and take the drift list)
        gen = torch.Generator(device="cpu").manual_seed(1)
        b = float(d.bin_size_x)
        pos_d = (pos32.double().cpu() + 0.5 * b * torch.randn(pos32.numel(), generator=gen,
                                                               dtype=torch.float64)).to(pos32)
        ms_d, _ = time_steps(torch, obj32, pos_d, max(args.steps // 2, 5), 2)
        del pos_d
        result.update(ms_per_step_drifted_half_bin=round(ms_d, 4))
        log("[bench] fp32/%s at positions drifted by 0.5 bin: %.3f ms/step" % (args.precision, ms_d))
    if not args.skip_fp64:
        dd64 = DeviceDesign(d, dev, dtype=torch.float64)  # already in dd32's order
        obj64 = PlacementObjective(dd64, gamma, density_weight=lam, precision="fp64",
                                   algorithm=args.algorithm, schedule=args.schedule,
                                   energy_mode=args.energy_mode)
        if args.filler_order == "tile":
            obj64.update_filler_ranges(dd64.pos)
        ms64, _ = time_steps(torch, obj64, dd64.pos, max(args.steps // 2, 5), 2)
        log("[bench] fp64: %.3f ms/step" % ms64)
        obj64.evaluate(dd64.pos)
        obj32.evaluate(pos32)
        result.update(fp64_ms_per_step=round(ms64, 4), speedup_vs_fp64=round(ms64 / ms32, 3))
        result.update(accuracy(torch, obj32, obj64, dd64))
        if args.arith == "accurate":
            # the same FP32 form in the reference's arithmetic (bit-exact with
            # its FP32 map): its time and its whole-step errors
            objr = PlacementObjective(dd32, gamma, density_weight=lam, precision=args.precision,
                                      algorithm=args.algorithm, schedule=args.schedule,
                                      energy_mode=args.energy_mode, **arith["reference"])
            if args.filler_order == "tile":
                objr.update_filler_ranges(pos32)
            msr, _ = time_steps(torch, objr, pos32, max(args.steps // 2, 5), 2)
            log("[bench] fp32/%s reference arithmetic: %.3f ms/step" % (args.precision, msr))
            objr.evaluate(pos32)
            result.update(ms_per_step_reference_arith=round(msr, 4))
            result.update(accuracy(torch, objr, obj64, dd64, suffix="_reference_arith"))
            del objr
        log("[bench] accuracy: " + ", ".join("%s %.2e" % (k, v) for k, v in result.items()
                                               if k.startswith("l2err")))
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
            "form": "fp32 + %s spectral solve, %s schedule, %s density, %s energy, %s arithmetic" % (
                args.precision, args.schedule, args.algorithm, args.energy_mode, args.arith),
            "filler_order": args.filler_order,
            "movable_order": args.movable_order,
            "ops": ops,
            **result,
        },
        "roofline": {
            "kernel": "%s op (%s)" % (dom, ", ".join(OP_KERNELS.get(dom, []))),
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": measured_traffic(dom) if args.config == 1 and args.scale == 1.0 else None,
            "kernel_ms": round(dom_ms, 4),
            "algorithmic_bytes": dom_bytes,
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
