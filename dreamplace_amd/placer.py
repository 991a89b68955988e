"""Global placement driver: the default-parameter loop of NonLinearPlace
(dreamplace/NonLinearPlace.py:259-271, 377-800) around the device objective.

Per iteration, in the reference's order, with one Lsub and one Llambda step
per Lgamma step:
  divergence check on the last 50 (HPWL, overflow) pairs, with roll-back to
  the best position and a stop after this iteration (NonLinearPlace.py:
  534-570, 633-651);
  move_boundary; (first iteration) initialise the density weight
  (PlaceObj.initialize_density_weight, dreamplace/PlaceObj.py:767-815);
  metrics at the current positions (HPWL over net_mask_all, overflow and
  max density of the movables, FP64 reductions; EvalMetrics.py:116-128);
  one Nesterov step on obj_and_grad_fn = objective + preconditioner
  (PlaceObj.py:69-132, 394-411); best-overflow bookkeeping (:519-526);
  plateau handling: quadratic density penalty, doubled density factor and
  entropy injection (:657-693); from the second iteration the HPWL-mode
  density weight update (PlaceObj.py:830-848); update_gamma (PlaceObj.py:
  893-909); the Lgamma stop tests (NonLinearPlace.py:286-326).
Routability, timing, fence regions and macro halos are not part of this loop.
"""
import math

import numpy as np
import torch

from . import _native as N
from .engine import DeviceDesign, PlacementObjective, can_track_metrics, movable_netlist_order
from .nesterov import NesterovAcceleratedGradientOptimizer

DEFAULTS = dict(gamma=4.0, density_weight=8e-5, stop_overflow=0.1, learning_rate=0.01,
                ref_hpwl=350000.0, lower_pcof=0.95, upper_pcof=1.05, noise_ratio=0.025,
                use_bb=False, precondition_alpha=1.0, quad_penalty=False, density_quad_coeff=2000.0,
                density_factor=1.0, early_stop=True, filler_sort_interval=25, movable_order="netlist",
                fuse_move_boundary=True, fused_metrics=True, fused_norms=True, timing_opt=False,
                net_weighting_scheme="lilith", momentum_decay_factor=0.5, max_net_weight=math.inf,
                timing_start=500, timing_interval=15)


def noise(d, ratio, seed):
    """The noise op (PlaceObj.py:928-943): (r - 0.5) * size * ratio on
    movables and fillers, r uniform from a seeded numpy generator (FP64)."""
    n = d["num_nodes"]
    r = np.random.default_rng(seed).random(2 * n)
    size = np.concatenate([d["node_size_x"], d["node_size_y"]]).astype(np.float64)
    z = (r - 0.5) * size * ratio
    m0, m1 = d["num_movable_nodes"], n - d["num_filler_nodes"]
    z[m0:m1] = 0
    z[n + m0:n + m1] = 0
    return z


def injection_noise(num_nodes, seed, count):
    """Standard normal draws for the `count`-th entropy injection (FP64)."""
    return np.random.default_rng([seed, 7919, count]).standard_normal(2 * num_nodes)


def check_plateau(x, window=10, threshold=0.001):
    """NonLinearPlace.check_plateau (NonLinearPlace.py:527-532)."""
    if len(x) < window:
        return False
    x = x[-window:]
    return (np.max(x) - np.min(x)) / np.mean(x) < threshold


def check_divergence(x, best, stop_overflow, window=50, threshold=0.05):
    """NonLinearPlace.check_divergence (NonLinearPlace.py:534-570) without
    routability: `x` is the list of (hpwl, overflow) pairs, `best` the
    best-overflow metrics dict (or None)."""
    if len(x) < window or best is None:
        return False
    x = np.array(x[-window:])
    overflow_mean = np.mean(x[:, 1])
    overflow_diff = np.maximum(0, np.sign(x[1:, 1] - x[:-1, 1])).astype(np.float32)
    overflow_diff = np.sum(overflow_diff) / overflow_diff.shape[0]
    overflow_range = np.max(x[:, 1]) - np.min(x[:, 1])
    wl_mean = np.mean(x[:, 0])
    wl_ratio = (wl_mean - best["hpwl"]) / best["hpwl"]
    overflow_ratio = (overflow_mean - max(stop_overflow, best["overflow"])) / best["overflow"]
    if wl_ratio > threshold * 1.2:
        if overflow_ratio > threshold:
            return True
        if overflow_range / overflow_mean < threshold:
            return True
        return overflow_diff > 0.6
    return False


def density_weight_factor(cur_hpwl, prev_hpwl, flat_iteration, p):
    """mu of the HPWL-mode density weight update (PlaceObj.py:830-841)."""
    delta = cur_hpwl - prev_hpwl
    if delta < 0:
        return p["upper_pcof"] * max(0.9999 ** float(flat_iteration), 0.98)
    return p["upper_pcof"] * min(max(p["upper_pcof"] ** (-delta / p["ref_hpwl"]), p["lower_pcof"]),
                                 p["upper_pcof"])


class BestPositions:
    """best_pos of NonLinearPlace.py:518-524,640-647 (the positions after the step of
    the iteration with the lowest overflow so far, restored when the loop
    diverges), kept without copying them while they are the current ones.

    The reference copies the positions at every improving iteration, which
    is most of them.  Here an improving iteration only notes that the best
    positions are the parameter's values (`pending`).  After the next step,
    step_nobb's storage rotation has moved those values into its v_k_1,
    which it does not read after its first step; if that step did not
    improve, this class takes that storage and gives the optimizer its old
    buffer instead.  Whoever changes the positions outside the optimizer
    calls `materialize` first.  With the Barzilai-Borwein step (which reads
    v_k_1 in every step) it copies, as the reference does."""

    def __init__(self, optimizer, defer):
        self.opt = optimizer
        self.defer = defer
        self.buf = None
        self.pending = False

    def improved(self, x):
        """The iteration just stepped improved: the best positions are x."""
        if self.defer:
            self.pending = True
        else:
            self._copy(x)

    def not_improved(self):
        """The iteration just stepped did not improve."""
        if self.pending:
            group = self.opt.param_groups[0]
            prev = group["v_k_1"][0]  # the parameter's values before this step
            if self.buf is None:
                self.buf = torch.empty_like(prev)
            self.buf, group["v_k_1"][0] = prev, self.buf
            self.pending = False

    def materialize(self, x):
        """Before x changes outside the optimizer."""
        if self.pending:
            self._copy(x)
            self.pending = False

    def restore(self, x):
        """x <- the best positions."""
        if not self.pending:
            x.copy_(self.buf)

    def permute_(self, fn):
        if self.buf is not None and not self.pending:
            fn(self.buf)

    def value(self, x):
        """The best positions (x itself while they are the current ones)."""
        return x if self.pending else self.buf

    def _copy(self, x):
        if self.buf is None:
            self.buf = x.clone()
        else:
            self.buf.copy_(x)


class GlobalPlacer:
    """`d`: a design dict (dreamplace_amd.synth layout).  Objective options
    (precision, algorithm, schedule) go to PlacementObjective.  `params`
    override DEFAULTS; early_stop=False skips the stop tests and the
    divergence roll-back, so a run takes exactly the iterations asked for
    (to compare runs step for step).

    filler_sort_interval > 0: the (interchangeable) fillers are kept in the
    tile order of the density kernels, sorted at set-up and by `run` every
    that many iterations (DeviceDesign.reorder_fillers).  Positions
    are then in that order; `positions()` gives them in the design's.
    fuse_move_boundary: the optimizer clamps its extrapolated positions to
    the die in the extrapolation pass instead of a move_boundary pass after
    it (the same bits).
    fused_norms: the step sizes' FP64 norms sum((v_kp1 - v_k)^2) and
    sum((g_kp1 - g_k)^2) are formed in the extrapolation and node-gradient
    passes instead of a pass of their own (the same bits).
    fused_metrics: HPWL and overflow at positions the objective has just
    evaluated come from that evaluation (PlacementObjective track_metrics)
    instead of a pin-position pass, an HPWL pass and a second scatter (the
    same bits); where the objective cannot track them, the separate passes
    run.
    timing_opt: timing-driven net weighting (NonLinearPlace.py:463-510):
    after the step of every iteration above timing_start (500) that is a
    multiple of timing_interval (15), `timer(placer)` reports per-net slacks
    and the WNS (LILITH) or the critical nets (ADAMS), and the net weights
    are updated on the device (net_weighting.NetWeighting, params
    net_weighting_scheme, momentum_decay_factor, max_net_weight); the
    objective, the preconditioner's pin weights and the HPWL metric follow
    the new weights, as the reference's in-place update of its net weights
    makes them.  The default timer is net_weighting.WirelengthTimer, whose
    net delays are the net-weight adjoint d obj / d w_n."""

    def __init__(self, d, device, params=None, dtype=None, seed=0, timer=None, **objective_kw):
        p = dict(DEFAULTS)
        p.update(params or {})
        self.p = p
        self.seed = seed
        d = dict(d)
        pos0 = d["pos"].astype(np.float64)
        if p["noise_ratio"] > 0:
            pos0 = pos0 + noise(d, p["noise_ratio"], seed)
        d["pos"] = pos0.astype(d["pos"].dtype)
        self.d = d
        self.dd = DeviceDesign(d, device, dtype=dtype)
        dd = self.dd
        if p["movable_order"] == "netlist":
            # same objective bits; positions() returns the generated order
            dd.reorder_movables(movable_netlist_order(dd))
        self.base_gamma = p["gamma"] * (d["bin_size_x"] + d["bin_size_y"])
        self.gamma = 10 * self.base_gamma
        track = False
        if p["fused_metrics"]:
            # where the objective cannot track the metrics (another schedule
            # or algorithm, too many tiles), the separate passes run instead
            kw = dict(objective_kw)
            track = can_track_metrics(kw.get("schedule"), kw.get("algorithm"), d["num_bins_x"],
                                      d["num_bins_y"], 0,
                                      dd.num_filler_nodes == 0 or dd.uniform_fillers())
        self.objective = PlacementObjective(dd, self.gamma, density_weight=0.0, track_metrics=track,
                                            **objective_kw)
        self.lam = 0.0
        self.quad_penalty = bool(p["quad_penalty"])
        self.density_factor = float(p["density_factor"])
        self.init_density = None
        n, dev, code = dd.num_nodes, dd.device, N.dtype_code(dd.pos)
        self.code = code
        self.areas = (dd.node_size_x * dd.node_size_y).contiguous()
        # fillers of one size have one area (and no pins): the preconditioned
        # gradient pass then reads neither array for them
        nf = dd.num_filler_nodes
        self.filler_area = (float(self.areas[n - nf].item()) if nf > 0 and dd.uniform_fillers()
                            else -1.0)
        self.pws = torch.empty(n, dtype=dd.dtype, device=dev)
        N.call("dp_pin_weight_sum", code, N.ptr(dd.net_weights), N.ptr(dd.flat_node2pin_map),
               N.ptr(dd.flat_node2pin_start_map), N.ptr(dd.pin2net_map), dd.num_physical_nodes, n,
               N.ptr(self.pws), N.stream_ptr(dev))
        self.pin_pos = torch.empty(2 * dd.num_pins, dtype=dd.dtype, device=dev)
        self.hpwl_ws = torch.empty(N.query("dp_wa_workspace_bytes", dd.num_nets), dtype=torch.uint8,
                                   device=dev)
        self.metric_buf = torch.empty(3, dtype=torch.float64, device=dev)
        self.opt_ws = torch.empty(N.query("dp_optimizer_workspace_bytes"), dtype=torch.uint8,
                                  device=dev)
        self.l1_out = torch.empty(2, dtype=torch.float64, device=dev)
        self.norm_out = torch.empty(3, dtype=torch.float64, device=dev)
        self.movable_area = float(np.sum((d["node_size_x"][:d["num_movable_nodes"]] *
                                          d["node_size_y"][:d["num_movable_nodes"]]).astype(np.float64)))
        self.pos = dd.pos
        self.net_weighting = None
        self.timer = None
        if p["timing_opt"]:
            from .net_weighting import NetWeighting, WirelengthTimer
            self.timer = timer if timer is not None else WirelengthTimer()
            if isinstance(self.timer, WirelengthTimer):
                self.objective.keep_partial_wl = True
            self.net_weighting = NetWeighting(
                dd.flat_net2pin_start_map, scheme=p["net_weighting_scheme"],
                momentum_decay_factor=p["momentum_decay_factor"],
                max_net_weight=p["max_net_weight"], ignore_net_degree=d.get("ignore_net_degree", 100),
                dtype=dd.dtype)
        self.history = []
        self.best = None
        self.best_positions = BestPositions(None, defer=False)
        self.rolled_back = False
        self.perturbations = 0
        self.optimizer = None
        self.filler_sorts = 0
        self.net_weight_updates = 0
        interval = int(p["filler_sort_interval"])
        self.sorting = interval > 0 and dd.num_filler_nodes > 0 and dd.uniform_fillers()
        if self.sorting:
            self.sort_fillers()  # at set-up; `run` sorts again every `interval` iterations

    def apply_net_weighting(self):
        """One timing-driven net-weight update: the timer's report into the
        device weights, then the pin weights of the preconditioner and the
        objective's per-slot weights from them."""
        dd = self.dd
        rep = self.timer(self)
        w = dd.net_weights
        self.net_weighting.update(w, net_slack=rep.get("net_slack"), wns=rep.get("wns"),
                                  net_critical=rep.get("net_critical"))
        N.call("dp_pin_weight_sum", self.code, N.ptr(w), N.ptr(dd.flat_node2pin_map),
               N.ptr(dd.flat_node2pin_start_map), N.ptr(dd.pin2net_map), dd.num_physical_nodes,
               dd.num_nodes, N.ptr(self.pws), N.stream_ptr(dd.device))
        self.objective.set_net_weights(w)
        self.net_weight_updates += 1

    def positions(self):
        """The current positions with the fillers in the design's order."""
        return self.dd.generation_order(self.pos.data)

    def sort_fillers(self):
        """Relabel the fillers into the density kernels' tile order: the
        positions (the optimizer's parameter), the optimizer's per-node state
        and the best positions take the same permutation, and the objective
        takes the fillers by tile range from then on (no binning of them)."""
        dd = self.dd
        ob = self.objective
        # the optimizer's per-node state and the best positions (when they
        # are not the current ones) take the permutation in the same passes
        # as dd.pos, which is self.pos
        extra = []
        if self.optimizer is not None:
            self.optimizer.permute_state_(extra.append, include_param=False)
        self.best_positions.permute_(extra.append)
        dd.reorder_fillers(origin_offset=ob.filler_origin_offset(), extra=extra)
        ob.update_filler_ranges(dd.pos)
        self.filler_sorts += 1

    def boundary_box(self):
        """The move_boundary box of `constraint` for the optimizer's fused
        extrapolation; fillers of one size pass that size instead of their
        per-node entries."""
        dd = self.dd
        fs = None
        if dd.num_filler_nodes > 0 and dd.uniform_fillers():
            fs = (float(dd.node_size_x[-1].item()), float(dd.node_size_y[-1].item()))
        return dict(node_size_x=dd.node_size_x, node_size_y=dd.node_size_y, xl=dd.xl, yl=dd.yl,
                    xh=dd.xh, yh=dd.yh, num_movable_nodes=dd.num_movable_nodes,
                    num_filler_nodes=dd.num_filler_nodes, filler_size=fs)

    # -- pieces of the reference loop ----------------------------------------
    def constraint(self, x):
        self.objective.metric_ptr = None  # positions change outside an evaluation
        dd = self.dd
        N.call("dp_move_boundary", self.code, N.ptr(x), dd.num_nodes, N.ptr(dd.node_size_x),
               N.ptr(dd.node_size_y), float(dd.xl), float(dd.yl), float(dd.xh), float(dd.yh),
               dd.num_movable_nodes, dd.num_filler_nodes, N.stream_ptr(x.device))

    def _penalty(self):
        c = None
        if self.quad_penalty:
            # quad_penalty_coeff = coeff / 2 * (1 / init_density), 0 when the
            # initial density is not positive (PlaceObj.py:306-311)
            T = np.float32 if self.dd.dtype == torch.float32 else np.float64
            inv = 1.0 / self.init_density if self.init_density > 0 else self.init_density
            c = float(T(self.p["density_quad_coeff"] / 2) * T(inv))
        self.objective.set_density_penalty(self.density_factor, c)

    def obj_and_grad(self, x, grad_out=None, grad_ref=None):
        """PlaceObj.obj_and_grad_fn: objective, gradient, preconditioner
        (the gradient into grad_out when given).  With grad_ref it returns
        (objective, gradient, sum((gradient - grad_ref)^2) as a device FP64
        scalar formed in the gradient's own pass, or None where it is not)."""
        pre = (self.pws, self.areas, self.p["precondition_alpha"], self.filler_area)
        f, g = self.objective.evaluate(x, precondition=pre, grad_out=grad_out, grad_ref=grad_ref)
        if self.init_density is None:
            # PlaceObj.obj_fn records the density of its first evaluation
            self.init_density = float(self.objective._energy(x).item())
            if self.quad_penalty:
                self._penalty()
                f, g = self.objective.evaluate(x, precondition=pre, grad_out=grad_out,
                                               grad_ref=grad_ref)
        if grad_ref is not None:
            return f, g, self.objective.last_dg2
        return f, g

    def initialize_density_weight(self, x):
        """lambda = density_weight * |grad WL|_1 / |grad D|_1 (no preconditioner)."""
        ob = self.objective
        ob.set_density_penalty(1.0, None)
        ob.set_density_weight(0.0)
        _, g = ob.evaluate(x)  # the wirelength gradient alone
        gw = g.clone()
        ob.set_density_weight(1.0)
        ob.evaluate(x)
        gd = ob.density_grad
        n2 = 2 * self.dd.num_nodes
        for k, v in enumerate((gw, gd)):
            N.call("dp_l1_norm", self.code, n2, N.ptr(v.contiguous()), N.ptr(self.opt_ws),
                   N.ptr(self.l1_out[k:k + 1]), N.stream_ptr(x.device))
        wl_norm, d_norm = (float(v) for v in self.l1_out.tolist())
        T = np.float32 if self.dd.dtype == torch.float32 else np.float64
        self.lam = float(T(self.p["density_weight"] * wl_norm / d_norm))
        ob.set_density_weight(self.lam)
        self._penalty()

    def metrics(self, x):
        dd, dev = self.dd, x.device
        ob = self.objective
        ep = ob.density
        bin_area = ep.bin_size_x * ep.bin_size_y
        T = np.float32 if x.dtype == torch.float32 else np.float64
        if ob.track_metrics and ob.metric_ptr is not None and ob.metric_ptr == x.data_ptr():
            # the objective's last evaluation was at x: its metrics, one read-back
            h, cost, mx = ob.metric_out.tolist()
            return dict(hpwl=h, overflow=float(T(cost)) / self.movable_area,
                        max_density=float(T(mx / bin_area)), density_weight=self.lam,
                        gamma=self.gamma)
        N.call("dp_pin_pos_forward", self.code, N.ptr(x), dd.num_nodes, N.ptr(dd.pin_offset_x),
               N.ptr(dd.pin_offset_y), N.ptr(dd.pin2node_map), N.DP_INDEX_INT32, dd.num_pins,
               N.ptr(self.pin_pos), N.stream_ptr(dev))
        buf = self.metric_buf  # {HPWL, overflow cost, max bin density}, FP64, one read-back
        N.call("dp_hpwl", self.code, N.ptr(self.pin_pos), dd.num_pins, N.ptr(dd.flat_net2pin_map),
               N.ptr(dd.flat_net2pin_start_map), dd.num_nets, N.ptr(dd.net_weights),
               N.ptr(dd.net_mask_all), N.ptr(self.hpwl_ws), N.ptr(buf[0:1]), N.stream_ptr(dev))
        m = ep.density_map(x, num_filler_nodes=0)
        N.call("dp_density_overflow", N.dtype_code(m), N.ptr(m), m.numel(),
               float(ep.target_density) * bin_area, N.ptr(ep.reduce_workspace(dev)), N.ptr(buf[1:3]),
               N.stream_ptr(dev))
        h, cost, mx = buf.tolist()
        # ElectricOverflow returns the cost and the peak in the map's dtype
        return dict(hpwl=h, overflow=float(T(cost)) / self.movable_area,
                    max_density=float(T(mx / bin_area)), density_weight=self.lam, gamma=self.gamma)

    def estimate_learning_rate(self, x):
        """PlaceObj.estimate_initial_learning_rate (PlaceObj.py:443-454)."""
        _, g = self.obj_and_grad(x)
        g0 = g.clone()
        x1 = x - self.p["learning_rate"] * g0
        _, g1 = self.obj_and_grad(x1)
        N.call("dp_step_norms", self.code, x.numel(), N.ptr(x), N.ptr(x1), N.ptr(g0), N.ptr(g1),
               N.ptr(self.opt_ws), N.ptr(self.norm_out), N.stream_ptr(x.device))
        s = self.norm_out.tolist()
        return math.sqrt(s[0]) / math.sqrt(s[1])

    def entropy_injection(self, x, shrink_factor, noise_intensity):
        """entropy_injection, mode "random" (NonLinearPlace.py:572-605):
        shrink every node towards the movables' centroid, add Gaussian noise,
        then put the fixed nodes back."""
        dd = self.dd
        self.objective.metric_ptr = None
        n, m0 = dd.num_nodes, dd.num_movable_nodes
        m1 = n - dd.num_filler_nodes
        xc = x[:m0].mean()
        yc = x[n:n + m0].mean()
        fx, fy = x[m0:m1].clone(), x[n + m0:n + m1].clone()
        if shrink_factor != 1:
            x[:n] = (x[:n] - xc) * shrink_factor + xc
            x[n:] = (x[n:] - yc) * shrink_factor + yc
        if noise_intensity > 0.01:
            z = torch.from_numpy(injection_noise(n, self.seed, self.perturbations)).to(x.device, x.dtype)
            x.add_(noise_intensity * dd.current_order(z))  # each filler keeps its draw
        x[m0:m1] = fx
        x[n + m0:n + m1] = fy

    # -- the loop -------------------------------------------------------------
    def run(self, iterations, log_every=0):
        """At most `iterations` steps; returns (positions in the design's
        node order, as positions() gives them; metrics before each step).
        The loop's own vector, in its relabelled order, stays `self.pos`.
        `log_every` > 0 prints every that many iterations' metrics."""
        p, x = self.p, self.pos
        interval, sorting = int(p["filler_sort_interval"]), self.sorting
        lr = self.estimate_learning_rate(x)
        self.optimizer = NesterovAcceleratedGradientOptimizer(
            [x], lr=lr, obj_and_grad_fn=self.obj_and_grad, constraint_fn=self.constraint,
            use_bb=p["use_bb"], bounds=self.boundary_box() if p["fuse_move_boundary"] else None,
            grad_into=True, fused_norms=p["fused_norms"])
        # a best kept from an earlier run lives in storage of that run's optimizer
        self.best_positions.materialize(x)
        self.best_positions.opt = self.optimizer
        self.best_positions.defer = not p["use_bb"]
        hist = self.history
        T = np.float32 if self.dd.dtype == torch.float32 else np.float64
        overflow_list, divergence_list = [1], []
        last_perturb, min_perturb_interval = -50, 50
        stop = False
        # move_boundary is idempotent: positions the optimizer left (its last
        # v_kp1, already constrained) need no second pass
        constrained = False
        for it in range(iterations):
            if sorting and it > 0 and it % interval == 0:
                self.sort_fillers()
            if (p["early_stop"] and p["stop_overflow"] * 1.1 < overflow_list[-1] < p["stop_overflow"] * 4
                    and check_divergence(divergence_list, self.best, p["stop_overflow"], window=50,
                                         threshold=overflow_list[-1])):
                self.best_positions.restore(x)
                self.objective.metric_ptr = None
                constrained = False
                stop = True
                self.rolled_back = True
            if not constrained:
                self.constraint(x)
            if self.lam == 0.0:
                self.initialize_density_weight(x)
            cur = self.metrics(x)
            hist.append(cur)
            if log_every and it % log_every == 0:
                print("iteration %4d  HPWL %.6E  overflow %.4f  max density %.3f  lambda %.3E  gamma %.3E"
                      % (it, cur["hpwl"], cur["overflow"], cur["max_density"], cur["density_weight"],
                         cur["gamma"]), flush=True)
            self.optimizer.step()
            constrained = True
            if (self.net_weighting is not None and it > p["timing_start"]
                    and it % p["timing_interval"] == 0):
                self.apply_net_weighting()
            if self.best is None or self.best["overflow"] > cur["overflow"]:
                self.best = cur
                self.best_positions.improved(x)
            else:
                self.best_positions.not_improved()
            overflow_list.append(cur["overflow"])
            divergence_list.append([cur["hpwl"], cur["overflow"]])
            if it - last_perturb > min_perturb_interval and check_plateau(overflow_list, 15, 0.001):
                if overflow_list[-1] > 0.9:
                    self.quad_penalty = True
                    self.density_factor *= 2
                    self._penalty()
                    if overflow_list[-1] > 0.95:
                        intensity = min(max(40 + (120 - 40) * (overflow_list[-1] - 0.95) * 10, 40), 90)
                        self.best_positions.materialize(x)
                        self.entropy_injection(x, 0.996, intensity)
                        constrained = False
                    last_perturb = it
                    self.perturbations += 1
            if len(hist) > 1:
                mu = density_weight_factor(cur["hpwl"], hist[-2]["hpwl"], it + 1, p)
                self.lam = float(T(self.lam) * T(mu))
                self.objective.set_density_weight(self.lam)
            self.gamma = self.base_gamma * 10.0 ** ((cur["overflow"] - 0.1) * 20 / 9 - 1)
            self.objective.set_gamma(self.gamma)
            if not p["early_stop"]:
                continue
            if len(hist) > 1 and it > 100 and (
                    (cur["overflow"] < p["stop_overflow"] and cur["hpwl"] > hist[-2]["hpwl"])
                    or cur["max_density"] < self.d["target_density"]):
                break
            if len(hist) > 50 and (cur["overflow"] > hist[-50]["overflow"]
                                   and cur["hpwl"] > self.best["hpwl"] * 2):
                break
            if stop:
                break
        return self.positions(), hist


def main(argv=None):
    """Run global placement on a synthetic design and print one JSON line
    with the final metrics (at the last positions) and the timing: the loop's
    wall time after a short throw-away run has loaded every kernel."""
    import argparse
    import json
    import time

    from . import synth
    ap = argparse.ArgumentParser(description=main.__doc__)
    ap.add_argument("--config", type=int, default=0)
    ap.add_argument("--scale", type=float, default=1.0)
    ap.add_argument("--iterations", type=int, default=1000)
    ap.add_argument("--precision", default="bf16x3", choices=["fp64", "fp32", "bf16x3"])
    ap.add_argument("--arith", default=None, choices=["accurate", "reference"],
                    help="arithmetic of the FP32 forms: accurate (centred wirelength, bin-relative "
                         "density overlaps; the default) or the reference's own; FP64 runs take "
                         "the reference's")
    ap.add_argument("--stop-overflow", type=float, default=None)
    ap.add_argument("--fixed-iterations", action="store_true",
                    help="no stop tests: run exactly --iterations steps")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--history", default=None, help="write per-iteration metrics (JSON) here")
    ap.add_argument("--log-every", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=3,
                    help="iterations of a throw-away run before the timed one: a process's first "
                         "use of each kernel loads it once; 0 times the first run")
    a = ap.parse_args(argv)
    dt = np.float64 if a.precision == "fp64" else np.float32
    arith = a.arith or ("reference" if a.precision == "fp64" else "accurate")
    akw = (dict(wa_arith="centred", density_arith="bin_relative") if arith == "accurate"
           else dict(wa_arith="reference", density_arith="reference"))
    t0 = time.time()
    d = synth.make_design(a.config, seed=a.seed, dtype=dt, scale=a.scale)
    t_gen = time.time() - t0
    dev = torch.device("cuda", torch.cuda.current_device())
    prm = {} if a.stop_overflow is None else dict(stop_overflow=a.stop_overflow)
    if a.fixed_iterations:
        prm["early_stop"] = False
    t_warm = 0.0
    if a.warmup > 0:
        t0 = time.time()
        gp = GlobalPlacer(d, dev, params=dict(prm, early_stop=False), precision=a.precision,
                          algorithm="tiled", schedule="pins", energy_mode="spectral", seed=a.seed,
                          **akw)
        gp.run(a.warmup)
        torch.cuda.synchronize()
        del gp
        torch.cuda.empty_cache()
        t_warm = time.time() - t0
    gp = GlobalPlacer(d, dev, params=prm, precision=a.precision, algorithm="tiled", schedule="pins",
                      energy_mode="spectral", seed=a.seed, **akw)
    torch.cuda.synchronize()
    t1 = time.time()
    _, hist = gp.run(a.iterations, log_every=a.log_every)
    torch.cuda.synchronize()
    t_run = time.time() - t1
    x = gp.pos
    gp.constraint(x)
    fin = gp.metrics(x)
    out = dict(config=a.config, scale=a.scale, precision=a.precision, arith=arith,
               dtype=str(np.dtype(dt)),
               num_nodes=int(d["num_nodes"]), num_filler_nodes=int(d["num_filler_nodes"]),
               num_pins=int(d["num_pins"]), bins=int(round((d["xh"] - d["xl"]) / d["bin_size_x"])),
               iterations=len(hist), evaluations=int(gp.optimizer.param_groups[0]["obj_eval_count"]),
               final_hpwl=fin["hpwl"], final_overflow=fin["overflow"],
               final_max_density=fin["max_density"], best_overflow=gp.best["overflow"],
               rolled_back=gp.rolled_back, perturbations=gp.perturbations,
               seconds=t_run, ms_per_iteration=1e3 * t_run / max(len(hist), 1),
               generate_seconds=t_gen, warmup_iterations=a.warmup, warmup_seconds=t_warm)
    if a.history:
        with open(a.history, "w") as f:
            json.dump(hist, f)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
