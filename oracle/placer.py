"""CPU restatement of the global-placement driver around the objective
(test infrastructure; only tests/, smoke() and the bench's CPU leg use the
oracle):

* NesterovAcceleratedGradientOptimizer.step_nobb / step_bb
  (dreamplace/NesterovAcceleratedGradientOptimizer.py:65-166, 168-240);
* PreconditionOp without fence regions (dreamplace/PlaceObj.py:69-132);
* PlaceObj.initialize_density_weight (dreamplace/PlaceObj.py:767-815, no
  fence regions), the HPWL-mode density-weight update (:830-848),
  update_gamma (:893-909), estimate_initial_learning_rate (:443-454) and the
  quadratic density penalty with its density factor (:293-320);
* the noise op (dreamplace/PlaceObj.py:917-945);
* the default-parameter global placement loop of NonLinearPlace
  (dreamplace/NonLinearPlace.py:259-271, 377-800): one Lsub and one Llambda
  step per Lgamma step, the divergence test with roll-back to the best
  position, metrics before each step, best-overflow bookkeeping, plateau
  handling (quadratic penalty, doubled density factor, entropy injection),
  density weight and gamma updates after the step, the Lgamma stop tests.
  Routability, timing, fence regions and macro halos are not part of it.

Vectors are numpy arrays of the positions' dtype and follow the reference's
torch operation order; the step-size norms are accumulated in FP64 (the
reference's FP64 form does the same; its FP32 form accumulates in FP32).
Entropy injection takes its normal draws from a seeded numpy generator
where the reference calls torch.randn.
"""
import math

import numpy as np

from . import oracle as O


# --- Nesterov ----------------------------------------------------------------
def _norm2(a):
    a = np.asarray(a, dtype=np.float64)
    return math.sqrt(float(np.dot(a, a)))


class Nesterov:
    """State and steps of NesterovAcceleratedGradientOptimizer for one
    parameter vector.  obj_and_grad(x) -> (obj, grad); constraint(x) edits x
    in place."""

    def __init__(self, x, lr, obj_and_grad, constraint, use_bb=False):
        self.v_k = x  # the parameter itself, updated in place
        self.lr = lr
        self.f = obj_and_grad
        self.constraint = constraint
        self.use_bb = use_bb
        self.u_k = None
        self.obj_eval_count = 0
        self.backtracks = []

    def step(self):
        return self.step_bb() if self.use_bb else self.step_nobb()

    def step_nobb(self):
        T = self.v_k.dtype.type
        if self.u_k is None:
            self.u_k = self.v_k.copy()
            obj, g = self.f(self.v_k)
            self.g_k, self.obj_k = g.copy(), obj
            self.a_k = T(1)
            self.v_k_1 = (self.v_k - T(self.lr) * self.g_k).astype(self.v_k.dtype)
            obj, g = self.f(self.v_k_1)
            self.g_k_1, self.obj_k_1 = g.copy(), obj
            self.alpha_k = T(_norm2(self.v_k - self.v_k_1) / _norm2(self.g_k - self.g_k_1))
        a_k = self.a_k
        a_kp1 = (T(1) + np.sqrt(T(4) * (a_k * a_k) + T(1))) / T(2)
        coef = (a_k - T(1)) / a_kp1
        cnt = 0
        while True:
            u_kp1 = self.v_k - self.alpha_k * self.g_k
            v_kp1 = u_kp1 + coef * (u_kp1 - self.u_k)
            self.constraint(v_kp1)
            f_kp1, g_kp1 = self.f(v_kp1)
            g_kp1 = g_kp1.copy()
            dv = (v_kp1 - self.v_k).astype(np.float64)
            dg = (g_kp1 - self.g_k).astype(np.float64)
            alpha_kp1 = T(math.sqrt(float(np.dot(dv, dv)) / float(np.dot(dg, dg))))
            cnt += 1
            self.obj_eval_count += 1
            if alpha_kp1 > T(0.95) * self.alpha_k or cnt >= 10:
                self.alpha_k = alpha_kp1
                break
            self.alpha_k = alpha_kp1
        self.backtracks.append(cnt)
        self.v_k_1 = self.v_k.copy()
        self.g_k_1 = self.g_k
        self.obj_k_1 = self.obj_k
        self.u_k = u_kp1
        self.v_k[:] = v_kp1
        self.g_k = g_kp1
        self.obj_k = f_kp1
        self.a_k = a_kp1

    def step_bb(self):
        T = self.v_k.dtype.type
        if self.u_k is None:
            self.u_k = self.v_k.copy()
        obj_k, g_k = self.f(self.v_k)
        g_k = g_k.copy()
        if not hasattr(self, "a_k"):
            self.a_k = T(1)
            self.v_k_1 = (self.v_k - T(self.lr) * g_k).astype(self.v_k.dtype)
        obj_k_1, g_k_1 = self.f(self.v_k_1)
        g_k_1 = g_k_1.copy()
        s = (self.v_k - self.v_k_1).astype(np.float64)
        y = (g_k - g_k_1).astype(np.float64)
        ss, yy, sy = float(np.dot(s, s)), float(np.dot(y, y)), float(np.dot(s, y))
        if not hasattr(self, "alpha_k"):
            self.alpha_k = T(math.sqrt(ss) / math.sqrt(yy))
        a_k = self.a_k
        a_kp1 = (T(1) + np.sqrt(T(4) * (a_k * a_k) + T(1))) / T(2)
        coef = (a_k - T(1)) / a_kp1
        bb_short = T(sy / yy)
        lip = T(math.sqrt(ss) / math.sqrt(yy))
        step = bb_short if bb_short > 0 else min(lip, self.alpha_k)
        u_kp1 = self.v_k - step * g_k
        v_kp1 = u_kp1 + coef * (u_kp1 - self.u_k)
        self.constraint(v_kp1)
        self.obj_eval_count += 1
        self.v_k_1 = self.v_k.copy()
        self.alpha_k = step
        self.u_k = u_kp1
        self.v_k[:] = v_kp1
        self.a_k = a_kp1


# --- placement driver ---------------------------------------------------------
DEFAULTS = dict(gamma=4.0, density_weight=8e-5, stop_overflow=0.1, learning_rate=0.01,
                ref_hpwl=350000.0, lower_pcof=0.95, upper_pcof=1.05, noise_ratio=0.025,
                use_bb=False, precondition_alpha=1.0, quad_penalty=False, density_quad_coeff=2000.0,
                density_factor=1.0)


def noise(d, ratio, seed):
    """Uniform noise of the noise op (PlaceObj.py:928-943): (r - 0.5) * size *
    ratio on movables and fillers, from a seeded numpy generator."""
    n = d["num_nodes"]
    r = np.random.default_rng(seed).random(2 * n)
    size = np.concatenate([d["node_size_x"], d["node_size_y"]]).astype(np.float64)
    z = (r - 0.5) * size * ratio
    m0, m1 = d["num_movable_nodes"], n - d["num_filler_nodes"]
    z[m0:m1] = 0
    z[n + m0:n + m1] = 0
    return z


def plateau(values, window, threshold):
    """check_plateau (NonLinearPlace.py:527-532): relative range of the last
    `window` values below `threshold`."""
    if len(values) < window:
        return False
    tail = values[len(values) - window:]
    lo, hi = min(tail), max(tail)
    return (hi - lo) / (sum(tail) / window) < threshold


def diverging(pairs, best, stop_overflow, window, threshold):
    """check_divergence (NonLinearPlace.py:534-570), no routability: from the
    last `window` (hpwl, overflow) pairs and the best-overflow record."""
    if best is None or len(pairs) < window:
        return False
    tail = pairs[len(pairs) - window:]
    wl = [a for a, _ in tail]
    ov = [b for _, b in tail]
    ov_mean = sum(ov) / window
    rises = sum(1 for i in range(1, window) if ov[i] > ov[i - 1]) / (window - 1)
    if (sum(wl) / window - best["hpwl"]) / best["hpwl"] <= threshold * 1.2:
        return False
    if (ov_mean - max(stop_overflow, best["overflow"])) / best["overflow"] > threshold:
        return True
    if (max(ov) - min(ov)) / ov_mean < threshold:
        return True
    return rises > 0.6


def injection_noise(num_nodes, seed, count):
    """Standard normal draws of the `count`-th entropy injection."""
    return np.random.default_rng([seed, 7919, count]).standard_normal(2 * num_nodes)


def entropy_injection(d, pos, shrink, intensity, z):
    """entropy_injection, mode "random" (NonLinearPlace.py:572-605), with the
    normal draws `z` given: shrink all nodes to the movables' centroid, add
    intensity * z, restore the fixed nodes."""
    n, m0 = d["num_nodes"], d["num_movable_nodes"]
    m1 = n - d["num_filler_nodes"]
    T = pos.dtype.type
    keep = pos.copy()
    xc, yc = T(np.mean(pos[:m0])), T(np.mean(pos[n:n + m0]))
    if shrink != 1:
        pos[:n] = (pos[:n] - xc) * T(shrink) + xc
        pos[n:] = (pos[n:] - yc) * T(shrink) + yc
    if intensity > 0.01:
        pos += (intensity * z).astype(pos.dtype)
    pos[m0:m1] = keep[m0:m1]
    pos[n + m0:n + m1] = keep[n + m0:n + m1]


def place(d, iterations, params=None, seed=0, objective=None, log_every=0):
    """Run the default-parameter global placement loop on design `d` (a
    synth.make_design dict; positions in its dtype).  Returns (pos, metrics,
    info): metrics a list of dicts (hpwl, overflow, max_density,
    density_weight, gamma) taken before each step, as EvalMetrics records
    them; info has the best-overflow record, whether the divergence test
    rolled back, and the number of plateau perturbations.
    `objective(d, density_weight, gamma, density_factor=, quad_coeff=)`
    returns (obj, grad, parts with "energy", "wl_grad", "density_grad"); the
    default is the oracle's own composition (oracle.objective_and_gradient),
    ref_iteration.reference_objective drives the loop with the reference's
    CPU ops instead."""
    objective = objective or O.objective_and_gradient
    p = dict(DEFAULTS)
    p.update(params or {})
    d = dict(d)
    pos = d["pos"].copy()
    dtype = pos.dtype
    T = dtype.type
    n = d["num_nodes"]
    bx, by = d["bin_size_x"], d["bin_size_y"]
    base_gamma = p["gamma"] * (bx + by)
    st = dict(gamma=10 * base_gamma, lam=0.0, factor=float(p["density_factor"]),
              quad=bool(p["quad_penalty"]), init=None)
    areas = (d["node_size_x"] * d["node_size_y"]).astype(dtype)
    pws = O.pin_weight_sum(d["net_weights"], d["flat_node2pin_map"], d["flat_node2pin_start_map"],
                           d["pin2net_map"], n).astype(dtype)
    dp = O.DensityParams(d["node_size_x"], d["node_size_y"], d["xl"], d["yl"], d["xh"], d["yh"], bx, by,
                         d["num_movable_nodes"], d["num_terminals"], d["num_filler_nodes"],
                         d["target_density"], d["movable_macro_mask"])
    init_map = O.fixed_density_map(dp, pos, d["bin_center_x"], d["bin_center_y"])
    movable_area = float(np.sum(areas[:d["num_movable_nodes"]].astype(np.float64)))

    def constraint(x):
        x[:] = O.move_boundary(x, d["node_size_x"], d["node_size_y"], d["xl"], d["yl"], d["xh"], d["yh"],
                               d["num_movable_nodes"], d["num_filler_nodes"])

    def quad_coeff():
        if not st["quad"]:
            return None
        e = st["init"]
        return float(T(p["density_quad_coeff"] / 2) * T(1.0 / e if e > 0 else e))

    def obj_and_grad(x):
        dd = dict(d)
        dd["pos"] = x
        if st["init"] is None:
            # PlaceObj.obj_fn keeps the density of its first evaluation
            st["init"] = float(T(objective(dd, st["lam"], st["gamma"])[2]["energy"]))
        obj, g, _ = objective(dd, st["lam"], st["gamma"], density_factor=st["factor"],
                                             quad_coeff=quad_coeff())
        g = g.astype(dtype)
        pre = np.maximum(pws + (T(p["precondition_alpha"]) * T(st["lam"])) * areas, T(1))
        g[:n] = g[:n] / pre
        g[n:] = g[n:] / pre
        return obj, g

    def metrics(x):
        pin = O.pin_pos_forward(x, d["pin_offset_x"], d["pin_offset_y"], d["pin2node_map"])
        h = O.hpwl(pin, d["flat_net2pin_map"], d["flat_net2pin_start_map"], d["net_weights"],
                   d["net_mask_all"])
        m = O.density_map(dp, x, init_map, num_filler_nodes=0)
        of, md = O.overflow(m, d["target_density"], bx, by, movable_area)
        return dict(hpwl=float(h), overflow=of, max_density=md, density_weight=st["lam"],
                    gamma=st["gamma"])

    if p["noise_ratio"] > 0:
        pos = (pos.astype(np.float64) + noise(d, p["noise_ratio"], seed)).astype(dtype)
    # estimate_initial_learning_rate (PlaceObj.py:443-454), at lambda 0
    _, g0 = obj_and_grad(pos)
    x1 = (pos - T(p["learning_rate"]) * g0).astype(dtype)
    _, g1 = obj_and_grad(x1)
    lr = _norm2(pos - x1) / _norm2(g0 - g1)
    opt = Nesterov(pos, lr, obj_and_grad, constraint, use_bb=p["use_bb"])
    hist = []
    info = dict(best=None, best_pos=None, rolled_back=False, perturbations=0, optimizer=opt)
    ovf, pairs = [1], []
    last_perturb = -50
    stop = False
    so = p["stop_overflow"]
    for it in range(iterations):
        if so * 1.1 < ovf[-1] < so * 4 and diverging(pairs, info["best"], so, 50, ovf[-1]):
            pos[:] = info["best_pos"]
            stop = True
            info["rolled_back"] = True
        constraint(pos)
        if st["lam"] == 0.0:
            dd = dict(d)
            dd["pos"] = pos
            _, _, parts = objective(dd, 1.0, st["gamma"])
            wl_norm = float(np.sum(np.abs(parts["wl_grad"].astype(np.float64))))
            d_norm = float(np.sum(np.abs(parts["density_grad"].astype(np.float64))))
            st["lam"] = float(T(p["density_weight"] * wl_norm / d_norm))
        cur = metrics(pos)
        hist.append(cur)
        if log_every and it % log_every == 0:
            print("iteration %4d  HPWL %.6E  overflow %.4f  lambda %.3E  gamma %.3E"
                  % (it, cur["hpwl"], cur["overflow"], cur["density_weight"], cur["gamma"]), flush=True)
        opt.step()
        if info["best"] is None or info["best"]["overflow"] > cur["overflow"]:
            info["best"] = cur
            info["best_pos"] = pos.copy()
        ovf.append(cur["overflow"])
        pairs.append((cur["hpwl"], cur["overflow"]))
        if it - last_perturb > 50 and plateau(ovf, 15, 0.001) and ovf[-1] > 0.9:
            st["quad"] = True
            st["factor"] *= 2
            if ovf[-1] > 0.95:
                intensity = min(max(40 + 80 * (ovf[-1] - 0.95) * 10, 40), 90)
                entropy_injection(d, pos, 0.996, intensity, injection_noise(n, seed, info["perturbations"]))
            last_perturb = it
            info["perturbations"] += 1
        if len(hist) > 1:
            delta = cur["hpwl"] - hist[-2]["hpwl"]
            if delta < 0:
                mu = p["upper_pcof"] * max(0.9999 ** float(it + 1), 0.98)
            else:
                mu = p["upper_pcof"] * min(max(p["upper_pcof"] ** (-delta / p["ref_hpwl"]),
                                               p["lower_pcof"]), p["upper_pcof"])
            st["lam"] = float(T(st["lam"]) * T(mu))
        st["gamma"] = base_gamma * 10.0 ** ((cur["overflow"] - 0.1) * 20 / 9 - 1)
        if len(hist) > 1 and it > 100 and (
                (cur["overflow"] < so and cur["hpwl"] > hist[-2]["hpwl"])
                or cur["max_density"] < d["target_density"]):
            break
        if len(hist) > 50 and cur["overflow"] > hist[-50]["overflow"] and \
                cur["hpwl"] > info["best"]["hpwl"] * 2:
            break
        if stop:
            break
    return pos, hist, info
