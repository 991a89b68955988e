"""Nesterov's accelerated gradient method of e-place, as
NesterovAcceleratedGradientOptimizer (dreamplace/NesterovAcceleratedGradientOptimizer.py:18-240),
with the vector work on the device (dp_nesterov_extrapolate) and the step
sizes from FP64 norms (dp_step_norms).

Same constructor and state layout as the reference; obj_and_grad_fn(x)
returns (obj, grad) and may return the same gradient buffer on every call
(the optimizer copies what it keeps).  Parameters need no .grad.  The
line search of step_nobb compares step sizes on the host, as the
reference's tensor comparison does; step_bb needs no host round trip.
"""
import torch
from torch.optim.optimizer import Optimizer, required

from . import _native as N


class _Norms:
    """FP64 {sum dv^2, sum dg^2, sum dv*dg} of two (position, gradient) pairs."""

    def __init__(self, device):
        self.ws = torch.empty(N.query("dp_optimizer_workspace_bytes"), dtype=torch.uint8,
                              device=device)
        self.out = torch.empty(3, dtype=torch.float64, device=device)

    def __call__(self, v1, v0, g1, g0):
        N.call("dp_step_norms", N.dtype_code(v1), v1.numel(), N.ptr(v1), N.ptr(v0), N.ptr(g1),
               N.ptr(g0), N.ptr(self.ws), N.ptr(self.out), N.stream_ptr(v1.device))
        return self.out


class NesterovAcceleratedGradientOptimizer(Optimizer):
    def __init__(self, params, lr=required, obj_and_grad_fn=required, constraint_fn=None,
                 use_bb=True, bounds=None, grad_into=False, fused_norms=False):
        """bounds (new, optional): the move_boundary box as a dict with
        node_size_x, node_size_y (device), xl, yl, xh, yh, num_movable_nodes,
        num_filler_nodes and filler_size ((w, h) shared by all fillers, or
        None).  The extrapolation then clamps v_kp1 in the same pass
        (dp_nesterov_extrapolate_bounded) and constraint_fn, which must be
        that same clamp, is not called after it.
        grad_into (new, optional): obj_and_grad_fn also takes grad_out=, a
        buffer that receives the gradient.  step_nobb then evaluates into the
        storage of g_k_1 (which the line search no longer reads) and keeps
        it as g_k, with no copy of the gradient.
        fused_norms (new, optional; with bounds and grad_into): step_nobb
        takes sum((v_kp1 - v_k)^2) from the extrapolation pass and asks
        obj_and_grad_fn(x, grad_out=, grad_ref=g_k) for sum((g_kp1 - g_k)^2)
        as a third result (a device FP64 scalar, or None to fall back to a
        pass of its own); the same bits as dp_step_norms."""
        if lr is not required and lr < 0.0:
            raise ValueError("the learning rate must be non-negative, got {}".format(lr))
        defaults = dict(lr=lr, u_k=[], v_k=[], g_k=[], obj_k=[], a_k=[], alpha_k=[], v_k_1=[],
                        g_k_1=[], obj_k_1=[], v_kp1=[None], obj_eval_count=0)
        super(NesterovAcceleratedGradientOptimizer, self).__init__(params, defaults)
        self.obj_and_grad_fn = obj_and_grad_fn
        self.constraint_fn = constraint_fn
        self.use_bb = use_bb
        self.bounds = bounds
        self.grad_into = grad_into
        self.fused_norms = bool(fused_norms) and bounds is not None and grad_into
        if len(self.param_groups) != 1 or len(self.param_groups[0]["params"]) != 1:
            raise ValueError("the optimizer takes exactly one parameter tensor (the [2N] positions)")
        self._norms = None
        self.backtracks = []

    def step(self, closure=None):
        if self.use_bb:
            return self.step_bb(closure)
        return self.step_nobb(closure)

    def _setup(self, p):
        N.require_gpu(p, name="NesterovAcceleratedGradientOptimizer")
        if self._norms is None:
            self._norms = _Norms(p.device)
            self._u_kp1 = torch.empty_like(p.data)
            self._a_kp1 = torch.empty(1, dtype=p.dtype, device=p.device)
            self._dv2 = torch.empty(1, dtype=torch.float64, device=p.device)

    def _constrain(self, x):
        if self.constraint_fn is not None and self.bounds is None:
            self.constraint_fn(x)

    def _extrapolate(self, v_k, g_k, u_k, alpha, a_k, v_kp1, dv2=None):
        b = self.bounds
        if b is None:
            N.call("dp_nesterov_extrapolate", N.dtype_code(v_k), v_k.numel(), N.ptr(v_k),
                   N.ptr(g_k), N.ptr(u_k), N.ptr(alpha), N.ptr(a_k), N.ptr(self._u_kp1),
                   N.ptr(v_kp1), N.ptr(self._a_kp1), N.stream_ptr(v_k.device))
            return
        fs = b.get("filler_size")
        N.call("dp_nesterov_extrapolate_bounded", N.dtype_code(v_k), v_k.numel() // 2, N.ptr(v_k),
               N.ptr(g_k), N.ptr(u_k), N.ptr(alpha), N.ptr(a_k), N.ptr(self._u_kp1),
               N.ptr(v_kp1), N.ptr(self._a_kp1), N.ptr(b["node_size_x"]), N.ptr(b["node_size_y"]),
               float(b["xl"]), float(b["yl"]), float(b["xh"]), float(b["yh"]),
               int(b["num_movable_nodes"]), int(b["num_filler_nodes"]),
               N.host_doubles(fs) if fs is not None else None,
               N.ptr(self._norms.ws) if dv2 is not None else None, N.ptr(dv2),
               N.stream_ptr(v_k.device))

    def step_nobb(self, closure=None):
        """Lines 65-166: Lipschitz step size with backtracking (at most 10
        evaluations), sqrt(sum dv^2 / sum dg^2) in FP64."""
        loss = closure() if closure is not None else None
        group = self.param_groups[0]
        p = group["params"][0]
        self._setup(p)
        f = self.obj_and_grad_fn
        if not group["u_k"]:
            group["u_k"].append(p.data.clone())
            group["v_k"].append(p)
            obj, grad = f(p)
            group["g_k"].append(grad.detach().clone())
            group["obj_k"].append(obj.detach().clone())
        u_k, v_k, g_k = group["u_k"][0], group["v_k"][0], group["g_k"][0]
        if not group["a_k"]:
            group["a_k"].append(torch.ones(1, dtype=g_k.dtype, device=g_k.device))
            v_k_1 = (v_k.data - group["lr"] * g_k).detach()
            group["v_k_1"].append(v_k_1)
            obj, grad = f(v_k_1)
            group["g_k_1"].append(grad.detach().clone())
            group["obj_k_1"].append(obj.detach().clone())
        a_k, v_k_1, g_k_1 = group["a_k"][0], group["v_k_1"][0], group["g_k_1"][0]
        if not group["alpha_k"]:
            s = self._norms(v_k.data, v_k_1, g_k, g_k_1)
            group["alpha_k"].append((s[0].sqrt() / s[1].sqrt()).to(g_k.dtype).reshape(1))
        alpha_k = group["alpha_k"][0]
        if group["v_kp1"][0] is None:
            group["v_kp1"][0] = torch.zeros_like(v_k.data)
        v_kp1 = group["v_kp1"][0]
        cnt = 0
        while True:
            dv2 = self._dv2 if self.fused_norms else None
            self._extrapolate(v_k.data, g_k, u_k, alpha_k, a_k, v_kp1, dv2)
            self._constrain(v_kp1)
            dg2 = None
            if self.fused_norms:
                f_kp1, g_kp1, dg2 = f(v_kp1, grad_out=g_k_1, grad_ref=g_k)
            elif self.grad_into:
                f_kp1, g_kp1 = f(v_kp1, grad_out=g_k_1)
            else:
                f_kp1, g_kp1 = f(v_kp1)
            if dg2 is not None:
                alpha_kp1 = (dv2 / dg2).sqrt().to(g_k.dtype).reshape(1)
            else:
                s = self._norms(v_kp1, v_k.data, g_kp1, g_k)
                alpha_kp1 = (s[0] / s[1]).sqrt().to(g_k.dtype).reshape(1)
            cnt += 1
            group["obj_eval_count"] += 1
            done = bool((alpha_kp1 > 0.95 * alpha_k).item()) or cnt >= 10
            # alpha_k <- alpha_kp1 (a fresh tensor: taken, not copied)
            alpha_k = alpha_kp1
            group["alpha_k"][0] = alpha_k
            if done:
                break
        self.backtracks.append(cnt)
        # obj_k_1 <- obj_k, obj_k <- f_kp1 (the objective's reused buffer),
        # a_k <- a_kp1: the optimizer's own buffers swap instead of copying
        group["obj_k_1"][0], group["obj_k"][0] = group["obj_k"][0], group["obj_k_1"][0]
        group["obj_k"][0].copy_(f_kp1)
        group["a_k"][0], self._a_kp1 = self._a_kp1, a_k
        # v_k_1 <- v_k, v_k <- v_kp1, g_k_1 <- g_k, g_k <- g_kp1, u_k <- u_kp1
        # (lines 151-162) by rotating storage: only the objective's reused
        # gradient buffer is copied, and with grad_into the gradient is
        # already in g_k_1's storage.  The parameter keeps its identity and
        # takes v_kp1's storage.
        self._rotate_positions(group, v_k, v_kp1)
        group["g_k_1"][0], group["g_k"][0] = g_k, g_k_1
        if g_kp1.data_ptr() != g_k_1.data_ptr():
            g_k_1.copy_(g_kp1)
        group["u_k"][0], self._u_kp1 = self._u_kp1, u_k
        return loss

    def permute_state_(self, fn, include_param=True):
        """Apply an in-place relabelling fn(v) of the nodes to the parameter
        (unless the caller has) and to every per-node vector the state
        carries from one step to the next (u_k, g_k, v_k_1, g_k_1).  The
        extrapolation's outputs (v_kp1 and u_kp1) are written whole before
        they are read, so they need none."""
        group = self.param_groups[0]
        if include_param:
            fn(group["params"][0].data)
        for key in ("u_k", "g_k", "v_k_1", "g_k_1"):
            if group[key]:
                fn(group[key][0])

    @staticmethod
    def _rotate_positions(group, v_k, v_kp1):
        old = v_k.data
        v_k.data = v_kp1
        group["v_kp1"][0] = group["v_k_1"][0]
        group["v_k_1"][0] = old

    def step_bb(self, closure=None):
        """Lines 168-240: Barzilai-Borwein short step, else min(Lipschitz,
        previous); the dot products in FP64, the choice on the device."""
        loss = closure() if closure is not None else None
        group = self.param_groups[0]
        p = group["params"][0]
        self._setup(p)
        f = self.obj_and_grad_fn
        if not group["u_k"]:
            group["u_k"].append(p.data.clone())
            group["v_k"].append(p)
        u_k, v_k = group["u_k"][0], group["v_k"][0]
        obj_k, g = f(v_k)
        if not group["g_k"]:
            group["g_k"].append(g.detach().clone())
        g_k = group["g_k"][0]
        g_k.copy_(g)
        if not group["obj_k"]:
            group["obj_k"].append(None)
        group["obj_k"][0] = obj_k.detach().clone()
        if not group["a_k"]:
            group["a_k"].append(torch.ones(1, dtype=g_k.dtype, device=g_k.device))
            group["v_k_1"].append((v_k.data - group["lr"] * g_k).detach())
        a_k, v_k_1 = group["a_k"][0], group["v_k_1"][0]
        obj_k_1, g_k_1 = f(v_k_1)
        if not group["obj_k_1"]:
            group["obj_k_1"].append(None)
        group["obj_k_1"][0] = obj_k_1.detach().clone()
        if group["v_kp1"][0] is None:
            group["v_kp1"][0] = torch.zeros_like(v_k.data)
        v_kp1 = group["v_kp1"][0]
        s = self._norms(v_k.data, v_k_1, g_k, g_k_1)
        lip = (s[0].sqrt() / s[1].sqrt()).to(g_k.dtype).reshape(1)
        if not group["alpha_k"]:
            group["alpha_k"].append(lip.clone())
        alpha_k = group["alpha_k"][0]
        short = (s[2] / s[1]).to(g_k.dtype).reshape(1)
        step = torch.where(short > 0, short, torch.minimum(lip, alpha_k))
        self._extrapolate(v_k.data, g_k, u_k, step, a_k, v_kp1)
        self._constrain(v_kp1)
        group["obj_eval_count"] += 1
        alpha_k.copy_(step)
        a_k.copy_(self._a_kp1)
        self._rotate_positions(group, v_k, v_kp1)
        group["u_k"][0], self._u_kp1 = self._u_kp1, u_k
        return loss
