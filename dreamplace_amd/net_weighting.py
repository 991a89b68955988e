"""Timing-driven net weighting on the device.

``NetWeighting`` is the weighting step of the reference's timing op
(``TimingOpt.update_net_weights``, dreamplace/ops/timing/timing.py:182-206,
and net_weighting_scheme.h:106-190): the ADAMS and LILITH schemes run as one
HIP kernel (dp_net_weighting) over the nets, updating the per-net
criticality and the net weights in place.  What the reference takes from a
static timing analyser (OpenTimer or HeteroSTA: per-net worst sink slack and
the WNS, or the nets on the most critical paths) comes from a *timer*
callable here, so any timer can drive it.

``WirelengthTimer`` is a timer for runs without a timing library: a net's
delay is taken as its wirelength WL_n, which is exactly the net-weight
adjoint d obj / d w_n of the weighted wirelength objective (the per-net
pre-weight WA wirelength that PlacementObjective stores with
keep_partial_wl).  Its slack is required - WL_n with the required time at a
quantile of the delays, so the longest nets are the critical ones.

The placer calls the timer and the weighting at the reference's schedule
(NonLinearPlace.py:463-510: iteration > 500, every 15 iterations).
"""
import math

import torch

from . import _native as N

SCHEMES = {"adams": N.DP_NET_WEIGHTING_ADAMS, "lilith": N.DP_NET_WEIGHTING_LILITH}


class NetWeighting:
    """Per-net criticality state and the weighting update.  `netpin_start`:
    the net CSR start array (device int32, num_nets + 1); degrees come from
    it (the reference's degree_map).  max_net_weight: cap of LILITH's
    weights (params.json "inf" by default); momentum_decay_factor: LILITH's
    decay (params.json 0.5)."""

    def __init__(self, netpin_start, scheme="lilith", momentum_decay_factor=0.5,
                 max_net_weight=math.inf, ignore_net_degree=100, dtype=torch.float32):
        if scheme not in SCHEMES:
            raise ValueError("net weighting scheme must be 'adams' or 'lilith', got %r" % (scheme,))
        N.require_gpu(netpin_start, name="NetWeighting")
        self.scheme = scheme
        self.netpin_start = N.as_i32(netpin_start).contiguous()
        self.num_nets = self.netpin_start.numel() - 1
        self.decay = float(momentum_decay_factor)
        self.max_net_weight = float(max_net_weight)
        self.ignore_net_degree = int(ignore_net_degree)
        self.net_criticality = torch.zeros(self.num_nets, dtype=dtype,
                                           device=self.netpin_start.device)

    def update(self, net_weights, net_slack=None, wns=None, net_critical=None):
        """Update `net_weights` (device, per net, the criticality's dtype) in
        place from a timer's report: LILITH takes net_slack (per net) and wns,
        ADAMS net_critical (per net, nonzero on critical paths)."""
        N.require_gpu(net_weights, name="NetWeighting")
        if net_weights.dtype != self.net_criticality.dtype or not net_weights.is_contiguous():
            raise RuntimeError("NetWeighting: net_weights must be a contiguous %s tensor"
                               % self.net_criticality.dtype)
        if self.scheme == "lilith":
            if net_slack is None or wns is None:
                raise ValueError("the lilith scheme needs net_slack and wns")
            net_slack = net_slack.to(net_weights.dtype).contiguous()
            crit = None
        else:
            if net_critical is None:
                raise ValueError("the adams scheme needs net_critical")
            crit = N.as_u8(net_critical).contiguous()
        N.call("dp_net_weighting", N.dtype_code(net_weights), SCHEMES[self.scheme], self.num_nets,
               N.ptr(net_slack), N.ptr(crit), float(wns if wns is not None else 0.0),
               N.ptr(self.netpin_start), self.ignore_net_degree, self.decay, self.max_net_weight,
               N.ptr(self.net_criticality), N.ptr(net_weights), N.stream_ptr(net_weights.device))
        return net_weights


class WirelengthTimer:
    """A timer stand-in: per-net delay = the net's wirelength WL_n, i.e. the
    net-weight adjoint d obj / d w_n of the objective's last evaluation
    (PlacementObjective.partial_wl, kept with keep_partial_wl=True); slack =
    required - WL_n with required the `quantile` of the delays of the nets
    the wirelength counts.  Returns what NetWeighting.update takes: net_slack
    and wns for LILITH, and net_critical (slack < 0) for ADAMS."""

    def __init__(self, quantile=0.97):
        self.quantile = float(quantile)

    def __call__(self, placer):
        ob = placer.objective
        if not ob.keep_partial_wl:
            raise RuntimeError("WirelengthTimer needs the objective's per-net wirelength "
                               "(PlacementObjective(keep_partial_wl=True))")
        wl = ob.partial_wl
        live = ob.net_mask.bool()
        vals = wl[live].float()
        if vals.numel() == 0:
            required = 0.0
        else:
            k = max(1, min(vals.numel(), int(math.ceil(self.quantile * vals.numel()))))
            required = float(torch.kthvalue(vals, k).values.item())
        slack = torch.where(live, required - wl, torch.full_like(wl, math.inf))
        wns = min(float(slack.min().item()), 0.0) if slack.numel() else 0.0
        return dict(net_slack=slack, wns=wns, net_critical=(slack < 0).to(torch.uint8))
