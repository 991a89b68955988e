"""Central finite differences of the merged WA wirelength in FP64, for the
differentiation study (random pin coordinates and random net weights, step
h = 1e-4 x bin width).

Nets are independent: the per-net wirelength WL_n depends on the pins of net
n only.  Perturbing one coordinate in each of K distinct nets and reading the
per-net partials gives K central differences from two forward passes,
d(sum_m w_m WL_m)/dx_p = w_n (WL_n(x + h) - WL_n(x - h)) / 2h.  The net
weights enter the total linearly, so each weight gets its own pair of passes
over the weighted total.  All passes run dp_wa_merged_forward with no
per-pin output.
"""
import ctypes

import numpy as np
import torch

from . import _native as N


class WaForward64:
    """FP64 forward passes of the merged WA model on one netlist."""

    def __init__(self, flat_netpin, netpin_start, net_mask, gamma):
        self.fnp = N.as_i32(flat_netpin).contiguous()
        self.nps = N.as_i32(netpin_start).contiguous()
        self.mask = N.as_u8(net_mask).contiguous()
        self.num_nets = self.nps.numel() - 1
        dev = self.fnp.device
        self.inv_gamma = torch.tensor(1.0 / float(gamma), dtype=torch.float64, device=dev)
        self.ws = N.workspace.get("wa_fd", N.query("dp_wa_workspace_bytes", self.num_nets), dev)

    def __call__(self, pin_pos, w, partial=None, wl_out=None, wl_index=0):
        """Writes sum_n w_n WL_n to wl_out[wl_index] (a fresh 0-dim tensor by
        default, returned) and the unweighted per-net WL to `partial`."""
        if wl_out is None:
            wl_out = torch.empty((), dtype=torch.float64, device=pin_pos.device)
        out_ptr = ctypes.c_void_p(wl_out.data_ptr() + 8 * int(wl_index))
        N.call("dp_wa_merged_forward", N.DP_FLOAT64, N.ptr(pin_pos), pin_pos.numel() // 2,
               N.ptr(self.fnp), N.ptr(self.nps), self.num_nets, N.ptr(w), N.ptr(self.mask),
               N.ptr(self.inv_gamma), N.ptr(partial), None, out_ptr, N.ptr(self.ws),
               N.stream_ptr(pin_pos.device))
        return wl_out


def pick_pin_coordinates(rng, netpin_start, k):
    """k coordinates on k distinct nets: (net, pin, axis) arrays, the pin
    uniform among the net's pins (slots of flat_netpin) and the axis uniform."""
    nps = np.asarray(netpin_start)
    num_nets = len(nps) - 1
    nets = rng.choice(num_nets, size=min(k, num_nets), replace=False)
    deg = nps[nets + 1] - nps[nets]
    slots = nps[nets] + (rng.random(len(nets)) * deg).astype(np.int64)
    axes = rng.integers(0, 2, size=len(nets))
    return nets, slots, axes


def pin_fd(fwd, pin_pos64, w64, nets, pins, axes, h):
    """FD of the weighted total with respect to pin coordinate axes[i] of pin
    pins[i] (which belongs to net nets[i]; all nets distinct)."""
    P = pin_pos64.numel() // 2
    dev = pin_pos64.device
    idx = torch.from_numpy(np.asarray(axes, dtype=np.int64) * P + np.asarray(pins, dtype=np.int64)).to(dev)
    nets_t = torch.from_numpy(np.asarray(nets, dtype=np.int64)).to(dev)
    parts = []
    for sign in (1.0, -1.0):
        x = pin_pos64.clone()
        x[idx] += sign * h
        partial = torch.empty(fwd.num_nets, dtype=torch.float64, device=dev)
        fwd(x, w64, partial=partial)
        parts.append(partial[nets_t])
    wn = w64[nets_t] if w64 is not None else torch.ones_like(parts[0])
    return (wn * (parts[0] - parts[1]) / (2.0 * h)).cpu().numpy()


def weight_fd(fwd, pin_pos64, w64, nets, h):
    """FD of the weighted total with respect to net weight nets[i]."""
    dev = pin_pos64.device
    out = torch.empty(2 * len(nets), dtype=torch.float64, device=dev)
    wv = w64.clone()
    for i, n in enumerate(nets):
        n = int(n)
        wv[n] = w64[n] + h
        fwd(pin_pos64, wv, wl_out=out, wl_index=2 * i)
        wv[n] = w64[n] - h
        fwd(pin_pos64, wv, wl_out=out, wl_index=2 * i + 1)
        wv[n] = w64[n]
    o = out.view(-1, 2)
    return ((o[:, 0] - o[:, 1]) / (2.0 * h)).cpu().numpy()
