"""The oracle: a CPU restatement of DREAMPlace's global-placement inner loop.

TEST INFRASTRUCTURE ONLY.  Only tests/, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of bench.py may import this module, and only as the
checker.  The product package ``dreamplace_amd`` never imports it; its HIP
path fails loudly when the native library is missing instead of falling back
here.

Two layers:

* ``dp_oracle.c`` (built into ``oracle/libdp_oracle.so`` by :func:`build`),
  bound with ctypes: the integer / fixed-point scatter, the force gather, the
  merged WA wirelength, pin_pos, HPWL, move_boundary and the pin-weight sum,
  each following the reference loop it cites.
* numpy FP64 closed forms for the spectral transforms and the Poisson solve,
  and the host-side array preparation of ``ElectricOverflow.reset``.

How each function is pinned (golden vectors, oracle/_ref, or mathematics) is
stated at the function; a function pinned by none says "parity unpinned".
"""
import ctypes
import math
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(HERE, "libdp_oracle.so")
_lib = None


def build(force=False):
    """Compile dp_oracle.c with gcc (-O2 -fopenmp -ffp-contract=off)."""
    src = os.path.join(HERE, "dp_oracle.c")
    if not force and os.path.exists(_LIB_PATH) and os.path.getmtime(_LIB_PATH) >= os.path.getmtime(src):
        return _LIB_PATH
    cmd = ["gcc", "-O2", "-fopenmp", "-ffp-contract=off", "-shared", "-fPIC",
           "-Wno-unknown-pragmas", src, "-o", _LIB_PATH + ".tmp", "-lm"]
    subprocess.check_call(cmd)
    os.replace(_LIB_PATH + ".tmp", _LIB_PATH)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        _lib.dpo_fixed_scale.restype = ctypes.c_int64
        _lib.dpo_fixed_scale.argtypes = [ctypes.c_double] * 4
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _sfx(dtype):
    dtype = np.dtype(dtype)
    if dtype == np.float32:
        return "f32", ctypes.c_float
    if dtype == np.float64:
        return "f64", ctypes.c_double
    raise TypeError("oracle supports float32/float64, got %s" % dtype)


def _c(a, dtype=None):
    return np.ascontiguousarray(a if dtype is None else np.asarray(a, dtype=dtype))


# ---------------------------------------------------------------------------
# Fixed-point density maps
# ---------------------------------------------------------------------------

def fixed_scale(xl, yl, xh, yh):
    """2^(64 - max(ceil(log2(area))+1, 32)); electric_density_map.cpp:117-123.
    Pinned by: closed form (2^32 for every die up to 2^31 area)."""
    return int(lib().dpo_fixed_scale(float(xl), float(yl), float(xh), float(yh)))


def _arith(arithmetic):
    if arithmetic not in ("reference", "bin_relative"):
        raise ValueError("arithmetic must be 'reference' or 'bin_relative'")
    return "rel_" if arithmetic == "bin_relative" else ""


def triangle_density_fixed(x, y, wxc, wyc, offx, offy, ratio, num_bins_x, num_bins_y,
                           xl, yl, bin_size_x, bin_size_y, scale, buf=None,
                           arithmetic="reference"):
    """int64 map of sum_(cell,bin) trunc(((px*ratio)*py)*scale) over the given cells.
    electric_density_map.cpp:254-309 + atomic_ops.h:29-45.
    Pinned by: oracle/_ref density_map (bit-exact), brute-force overlap, mass
    conservation (tests/test_oracle.py).  arithmetic="bin_relative": the same
    bins and recipe with bin-relative overlaps (dp_oracle.c DPO_DEFINE_REL; no
    reference counterpart), pinned by FP64 agreement with the reference
    arithmetic and FP32 accuracy against FP64 (tests/test_oracle.py)."""
    x = _c(x)
    sfx, _ = _sfx(x.dtype)
    arrs = [_c(a, x.dtype) for a in (x, y, wxc, wyc, offx, offy, ratio)]
    if buf is None:
        buf = np.zeros(num_bins_x * num_bins_y, dtype=np.int64)
    getattr(lib(), "dpo_triangle_density_" + _arith(arithmetic) + sfx)(
        *[_p(a) for a in arrs], ctypes.c_int64(len(x)), ctypes.c_int(num_bins_x),
        ctypes.c_int(num_bins_y), ctypes.c_double(xl), ctypes.c_double(yl),
        ctypes.c_double(bin_size_x), ctypes.c_double(bin_size_y), ctypes.c_int64(scale), _p(buf))
    return buf


def exact_density_fixed(x, y, wx, wy, bin_center_x, bin_center_y, num_bins_x, num_bins_y,
                        xl, yl, xh, yh, bin_size_x, bin_size_y, scale, fixed_flag=True, buf=None,
                        arithmetic="reference"):
    """int64 exact-overlap map of fixed cells; electric_density_map.cpp:311-358.
    Pinned by: oracle/_ref fixed_density_map (bit-exact), brute force.
    arithmetic="bin_relative" (fixed_flag only): overlaps measured from each
    bin's lower edge (dp_oracle.c DPO_DEFINE_REL)."""
    x = _c(x)
    sfx, _ = _sfx(x.dtype)
    arrs = [_c(a, x.dtype) for a in (x, y, wx, wy, bin_center_x, bin_center_y)]
    if buf is None:
        buf = np.zeros(num_bins_x * num_bins_y, dtype=np.int64)
    if arithmetic == "bin_relative":
        if not fixed_flag:
            raise ValueError("the bin-relative exact overlap has no boundary stretch")
        getattr(lib(), "dpo_exact_density_rel_" + sfx)(
            *[_p(a) for a in arrs], ctypes.c_int64(len(x)), ctypes.c_int(num_bins_x),
            ctypes.c_int(num_bins_y), *[ctypes.c_double(v) for v in (xl, yl, bin_size_x, bin_size_y)],
            ctypes.c_int64(scale), _p(buf))
        return buf
    _arith(arithmetic)
    getattr(lib(), "dpo_exact_density_" + sfx)(
        *[_p(a) for a in arrs], ctypes.c_int64(len(x)), ctypes.c_int(num_bins_x),
        ctypes.c_int(num_bins_y), *[ctypes.c_double(v) for v in (xl, yl, xh, yh, bin_size_x, bin_size_y)],
        ctypes.c_int(1 if fixed_flag else 0), ctypes.c_int64(scale), _p(buf))
    return buf


def fixed_add_cpu(buf, scale, dst):
    """dst += buf * (1.0/scale) in double (atomic_ops.h:48-54).  In place."""
    sfx, _ = _sfx(dst.dtype)
    getattr(lib(), "dpo_fixed_add_cpu_" + sfx)(_p(buf), ctypes.c_int64(buf.size), ctypes.c_int64(scale), _p(dst))
    return dst


def fixed_quantize_gpu(init, scale):
    """u64(init*T(scale)) (atomic_ops.cuh:53-59, first copyScaleArray)."""
    init = _c(init)
    sfx, _ = _sfx(init.dtype)
    buf = np.zeros(init.size, dtype=np.int64)
    getattr(lib(), "dpo_fixed_quantize_gpu_" + sfx)(_p(init), ctypes.c_int64(init.size), ctypes.c_int64(scale), _p(buf))
    return buf


def fixed_to_map_gpu(buf, scale, dtype):
    """T(u64) * T(1/scale) (atomic_ops.cuh:53-59, second copyScaleArray)."""
    sfx, _ = _sfx(dtype)
    out = np.zeros(buf.size, dtype=dtype)
    getattr(lib(), "dpo_fixed_to_map_gpu_" + sfx)(_p(buf), ctypes.c_int64(buf.size), ctypes.c_int64(scale), _p(out))
    return out


# ---------------------------------------------------------------------------
# ElectricOverflow host-side preparation (electric_overflow.py:162-227)
# ---------------------------------------------------------------------------

class DensityParams:
    """The derived per-node arrays and scalars of ElectricOverflow.reset.
    electric_overflow.py:162-227.  Pinned by: closed form (the stretched area
    w_c*h_c*ratio equals w*h; tests/test_oracle.py)."""

    def __init__(self, node_size_x, node_size_y, xl, yl, xh, yh, bin_size_x, bin_size_y,
                 num_movable_nodes, num_terminals, num_filler_nodes, target_density,
                 movable_macro_mask=None):
        dt = node_size_x.dtype
        sqrt2 = math.sqrt(2)
        self.node_size_x = node_size_x
        self.node_size_y = node_size_y
        self.wxc = np.maximum(node_size_x, dt.type(bin_size_x * sqrt2)).astype(dt)
        self.wyc = np.maximum(node_size_y, dt.type(bin_size_y * sqrt2)).astype(dt)
        self.offx = ((node_size_x - self.wxc) * dt.type(0.5)).astype(dt)
        self.offy = ((node_size_y - self.wyc) * dt.type(0.5)).astype(dt)
        self.ratio = ((node_size_x * node_size_y) / (self.wxc * self.wyc)).astype(dt)
        if target_density < 1 and movable_macro_mask is not None:
            r = self.ratio[:num_movable_nodes]
            r[np.asarray(movable_macro_mask, dtype=bool)] = dt.type(target_density)
        self.xl, self.yl, self.xh, self.yh = xl, yl, xh, yh
        self.bin_size_x, self.bin_size_y = bin_size_x, bin_size_y
        self.num_bins_x = int(round((xh - xl) / bin_size_x))
        self.num_bins_y = int(round((yh - yl) / bin_size_y))
        self.num_movable_nodes = num_movable_nodes
        self.num_terminals = num_terminals
        self.num_filler_nodes = num_filler_nodes
        self.target_density = target_density
        self.scale = fixed_scale(xl, yl, xh, yh)
        self.dtype = dt


def bin_centers(l, h, bin_size, dtype=np.float64):
    """PlaceDB.bin_centers, dreamplace/PlaceDB.py:331-346."""
    n = int(np.ceil((h - l) / bin_size))
    c = np.zeros(n, dtype=dtype)
    for i in range(n):
        bl = l + i * bin_size
        bh = min(bl + bin_size, h)
        c[i] = (bl + bh) / 2
    return c


def fixed_density_map(dp, pos, bin_center_x, bin_center_y, deterministic=True,
                      arithmetic="reference"):
    """initial_density_map: exact map of fixed cells, clamp_max(bin area), * t.
    electric_density_map.cpp:155-213 (deterministic branch) and
    electric_overflow.py:229-258.  Returns [Mx, My] in dp.dtype."""
    n = pos.size // 2
    m0, m1 = dp.num_movable_nodes, dp.num_movable_nodes + dp.num_terminals
    out = np.zeros(dp.num_bins_x * dp.num_bins_y, dtype=dp.dtype)
    if dp.num_terminals:
        buf = exact_density_fixed(pos[m0:m1], pos[n + m0:n + m1], dp.node_size_x[m0:m1],
                                  dp.node_size_y[m0:m1], bin_center_x, bin_center_y,
                                  dp.num_bins_x, dp.num_bins_y, dp.xl, dp.yl, dp.xh, dp.yh,
                                  dp.bin_size_x, dp.bin_size_y, dp.scale, True,
                                  arithmetic=arithmetic)
        fixed_add_cpu(buf, dp.scale, out)
        np.minimum(out, dp.dtype.type(dp.bin_size_x * dp.bin_size_y), out=out)
    out *= dp.dtype.type(dp.target_density)
    return out.reshape(dp.num_bins_x, dp.num_bins_y)


def density_fixed_sum(dp, pos, num_filler_nodes=None, arithmetic="reference"):
    """int64 scatter of movables and fillers (no initial map)."""
    n = pos.size // 2
    nf = dp.num_filler_nodes if num_filler_nodes is None else num_filler_nodes
    buf = np.zeros(dp.num_bins_x * dp.num_bins_y, dtype=np.int64)
    rng = [(0, dp.num_movable_nodes)]
    if nf:
        rng.append((n - nf, n))
    for b, e in rng:
        triangle_density_fixed(pos[b:e], pos[n + b:n + e], dp.wxc[b:e], dp.wyc[b:e], dp.offx[b:e],
                               dp.offy[b:e], dp.ratio[b:e], dp.num_bins_x, dp.num_bins_y,
                               dp.xl, dp.yl, dp.bin_size_x, dp.bin_size_y, dp.scale, buf,
                               arithmetic=arithmetic)
    return buf


def density_map(dp, pos, initial_density_map, num_filler_nodes=None, convention="cpu",
                arithmetic="reference"):
    """Deterministic density map of movables+fillers over the initial map.
    convention "cpu": init + (double)sum*2^-32 -> T  (electric_density_map.cpp:125-140);
    convention "gpu": T(u64(quant(init) + sum)) * T(2^-32)
    (electric_density_map_cuda_kernel.cu:433-463)."""
    buf = density_fixed_sum(dp, pos, num_filler_nodes, arithmetic=arithmetic)
    init = np.ascontiguousarray(initial_density_map, dtype=dp.dtype).reshape(-1)
    if convention == "cpu":
        out = init.copy()
        fixed_add_cpu(buf, dp.scale, out)
    else:
        q = fixed_quantize_gpu(init, dp.scale)
        out = fixed_to_map_gpu(q + buf, dp.scale, dp.dtype)
    return out.reshape(dp.num_bins_x, dp.num_bins_y)


def electric_force(dp, pos, field_x, field_y, grad=1.0, arithmetic="reference"):
    """-(electric_force(...))*grad for movables and fillers, 0 for fixed cells:
    the density op's backward (electric_potential.py:219-245,
    electric_force.cpp:76-172).  Returns [2N] in dp.dtype."""
    n = pos.size // 2
    T = dp.dtype
    sfx, ct = _sfx(T)
    out = np.zeros(2 * n, dtype=T)
    fx = _c(field_x, T).reshape(-1)
    fy = _c(field_y, T).reshape(-1)
    rng = [(0, dp.num_movable_nodes)]
    if dp.num_filler_nodes:
        rng.append((n - dp.num_filler_nodes, n))
    for b, e in rng:
        gx = np.zeros(e - b, dtype=T)
        gy = np.zeros(e - b, dtype=T)
        arrs = [_c(a, T) for a in (pos[b:e], pos[n + b:n + e], dp.wxc[b:e], dp.wyc[b:e],
                                   dp.offx[b:e], dp.offy[b:e], dp.ratio[b:e])]
        getattr(lib(), "dpo_electric_force_" + _arith(arithmetic) + sfx)(
            ct(grad), _p(fx), _p(fy), *[_p(a) for a in arrs], ctypes.c_int64(e - b),
            ctypes.c_int(dp.num_bins_x), ctypes.c_int(dp.num_bins_y), ctypes.c_double(dp.xl),
            ctypes.c_double(dp.yl), ctypes.c_double(dp.bin_size_x), ctypes.c_double(dp.bin_size_y),
            _p(gx), _p(gy))
        out[b:e] = -gx
        out[n + b:n + e] = -gy
    return out


# ---------------------------------------------------------------------------
# Spectral transforms (closed forms), FP64
# ---------------------------------------------------------------------------

def cos_matrix(M):
    """C_M[u, i] = cos(pi*u*(2i+1)/(2M))."""
    u = np.arange(M, dtype=np.float64)[:, None]
    i = np.arange(M, dtype=np.float64)[None, :]
    return np.cos(np.pi * u * (2 * i + 1) / (2 * M))


def sin_matrix(M):
    """S_M[u, i] = sin(pi*u*(2i+1)/(2M))."""
    u = np.arange(M, dtype=np.float64)[:, None]
    i = np.arange(M, dtype=np.float64)[None, :]
    return np.sin(np.pi * u * (2 * i + 1) / (2 * M))


def _cw(M):
    c = np.full(M, 2.0)
    c[0] = 1.0
    return c


def dct2(x):
    """DCT2(x)[u,v] = 4/(MN) sum_ij x[i,j] C_M[u,i] C_N[v,j].
    Convention of dreamplace/ops/dct/dct2_fft2.py DCT2 (tested there against
    dct2_N, unittest/ops/dct_unittest.py:170-248).  Pinned by: oracle/_ref
    dct2_fft2, brute-force sums."""
    M, N = x.shape
    return (4.0 / (M * N)) * cos_matrix(M) @ np.asarray(x, np.float64) @ cos_matrix(N).T


def idct2(x):
    """IDCT2(x)[m,n] = sum_uv c_u c_v x[u,v] C_M[u,m] C_N[v,n], c_0=1, c_k=2."""
    M, N = x.shape
    return (_cw(M)[:, None] * cos_matrix(M)).T @ np.asarray(x, np.float64) @ (_cw(N)[:, None] * cos_matrix(N))


def idxst_idct(x):
    """IDXST_IDCT(x)[m,n] = 2 sum_uv c_v x[u,v] S_M[u,m] C_N[v,n]."""
    M, N = x.shape
    return 2.0 * sin_matrix(M).T @ np.asarray(x, np.float64) @ (_cw(N)[:, None] * cos_matrix(N))


def idct_idxst(x):
    """IDCT_IDXST(x)[m,n] = 2 sum_uv c_u x[u,v] C_M[u,m] S_N[v,n]."""
    M, N = x.shape
    return 2.0 * (_cw(M)[:, None] * cos_matrix(M)).T @ np.asarray(x, np.float64) @ sin_matrix(N)


def spectral_weights(M, N, bin_size_x, bin_size_y):
    """wu, wv and the three spectral multipliers of electric_potential.py:504-519."""
    wu = (np.arange(M, dtype=np.float64) * (2 * np.pi / M))[:, None]
    wv = (np.arange(N, dtype=np.float64) * (2 * np.pi / N))[None, :] * (bin_size_x / bin_size_y)
    w2 = wu ** 2 + wv ** 2
    w2[0, 0] = 1.0
    inv = 1.0 / w2
    inv[0, 0] = 0.0
    return inv, wu * inv * 0.5, wv * inv * 0.5


def poisson_solve(density_map, bin_size_x, bin_size_y):
    """The spectral solve of ElectricPotentialFunction.forward
    (electric_potential.py:137-186): rho = map/(bx*by); a = DCT2(rho);
    Ex = IDXST_IDCT(a*wu/(2(wu^2+wv^2))); Ey = IDCT_IDXST(a*wv/(2(...)));
    phi = IDCT2(a/(wu^2+wv^2)); energy = sum(phi*rho).  All FP64.
    Pinned by: the closed-form checks of tests/test_oracle.py (zero field for a
    constant map, E = -d(phi)/d(bin index) of the cosine interpolant, the
    spectral form of the energy) and oracle/_ref dct2_fft2."""
    M, N = density_map.shape
    rho = np.asarray(density_map, np.float64) * (1.0 / (bin_size_x * bin_size_y))
    a = dct2(rho)
    inv, wu_h, wv_h = spectral_weights(M, N, bin_size_x, bin_size_y)
    ex = idxst_idct(a * wu_h)
    ey = idct_idxst(a * wv_h)
    phi = idct2(a * inv)
    energy = float(np.sum(phi * rho))
    return dict(rho=rho, auv=a, phi=phi, field_x=ex, field_y=ey, energy=energy)


# ---------------------------------------------------------------------------
# Wirelength, pin positions, metrics
# ---------------------------------------------------------------------------

def wa_merged_forward(pin_pos, flat_netpin, netpin_start, net_weights, net_mask, gamma):
    """Merged WA forward (weighted_average_wirelength_merged.cpp:31-180).
    Returns (wl = sum w_n*WL_n, partial_wl[E] pre-weight, grad_intermediate[2P]).
    Pinned by: the reference WA golden (unittest/ops/weighted_average_wirelength_unittest.py:92-162),
    oracle/_ref merged forward, central finite differences, gamma->0 HPWL limit."""
    pin_pos = _c(pin_pos)
    T = pin_pos.dtype
    sfx, ct = _sfx(T)
    P = pin_pos.size // 2
    E = len(netpin_start) - 1
    px, py = _c(pin_pos[:P]), _c(pin_pos[P:])
    fnp = _c(flat_netpin, np.int32)
    nps = _c(netpin_start, np.int32)
    nm = _c(net_mask, np.uint8)
    partial = np.zeros(E, dtype=T)
    gi = np.zeros(2 * P, dtype=T)
    gix, giy = np.zeros(P, dtype=T), np.zeros(P, dtype=T)
    getattr(lib(), "dpo_wa_merged_fwd_" + sfx)(
        _p(px), _p(py), _p(fnp), _p(nps), _p(nm), ctypes.c_int64(E), ct(T.type(1.0) / T.type(gamma)),
        _p(partial), _p(gix), _p(giy))
    gi[:P], gi[P:] = gix, giy
    w = np.asarray(net_weights, dtype=T)
    wl = float(np.sum((partial * w).astype(np.float64))) if w.size else float(np.sum(partial.astype(np.float64)))
    return wl, partial, gi


def lse_merged_forward(pin_pos, flat_netpin, netpin_start, net_weights, net_mask, gamma):
    """Merged log-sum-exp wirelength forward
    (logsumexp_wirelength_merged.cpp:14-85, 99-146), per unmasked net and axis:
    WL = (log S+ + log S-)*gamma + max - min with S+ = sum exp((x-max)/gamma),
    S- = sum exp((min-x)/gamma); d/dx_p = e+_p/S+ - e-_p/S-.  Returns
    (wl = sum w_n*WL_n in FP64, partial_wl[E] pre-weight, grad_intermediate[2P]).
    Sums run in pin order in T; exponentials and logs are numpy's in T.
    Pinned by: oracle/_ref merged forward, the two-pin closed form
    |d| + 2*gamma*log(1 + exp(-|d|/gamma)), the gamma->0 HPWL limit and
    central finite differences."""
    pin_pos = np.asarray(pin_pos)
    T = pin_pos.dtype.type
    P = pin_pos.size // 2
    fnp = np.asarray(flat_netpin, dtype=np.int64)
    nps = np.asarray(netpin_start, dtype=np.int64)
    E = len(nps) - 1
    mask = np.asarray(net_mask, dtype=bool)
    g = T(gamma)
    ig = T(1) / g
    partial = np.zeros(E, dtype=T)
    gi = np.zeros(2 * P, dtype=T)
    deg = np.diff(nps)
    live = mask & (deg > 0)
    net_of = np.repeat(np.arange(E), deg)
    for ax in (0, 1):
        v = pin_pos[ax * P:(ax + 1) * P][fnp]
        vmax = np.full(E, -np.finfo(T).max, dtype=T)
        vmin = np.full(E, np.finfo(T).max, dtype=T)
        np.maximum.at(vmax, net_of, v)
        np.minimum.at(vmin, net_of, v)
        ep = np.exp((v - vmax[net_of]) * ig)
        en = np.exp((vmin[net_of] - v) * ig)
        sp = np.zeros(E, dtype=T)
        sn = np.zeros(E, dtype=T)
        np.add.at(sp, net_of, ep)   # unbuffered, in pin order
        np.add.at(sn, net_of, en)
        with np.errstate(divide="ignore", invalid="ignore"):
            term = (np.log(sp) + np.log(sn)) * g
            rp, rn = T(1) / sp, T(1) / sn
        if ax == 0:
            partial = np.where(live, (term + vmax) - vmin, T(0)).astype(T)
        else:
            partial = np.where(live, ((partial + term) + vmax) - vmin, T(0)).astype(T)
        d = ep * rp[net_of] - en * rn[net_of]
        on = live[net_of]
        gi[ax * P + fnp[on]] = d[on]
    w = np.asarray(net_weights, dtype=T)
    wl = float(np.sum((partial * w).astype(np.float64))) if w.size else float(np.sum(partial.astype(np.float64)))
    return wl, partial, gi


def wa_merged_backward(grad, grad_intermediate, pin2net_map, net_weights, net_mask, pin_mask):
    """WA backward (weighted_average_wirelength.py:197-213,
    weighted_average_wirelength_merged.cpp:193-230, functional.h:10-34):
    g = grad_intermediate*grad, g[p] *= w[net(p)] for unmasked nets, g[pin_mask]=0."""
    gi = np.asarray(grad_intermediate)
    T = gi.dtype
    P = gi.size // 2
    g = gi * T.type(grad)
    p2n = np.asarray(pin2net_map, dtype=np.int64)
    if np.asarray(net_weights).size:
        w = np.asarray(net_weights, dtype=T)[p2n]
        on = np.asarray(net_mask, dtype=bool)[p2n]
        w = np.where(on, w, T.type(1))
        g[:P] *= w
        g[P:] *= w
    pm = np.asarray(pin_mask, dtype=bool)
    g[:P][pm] = 0
    g[P:][pm] = 0
    return g


def wa_net_weight_grad(grad, partial_wl, net_mask):
    """d(sum_n w_n WL_n)/d(w_n) = grad*WL_n for unmasked nets (new adjoint; the
    reference returns None, weighted_average_wirelength.py:213).
    Pinned by: closed form and central finite differences in net weights."""
    return np.where(np.asarray(net_mask, dtype=bool), partial_wl * grad, 0).astype(partial_wl.dtype)


def pin_pos_forward(pos, pin_offset_x, pin_offset_y, pin2node_map):
    """pin_pos.cpp:133-146.  Pinned by: reference pin_pos golden
    (unittest/ops/pin_pos_unittest.py:21-33 index_select form) and oracle/_ref."""
    pos = _c(pos)
    T = pos.dtype
    sfx, _ = _sfx(T)
    n = pos.size // 2
    P = len(pin2node_map)
    x, y = _c(pos[:n]), _c(pos[n:])
    out_x, out_y = np.zeros(P, dtype=T), np.zeros(P, dtype=T)
    getattr(lib(), "dpo_pin_pos_fwd_" + sfx)(
        _p(x), _p(y), _p(_c(pin_offset_x, T)), _p(_c(pin_offset_y, T)), _p(_c(pin2node_map, np.int64)),
        ctypes.c_int64(P), _p(out_x), _p(out_y))
    return np.concatenate([out_x, out_y])


def pin_pos_backward(grad_pin, flat_node2pin, node2pin_start, num_physical_nodes, num_nodes):
    """pin_pos.cpp:148-167; fillers receive 0.  Pinned by: adjoint identity
    <pin_pos(dx), g> = <dx, pin_pos^T g> and oracle/_ref."""
    grad_pin = _c(grad_pin)
    T = grad_pin.dtype
    sfx, _ = _sfx(T)
    P = grad_pin.size // 2
    gx, gy = np.zeros(num_physical_nodes, dtype=T), np.zeros(num_physical_nodes, dtype=T)
    getattr(lib(), "dpo_pin_pos_bwd_" + sfx)(
        _p(_c(grad_pin[:P])), _p(_c(grad_pin[P:])), _p(_c(flat_node2pin, np.int32)),
        _p(_c(node2pin_start, np.int32)), ctypes.c_int64(num_physical_nodes), _p(gx), _p(gy))
    out = np.zeros(2 * num_nodes, dtype=T)
    out[:num_physical_nodes] = gx
    out[num_nodes:num_nodes + num_physical_nodes] = gy
    return out


def hpwl_per_net(pin_pos, flat_netpin, netpin_start, net_mask):
    """Unweighted per-net HPWL (hpwl.cpp:60-84).  Pinned by: the reference HPWL
    golden (unittest/ops/hpwl_unittest.py:39-96) and oracle/_ref."""
    pin_pos = _c(pin_pos)
    T = pin_pos.dtype
    sfx, _ = _sfx(T)
    P = pin_pos.size // 2
    E = len(netpin_start) - 1
    out = np.zeros(E, dtype=T)
    getattr(lib(), "dpo_hpwl_" + sfx)(
        _p(_c(pin_pos[:P])), _p(_c(pin_pos[P:])), _p(_c(flat_netpin, np.int32)),
        _p(_c(netpin_start, np.int32)), _p(_c(net_mask, np.uint8)), ctypes.c_int64(E), _p(out))
    return out


def hpwl(pin_pos, flat_netpin, netpin_start, net_weights, net_mask):
    """sum_n w_n*HPWL_n accumulated in FP64 (hpwl.cpp:27-52 multiplies by the
    weights then sums)."""
    h = hpwl_per_net(pin_pos, flat_netpin, netpin_start, net_mask).astype(np.float64)
    w = np.asarray(net_weights, dtype=np.float64)
    return float(np.sum(h * w)) if w.size else float(np.sum(h))


def move_boundary(pos, node_size_x, node_size_y, xl, yl, xh, yh, num_movable_nodes, num_filler_nodes):
    """In-place clamp of movables and fillers (move_boundary.cpp:42-62).
    Pinned by: closed form and oracle/_ref."""
    T = pos.dtype
    sfx, _ = _sfx(T)
    n = pos.size // 2
    x, y = _c(pos[:n]), _c(pos[n:])
    getattr(lib(), "dpo_move_boundary_" + sfx)(
        _p(x), _p(y), _p(_c(node_size_x, T)), _p(_c(node_size_y, T)),
        *[ctypes.c_double(v) for v in (xl, yl, xh, yh)], ctypes.c_int64(n),
        ctypes.c_int64(num_movable_nodes), ctypes.c_int64(num_filler_nodes))
    pos[:n], pos[n:] = x, y
    return pos


def pin_weight_sum(net_weights, flat_nodepin, nodepin_start, pin2net_map, num_nodes):
    """Per-node sum of the weights of its pins' nets (pws.cpp:12-48).
    Pinned by: closed form (np.add.at over pins) and oracle/_ref."""
    w = _c(net_weights)
    T = w.dtype
    sfx, _ = _sfx(T)
    nphys = len(nodepin_start) - 1
    out = np.zeros(num_nodes, dtype=T)
    getattr(lib(), "dpo_pws_" + sfx)(
        _p(w), _p(_c(flat_nodepin, np.int32)), _p(_c(nodepin_start, np.int32)),
        _p(_c(pin2net_map, np.int32)), ctypes.c_int64(nphys), _p(out))
    return out


def objective_and_gradient(d, density_weight, gamma, density_factor=1.0, quad_coeff=None,
                           arithmetic="reference"):
    """obj = WA(pin_pos(pos)) + lambda*energy and its gradient for a design
    dict (dreamplace_amd.synth layout), composed from the functions above as
    PlaceObj.obj_fn / obj_and_grad_fn compose the ops
    (dreamplace/PlaceObj.py:293-320,394-411), FP64 spectral solve.
    With `density_factor` f and the quadratic penalty coefficient c
    (PlaceObj.py:308-318): obj = WA + (f*lambda) * E*(1 + c*E), whose
    derivative in E, g*(1 + c*E) + (g*E)*c with g = f*lambda, scales the field
    gather.  arithmetic: the overlap arithmetic of the maps and the force
    ("reference" or "bin_relative").  Returns (obj, grad[2N], parts)."""
    dp = DensityParams(d["node_size_x"], d["node_size_y"], d["xl"], d["yl"], d["xh"], d["yh"],
                       d["bin_size_x"], d["bin_size_y"], d["num_movable_nodes"], d["num_terminals"],
                       d["num_filler_nodes"], d["target_density"], d["movable_macro_mask"])
    pos = d["pos"]
    T = pos.dtype
    pin_pos = pin_pos_forward(pos, d["pin_offset_x"], d["pin_offset_y"], d["pin2node_map"])
    wl, partial, gi = wa_merged_forward(pin_pos, d["flat_net2pin_map"], d["flat_net2pin_start_map"],
                                        d["net_weights"], d["net_mask_ignore_large_degrees"], gamma)
    init = fixed_density_map(dp, pos, d["bin_center_x"], d["bin_center_y"], arithmetic=arithmetic)
    dmap = density_map(dp, pos, init, arithmetic=arithmetic)
    sol = poisson_solve(dmap, d["bin_size_x"], d["bin_size_y"])
    g = density_weight
    energy = sol["energy"]
    if quad_coeff is not None or density_factor != 1.0:
        g = T.type(T.type(density_factor) * T.type(density_weight))
        e = T.type(energy)
        if quad_coeff is not None:
            c = T.type(quad_coeff)
            g = g * (T.type(1) + c * e) + (g * e) * c
            energy = e * (T.type(1) + c * e)
    dgrad = electric_force(dp, pos, sol["field_x"].astype(T), sol["field_y"].astype(T), grad=g,
                           arithmetic=arithmetic)
    gpin = wa_merged_backward(1.0, gi, d["pin2net_map"], d["net_weights"],
                              d["net_mask_ignore_large_degrees"], d["pin_mask_ignore_fixed_macros"])
    wgrad = pin_pos_backward(gpin, d["flat_node2pin_map"], d["flat_node2pin_start_map"],
                             d["num_physical_nodes"], d["num_nodes"])
    obj = wl + (density_factor * density_weight) * energy
    parts = dict(sol)
    parts.update(wl=wl, density_map=dmap, density_grad=dgrad, wl_grad=wgrad)
    return obj, wgrad + dgrad, parts


def energy_gradient(dp, pos, potential, grad=1.0):
    """The exact gradient of the density energy sum(phi*rho) with respect to
    the positions (dp_oracle.c dpo_energy_gradient): grad * 2/(bx*by) *
    ratio * sum (d overlap) * phi for movables and fillers, 0 for fixed
    cells.  Returns [2N] in dp.dtype.  Pinned by central finite differences
    of the oracle's energy (tests/test_oracle.py)."""
    n = pos.size // 2
    T = dp.dtype
    sfx, ct = _sfx(T)
    out = np.zeros(2 * n, dtype=T)
    phi = _c(potential, T).reshape(-1)
    rng = [(0, dp.num_movable_nodes)]
    if dp.num_filler_nodes:
        rng.append((n - dp.num_filler_nodes, n))
    for b, e in rng:
        gx = np.zeros(e - b, dtype=T)
        gy = np.zeros(e - b, dtype=T)
        arrs = [_c(a, T) for a in (pos[b:e], pos[n + b:n + e], dp.wxc[b:e], dp.wyc[b:e],
                                   dp.offx[b:e], dp.offy[b:e], dp.ratio[b:e])]
        getattr(lib(), "dpo_energy_gradient_" + sfx)(
            ct(grad), _p(phi), *[_p(a) for a in arrs], ctypes.c_int64(e - b),
            ctypes.c_int(dp.num_bins_x), ctypes.c_int(dp.num_bins_y), ctypes.c_double(dp.xl),
            ctypes.c_double(dp.yl), ctypes.c_double(dp.bin_size_x), ctypes.c_double(dp.bin_size_y),
            _p(gx), _p(gy))
        out[b:e] = gx
        out[n + b:n + e] = gy
    return out


def net_weighting(scheme, net_weights, net_criticality, netpin_start, ignore_net_degree,
                  net_slack=None, wns=None, net_critical=None, decay=0.5, max_net_weight=np.inf):
    """The weighting step of timing_cpp.update_net_weights, in the arrays'
    dtype: "adams" (net_weighting_scheme.h:106-148) halves each net's
    criticality and adds 0.5 for nets on a reported critical path, then
    multiplies the weight by 1 + criticality; "lilith" (:150-188) blends the
    criticality with nc = max(0, slack/wns) of negative-slack nets as
    (1+c)^decay * (1+nc)^(1-decay) - 1 when wns < 0, then multiplies the
    weight by 1 + criticality and caps it at max_net_weight.  Nets of degree
    above ignore_net_degree keep their weight (and, for adams, their
    criticality).  Returns (weights, criticality).  Pinned by hand-derived
    updates (tests/test_oracle.py)."""
    T = np.asarray(net_weights).dtype.type
    w = np.array(net_weights, copy=True)
    c = np.array(net_criticality, dtype=w.dtype, copy=True)
    keep = np.diff(np.asarray(netpin_start, dtype=np.int64)) <= ignore_net_degree
    if scheme == "adams":
        flag = np.asarray(net_critical) != 0
        c[keep] = c[keep] * T(0.5) + np.where(flag[keep], T(0.5), T(0))
        w[keep] = w[keep] * (T(1) + c[keep])
        return w, c
    if scheme != "lilith":
        raise ValueError("scheme must be adams or lilith")
    if wns < 0:
        sl = np.asarray(net_slack, dtype=w.dtype)
        nc = np.where(sl < 0, np.maximum(T(0), sl / T(wns)), T(0)).astype(w.dtype)
        c = (np.power(T(1) + c, T(decay)) * np.power(T(1) + nc, T(1) - T(decay)) - T(1)).astype(w.dtype)
    nw = w[keep] * (T(1) + c[keep])
    w[keep] = np.where(nw > T(max_net_weight), T(max_net_weight), nw)
    return w, c


def overflow(density_map_movable, target_density, bin_size_x, bin_size_y, total_movable_area):
    """ElectricOverflow.forward metric (electric_overflow.py:260-278,
    EvalMetrics.py:116-128): sum(max(D - t*A_bin, 0)) / movable area, and
    max(D)/A_bin, in FP64."""
    d = np.asarray(density_map_movable, dtype=np.float64)
    a = bin_size_x * bin_size_y
    return float(np.sum(np.maximum(d - target_density * a, 0.0))) / total_movable_area, float(d.max() / a)
