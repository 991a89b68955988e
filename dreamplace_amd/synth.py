"""Synthetic placement designs for the five benchmark workloads.

The generator produces the arrays a DREAMPlace placer hands to the density and
wirelength ops (the ``PlaceDataCollection`` fields of
``dreamplace/BasicPlace.py:58-194``), from a seeded ``numpy.random.default_rng``.
It mirrors the reference's formulas where the reference fixes them:

* filler cells: width = mean of the 5%..95% width quantile of movables, height =
  row height (1), count = round(max(space*t - movable cell area, 0) / filler area)
  (``dreamplace/PlaceDB.py:917-936``);
* ``net_mask_ignore_large_degrees = 2 <= degree < ignore_net_degree``,
  ``net_mask_all = degree >= 2``, ``pin_mask_ignore_fixed_macros = pin on a node >=
  num_movable`` (``dreamplace/BasicPlace.py:159-183``);
* ``sorted_node_map = argsort(movable widths)`` (``dreamplace/BasicPlace.py:186-188``);
* movable macros: area > 10x mean movable area and height > 2 rows
  (``dreamplace/PlaceDB.py:803``);
* fillers start uniform in the die (``dreamplace/BasicPlace.py:351-363``).

Net degrees follow 2 (60%), 3 (20%), U{4..10} (15%), U{11..50} (5%), capped at
``max_degree``; pins go on distinct physical nodes, at offsets uniform inside the
node measured from its lower-left corner.  Pins are numbered net-major, as a
Bookshelf ``.nets`` file numbers them, so ``flat_net2pin_map`` is the identity.
"""
import math

import numpy as np


class Design(dict):
    """A dict of arrays and scalars with attribute access."""

    __getattr__ = dict.__getitem__

    def __setattr__(self, k, v):
        self[k] = v


def _degrees(rng, num_nets, max_degree):
    u = rng.random(num_nets)
    deg = np.empty(num_nets, dtype=np.int64)
    deg[u < 0.60] = 2
    m3 = (u >= 0.60) & (u < 0.80)
    deg[m3] = 3
    m4 = (u >= 0.80) & (u < 0.95)
    deg[m4] = rng.integers(4, 11, size=int(m4.sum()))
    m5 = u >= 0.95
    deg[m5] = rng.integers(11, 51, size=int(m5.sum()))
    return np.minimum(deg, max_degree)


def _distinct_nodes(rng, deg, num_nodes):
    """One node per pin, distinct within each net."""
    P = int(deg.sum())
    net_of_pin = np.repeat(np.arange(len(deg)), deg)
    nodes = rng.integers(0, num_nodes, size=P)
    for _ in range(64):
        order = np.lexsort((nodes, net_of_pin))
        s_net, s_node = net_of_pin[order], nodes[order]
        dup = np.zeros(P, dtype=bool)
        dup[1:] = (s_net[1:] == s_net[:-1]) & (s_node[1:] == s_node[:-1])
        if not dup.any():
            return nodes
        idx = order[dup]
        nodes[idx] = rng.integers(0, num_nodes, size=len(idx))
    raise RuntimeError("could not draw distinct pin nodes")


def _place_blocks(rng, sizes_x, sizes_y, xl, yl, xh, yh, max_tries=200):
    """Random non-overlapping sites for blocks, largest first (rejection sampling
    on a uniform grid hash).  Blocks that find no site after max_tries are
    dropped; the caller reports the count it got."""
    n = len(sizes_x)
    order = np.argsort(-(sizes_x * sizes_y), kind="stable")
    cell = float(max(sizes_x.max(), sizes_y.max()))
    grid = {}
    px = np.zeros(n)
    py = np.zeros(n)
    ok = np.zeros(n, dtype=bool)
    for i in order:
        w, h = sizes_x[i], sizes_y[i]
        for _ in range(max_tries):
            x = rng.uniform(xl, xh - w)
            y = rng.uniform(yl, yh - h)
            gx0, gx1 = int((x - xl) // cell), int((x + w - xl) // cell)
            gy0, gy1 = int((y - yl) // cell), int((y + h - yl) // cell)
            clash = False
            for gx in range(gx0 - 1, gx1 + 1):
                for gy in range(gy0 - 1, gy1 + 1):
                    for j in grid.get((gx, gy), ()):
                        if x < px[j] + sizes_x[j] and px[j] < x + w and y < py[j] + sizes_y[j] and py[j] < y + h:
                            clash = True
                            break
                    if clash:
                        break
                if clash:
                    break
            if not clash:
                px[i], py[i], ok[i] = x, y, True
                for gx in range(gx0, gx1 + 1):
                    for gy in range(gy0, gy1 + 1):
                        grid.setdefault((gx, gy), []).append(i)
                break
    return px, py, ok


def _power_law_sides(rng, n, lo, hi, target_area):
    """Integer sides in [lo, hi] from p(s) ~ s^-alpha with alpha chosen so the
    expected total area n*E[s^2] is target_area."""
    s = np.arange(lo, hi + 1, dtype=np.float64)

    def e2(alpha):
        p = s ** -alpha
        return float((p * s * s).sum() / p.sum())

    a0, a1 = 0.0, 6.0
    for _ in range(80):
        am = 0.5 * (a0 + a1)
        if n * e2(am) > target_area:
            a0 = am
        else:
            a1 = am
    p = s ** -(0.5 * (a0 + a1))
    p /= p.sum()
    return rng.choice(s, size=n, p=p)


CONFIGS = {
    # die side, bins per side, target density, #movable cells, cell width range,
    # #fixed blocks, fixed side range, #nets, max degree, movable position model
    0: dict(die=512.0, bins=256, t=1.0, cells=10_000, w=(1, 8), fixed=200, fixed_side=(10, 40),
            fixed_model="uniform", nets=10_000, max_degree=50, init="uniform", macros=0, ios=0),
    1: dict(die=10_000.0, bins=1024, t=0.9, cells=2_200_000, w=(1, 10), fixed=25_000, fixed_side=(5, 200),
            fixed_model="powerlaw20", nets=2_200_000, max_degree=200, init="gauss0.05", macros=0, ios=0),
    3: dict(die=10_000.0, bins=1024, t=0.9, cells=2_200_000, w=(1, 10), fixed=25_000, fixed_side=(5, 200),
            fixed_model="powerlaw20", nets=2_500_000, max_degree=50, init="gauss0.05", macros=0, ios=0,
            pins=10_000_000, weights="loguniform"),
    4: dict(die=20_000.0, bins=2048, t=0.8, cells=1_000_000, w=(1, 10), fixed=0, fixed_side=(5, 20),
            fixed_model="uniform", nets=1_100_000, max_degree=50, init="gauss0.05", macros=600,
            macro_side=(50, 400), ios=1000, io_side=(5, 20)),
}
CONFIGS[2] = dict(CONFIGS[1])


def make_design(config=0, seed=0, dtype=np.float64, scale=1.0, bins=None, ignore_net_degree=100,
                **overrides):
    """Build a synthetic design.

    ``scale`` shrinks every count (cells, blocks, nets) by the given factor and
    the die side by sqrt(scale), keeping densities, so tests can run a config's
    shape at a small size.  ``bins`` overrides the bin count per side.
    """
    c = dict(CONFIGS[config])
    c.update(overrides)
    rng = np.random.default_rng(seed)
    die = c["die"] * math.sqrt(scale)
    xl, yl, xh, yh = 0.0, 0.0, die, die
    nb = bins if bins is not None else max(4, int(round(c["bins"] * math.sqrt(scale))))
    t = c["t"]
    nm_cells = max(1, int(round(c["cells"] * scale)))
    n_macros = int(round(c.get("macros", 0) * scale))
    n_fixed = int(round(c["fixed"] * scale))
    n_ios = int(round(c.get("ios", 0) * scale))

    # movable std cells (h = 1, integer widths) and movable macros
    wx = rng.integers(c["w"][0], c["w"][1] + 1, size=nm_cells).astype(np.float64)
    wy = np.ones(nm_cells)
    if n_macros:
        ms = rng.integers(c["macro_side"][0], c["macro_side"][1] + 1, size=n_macros).astype(np.float64)
        wx = np.concatenate([wx, ms])
        wy = np.concatenate([wy, ms])
    num_movable = len(wx)

    # fixed blocks (+ IO blocks)
    if n_fixed:
        if c["fixed_model"] == "powerlaw20":
            sides = _power_law_sides(rng, n_fixed, c["fixed_side"][0], c["fixed_side"][1], 0.2 * die * die)
        else:
            sides = rng.integers(c["fixed_side"][0], c["fixed_side"][1] + 1, size=n_fixed).astype(np.float64)
        fx_, fy_, ok = _place_blocks(rng, sides, sides, xl, yl, xh, yh)
        fsx, fsy, fx_, fy_ = sides[ok], sides[ok], fx_[ok], fy_[ok]
    else:
        fsx = fsy = fx_ = fy_ = np.zeros(0)
    if n_ios:
        ios = rng.integers(c["io_side"][0], c["io_side"][1] + 1, size=n_ios).astype(np.float64)
        # IO blocks sit on the die boundary, spread over the four sides
        side = rng.integers(0, 4, size=n_ios)
        along = rng.uniform(0, 1, size=n_ios)
        ix = np.where(side < 2, along * (die - ios), np.where(side == 2, 0.0, die - ios))
        iy = np.where(side < 2, np.where(side == 0, 0.0, die - ios), along * (die - ios))
        fsx, fsy = np.concatenate([fsx, ios]), np.concatenate([fsy, ios])
        fx_, fy_ = np.concatenate([fx_, ix]), np.concatenate([fy_, iy])
    num_terminals = len(fsx)

    # movable macro mask (PlaceDB.py:801-804) and areas
    mov_area = wx * wy
    mean_area = mov_area.sum() / num_movable
    macro_mask = (mov_area > mean_area * 10) & (wy > 2.0)
    total_movable_area = float(mov_area.sum())
    macro_area = float(mov_area[macro_mask].sum())
    ov_x = np.maximum(np.minimum(fx_ + fsx, xh) - np.maximum(fx_, xl), 0)
    ov_y = np.maximum(np.minimum(fy_ + fsy, yh) - np.maximum(fy_, yl), 0)
    total_fixed_area = float((ov_x * ov_y).sum())
    total_space_area = (xh - xl) * (yh - yl) - total_fixed_area

    # fillers (PlaceDB.py:917-936)
    order = np.argsort(wx, kind="stable")
    lb, ub = int(num_movable * 0.05), int(num_movable * 0.95)
    filler_w = float(np.mean(wx[order[lb:ub]])) if lb < ub else 0.0
    filler_h = 1.0
    cell_space = total_space_area - macro_area
    filler_area_total = max(cell_space * t - (total_movable_area - macro_area), 0.0)
    num_filler = int(round(filler_area_total / (filler_w * filler_h))) if filler_w > 0 else 0
    if "max_fillers" in c and c["max_fillers"] is not None:
        num_filler = min(num_filler, int(c["max_fillers"]))

    num_physical = num_movable + num_terminals
    num_nodes = num_physical + num_filler
    node_size_x = np.concatenate([wx, fsx, np.full(num_filler, filler_w)])
    node_size_y = np.concatenate([wy, fsy, np.full(num_filler, filler_h)])

    # initial positions
    x = np.zeros(num_nodes)
    y = np.zeros(num_nodes)
    if c["init"] == "uniform":
        x[:num_movable] = rng.uniform(xl, xh - wx)
        y[:num_movable] = rng.uniform(yl, yh - wy)
    else:
        sig = float(c["init"][5:]) * (xh - xl)
        x[:num_movable] = rng.normal((xl + xh) / 2, sig, size=num_movable)
        y[:num_movable] = rng.normal((yl + yh) / 2, sig, size=num_movable)
    x[num_movable:num_physical] = fx_
    y[num_movable:num_physical] = fy_
    if num_filler:
        x[num_physical:] = rng.uniform(xl, xh - filler_w, size=num_filler)
        y[num_physical:] = rng.uniform(yl, yh - filler_h, size=num_filler)

    # nets and pins
    num_nets = max(1, int(round(c["nets"] * scale)))
    if "pins" in c:
        # scale the degree distribution so the total pin count hits the target
        deg = _degrees(rng, num_nets, c["max_degree"])
        target = int(round(c["pins"] * scale))
        while deg.sum() < target:
            k = min(target - int(deg.sum()), num_nets)
            idx = rng.integers(0, num_nets, size=k)
            deg[idx] = np.minimum(deg[idx] + 1, c["max_degree"])
    else:
        deg = _degrees(rng, num_nets, c["max_degree"])
    if c.get("big_nets"):
        # a few nets above 64 pins (some at or above the ignore threshold)
        k = min(int(c["big_nets"]), num_nets)
        deg[rng.choice(num_nets, size=k, replace=False)] = rng.integers(65, 151, size=k)
    deg = np.minimum(deg, num_physical)
    num_pins = int(deg.sum())
    pin2node = _distinct_nodes(rng, deg, num_physical).astype(np.int32)
    pin2net = np.repeat(np.arange(num_nets, dtype=np.int32), deg)
    netpin_start = np.zeros(num_nets + 1, dtype=np.int32)
    netpin_start[1:] = np.cumsum(deg)
    flat_netpin = np.arange(num_pins, dtype=np.int32)
    pin_offset_x = rng.random(num_pins) * node_size_x[pin2node]
    pin_offset_y = rng.random(num_pins) * node_size_y[pin2node]
    node_order = np.argsort(pin2node, kind="stable").astype(np.int32)
    node2pin_start = np.zeros(num_physical + 1, dtype=np.int32)
    node2pin_start[1:] = np.cumsum(np.bincount(pin2node, minlength=num_physical))
    if c.get("weights") == "loguniform":
        net_weights = np.exp(rng.uniform(math.log(0.5), math.log(4.0), size=num_nets))
    else:
        net_weights = np.ones(num_nets)

    degrees = netpin_start[1:] - netpin_start[:-1]
    bsx = (xh - xl) / nb
    bsy = (yh - yl) / nb
    d = Design(
        config=config, seed=seed, dtype=np.dtype(dtype),
        xl=xl, yl=yl, xh=xh, yh=yh, num_bins_x=nb, num_bins_y=nb,
        bin_size_x=bsx, bin_size_y=bsy, target_density=t,
        num_movable_nodes=num_movable, num_terminals=num_terminals, num_filler_nodes=num_filler,
        num_physical_nodes=num_physical, num_nodes=num_nodes, num_nets=num_nets, num_pins=num_pins,
        pos=np.concatenate([x, y]).astype(dtype),
        node_size_x=node_size_x.astype(dtype), node_size_y=node_size_y.astype(dtype),
        pin_offset_x=pin_offset_x.astype(dtype), pin_offset_y=pin_offset_y.astype(dtype),
        pin2node_map=pin2node, pin2net_map=pin2net,
        flat_net2pin_map=flat_netpin, flat_net2pin_start_map=netpin_start,
        flat_node2pin_map=node_order, flat_node2pin_start_map=node2pin_start,
        net_weights=net_weights.astype(dtype),
        net_mask_all=(degrees >= 2).astype(np.uint8),
        net_mask_ignore_large_degrees=((degrees >= 2) & (degrees < ignore_net_degree)).astype(np.uint8),
        pin_mask_ignore_fixed_macros=(pin2node >= num_movable),
        movable_macro_mask=macro_mask,
        sorted_node_map=np.argsort(wx, kind="stable").astype(np.int32),
        total_movable_node_area=total_movable_area, total_fixed_node_area=total_fixed_area,
        total_space_area=total_space_area, filler_size_x=filler_w, filler_size_y=filler_h,
        wa_gamma=4.0 * bsx,
    )
    d.bin_center_x = _bin_centers(xl, xh, bsx, nb).astype(dtype)
    d.bin_center_y = _bin_centers(yl, yh, bsy, nb).astype(dtype)
    return d


def _bin_centers(l, h, bs, n):
    """PlaceDB.bin_centers (dreamplace/PlaceDB.py:331-346), vectorised."""
    bl = l + np.arange(n) * bs
    bh = np.minimum(bl + bs, h)
    return (bl + bh) / 2


def summary(d):
    return ("config %d: nodes %d (movable %d, fixed %d, filler %d), nets %d, pins %d, bins %dx%d"
            % (d.config, d.num_nodes, d.num_movable_nodes, d.num_terminals, d.num_filler_nodes,
               d.num_nets, d.num_pins, d.num_bins_x, d.num_bins_y))
