"""Bookshelf designs: reading (native reader behind dp_bookshelf_*), the
PlaceDB-level arrays, conversion to the placer's design layout, and .pl
output.

* `read(aux)` returns a `PyPlaceDB` with the fields and orders of the
  reference's PyPlaceDB (dreamplace/ops/place_io/src/PyPlaceDB.cpp:161-440);
* `to_design(db, ...)` follows PlaceDB.initialize and BasicPlace's initial
  positions (dreamplace/PlaceDB.py:752-960, 617-655; BasicPlace.py:268-345):
  shift to the origin, scale by 1/site_width, movable-macro mask, fillers,
  bin counts, initial positions;
* `write_pl(path, db, x, y)` writes PlaceDB.write_pl's format
  (dreamplace/PlaceDB.py:1039-1074).

Fixed nodes and IO pins placed in an orientation other than N have their
sizes and pin offsets restated for N by the reader, as PyPlaceDB does
(PyPlaceDB.cpp:669-740); `node_orient` keeps the orientation read.
"""
import ctypes
import math

import numpy as np

from . import _native as N
from .synth import Design, _bin_centers

ORIENTS = ["N", "S", "W", "E", "FN", "FS", "FW", "FE", "UNKNOWN"]
DIRECTS = ["INPUT", "OUTPUT", "INOUT"]
STATUS_FIXED = 3


def _names(buf, count):
    out = buf.raw.split(b"\0")[:count] if count else []
    return [s.decode() for s in out]


class PyPlaceDB:
    """The arrays of one Bookshelf design, in the reference's orders."""

    def __repr__(self):
        return "PyPlaceDB(nodes=%d, terminals=%d, terminal_NIs=%d, nets=%d, pins=%d)" % (
            self.num_nodes, self.num_terminals, self.num_terminal_NIs, self.num_nets, self.num_pins)


def read(aux_path, sort_nets_by_degree=True):
    """Parse the Bookshelf design an .aux file names (PlaceIOFunction.read +
    PlaceIOFunction.pydb, dreamplace/ops/place_io/place_io.py:15-44)."""
    L = N.lib()
    h = ctypes.c_void_p()
    N.call("dp_bookshelf_read", str(aux_path).encode(), int(bool(sort_nets_by_degree)), ctypes.byref(h))
    try:
        s = (ctypes.c_int64 * 9)()
        N.call("dp_bookshelf_sizes", h, s)
        n, nt, nio, E, P, R, nb, eb, mp = (int(v) for v in s)
        f64 = lambda k: np.zeros(k, dtype=np.float64)
        i32 = lambda k: np.zeros(k, dtype=np.int32)
        node_x, node_y, sx, sy = f64(n), f64(n), f64(n), f64(n)
        orient, status = i32(n), i32(n)
        node_names = ctypes.create_string_buffer(max(nb, 1))
        pox, poy, pdir, p2n, p2e = f64(P), f64(P), i32(P), i32(P), i32(P)
        fnp, fnps, fodp, fodps = i32(P), i32(E + 1), i32(P), i32(n + 1)
        w = f64(E)
        ign = np.zeros(E, dtype=np.uint8)
        net_names = ctypes.create_string_buffer(max(eb, 1))
        rows = f64(4 * R)
        die = f64(6)
        p = N.ptr_np
        N.call("dp_bookshelf_export", h, p(node_x), p(node_y), p(sx), p(sy), p(orient), p(status),
               node_names, p(pox), p(poy), p(pdir), p(p2n), p(p2e), p(fnp), p(fnps), p(fodp), p(fodps),
               p(w), p(ign), net_names, p(rows), p(die))
    finally:
        L.dp_bookshelf_free(h)
    db = PyPlaceDB()
    db.num_nodes, db.num_terminals, db.num_terminal_NIs = n, nt, nio
    db.num_nets, db.num_pins, db.num_movable_pins = E, P, mp
    db.node_names = _names(node_names, n)
    db.node_name2id_map = {k: i for i, k in enumerate(db.node_names)}
    db.node_x, db.node_y, db.node_size_x, db.node_size_y = node_x, node_y, sx, sy
    db.node_orient = [ORIENTS[o] for o in orient]
    db.node_status = status
    db.pin_direct = [DIRECTS[v] for v in pdir]
    db.pin_offset_x, db.pin_offset_y = pox, poy
    db.net_names = _names(net_names, E)
    db.net_name2id_map = {k: i for i, k in enumerate(db.net_names)}
    db.flat_net2pin_map, db.flat_net2pin_start_map = fnp, fnps
    db.net2pin_map = [fnp[fnps[i]:fnps[i + 1]] for i in range(E)]
    db.flat_node2pin_map, db.flat_node2pin_start_map = fodp, fodps
    db.node2pin_map = [fodp[fodps[i]:fodps[i + 1]] for i in range(n)]
    db.pin2node_map, db.pin2net_map = p2n, p2e
    db.net_weights, db.net_ignore = w, ign
    db.rows = rows.reshape(R, 4)
    db.xl, db.yl, db.xh, db.yh, db.row_height, db.site_width = (float(v) for v in die)
    db.num_movable_nodes = n - nt - nio
    fixed = slice(db.num_movable_nodes, n - nio)
    clip = np.array([db.xl, db.yl, db.xh, db.yh])
    fxl, fyl = node_x[fixed].copy(), node_y[fixed].copy()
    fxh, fyh = fxl + sx[fixed], fyl + sy[fixed]
    union = L.dp_rect_union_area(fxl.size, p(fxl), p(fyl), p(fxh), p(fyh), p(clip))
    total = float(np.sum(sx[fixed] * sy[fixed]))
    db.total_space_area = (db.xh - db.xl) * (db.yh - db.yl) - min(union, total)
    return db


def num_bins(area, total_movable_area, num_movable, target_density, width, height):
    """PlaceDB.initialize_num_bins (dreamplace/PlaceDB.py:617-655): powers of
    two keeping the aspect ratio, 2 to 4096 per side, about one movable cell
    per bin at the target density."""
    ideal = max(area / (total_movable_area / num_movable / target_density), 4)
    aspect = height / width
    y_over_x = True
    if aspect < 1:
        aspect, y_over_x = 1.0 / aspect, False
    aspect = int(math.pow(2, round(math.log2(aspect))))
    nb = 2
    while nb <= 4096:
        found = nb * nb * aspect
        if ideal / 4 < found <= ideal:
            break
        nb *= 2
    return (nb, nb * aspect) if y_over_x else (nb * aspect, nb)


def to_design(db, target_density=1.0, num_bins_x=0, num_bins_y=0, dtype=np.float64,
              enable_fillers=True, ignore_net_degree=100, random_center_init=True, seed=1000,
              gamma=4.0):
    """The placer's design layout (as synth.make_design returns) for a
    Bookshelf design, per PlaceDB.initialize and BasicPlace's initial
    positions.  Coordinates are shifted so the row box starts at (0, 0) and
    scaled by 1/site_width."""
    n_phys = db.num_nodes
    nm, nt, nio = db.num_movable_nodes, db.num_terminals, db.num_terminal_NIs
    sf = 1.0 / db.site_width
    shx, shy = db.xl, db.yl
    x = (db.node_x - shx) * sf
    y = (db.node_y - shy) * sf
    sx, sy = db.node_size_x * sf, db.node_size_y * sf
    pox, poy = db.pin_offset_x * sf, db.pin_offset_y * sf
    xl, yl = 0.0, 0.0
    xh, yh = (db.xh - shx) * sf, (db.yh - shy) * sf
    row_height = db.row_height * sf
    space = db.total_space_area * sf * sf
    area = (xh - xl) * (yh - yl)
    mov_area = sx[:nm] * sy[:nm]
    total_movable = float(np.sum(mov_area))
    fixed = slice(nm, n_phys - nio)
    ox = np.maximum(np.minimum(x[fixed] + sx[fixed], xh) - np.maximum(x[fixed], xl), 0.0)
    oy = np.maximum(np.minimum(y[fixed] + sy[fixed], yh) - np.maximum(y[fixed], yl), 0.0)
    total_fixed = float(np.sum(ox * oy))
    macro = (mov_area > total_movable / max(nm, 1) * 10) & (sy[:nm] > row_height * 2)
    macro_area = float(mov_area[macro].sum())
    cell_area = total_movable - macro_area
    cell_space = space - macro_area
    t = max(float(target_density), min(cell_area / cell_space, 1.0))
    # fillers (PlaceDB.py:902-935)
    nf, fw, fh = 0, 0.0, row_height
    if enable_fillers:
        order = np.argsort(sx[:nm], kind="stable")
        lb, ub = int(nm * 0.05), int(nm * 0.95)
        fw = float(np.mean(sx[order[lb:ub]])) if lb < ub else 0.0
        if fw * fh > 0:
            nf = int(round(max(cell_space * t - cell_area, 0.0) / (fw * fh)))
    if num_bins_x <= 1 or num_bins_y <= 1:
        num_bins_x, num_bins_y = num_bins(area, total_movable, nm, t, xh - xl, yh - yl)
    bsx, bsy = (xh - xl) / num_bins_x, (yh - yl) / num_bins_y
    n = n_phys + nf
    pos_x = np.zeros(n)
    pos_y = np.zeros(n)
    pos_x[:n_phys], pos_y[:n_phys] = x, y
    rng = np.random.RandomState(seed)
    if random_center_init:
        pos_x[:nm] = rng.normal(loc=(xl + xh) / 2, scale=(xh - xl) * 0.001, size=nm)
        pos_y[:nm] = rng.normal(loc=(yl + yh) / 2, scale=(yh - yl) * 0.001, size=nm)
    if nf:
        pos_x[n_phys:] = rng.uniform(low=xl, high=xh - fw, size=nf)
        pos_y[n_phys:] = rng.uniform(low=yl, high=yh - fh, size=nf)
    degrees = np.diff(db.flat_net2pin_start_map)
    node_size_x = np.concatenate([sx, np.full(nf, fw)])
    node_size_y = np.concatenate([sy, np.full(nf, fh)])
    d = Design(
        config=-1, seed=seed, dtype=np.dtype(dtype),
        xl=xl, yl=yl, xh=xh, yh=yh, num_bins_x=int(num_bins_x), num_bins_y=int(num_bins_y),
        bin_size_x=bsx, bin_size_y=bsy, target_density=t,
        num_movable_nodes=nm, num_terminals=nt + nio, num_filler_nodes=nf,
        num_physical_nodes=n_phys, num_nodes=n, num_nets=db.num_nets, num_pins=db.num_pins,
        pos=np.concatenate([pos_x, pos_y]).astype(dtype),
        node_size_x=node_size_x.astype(dtype), node_size_y=node_size_y.astype(dtype),
        pin_offset_x=pox.astype(dtype), pin_offset_y=poy.astype(dtype),
        pin2node_map=db.pin2node_map.copy(), pin2net_map=db.pin2net_map.copy(),
        flat_net2pin_map=db.flat_net2pin_map.copy(),
        flat_net2pin_start_map=db.flat_net2pin_start_map.copy(),
        flat_node2pin_map=db.flat_node2pin_map.copy(),
        flat_node2pin_start_map=db.flat_node2pin_start_map.copy(),
        net_weights=db.net_weights.astype(dtype),
        net_mask_all=np.ones(db.num_nets, dtype=np.uint8),
        net_mask_ignore_large_degrees=((degrees >= 2) & (degrees < ignore_net_degree)).astype(np.uint8),
        pin_mask_ignore_fixed_macros=(db.pin2node_map >= nm),
        movable_macro_mask=macro,
        sorted_node_map=np.argsort(sx[:nm], kind="stable").astype(np.int32),
        total_movable_node_area=total_movable, total_fixed_node_area=total_fixed,
        total_space_area=space, filler_size_x=fw, filler_size_y=fh,
        wa_gamma=gamma * bsx, node_names=db.node_names,
    )
    d.bin_center_x = _bin_centers(xl, xh, bsx, int(num_bins_x)).astype(dtype)
    d.bin_center_y = _bin_centers(yl, yh, bsy, int(num_bins_y)).astype(dtype)
    d.shift = (shx, shy)
    d.scale = sf
    return d


def write_pl(path, db, node_x, node_y):
    """PlaceDB.write_pl (dreamplace/PlaceDB.py:1039-1074): movable nodes at
    the given positions, fixed nodes at their read positions with /FIXED,
    IO pins with /FIXED_NI; positions in the file's units."""
    nm, nt, nio = db.num_movable_nodes, db.num_terminals, db.num_terminal_NIs
    out = ["UCLA pl 1.0\n"]
    for i in range(nm):
        out.append("\n%s %g %g : %s" % (db.node_names[i], node_x[i], node_y[i], db.node_orient[i]))
    for i in range(nm, nm + nt):
        out.append("\n%s %g %g : %s /FIXED" % (db.node_names[i], float(db.node_x[i]), float(db.node_y[i]),
                                               "N"))
    for i in range(nm + nt, nm + nt + nio):
        out.append("\n%s %g %g : %s /FIXED_NI" % (db.node_names[i], node_x[i], node_y[i],
                                                  db.node_orient[i]))
    with open(path, "w") as f:
        f.write("".join(out))


def positions_in_file_units(d, pos):
    """Positions of a to_design() design mapped back to the .pl file's
    coordinates (undo the shift and scale)."""
    n = d["num_physical_nodes"]
    nn = d["num_nodes"]
    x = np.asarray(pos[:n], dtype=np.float64) / d["scale"] + d["shift"][0]
    y = np.asarray(pos[nn:nn + n], dtype=np.float64) / d["scale"] + d["shift"][1]
    return x, y
