"""Bookshelf reader: the reference's place_io fixture and golden
(unittest/ops/place_io_unittest/simple/, simple.golden, data files kept
under tests/golden/bookshelf_simple/), order invariants, the space area,
conversion to the placer's layout and the .pl round trip.  Host only."""
import ast
import os

import numpy as np
import pytest

from dreamplace_amd import bookshelf as B

HERE = os.path.dirname(os.path.abspath(__file__))
SIMPLE = os.path.join(HERE, "golden", "bookshelf_simple")


def _golden():
    """simple.golden lines `key = value`; values are Python-like lists with
    unquoted names."""
    out = {}
    with open(os.path.join(SIMPLE, "simple.golden")) as f:
        for line in f:
            if " = " not in line:
                continue
            k, v = line.rstrip("\n").split(" = ", 1)
            out[k] = v
    return out


def _names(v):
    return [s.strip() for s in v.strip()[1:-1].split(",")]


def _name_map(v):
    items = [s.split(":") for s in v.strip()[1:-1].split(",")]
    return {a.strip(): int(b) for a, b in items}


@pytest.fixture(scope="module")
def db():
    return B.read(os.path.join(SIMPLE, "simple.aux"))


def test_simple_matches_reference_golden(db):
    g = _golden()
    assert db.num_nodes == int(g["num_nodes"])
    assert db.num_terminals == int(g["num_terminals"])
    assert db.node_name2id_map == _name_map(g["node_name2id_map"])
    assert db.node_names == _names(g["node_names"])
    assert db.node_orient == _names(g["node_orient"])
    assert db.pin_direct == _names(g["pin_direct"])
    assert db.net_name2id_map == _name_map(g["net_name2id_map"])
    assert db.net_names == _names(g["net_names"])
    for k in ("node_x", "node_y", "node_size_x", "node_size_y", "pin_offset_x", "pin_offset_y",
              "flat_net2pin_map", "flat_net2pin_start_map", "flat_node2pin_map",
              "flat_node2pin_start_map", "pin2node_map", "pin2net_map"):
        assert np.array_equal(np.asarray(getattr(db, k)), np.array(ast.literal_eval(g[k]))), k
    assert [list(v) for v in db.net2pin_map] == ast.literal_eval(g["net2pin_map"])
    assert [list(v) for v in db.node2pin_map] == ast.literal_eval(g["node2pin_map"])
    assert [tuple(r) for r in db.rows] == [tuple(float(x) for x in r) for r in ast.literal_eval(g["rows"])]
    for k in ("xl", "yl", "xh", "yh", "row_height", "site_width"):
        assert getattr(db, k) == float(g[k]), k
    assert db.num_movable_pins == int(g["num_movable_pins"])


def test_order_invariants(db):
    """Nets ascend in degree; pins are grouped by net in id order; every
    node's pin list and every net's pin list name each pin once."""
    deg = np.diff(db.flat_net2pin_start_map)
    assert np.all(deg[1:] >= deg[:-1])
    assert np.all(np.diff(db.pin2net_map) >= 0)
    assert sorted(db.flat_net2pin_map.tolist()) == list(range(db.num_pins))
    assert sorted(db.flat_node2pin_map.tolist()) == list(range(db.num_pins))
    for e in range(db.num_nets):
        pins = db.net2pin_map[e]
        assert np.all(db.pin2net_map[pins] == e)
        # one pin per node, an output pin (if any) first
        assert len(set(db.pin2node_map[pins].tolist())) == len(pins)
        dirs = [db.pin_direct[p] for p in pins]
        if "OUTPUT" in dirs:
            assert dirs[0] == "OUTPUT"
    for i in range(db.num_nodes):
        assert np.all(db.pin2node_map[db.node2pin_map[i]] == i)


def test_space_area_closed_form(db):
    """Die 96 x 96 = 9216; fixed o8 (459,470)-(499,510) lies inside (1600),
    o9 (500,555)-(548,579) lies above the die (0 inside): 9216 - 1600."""
    assert db.total_space_area == 9216.0 - 1600.0


def test_rect_union_area_brute_force():
    """Union area against a unit-grid count on random integer rectangles."""
    from dreamplace_amd import _native as N
    rng = np.random.default_rng(0)
    for _ in range(20):
        k = int(rng.integers(1, 12))
        a = rng.integers(0, 30, size=(k, 2)).astype(np.float64)
        s = rng.integers(1, 15, size=(k, 2)).astype(np.float64)
        xl, yl = a[:, 0].copy(), a[:, 1].copy()
        xh, yh = xl + s[:, 0], yl + s[:, 1]
        clip = np.array([3.0, 2.0, 35.0, 33.0])
        grid = np.zeros((50, 50), dtype=bool)
        for i in range(k):
            grid[int(xl[i]):int(xh[i]), int(yl[i]):int(yh[i])] = True
        want = grid[3:35, 2:33].sum()
        got = N.lib().dp_rect_union_area(k, N.ptr_np(xl), N.ptr_np(yl), N.ptr_np(xh), N.ptr_np(yh),
                                         N.ptr_np(clip))
        assert got == want


def test_unsorted_nets_keep_file_order():
    db = B.read(os.path.join(SIMPLE, "simple.aux"), sort_nets_by_degree=False)
    assert db.net_names == ["n0", "n1", "n2", "n3", "n4", "n5", "n6", "n7"]


def test_to_design_layout(db):
    """Shifted to the origin; fillers by the reference formula; positions in
    the die; maps carried over."""
    d = B.to_design(db, target_density=1.0, num_bins_x=8, num_bins_y=8)
    assert (d.xl, d.yl, d.xh, d.yh) == (0.0, 0.0, 96.0, 96.0)
    nm = d.num_movable_nodes
    assert nm == 8 and d.num_terminals == 2
    cell_area = float(np.sum(d.node_size_x[:nm] * d.node_size_y[:nm]))
    # filler width: mean movable width between the 5th and 95th percentile ranks
    assert d.filler_size_x == pytest.approx(np.mean(np.sort(db.node_size_x[:8])[0:7]))
    want = int(round(max(d.total_space_area * d.target_density - cell_area, 0) /
                     (d.filler_size_x * d.filler_size_y)))
    assert d.num_filler_nodes == want
    n = d.num_nodes
    fill = slice(d.num_physical_nodes, n)
    assert np.all(d.pos[fill] >= 0) and np.all(d.pos[fill] <= 96 - d.filler_size_x)
    assert np.array_equal(d.pos[nm:d.num_physical_nodes], db.node_x[nm:] - 459)
    assert np.array_equal(d.flat_net2pin_map, db.flat_net2pin_map)


def test_pl_round_trip(db, tmp_path):
    """write_pl, then reading the design with that .pl, gives the same
    positions (movable nodes moved, fixed nodes kept)."""
    x = db.node_x.copy()
    y = db.node_y.copy()
    x[:8] = np.arange(8) * 10 + 460
    y[:8] = np.arange(8) * 7 + 461
    for name in ("simple.nodes", "simple.nets", "simple.scl"):
        (tmp_path / name).write_text(open(os.path.join(SIMPLE, name)).read())
    B.write_pl(str(tmp_path / "simple.pl"), db, x, y)
    (tmp_path / "simple.aux").write_text("RowBasedPlacement : simple.nodes simple.nets simple.pl simple.scl\n")
    db2 = B.read(str(tmp_path / "simple.aux"))
    assert db2.node_names == db.node_names
    assert np.array_equal(db2.node_x, x) and np.array_equal(db2.node_y, y)
    assert np.array_equal(db2.flat_net2pin_map, db.flat_net2pin_map)


def test_missing_file_raises(tmp_path):
    (tmp_path / "x.aux").write_text("RowBasedPlacement : x.nodes x.nets x.pl x.scl\n")
    with pytest.raises(RuntimeError, match="cannot read"):
        B.read(str(tmp_path / "x.aux"))


ORIENT_CASES = {
    # orientation: (size in N, lower-left offset of the pin in N), derived by
    # hand from PyPlaceDB.cpp:545-740 for a 6x4 node whose pin sits at (+2, -1)
    # from its centre, i.e. (5, 1) from its lower-left corner in the file
    "N": ((6, 4), (5, 1)),
    "S": ((6, 4), (1, 3)),    # rotate 180: (w - x, h - y)
    "W": ((4, 6), (3, 5)),    # rotate 270: (h - y, x), sizes swapped
    "E": ((4, 6), (1, 1)),    # rotate 90: (y, w - x), sizes swapped
    "FN": ((6, 4), (1, 1)),   # mirror: (w - x, y)
    "FS": ((6, 4), (5, 3)),   # rotate 180, then mirror
    "FW": ((4, 6), (1, 5)),   # rotate 270, then mirror with the new width 4
    "FE": ((4, 6), (3, 1)),   # rotate 90, then mirror with the new width 4
}


def _orient_design(tmp_path):
    names = list(ORIENT_CASES)
    nodes = ["UCLA nodes 1.0", "NumNodes : %d" % (len(names) + 1),
             "NumTerminals : %d" % len(names), "c0 2 4"]
    nodes += ["f_%s 6 4 terminal" % o for o in names]
    nets = ["UCLA nets 1.0", "NumNets : %d" % len(names), "NumPins : %d" % (2 * len(names))]
    for k, o in enumerate(names):
        nets += ["NetDegree : 2 n%d" % k, "c0 I : 0 0", "f_%s O : 2 -1" % o]
    pl = ["UCLA pl 1.0", "c0 40 40 : FS"]
    pl += ["f_%s %d %d : %s /FIXED" % (o, 10 * k, 80, o) for k, o in enumerate(names)]
    scl = ["UCLA scl 1.0", "NumRows : 25"]
    for r in range(25):
        scl += ["CoreRow Horizontal", " Coordinate : %d" % (4 * r), " Height : 4",
                " Sitewidth : 1", " Sitespacing : 1", " Siteorient : 1", " Sitesymmetry : 1",
                " SubrowOrigin : 0 NumSites : 100", "End"]
    for ext, lines in (("nodes", nodes), ("nets", nets), ("pl", pl), ("scl", scl)):
        (tmp_path / ("o.%s" % ext)).write_text("\n".join(lines) + "\n")
    (tmp_path / "o.aux").write_text("RowBasedPlacement : o.nodes o.nets o.pl o.scl\n")
    return B.read(str(tmp_path / "o.aux"))


def test_fixed_node_orientations(tmp_path):
    """Fixed nodes placed in each of the eight orientations: sizes and pin
    offsets restated for N (PyPlaceDB::convertOrient); the recorded
    orientation is kept; a movable node's orientation changes nothing; every
    pin stays inside its (rotated) node."""
    db = _orient_design(tmp_path)
    for o, (size, off) in ORIENT_CASES.items():
        i = db.node_name2id_map["f_" + o]
        assert db.node_orient[i] == o
        assert (db.node_size_x[i], db.node_size_y[i]) == size, o
        (p,) = db.node2pin_map[i]
        assert (db.pin_offset_x[p], db.pin_offset_y[p]) == off, o
        assert 0 <= off[0] <= size[0] and 0 <= off[1] <= size[1]
    c = db.node_name2id_map["c0"]
    assert db.node_orient[c] == "FS"
    assert (db.node_size_x[c], db.node_size_y[c]) == (2, 4)
    assert all((db.pin_offset_x[p], db.pin_offset_y[p]) == (1, 2) for p in db.node2pin_map[c])
    # the space area uses the rotated sizes: 4 nodes of 6x4 and 4 of 4x6, apart
    assert db.total_space_area == pytest.approx(100 * 100 - 8 * 24)
