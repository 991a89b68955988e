"""Write tests/golden/ops_golden.json and tests/golden/electric_potential_84.npz.

Fixtures are the input data held by the reference's unit tests; the expected
values are evaluated here from the golden definitions those tests state (not
from the oracle, and without importing the reference):

* WA wirelength: fixture of unittest/ops/weighted_average_wirelength_unittest.py:92-134
  (5 pins, 2 nets, weights [1, 2], gamma 0.5, ignore_net_degree 4); expected
  value = sum_n w_n * (sum x e^{x/g}/sum e^{x/g} - sum x e^{-x/g}/sum e^{-x/g}
  + the same in y), the un-shifted form of the test's golden (:38-86),
  evaluated in float64.
* HPWL: fixture of unittest/ops/hpwl_unittest.py:39-96; expected = sum_n w_n
  (max-min in x + max-min in y).
* pin_pos: fixture of unittest/ops/pin_pos_unittest.py:47-80; expected pin
  positions = offset + node position (the test's index_select golden), and the
  node gradient of a given pin gradient = its CSR sum.
* ElectricPotential: the 84-node input fixture of
  unittest/ops/electric_potential_unittest.py:43-110 (inputs only; the
  reference test holds no outputs).  Outputs for it are checked against
  oracle/_ref and the GPU path in the tests.

    python tests/golden/make_golden.py
"""
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def wa_golden():
    pin = np.array([[0.0, 0.0], [1.0, 2.0], [1.5, 0.2], [0.5, 3.1], [0.6, 1.1]], dtype=np.float64)
    nets = [[0, 4], [1, 2, 3]]
    w = np.array([1.0, 2.0])
    g = 0.5
    total = 0.0
    for n, pins in enumerate(nets):
        wl = 0.0
        for axis in (0, 1):
            v = pin[pins, axis]
            ep, en = np.exp(v / g), np.exp(-v / g)
            wl += (v * ep).sum() / ep.sum() - (v * en).sum() / en.sum()
        total += w[n] * wl
    return dict(pin_pos=pin.tolist(), net2pin=nets, net_weights=w.tolist(), gamma=g,
                ignore_net_degree=4, expected=total)


def hpwl_golden():
    pin = np.array([[0.0, 0.0], [1.0, 2.0], [1.5, 0.2], [0.5, 3.1], [0.6, 1.1]], dtype=np.float64)
    nets = [[0, 4], [1, 2, 3]]
    w = np.array([1.0, 2.0])
    total = 0.0
    for n, pins in enumerate(nets):
        total += w[n] * (np.ptp(pin[pins, 0]) + np.ptp(pin[pins, 1]))
    return dict(pin_pos=pin.tolist(), net2pin=nets, net_weights=w.tolist(), expected=total)


def pin_pos_golden():
    pos = np.array([[1, 10], [2, 20], [3, 30]], dtype=np.float64)
    node2pin = [[0, 4], [1, 2, 3]]
    offx = [0.1, 0.2, 0.3, 0.4, 0.5]
    offy = [0.01, 0.02, 0.03, 0.04, 0.05]
    pin2node = [0] * 5
    for nd, pins in enumerate(node2pin):
        for p in pins:
            pin2node[p] = nd
    px = [offx[p] + pos[pin2node[p], 0] for p in range(5)]
    py = [offy[p] + pos[pin2node[p], 1] for p in range(5)]
    gpin = [1.0, -2.0, 0.5, 4.0, 3.0, 0.25, 0.75, -1.0, 2.0, 1.5]
    gx = [sum(gpin[p] for p in pins) for pins in node2pin] + [0.0]
    gy = [sum(gpin[5 + p] for p in pins) for pins in node2pin] + [0.0]
    return dict(pos=pos.tolist(), node2pin=node2pin, pin_offset_x=offx, pin_offset_y=offy,
                pin2node=pin2node, num_physical_nodes=2, expected_pin_pos=px + py,
                grad_pin_pos=gpin, expected_grad_pos=gx + gy)


def electric_potential_84():
    xx = [1000, 11148, 11148, 11148, 11148, 11148, 11124, 11148, 11148, 11137, 11126, 11148, 11130,
          11148, 11148, 11148, 11148, 11148, 11148, 0, 11148, 11148, 11150, 11134, 11148, 11148, 11148,
          10550, 11148, 11148, 11144, 11148, 11148, 11148, 11148, 11140, 11120, 11154, 11148, 11133,
          11148, 11148, 11134, 11125, 11148, 11148, 11148, 11155, 11127, 11148, 11148, 11148, 11148,
          11131, 11148, 11148, 11148, 11148, 11136, 11148, 11146, 11148, 11135, 11148, 11125, 11150,
          11148, 11139, 11148, 11148, 11130, 11148, 11128, 11148, 11138, 11148, 11148, 11148, 11130,
          11148, 11132, 11148, 11148, 11090]
    yy = [1000, 11178, 11178, 11190, 11400, 11178, 11172, 11178, 11178, 11418, 11418, 11178, 11418,
          11178, 11178, 11178, 11178, 11178, 11178, 11414, 11178, 11178, 11172, 11418, 11406, 11184,
          11178, 10398, 11178, 11178, 11172, 11178, 11178, 11178, 11178, 11418, 11418, 11172, 11178,
          11418, 11178, 11178, 11172, 11418, 11178, 11178, 11178, 11418, 11418, 11178, 11178, 11178,
          11178, 11418, 11178, 11178, 11394, 11178, 11418, 11178, 11418, 11178, 11418, 11178, 11418,
          11418, 11178, 11172, 11178, 11178, 11418, 11178, 11418, 11178, 11418, 11412, 11178, 11178,
          11172, 11178, 11418, 11178, 11178, 11414]
    sx = [6, 3, 3, 3, 3, 3, 5, 3, 3, 1, 1, 3, 1, 3, 3, 3, 3, 3, 3, 16728, 3, 3, 5, 1, 3, 3, 3, 740,
          3, 3, 5, 3, 3, 3, 3, 5, 5, 5, 3, 1, 3, 3, 5, 1, 3, 3, 3, 5, 1, 3, 3, 3, 3, 1, 3, 3, 3, 3, 5,
          3, 5, 3, 1, 3, 5, 5, 3, 5, 3, 3, 5, 3, 1, 3, 1, 3, 3, 3, 5, 3, 1, 3, 3, 67]
    sy = [6, 240, 240, 6, 6, 240, 6, 240, 240, 6, 6, 240, 6, 240, 240, 240, 240, 240, 240, 10, 240, 6,
          6, 6, 6, 6, 240, 780, 240, 240, 6, 240, 240, 240, 240, 6, 6, 6, 240, 6, 240, 240, 6, 6, 240,
          240, 240, 6, 6, 240, 240, 240, 240, 6, 240, 240, 6, 240, 6, 240, 6, 240, 6, 240, 6, 6, 240,
          6, 240, 240, 6, 240, 6, 240, 6, 6, 240, 240, 6, 240, 6, 240, 240, 10]
    return dict(x=np.array(xx, np.float64), y=np.array(yy, np.float64),
                node_size_x=np.array(sx, np.float64), node_size_y=np.array(sy, np.float64),
                num_movable_nodes=np.int64(1), xl=np.float64(0.0), yl=np.float64(6.0),
                xh=np.float64(16728.0), yh=np.float64(11430.0), target_density=np.float64(0.7),
                num_bins_x=np.int64(1024), num_bins_y=np.int64(1024))


def main():
    out = dict(weighted_average_wirelength=wa_golden(), hpwl=hpwl_golden(), pin_pos=pin_pos_golden())
    with open(os.path.join(HERE, "ops_golden.json"), "w") as f:
        json.dump(out, f, indent=1)
    np.savez(os.path.join(HERE, "electric_potential_84.npz"), **electric_potential_84())
    print("wrote ops_golden.json and electric_potential_84.npz")


if __name__ == "__main__":
    main()
