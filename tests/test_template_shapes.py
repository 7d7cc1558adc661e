"""BOX, CIRCLE and ARC template vendors (okano_3_0/TemplateFactory.cu:340-402,
:501-666, :771-936): the oracle (oracle/tpl_oracle.c) is pinned by hand-traced
small cases and geometric properties, and the product generators
(bf_box_template, bf_circle_template, bf_arc_template; TemplateFactory pools
the same arrays) equal it bit for bit."""
import ctypes

import numpy as np
import pytest

import opencuda_amd as oc
import oracle


def _product(shape, *size):
    L = oc.lib()
    f = getattr(L, f"bf_{shape}_template")
    f.argtypes = [ctypes.c_int32] * len(size) + [ctypes.c_void_p, ctypes.c_void_p]
    f.restype = ctypes.c_int
    n = f(*size, None, None)
    assert n > 0
    pts = np.zeros((n, 2), np.int32)
    w = np.zeros(n, np.float32)
    assert f(*size, pts.ctypes.data, w.ctypes.data) == n
    return pts, w


# ---------------------------------------------------------------- BOX

@pytest.mark.parametrize("sx,sy", [(1, 1), (3, 3), (5, 3), (4, 6), (9, 2)])
def test_box_closed_form(sx, sy):
    pts, w = oracle.shape_template("box", sx, sy)
    n = sx * sy
    assert len(pts) == n
    box = [(c, r) for r in range(-((sy - 1) // 2), sy // 2)
           for c in range(-((sx - 1) // 2), sx // 2)]
    k = len(box)
    assert k == (sx - 1) * (sy - 1)
    assert [tuple(p) for p in pts[:k]] == box
    assert np.all(pts[k:] == 0) and np.all(w[k:] == 0)
    assert np.all(w[:k] == np.float32(1.0) / np.float32(n))
    p2, w2 = _product("box", sx, sy)
    assert np.array_equal(p2, pts) and np.array_equal(w2.view(np.uint32), w.view(np.uint32))


# -------------------------------------------------------------- CIRCLE

def test_circle_hand_traced():
    pts, w = oracle.shape_template("circle", 3)
    assert [tuple(p) for p in pts] == [(0, 1), (0, -1), (-1, 0), (0, 0), (1, 0)]
    pts, _ = oracle.shape_template("circle", 5)
    assert len(pts) == 21                       # 5 x 5 minus the 4 corners
    s = {tuple(p) for p in pts}
    assert s == {(x, y) for x in range(-2, 3) for y in range(-2, 3)
                 if (abs(x), abs(y)) != (2, 2)}


def _symmetric(s):
    return all((sx * x, sy * y) in s and (sx * y, sy * x) in s
               for (x, y) in s for sx in (1, -1) for sy in (1, -1))


@pytest.mark.parametrize("size", [7, 9, 15, 31, 41, 61])
def test_circle_properties(size):
    pts, w = oracle.shape_template("circle", size)
    r = (size - 1) // 2
    s = {tuple(p) for p in pts}
    assert len(s) == len(pts)                   # no point twice
    assert _symmetric(s)
    d2 = pts[:, 0].astype(int) ** 2 + pts[:, 1].astype(int) ** 2
    assert d2.max() <= r * r + r                # inside the midpoint boundary
    inner = {(x, y) for x in range(-r, r + 1) for y in range(-r, r + 1) if x * x + y * y <= r * r - r}
    assert inner <= s                           # filled
    for y in range(-r, r + 1):                  # every row one interval
        xs = sorted(x for (x, yy) in s if yy == y)
        assert xs == list(range(xs[0], xs[-1] + 1))
    assert np.all(w == np.float32(1.0) / np.float32(len(pts)))
    p2, w2 = _product("circle", size)
    assert np.array_equal(p2, pts) and np.array_equal(w2, w)


# ----------------------------------------------------------------- ARC

def test_arc_hand_traced():
    pts, _ = oracle.shape_template("arc", 3, 5)
    assert {tuple(p) for p in pts} == {(x, y) for x in (-1, 0, 1) for y in (-1, 0, 1)} - {(0, 0)}
    assert len(pts) == 8


@pytest.mark.parametrize("sx,sy", [(3, 7), (5, 11), (9, 21), (1, 9), (15, 41)])
def test_arc_properties(sx, sy):
    pts, w = oracle.shape_template("arc", sx, sy)
    sr = max(1, (sx - 1) // 2)
    lr = max((sy - 1) // 2, sr + 1)
    s = {tuple(p) for p in pts}
    assert len(s) == len(pts)
    assert _symmetric(s)
    d2 = pts[:, 0].astype(int) ** 2 + pts[:, 1].astype(int) ** 2
    assert d2.max() <= lr * lr + lr and d2.min() >= sr * sr - sr
    assert (0, 0) not in s
    # the ring's area is within a perimeter's width of pi (lr^2 - sr^2)
    assert abs(len(pts) - np.pi * (lr * lr - sr * sr)) <= 2 * np.pi * (lr + sr) + 8
    p2, w2 = _product("arc", sx, sy)
    assert np.array_equal(p2, pts) and np.array_equal(w2, w)


def test_illegal_sizes():
    L = oc.lib()
    assert L.bf_circle_template(8, None, None) < 0
    assert L.bf_arc_template(5, 3, None, None) < 0
    assert L.bf_arc_template(4, 9, None, None) < 0
    assert L.bf_box_template(0, 3, None, None) < 0
