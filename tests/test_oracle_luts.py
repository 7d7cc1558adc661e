"""Pins the oracle's (and the product's) weight tables.

Reference: okano_3_0/TemplateFactory.cu:1092-1168 (GAUSS, ring order and
weights) and :1323-1373 (EUCLIDE).  Checks: bit-exact against glibc expf
called directly, the closed-form ring enumeration, <= 1 ulp against exp()
in double, special cases, and the C-ABI generators of the product equal to
the oracle bit for bit.
"""
import math

import numpy as np
import pytest

import oracle
from tests import _brute


@pytest.mark.parametrize("sigma", [0.5, 1.0, 3.0, 5.0, 10.0, 6.6666665, 33.3])
def test_gauss_template_bits_against_glibc(sigma):
    off, w = oracle.gauss_template(sigma)
    assert off.shape == (1681, 2)
    assert w[0] == np.float32(1.0) and tuple(off[0]) == (0, 0)
    for n in range(1681):
        dy, dx = int(off[n, 0]), int(off[n, 1])
        assert w[n].view(np.uint32) == _brute.gauss_weight(dy, dx, sigma).view(np.uint32)


def test_ring_order_closed_form():
    off, _ = oracle.gauss_template(2.0)
    assert [tuple(p) for p in off] == _brute.ring_offsets(20)
    # The ring loop's last step (j = i) emits (i,i), (-i,i), (i,i), (i,-i):
    # corner (i,i) twice and (-i,-i) never (TemplateFactory.cu:1141-1164).
    # So the first (2r+1)^2 points are the square window with every
    # up-left corner (-i,-i) replaced by a second copy of (i,i); point
    # (2r+1)^2 is the first point of ring r+1, (r+1, -r).
    from collections import Counter
    for r in range(1, 20):
        G = (2 * r + 1) ** 2
        expect = Counter((a, b) for a in range(-r, r + 1) for b in range(-r, r + 1))
        for i in range(1, r + 1):
            expect[(-i, -i)] -= 1
            expect[(i, i)] += 1
        expect = +expect
        assert Counter(tuple(int(v) for v in p) for p in off[:G]) == expect
        assert tuple(off[G]) == (r + 1, -r)
    # Ring i holds 8i points
    rings = np.maximum(np.abs(off[:, 0]), np.abs(off[:, 1]))
    assert all((rings == i).sum() == 8 * i for i in range(1, 21))


@pytest.mark.parametrize("sigma", [1.0, 5.0, 10.0, 25.0, 30.0])
def test_gauss_within_one_ulp_of_double_exp(sigma):
    off, w = oracle.gauss_template(sigma)
    s22 = np.float32(np.float32(2 * np.float32(sigma)) * np.float32(sigma))
    for n in range(1, 1681, 7):
        q = int(off[n, 0]) ** 2 + int(off[n, 1]) ** 2
        arg = np.float32(-q) / s22
        ref = np.float32(math.exp(float(arg)))
        if ref == 0 and w[n] == 0:
            continue
        ulp = np.spacing(ref) if ref != 0 else np.float32(1.4e-45)
        assert abs(float(w[n]) - float(ref)) <= float(ulp)


@pytest.mark.parametrize("sigma", [10.0, 15.0, 25.0, 30.0, 45.0, 50.0])
def test_euclid_bits_against_glibc(sigma):
    E = oracle.euclid_template(sigma)
    assert E[0] == np.float32(1.0)
    for d in range(256):
        assert E[d].view(np.uint32) == _brute.range_weight(d, sigma).view(np.uint32)
    assert np.all(np.diff(E.astype(np.float64)) <= 0)


def test_subnormal_range_weights_are_kept():
    # sigma 10: exp(-d^2/200) is subnormal for d ~ 133..144 and zero beyond.
    E = oracle.euclid_template(10.0)
    tiny = np.finfo(np.float32).tiny
    sub = (E > 0) & (E < tiny)
    assert sub.any(), "expected subnormal weights"
    assert E[255] == 0.0


def test_huge_sigma_range_gives_unit_weights():
    E = oracle.euclid_template(1.0e7)
    assert np.all(E == np.float32(1.0))


def test_illegal_parameters():
    with pytest.raises(ValueError):
        oracle.gauss_template(0.0)
    with pytest.raises(ValueError):
        oracle.gauss_template(1.0, size=4)
    with pytest.raises(ValueError):
        oracle.euclid_template(0.0)


@pytest.mark.parametrize("ss,sr", [(3.0, 30.0), (5.0, 25.0), (10.0, 30.0), (2.0, 10.0), (6.0, 45.0)])
def test_product_generators_match_oracle(ss, sr):
    import opencuda_amd as oc
    off, w = oc.gauss_template(ss)
    off2, w2 = oracle.gauss_template(ss)
    assert np.array_equal(off, off2)
    assert np.array_equal(w.view(np.uint32), w2.view(np.uint32))
    E, E2 = oc.euclid_template(sr), oracle.euclid_template(sr)
    assert np.array_equal(E.view(np.uint32), E2.view(np.uint32))


def test_distinct_weight_rows_per_radius():
    """The kernels build one fused-table row per distinct spatial weight; at
    r=15 (REF, 962 taps) there are 121 rows."""
    off, w = oracle.gauss_template(10.0)
    expect = {3: 11, 7: 35, 9: 52, 15: 121}
    for r, rows in expect.items():
        G = (2 * r + 1) ** 2 + 1
        q = {int(a) ** 2 + int(b) ** 2 for a, b in off[:G]}
        assert len(q) == rows
        assert len({x.tobytes() for x in w[:G]}) <= rows
