"""Pins the oracle's filter (REF and STD modes).

Reference: okano_3_0/BilateralFilter.cu:71-199 (kernel), :312-320 (trim to
common size), :18/:61-64 (clamp-to-edge over the input ROI).  Checks:
  * brute force: an independent per-pixel Python statement, bit-exact, on
    tiny images with offset ROIs and trimmed outputs;
  * FP64 bounds: every u8 lies within the truncation of the real-valued
    result computed in double;
  * invariants: constant images, window min/max, ROI isolation;
  * special case: sigma_range -> infinity is a normalised Gaussian blur
    (numpy convolution in float64);
  * REF/STD relations (row aliasing);
  * planted mistakes (oracle mutations) each fail at least one check.
"""
import numpy as np
import pytest

import oracle
from tests import _brute


def _noisy(h, w, seed, lo=0, hi=255):
    rng = np.random.default_rng(seed)
    base = np.linspace(20, 220, w)[None, :].repeat(h, 0)
    img = base + rng.normal(0, 25, (h, w))
    img[h // 3: h // 2, w // 4: w // 2] = 250
    img[h // 2:, : w // 3] = 5
    return np.clip(np.rint(img), lo, hi).astype(np.uint8)


CASES = [
    # (h, w, roi, out_size, radius, ss, sr)
    (13, 11, None, None, 1, 1.0, 20.0),
    (14, 12, (1, 2, 11, 13), None, 2, 1.5, 30.0),
    (16, 15, (2, 1, 14, 15), (10, 11), 3, 3.0, 30.0),
    (11, 17, (0, 0, 17, 11), (17, 9), 3, 2.0, 10.0),
    (18, 13, (3, 3, 12, 17), None, 4, 2.5, 45.0),
]


@pytest.mark.parametrize("mode", [oracle.REF, oracle.STD])
@pytest.mark.parametrize("case", CASES)
def test_bit_exact_vs_brute_force(case, mode):
    h, w, roi, out_size, r, ss, sr = case
    img = _noisy(h, w, seed=h * 31 + w)
    roi = roi or (0, 0, w, h)
    got, gq = oracle.filter_q(img, r, ss, sr, mode=mode, roi=roi, out_size=out_size)
    exp, eq = _brute.filter_pixelwise(img, roi, out_size, r, ss, sr, mode, return_q=True)
    assert got.shape == exp.shape
    assert np.array_equal(got, exp)
    assert np.array_equal(gq.view(np.uint32), eq.view(np.uint32))   # every rounding step


@pytest.mark.parametrize("mode", [oracle.REF, oracle.STD])
@pytest.mark.parametrize("r,ss,sr", [(1, 1.0, 20.0), (3, 3.0, 30.0), (7, 5.0, 25.0), (15, 10.0, 30.0)])
def test_fp64_bounds(mode, r, ss, sr):
    img = _noisy(64, 48, seed=r)
    roi = (3, 5, 45, 61)
    o = oracle.filter(img, r, ss, sr, mode=mode, roi=roi)
    v = oracle.filter_f64(img, r, ss, sr, mode=mode, roi=roi)
    d = 0.05   # |fl32(t/s) - t64/s64| < 0.05 for up to 962 taps (relative ~1e-4)
    assert np.all(o >= np.floor(v - d)) and np.all(o <= np.floor(v + d))
    err = o - v
    # truncation biases the mean by about -0.5 LSB
    assert -0.75 < err.mean() < -0.25
    assert np.abs(err).max() < 1.0 + d


@pytest.mark.parametrize("mode", [oracle.REF, oracle.STD])
def test_constant_image(mode):
    for c in (0, 1, 17, 100, 128, 200, 254, 255):
        img = np.full((20, 24), c, np.uint8)
        for r, ss, sr in [(3, 3.0, 30.0), (7, 5.0, 25.0), (15, 10.0, 30.0)]:
            o = oracle.filter(img, r, ss, sr, mode=mode)
            assert set(np.unique(o)) <= {max(c - 1, 0), c}


def test_window_min_max_ref():
    img = _noisy(40, 40, seed=7)
    r = 3
    o = oracle.filter(img, r, 3.0, 30.0, mode=oracle.REF)
    pad = np.pad(img, r + 1, mode="edge").astype(int)
    for y in range(40):
        y0 = y - y % 4
        for x in range(0, 40, 3):
            win = pad[y0 + 1: y0 + 2 * r + 3, x + 1: x + 2 * r + 2]
            assert win.min() - 1 <= o[y, x] <= win.max()


def test_input_outside_roi_never_read():
    img = _noisy(30, 30, seed=3)
    roi = (5, 6, 25, 24)
    a = oracle.filter(img, 4, 2.0, 30.0, roi=roi)
    img2 = img.copy()
    mask = np.ones_like(img2, bool)
    mask[6:24, 5:25] = False
    img2[mask] = 255 - img2[mask]
    b = oracle.filter(img2, 4, 2.0, 30.0, roi=roi)
    assert np.array_equal(a, b)


@pytest.mark.parametrize("mode", [oracle.REF, oracle.STD])
def test_huge_sigma_range_is_gaussian_blur(mode):
    """E == 1 exactly, so the filter is sum(g p) / sum(g) over the taps:
    compare with a float64 convolution (edge padding) of the same taps."""
    r, ss = 3, 2.0
    img = _noisy(24, 28, seed=11)
    o = oracle.filter(img, r, ss, 1.0e7, mode=mode)
    taps = _brute.ring_offsets(r + 1)[: (2 * r + 1) ** 2 + (1 if mode == oracle.REF else 0)]
    pad = np.pad(img.astype(np.float64), r + 2, mode="edge")
    num = np.zeros(img.shape)
    den = 0.0
    H, W = img.shape
    rows = np.arange(H) - (np.arange(H) % 4 if mode == oracle.REF else 0)
    for dy, dx in taps:
        g = np.exp(-(dy * dy + dx * dx) / (2 * ss * ss))
        num += g * pad[rows[:, None] + dy + r + 2, np.arange(W)[None, :] + dx + r + 2]
        den += g
    v = num / den
    assert np.all(np.abs(o - np.floor(v)) <= 1)
    assert np.mean(o == np.floor(v)) > 0.95


def test_ref_rows_alias_first_row_of_group():
    """REF row y uses the window of row y - y%4 with its own centre: on an
    image whose rows are constant and equal within each 4-row group, REF ==
    STD shifted by the in-group offset except the extra tap."""
    img = _noisy(32, 32, seed=5)
    ref = oracle.filter(img, 3, 3.0, 30.0, mode=oracle.REF)
    # Row 4k of REF and STD differ only by the extra tap; rows 4k+1.. differ more.
    std = oracle.filter(img, 3, 3.0, 30.0, mode=oracle.STD)
    d0 = np.abs(ref[0::4].astype(int) - std[0::4]).mean()
    d1 = np.abs(ref[1::4].astype(int) - std[1::4]).mean()
    assert d0 < d1
    # Row group images constant along y: REF row 4k+m equals the STD result
    # of a pixel with the same window and centre, which the brute force pins.
    col = _noisy(1, 20, seed=9)
    img2 = np.repeat(col, 16, axis=0)
    a = oracle.filter(img2, 2, 1.5, 20.0, mode=oracle.REF)
    assert np.all(a == a[0:1])   # every row identical


def test_trimmed_size_and_untouched_pixels():
    img = _noisy(20, 20, seed=2)
    o = oracle.filter(img, 2, 1.0, 30.0, roi=(2, 2, 18, 18), out_size=(10, 7))
    assert o.shape == (7, 10)


MUTATIONS = [oracle.MUT_DROP_EXTRA_TAP, oracle.MUT_SWAP_DYDX, oracle.MUT_OWN_ROW_WINDOW,
             oracle.MUT_ROUND_NEAREST, oracle.MUT_FUSED_MULADD, oracle.MUT_CENTRE_ROW0,
             oracle.MUT_CLAMP_TRIMMED]


@pytest.mark.parametrize("mutation", MUTATIONS)
def test_planted_mistakes_are_detected(mutation):
    """Each mutation makes the oracle disagree with the brute force (u8
    output or the float quotient before truncation)."""
    h, w = 17, 19
    img = _noisy(h, w, seed=21)
    roi, out_size, r, ss, sr = (1, 1, 18, 16), (13, 12), 3, 2.0, 25.0
    exp, eq = _brute.filter_pixelwise(img, roi, out_size, r, ss, sr, 0, return_q=True)
    try:
        oracle.set_mutation(mutation)
        got, gq = oracle.filter_q(img, r, ss, sr, mode=oracle.REF, roi=roi, out_size=out_size)
    finally:
        oracle.set_mutation(oracle.MUT_NONE)
    assert not (np.array_equal(got, exp) and np.array_equal(gq.view(np.uint32), eq.view(np.uint32)))
    # and the faithful oracle agrees again
    got, gq = oracle.filter_q(img, r, ss, sr, mode=oracle.REF, roi=roi, out_size=out_size)
    assert np.array_equal(got, exp) and np.array_equal(gq.view(np.uint32), eq.view(np.uint32))


def test_threads_do_not_change_results():
    img = _noisy(70, 66, seed=4)
    a = oracle.filter(img, 5, 3.0, 20.0, nthreads=1)
    b = oracle.filter(img, 5, 3.0, 20.0, nthreads=4)
    assert np.array_equal(a, b)


@pytest.mark.parametrize("mode", [oracle.REF, oracle.STD])
def test_row_band_equals_full(mode):
    img = _noisy(53, 47, seed=12)
    full = oracle.filter(img, 4, 2.0, 30.0, mode=mode, roi=(2, 1, 45, 52))
    for y0, y1 in [(0, 4), (8, 30), (12, 51), (48, 51)]:
        band = oracle.filter(img, 4, 2.0, 30.0, mode=mode, roi=(2, 1, 45, 52), rows=(y0, y1))
        assert np.array_equal(band, full[y0:y1])
    v = oracle.filter_f64(img, 4, 2.0, 30.0, mode=mode)
    vb = oracle.filter_f64(img, 4, 2.0, 30.0, mode=mode, rows=(20, 33))
    assert np.array_equal(v[20:33], vb)
