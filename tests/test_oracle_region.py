"""The oracle of the stages around the filter (oracle/region_oracle.c),
pinned without running the reference:

* ConnectRegion against scipy.ndimage.label (8-connectivity) plus an
  independent numpy statement of the reference's numbering: every pixel
  outside [10, 250] is its own region, regions are numbered in raster order
  of their first pixel after the area filter, areas count only the rows the
  reference's area launch covers, numbers >= 255 map to 0.
* LocalCluster on closed forms: pixels in the guard band or outside
  (proBlack, proWhite) are copied; constant images are fixed points; on a
  horizontal ramp with open thresholds every directional average is known
  exactly, which fixes the walk's step pattern, the divisor and the
  replacement's tie-break.
"""
import numpy as np
import pytest
from scipy import ndimage

import oracle


def _expected_regions(img, min_area, max_area, mingray=10, maxgray=250):
    H, W = img.shape
    fg = (img >= mingray) & (img <= maxgray)
    lab, n = ndimage.label(fg, structure=np.ones((3, 3), int))
    # region id per pixel: fg components 1..n, every other pixel its own id
    rid = np.where(fg, lab, 0).astype(np.int64)
    bg = ~fg
    rid[bg] = n + 1 + np.arange(int(bg.sum()))
    flat = rid.reshape(-1)
    first = {}
    for i, r in enumerate(flat):
        first.setdefault(int(r), i)
    packed = (H + 4) // 32
    covered = min(H, 256 * ((packed + 7) // 8))
    area = {}
    for r in flat[:covered * W]:
        area[int(r)] = area.get(int(r), 0) + 1
    kept = []
    for r, f in sorted(first.items(), key=lambda kv: kv[1]):
        a = area.get(r, 0)
        if a != 0 and (a < min_area or a > max_area):
            continue
        kept.append(r)
    num = {r: k for k, r in enumerate(kept)}
    out = np.zeros(H * W, np.uint8)
    for i, r in enumerate(flat):
        k = num.get(int(r), -1)
        out[i] = (k + 1) % 256 if 0 <= k < 256 else 0
    return out.reshape(H, W)


def _blobs(H, W, seed, n=30):
    rng = np.random.default_rng(seed)
    img = rng.integers(0, 10, (H, W)).astype(np.uint8)          # background below 10
    for _ in range(n):
        y, x = rng.integers(0, H), rng.integers(0, W)
        h, w = rng.integers(1, 12), rng.integers(1, 12)
        img[y:y + h, x:x + w] = rng.integers(10, 251)
    img[rng.random((H, W)) < 0.01] = 255                        # above the range
    return img


@pytest.mark.parametrize("shape,min_area,max_area", [
    ((40, 57), 1, 100000),
    ((64, 64), 5, 60),
    ((97, 33), 20, 100000),
])
def test_connect_region_against_scipy(shape, min_area, max_area):
    img = _blobs(*shape, seed=shape[0])
    got = oracle.connect_region(img, min_area=min_area, max_area=max_area)
    assert np.array_equal(got, _expected_regions(img, min_area, max_area))


def test_connect_region_area_rows_quirk():
    # H = 270: the reference's area launch covers rows 0..255 only
    H, W = 270, 20
    img = np.zeros((H, W), np.uint8)
    img[250:265, 2:6] = 100      # 60 px, 24 of them in counted rows
    img[258:268, 10:15] = 100    # 50 px, none counted: area 0, never dropped
    img[10:12, 10:12] = 100      # 4 px
    got = oracle.connect_region(img, min_area=30, max_area=1000)
    assert np.array_equal(got, _expected_regions(img, 30, 1000))
    assert got[250, 2] == 0                 # counted area 24 < 30: dropped
    # uncounted (area 0): kept -- as is every background pixel of rows >= 256
    blob = got[258:268, 10:15]
    assert blob[0, 0] != 0 and np.all(blob == blob[0, 0])
    assert got[256, 0] != 0 and got[255, 0] == 0   # background singletons: area 0 vs 1
    assert got[10, 10] == 0
    # the same picture 256 rows high counts everything
    small = img[14:].copy()
    g2 = oracle.connect_region(small, min_area=30, max_area=1000)
    assert g2[250 - 14, 2] != 0


def test_connect_region_numbers_wrap():
    # 400 isolated two-pixel regions (every other pixel is a one-pixel
    # background region, dropped by the area filter [2, 2])
    H, W = 40, 60
    img = np.zeros((H, W), np.uint8)
    img[0::2, 0::3] = 50
    img[0::2, 1::3] = 50
    got = oracle.connect_region(img, min_area=2, max_area=2)
    assert np.array_equal(got, _expected_regions(img, 2, 2))
    vals = got[0::2, 0::3].reshape(-1)      # raster order = number order
    assert len(vals) == 400
    assert list(vals[:255]) == list(range(1, 256))
    assert vals[255] == 0 and np.all(vals[256:] == 0)
    assert np.all(got[1::2] == 0)


def test_local_cluster_copies_guard_band_and_extremes():
    rng = np.random.default_rng(3)
    img = rng.integers(0, 256, (60, 70)).astype(np.uint8)
    pr = 8
    out = oracle.local_cluster(img, pntrange=pr)
    band = np.ones_like(img, bool)
    band[pr + 2:60 - pr - 2 + 1, pr + 2:70 - pr - 2 + 1] = False
    assert np.array_equal(out[band], img[band])
    ext = (img <= 10) | (img >= 250)
    assert np.array_equal(out[ext], img[ext])


@pytest.mark.parametrize("c", [11, 100, 249])
def test_local_cluster_constant_image_is_fixed(c):
    img = np.full((50, 50), c, np.uint8)
    for pc in (1, 5, 8):
        assert np.all(oracle.local_cluster(img, pntrange=8, pntcount=pc) == c)


def test_local_cluster_ramp_closed_form():
    # I(x, y) = x + 20.  Along direction dx the walk adds the pixels at 2, 2,
    # 4, 6, ..., 2(R-2) steps and divides by R = pntRange, so each average is
    # cv + dx * (2 + (R-1)(R-2)) / R = cv + 5.5 dx for R = 8 (exact in FP32).
    W, H = 120, 60
    img = np.tile((np.arange(W) + 20).astype(np.uint8), (H, 1))
    R = 8
    inner = np.s_[2 * R:H - 2 * R, 2 * R:W - 2 * R]       # walks stay inside
    out8 = oracle.local_cluster(img, pntrange=R, pntcount=8, gapthred=255, diffethred=255)
    assert np.array_equal(out8[inner], img[inner])
    # pntCount 7: the last direction at the maximal distance (z = 7, dx = +1)
    # is replaced: (8 cv - 5.5 - cv) / 7 = cv - 0.786, truncated to cv - 1
    out7 = oracle.local_cluster(img, pntrange=R, pntcount=7, gapthred=255, diffethred=255)
    assert np.array_equal(out7[inner], img[inner] - 1)


def test_local_cluster_walk_stops_at_an_edge():
    # a vertical step of 100 gray levels: a pixel left of it keeps its own
    # value in the +x directions (the first gap check fails), so with
    # pntCount 8 every walk average is either 60 (no step crossed) or 60.
    img = np.full((40, 40), 60, np.uint8)
    img[:, 25:] = 160
    out = oracle.local_cluster(img, pntrange=4, pntcount=8, gapthred=10, diffethred=10)
    assert out[20, 24] == 60
    assert out[20, 25] == 160
