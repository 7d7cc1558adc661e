"""The union and area rules of the ConnectRegion kernels
(opencuda_amd/csrc/region.hip), restated on the CPU and checked against
independent references on random images:

* tile pass + border pass: a run of foreground pixels (8-connectivity) is
  united with a neighbouring line once per run it touches there -- through
  the straight neighbour at the first of its pixels under each segment,
  through the diagonal at its ends when the straight neighbour is
  background; inside 64x64 tiles along rows, then along every tile edge
  (top row against the row above, left column against the column to the
  left, lines extended by the corner diagonals).  The components and their
  minimum-index roots must equal scipy.ndimage.label's;
* area pass: a wave walks a row in 64-pixel chunks, the first lane of a run
  inside a chunk adds it, the run reaching the chunk's end is carried, and
  the row's first run of >= 32 pixels names a label whose later runs are
  summed before one add.  The areas must equal a plain count.

The GPU kernels themselves are pinned against the oracle by
tests/test_gpu_region.py; this file pins the rules they implement."""
import numpy as np
import pytest
from scipy import ndimage

T = 64


def _find(L, x):
    while L[x] != x:
        x = L[x]
    return x


def _unite(L, a, b):
    a, b = _find(L, a), _find(L, b)
    if a != b:
        if a < b:
            L[b] = a
        else:
            L[a] = b


def _run_rule(L, idx, io, step, f, o, ob, oa, start, end):
    if not f:
        return
    if o and (start or not ob):
        _unite(L, idx, io)
    if not o and start and ob:
        _unite(L, idx, io - step)
    if not o and end and oa:
        _unite(L, idx, io + step)


def _label(fg):
    H, W = fg.shape
    L = list(range(H * W))

    def f(x, y):
        return 0 <= x < W and 0 <= y < H and bool(fg[y, x])

    for ty in range(0, H, T):
        for tx in range(0, W, T):
            x1, y1 = min(tx + T, W), min(ty + T, H)
            for y in range(ty, y1):                  # runs linked to their first pixel
                for x in range(tx, x1):
                    if fg[y, x]:
                        s = x
                        while s > tx and fg[y, s - 1]:
                            s -= 1
                        L[y * W + x] = y * W + s
            for y in range(ty + 1, y1):              # rows inside the tile
                for x in range(tx, x1):
                    _run_rule(L, y * W + x, (y - 1) * W + x, 1, fg[y, x], fg[y - 1, x],
                              x > tx and fg[y - 1, x - 1], x < x1 - 1 and fg[y - 1, x + 1],
                              x == tx or not fg[y, x - 1], x == x1 - 1 or not fg[y, x + 1])
    for ty in range(0, H, T):                        # tile edges
        for tx in range(0, W, T):
            for top in (True, False):
                if (ty == 0) if top else (tx == 0):
                    continue
                dx, dy = (1, 0) if top else (0, 1)
                for lane in range(T):
                    x, y = (tx + lane, ty) if top else (tx, ty + lane)
                    xo, yo = (x, y - 1) if top else (x - 1, y)
                    _run_rule(L, y * W + x, yo * W + xo, dy * W + dx, f(x, y), f(xo, yo),
                              f(xo - dx, yo - dy), f(xo + dx, yo + dy),
                              lane == 0 or not f(x - dx, y - dy),
                              lane == T - 1 or not f(x + dx, y + dy))
    return np.array([_find(L, i) for i in range(H * W)]).reshape(H, W)


@pytest.mark.parametrize("seed", range(4))
def test_union_rules_match_scipy(seed):
    rng = np.random.default_rng(seed)
    H, W = int(rng.integers(70, 150)), int(rng.integers(70, 150))
    fg = rng.random((H, W)) < rng.uniform(0.35, 0.65)
    got = _label(fg).reshape(-1)
    lab, n = ndimage.label(fg, structure=np.ones((3, 3), int))
    flat = lab.reshape(-1)
    for c in range(1, n + 1):
        idx = np.flatnonzero(flat == c)
        assert np.all(got[idx] == idx.min())
    bg = np.flatnonzero(flat == 0)
    assert np.all(got[bg] == bg)


def _area_row(row, area):
    W = len(row)
    cur, cnt, hot, hot_cnt = -1, 0, -1, 0

    def flush(lab, c):
        nonlocal hot, hot_cnt
        if lab < 0 or c <= 0:
            return
        if lab == hot:
            hot_cnt += c
        elif hot < 0 and c >= 32:
            hot, hot_cnt = lab, c
        else:
            area[lab] += c

    for x0 in range(0, W, 64):
        n = min(64, W - x0)
        lab = [int(row[x0 + i]) if i < n else -1 for i in range(64)]
        prev = [cur] + lab[:63]
        starts = [i for i in range(64) if i < n and lab[i] != prev[i]]
        for k, i in enumerate(starts[:-1]):
            ln = starts[k + 1] - i
            if lab[i] == hot:
                hot_cnt += ln
            else:
                area[lab[i]] += ln
        cnt += starts[0] if starts else n
        if starts:
            flush(cur, cnt)
            cur, cnt = lab[starts[-1]], n - starts[-1]
    flush(cur, cnt)
    if hot >= 0:
        area[hot] += hot_cnt


@pytest.mark.parametrize("seed", range(4))
def test_area_rule_matches_count(seed):
    rng = np.random.default_rng(100 + seed)
    H, W = int(rng.integers(1, 6)), int(rng.integers(1, 400))
    lab = rng.integers(0, 6, (H, W))
    lab = np.repeat(lab, rng.integers(1, 90), axis=1)[:, :W]
    area = np.zeros(6, int)
    for r in lab:
        _area_row(r, area)
    assert np.array_equal(area, np.bincount(lab.reshape(-1), minlength=6))
