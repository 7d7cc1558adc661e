"""Mixed workload (c4 subset): exact batch for small radii and surrogate +
sampled check + exact fallback for r = 15, captured in one graph.

* exact images: bit-exact to the oracle;
* surrogate images (spread over several streams in the graph): identical
  to a plain single-stream bf_sur_filter call, the tile count is every
  image's tiles, replays are identical to the eager run; error against the
  exact filter and the FP64 oracle is reported;
* routing: by a minimum image size, and by a one-off calibration of each
  image's fallback fraction (images above the threshold join the batch)."""
import numpy as np
import pytest

import opencuda_amd as oc
import oracle
from opencuda_amd import mixed
from opencuda_amd import surrogate as sg
from opencuda_amd import workloads as wl

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("sur_min_pixels", [0, 300_000])
def test_mixed_graph_subset(device, sur_min_pixels):
    import torch
    items = [it for it in wl.c4(60) if it.pixels <= 1_600_000][:16]
    assert {it.radius for it in items} >= {15}
    net = sg.Surrogate(seed=1)
    net.load(sg.DEFAULT_WEIGHTS)
    mx = mixed.Mixed(items, net, mode=oc.MODE_STD, streams=3, sur_min_pixels=sur_min_pixels)
    assert mx.sur_idx, "no image took the surrogate route"
    # images below the size threshold take the exact batch, whatever their radius
    small15 = [i for i, it in enumerate(items) if it.radius == 15 and it.pixels < sur_min_pixels]
    assert set(small15) <= set(mx.ex_idx)
    if sur_min_pixels:
        assert small15, "the threshold routes no r = 15 image to the batch"
    mx.capture()
    eager = [t.cpu().numpy() for t in mx.tout]
    for t in mx.tout:
        t.fill_(0)
    mx.reset_counts()
    mx.replay()
    tiles, fb = mx.counts()
    got = [t.cpu().numpy() for t in mx.tout]
    for a, b in zip(got, eager):
        assert np.array_equal(a, b)
    n_tiles = sum(((it.roi[2] - it.roi[0] + 63) // 64) * ((it.roi[3] - it.roi[1] + 63) // 64)
                  for it in (items[i] for i in mx.sur_idx))
    assert tiles == n_tiles and 0 <= fb <= tiles
    for i in mx.ex_idx:
        it = items[i]
        exp = oracle.filter(mx.tin[i].cpu().numpy(), it.radius, it.sigma_space, it.sigma_range,
                            mode=oracle.STD, roi=it.roi)
        assert np.array_equal(got[i], exp)
    worst, fp64 = 0, []
    for i in mx.sur_idx:
        it = items[i]
        ref = mx.exact_reference(i).cpu().numpy()
        d = np.abs(got[i].astype(int) - ref)
        worst = max(worst, int(d.max()))
        # the graph's multi-stream run equals a plain single-stream call
        single = torch.zeros_like(mx.tout[i])
        net.filter(oc.view(mx.tin[i], it.roi), oc.view(single),
                   mx.luts[(it.sigma_space, it.sigma_range)], it.radius, oc.MODE_STD,
                   it.sigma_range, 2)
        torch.cuda.synchronize()
        assert np.array_equal(got[i], single.cpu().numpy())
        h, w = ref.shape
        y0 = h // 2 // 4 * 4
        v = oracle.filter_f64(mx.tin[i].cpu().numpy(), it.radius, it.sigma_space, it.sigma_range,
                              mode=oracle.STD, roi=it.roi, rows=(y0, y0 + 4))
        fp64.append(float(np.abs(got[i][y0:y0 + 4] - v).max()))
    print(f"mixed subset: {len(mx.ex_idx)} exact + {len(mx.sur_idx)} surrogate images, "
          f"{fb}/{tiles} tiles fell back, max |err| vs exact {worst}, "
          f"max |err| vs FP64 on sampled rows {max(fp64):.3f}")


def test_mixed_calibration_routes_by_fallback(device):
    """max_fallback: each r >= 9 image is filtered once through the network
    at construction; those whose check sent more than the threshold of their
    tiles to the exact filter join the exact batch (and are then bit-exact
    to the oracle), the others keep the network route, whose counts in the
    graph equal the calibration's (the sampled check is deterministic)."""
    items = [it for it in wl.c4(60) if it.pixels <= 1_600_000][:16]
    net = sg.Surrogate(seed=1)
    net.load(sg.DEFAULT_WEIGHTS)
    thr = 0.35
    mx = mixed.Mixed(items, net, mode=oc.MODE_STD, streams=2, max_fallback=thr)
    cands = [i for i, it in enumerate(items) if it.radius >= 9]
    assert cands and sorted(mx.calibration) == cands
    for i, (total, fb) in mx.calibration.items():
        assert total > 0 and 0 <= fb <= total
        assert (i in mx.sur_idx) == (fb <= thr * total)
    assert sorted(mx.sur_idx + mx.ex_idx) == list(range(len(items)))
    mx.capture()
    mx.reset_counts()
    mx.replay()
    tiles, fb = mx.counts()
    assert tiles == sum(mx.calibration[i][0] for i in mx.sur_idx)
    assert fb == sum(mx.calibration[i][1] for i in mx.sur_idx)
    got = [t.cpu().numpy() for t in mx.tout]
    for i in mx.ex_idx:
        it = items[i]
        exp = oracle.filter(mx.tin[i].cpu().numpy(), it.radius, it.sigma_space, it.sigma_range,
                            mode=oracle.STD, roi=it.roi)
        assert np.array_equal(got[i], exp), f"image {i} (r={it.radius})"
    moved = [i for i in cands if i in mx.ex_idx]
    print(f"calibration: {len(moved)} of {len(cands)} network candidates moved to the batch")
