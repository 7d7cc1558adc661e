"""Parity at the benchmark's full sizes.

* c3 (8192^2, r = 15), exact STD -- the surrogate's judge and fallback
  engine -- bit-exact to the oracle on all 67M pixels (REF is covered by
  test_gpu_exact.py::test_c3_parity_full_size).
* c3 surrogate + sampled check + fallback: the fallback set equals the
  host restatement of the sampled check, the output equals the surrogate's
  with exactly those tiles replaced by the exact filter, and the residual
  against exact STD (largest error, share of pixels beyond 2 LSB) is printed
  and bounded.
* c4 (all 512 images, 128..4096 per side, r in {3, 7, 15}, 2.28 Gpx) in one
  captured multi-stream graph: a replay equals the eager run, every image on
  the exact batch is bit-exact to the oracle, every image on the network
  route equals a plain single-stream bf_sur_filter call, and the graph's
  tile counts equal the sum over the network images.
"""
import sys
import time

import numpy as np
import pytest

import opencuda_amd as oc
import oracle
from opencuda_amd import mixed
from opencuda_amd import surrogate as sg
from opencuda_amd import workloads as wl
from tests._check import apply_fallback, expected_fallback

pytestmark = pytest.mark.gpu

# The residual of the shipped network at c3 after the fallback (measured:
# largest error 28 LSB, 0.129% of pixels beyond 2 LSB).  A change to the
# network or the check that makes it worse fails here.
C3_RESIDUAL_MAX = 28
C3_RESIDUAL_BEYOND_2 = 0.0013


def _log(msg):
    print(f"[fullsize] {msg}", file=sys.stderr, flush=True)


@pytest.fixture(scope="module")
def c3():
    (it,) = wl.c3()
    _log("generating the 8192^2 image")
    return it, it.image()


def test_c3_std_bit_exact_full_size(device, c3):
    import torch
    it, img = c3
    lut = oc.Lut.from_sigmas(it.sigma_space, it.sigma_range)
    tin = torch.from_numpy(img).cuda()
    tout = torch.zeros_like(tin)
    oc.exact(oc.view(tin), oc.view(tout), lut, it.radius, oc.MODE_STD)
    torch.cuda.synchronize()
    got = tout.cpu().numpy()
    t0 = time.time()
    exp = oracle.filter(img, it.radius, it.sigma_space, it.sigma_range, mode=oracle.STD)
    _log(f"oracle STD 8192^2 r=15: {time.time() - t0:.1f} s")
    assert np.array_equal(got, exp), f"{(got != exp).sum()} pixels differ"
    v = oracle.filter_f64(img, it.radius, it.sigma_space, it.sigma_range, mode=oracle.STD,
                          rows=(4096, 4160))
    err = got[4096:4160].astype(np.float64) - v
    assert np.abs(err).max() < 1.05 and -0.75 < err.mean() < -0.25


def test_c3_surrogate_fallback_set_and_residual(device, c3):
    import torch
    it, img = c3
    net = sg.Surrogate(seed=1)
    net.load(sg.DEFAULT_WEIGHTS)
    lut = oc.Lut.from_sigmas(it.sigma_space, it.sigma_range)
    tin = torch.from_numpy(img).cuda()
    sur, ex, out = (torch.zeros_like(tin) for _ in range(3))
    net.infer(oc.view(tin), oc.view(sur), it.sigma_range)
    oc.exact(oc.view(tin), oc.view(ex), lut, it.radius, oc.MODE_STD)
    total, fb = net.filter(oc.view(tin), oc.view(out), lut, it.radius, oc.MODE_STD,
                           it.sigma_range, tol_lsb=2)
    torch.cuda.synchronize()
    s_np, e_np, o_np = sur.cpu().numpy(), ex.cpu().numpy(), out.cpu().numpy()
    bad, tx = expected_fallback(s_np, e_np, 2)
    assert total == bad.size == 128 * 128
    assert fb == int(bad.sum())
    assert np.array_equal(o_np, apply_fallback(s_np, e_np, bad, tx))
    d = np.abs(o_np.astype(np.int16) - e_np.astype(np.int16))
    beyond = float((d > 2).mean())
    _log(f"c3 surrogate: {fb}/{total} tiles fell back ({fb / total:.2%}); residual after "
         f"fallback: max {int(d.max())} LSB, {beyond:.4%} of pixels beyond 2 LSB, mean "
         f"{float(d.mean()):.4f}")
    assert int(d.max()) <= C3_RESIDUAL_MAX and beyond <= C3_RESIDUAL_BEYOND_2


def test_c4_all_512_images_in_one_graph(device):
    import torch
    t0 = time.time()
    items = wl.c4()
    assert len(items) == 512
    net = sg.Surrogate(seed=1)
    net.load(sg.DEFAULT_WEIGHTS)
    mx = mixed.Mixed(items, net, mode=oc.MODE_STD, streams=4, log=_log)
    assert sorted(mx.sur_idx) == [i for i, it in enumerate(items) if it.radius == 15]
    _log(f"{len(mx.ex_idx)} exact + {len(mx.sur_idx)} network images uploaded "
         f"({time.time() - t0:.0f} s); capturing")
    mx.capture()                                   # eager run + capture
    eager = [t.clone() for t in mx.tout]
    for t in mx.tout:
        t.fill_(0)
    mx.reset_counts()
    mx.replay()
    tiles, fb = mx.counts()
    for i, (a, b) in enumerate(zip(mx.tout, eager)):
        assert torch.equal(a, b), f"image {i}: replay differs from the eager run"
    del eager
    n_tiles = sum(((items[i].roi[2] - items[i].roi[0] + 63) // 64) *
                  ((items[i].roi[3] - items[i].roi[1] + 63) // 64) for i in mx.sur_idx)
    assert tiles == n_tiles and 0 < fb < tiles
    _log(f"replay == eager; {fb}/{tiles} network tiles fell back ({time.time() - t0:.0f} s)")
    # network images: the graph's multi-stream run equals a plain call
    worst = 0
    for i in mx.sur_idx:
        it = items[i]
        single = torch.zeros_like(mx.tout[i])
        net.filter(oc.view(mx.tin[i], it.roi), oc.view(single),
                   mx.luts[(it.sigma_space, it.sigma_range)], it.radius, oc.MODE_STD,
                   it.sigma_range, 2)
        assert torch.equal(single, mx.tout[i]), f"image {i}"
        ref = mx.exact_reference(i)
        worst = max(worst, int((single.int() - ref.int()).abs().max()))
    _log(f"network images equal single-stream calls; largest error vs exact {worst} "
         f"({time.time() - t0:.0f} s)")
    # exact images: bit-exact to the oracle
    for k, i in enumerate(mx.ex_idx):
        it = items[i]
        got = mx.tout[i].cpu().numpy()
        exp = oracle.filter(mx.tin[i].cpu().numpy(), it.radius, it.sigma_space, it.sigma_range,
                            mode=oracle.STD, roi=it.roi)
        assert np.array_equal(got, exp), f"image {i} (r={it.radius}, {it.roi})"
        if (k + 1) % 64 == 0:
            _log(f"{k + 1}/{len(mx.ex_idx)} exact images checked ({time.time() - t0:.0f} s)")
