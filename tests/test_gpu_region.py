"""The stages around the filter on the GPU (okano_3_0/SimpleRegionDetect.cu
:287-312): LocalCluster and ConnectRegion bit-exact against the oracle, and
the chain LocalCluster -> BilateralFilter -> ConnectRegion through the
reference-compatible C++ classes (tests/cpp/chain_driver)."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

import opencuda_amd as oc
import oracle

from tests._bmp import read_bmp, write_bmp

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "cpp", "chain_driver")


def _regions_image(W, H, seed, cut=60):
    img = oc.synth_image(W, H, max(4, W * H // 4000), 4, max(8, min(W, H) // 4), 6.0, seed)
    img = img.copy()
    img[img < cut] = 0          # background below the 10..250 range
    img[img > 240] = 255        # and some above it
    return img


def _gpu_connect(img, roi, out_size, min_area, max_area):
    import torch
    tin = torch.from_numpy(img).cuda()
    x1, y1, x2, y2 = roi or (0, 0, img.shape[1], img.shape[0])
    ow, oh = out_size or (x2 - x1, y2 - y1)
    tout = torch.full((oh, ow), 7, dtype=torch.uint8, device="cuda")
    oc.connect_region(oc.view(tin, roi), oc.view(tout), min_area, max_area)
    torch.cuda.synchronize()
    return tout.cpu().numpy()


@pytest.mark.parametrize("shape,roi,out_size,min_area,max_area", [
    ((300, 270), None, None, 20, 100000),            # H = 270: area rows quirk
    ((1000, 60), (3, 5, 998, 57), None, 5, 400),     # pitched ROI
    ((512, 512), None, (500, 480), 100, 100000),     # trimmed output
    ((2048, 1536), None, None, 50, 20000),
    ((64, 40), None, None, 1, 100000),               # every background pixel a region
])
def test_connect_region_matches_oracle(device, shape, roi, out_size, min_area, max_area):
    W, H = shape
    img = _regions_image(W, H, 11 + W)
    got = _gpu_connect(img, roi, out_size, min_area, max_area)
    ref = oracle.connect_region(img, roi, out_size, min_area, max_area)
    assert got.shape == ref.shape
    assert np.array_equal(got, ref)


def test_connect_region_number_wrap(device):
    H, W = 40, 60
    img = np.zeros((H, W), np.uint8)
    img[0::2, 0::3] = 50
    img[0::2, 1::3] = 50
    got = _gpu_connect(img, None, None, 2, 2)
    assert np.array_equal(got, oracle.connect_region(img, min_area=2, max_area=2))
    vals = got[0::2, 0::3].reshape(-1)
    assert list(vals[:255]) == list(range(1, 256)) and np.all(vals[255:] == 0)


def test_connect_region_one_large_region(device):
    # a single region covering almost the image: long union chains
    W, H = 1500, 1100
    img = np.full((H, W), 100, np.uint8)
    img[::7, ::5] = 0
    got = _gpu_connect(img, None, None, 1, 10 ** 7)
    assert np.array_equal(got, oracle.connect_region(img, min_area=1, max_area=10 ** 7))


@pytest.mark.parametrize("params", [
    dict(pnt_range=8, pnt_count=5, gap_thred=10, diffe_thred=10),     # parameterised ctor
    dict(pnt_range=16, pnt_count=7, gap_thred=40, diffe_thred=40),    # default ctor
    dict(pnt_range=3, pnt_count=8, gap_thred=255, diffe_thred=255),
    # halo 78: the LDS tile (220 x 172 bytes) near its size limit
    dict(pnt_range=40, pnt_count=6, gap_thred=60, diffe_thred=60),
    # halo 118: beyond the LDS limit, every read from global memory
    dict(pnt_range=60, pnt_count=4, gap_thred=255, diffe_thred=255),
])
@pytest.mark.parametrize("layout", ["full", "roi"])
def test_local_cluster_matches_oracle(device, params, layout):
    import torch
    W, H = 301, 222
    img = oc.synth_image(W, H, 40, 4, 60, 8.0, 5 + params["pnt_range"])
    roi = (17, 9, 290, 211) if layout == "roi" else None
    tin = torch.from_numpy(img).cuda()
    x1, y1, x2, y2 = roi or (0, 0, W, H)
    tout = torch.full((y2 - y1, x2 - x1), 3, dtype=torch.uint8, device="cuda")
    alloc = oc.alloc_range(tin, roi or (0, 0, W, H))
    oc.local_cluster(oc.view(tin, roi), oc.view(tout), alloc=alloc, **params)
    torch.cuda.synchronize()
    ref = oracle.local_cluster(img, roi, pntrange=params["pnt_range"],
                               pntcount=params["pnt_count"], gapthred=params["gap_thred"],
                               diffethred=params["diffe_thred"])
    assert np.array_equal(tout.cpu().numpy(), ref)


def test_chain_through_cpp_classes(device):
    if not os.path.exists(DRIVER):
        pytest.fail("tests/cpp/chain_driver missing: run __graft_entry__.build()")
    W, H = 333, 250
    img = oc.synth_image(W, H, 30, 6, 70, 10.0, 77)
    pr, pc, gap, diffe = 3, 5, 12, 14      # walks stay inside the 5-pixel guard band
    ss, sr, r = 3.0, 30.0, 3
    min_area, max_area = 15, 100000
    with tempfile.TemporaryDirectory() as d:
        paths = [os.path.join(d, n) for n in ("in.bmp", "lc.bmp", "bf.bmp", "cr.bmp")]
        write_bmp(paths[0], img)
        res = subprocess.run([DRIVER, *paths, str(pr), str(pc), str(gap), str(diffe), str(ss),
                              str(sr), str(r), str(min_area), str(max_area)],
                             capture_output=True, text=True, timeout=120)
        print(res.stdout, res.stderr)
        assert res.returncode == 0
        lc, bf, cr = (read_bmp(p) for p in paths[1:])
    ref_lc = oracle.local_cluster(img, pntrange=pr, pntcount=pc, gapthred=gap, diffethred=diffe)
    assert np.array_equal(lc, ref_lc)
    ref_bf = oracle.filter(ref_lc, r, ss, sr, mode=oracle.REF)
    assert np.array_equal(bf, ref_bf)
    ref_cr = oracle.connect_region(ref_bf, min_area=min_area, max_area=max_area)
    assert np.array_equal(cr, ref_cr)
    assert cr.max() > 1          # the chain found several regions
