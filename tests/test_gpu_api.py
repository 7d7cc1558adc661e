"""The reference-compatible C++ API on the GPU: tests/cpp/api_driver reads a
BMP (ImageBasicOp::readFromFile), sets a ROI, runs BilateralFilter::doFilter
and writes the result (ImageBasicOp::writeToFile); results are compared with
the oracle."""
import os
import subprocess

import numpy as np
import pytest

import opencuda_amd as oc
import oracle
from tests._bmp import read_bmp, write_bmp

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "cpp", "api_driver")


def _run(tmp_path, img, ss, sr, r, repeat, roi, out_size=(0, 0), mode=0, inplace=0):
    src, dst = tmp_path / "in.bmp", tmp_path / "out.bmp"
    write_bmp(src, img)
    args = [DRIVER, str(src), str(dst), repr(ss), repr(sr), str(r), str(repeat), *map(str, roi),
            *map(str, out_size), str(mode), str(inplace)]
    p = subprocess.run(args, capture_output=True, text=True, timeout=120)
    status = int(p.stdout.split("status")[1].split()[0]) if "status" in p.stdout else None
    out = read_bmp(dst) if p.returncode == 0 else None
    return status, out, p


@pytest.fixture(scope="module")
def driver(device):
    if not os.path.exists(DRIVER):
        pytest.fail("tests/cpp/api_driver missing: run __graft_entry__.build()")
    return DRIVER


def test_bmp_round_trip_through_library(tmp_path, driver):
    img = oc.synth_image(37, 23, 5, 2, 9, 10.0, 1)   # odd width: row padding
    status, out, p = _run(tmp_path, img, 2.0, 30.0, 2, 1, (0, 0, 37, 23))
    assert status == 0, p.stdout + p.stderr
    assert out.shape == img.shape


@pytest.mark.parametrize("mode", [0, 1])
def test_doFilter_out_of_place_full(tmp_path, driver, mode):
    img = oc.synth_image(203, 157, 20, 4, 40, 10.0, 2)
    status, out, p = _run(tmp_path, img, 3.0, 30.0, 3, 1, (0, 0, 203, 157), mode=mode)
    assert status == 0, p.stdout + p.stderr
    assert np.array_equal(out, oracle.filter(img, 3, 3.0, 30.0, mode=mode))


def test_doFilter_roi_creates_output_of_roi_size(tmp_path, driver):
    img = oc.synth_image(300, 200, 20, 4, 60, 12.0, 3)
    roi = (17, 9, 290, 187)
    status, out, p = _run(tmp_path, img, 5.0, 25.0, 7, 1, roi)
    assert status == 0, p.stdout + p.stderr
    assert out.shape == (roi[3] - roi[1], roi[2] - roi[0])
    assert np.array_equal(out, oracle.filter(img, 7, 5.0, 25.0, roi=roi))


def test_doFilter_trims_to_smaller_output(tmp_path, driver):
    img = oc.synth_image(160, 120, 12, 4, 40, 10.0, 4)
    roi = (10, 10, 150, 110)
    status, out, p = _run(tmp_path, img, 3.0, 30.0, 3, 1, roi, out_size=(100, 64))
    assert status == 0, p.stdout + p.stderr
    exp = oracle.filter(img, 3, 3.0, 30.0, roi=roi, out_size=(100, 64))
    assert np.array_equal(out[:64, :100], exp)


def test_doFilter_in_place_repeat(tmp_path, driver):
    img = oc.synth_image(128, 96, 12, 4, 40, 10.0, 5)
    roi = (8, 4, 120, 90)
    status, out, p = _run(tmp_path, img, 3.0, 20.0, 3, 2, roi, inplace=1)
    assert status == 0, p.stdout + p.stderr
    exp = img.copy()
    sub = np.ascontiguousarray(exp[4:90, 8:120])
    for _ in range(2):
        sub = oracle.filter(sub, 3, 3.0, 20.0)
    exp[4:90, 8:120] = sub
    assert np.array_equal(out, exp)


def test_invalid_sigma_is_reported_by_doFilter(tmp_path, driver):
    """The constructor swallows setter errors (BilateralFilter.h:88-91); with
    no spatial template doFilter returns INVALID_DATA (.cu:269-270)."""
    img = oc.synth_image(40, 30, 3, 2, 9, 10.0, 6)
    status, out, p = _run(tmp_path, img, -1.0, 30.0, 3, 1, (0, 0, 40, 30))
    assert status == oc.INVALID_DATA


def test_zero_repeat_is_invalid(tmp_path, driver):
    img = oc.synth_image(40, 30, 3, 2, 9, 10.0, 6)
    status, out, p = _run(tmp_path, img, 3.0, 30.0, 3, 0, (0, 0, 40, 30))
    assert status == oc.INVALID_DATA


def test_device_allocation_failure_is_reported(driver):
    """A device image larger than the device (1 TiB) fails with CUDA_ERROR,
    as the reference's cudaMallocPitch path does (Image.cu:94-137); the image
    stays empty and the next allocation works.  No kernel runs."""
    p = subprocess.run([DRIVER, "alloc-fail"], capture_output=True, text=True, timeout=60)
    assert p.returncode == 0, p.stderr
    res = dict(line.split() for line in p.stdout.splitlines())
    assert int(res["device_huge"]) == oc.CUDA_ERROR
    assert int(res["device_empty"]) == 1
    assert int(res["device_after"]) == 0
