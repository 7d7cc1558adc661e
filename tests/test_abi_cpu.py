"""CPU-side checks of the product library: it loads, exports every entry
point the public headers declare, validates arguments without touching a
GPU, and its host-only utilities behave."""
import ctypes
import os
import re

import numpy as np
import pytest

import opencuda_amd as oc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_c_functions():
    src = open(os.path.join(ROOT, "include", "opencuda_bf.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(bf_[a-z0-9_]+)\s*\(", src)))


def test_library_loads_and_exports_every_declared_function():
    L = oc.lib()
    names = _declared_c_functions()
    assert len(names) >= 25
    for n in names:
        assert hasattr(L, n), f"{n} declared in include/opencuda_bf.h but not exported"


CXX_SYMBOLS = [
    "_ZN12LocalCluster12localClusterEP8Image_stS1_",
    "_ZN13ConnectRegion13connectRegionEP8Image_stS1_",
    "_ZN12ImageBasicOp8newImageEPP8Image_st",
    "_ZN12ImageBasicOp11deleteImageEP8Image_st",
    "_ZN12ImageBasicOp19makeAtCurrentDeviceEP8Image_stmm",
    "_ZN12ImageBasicOp10makeAtHostEP8Image_stmm",
    "_ZN12ImageBasicOp12readFromFileEPKcP8Image_st",
    "_ZN12ImageBasicOp11writeToFileEPKcP8Image_st",
    "_ZN12ImageBasicOp19copyToCurrentDeviceEP8Image_st",
    "_ZN12ImageBasicOp19copyToCurrentDeviceEP8Image_stS1_",
    "_ZN12ImageBasicOp10copyToHostEP8Image_st",
    "_ZN12ImageBasicOp10copyToHostEP8Image_stS1_",
    "_ZN15TemplateBasicOp11newTemplateEPP11Template_st",
    "_ZN15TemplateBasicOp14deleteTemplateEP11Template_st",
    "_ZN15TemplateBasicOp10makeAtHostEP11Template_stm",
    "_ZN15TemplateBasicOp19makeAtCurrentDeviceEP11Template_stm",
    "_ZN15TemplateBasicOp12readFromFileEPKcP11Template_st",
    "_ZN15TemplateBasicOp11writeToFileEPKcP11Template_st",
    "_ZN15TemplateBasicOp19copyToCurrentDeviceEP11Template_st",
    "_ZN15TemplateBasicOp10copyToHostEP11Template_st",
    "_ZN15TemplateFactory11getTemplateEPP11Template_sti6TFSizePv",
    "_ZN15TemplateFactory11putTemplateEP11Template_st",
    "_ZN15BilateralFilter8doFilterEP8Image_st",
    "_ZN15BilateralFilter8doFilterEP8Image_stS1_",
]


def test_cxx_object_model_exported():
    L = oc.lib()
    for s in CXX_SYMBOLS:
        assert hasattr(L, s), s


def test_argument_validation_without_gpu():
    L = oc.lib()
    v = oc.View(None, 4, 4, 4)
    # null lut / null views are rejected before any device call
    assert L.bf_exact(None, None, None, 3, 0, None) == oc.NULL_POINTER
    assert L.bf_exact(ctypes.byref(v), ctypes.byref(v), None, 3, 0, None) == oc.NULL_POINTER
    assert L.bf_lut_create(None, None, 10, None, None) == oc.NULL_POINTER
    assert L.bf_exact_inplace(None, None, 3, 0, 1, None, None) == oc.NULL_POINTER
    assert L.bf_batch_create(None, 0, -1, None) == oc.NULL_POINTER
    assert L.bf_gauss_template(0.0, 41, None, None) == oc.NULL_POINTER
    off = np.zeros((41 * 41, 2), np.int32)
    w = np.zeros(41 * 41, np.float32)
    assert L.bf_gauss_template(0.0, 41, off.ctypes.data, w.ctypes.data) == oc.INVALID_DATA
    assert L.bf_gauss_template(1.0, 40, off.ctypes.data, w.ctypes.data) == oc.INVALID_DATA
    E = np.zeros(256, np.float32)
    assert L.bf_euclid_template(0.0, 256, E.ctypes.data) == oc.INVALID_DATA
    assert L.bf_euclid_template(10.0, 256, E.ctypes.data) == 256
    # the region stages reject null and malformed views before any device call
    p = oc.LocalClusterParams(8, 5, 10, 10, 10, 250)
    assert L.bf_local_cluster(None, 0, 0, None, ctypes.byref(p), None) == oc.NULL_POINTER
    assert L.bf_connect_region(None, None, 1, 10, 10, 250, None) == oc.NULL_POINTER
    bad = oc.View(1234, 8, 8, 4)           # pitch < width
    assert L.bf_connect_region(ctypes.byref(bad), ctypes.byref(bad), 1, 10, 10, 250,
                               None) == oc.INVALID_DATA
    good = oc.View(1234, 8, 8, 8)
    assert L.bf_local_cluster(ctypes.byref(good), 8, 100, ctypes.byref(good), ctypes.byref(p),
                              None) == oc.INVALID_DATA    # allocation must contain the ROI
    assert oc.last_error() is not None


def test_tile_configs_describe():
    cfgs = oc.tune_configs()
    assert len(cfgs) >= 4
    for c, tw, th, nt, cpl in cfgs:
        assert tw * th // 4 // cpl == nt and th % 4 == 0 and nt % 64 == 0 and tw % cpl == 0
    # pixels per lane is a tuned dimension: some configurations give a lane
    # two columns (8 pixels), the others one (4 pixels)
    assert {c[4] for c in cfgs} == {1, 2}
    assert oc.lib().bf_tune_describe(999, None, None, None, None) == oc.INVALID_DATA


def test_synth_image_is_deterministic():
    a = oc.synth_image(300, 200, 20, 8, 64, 10.0, 123)
    b = oc.synth_image(300, 200, 20, 8, 64, 10.0, 123)
    c = oc.synth_image(300, 200, 20, 8, 64, 10.0, 124)
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    # no shapes, no noise -> exactly the rounded gradient
    g = oc.synth_image(101, 3, 0, 1, 1, 0.0, 1)
    assert np.array_equal(g[0], np.rint(200.0 * np.arange(101) / 100).astype(np.uint8))
    # noise statistics of a flat region
    n = oc.synth_image(2000, 50, 0, 1, 1, 12.0, 7).astype(float) - np.rint(
        200.0 * np.arange(2000) / 1999)[None, :]
    inner = n[:, 200:1800]
    assert abs(inner.mean()) < 0.2 and 11.5 < inner.std() < 12.5


def test_workload_definitions():
    from opencuda_amd import workloads as wl
    (c0,) = wl.c0()
    assert (c0.width, c0.height, c0.radius, c0.sigma_space, c0.sigma_range) == (512, 512, 3, 3.0, 30.0)
    (c1,) = wl.c1()
    assert c1.roi == (100, 200, 3900, 3800) and c1.pixels == 3800 * 3600
    c2 = wl.c2()
    assert len(c2) == 256
    for it in c2:
        x1, y1, x2, y2 = it.roi
        assert 256 <= it.width <= 2048 and 256 <= it.height <= 2048
        assert 2 * (x2 - x1) >= it.width and 2 * (y2 - y1) >= it.height
        assert x2 <= it.width and y2 <= it.height
        assert it.radius in (3, 5, 9) and 10 <= it.sigma_range <= 50
        assert it.sigma_space == pytest.approx(it.radius / 1.5)
    (c3,) = wl.c3()
    assert c3.pixels == 8192 * 8192 and c3.radius == 15
    c4 = wl.c4()
    assert len(c4) == 512 and {it.radius for it in c4} == {3, 7, 15}
    assert wl.c2()[17] == c2[17]     # deterministic


def test_public_headers_compile_with_plain_gxx(tmp_path):
    """The C++ object model headers need no HIP toolchain."""
    import subprocess
    src = tmp_path / "t.cpp"
    src.write_text('#include "BilateralFilter.h"\n#include "LocalCluster.h"\n'
                   '#include "ConnectRegion.h"\n#include "opencuda_bf.h"\n'
                   'int main(){ BilateralFilter f; LocalCluster l; ConnectRegion c(0, 500, 20);\n'
                   '  return f.getRadius() + l.getPntRange() + c.getMinArea(); }\n')
    r = subprocess.run(["g++", "-std=c++11", "-fsyntax-only", "-I", os.path.join(ROOT, "include"),
                        str(src)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_tune_table_without_gpu():
    from opencuda_amd import autotune
    n = len(oc.tune_configs())
    try:
        cost = [0.5 + 0.1 * c for c in range(n)]
        cost[3] = 0.1
        autotune.install(7, oc.MODE_STD, cost)
        best, got = autotune.get(7, oc.MODE_STD)
        assert best == 3 and np.allclose(got, cost)
        assert autotune.get(7, oc.MODE_REF)[0] == -1
        autotune.install(7, oc.MODE_STD, [-1.0] * n)      # nothing usable -> untuned
        assert autotune.get(7, oc.MODE_STD)[0] == -1
    finally:
        autotune.reset()
