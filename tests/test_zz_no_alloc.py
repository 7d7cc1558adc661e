"""bf_sur_filter_noalloc (the allocation-free variant of
bf_sur_filter_async, whose fallback list lives in the library's static
device array) against bf_sur_filter_async, called through the C ABI:
identical output bytes and identical fallback/tile counts on two calls in a
row, the second on a smaller image (the array keeps the first call's list:
the count must start again from zero), plus an image beyond the array's
65536 tiles rejected before any launch."""
import ctypes

import numpy as np
import pytest

import opencuda_amd as oc
from opencuda_amd import surrogate as sg

pytestmark = pytest.mark.gpu


def _call(fn, net, tin, tout, lut, radius, mode, sr, fa, ta):
    return fn(net.handle, ctypes.byref(oc.view(tin)), ctypes.byref(oc.view(tout)), lut.handle,
              radius, mode, ctypes.c_float(sr), 2, ctypes.c_void_p(fa.data_ptr()),
              ctypes.c_void_p(ta.data_ptr()), oc._stream(None))


def test_noalloc_equals_the_allocating_entry_point(device):
    import torch
    L = oc.lib()
    net = sg.Surrogate(seed=1)
    net.load(sg.DEFAULT_WEIGHTS)
    for (h, w, seed, radius, mode, sr) in ((700, 900, 11, 15, oc.MODE_STD, 30.0),
                                          (150, 230, 12, 9, oc.MODE_STD, 22.0)):
        img = oc.synth_image(w, h, 30, 8, 160, 12.0, seed)
        lut = oc.Lut.from_sigmas(radius / 1.5, sr)
        tin = torch.from_numpy(img).cuda()
        a, b = torch.zeros_like(tin), torch.full_like(tin, 99)
        cnt = torch.zeros(4, dtype=torch.int32, device="cuda")
        assert _call(L.bf_sur_filter_async, net, tin, a, lut, radius, mode, sr,
                     cnt[0:1], cnt[1:2]) == 0
        assert _call(L.bf_sur_filter_noalloc, net, tin, b, lut, radius, mode, sr,
                     cnt[2:3], cnt[3:4]) == 0
        torch.cuda.synchronize()
        c = cnt.cpu().numpy()
        print(f"{h}x{w} r={radius}: fallback {c[0]}/{c[1]} (async), {c[2]}/{c[3]} (noalloc)")
        assert c[1] == c[3] == ((w + 63) // 64) * ((h + 63) // 64)
        assert c[0] == c[2] and 0 < c[0] < c[1]
        assert torch.equal(a, b)


def test_noalloc_rejects_images_beyond_its_array(device):
    import torch
    L = oc.lib()
    net = sg.Surrogate(seed=1)
    lut = oc.Lut.from_sigmas(10.0, 30.0)
    t = torch.zeros((64, 64), dtype=torch.uint8, device="cuda")
    o = torch.full((64, 64), 5, dtype=torch.uint8, device="cuda")
    # a view claiming 16448 x 16448 pixels (66049 tiles): rejected on the
    # tile count before anything reads or writes it
    big_in = oc.View(t.data_ptr(), 16448, 16448, 16448)
    big_out = oc.View(o.data_ptr(), 16448, 16448, 16448)
    cnt = torch.zeros(2, dtype=torch.int32, device="cuda")
    rc = L.bf_sur_filter_noalloc(net.handle, ctypes.byref(big_in), ctypes.byref(big_out),
                                 lut.handle, 15, oc.MODE_STD, ctypes.c_float(30.0), 2,
                                 ctypes.c_void_p(cnt.data_ptr()),
                                 ctypes.c_void_p(cnt.data_ptr() + 4), oc._stream(None))
    assert rc == oc.INVALID_DATA
    assert "65536" in oc.last_error()
    torch.cuda.synchronize()
    assert int(o.min()) == int(o.max()) == 5 and int(cnt.abs().sum()) == 0
