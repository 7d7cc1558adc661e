"""MI355X bilateral filter -- Python binding of libopencuda_bf.so.

The product is the C ABI in ``include/opencuda_bf.h`` (and the C++ object
model above it).  This module binds that ABI with ctypes so Python callers,
the tests and ``bench.py`` can drive it with device memory owned by torch.
It never falls back to a CPU implementation: if the HIP library is missing
it raises.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

MODE_REF = 0
MODE_STD = 1
MAX_RADIUS = 20

# ErrorCode.h
NO_ERROR = 0
INVALID_DATA = -1
NULL_POINTER = -2
OP_OVERFLOW = -3
NO_FILE = -4
WRONG_FILE = -5
OUT_OF_MEM = -6
CUDA_ERROR = -7
UNMATCH_IMG = -8
UNIMPLEMENT = -998
UNKNOW_ERROR = -999

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libopencuda_bf.so")
_lib = None


class BfError(RuntimeError):
    def __init__(self, code: int, where: str, msg: str = ""):
        super().__init__(f"{where} failed with status {code}" + (f": {msg}" if msg else ""))
        self.code = code


class View(ctypes.Structure):
    """bf_image_view: ROI view of a device image."""
    _fields_ = [("data", ctypes.c_void_p), ("width", ctypes.c_int32),
                ("height", ctypes.c_int32), ("pitch", ctypes.c_int64)]


class LocalClusterParams(ctypes.Structure):
    """bf_local_cluster_params (LocalCluster's members)."""
    _fields_ = [("pnt_range", ctypes.c_int32), ("pnt_count", ctypes.c_int32),
                ("gap_thred", ctypes.c_uint8), ("diffe_thred", ctypes.c_uint8),
                ("pro_black", ctypes.c_uint8), ("pro_white", ctypes.c_uint8)]


class Job(ctypes.Structure):
    """bf_job: one image of a batch."""
    _fields_ = [("inp", View), ("out", View), ("lut", ctypes.c_void_p),
                ("radius", ctypes.c_int32), ("mode", ctypes.c_int32)]


def lib() -> ctypes.CDLL:
    """Load libopencuda_bf.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with "
                          "`python -c 'import __graft_entry__ as g; g.build()'`")
    L = ctypes.CDLL(LIB_PATH)
    P, I32, I64, F = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_float
    VP = ctypes.POINTER(View)
    sig = {
        "bf_gauss_template": ([F, I32, P, P], ctypes.c_int),
        "bf_euclid_template": ([F, I32, P], ctypes.c_int),
        "bf_lut_create": ([P, P, I32, P, ctypes.POINTER(P)], ctypes.c_int),
        "bf_lut_destroy": ([P], ctypes.c_int),
        "bf_lut_rows": ([P, I32], ctypes.c_int),
        "bf_exact": ([VP, VP, P, I32, I32, P], ctypes.c_int),
        "bf_exact_config": ([VP, VP, P, I32, I32, I32, P], ctypes.c_int),
        "bf_exact_inplace": ([VP, P, I32, I32, I32, P, P], ctypes.c_int),
        "bf_tune_count": ([], ctypes.c_int),
        "bf_tune_describe": ([I32, P, P, P, P], ctypes.c_int),
        "bf_tune_radius": ([I32, I32, I32, I32, I32, P, P], ctypes.c_int),
        "bf_tune_set": ([I32, I32, P], ctypes.c_int),
        "bf_tune_get": ([I32, I32, P], ctypes.c_int),
        "bf_tune_reset": ([], ctypes.c_int),
        "bf_batch_create": ([P, I32, I32, ctypes.POINTER(P)], ctypes.c_int),
        "bf_batch_destroy": ([P], ctypes.c_int),
        "bf_batch_run": ([P, P], ctypes.c_int),
        "bf_batch_run_graph": ([P, P], ctypes.c_int),
        "bf_batch_pixels": ([P], I64),
        "bf_batch_unrolled": ([P], ctypes.c_int),
        "bf_batch_config": ([P], ctypes.c_int),
        "bf_sur_param_count": ([], I64),
        "bf_sur_create": ([ctypes.c_uint64, ctypes.POINTER(P)], ctypes.c_int),
        "bf_sur_destroy": ([P], ctypes.c_int),
        "bf_sur_get_params": ([P, P], ctypes.c_int),
        "bf_sur_set_params": ([P, P], ctypes.c_int),
        "bf_sur_train_step": ([P, P, P, P, P, I32, I32, I32, F, F, F, F, P, P], ctypes.c_int),
        "bf_sur_set_loss_tail": ([P, F, F], ctypes.c_int),
        "bf_sur_forward_backward": ([P, P, P, P, P, I32, I32, I32, P, P, P, P], ctypes.c_int),
        "bf_sur_infer": ([P, VP, VP, F, P], ctypes.c_int),
        "bf_sur_filter": ([P, VP, VP, P, I32, I32, F, I32, P, P, P], ctypes.c_int),
        "bf_sur_filter_async": ([P, VP, VP, P, I32, I32, F, I32, P, P, P], ctypes.c_int),
        "bf_sur_filter_noalloc": ([P, VP, VP, P, I32, I32, F, I32, P, P, P], ctypes.c_int),
        "bf_last_error": ([], ctypes.c_char_p),
        "bf_version": ([], ctypes.c_int),
        "bf_synth_image": ([P, I32, I32, I32, I32, I32, ctypes.c_double, ctypes.c_uint64],
                           ctypes.c_int),
        "bf_local_cluster": ([VP, I64, I64, VP, ctypes.POINTER(LocalClusterParams), P],
                             ctypes.c_int),
        "bf_connect_region": ([VP, VP, I32, I32, I32, I32, P], ctypes.c_int),
        "bf_box_template": ([I32, I32, P, P], ctypes.c_int),
        "bf_circle_template": ([I32, P, P], ctypes.c_int),
        "bf_arc_template": ([I32, I32, P, P], ctypes.c_int),
    }
    for name, (args, res) in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    _lib = L
    return L


def _check(rc: int, where: str) -> int:
    if rc < 0:
        msg = lib().bf_last_error()
        raise BfError(rc, where, msg.decode() if msg else "")
    return rc


def last_error() -> str:
    m = lib().bf_last_error()
    return m.decode() if m else ""


# ------------------------------------------------------------------ tables

def gauss_template(sigma: float, size: int = 2 * MAX_RADIUS + 1):
    """Spatial template: (offsets int32[n,2] as (dy,dx), weights float32[n])."""
    n = size * size
    off = np.zeros((n, 2), np.int32)
    w = np.zeros(n, np.float32)
    _check(lib().bf_gauss_template(float(sigma), int(size), off.ctypes.data, w.ctypes.data),
           "bf_gauss_template")
    return off, w


def euclid_template(sigma: float, size: int = 256) -> np.ndarray:
    w = np.zeros(size, np.float32)
    _check(lib().bf_euclid_template(float(sigma), int(size), w.ctypes.data), "bf_euclid_template")
    return w


class Lut:
    """Device filter tables (bf_lut) for one spatial template + range table."""

    def __init__(self, offsets: np.ndarray, weights: np.ndarray, range_lut: np.ndarray):
        off = np.ascontiguousarray(offsets, np.int32).reshape(-1, 2)
        w = np.ascontiguousarray(weights, np.float32)
        r = np.ascontiguousarray(range_lut, np.float32)
        if r.size < 256 or w.size != off.shape[0]:
            raise ValueError("need 256 range weights and one weight per offset")
        h = ctypes.c_void_p()
        _check(lib().bf_lut_create(off.ctypes.data, w.ctypes.data, int(w.size), r.ctypes.data,
                                   ctypes.byref(h)), "bf_lut_create")
        self.handle = h
        self.count = int(w.size)

    @classmethod
    def from_sigmas(cls, sigma_space: float, sigma_range: float) -> "Lut":
        off, w = gauss_template(sigma_space)
        return cls(off, w, euclid_template(sigma_range))

    def rows(self, ntaps: int) -> int:
        return _check(lib().bf_lut_rows(self.handle, int(ntaps)), "bf_lut_rows")

    def close(self):
        if self.handle:
            lib().bf_lut_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ------------------------------------------------------------------- views

def view(t, roi=None) -> View:
    """View of a 2-D uint8 device tensor (torch), optionally of a ROI
    (x1, y1, x2, y2) with exclusive x2/y2."""
    import torch
    if not (isinstance(t, torch.Tensor) and t.dtype == torch.uint8 and t.dim() == 2
            and t.is_cuda and t.stride(1) == 1):
        raise ValueError("expected a row-contiguous 2-D uint8 device tensor")
    h, w = t.shape
    x1, y1, x2, y2 = roi if roi is not None else (0, 0, w, h)
    if not (0 <= x1 <= x2 <= w and 0 <= y1 <= y2 <= h):
        raise ValueError("ROI outside the image")
    pitch = t.stride(0)
    return View(t.data_ptr() + y1 * pitch + x1, x2 - x1, y2 - y1, pitch)


def _stream(stream):
    if stream is None:
        import torch
        return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    if isinstance(stream, int):
        return ctypes.c_void_p(stream)
    return ctypes.c_void_p(stream.cuda_stream)


def exact(vin: View, vout: View, lut: Lut, radius: int, mode: int = MODE_REF,
          stream=None, config: int | None = None) -> None:
    """bf_exact / bf_exact_config (asynchronous on `stream`)."""
    s = _stream(stream)
    if config is None:
        rc = lib().bf_exact(ctypes.byref(vin), ctypes.byref(vout), lut.handle, int(radius),
                            int(mode), s)
    else:
        rc = lib().bf_exact_config(ctypes.byref(vin), ctypes.byref(vout), lut.handle,
                                   int(radius), int(mode), int(config), s)
    _check(rc, "bf_exact")


def exact_inplace(v: View, lut: Lut, radius: int, mode: int = MODE_REF, repeat: int = 1,
                  work=None, stream=None) -> None:
    wp = None if work is None else ctypes.c_void_p(work.data_ptr())
    _check(lib().bf_exact_inplace(ctypes.byref(v), lut.handle, int(radius), int(mode),
                                  int(repeat), wp, _stream(stream)), "bf_exact_inplace")


def tune_configs():
    """[(id, tile_w, tile_h, threads, cols_per_lane)] of the exact kernel's
    tile configurations."""
    out = []
    for c in range(lib().bf_tune_count()):
        tw, th, nt, cpl = (ctypes.c_int32() for _ in range(4))
        _check(lib().bf_tune_describe(c, ctypes.byref(tw), ctypes.byref(th), ctypes.byref(nt),
                                      ctypes.byref(cpl)), "bf_tune_describe")
        out.append((c, tw.value, th.value, nt.value, cpl.value))
    return out


class Batch:
    """bf_batch: many jobs in one persistent launch."""

    def __init__(self, jobs: list[Job], config: int = -1):
        arr = (Job * len(jobs))(*jobs)
        h = ctypes.c_void_p()
        _check(lib().bf_batch_create(ctypes.cast(arr, ctypes.c_void_p), len(jobs), int(config),
                                     ctypes.byref(h)), "bf_batch_create")
        self.handle = h
        self.pixels = int(lib().bf_batch_pixels(h))
        self.unrolled = bool(_check(lib().bf_batch_unrolled(h), "bf_batch_unrolled"))
        self.config = int(_check(lib().bf_batch_config(h), "bf_batch_config"))

    def run(self, stream=None):
        _check(lib().bf_batch_run(self.handle, _stream(stream)), "bf_batch_run")

    def run_graph(self, stream=None):
        _check(lib().bf_batch_run_graph(self.handle, _stream(stream)), "bf_batch_run_graph")

    def close(self):
        if self.handle:
            lib().bf_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ------------------------------------------------------- region stages

def local_cluster(vin: View, vout: View, pnt_range: int = 8, pnt_count: int = 5,
                  gap_thred: int = 10, diffe_thred: int = 10, pro_black: int = 10,
                  pro_white: int = 250, alloc=(0, 0), stream=None) -> None:
    """bf_local_cluster: LocalCluster (okano_3_0/LocalCluster.cu) on a ROI
    view; `alloc` = readable byte range [lo, hi) around vin.data (see
    opencuda_bf.h), (0, 0) = the ROI itself.  Defaults are the
    parameterised constructor's."""
    p = LocalClusterParams(int(pnt_range), int(pnt_count), int(gap_thred), int(diffe_thred),
                           int(pro_black), int(pro_white))
    _check(lib().bf_local_cluster(ctypes.byref(vin), int(alloc[0]), int(alloc[1]),
                                  ctypes.byref(vout), ctypes.byref(p), _stream(stream)),
           "bf_local_cluster")


def alloc_range(t, roi) -> tuple:
    """[lo, hi) byte offsets of tensor t's storage around the ROI origin."""
    x1, y1 = roi[0], roi[1]
    origin = y1 * t.stride(0) + x1
    return -origin, t.shape[0] * t.stride(0) - origin


def connect_region(vin: View, vout: View, min_area: int = 100, max_area: int = 100000,
                   min_gray: int = 10, max_gray: int = 250, stream=None) -> None:
    """bf_connect_region: ConnectRegion (okano_3_0/ConnectRegion.cu)."""
    _check(lib().bf_connect_region(ctypes.byref(vin), ctypes.byref(vout), int(min_area),
                                   int(max_area), int(min_gray), int(max_gray), _stream(stream)),
           "bf_connect_region")


# ------------------------------------------------------------ synthetic data

def synth_image(width: int, height: int, nshapes: int, smin: int, smax: int,
                noise_sigma: float, seed: int) -> np.ndarray:
    """bf_synth_image: deterministic synthetic test image (host uint8)."""
    out = np.empty((height, width), np.uint8)
    _check(lib().bf_synth_image(out.ctypes.data, int(width), int(height), int(nshapes),
                                int(smin), int(smax), float(noise_sigma),
                                int(seed) & 0xFFFFFFFFFFFFFFFF), "bf_synth_image")
    return out


from . import autotune, workloads  # noqa: E402,F401
