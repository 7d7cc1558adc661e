"""CPU oracle for the bilateral filter -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s
``cpu_baseline`` leg may import this package, and only as the checker (or
the timed CPU baseline), never as the thing measured or shipped.  The
product (``opencuda_amd``) never imports it.

It wraps ``liborc.so`` built from ``bf_oracle.c`` (see that file for the
reference lines each function restates).  Every function here is pinned by
tests in ``tests/test_oracle_*.py``:

* ``gauss_template`` / ``euclid_template``: bit-exact against glibc
  ``expf`` called directly, against the closed-form ring enumeration, and
  within 1 ulp of exp() evaluated in double.
* ``filter`` (REF and STD): brute force on tiny images (an independent
  Python restatement of the per-pixel semantics), FP64 bounds
  (|u8 - v64| < 1 + truncation), invariants (constant images,
  window min/max, ROI isolation), the sigma_range -> infinity special case
  (a normalised Gaussian blur, checked against a numpy convolution), and
  the row-aliasing relations between REF and STD.  Planted mistakes
  (``set_mutation``) each fail at least one of those checks.
* ``filter_f64``: closed forms on constant/linear images and agreement with
  a float64 numpy evaluation on small inputs.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

REF = 0
STD = 1

MUT_NONE = 0
MUT_DROP_EXTRA_TAP = 1
MUT_SWAP_DYDX = 2
MUT_OWN_ROW_WINDOW = 3
MUT_ROUND_NEAREST = 4
MUT_FUSED_MULADD = 5
MUT_CENTRE_ROW0 = 6
MUT_CLAMP_TRIMMED = 7

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liborc.so")
_lib = None


def build() -> None:
    """Compile liborc.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-B", "-C", _HERE], check=True)


def build_native(out_path: str) -> str:
    """Compile the same sources with -march=native into out_path (the CPU
    baseline's timed build on the host that runs it); returns the flags."""
    subprocess.run(["make", "-s", "-C", _HERE, "native", f"OUT={os.path.abspath(out_path)}"],
                   check=True, capture_output=True, timeout=300)
    return "-O3 -march=native -fopenmp -ffp-contract=off -fno-fast-math"


SHIPPED_FLAGS = "-O3 -march=x86-64-v3 -fopenmp -ffp-contract=off -fno-fast-math"


def use(path: str) -> None:
    """Load the oracle from another build of the same sources (the native
    CPU-baseline build); every function below then calls it."""
    global _LIB_PATH, _lib
    _LIB_PATH = os.path.abspath(path)
    _lib = None


def max_threads() -> int:
    """omp_get_max_threads() of the loaded oracle build."""
    return int(lib().orc_max_threads())


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        P = ctypes.c_void_p
        L.orc_gauss_template.argtypes = [ctypes.c_float, ctypes.c_int, P, P]
        L.orc_gauss_template.restype = ctypes.c_int
        L.orc_euclid_template.argtypes = [ctypes.c_float, ctypes.c_int, P]
        L.orc_euclid_template.restype = ctypes.c_int
        L.orc_filter.argtypes = [P, ctypes.c_size_t, ctypes.c_int, ctypes.c_int,
                                 P, ctypes.c_size_t, ctypes.c_int, ctypes.c_int,
                                 ctypes.c_int, P, P, P, ctypes.c_int, ctypes.c_int]
        L.orc_filter.restype = ctypes.c_int
        L.orc_filter_rows.argtypes = L.orc_filter.argtypes + [ctypes.c_int, ctypes.c_int]
        L.orc_filter_rows.restype = ctypes.c_int
        L.orc_filter_q.argtypes = [P, ctypes.c_size_t, ctypes.c_int, ctypes.c_int,
                                   P, ctypes.c_size_t, ctypes.c_int, ctypes.c_int,
                                   ctypes.c_int, P, P, P, ctypes.c_int, P, ctypes.c_size_t]
        L.orc_filter_q.restype = ctypes.c_int
        L.orc_filter_f64.argtypes = [P, ctypes.c_size_t, ctypes.c_int, ctypes.c_int,
                                     P, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                     P, ctypes.c_double, ctypes.c_double,
                                     ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.orc_filter_f64.restype = ctypes.c_int
        L.orc_max_threads.argtypes = []
        L.orc_max_threads.restype = ctypes.c_int
        L.orc_set_mutation.argtypes = [ctypes.c_int]
        L.orc_set_mutation.restype = None
        I64 = ctypes.c_int64
        L.orc_local_cluster.argtypes = [P, I64, I64, I64, ctypes.c_int, ctypes.c_int, P, I64] + \
            [ctypes.c_int] * 6
        L.orc_local_cluster.restype = ctypes.c_int
        L.orc_connect_region.argtypes = [P, I64, ctypes.c_int, ctypes.c_int, P, I64] + \
            [ctypes.c_int] * 4
        L.orc_connect_region.restype = ctypes.c_int
        for name, nargs in (("orc_box_template", 2), ("orc_circle_template", 1),
                            ("orc_arc_template", 2)):
            f = getattr(L, name)
            f.argtypes = [ctypes.c_int] * nargs + [P, P]
            f.restype = ctypes.c_int
        _lib = L
    return _lib


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


def set_mutation(m: int) -> None:
    lib().orc_set_mutation(int(m))


def gauss_template(sigma: float, size: int = 41):
    """(offsets int32[n,2] as (dy,dx), weights float32[n])."""
    n = size * size
    off = np.zeros((n, 2), np.int32)
    w = np.zeros(n, np.float32)
    got = lib().orc_gauss_template(float(sigma), int(size), _ptr(off), _ptr(w))
    if got < 0:
        raise ValueError("illegal gauss template parameters")
    assert got == n
    return off, w


def euclid_template(sigma: float, size: int = 256):
    w = np.zeros(size, np.float32)
    got = lib().orc_euclid_template(float(sigma), int(size), _ptr(w))
    if got < 0:
        raise ValueError("illegal euclid template parameters")
    return w


def _view(img: np.ndarray, roi):
    """ROI sub-image view: returns (base ndarray slice, pitch, w, h)."""
    assert img.dtype == np.uint8 and img.ndim == 2 and img.flags.c_contiguous
    if roi is None:
        roi = (0, 0, img.shape[1], img.shape[0])
    x1, y1, x2, y2 = roi
    pitch = img.strides[0]
    return img, y1 * pitch + x1, pitch, x2 - x1, y2 - y1


def filter(img: np.ndarray, radius: int, sigma_space: float, sigma_range: float,
           mode: int = REF, roi=None, out_size=None, nthreads: int = 0,
           tpl=None, E=None, rows=None) -> np.ndarray:
    """Filter the ROI of ``img``; returns an array of the output size
    (min of in-ROI and ``out_size``) like doFilter(in, out).  ``rows=(y0,
    y1)`` computes only output rows y0..y1-1 (y0 a multiple of 4 in REF
    mode) and returns them as a (y1-y0) x W array."""
    base, off, pitch, w, h = _view(img, roi)
    W, H = (w, h) if out_size is None else (min(w, out_size[0]), min(h, out_size[1]))
    if tpl is None:
        tpl, g = gauss_template(sigma_space)
    else:
        tpl, g = tpl
    if E is None:
        E = euclid_template(sigma_range)
    tpl = np.ascontiguousarray(tpl, np.int32)
    g = np.ascontiguousarray(g, np.float32)
    E = np.ascontiguousarray(E, np.float32)
    y0, y1 = rows if rows is not None else (0, H)
    out = np.zeros((y1 - y0, W), np.uint8)
    rc = lib().orc_filter_rows(base.ctypes.data + off, pitch, w, h, _ptr(out), W, W, H,
                               int(radius), _ptr(tpl), _ptr(g), _ptr(E), int(mode),
                               int(nthreads), int(y0), int(y1))
    if rc != 0:
        raise ValueError(f"orc_filter failed: {rc}")
    return out


def filter_q(img: np.ndarray, radius: int, sigma_space: float, sigma_range: float,
             mode: int = REF, roi=None, out_size=None):
    """Like filter(), also returning the float32 quotient t/s per pixel."""
    base, off, pitch, w, h = _view(img, roi)
    W, H = (w, h) if out_size is None else (min(w, out_size[0]), min(h, out_size[1]))
    tpl, g = gauss_template(sigma_space)
    E = euclid_template(sigma_range)
    out = np.zeros((H, W), np.uint8)
    q = np.zeros((H, W), np.float32)
    rc = lib().orc_filter_q(base.ctypes.data + off, pitch, w, h, _ptr(out), W, W, H,
                            int(radius), _ptr(tpl), _ptr(g), _ptr(E), int(mode), _ptr(q), W)
    if rc != 0:
        raise ValueError(f"orc_filter_q failed: {rc}")
    return out, q


def filter_f64(img: np.ndarray, radius: int, sigma_space: float, sigma_range: float,
               mode: int = REF, roi=None, out_size=None, nthreads: int = 0,
               rows=None) -> np.ndarray:
    """Real-valued result in float64; ``rows=(y0, y1)`` computes only those
    output rows (returned as a (y1-y0) x W array)."""
    base, off, pitch, w, h = _view(img, roi)
    W, H = (w, h) if out_size is None else (min(w, out_size[0]), min(h, out_size[1]))
    y0, y1 = rows if rows is not None else (0, H)
    tpl, _ = gauss_template(1.0)   # offsets do not depend on sigma
    out = np.zeros((y1 - y0, W), np.float64)
    rc = lib().orc_filter_f64(base.ctypes.data + off, pitch, w, h, _ptr(out), W, H,
                              int(radius), _ptr(tpl), float(sigma_space),
                              float(sigma_range), int(mode), int(nthreads), int(y0), int(y1))
    if rc != 0:
        raise ValueError(f"orc_filter_f64 failed: {rc}")
    return out


# ------------------------------------------------------------ region stages

def local_cluster(img: np.ndarray, roi=None, out_size=None, pntrange: int = 8,
                  gapthred: int = 10, diffethred: int = 10, problack: int = 10,
                  prowhite: int = 250, pntcount: int = 5) -> np.ndarray:
    """okano_3_0/LocalCluster.cu restated (region_oracle.c).  The image array
    is the whole allocation; reads the reference would make outside it are
    clamped into it.  Defaults are the parameterised constructor's
    (LocalCluster.h)."""
    base, off, pitch, w, h = _view(img, roi)
    if out_size is not None:
        w, h = min(w, out_size[0]), min(h, out_size[1])
    out = np.zeros((h, w), np.uint8)
    lo, hi = -off, img.size - off
    flat = base.reshape(-1)
    rc = lib().orc_local_cluster(flat.ctypes.data + off, pitch, lo, hi, w, h, _ptr(out), w,
                                 pntrange, gapthred, diffethred, problack, prowhite, pntcount)
    if rc != 0:
        raise ValueError(f"orc_local_cluster: {rc}")
    return out


def connect_region(img: np.ndarray, roi=None, out_size=None, min_area: int = 100,
                   max_area: int = 100000, mingray: int = 10, maxgray: int = 250) -> np.ndarray:
    """okano_3_0/ConnectRegion.cu restated (region_oracle.c): u8 region
    numbers (1..255, 0 = no region)."""
    base, off, pitch, w, h = _view(img, roi)
    if out_size is not None:
        w, h = min(w, out_size[0]), min(h, out_size[1])
    out = np.zeros((h, w), np.uint8)
    flat = base.reshape(-1)
    rc = lib().orc_connect_region(flat.ctypes.data + off, pitch, w, h, _ptr(out), w,
                                  min_area, max_area, mingray, maxgray)
    if rc != 0:
        raise ValueError(f"orc_connect_region: {rc}")
    return out


# ------------------------------------------------------------ shape templates

def shape_template(shape: str, *size):
    """(points int32[n, 2] as (x, y), weights float32[n]) of the BOX, CIRCLE
    or ARC vendor (tpl_oracle.c)."""
    f = getattr(lib(), f"orc_{shape}_template")
    n = f(*size, None, None)
    if n < 0:
        raise ValueError(f"illegal {shape} size {size}")
    pts = np.zeros((n, 2), np.int32)
    w = np.zeros(n, np.float32)
    assert f(*size, _ptr(pts), _ptr(w)) == n
    return pts, w
