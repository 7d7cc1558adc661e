"""Independent pure-Python statement of the filter for tiny inputs.

Written per pixel from the semantics of okano_3_0/BilateralFilter.cu:71-199
rather than per thread: pixel (x, y) of the output uses the window of row
y0 = y - y % 4 (REF) or y (STD), the centre value of its own position, the
first (2r+1)^2 (+1 in REF) points of the ring-ordered template, clamp-to-edge
sampling over the whole input ROI, float32 arithmetic one rounding per
operation, and truncation.  Weights come from glibc expf called through
ctypes.  Used only on images of a few hundred pixels.
"""
import ctypes
import ctypes.util

import numpy as np

_libm = ctypes.CDLL(ctypes.util.find_library("m"))
_libm.expf.argtypes = [ctypes.c_float]
_libm.expf.restype = ctypes.c_float

f32 = np.float32


def expf(x) -> np.float32:
    return f32(_libm.expf(float(f32(x))))


def ring_offsets(radius: int):
    """Template order: centre, then per ring i the 2i values j = -i+1..i,
    each giving (i,j), (-i,j), (j,i), (j,-i)."""
    keyed = [((0, 0, 0), (0, 0))]
    for i in range(1, radius + 1):
        for j in range(-i + 1, i + 1):
            for k, p in enumerate(((i, j), (-i, j), (j, i), (j, -i))):
                keyed.append(((i, j + i, k), p))
    keyed.sort(key=lambda e: e[0])
    return [p for _, p in keyed]


def gauss_weight(dy: int, dx: int, sigma: float) -> np.float32:
    if dy == 0 and dx == 0:
        return f32(1.0)
    s22 = f32(f32(f32(2.0) * f32(sigma)) * f32(sigma))
    return expf(f32(-(dy * dy + dx * dx)) / s22)


def range_weight(d: int, sigma: float) -> np.float32:
    s22 = f32(f32(f32(2.0) * f32(sigma)) * f32(sigma))
    return expf(f32(-(d * d)) / s22)


def to_u8(q: np.float32) -> int:
    if not (q > 0):
        return 0
    if q >= 255:
        return 255
    return int(q)


def filter_pixelwise(img, roi, out_size, radius, ss, sr, mode, return_q=False):
    x1, y1, x2, y2 = roi
    w, h = x2 - x1, y2 - y1
    W = w if out_size is None else min(w, out_size[0])
    H = h if out_size is None else min(h, out_size[1])
    sub = img[y1:y2, x1:x2]

    def T(x, y):
        return int(sub[min(max(y, 0), h - 1), min(max(x, 0), w - 1)])

    taps = ring_offsets(radius + 1)[: (2 * radius + 1) ** 2 + (1 if mode == 0 else 0)]
    g = [gauss_weight(dy, dx, ss) for dy, dx in taps]
    E = [range_weight(d, sr) for d in range(256)]
    out = np.zeros((H, W), np.uint8)
    qs = np.zeros((H, W), np.float32)
    for y in range(H):
        y0 = y - y % 4 if mode == 0 else y
        for x in range(W):
            c = T(x, y)
            t = f32(0.0)
            s = f32(0.0)
            for (dy, dx), gw in zip(taps, g):
                p = T(x + dx, y0 + dy)
                f = f32(gw * E[abs(p - c)])
                t = f32(t + f32(f * f32(p)))
                s = f32(s + f)
            with np.errstate(invalid="ignore", divide="ignore"):
                q = f32(t / s)
            qs[y, x] = q
            out[y, x] = to_u8(q)
    return (out, qs) if return_q else out
