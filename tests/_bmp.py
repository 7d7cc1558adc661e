"""Minimal 8-bit grayscale BMP writer/reader for tests (bottom-up rows,
4-byte row padding, 256-entry grey palette)."""
import struct

import numpy as np


def write_bmp(path, img: np.ndarray) -> None:
    h, w = img.shape
    pad = (4 - (w & 3)) & 3
    dataoff = 54 + 1024
    datalen = h * (w + pad)
    with open(path, "wb") as f:
        f.write(struct.pack("<2sIHHI", b"BM", dataoff + datalen, 0, 0, dataoff))
        f.write(struct.pack("<IiiHHIIiiII", 40, w, h, 1, 8, 0, datalen, 0, 0, 256, 0))
        for i in range(256):
            f.write(bytes((i, i, i, 0)))
        for r in range(h - 1, -1, -1):
            f.write(img[r].tobytes() + b"\0" * pad)


def read_bmp(path) -> np.ndarray:
    data = open(path, "rb").read()
    assert data[:2] == b"BM"
    dataoff = struct.unpack_from("<I", data, 10)[0]
    w, h = struct.unpack_from("<ii", data, 18)
    bpp = struct.unpack_from("<H", data, 28)[0]
    assert bpp == 8
    pad = (4 - (w & 3)) & 3
    out = np.zeros((h, w), np.uint8)
    for r in range(h):
        off = dataoff + r * (w + pad)
        out[h - 1 - r] = np.frombuffer(data, np.uint8, w, off)
    return out
