"""The two on-disk formats, pinned to the reference's byte layout rather than
to the library's own round trip.

* TPLT (okano_3_0/Template.cu:174-295 reads, :298-382 writes): "TPLT", a
  4-byte little-endian point count, 20 reserved bytes (written as zeros,
  ignored on read), 2*count int32 coordinates (point i = tplData[2i],
  tplData[2i+1]), count float32 weights.
* 8-bit BMP (okano_3_0/Image.cu:178-290 reads, :293-432 writes): "BM" at 0,
  file size at 2, pixel-data offset at 0x0A (the reader seeks there), a
  40-byte info header with width at 0x12, height at 0x16, planes 1, bpp 8 at
  0x1C, image size, 256 palette entries (i, i, i, 0xFF); rows bottom-up, each
  padded with zeros to 4 bytes; the writer writes the whole image whatever
  the ROI.

The expected bytes are built here with `struct` from those lines; the
library is driven through tests/cpp/host_api_driver (no GPU needed).
"""
import os
import struct
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "cpp", "host_api_driver")

NO_FILE, WRONG_FILE = -4, -5


def run(cmd, path):
    if not os.path.exists(DRIVER):
        import __graft_entry__
        __graft_entry__._build_api_driver()
    p = subprocess.run([DRIVER, cmd, str(path)], capture_output=True, text=True, timeout=60)
    assert p.returncode == 0, p.stderr
    return p.stdout.splitlines()


def tplt_bytes(points, weights, reserved=b"\0" * 20, magic=b"TPLT", count=None):
    n = len(points) if count is None else count
    body = b"".join(struct.pack("<ii", x, y) for x, y in points)
    return magic + struct.pack("<I", n) + reserved + body + np.asarray(weights, "<f4").tobytes()


def parse_tplt_dump(lines):
    kv = {}
    pts, ws = [], []
    for ln in lines:
        f = ln.split()
        if f[0] == "p":
            pts.append((int(f[1]), int(f[2])))
        elif f[0] == "w":
            ws.append(int(f[1], 16))
        else:
            kv[f[0]] = int(f[1])
    return kv, pts, ws


def test_tplt_read_of_a_stream_built_from_the_layout(tmp_path):
    rng = np.random.default_rng(5)
    n = 37
    pts = [(int(a), int(b)) for a, b in rng.integers(-(2**31), 2**31 - 1, (n, 2))]
    w = rng.standard_normal(n).astype(np.float32)
    w[3] = np.float32(1e-41)          # subnormal weight kept bit for bit
    w[4] = -0.0
    # reserved bytes are not zero: the reader must skip them
    path = tmp_path / "ext.tplt"
    path.write_bytes(tplt_bytes(pts, w, reserved=bytes(range(1, 21))))
    kv, got_p, got_w = parse_tplt_dump(run("tplt-read", path))
    assert kv["rc"] == 0 and kv["count"] == n
    assert got_p == pts
    assert got_w == [int(b) for b in w.view("<u4")]


def test_tplt_write_matches_the_layout_byte_for_byte(tmp_path):
    path = tmp_path / "lib.tplt"
    (line,) = run("tplt-write", path)
    assert line == "rc 0"
    pts = [(0, 0), (-3, 7), (12, -1), (-128, 65535), (2147483647, -2147483648)]
    w = np.array([1.0, 0.5, -2.25, 1e-40, 3.4e38], np.float32)
    data = path.read_bytes()
    assert data == tplt_bytes(pts, w)
    # field by field, as Template.cu:331-382 writes them
    assert data[:4] == b"TPLT"
    assert struct.unpack_from("<I", data, 4)[0] == 5
    assert data[8:28] == b"\0" * 20
    assert list(struct.unpack_from("<10i", data, 28)) == [c for p in pts for c in p]
    assert np.frombuffer(data, "<f4", 5, 68).tobytes() == w.tobytes()
    assert len(data) == 4 + 4 + 20 + 5 * 8 + 5 * 4


@pytest.mark.parametrize("case", ["magic", "zero_count", "missing"])
def test_tplt_errors(tmp_path, case):
    path = tmp_path / "bad.tplt"
    if case == "magic":
        path.write_bytes(tplt_bytes([(1, 2)], [1.0], magic=b"TPLX"))
    elif case == "zero_count":
        path.write_bytes(tplt_bytes([(1, 2)], [1.0], count=0))
    kv, _, _ = parse_tplt_dump(run("tplt-read", path))
    assert kv["rc"] == (NO_FILE if case == "missing" else WRONG_FILE)


def bmp_bytes(img, dataoff=None, bpp=8, magic=b"BM", gap=b""):
    h, w = img.shape
    pad = (4 - (w & 3)) & 3
    datalen = h * (w + pad)
    off = 54 + 1024 + len(gap) if dataoff is None else dataoff
    head = struct.pack("<2sIHHI", magic, off + datalen, 0, 0, off)
    info = struct.pack("<IiiHHIIiiII", 40, w, h, 1, bpp, 0, datalen, 0, 0, 256, 0)
    pal = b"".join(bytes((i, i, i, 0xFF)) for i in range(256))
    rows = b"".join(img[r].tobytes() + b"\xee" * pad for r in range(h - 1, -1, -1))
    return head + info + pal + gap + rows


def parse_bmp_dump(lines):
    kv, rows = {}, []
    for ln in lines:
        k, v = ln.split()
        if k == "row":
            rows.append(np.frombuffer(bytes.fromhex(v), np.uint8))
        else:
            kv[k] = int(v)
    return kv, (np.stack(rows) if rows else None)


@pytest.mark.parametrize("w,h,gap", [(5, 3, b""), (8, 4, b""), (13, 7, b"\x55" * 36)])
def test_bmp_read_of_a_file_built_from_the_layout(tmp_path, w, h, gap):
    img = np.random.default_rng(w * h).integers(0, 256, (h, w), dtype=np.uint8)
    path = tmp_path / "ext.bmp"
    # padding bytes are not zero and the pixel data may start after a gap:
    # the reader seeks to the 0x0A offset and skips each row's padding
    path.write_bytes(bmp_bytes(img, gap=gap))
    kv, got = parse_bmp_dump(run("bmp-read", path))
    assert kv["rc"] == 0 and kv["w"] == w and kv["h"] == h
    assert kv["roi"] == w * 100 + h          # ROI reset to the whole image
    assert np.array_equal(got, img)


@pytest.mark.parametrize("case", ["magic", "bpp24", "width0", "missing"])
def test_bmp_errors(tmp_path, case):
    img = np.zeros((3, 4), np.uint8)
    path = tmp_path / "bad.bmp"
    if case == "magic":
        path.write_bytes(bmp_bytes(img, magic=b"MB"))
    elif case == "bpp24":
        path.write_bytes(bmp_bytes(img, bpp=24))
    elif case == "width0":
        data = bytearray(bmp_bytes(img))
        struct.pack_into("<i", data, 0x12, 0)
        path.write_bytes(bytes(data))
    kv, _ = parse_bmp_dump(run("bmp-read", path))
    assert kv["rc"] == (NO_FILE if case == "missing" else WRONG_FILE)


def test_bmp_write_matches_the_layout_field_by_field(tmp_path):
    path = tmp_path / "lib.bmp"
    (line,) = run("bmp-write", path)
    assert line == "rc 0"
    data = path.read_bytes()
    w, h = 37, 11
    pad = 3
    datalen = h * (w + pad)
    assert data[:2] == b"BM"
    assert struct.unpack_from("<IHHI", data, 2) == (datalen + 1078, 0, 0, 1078)
    assert struct.unpack_from("<IiiHHIIiiII", data, 14) == (40, w, h, 1, 8, 0, datalen, 0, 0, 256, 0)
    pal = np.frombuffer(data, np.uint8, 1024, 54).reshape(256, 4)
    assert np.array_equal(pal[:, 0], np.arange(256)) and np.array_equal(pal[:, 1], pal[:, 0])
    assert np.array_equal(pal[:, 2], pal[:, 0]) and np.all(pal[:, 3] == 0xFF)
    assert len(data) == 1078 + datalen
    # the whole image, bottom-up, zero padding -- not the (2,3)-(30,9) ROI
    img = ((np.arange(w * h) * 7 + 3) & 0xFF).astype(np.uint8).reshape(h, w)
    for k in range(h):
        off = 1078 + k * (w + pad)
        assert data[off:off + w] == img[h - 1 - k].tobytes()
        assert data[off + w:off + w + pad] == b"\0" * pad
