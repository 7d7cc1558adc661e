"""Writes tests/golden/fixtures.npz: small inputs, parameters and expected
u8 outputs of the bilateral filter, with the weight tables they use.

The reference ships no test vectors for this path and cannot run here, so
these are regression pins produced by the oracle (oracle/bf_oracle.c, itself
pinned by tests/test_oracle_*.py); the product's HIP path must reproduce
them bit for bit.

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402

CASES = [
    # name, h, w, roi, out_size, radius, ss, sr, mode, content
    ("gradient_r3_ref", 48, 64, (2, 3, 61, 45), None, 3, 3.0, 30.0, 0, "synth"),
    ("gradient_r3_std", 48, 64, (2, 3, 61, 45), None, 3, 3.0, 30.0, 1, "synth"),
    ("trimmed_r7_ref", 40, 40, (0, 0, 40, 40), (33, 27), 7, 5.0, 25.0, 0, "synth"),
    ("edges_r5_ref", 37, 41, (1, 1, 40, 36), None, 5, 3.3, 10.0, 0, "checker"),
    ("tiny_sigma_r1_ref", 24, 20, (0, 0, 20, 24), None, 1, 0.05, 5.0, 0, "random"),
    ("constant_r9_std", 30, 30, (0, 0, 30, 30), None, 9, 6.0, 30.0, 1, "constant"),
    ("wide_r15_ref", 20, 96, (0, 0, 96, 20), None, 15, 10.0, 30.0, 0, "synth"),
]


def content(kind, h, w, seed):
    rng = np.random.default_rng(seed)
    if kind == "random":
        return rng.integers(0, 256, (h, w), dtype=np.uint8)
    if kind == "constant":
        return np.full((h, w), 117, np.uint8)
    if kind == "checker":
        yy, xx = np.mgrid[0:h, 0:w]
        return np.where(((yy // 5) + (xx // 7)) % 2 == 0, 20, 235).astype(np.uint8)
    base = np.linspace(0, 200, w)[None, :] + rng.normal(0, 10, (h, w))
    base[h // 4: h // 2, w // 3: 2 * w // 3] = 240
    return np.clip(np.rint(base), 0, 255).astype(np.uint8)


def main():
    out = {}
    for k, (name, h, w, roi, osz, r, ss, sr, mode, kind) in enumerate(CASES):
        img = content(kind, h, w, 100 + k)
        exp = oracle.filter(img, r, ss, sr, mode=mode, roi=roi, out_size=osz, nthreads=1)
        _, g = oracle.gauss_template(ss)
        E = oracle.euclid_template(sr)
        out[f"{name}/input"] = img
        out[f"{name}/params"] = np.array([roi[0], roi[1], roi[2], roi[3],
                                          osz[0] if osz else -1, osz[1] if osz else -1, r, mode],
                                         np.int32)
        out[f"{name}/sigmas"] = np.array([ss, sr], np.float32)
        out[f"{name}/expected"] = exp
        out[f"{name}/gauss_bits"] = g.view(np.uint32)
        out[f"{name}/range_bits"] = E.view(np.uint32)
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fixtures.npz")
    np.savez_compressed(path, **out)
    print(f"wrote {path} ({len(CASES)} cases)")


if __name__ == "__main__":
    main()
