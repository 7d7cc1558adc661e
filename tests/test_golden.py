"""Committed golden fixtures (tests/golden/fixtures.npz, written by
tests/golden/make_golden.py): the oracle and the product's table generators
reproduce them on the CPU; the HIP path reproduces them on the GPU."""
import os

import numpy as np
import pytest

import opencuda_amd as oc
import oracle

PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "fixtures.npz")


def _cases():
    with np.load(PATH, allow_pickle=False) as z:
        names = sorted({k.split("/")[0] for k in z.files})
        return [(n, {f: z[f"{n}/{f}"] for f in ("input", "params", "sigmas", "expected",
                                                    "gauss_bits", "range_bits")}) for n in names]


CASES = _cases()


def _unpack(c):
    x1, y1, x2, y2, ow, oh, r, mode = (int(v) for v in c["params"])
    osz = None if ow < 0 else (ow, oh)
    ss, sr = (float(v) for v in c["sigmas"])
    return (x1, y1, x2, y2), osz, r, mode, ss, sr


@pytest.mark.parametrize("name,c", CASES, ids=[n for n, _ in CASES])
def test_oracle_reproduces_fixture(name, c):
    roi, osz, r, mode, ss, sr = _unpack(c)
    got = oracle.filter(c["input"], r, ss, sr, mode=mode, roi=roi, out_size=osz)
    assert np.array_equal(got, c["expected"])


@pytest.mark.parametrize("name,c", CASES, ids=[n for n, _ in CASES])
def test_product_tables_match_fixture(name, c):
    roi, osz, r, mode, ss, sr = _unpack(c)
    _, g = oc.gauss_template(ss)
    assert np.array_equal(g.view(np.uint32), c["gauss_bits"])
    assert np.array_equal(oc.euclid_template(sr).view(np.uint32), c["range_bits"])


@pytest.mark.gpu
@pytest.mark.parametrize("name,c", CASES, ids=[n for n, _ in CASES])
def test_hip_reproduces_fixture(device, name, c):
    from tests._gpu import run_exact
    roi, osz, r, mode, ss, sr = _unpack(c)
    got = run_exact(c["input"], r, ss, sr, mode, roi=roi, out_size=osz, pad_in=3)
    eh, ew = c["expected"].shape
    assert np.array_equal(got[:eh, :ew], c["expected"])
