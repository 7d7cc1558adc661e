"""Helpers for GPU tests: device images via torch, runs through the C ABI."""
import numpy as np

import opencuda_amd as oc

SENTINEL = 77


def to_dev(img: np.ndarray, pad: int = 0):
    import torch
    H, W = img.shape
    t = torch.full((H, W + pad), SENTINEL, dtype=torch.uint8, device="cuda")
    t[:, :W] = torch.from_numpy(np.ascontiguousarray(img)).cuda()
    return t


def out_tensor(H: int, W: int, pad: int = 0):
    import torch
    return torch.full((H, W + pad), SENTINEL, dtype=torch.uint8, device="cuda")


def run_exact(img, radius, ss, sr, mode=oc.MODE_REF, roi=None, out_size=None, config=None,
              pad_in=0, pad_out=0, lut=None):
    """Filter img's ROI into a fresh output of out_size (default: ROI size)
    prefilled with SENTINEL; returns the whole output (host)."""
    import torch
    H, W = img.shape
    roi = roi or (0, 0, W, H)
    rw, rh = roi[2] - roi[0], roi[3] - roi[1]
    ow, oh = out_size or (rw, rh)
    lut = lut or oc.Lut.from_sigmas(ss, sr)
    tin = to_dev(img, pad_in)
    tout = out_tensor(oh, ow, pad_out)
    oc.exact(oc.view(tin, roi), oc.view(tout, (0, 0, ow, oh)), lut, radius, mode, config=config)
    torch.cuda.synchronize()
    return tout[:, :ow].cpu().numpy()


def check_against_oracle(img, radius, ss, sr, mode, roi=None, out_size=None, got=None):
    import oracle
    H, W = img.shape
    roi = roi or (0, 0, W, H)
    exp = oracle.filter(img, radius, ss, sr, mode=mode, roi=roi, out_size=out_size)
    eh, ew = exp.shape
    assert np.array_equal(got[:eh, :ew], exp), (
        f"mismatch: {(got[:eh, :ew] != exp).sum()} of {exp.size} pixels")
    rest = got.copy()
    rest[:eh, :ew] = SENTINEL
    assert np.all(rest == SENTINEL), "pixels outside the filtered region were written"
    return exp
