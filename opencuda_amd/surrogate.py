"""Learned BF16 surrogate for large radii: binding, training data, trainer.

The network and all its kernels (forward, backward, Adam, inference, the
sampled exact check and the exact fallback) live in libopencuda_bf.so
(csrc/surrogate.hip); this module drives them with torch-owned device
memory.

Training data: seeded synthetic images (bf_synth_image), filtered on the
GPU by the exact kernel (bf_exact, STD mode by default) at r=15,
sigma_space=10, sigma_range uniform in [15, 45]; batches of random crops.

    python -m opencuda_amd.surrogate --out opencuda_amd/data/surrogate_r15.npz
"""
from __future__ import annotations

import argparse
import ctypes
import os
import sys
import time

import numpy as np

from . import (BfError, Lut, MODE_REF, MODE_STD, View, _check, _stream, exact, lib,
               synth_image, view)

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_WEIGHTS = os.path.join(HERE, "data", "surrogate_r15.npz")


class Surrogate:
    """bf_sur: FP32 master weights + Adam state on the device."""

    def __init__(self, seed: int = 1):
        h = ctypes.c_void_p()
        _check(lib().bf_sur_create(int(seed), ctypes.byref(h)), "bf_sur_create")
        self.handle = h
        self.nparams = int(lib().bf_sur_param_count())

    # ---- parameters
    def get_params(self) -> np.ndarray:
        p = np.zeros(self.nparams, np.float32)
        _check(lib().bf_sur_get_params(self.handle, p.ctypes.data), "bf_sur_get_params")
        return p

    def set_params(self, p: np.ndarray) -> None:
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == self.nparams
        _check(lib().bf_sur_set_params(self.handle, p.ctypes.data), "bf_sur_set_params")

    def save(self, path: str, **meta) -> None:
        os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
        np.savez(path, params=self.get_params(), **{k: np.asarray(v) for k, v in meta.items()})

    def load(self, path: str) -> dict:
        with np.load(path, allow_pickle=False) as z:
            self.set_params(z["params"])
            return {k: z[k] for k in z.files if k != "params"}

    # ---- training
    def train_step(self, x, y, sigma, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8,
                   loss=None, stream=None, edges=None) -> None:
        """x, y: uint8 device tensors (n, h, w); sigma: float32 device (n);
        edges: optional uint8 device (n), the crop sides on an image edge
        (bit 0 top, 1 bottom, 2 left, 3 right), trained without margin."""
        n, h, w = x.shape
        lp = ctypes.c_void_p(loss.data_ptr()) if loss is not None else None
        ep = ctypes.c_void_p(edges.data_ptr()) if edges is not None else None
        _check(lib().bf_sur_train_step(self.handle, ctypes.c_void_p(x.data_ptr()),
                                       ctypes.c_void_p(y.data_ptr()),
                                       ctypes.c_void_p(sigma.data_ptr()), ep, n, h, w, lr, beta1,
                                       beta2, eps, lp, _stream(stream)), "bf_sur_train_step")

    def set_loss_tail(self, weight: float, start_lsb: float) -> None:
        """Loss of train_step and forward_backward: MSE + weight *
        max(|pred - y| - start_lsb/255, 0)^2 per pixel (weight 0: MSE)."""
        _check(lib().bf_sur_set_loss_tail(self.handle, weight, start_lsb), "bf_sur_set_loss_tail")

    def forward_backward(self, x, y, sigma, stream=None, edges=None):
        """Returns (pred float32 (n,h,w), grads float32 (nparams), loss) tensors."""
        import torch
        n, h, w = x.shape
        pred = torch.empty((n, h, w), dtype=torch.float32, device=x.device)
        grads = torch.empty(self.nparams, dtype=torch.float32, device=x.device) if y is not None else None
        loss = torch.zeros(1, dtype=torch.float32, device=x.device)
        _check(lib().bf_sur_forward_backward(
            self.handle, ctypes.c_void_p(x.data_ptr()),
            ctypes.c_void_p(y.data_ptr()) if y is not None else None,
            ctypes.c_void_p(sigma.data_ptr()),
            ctypes.c_void_p(edges.data_ptr()) if edges is not None else None,
            n, h, w, ctypes.c_void_p(pred.data_ptr()),
            ctypes.c_void_p(grads.data_ptr()) if grads is not None else None,
            ctypes.c_void_p(loss.data_ptr()), _stream(stream)), "bf_sur_forward_backward")
        return pred, grads, loss

    # ---- inference
    def infer(self, vin: View, vout: View, sigma_range: float, stream=None) -> None:
        _check(lib().bf_sur_infer(self.handle, ctypes.byref(vin), ctypes.byref(vout),
                                  float(sigma_range), _stream(stream)), "bf_sur_infer")

    def filter(self, vin: View, vout: View, lut: Lut, radius: int, mode: int, sigma_range: float,
               tol_lsb: int = 2, stream=None, counts: bool = True):
        """Surrogate + sampled exact check + exact fallback of failing 64x64
        tiles.  Returns (tiles_total, tiles_fallback) when counts=True (that
        synchronises the stream)."""
        tt, tf = ctypes.c_int32(), ctypes.c_int32()
        _check(lib().bf_sur_filter(self.handle, ctypes.byref(vin), ctypes.byref(vout), lut.handle,
                                   int(radius), int(mode), float(sigma_range), int(tol_lsb),
                                   ctypes.byref(tt) if counts else None,
                                   ctypes.byref(tf) if counts else None, _stream(stream)),
               "bf_sur_filter")
        return (tt.value, tf.value) if counts else None

    def filter_async(self, vin: View, vout: View, lut: Lut, radius: int, mode: int,
                     sigma_range: float, tol_lsb: int = 2, fallback_accum=None,
                     tiles_accum=None, stream=None) -> None:
        """bf_sur_filter_async: no host synchronisation (graph-capturable);
        fallback/tile counts are added to the given device int32 tensors."""
        fa = ctypes.c_void_p(fallback_accum.data_ptr()) if fallback_accum is not None else None
        ta = ctypes.c_void_p(tiles_accum.data_ptr()) if tiles_accum is not None else None
        _check(lib().bf_sur_filter_async(self.handle, ctypes.byref(vin), ctypes.byref(vout),
                                         lut.handle, int(radius), int(mode), float(sigma_range),
                                         int(tol_lsb), fa, ta, _stream(stream)),
               "bf_sur_filter_async")

    def close(self):
        if self.handle:
            lib().bf_sur_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ------------------------------------------------------------ training data

class Dataset:
    """Seeded synthetic images and their exact-filter targets, on the GPU."""

    def __init__(self, n_images: int = 2000, size: int = 512, radius: int = 15,
                 sigma_space: float = 10.0, sr_range=(15.0, 45.0), mode: int = MODE_STD,
                 seed: int = 2024, log=None, c3_frac: float = 0.0):
        """c3_frac: fraction of the images drawn with the 8192^2 benchmark
        image's generator settings scaled to `size` (its shape density,
        shape sizes 16..1024, noise 12), the others with random settings
        (10-60 shapes, sizes from 8-24 up to 16-320, noise 10 or 12)."""
        import torch
        rng = np.random.default_rng(seed)
        self.size = size
        self.radius = radius
        self.mode = mode
        self.x = torch.empty((n_images, size, size), dtype=torch.uint8, device="cuda")
        self.y = torch.empty_like(self.x)
        self.sigma = torch.empty(n_images, dtype=torch.float32, device="cuda")
        srs = rng.uniform(sr_range[0], sr_range[1], n_images).astype(np.float32)
        luts = {}
        t0 = time.time()
        for i in range(n_images):
            nshapes = int(rng.integers(10, 60))
            smin = int(rng.integers(8, 24))
            smax = int(rng.integers(smin + 8, 320))
            noise = float(rng.choice([10.0, 12.0]))
            if rng.random() < c3_frac:
                nshapes = max(1, round(1600 * size * size / (8192 * 8192)))
                smin, smax, noise = 16, 1024, 12.0
            img = synth_image(size, size, nshapes, smin, smax, noise, int(rng.integers(1 << 62)))
            self.x[i].copy_(torch.from_numpy(img))
            sr = float(srs[i])
            key = round(sr, 3)
            if key not in luts:
                luts[key] = Lut.from_sigmas(sigma_space, sr)
            exact(view(self.x[i]), view(self.y[i]), luts[key], radius, mode)
            if log and (i + 1) % 250 == 0:
                log(f"dataset {i + 1}/{n_images} ({time.time() - t0:.1f} s)")
        self.sigma.copy_(torch.from_numpy(srs))
        torch.cuda.synchronize()
        self.luts = luts

    def batch(self, rng: np.random.Generator, n: int = 64, crop: int = 256,
              edge_frac: float = 0.0):
        """n random crops; with probability edge_frac a crop's row (and,
        independently, its column) offset is snapped to the image's first or
        last position, so image edges and corners are trained.  Returns
        x, y, sigma and the edge bits of each crop (bit 0 top, 1 bottom,
        2 left, 3 right)."""
        import torch
        idx = torch.from_numpy(rng.integers(0, self.x.shape[0], n)).cuda()
        last = self.size - crop
        oy = rng.integers(0, last + 1, n)
        ox = rng.integers(0, last + 1, n)
        if edge_frac > 0:
            sy = rng.random(n) < edge_frac
            sx = rng.random(n) < edge_frac
            oy = np.where(sy, np.where(rng.random(n) < 0.5, 0, last), oy)
            ox = np.where(sx, np.where(rng.random(n) < 0.5, 0, last), ox)
        edges = ((oy == 0) * 1 + (oy == last) * 2 + (ox == 0) * 4 + (ox == last) * 8).astype(np.uint8)
        ar = torch.arange(crop, device="cuda")
        rows = (torch.from_numpy(oy).cuda()[:, None] + ar[None, :])[:, :, None]
        cols = (torch.from_numpy(ox).cuda()[:, None] + ar[None, :])[:, None, :]
        ii = idx[:, None, None]
        return (self.x[ii, rows, cols].contiguous(), self.y[ii, rows, cols].contiguous(),
                self.sigma[idx].contiguous(), torch.from_numpy(edges).cuda())


def train(net: Surrogate, data: Dataset, steps: int = 1500, batch: int = 64, crop: int = 256,
          lr: float = 1e-3, seed: int = 7, log=None, log_every: int = 100, cosine: bool = False,
          edge_frac: float = 0.0):
    """Adam on random crops (learning rate `lr`, or `lr` decayed to 0 along
    a half cosine when `cosine`; `edge_frac` as in Dataset.batch); returns
    the list of (step, loss)."""
    import math

    import torch
    rng = np.random.default_rng(seed)
    loss = torch.zeros(1, dtype=torch.float32, device="cuda")
    hist = []
    t0 = time.time()
    for s in range(1, steps + 1):
        x, y, sg, edges = data.batch(rng, batch, crop, edge_frac)
        lr_s = lr * 0.5 * (1.0 + math.cos(math.pi * (s - 1) / steps)) if cosine else lr
        logged = s % log_every == 0 or s == 1 or s == steps
        # the loss is summed (atomics on one float) only on the logged steps
        net.train_step(x, y, sg, lr=lr_s, loss=loss if logged else None,
                       edges=edges if edge_frac > 0 else None)
        if logged:
            v = float(loss.item())
            hist.append((s, v))
            if log:
                log(f"step {s} loss {v:.3e} rmse {255 * v ** 0.5:.3f} LSB ({time.time() - t0:.1f} s)")
    torch.cuda.synchronize()
    return hist


def evaluate(net: Surrogate, img: np.ndarray, sigma_space: float, sigma_range: float,
             radius: int = 15, mode: int = MODE_STD, tol: int = 2) -> dict:
    """Accuracy of the pure surrogate against the exact filter on one image:
    fraction of pixels within `tol`, fraction of 64x64 tiles entirely within
    `tol`, and the fraction of tiles the sampled check lets through."""
    import torch
    lut = Lut.from_sigmas(sigma_space, sigma_range)
    tin = torch.from_numpy(img).cuda()
    ex, su, sf = torch.zeros_like(tin), torch.zeros_like(tin), torch.zeros_like(tin)
    exact(view(tin), view(ex), lut, radius, mode)
    net.infer(view(tin), view(su), sigma_range)
    total, fb = net.filter(view(tin), view(sf), lut, radius, mode, sigma_range, tol)
    torch.cuda.synchronize()
    d = (su.to(torch.int16) - ex.to(torch.int16)).abs().cpu().numpy()
    H, W = d.shape
    th, tw = (H + 63) // 64, (W + 63) // 64
    pad = np.zeros((th * 64, tw * 64), d.dtype)
    pad[:H, :W] = d
    tile_max = pad.reshape(th, 64, tw, 64).max(axis=(1, 3))
    return {"within": float((d <= tol).mean()), "mean_abs": float(d.mean()), "max": int(d.max()),
            "tiles_all_within": float((tile_max <= tol).mean()),
            "tiles_passing_check": 1.0 - fb / total, "tiles": total}


def _main() -> int:
    ap = argparse.ArgumentParser(description="train the r=15 surrogate on the GPU")
    ap.add_argument("--images", type=int, default=2000)
    ap.add_argument("--size", type=int, default=512)
    ap.add_argument("--steps", type=int, default=1500)
    ap.add_argument("--batch", type=int, default=64)
    ap.add_argument("--crop", type=int, default=256)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--mode", choices=["std", "ref"], default="std")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--out", default=DEFAULT_WEIGHTS)
    ap.add_argument("--cosine", action="store_true", help="decay lr along a half cosine")
    ap.add_argument("--eval", action="store_true", help="report accuracy on the c3 image")
    ap.add_argument("--init", default=None, help="start from these weights (.npz)")
    ap.add_argument("--log-every", type=int, default=100)
    ap.add_argument("--edge-frac", type=float, default=0.0,
                    help="fraction of crop offsets snapped to an image edge (trained unmasked)")
    ap.add_argument("--c3-frac", type=float, default=0.0,
                    help="fraction of training images with the 8192^2 benchmark image's settings")
    ap.add_argument("--data-seed", type=int, default=2024, help="seed of the training images")
    ap.add_argument("--crop-seed", type=int, default=7, help="seed of the crop sequence")
    ap.add_argument("--tail-weight", type=float, default=0.0,
                    help="weight of the loss term max(|err| - tail-start, 0)^2 (0: plain MSE)")
    ap.add_argument("--tail-start", type=float, default=2.0, help="start of that term, in LSB")
    args = ap.parse_args()

    def log(m):
        print(f"[surrogate] {m}", file=sys.stderr, flush=True)

    mode = MODE_STD if args.mode == "std" else MODE_REF
    data = Dataset(args.images, args.size, mode=mode, log=log, c3_frac=args.c3_frac,
                   seed=args.data_seed)
    net = Surrogate(args.seed)
    init_steps = 0
    if args.init:
        init_steps = int(net.load(args.init).get("steps", 0))
        log(f"starting from {args.init} ({init_steps} steps)")
    if args.tail_weight > 0:
        net.set_loss_tail(args.tail_weight, args.tail_start)
    hist = train(net, data, args.steps, args.batch, args.crop, args.lr, seed=args.crop_seed,
                 log=log, cosine=args.cosine, log_every=args.log_every, edge_frac=args.edge_frac)
    net.save(args.out, history=np.array(hist, np.float64), mode=mode, radius=15,
             sigma_space=10.0, images=args.images, steps=init_steps + args.steps, batch=args.batch,
             crop=args.crop, lr=args.lr, seed=args.seed, cosine=int(args.cosine),
             init_steps=init_steps, edge_frac=args.edge_frac, c3_frac=args.c3_frac,
             data_seed=args.data_seed, crop_seed=args.crop_seed, tail_weight=args.tail_weight,
             tail_start=args.tail_start)
    log(f"saved {args.out}")
    if args.eval:
        from . import workloads
        (it,) = workloads.c3()
        log(f"c3 accuracy: {evaluate(net, it.image(), it.sigma_space, it.sigma_range, mode=mode)}")
    return 0


if __name__ == "__main__":
    sys.exit(_main())
