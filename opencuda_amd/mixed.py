"""Mixed workload: exact filter for small radii and the learned surrogate
with exact fallback for large radii, all in one captured graph.

Images with r >= 9 (here r = 15, the radius the shipped network is trained
for) run bf_sur_filter_async (surrogate + sampled exact check + exact
fallback of failing 64x64 tiles); the others run in one persistent
bf_batch launch.  The batch and the surrogate images are spread over
several streams forked from the launching stream (each stream with its own
network instance, whose check/fallback workspace it uses), so small images
run side by side instead of one launch after another; the whole sequence is
captured once into a graph and replayed.  Every image uses the same filter
mode (STD by default: the mode the network approximates), so "exact" means
the same thing for both groups.  Optional routing: `sur_min_pixels` sends
smaller r >= 9 images to the batch; `max_fallback` filters every r >= 9
image once through the network at construction (one timed multi-stream
pass, `calibration_ms`) and sends those whose check failed more than that
fraction of their tiles to the batch.  A single pass over fresh images then
costs the calibration pass plus the exact batch (`one_pass_ms` in
`measure`), not the graph replay alone.

    python -m opencuda_amd.mixed [--count 512] [--steps 3] [--max-fallback 0.33]
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from . import Batch, Job, Lut, MODE_REF, MODE_STD, exact, view
from . import workloads as wl
from .surrogate import DEFAULT_WEIGHTS, Surrogate


class Mixed:
    def __init__(self, items, net: Surrogate, mode: int = MODE_STD, sur_min_radius: int = 9,
                 tol: int = 2, streams: int = 4, log=None, sur_min_pixels: int = 0,
                 max_fallback: float | None = None):
        import torch
        self.items = items
        self.mode = mode
        self.tol = tol
        self.net = net
        # one network instance (same weights) per stream: each owns the
        # check/fallback workspace its calls use
        params = net.get_params()
        self.nets = [net]
        for _ in range(max(1, streams) - 1):
            n = Surrogate(seed=1)
            n.set_params(params)
            self.nets.append(n)
        self.streams = [torch.cuda.Stream() for _ in range(len(self.nets))]
        self.batch_stream = torch.cuda.Stream()
        self.luts = {}
        self.tin, self.tout = [], []
        self.sur_idx, self.ex_idx = [], []
        t0 = time.time()
        for i, it in enumerate(items):
            img = it.image()
            x1, y1, x2, y2 = it.roi
            self.tin.append(torch.from_numpy(img).cuda())
            self.tout.append(torch.zeros((y2 - y1, x2 - x1), dtype=torch.uint8, device="cuda"))
            key = (it.sigma_space, it.sigma_range)
            if key not in self.luts:
                self.luts[key] = Lut.from_sigmas(*key)
            to_sur = it.radius >= sur_min_radius and it.pixels >= sur_min_pixels
            (self.sur_idx if to_sur else self.ex_idx).append(i)
            if log and (i + 1) % 128 == 0:
                log(f"{i + 1}/{len(items)} images uploaded ({time.time() - t0:.1f} s)")
        # Calibration (max_fallback set): an image whose sampled check sends
        # more than max_fallback of its tiles to the exact filter costs more
        # through the network than exactly, so it joins the exact batch.
        self.calibration = {}
        self.calibration_ms = None
        if max_fallback is not None:
            self._calibrate(max_fallback)
        jobs = [Job(view(self.tin[i], items[i].roi), view(self.tout[i]),
                    self.luts[(items[i].sigma_space, items[i].sigma_range)].handle,
                    items[i].radius, mode) for i in self.ex_idx]
        self.batch = Batch(jobs) if jobs else None
        self.fallback = torch.zeros(1, dtype=torch.int32, device="cuda")
        self.tiles = torch.zeros(1, dtype=torch.int32, device="cuda")
        self.graph = None
        self.pixels = sum(it.pixels for it in items)
        self.sur_pixels = sum(items[i].pixels for i in self.sur_idx)

    def _calibrate(self, max_fallback: float) -> None:
        """One pass of every network candidate through the network (check and
        fallback included, spread over the streams like the graph), each
        image's fallback count into its own device counter; then the images
        whose check failed more than max_fallback of their tiles join the
        exact batch.  The pass already leaves the final output of the images
        that stay on the network.  calibration_ms: its GPU time (events on
        the launching stream around the fork/join), host routing excluded."""
        import torch
        cand = sorted(self.sur_idx, key=lambda i: -self.items[i].pixels)
        fb = torch.zeros(max(1, len(cand)), dtype=torch.int32, device="cuda")
        tiles = torch.zeros_like(fb)
        main = torch.cuda.current_stream()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        # twice: the first pass grows each network's workspace to its largest
        # image; the second is timed (same counts: the check is deterministic)
        for rep in range(2):
            fb.zero_()
            tiles.zero_()
            e0.record(main)
            for st in self.streams:
                st.wait_stream(main)
            for k, i in enumerate(cand):
                it = self.items[i]
                j = k % len(self.nets)
                self.nets[j].filter_async(view(self.tin[i], it.roi), view(self.tout[i]),
                                          self.luts[(it.sigma_space, it.sigma_range)], it.radius,
                                          self.mode, it.sigma_range, self.tol, fb[k:k + 1],
                                          tiles[k:k + 1], self.streams[j])
            for st in self.streams:
                main.wait_stream(st)
            e1.record(main)
            torch.cuda.synchronize()
        self.calibration_ms = e0.elapsed_time(e1)
        fbh, th = fb.cpu().tolist(), tiles.cpu().tolist()
        keep = []
        for k, i in enumerate(cand):
            self.calibration[i] = (th[k], fbh[k])
            (keep if fbh[k] <= max_fallback * th[k] else self.ex_idx).append(i)
        self.sur_idx = sorted(keep)
        self.ex_idx.sort()

    def _sur_order(self):
        # largest first, so the network's workspace reaches its final size
        # on the first image of the eager warm-up
        return sorted(self.sur_idx, key=lambda i: -self.items[i].pixels)

    def run(self, stream=None):
        """The sequence the graph holds: the batch launch on its own stream,
        the surrogate images round-robin over the network streams, all
        forked from and joined back into `stream` (default: current)."""
        import torch
        main = stream if stream is not None else torch.cuda.current_stream()
        side = [self.batch_stream] + self.streams
        for st in side:
            st.wait_stream(main)
        if self.batch is not None:
            self.batch.run(self.batch_stream)
        for k, i in enumerate(self._sur_order()):
            it = self.items[i]
            j = k % len(self.nets)
            self.nets[j].filter_async(view(self.tin[i], it.roi), view(self.tout[i]),
                                      self.luts[(it.sigma_space, it.sigma_range)], it.radius,
                                      self.mode, it.sigma_range, self.tol, self.fallback,
                                      self.tiles, self.streams[j])
        for st in side:
            main.wait_stream(st)

    def capture(self):
        import torch
        self.run()                      # eager warm-up: grows every workspace
        torch.cuda.synchronize()
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            self.run()
        torch.cuda.synchronize()

    def replay(self):
        self.graph.replay()

    def reset_counts(self):
        self.fallback.zero_()
        self.tiles.zero_()

    def counts(self):
        import torch
        torch.cuda.synchronize()
        return int(self.tiles.item()), int(self.fallback.item())

    def exact_reference(self, i):
        """The exact filter of image i (same mode), computed on the GPU."""
        import torch
        it = self.items[i]
        ref = torch.zeros_like(self.tout[i])
        exact(view(self.tin[i], it.roi), view(ref), self.luts[(it.sigma_space, it.sigma_range)],
              it.radius, self.mode)
        torch.cuda.synchronize()
        return ref


def measure(count=512, steps=3, mode=MODE_STD, log=None, streams=4, weights=None,
            sur_min_pixels=0, max_fallback=None):
    import torch
    net = Surrogate(seed=1)
    net.load(weights or DEFAULT_WEIGHTS)
    mx = Mixed(wl.c4(count), net, mode=mode, streams=streams, log=log,
               sur_min_pixels=sur_min_pixels, max_fallback=max_fallback)
    if log:
        log(f"{len(mx.ex_idx)} exact images, {len(mx.sur_idx)} surrogate images, "
            f"{mx.pixels / 1e6:.1f} Mpx; capturing")
    mx.capture()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    mx.replay()
    torch.cuda.synchronize()
    mx.reset_counts()
    e0.record()
    for _ in range(steps):
        mx.replay()
    e1.record()
    tiles, fb = mx.counts()
    ms = e0.elapsed_time(e1) / steps
    # errors of the surrogate images against the exact filter (all pixels)
    errs, within = [], []
    for i in mx.sur_idx:
        got = mx.tout[i].cpu().numpy().astype(np.int16)
        ref = mx.exact_reference(i).cpu().numpy().astype(np.int16)
        d = np.abs(got - ref)
        errs.append((int(d.max()), float(d.sum()), d.size))
        within.append(int((d <= 2).sum()))
    npx = sum(e[2] for e in errs)
    # A single pass over fresh images: with routing by radius the graph is
    # that pass; with calibrated routing it is the calibration pass (which
    # already produced the final output of the images kept on the network)
    # plus the exact batch over every other image.
    if max_fallback is None:
        one_pass = ms
    else:
        bms = 0.0
        if mx.batch is not None:
            stream = torch.cuda.current_stream()
            mx.batch.run(stream)
            torch.cuda.synchronize()
            e0.record(stream)
            for _ in range(steps):
                mx.batch.run(stream)
            e1.record(stream)
            torch.cuda.synchronize()
            bms = e0.elapsed_time(e1) / steps
        one_pass = mx.calibration_ms + bms
    out = {
        "workload": f"c4: {count} images 128..4096 per side, r in {{3,7,15}}, "
                    f"{'STD' if mode == MODE_STD else 'REF'} mode",
        "images": count, "exact_images": len(mx.ex_idx), "surrogate_images": len(mx.sur_idx),
        "streams": len(mx.nets), "sur_min_pixels": sur_min_pixels,
        "max_fallback": max_fallback,
        "pixels": mx.pixels, "ms_per_graph": round(ms, 3),
        "one_pass_ms": round(one_pass, 3),
        "one_pass_mpix_s": round(mx.pixels / (one_pass / 1e3) / 1e6, 1),
        "calibration_ms": None if mx.calibration_ms is None else round(mx.calibration_ms, 3),
        "mpix_s": round(mx.pixels / (ms / 1e3) / 1e6, 1),
        "tiles_checked": tiles // steps, "tiles_fallback": fb // steps,
        "fallback_fraction": round(fb / max(tiles, 1), 4),
        "sur_err_vs_exact": {"max": max((e[0] for e in errs), default=0),
                             "mean": round(sum(e[1] for e in errs) / max(npx, 1), 4),
                             "within_2lsb": round(sum(within) / max(npx, 1), 6)},
    }
    return out, mx


def _main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--count", type=int, default=512)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--mode", choices=["std", "ref"], default="std")
    ap.add_argument("--streams", type=int, default=4)
    ap.add_argument("--sur-min-pixels", type=int, default=0)
    ap.add_argument("--max-fallback", type=float, default=None)
    args = ap.parse_args()

    def log(m):
        print(f"[mixed] {m}", file=sys.stderr, flush=True)

    res, _ = measure(args.count, args.steps, MODE_STD if args.mode == "std" else MODE_REF, log=log,
                     streams=args.streams, sur_min_pixels=args.sur_min_pixels,
                     max_fallback=args.max_fallback)
    print(json.dumps(res), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(_main())
