"""Throughput of the batched and mixed workloads on the current GPU.

    python -m opencuda_amd.measure c1 [--steps 20]    # one large ROI, r = 7, both modes
    python -m opencuda_amd.measure c2 [--steps 5]     # batched persistent launch
    python -m opencuda_amd.measure c4 [--steps 3]     # exact batch + surrogate, one graph
    python -m opencuda_amd.measure c3-surrogate       # the surrogate on the 8192^2 image
    python -m opencuda_amd.measure chain              # LocalCluster -> filter -> ConnectRegion

c1 (4096^2 image, ROI 3800x3600 at (100, 200), r = 7): the exact filter
in REF and STD mode, Mpix/s and algorithmic TFLOP/s (4T+1 FLOP/px).

c2 (256 images, 256..2048 per side, ROI 50..100 %, r in {3,5,9}, REF mode):
one bf_batch launch (eager and replayed from a captured graph) against the
same jobs issued as one bf_exact launch per image, all on one stream.

c4 (512 images, 128..4096 per side, r in {3,7,15}, STD mode): the mixed
graph (r >= 9 through the surrogate with sampled check and exact fallback)
against one bf_batch launch that filters every image exactly.

c3-surrogate (8192^2, r=15, STD target): fused BF16 inference alone, then
inference + sampled exact check + exact fallback of failing 64x64 tiles.

chain (4096^2 synthetic image): LocalCluster (pntRange 8, pntCount 5,
thresholds 10), the exact filter (REF, r = 7, sigma_space 5, sigma_range
25), ConnectRegion (areas 100..100000), each stage timed on its own.

Each prints one JSON line; progress goes to stderr.  Inputs are resident in
HBM before timing starts; times come from HIP events on the launch stream.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

from . import Batch, Job, Lut, MODE_REF, MODE_STD, exact, view
from . import workloads as wl


def _log(m):
    print(f"[measure] {m}", file=sys.stderr, flush=True)


def _flop(it, mode):
    T = (2 * it.radius + 1) ** 2 + (1 if mode == MODE_REF else 0)
    return it.pixels * (4 * T + 1)


def _time(fn, steps, stream):
    import torch
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(steps):
        fn()
    e1.record(stream)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / steps


def c1(steps: int = 20):
    import torch
    (it,) = wl.c1()
    tin = torch.from_numpy(it.image()).cuda()
    x1, y1, x2, y2 = it.roi
    tout = torch.zeros((y2 - y1, x2 - x1), dtype=torch.uint8, device="cuda")
    lut = Lut.from_sigmas(it.sigma_space, it.sigma_range)
    stream = torch.cuda.current_stream()
    px = (x2 - x1) * (y2 - y1)
    res = {"workload": f"c1: {x2 - x1}x{y2 - y1} ROI of a 4096^2 image, r={it.radius}",
           "pixels": px}
    for name, mode in (("ref", MODE_REF), ("std", MODE_STD)):
        f = lambda: exact(view(tin, it.roi), view(tout), lut, it.radius, mode, stream=stream)
        _time(f, 8, stream)                     # past the clock ramp
        ms = _time(f, steps, stream)
        taps = (2 * it.radius + 1) ** 2 + (1 if mode == MODE_REF else 0)
        res[name] = {"ms": round(ms, 4), "mpix_s": round(px / ms / 1e3, 1),
                     "tflops": round(px * (4 * taps + 1) / ms / 1e9, 2)}
        _log(f"c1 {name}: {ms:.3f} ms")
    return res


def c2(steps: int = 5, mode: int = MODE_REF, tune_table: str | None = None):
    import torch
    items = wl.c2()
    stream = torch.cuda.current_stream()
    luts, tin, tout, jobs = {}, [], [], []
    t0 = time.time()
    for it in items:
        key = (it.sigma_space, it.sigma_range)
        if key not in luts:
            luts[key] = Lut.from_sigmas(*key)
        tin.append(torch.from_numpy(it.image()).cuda())
        x1, y1, x2, y2 = it.roi
        tout.append(torch.zeros((y2 - y1, x2 - x1), dtype=torch.uint8, device="cuda"))
        jobs.append(Job(view(tin[-1], it.roi), view(tout[-1]), luts[key].handle, it.radius, mode))
    _log(f"{len(items)} images uploaded in {time.time() - t0:.1f} s")
    px = sum(it.pixels for it in items)
    flop = sum(_flop(it, mode) for it in items)

    def per_image():
        for it, j in zip(items, jobs):
            exact(j.inp, j.out, luts[(it.sigma_space, it.sigma_range)], it.radius, mode,
                  stream=stream)

    b = Batch(jobs)
    for _ in range(2):
        b.run(stream)
        b.run_graph(stream)
        per_image()
    torch.cuda.synchronize()
    ms_batch = _time(lambda: b.run(stream), steps, stream)
    ms_graph = _time(lambda: b.run_graph(stream), steps, stream)
    ms_loop = _time(per_image, steps, stream)
    _log(f"batch {ms_batch:.3f} ms, graph {ms_graph:.3f} ms, per-image launches {ms_loop:.3f} ms")
    tuned = None
    if tune_table:
        from . import autotune
        n = autotune.load(tune_table)
        bt = Batch(jobs)
        bt.run_graph(stream)
        torch.cuda.synchronize()
        ms_tuned = _time(lambda: bt.run_graph(stream), steps, stream)
        best, times = autotune.tune_batch(jobs, reps=steps, stream=stream)
        bb = Batch(jobs, best)
        bb.run_graph(stream)
        torch.cuda.synchronize()
        ms_best = _time(lambda: bb.run_graph(stream), steps, stream)
        tuned = {"entries": n, "table_config": bt.config, "graph_ms": round(ms_tuned, 4),
                 "graph_mpix_s": round(px / ms_tuned / 1e3, 1),
                 "batch_tuned": {"config": best,
                                 "ms_per_config": [None if t is None else round(t, 4)
                                                   for t in times],
                                 "graph_ms": round(ms_best, 4),
                                 "graph_mpix_s": round(px / ms_best / 1e3, 1)}}
        _log(f"with the tuned table ({n} entries, config {bt.config}): graph {ms_tuned:.3f} ms; "
             f"batch tuned on itself: config {best}, graph {ms_best:.3f} ms")
    by_r = {}
    for it in items:
        by_r[it.radius] = by_r.get(it.radius, 0) + 1
    return {
        "workload": "c2: 256 images 256..2048 per side, ROI 50..100 %, r in {3,5,9}, "
                    f"{'REF' if mode == MODE_REF else 'STD'} mode",
        "images": len(items), "images_per_radius": by_r, "pixels": px,
        "batch_ms": round(ms_batch, 4), "graph_ms": round(ms_graph, 4),
        "per_image_launches_ms": round(ms_loop, 4),
        "batch_mpix_s": round(px / ms_batch / 1e3, 1),
        "graph_mpix_s": round(px / ms_graph / 1e3, 1),
        "per_image_mpix_s": round(px / ms_loop / 1e3, 1),
        "batch_tflops": round(flop / ms_batch / 1e9, 2),
        "batch_vs_per_image": round(ms_loop / ms_batch, 3),
        "tuned": tuned,
    }


def c4(count: int = 512, steps: int = 3, mode: int = MODE_STD, streams: int = 4,
       weights: str | None = None, sur_min_pixels: int = 0, max_fallback: float | None = None):
    import torch
    from .mixed import measure
    out, mx = measure(count, steps, mode, log=_log, streams=streams, weights=weights,
                      sur_min_pixels=sur_min_pixels, max_fallback=max_fallback)
    stream = torch.cuda.current_stream()
    jobs = [Job(view(mx.tin[i], it.roi), view(mx.tout[i]),
                mx.luts[(it.sigma_space, it.sigma_range)].handle, it.radius, mode)
            for i, it in enumerate(mx.items)]
    b = Batch(jobs)
    b.run_graph(stream)
    torch.cuda.synchronize()
    ms_exact = _time(lambda: b.run_graph(stream), steps, stream)
    sur_px = mx.sur_pixels
    out.update({
        "all_exact_graph_ms": round(ms_exact, 3),
        "all_exact_mpix_s": round(mx.pixels / ms_exact / 1e3, 1),
        "mixed_vs_all_exact": round(ms_exact / out["ms_per_graph"], 4),
        "surrogate_pixels": sur_px,
    })
    _log(f"mixed {out['ms_per_graph']} ms, all exact {ms_exact:.3f} ms")
    return out


def c3_surrogate(steps: int = 5, weights: str | None = None):
    import torch
    from .surrogate import DEFAULT_WEIGHTS, Surrogate
    (it,) = wl.c3()
    img = it.image()
    tin = torch.from_numpy(img).cuda()
    tout = torch.zeros_like(tin)
    lut = Lut.from_sigmas(it.sigma_space, it.sigma_range)
    net = Surrogate(seed=1)
    net.load(weights or DEFAULT_WEIGHTS)
    stream = torch.cuda.current_stream()
    vin, vout = view(tin), view(tout)
    for _ in range(2):
        net.infer(vin, vout, it.sigma_range, stream=stream)
    torch.cuda.synchronize()
    ms_inf = _time(lambda: net.infer(vin, vout, it.sigma_range, stream=stream), steps, stream)
    total, fb = net.filter(vin, vout, lut, it.radius, MODE_STD, it.sigma_range, 2, stream=stream)
    ms_filt = _time(lambda: net.filter(vin, vout, lut, it.radius, MODE_STD, it.sigma_range, 2,
                                       stream=stream, counts=False), steps, stream)
    px = it.pixels
    _log(f"inference {ms_inf:.3f} ms, with check and fallback {ms_filt:.3f} ms "
         f"({fb}/{total} tiles fell back)")
    return {
        "workload": "c3: 8192x8192, r=15, sigma_range=30, surrogate (STD target)",
        "pixels": px, "inference_ms": round(ms_inf, 4),
        "inference_mpix_s": round(px / ms_inf / 1e3, 1),
        "inference_tflops_bf16": round(75456.0 * px / ms_inf / 1e9, 1),
        "filter_ms": round(ms_filt, 4), "filter_mpix_s": round(px / ms_filt / 1e3, 1),
        "tiles": total, "tiles_fallback": fb,
    }


def chain(steps: int = 5, size: int = 4096):
    import torch
    from . import alloc_range, connect_region, local_cluster
    img = wl.Item(size, size, 400, 16, 512, 10.0, wl.sm64(5), (0, 0, size, size), 7, 5.0,
                  25.0).image()
    tin = torch.from_numpy(img).cuda()
    lc, bf, cr = (torch.zeros_like(tin) for _ in range(3))
    lut = Lut.from_sigmas(5.0, 25.0)
    stream = torch.cuda.current_stream()
    full = (0, 0, size, size)
    stages = {
        "local_cluster": lambda: local_cluster(view(tin), view(lc), alloc=alloc_range(tin, full),
                                               stream=stream),
        "bilateral_filter": lambda: exact(view(lc), view(bf), lut, 7, MODE_REF, stream=stream),
        "connect_region": lambda: connect_region(view(bf), view(cr), 100, 100000, stream=stream),
    }
    for f in stages.values():
        f()
    torch.cuda.synchronize()
    px = size * size
    res = {"workload": f"chain on a {size}x{size} synthetic image", "pixels": px}
    total = 0.0
    for name, f in stages.items():
        ms = _time(f, steps, stream)
        total += ms
        res[name] = {"ms": round(ms, 4), "mpix_s": round(px / ms / 1e3, 1)}
        _log(f"{name}: {ms:.3f} ms")
    res["chain_ms"] = round(total, 4)
    res["regions"] = int(cr.max().item())
    return res


def _main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("workload", choices=["c1", "c2", "c4", "c3-surrogate", "chain"])
    ap.add_argument("--steps", type=int, default=0)
    ap.add_argument("--count", type=int, default=512, help="c4: number of images")
    ap.add_argument("--mode", choices=["ref", "std"], default=None)
    ap.add_argument("--tune", default=None, help="c2: also time with this autotuner table")
    ap.add_argument("--streams", type=int, default=4, help="c4: streams for the surrogate images")
    ap.add_argument("--weights", default=None, help="c4, c3-surrogate: surrogate weights (.npz)")
    ap.add_argument("--sur-min-pixels", type=int, default=0,
                    help="c4: r = 15 images smaller than this go to the exact batch")
    ap.add_argument("--max-fallback", type=float, default=None,
                    help="c4: calibrate, sending images whose fallback fraction exceeds this "
                         "to the exact batch")
    args = ap.parse_args()
    import torch
    torch.cuda.set_device(0)
    if args.workload == "c1":
        res = c1(args.steps or 20)
    elif args.workload == "c3-surrogate":
        res = c3_surrogate(args.steps or 5, args.weights)
    elif args.workload == "chain":
        res = chain(args.steps or 5)
    elif args.workload == "c2":
        mode = MODE_STD if args.mode == "std" else MODE_REF
        res = c2(args.steps or 5, mode, args.tune)
    else:
        mode = MODE_REF if args.mode == "ref" else MODE_STD
        res = c4(args.count, args.steps or 3, mode, args.streams, args.weights,
                 args.sur_min_pixels, args.max_fallback)
    print(json.dumps(res), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(_main())
