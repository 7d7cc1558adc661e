#!/usr/bin/env python3
"""Benchmark: exact FP32 bilateral filter, r=15, on an 8192x8192 u8 image.

One step = one bf_exact launch over the whole image (67.1 Mpix, REF mode =
the reference kernel's semantics) with input and output resident in HBM.
Prints ONE JSON line (see the keys below).  Progress goes to stderr.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--cpu-seconds S]

Roofline: the kernel is FP32-VALU bound (not a contraction, so no MFMA;
2 bytes/px of HBM traffic against 3849 FLOP/px).  Algorithmic work is
4*T + 1 FLOP per pixel with T = (2r+1)^2 + 1 = 962 taps (g*E, f*p, t+=, s+=
per tap, one divide).  `achieved` = that work per launch / the average
launch time measured with HIP events on the launch stream; `peak` = the
MI355X FP32 vector nameplate (157.3 TFLOP/s, which assumes packed FMA).

cpu_baseline: the oracle (host transcription, OpenMP), rebuilt with
-march=native on this host, on a band of rows of the same image: one
warm-up, then the median of 3 runs, with omp_get_max_threads() and nproc
recorded; the band is also compared bit-exactly with the GPU output.

surrogate: the BF16 MFMA surrogate with the sampled exact check and exact
fallback of failing 64x64 tiles (tolerance 2 LSB), against the exact filter
in STD mode -- the mode the network is trained on -- with error statistics
over all pixels.

chain: the stages around the filter in the reference's region detector
(LocalCluster -> exact filter r=7 -> ConnectRegion) on a 4096^2 image, each
stage timed on the GPU; the CPU leg times the oracle's three stages on a
1024^2 image.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

# The headline metric's name as the benchmark definition states it.  Its
# "(≤2 LSB)" is the surrogate check's tolerance on sampled pixels; the
# surrogate block reports whether every pixel met it (it does not).
METRIC = "Mpix/s r=15 on 8192² u8: exact FP32 vs BF16 surrogate+fallback (≤2 LSB)"
PEAK_FP32_TFLOPS = 157.3       # MI355X FP32 vector nameplate (packed FMA)
PEAK_FP32_NOFMA_TFLOPS = 78.6  # packed mul/add, no FMA
PEAK_HBM_GBS = 8000.0


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_baseline(img, it, gpu_out, seconds: float):
    """The oracle (the host transcription, OpenMP over 4-row groups) on a
    band of rows sized so that one run takes about seconds/4: one warm-up
    run, then the median of 3 timed runs, all threads; plus a 1-thread rate.
    The oracle is rebuilt here with -march=native for this host (the shipped
    x86-64-v3 build is the fallback, named in the result).  Returns the
    baseline dict, whether the band matched the GPU output bit for bit, and
    the band's rows."""
    import statistics
    import tempfile

    import numpy as np
    import oracle
    build = {"flags": oracle.SHIPPED_FLAGS, "native": False}
    try:
        path = os.path.join(tempfile.mkdtemp(prefix="orc_native_"), "liborc_native.so")
        build["flags"] = oracle.build_native(path)
        oracle.use(path)
        build["native"] = True
    except Exception as e:     # keep the shipped build
        build["fallback"] = f"native build failed: {str(e)[:160]}"
    omp_max = oracle.max_threads()
    nthreads = max(1, min(omp_max, 16))
    H = img.shape[0]
    y0 = (H // 2) // 4 * 4

    def run(rows, nt):
        t = time.perf_counter()
        out = oracle.filter(img, it.radius, it.sigma_space, it.sigma_range, mode=oracle.REF,
                            rows=(y0, y0 + rows), nthreads=nt)
        return out, time.perf_counter() - t

    # size the band: the first call pays thread start-up and cold caches, so
    # the rate comes from a second, larger call
    run(8, nthreads)
    _, dt = run(64, nthreads)
    rows = int(max(8, min(H - y0, seconds / 4 / max(dt, 1e-6) * 64)) // 4 * 4)
    log(f"cpu baseline: {rows} rows x {img.shape[1]} px on {nthreads} threads "
        f"({'native' if build['native'] else 'x86-64-v3'} build), warm-up + 3 runs")
    band, _ = run(rows, nthreads)                     # warm-up
    times = []
    for _ in range(3):
        band, t = run(rows, nthreads)
        times.append(t)
    px = band.size
    rates = [px / t / 1e6 for t in times]
    match = bool(np.array_equal(band, gpu_out[y0:y0 + rows]))
    # 1-thread rate on a band of at least 4 rows
    r1 = max(4, rows // (4 * nthreads) // 4 * 4)
    _, dt1 = run(r1, 1)
    base = {
        "value": round(statistics.median(rates), 3), "unit": "Mpix/s", "cores": nthreads,
        "kind": "port",
        "sample": f"oracle REF r=15 on rows {y0}..{y0 + rows} of the 8192^2 image "
                  f"({px} px per run): warm-up + median of 3",
        "runs_mpix_s": [round(r, 3) for r in rates],
        "one_thread_mpix_s": round(img.shape[1] * r1 / dt1 / 1e6, 4),
        "nproc": os.cpu_count(),
        "omp_max_threads": omp_max,
        "build": build,
    }
    return base, match, (y0, y0 + rows)


def chain_cpu(size: int = 1024):
    """The oracle's LocalCluster, filter (REF, r = 7) and ConnectRegion on a
    size^2 image with the chain leg's parameters: Mpix/s per stage."""
    import numpy as np
    import oracle
    from opencuda_amd import workloads as wl
    img = wl.Item(size, size, 40, 8, 128, 10.0, wl.sm64(5), (0, 0, size, size), 7, 5.0,
                  25.0).image()
    nthreads = min(os.cpu_count() or 1, 16)
    out = {}
    t = time.perf_counter()
    lc = oracle.local_cluster(img)
    out["local_cluster"] = time.perf_counter() - t
    t = time.perf_counter()
    bf = oracle.filter(lc, 7, 5.0, 25.0, mode=oracle.REF, nthreads=nthreads)
    out["bilateral_filter"] = time.perf_counter() - t
    t = time.perf_counter()
    oracle.connect_region(bf, min_area=100, max_area=100000)
    out["connect_region"] = time.perf_counter() - t
    px = img.size
    res = {k: {"mpix_s": round(px / v / 1e6, 3)} for k, v in out.items()}
    res["bilateral_filter"]["cores"] = nthreads
    res["local_cluster"]["cores"] = 1
    res["connect_region"]["cores"] = 1
    res["sample"] = f"oracle stages on a {size}x{size} synthetic image"
    return res


def fp64_error(img, it, gpu_out, rows):
    import numpy as np
    import oracle
    v = oracle.filter_f64(img, it.radius, it.sigma_space, it.sigma_range, mode=oracle.REF,
                          rows=rows, nthreads=min(os.cpu_count() or 1, 16))
    err = gpu_out[rows[0]:rows[1]].astype(np.float64) - v
    return {"rows": list(rows), "mean": round(float(err.mean()), 4),
            "mean_abs": round(float(np.abs(err).mean()), 4),
            "max_abs": round(float(np.abs(err).max()), 4)}


def surrogate_leg(args, it, img, tin, lut, std_out, std_ms, ref_ms, stream):
    """BF16 surrogate + sampled exact check + exact fallback (tolerance 2
    LSB, 64x64 tiles, 1/64 samples) against the exact filter in STD mode."""
    import numpy as np
    import torch

    import opencuda_amd as oc
    from opencuda_amd import surrogate as sg
    if not os.path.exists(sg.DEFAULT_WEIGHTS):
        return None
    net = sg.Surrogate(seed=1)
    meta = net.load(sg.DEFAULT_WEIGHTS)
    px = it.pixels
    sout = torch.zeros_like(tin)
    vin, vs = oc.view(tin), oc.view(sout)
    total, fb = net.filter(vin, vs, lut, it.radius, oc.MODE_STD, it.sigma_range, 2, stream=stream)
    steps = max(3, args.steps // 2)
    # as many warm-up calls as the exact leg: the fused kernel's first
    # dispatches in a process run up to 20% slower while the clock settles
    warm = max(2, args.warmup)
    for _ in range(warm):
        net.filter(vin, vs, lut, it.radius, oc.MODE_STD, it.sigma_range, 2, stream=stream,
                   counts=False)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(steps):
        net.filter(vin, vs, lut, it.radius, oc.MODE_STD, it.sigma_range, 2, stream=stream,
                   counts=False)
    e1.record(stream)
    torch.cuda.synchronize()
    sf_ms = e0.elapsed_time(e1) / steps
    got = sout.cpu().numpy()
    # pure surrogate (no check, no fallback)
    pout = torch.zeros_like(tin)
    for _ in range(warm):
        net.infer(vin, oc.view(pout), it.sigma_range, stream=stream)
    e0.record(stream)
    for _ in range(steps):
        net.infer(vin, oc.view(pout), it.sigma_range, stream=stream)
    e1.record(stream)
    torch.cuda.synchronize()
    inf_ms = e0.elapsed_time(e1) / steps
    pure = pout.cpu().numpy()
    d = np.abs(got.astype(np.int16) - std_out.astype(np.int16))
    dp = np.abs(pure.astype(np.int16) - std_out.astype(np.int16))
    inf_flop = 75456.0 * px
    log(f"surrogate+fallback: {sf_ms:.3f} ms ({px / sf_ms / 1e3:.0f} Mpix/s), "
        f"{fb}/{total} tiles fell back; pure inference {inf_ms:.3f} ms")
    return {
        "mpix_s": round(px / (sf_ms / 1e3) / 1e6, 2),
        "ms": round(sf_ms, 4),
        "vs_exact_std": round(std_ms / sf_ms, 4),
        "vs_exact_ref": round(ref_ms / sf_ms, 4),
        "tiles_total": total, "tiles_fallback": fb,
        "fallback_fraction": round(fb / total, 4),
        "err_vs_exact_std": {"max": int(d.max()), "mean": round(float(d.mean()), 4),
                             "within_2lsb": round(float((d <= 2).mean()), 6)},
        "pure_inference": {"mpix_s": round(px / (inf_ms / 1e3) / 1e6, 2), "ms": round(inf_ms, 4),
                           "tflops_bf16": round(inf_flop / (inf_ms / 1e3) / 1e12, 2),
                           "err_max": int(dp.max()), "err_mean": round(float(dp.mean()), 4),
                           "within_2lsb": round(float((dp <= 2).mean()), 6)},
        "target": "exact STD mode (own window, 961 taps)",
        # The tolerance is that of a sampled check (one hashed pixel per 8x8
        # cell of each 64x64 tile), not a per-pixel bound: pixels the samples
        # miss may stay further off, as err_vs_exact_std records.
        "check": "1/64 sampled exact check per 64x64 tile, tile refiltered if any sampled "
                 "|diff| > 2 LSB",
        "every_pixel_within_2lsb": bool(d.max() <= 2),
        "pixels_beyond_2lsb": int((d > 2).sum()),
        "weights": os.path.relpath(sg.DEFAULT_WEIGHTS, ROOT),
        "trained_steps": int(meta.get("steps", 0)),
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    # the first launches of a process run slower while clocks ramp (8.8,
    # 8.2, 7.9 ms, then 7.65 ms steady): warm up past them
    ap.add_argument("--warmup", type=int, default=8)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-surrogate", action="store_true", help="skip the surrogate leg")
    ap.add_argument("--no-chain", action="store_true", help="skip the region-chain leg")
    ap.add_argument("--copy-probe", action="store_true",
                    help="first run one plain 256 MiB device-to-device copy (torch copy_): "
                         "the known-size kernel that calibrates FETCH_SIZE/WRITE_SIZE counters")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}, "
              f"WORLD_SIZE {world}): this build runs on a single GPU; the filter "
              "would shard by row bands (read-only halo, no exchange). Run with --gpus 1.",
              file=sys.stderr)
        return 2

    import numpy as np
    import torch

    import opencuda_amd as oc
    from opencuda_amd import workloads as wl

    torch.cuda.set_device(0)
    if args.copy_probe:
        a = torch.empty(256 << 20, dtype=torch.uint8, device="cuda")
        b = torch.empty_like(a)
        a.fill_(1)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b
    (it,) = wl.c3()
    log("generating the 8192^2 workload image")
    img = it.image()
    H, W = img.shape
    lut = oc.Lut.from_sigmas(it.sigma_space, it.sigma_range)
    tin = torch.from_numpy(img).cuda()
    tout = torch.zeros_like(tin)
    vin, vout = oc.view(tin), oc.view(tout)
    stream = torch.cuda.current_stream()

    log(f"warmup x{args.warmup}")
    for _ in range(args.warmup):
        oc.exact(vin, vout, lut, it.radius, oc.MODE_REF, stream=stream)
    torch.cuda.synchronize()

    log(f"timing x{args.steps}")
    starts = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        starts[k].record(stream)
        oc.exact(vin, vout, lut, it.radius, oc.MODE_REF, stream=stream)
        ends[k].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    per_launch_ms = [s.elapsed_time(e) for s, e in zip(starts, ends)]
    span_ms = starts[0].elapsed_time(ends[-1])
    ms_per_step = span_ms / args.steps
    kern_ms = float(np.mean(per_launch_ms))
    px = it.pixels
    value = px / (ms_per_step / 1e3) / 1e6
    T = (2 * it.radius + 1) ** 2 + 1
    flop_px = 4 * T + 1
    achieved = px * flop_px / (kern_ms / 1e3) / 1e12
    log(f"exact REF: {ms_per_step:.3f} ms/step, {value:.0f} Mpix/s, {achieved:.2f} TFLOP/s "
        f"(wall {wall:.3f} s)")
    gpu_out = tout.cpu().numpy()

    # STD mode (the surrogate's target) on the same image, for reference.
    for _ in range(2):
        oc.exact(vin, vout, lut, it.radius, oc.MODE_STD, stream=stream)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    nstd = max(3, args.steps // 4)
    for _ in range(nstd):
        oc.exact(vin, vout, lut, it.radius, oc.MODE_STD, stream=stream)
    e1.record(stream)
    torch.cuda.synchronize()
    std_ms = e0.elapsed_time(e1) / nstd

    std_out = tout.cpu().numpy()     # exact STD result (the surrogate's target)
    sur = None if args.no_surrogate else surrogate_leg(args, it, img, tin, lut, std_out, std_ms,
                                                       ms_per_step, stream)

    traffic = None
    tpath = os.path.join(ROOT, "profiles", "exact_c3_traffic.json")
    if os.path.exists(tpath):
        try:
            traffic = json.load(open(tpath)).get("bytes_per_launch")
        except Exception:
            traffic = None

    result = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "Mpix/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "c3: 8192x8192 u8 synthetic (1600 rects + 1600 discs, noise 12), "
                        "full ROI, r=15, sigma_space=10, sigma_range=30, exact FP32 filter "
                        "in REF mode (the reference kernel's semantics), 962 taps/px",
            "path": "bf_exact",
            "pixels": px,
            "std_mode_mpix_s": round(px / (std_ms / 1e3) / 1e6, 2),
        },
        "roofline": {
            "bound": "valu",
            "achieved": round(achieved, 3),
            "peak": PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s",
            "frac": round(achieved / PEAK_FP32_TFLOPS, 4),
            "traffic": traffic,
            "frac_of_nofma_peak": round(achieved / PEAK_FP32_NOFMA_TFLOPS, 4),
            "flop_per_px": flop_px,
            "kernel_ms": round(kern_ms, 4),
            "hbm_gbs_algorithmic": round(2 * px / (kern_ms / 1e3) / 1e9, 1),
        },
    }
    result["surrogate"] = sur
    if not args.no_chain:
        from opencuda_amd import measure
        log("chain leg (LocalCluster -> filter -> ConnectRegion, 4096^2)")
        result["chain"] = measure.chain(steps=max(3, args.steps // 4))
    if not args.no_cpu:
        log("cpu baseline (oracle)")
        base, match, rows = cpu_baseline(img, it, gpu_out, args.cpu_seconds)
        result["cpu_baseline"] = base
        result["bit_exact_sample"] = match
        result["fp64_error"] = fp64_error(img, it, gpu_out, (rows[0], rows[0] + 16))
        if not args.no_chain:
            log("cpu chain stages (oracle, 1024^2)")
            result["cpu_baseline"]["chain"] = chain_cpu()
    else:
        result["cpu_baseline"] = None
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
