"""Autotuner for the exact filter's tile configuration: tile shape and
columns per lane (bf_tune_radius, bf_tune_describe).

For each (radius, mode) the library times every tile configuration on the
device and keeps the fastest in a process-wide table that bf_exact and
bf_batch_create consult (a batch picks the configuration with the least
estimated total time over its jobs).  Tables can be saved to JSON and
installed again in a later process (bf_tune_set).  `tune_batch` times a
given batch itself with every configuration.

    python -m opencuda_amd.autotune --out opencuda_amd/data/tune_gfx950.json
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_TABLE = os.path.join(HERE, "data", "tune_gfx950.json")


def _lib():
    from . import lib
    return lib()


def _nconf() -> int:
    return int(_lib().bf_tune_count())


def tune(radius: int, mode: int, width: int = 4096, height: int = 2048, reps: int = 3):
    """Time every configuration; returns (best_config, ms_per_mpix list)."""
    from . import _check
    cost = np.zeros(_nconf(), np.float32)
    best = ctypes.c_int32(-1)
    _check(_lib().bf_tune_radius(int(radius), int(mode), int(width), int(height), int(reps),
                                 ctypes.byref(best), cost.ctypes.data), "bf_tune_radius")
    return best.value, [float(c) for c in cost]


def install(radius: int, mode: int, ms_per_mpix) -> None:
    from . import _check
    c = np.ascontiguousarray(ms_per_mpix, np.float32)
    assert c.size == _nconf()
    _check(_lib().bf_tune_set(int(radius), int(mode), c.ctypes.data), "bf_tune_set")


def get(radius: int, mode: int):
    cost = np.zeros(_nconf(), np.float32)
    best = int(_lib().bf_tune_get(int(radius), int(mode), cost.ctypes.data))
    return (best, [float(c) for c in cost]) if best >= 0 else (-1, None)


def reset() -> None:
    _lib().bf_tune_reset()


def tune_batch(jobs, reps: int = 3, stream=None):
    """Time a batch of bf_job entries with every tile configuration that
    can run it (each as its own bf_batch, one warm-up run, `reps` timed
    runs); returns (best_config, [ms per run or None per configuration]).
    The per-radius table predicts a batch from single images; this measures
    the batch itself -- its mix of image sizes and partial tiles."""
    import torch
    from . import Batch, BfError
    s = stream if stream is not None else torch.cuda.current_stream()
    times = []
    for c in range(_nconf()):
        try:
            b = Batch(jobs, c)
        except BfError:
            times.append(None)
            continue
        b.run(s)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(s)
        for _ in range(reps):
            b.run(s)
        e1.record(s)
        torch.cuda.synchronize()
        times.append(e0.elapsed_time(e1) / reps)
        b.close()
    ok = [(t, c) for c, t in enumerate(times) if t is not None]
    return (min(ok)[1] if ok else -1), times


def tune_all(radii=(3, 5, 7, 9, 15), modes=(0, 1), width=4096, height=2048, reps=3, log=None):
    table = {}
    for mode in modes:
        for r in radii:
            best, cost = tune(r, mode, width, height, reps)
            table[f"{mode}:{r}"] = {"best": best, "ms_per_mpix": cost}
            if log:
                log(f"mode {mode} r {r}: best config {best}, ms/Mpix {[round(c, 4) for c in cost]}")
    return table


def save(table: dict, path: str = DEFAULT_TABLE) -> None:
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    with open(path, "w") as f:
        json.dump({"configs": _nconf(), "entries": table}, f, indent=1)


def load(path: str = DEFAULT_TABLE) -> int:
    """Install a saved table; returns the number of entries installed (0 if
    the file is missing or was made for another set of configurations)."""
    if not os.path.exists(path):
        return 0
    with open(path) as f:
        doc = json.load(f)
    if doc.get("configs") != _nconf():
        return 0
    n = 0
    for key, ent in doc["entries"].items():
        mode, r = (int(v) for v in key.split(":"))
        install(r, mode, ent["ms_per_mpix"])
        n += 1
    return n


def _main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=DEFAULT_TABLE)
    ap.add_argument("--radii", default="1,2,3,4,5,6,7,8,9,10,11,12,13,14,15,16,17,18,19")
    ap.add_argument("--width", type=int, default=4096)
    ap.add_argument("--height", type=int, default=2048)
    ap.add_argument("--reps", type=int, default=3)
    args = ap.parse_args()
    radii = [int(r) for r in args.radii.split(",")]

    def log(m):
        print(f"[autotune] {m}", file=sys.stderr, flush=True)

    table = tune_all(radii, (0, 1), args.width, args.height, args.reps, log=log)
    save(table, args.out)
    log(f"saved {args.out}")
    return 0


if __name__ == "__main__":
    sys.exit(_main())
