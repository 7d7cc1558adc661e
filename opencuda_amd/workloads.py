"""Synthetic workloads of the benchmark suite.

Each workload is a list of ``Item``s: a seeded synthetic u8 image (see
``bf_synth_image``), its ROI, the filter radius and the two sigmas.
Seeds: a workload's parameter stream is splitmix64-seeded with its index;
image ``i`` of workload ``c`` has content seed ``sm64(c ^ i)``.

  c0  512x512, full ROI, r=3, sigma_space=3, sigma_range=30 (40+40 shapes,
      sizes 8..96, noise 10)
  c1  4096x4096 (400+400 shapes, 16..512, noise 10), ROI (100,200) 3800x3600,
      r=7, sigma_space=5, sigma_range=25
  c2  256 images, width/height uniform in 256..2048, ROI 50..100% of each side
      at a random offset, r in {3,5,9}, sigma_space=r/1.5, sigma_range in
      [10,50]; content scaled by sqrt(W*H)/512
  c3  8192x8192 (1600+1600 shapes, 16..1024, noise 12), r=15,
      sigma_space=10, sigma_range=30
  c4  512 images, width/height uniform in 128..4096, full ROI, r in
      {3,7,15}, sigma_space=r/1.5, sigma_range in [15,45]
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

_GOLDEN = 0x9E3779B97F4A7C15
_M64 = (1 << 64) - 1


def sm64(x: int) -> int:
    z = (x + _GOLDEN) & _M64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
    return z ^ (z >> 31)


class Rng:
    def __init__(self, seed: int):
        self.s = seed & _M64

    def next(self) -> int:
        r = sm64(self.s)
        self.s = (self.s + _GOLDEN) & _M64
        return r

    def uniform(self, lo: int, hi: int) -> int:
        return lo + self.next() % (hi - lo + 1)

    def uniform01(self) -> float:
        return (self.next() >> 11) * 2.0 ** -53


@dataclass
class Item:
    width: int
    height: int
    nshapes: int
    smin: int
    smax: int
    noise: float
    seed: int
    roi: tuple            # (x1, y1, x2, y2), x2/y2 exclusive
    radius: int
    sigma_space: float
    sigma_range: float

    def image(self) -> np.ndarray:
        from . import synth_image
        return synth_image(self.width, self.height, self.nshapes, self.smin, self.smax,
                           self.noise, self.seed)

    @property
    def pixels(self) -> int:
        x1, y1, x2, y2 = self.roi
        return (x2 - x1) * (y2 - y1)


def _scaled(W: int, H: int, noise: float, seed: int, roi, r: int, ss: float, sr: float) -> Item:
    s = math.sqrt(W * H) / 512.0
    n = max(1, round(40 * s))
    smin = max(1, round(8 * s))
    smax = max(smin, round(96 * s))
    return Item(W, H, n, smin, smax, noise, seed, roi, r, ss, sr)


def c0() -> list[Item]:
    return [Item(512, 512, 40, 8, 96, 10.0, sm64(0), (0, 0, 512, 512), 3, 3.0, 30.0)]


def c1() -> list[Item]:
    return [Item(4096, 4096, 400, 16, 512, 10.0, sm64(1), (100, 200, 3900, 3800), 7, 5.0, 25.0)]


def c2(count: int = 256) -> list[Item]:
    rng = Rng(2)
    out = []
    for i in range(256):
        W, H = rng.uniform(256, 2048), rng.uniform(256, 2048)
        rw, rh = rng.uniform((W + 1) // 2, W), rng.uniform((H + 1) // 2, H)
        rx, ry = rng.uniform(0, W - rw), rng.uniform(0, H - rh)
        r = (3, 5, 9)[rng.uniform(0, 2)]
        sr = 10.0 + 40.0 * rng.uniform01()
        out.append(_scaled(W, H, 10.0, sm64(2 ^ i), (rx, ry, rx + rw, ry + rh), r, r / 1.5, sr))
    return out[:count]


def c3() -> list[Item]:
    return [Item(8192, 8192, 1600, 16, 1024, 12.0, sm64(3), (0, 0, 8192, 8192), 15, 10.0, 30.0)]


def c4(count: int = 512) -> list[Item]:
    rng = Rng(4)
    out = []
    for i in range(512):
        W, H = rng.uniform(128, 4096), rng.uniform(128, 4096)
        r = (3, 7, 15)[rng.uniform(0, 2)]
        sr = 15.0 + 30.0 * rng.uniform01()
        out.append(_scaled(W, H, 10.0, sm64(4 ^ i), (0, 0, W, H), r, r / 1.5, sr))
    return out[:count]


CONFIGS = {"c0": c0, "c1": c1, "c2": c2, "c3": c3, "c4": c4}
