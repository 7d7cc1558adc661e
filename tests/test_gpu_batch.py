"""Batched exact filter (bf_batch_*): one persistent launch over many images
of uneven size, ROI and radius; bit-exact per image to the oracle, also
when replayed from a hipGraph.  Canonical-template batches at radii 3..15
run the unrolled kernels (tile configs 0 and 1), others the tap-list
kernel; both are covered."""
import numpy as np
import pytest

import opencuda_amd as oc
import oracle
from opencuda_amd import workloads as wl
from tests._gpu import SENTINEL

pytestmark = pytest.mark.gpu


class BatchCase:
    def __init__(self, items, modes=None, config=-1):
        import torch
        self.items = items
        self.modes = modes or [oc.MODE_REF] * len(items)
        self.imgs = [it.image() for it in items]
        self.luts = {}
        self.tin, self.tout, jobs = [], [], []
        for it, img, mode in zip(items, self.imgs, self.modes):
            key = (it.sigma_space, it.sigma_range)
            if key not in self.luts:
                self.luts[key] = oc.Lut.from_sigmas(*key)
            x1, y1, x2, y2 = it.roi
            tin = torch.from_numpy(img).cuda()
            tout = torch.full((y2 - y1, x2 - x1 + 5), SENTINEL, dtype=torch.uint8, device="cuda")
            self.tin.append(tin)
            self.tout.append(tout)
            jobs.append(oc.Job(oc.view(tin, it.roi), oc.view(tout, (0, 0, x2 - x1, y2 - y1)),
                               self.luts[key].handle, it.radius, mode))
        self.batch = oc.Batch(jobs, config)

    def reset(self):
        for t in self.tout:
            t.fill_(SENTINEL)

    def results(self):
        import torch
        torch.cuda.synchronize()
        return [t.cpu().numpy() for t in self.tout]

    def check(self, results=None):
        results = results or self.results()
        for it, img, mode, got in zip(self.items, self.imgs, self.modes, results):
            exp = oracle.filter(img, it.radius, it.sigma_space, it.sigma_range, mode=mode,
                                roi=it.roi)
            h, w = exp.shape
            assert np.array_equal(got[:, :w], exp)
            assert np.all(got[:, w:] == SENTINEL)


def test_batch_small_c2_subset(device):
    bc = BatchCase(wl.c2(24))
    assert bc.batch.unrolled
    bc.batch.run()
    bc.check()


def test_batch_mixed_modes_and_table_modes(device):
    items = wl.c2(6)
    # add a large-radius STD job (separate-table mode) and reuse of one table
    big = wl.Item(300, 200, 10, 8, 60, 10.0, 99, (5, 5, 295, 190), 19, 12.0, 35.0)
    items = items + [big, items[0]]
    modes = [oc.MODE_REF, oc.MODE_STD] * 3 + [oc.MODE_STD, oc.MODE_STD]
    bc = BatchCase(items, modes)
    assert not bc.batch.unrolled          # r = 19 has no unrolled kernel
    bc.batch.run()
    bc.check()


def test_every_config_identical(device):
    items = wl.c2(10)
    for cfg, tw, th, nt, cpl in oc.tune_configs():
        bc = BatchCase(items, config=cfg)
        assert bc.batch.unrolled == (cfg in (0, 1))
        bc.batch.run()
        bc.check()


def test_graph_replays_match_eager(device):
    import torch
    bc = BatchCase(wl.c2(12))
    bc.batch.run()
    eager = bc.results()
    bc.check(eager)
    scratch = torch.zeros(1 << 20, device="cuda")
    for k in range(3):
        bc.reset()
        bc.batch.run_graph()
        got = bc.results()
        for a, b in zip(got, eager):
            assert np.array_equal(a, b), f"replay {k} differs"
        scratch.add_(1.0)              # eager work between replays
        torch.cuda.synchronize()


def test_full_c2_bit_exact(device):
    bc = BatchCase(wl.c2())
    assert bc.batch.pixels == sum(it.pixels for it in bc.items)
    bc.batch.run()
    bc.check()


@pytest.mark.parametrize("config", [-1, 0, 1])
def test_unrolled_batch_every_ring_radius_both_modes(device, config):
    # every radius with an unrolled kernel, REF and STD jobs in one launch
    # (halo 16: the r = 15 instantiations of the batch kernel)
    items, modes = [], []
    for k, r in enumerate((3, 5, 7, 9, 15)):
        for mode in (oc.MODE_REF, oc.MODE_STD):
            W, H = 150 + 37 * k, 120 + 23 * k
            items.append(wl.Item(W, H, 12, 6, 50, 10.0, 1000 + 10 * k + mode,
                                 (3, 2, W - 1, H - 4), r, r / 1.5, 18.0 + 4 * k))
            modes.append(mode)
    bc = BatchCase(items, modes, config=config)
    assert bc.batch.unrolled
    bc.batch.run()
    bc.check()
