"""GPU parity of the exact filter (bf_exact) against the oracle.

Bit-exact u8 on every pixel: small and ragged shapes, offset ROIs, pitched
buffers, trimmed outputs (pixels outside the filtered region untouched),
every radius, both modes, every tile configuration, and the full-size
configurations c0, c1 and c3."""
import numpy as np
import pytest

import opencuda_amd as oc
import oracle
from opencuda_amd import workloads as wl
from tests._gpu import check_against_oracle, run_exact

pytestmark = pytest.mark.gpu


def _img(h, w, seed):
    return oc.synth_image(w, h, max(1, (w * h) // 4000), 2, max(3, min(w, h) // 3), 12.0, seed)


@pytest.mark.parametrize("mode", [oc.MODE_REF, oc.MODE_STD])
def test_c0_parity(device, mode):
    (it,) = wl.c0()
    img = it.image()
    got = run_exact(img, it.radius, it.sigma_space, it.sigma_range, mode)
    check_against_oracle(img, it.radius, it.sigma_space, it.sigma_range, mode, got=got)


RAGGED = [
    # (h, w, roi, out_size, radius, ss, sr, pad_in, pad_out)
    (1, 1, None, None, 1, 1.0, 30.0, 0, 0),
    (1, 9, None, None, 2, 1.0, 30.0, 3, 0),
    (9, 1, None, None, 3, 2.0, 30.0, 0, 5),
    (2, 3, None, None, 5, 2.0, 20.0, 1, 1),
    (5, 5, (1, 1, 4, 5), None, 4, 1.5, 25.0, 0, 0),
    (37, 29, (3, 2, 28, 35), (20, 40), 3, 3.0, 30.0, 7, 3),
    (70, 131, (5, 9, 130, 66), (200, 30), 7, 5.0, 25.0, 13, 0),
    (129, 67, (0, 1, 67, 129), None, 9, 6.0, 15.0, 0, 9),
    (100, 300, (17, 3, 290, 97), (273, 94), 15, 10.0, 30.0, 5, 5),
    (64, 64, None, (63, 61), 15, 10.0, 30.0, 0, 0),
    (33, 200, (1, 0, 199, 33), None, 19, 12.0, 40.0, 3, 1),
]


@pytest.mark.parametrize("mode", [oc.MODE_REF, oc.MODE_STD])
@pytest.mark.parametrize("case", RAGGED)
def test_ragged_shapes(device, case, mode):
    h, w, roi, out_size, r, ss, sr, pi, po = case
    img = _img(h, w, seed=h * 1000 + w)
    got = run_exact(img, r, ss, sr, mode, roi=roi, out_size=out_size, pad_in=pi, pad_out=po)
    check_against_oracle(img, r, ss, sr, mode, roi=roi, out_size=out_size, got=got)


@pytest.mark.parametrize("mode", [oc.MODE_REF, oc.MODE_STD])
def test_every_radius(device, mode):
    img = _img(61, 83, seed=5)
    roi = (3, 4, 80, 58)
    top = 19 if mode == oc.MODE_REF else 20
    for r in range(1, top + 1):
        ss, sr = max(0.5, r / 1.5), 10.0 + 2 * r
        got = run_exact(img, r, ss, sr, mode, roi=roi)
        check_against_oracle(img, r, ss, sr, mode, roi=roi, got=got)


@pytest.mark.parametrize("radius,ss,sr", [(3, 3.0, 30.0), (7, 5.0, 25.0), (15, 10.0, 30.0)])
def test_every_tile_config_identical(device, radius, ss, sr):
    img = _img(150, 210, seed=radius)
    roi = (7, 5, 205, 148)
    ref = None
    for cfg, tw, th, nt, cpl in oc.tune_configs():
        for mode in (oc.MODE_REF, oc.MODE_STD):
            try:
                got = run_exact(img, radius, ss, sr, mode, roi=roi, config=cfg)
            except oc.BfError as e:       # config does not fit the LDS at this radius
                assert e.code == oc.INVALID_DATA
                continue
            check_against_oracle(img, radius, ss, sr, mode, roi=roi, got=got)
    del ref


def test_subnormal_and_zero_weights(device):
    """sigma_range 10 makes E[d] subnormal for d ~ 133..144 and 0 beyond:
    high-contrast content exercises them."""
    rng = np.random.default_rng(1)
    img = np.where(rng.random((90, 90)) < 0.5, 10, 250).astype(np.uint8)
    img[::7] = 140
    for mode in (oc.MODE_REF, oc.MODE_STD):
        got = run_exact(img, 5, 3.0, 10.0, mode)
        check_against_oracle(img, 5, 3.0, 10.0, mode, got=got)


def test_zero_weight_sum_gives_zero(device):
    """Tiny sigma_space: in REF mode rows 1..3 of a group can have every
    weight underflow to 0 (0/0 = NaN -> 0), as in the oracle."""
    rng = np.random.default_rng(2)
    img = rng.integers(0, 256, (32, 32), dtype=np.uint8)
    got = run_exact(img, 1, 0.05, 5.0, oc.MODE_REF)
    exp = check_against_oracle(img, 1, 0.05, 5.0, oc.MODE_REF, got=got)
    assert (exp[1::4] == 0).any()


def test_deterministic(device):
    img = _img(300, 400, seed=3)
    a = run_exact(img, 9, 6.0, 20.0, oc.MODE_REF)
    b = run_exact(img, 9, 6.0, 20.0, oc.MODE_REF)
    assert np.array_equal(a, b)


def test_inplace_repeat_matches_sequential_passes(device):
    import torch
    img = _img(90, 120, seed=8)
    lut = oc.Lut.from_sigmas(3.0, 20.0)
    t = torch.from_numpy(img).cuda().contiguous()
    oc.exact_inplace(oc.view(t, (4, 3, 110, 85)), lut, 3, oc.MODE_REF, repeat=3)
    torch.cuda.synchronize()
    got = t.cpu().numpy()
    exp = img.copy()
    sub = exp[3:85, 4:110].copy()
    for _ in range(3):
        sub = oracle.filter(np.ascontiguousarray(sub), 3, 3.0, 20.0, mode=oracle.REF)
    exp[3:85, 4:110] = sub
    assert np.array_equal(got, exp)


def test_invalid_arguments(device):
    import torch
    lut = oc.Lut.from_sigmas(3.0, 30.0)
    t = torch.zeros((16, 16), dtype=torch.uint8, device="cuda")
    o = torch.zeros((16, 16), dtype=torch.uint8, device="cuda")
    vin, vout = oc.view(t), oc.view(o)
    for r in (0, -1, 21):
        with pytest.raises(oc.BfError) as e:
            oc.exact(vin, vout, lut, r, oc.MODE_REF)
        assert e.value.code == oc.INVALID_DATA
    with pytest.raises(oc.BfError) as e:     # REF r=20 would read past the 41x41 template
        oc.exact(vin, vout, lut, 20, oc.MODE_REF)
    assert e.value.code == oc.INVALID_DATA
    oc.exact(vin, vout, lut, 20, oc.MODE_STD)  # STD r=20 fits
    with pytest.raises(oc.BfError) as e:     # overlapping input and output
        oc.exact(vin, oc.view(t, (2, 2, 10, 10)), lut, 3, oc.MODE_REF)
    assert e.value.code == oc.INVALID_DATA
    with pytest.raises(oc.BfError) as e:
        oc.exact(vin, vout, lut, 3, 7)
    assert e.value.code == oc.INVALID_DATA


def test_c1_parity(device):
    (it,) = wl.c1()
    img = it.image()
    got = run_exact(img, it.radius, it.sigma_space, it.sigma_range, oc.MODE_REF, roi=it.roi,
                    pad_in=64)
    check_against_oracle(img, it.radius, it.sigma_space, it.sigma_range, oc.MODE_REF,
                         roi=it.roi, got=got)


def test_c3_parity_full_size(device):
    """The headline workload, 8192^2 at r=15: bit-exact on all 67M pixels,
    and the error against the FP64 oracle on a band of 64 rows."""
    (it,) = wl.c3()
    img = it.image()
    got = run_exact(img, it.radius, it.sigma_space, it.sigma_range, oc.MODE_REF)
    check_against_oracle(img, it.radius, it.sigma_space, it.sigma_range, oc.MODE_REF, got=got)
    v = oracle.filter_f64(img, it.radius, it.sigma_space, it.sigma_range, mode=oracle.REF,
                          rows=(4096, 4160))
    err = got[4096:4160].astype(np.float64) - v
    assert np.abs(err).max() < 1.05 and -0.75 < err.mean() < -0.25


def test_autotuner_installs_a_working_config(device):
    from opencuda_amd import autotune
    try:
        best, cost = autotune.tune(5, oc.MODE_REF, 512, 256, reps=1)
        assert 0 <= best < len(oc.tune_configs())
        assert cost[best] == min(c for c in cost if c >= 0)
        assert autotune.get(5, oc.MODE_REF)[0] == best
        img = _img(120, 130, seed=17)
        got = run_exact(img, 5, 3.3, 22.0, oc.MODE_REF)      # uses the tuned config
        check_against_oracle(img, 5, 3.3, 22.0, oc.MODE_REF, got=got)
    finally:
        autotune.reset()
    assert autotune.get(5, oc.MODE_REF)[0] == -1
