"""The learned surrogate's MFMA kernels on the GPU.

* forward and gradients against an independent float64 CPU implementation
  (torch autograd on the CPU) with the same BF16 rounding points: weights,
  input channels and post-ReLU activations rounded to BF16;
* training lowers the loss and is deterministic;
* the fused inference kernel (all six layers in one launch) against the
  per-layer kernels, bit for bit, on ragged, multi-band and pitched images;
* the check-and-fallback path: tolerance -1 makes every tile exact
  (bit-exact to bf_exact), tolerance 255 leaves the pure surrogate output;
  at tolerance 2 the fallback set is exactly the tiles whose 64 hashed
  samples (restated here) hold a difference > 2, for the unrolled check of
  the canonical template (r = 9, 15) and the run-time tap list;
* the trained network against the method without it (identity / untrained
  network): what fraction of pixels it gets within 2 LSB of the exact
  filter, and how many 64x64 tiles survive the sampled check.
"""
import os

import numpy as np
import pytest

import opencuda_amd as oc
import oracle
from opencuda_amd import surrogate as sg
from tests._check import apply_fallback, expected_fallback

pytestmark = pytest.mark.gpu

LAYERS = [(2, 32), (32, 32), (32, 32), (32, 32), (32, 32), (32, 1)]


def _unpack(p):
    out, o = [], 0
    for cin, cout in LAYERS:
        W = p[o:o + 9 * cin * cout].reshape(9, cin, cout)
        o += 9 * cin * cout
        b = p[o:o + cout]
        o += cout
        out.append((W, b))
    assert o == p.size
    return out


def _reference(params, x_u8, y_u8, sigma, margin=6, edges=None, tail=(0.0, 0.0)):
    """float64 CPU network with the GPU's BF16 rounding points; the loss
    leaves out `margin` pixels along each crop side except the sides that
    `edges[n]` marks as image edges (bit 0 top, 1 bottom, 2 left, 3 right),
    normalised by the interior pixel count.  tail = (weight, start in LSB)
    adds weight * max(|d| - start/255, 0)^2 per pixel."""
    import torch
    import torch.nn.functional as F

    def bf(t):
        return t.to(torch.bfloat16).to(torch.float64)

    class RoundSTE(torch.autograd.Function):
        @staticmethod
        def forward(ctx, t):
            return bf(t)

        @staticmethod
        def backward(ctx, g):
            return g

    layers = _unpack(params.astype(np.float64))
    ws, bs = [], []
    for (W, b), (cin, cout) in zip(layers, LAYERS):
        w = torch.tensor(W, dtype=torch.float64).permute(2, 1, 0).reshape(cout, cin, 3, 3)
        ws.append(torch.nn.Parameter(w.clone()))
        bs.append(torch.nn.Parameter(torch.tensor(b, dtype=torch.float64)))
    x = torch.tensor(x_u8, dtype=torch.float64)
    n, h, w = x.shape
    inp = torch.stack([x / 255.0, torch.tensor(sigma, dtype=torch.float64)[:, None, None]
                       .expand(n, h, w) / 100.0], 1)
    a = bf(inp)
    for l in range(5):
        a = RoundSTE.apply(F.relu(F.conv2d(a, RoundSTE.apply(ws[l]), bs[l], padding=1)))
    out = F.conv2d(a, RoundSTE.apply(ws[5]), bs[5], padding=1)[:, 0]
    pred = x / 255.0 + out
    y = torch.tensor(y_u8, dtype=torch.float64) / 255.0
    mask = torch.zeros((n, h, w), dtype=torch.float64)
    for i in range(n):
        e = 0 if edges is None else int(edges[i])
        t, b = (0 if e & 1 else margin), (0 if e & 2 else margin)
        le, r = (0 if e & 4 else margin), (0 if e & 8 else margin)
        mask[i, t:h - b, le:w - r] = 1.0
    d = (pred - y) * mask
    over = torch.clamp(d.abs() - tail[1] / 255.0, min=0.0)
    loss = ((d * d).sum() + tail[0] * (over * over).sum()) / (n * (h - 2 * margin) * (w - 2 * margin))
    loss.backward()
    grads = []
    for (cin, cout), wp, bp in zip(LAYERS, ws, bs):
        grads.append(wp.grad.reshape(cout, cin, 9).permute(2, 1, 0).reshape(-1).numpy())
        grads.append(bp.grad.numpy())
    return pred.detach().numpy(), np.concatenate(grads), float(loss)


def test_param_count(device):
    assert oc.lib().bf_sur_param_count() == 37889


@pytest.mark.parametrize("edges,tail", [(None, None), ((1 | 4, 2 | 8), None), (None, (6.0, 3.0))])
def test_forward_and_gradients_match_cpu_float64(device, edges, tail):
    import torch
    rng = np.random.default_rng(0)
    net = sg.Surrogate(seed=3)
    p = net.get_params()
    p = p + rng.normal(0, 0.02, p.size).astype(np.float32)    # non-zero biases too
    net.set_params(p)
    n, h, w = 2, 30, 70
    x = np.stack([oc.synth_image(w, h, 4, 3, 12, 10.0, 11 + i) for i in range(n)])
    y = np.clip(x.astype(int) + rng.integers(-9, 10, x.shape), 0, 255).astype(np.uint8)
    sigma = np.array([20.0, 37.5], np.float32)
    tx, ty = torch.from_numpy(x).cuda(), torch.from_numpy(y).cuda()
    ts = torch.from_numpy(sigma).cuda()
    te = torch.tensor(edges, dtype=torch.uint8).cuda() if edges is not None else None
    if tail is not None:
        net.set_loss_tail(*tail)
    pred, grads, loss = net.forward_backward(tx, ty, ts, edges=te)
    torch.cuda.synchronize()
    pred, grads, loss = pred.cpu().numpy(), grads.cpu().numpy(), float(loss.item())
    rpred, rgrads, rloss = _reference(p, x, y, sigma, edges=edges, tail=tail or (0.0, 0.0))
    if tail is not None:
        # the term is active on a good share of the pixels and moves the loss
        _, _, mse = _reference(p, x, y, sigma, edges=edges)
        assert rloss > 1.2 * mse
    assert np.abs(pred - rpred).max() < 2e-3
    assert abs(loss - rloss) / rloss < 1e-2
    o = 0
    for l, (cin, cout) in enumerate(LAYERS):
        for part, size in (("W", 9 * cin * cout), ("b", cout)):
            g, r = grads[o:o + size], rgrads[o:o + size]
            o += size
            rel = np.linalg.norm(g - r) / max(np.linalg.norm(r), 1e-30)
            cos = float(np.dot(g, r) / max(np.linalg.norm(g) * np.linalg.norm(r), 1e-30))
            assert rel < 0.05 and cos > 0.998, f"layer {l + 1} {part}: rel {rel:.4f} cos {cos:.5f}"


def test_loss_tail_arguments_are_checked(device):
    net = sg.Surrogate(seed=3)
    for bad in ((-1.0, 2.0), (float("nan"), 2.0), (float("inf"), 2.0), (1.0, -0.5), (1.0, 300.0)):
        with pytest.raises(oc.BfError):
            net.set_loss_tail(*bad)
    net.set_loss_tail(0.0, 0.0)
    net.close()


def _small_data(n=48, size=128, seed=5):
    return sg.Dataset(n_images=n, size=size, seed=seed)


def test_training_lowers_loss_and_is_deterministic(device):
    import torch
    data = _small_data()
    hist_a = sg.train(sg.Surrogate(seed=9), data, steps=120, batch=16, crop=96, log_every=10)
    assert hist_a[-1][1] < 0.25 * hist_a[0][1], hist_a
    a, b = sg.Surrogate(seed=4), sg.Surrogate(seed=4)
    sg.train(a, data, steps=5, batch=8, crop=64, seed=3)
    sg.train(b, data, steps=5, batch=8, crop=64, seed=3)
    torch.cuda.synchronize()
    assert np.array_equal(a.get_params(), b.get_params())


def _image_and_exact(size=320, seed=77, sr=30.0, mode=oc.MODE_STD):
    import torch
    img = oc.synth_image(size, size, 12, 16, 160, 12.0, seed)
    lut = oc.Lut.from_sigmas(10.0, sr)
    tin = torch.from_numpy(img).cuda()
    tex = torch.zeros_like(tin)
    oc.exact(oc.view(tin), oc.view(tex), lut, 15, mode)
    torch.cuda.synchronize()
    return img, tin, tex, lut


def test_fallback_everywhere_is_exact(device):
    import torch
    img, tin, tex, lut = _image_and_exact()
    net = sg.Surrogate(seed=1)
    out = torch.zeros_like(tin)
    total, fb = net.filter(oc.view(tin), oc.view(out), lut, 15, oc.MODE_STD, 30.0, tol_lsb=-1)
    assert total == fb == 25
    assert torch.equal(out, tex)


def test_no_fallback_is_pure_surrogate(device):
    import torch
    img, tin, tex, lut = _image_and_exact()
    net = sg.Surrogate(seed=1)
    a, b = torch.zeros_like(tin), torch.zeros_like(tin)
    total, fb = net.filter(oc.view(tin), oc.view(a), lut, 15, oc.MODE_STD, 30.0, tol_lsb=255)
    assert fb == 0
    net.infer(oc.view(tin), oc.view(b), 30.0)
    torch.cuda.synchronize()
    assert torch.equal(a, b)


@pytest.mark.parametrize("radius,mode,table", [
    (15, oc.MODE_STD, "gauss"),      # unrolled check + unrolled fallback
    (15, oc.MODE_REF, "gauss"),
    (9, oc.MODE_STD, "gauss"),
    (11, oc.MODE_STD, "gauss"),      # run-time tap list
    (15, oc.MODE_STD, "custom"),     # not the canonical template: run-time tap list
])
def test_check_flags_exactly_the_sampled_tiles(device, radius, mode, table):
    import torch
    net = sg.Surrogate(seed=1)
    if os.path.exists(sg.DEFAULT_WEIGHTS):
        net.load(sg.DEFAULT_WEIGHTS)
    H, W = 300, 457
    img = oc.synth_image(W, H, 30, 8, 120, 12.0, 40 + radius)
    if table == "gauss":
        lut = oc.Lut.from_sigmas(10.0, 30.0)
    else:
        off, wt = oracle.gauss_template(10.0)
        wt = wt.copy()
        wt[5] *= np.float32(0.5)          # one ring point weighted apart from its q
        lut = oc.Lut(off, wt, oracle.euclid_template(30.0))
    tin = torch.from_numpy(img).cuda()
    sur, ex, out = (torch.zeros_like(tin) for _ in range(3))
    net.infer(oc.view(tin), oc.view(sur), 30.0)
    oc.exact(oc.view(tin), oc.view(ex), lut, radius, mode)
    total, fb = net.filter(oc.view(tin), oc.view(out), lut, radius, mode, 30.0, tol_lsb=2)
    torch.cuda.synchronize()
    s_np, e_np, o_np = sur.cpu().numpy(), ex.cpu().numpy(), out.cpu().numpy()
    bad, tx = expected_fallback(s_np, e_np, 2)
    assert total == bad.size and fb == int(bad.sum())
    # the network approximates STD: in REF mode (4-row windows) every tile
    # may fail, in STD some must pass
    assert fb > 0 and (mode == oc.MODE_REF or fb < total)
    assert np.array_equal(o_np, apply_fallback(s_np, e_np, bad, tx))


@pytest.mark.parametrize("shape,roi,sr", [
    ((217, 300), None, 27.5),              # one band, last strip partial
    ((530, 613), None, 40.0),              # three bands of 256 rows
    ((40, 1000), (3, 5, 998, 37), 15.0),   # pitched ROI view, fewer rows than a band
    ((13, 13), None, 30.0),                # smaller than one strip
    ((1111, 3200), None, 22.0),            # fills a 256-CU device: bands of 278 rows, the last 277
])
def test_fused_inference_equals_per_layer_path(device, shape, roi, sr):
    import torch
    net = sg.Surrogate(seed=3)
    if os.path.exists(sg.DEFAULT_WEIGHTS):
        net.load(sg.DEFAULT_WEIGHTS)
    H, W = shape
    img = oc.synth_image(W, H, 20, 3, 80, 10.0, 5 + H)
    tin = torch.from_numpy(img).cuda()
    x1, y1, x2, y2 = roi or (0, 0, W, H)
    out = torch.full((y2 - y1, x2 - x1), 77, dtype=torch.uint8, device="cuda")
    net.infer(oc.view(tin, roi), oc.view(out), sr)
    crop = tin[y1:y2, x1:x2].contiguous()
    sigma = torch.tensor([sr], dtype=torch.float32, device="cuda")
    pred, _, _ = net.forward_backward(crop[None], None, sigma)
    ref = torch.clamp(torch.round(pred[0] * 255.0), 0, 255).to(torch.uint8)
    torch.cuda.synchronize()
    assert torch.equal(out, ref)


def _within(a, b, tol=2):
    return float((np.abs(a.astype(int) - b.astype(int)) <= tol).mean())


def test_trained_network_against_the_method_without_it(device):
    """Trained net vs identity (no network: output = input) and vs the
    untrained network, on held-out images (r=15, STD target)."""
    import torch
    net = sg.Surrogate(seed=2)
    if os.path.exists(sg.DEFAULT_WEIGHTS):
        net.load(sg.DEFAULT_WEIGHTS)
    else:
        sg.train(net, _small_data(n=96, size=160, seed=8), steps=300, batch=16, crop=128)
    untrained = sg.Surrogate(seed=2)
    img, tin, tex, lut = _image_and_exact(size=512, seed=1234, sr=25.0)
    ex = tex.cpu().numpy()
    outs = {}
    for name, model in (("trained", net), ("untrained", untrained)):
        o = torch.zeros_like(tin)
        model.infer(oc.view(tin), oc.view(o), 25.0)
        torch.cuda.synchronize()
        outs[name] = o.cpu().numpy()
    outs["identity"] = img
    inner = (slice(16, -16), slice(16, -16))
    score = {k: _within(v[inner], ex[inner]) for k, v in outs.items()}
    mae = {k: float(np.abs(v[inner].astype(int) - ex[inner]).mean()) for k, v in outs.items()}
    print("within 2 LSB:", score, "MAE:", mae)
    assert score["trained"] > score["identity"] + 0.2
    assert score["trained"] > score["untrained"] + 0.2
    assert mae["trained"] < 0.5 * mae["identity"]


@pytest.mark.parametrize("entry", ["infer", "filter", "filter_async"])
def test_overlapping_views_are_rejected(device, entry):
    """In place (in == out) or partly overlapping views: the fused kernel
    would overwrite halo pixels other workgroups still read, and the check
    and fallback would read a modified input.  Rejected with INVALID_DATA
    before any launch; the image is untouched."""
    import torch
    net = sg.Surrogate(seed=1)
    lut = oc.Lut.from_sigmas(10.0, 30.0)
    img = oc.synth_image(200, 150, 10, 8, 60, 10.0, 3)
    t = torch.from_numpy(img).cuda()
    before = t.clone()
    for vin, vout in ((oc.view(t), oc.view(t)),
                      (oc.view(t, (0, 0, 120, 100)), oc.view(t, (50, 40, 200, 150)))):
        with pytest.raises(oc.BfError) as ei:
            if entry == "infer":
                net.infer(vin, vout, 30.0)
            elif entry == "filter":
                net.filter(vin, vout, lut, 15, oc.MODE_STD, 30.0)
            else:
                net.filter_async(vin, vout, lut, 15, oc.MODE_STD, 30.0)
        assert ei.value.code == oc.INVALID_DATA
    torch.cuda.synchronize()
    assert torch.equal(t, before)
