"""The learned denoiser's U-Net on the GPU (bf16 MFMA kernels, FP32
accumulation), through the C ABI, against the float64 oracle (oracle/unet_ref.py,
pinned by finite differences).

- The loss and every layer's weight and bias gradients of a training step equal
  the oracle's evaluated at the same bf16-rounded weights and input: the loss
  within 1 %, each layer's gradient with cosine > 0.98 and norm within 10 %
  (activations are rounded to bf16 at every layer on the GPU).
- The denoised frame equals the oracle's forward pass within bf16 rounding, on a
  small frame and on a 380 x 400 one (2-D tiled convolutions, partial tiles).
- Layers of at most 32 input and output channels on large frames run in
  persistent workgroups that walk tiles; the frame equals, bit for bit, the one
  computed with one workgroup per tile (VVR_UNET_STREAM=0).
- Training on the GPU fits a fixed batch of path-traced frames: the loss drops
  by more than a third in 150 Adam steps.
"""
import numpy as np
import pytest

import unet_ref as U  # oracle/unet_ref.py
from volview_amd import phantom, workloads
from volview_amd.renderer import Renderer, TransferFunction
from volview_amd.unet import Unet

pytestmark = pytest.mark.gpu


def bf16(a):
    a = np.ascontiguousarray(a, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    return (((u + 0x7fff + ((u >> 16) & 1)) >> 16) << 16).astype(np.uint32).view(np.float32).astype(np.float64)


def synthetic(rng, b, h, w, n=4.0):
    accum = np.zeros((b, h, w, 4), np.float32)
    accum[..., :3] = rng.gamma(1.0, 0.5, (b, h, w, 3)) * n
    accum[..., 3] = (rng.random((b, h, w)) < 0.1) * n
    albedo = np.zeros((b, h, w, 4), np.float32)
    albedo[..., :3] = rng.uniform(0.1, 0.9, (b, h, w, 3)) * n
    albedo[..., 3] = rng.uniform(0, 1, (b, h, w)) * n
    nrm = rng.normal(size=(b, h, w, 3))
    nrm /= np.linalg.norm(nrm, axis=-1, keepdims=True)
    normal = np.zeros((b, h, w, 4), np.float32)
    normal[..., :3] = nrm * n
    normal[..., 3] = rng.uniform(50, 200, (b, h, w)) * n
    target = np.zeros((b, h, w, 4), np.float32)
    target[..., :3] = rng.gamma(2.0, 0.25, (b, h, w, 3))
    return accum, albedo, normal, n, target


def oracle_inputs(params, channels, accum, albedo, normal, n, target):
    x, alb = U.prepare(accum.astype(np.float64), albedo.astype(np.float64), normal.astype(np.float64), n)
    layers = U.unflatten(params.astype(np.float64), channels)
    p = np.concatenate([np.concatenate([bf16(w).ravel(), b]) for w, b in layers])
    tlog = np.log1p(np.maximum(target[..., :3], 0.0) / alb)
    return p, bf16(x), alb, tlog


def test_gradients_equal_oracle(ctx):
    rng = np.random.default_rng(11)
    channels = (16, 32, 48)
    net = Unet(ctx, channels, seed=5)
    batch = synthetic(rng, 2, 16, 24)
    loss = net.train_step(*batch, lr=0.0)
    g = net.grads()
    p, x, alb, tlog = oracle_inputs(net.params(), channels, *batch)
    want_loss, want_g, _ = U.loss_and_grads(p, channels, x, tlog)
    print(f"loss {loss:.6f} oracle {want_loss:.6f}")
    assert abs(loss - want_loss) <= 0.01 * want_loss
    k = 0
    for i, (cin, cout) in enumerate(U.layer_shapes(channels)):
        for part, size in (("W", cout * 9 * cin), ("b", cout)):
            a, b = g[k:k + size].astype(np.float64), want_g[k:k + size]
            k += size
            cos = a @ b / (np.linalg.norm(a) * np.linalg.norm(b) + 1e-30)
            ratio = np.linalg.norm(a) / (np.linalg.norm(b) + 1e-30)
            print(f"layer {i} {part}: cos {cos:.4f} norm ratio {ratio:.4f}")
            assert cos > 0.98 and 0.9 < ratio < 1.1, (i, part, cos, ratio)


@pytest.mark.parametrize("h,w", [(32, 48), (380, 400)])  # the larger frame: the 2-D tiled kernel, partial tiles
def test_denoised_frame_equals_oracle_forward(ctx, h, w):
    rng = np.random.default_rng(12)
    channels = (16, 32)
    net = Unet(ctx, channels, seed=7)
    accum, albedo, normal, n, target = synthetic(rng, 1, h, w)
    got = net.denoise(accum[0], albedo[0], normal[0], n)
    p, x, alb, _ = oracle_inputs(net.params(), channels, accum, albedo, normal, n, target)
    out, _ = U.forward(p, channels, x)
    want = np.expm1(np.maximum(out, 0.0))[0] * alb[0]
    err = np.abs(got[..., :3] - want)
    bad = err > 0.03 * np.abs(want) + 2e-3
    print(f"forward: {bad.mean():.4f} of channels beyond bf16 tolerance, worst {err.max():.4g}")
    assert bad.mean() < 0.01
    assert np.allclose(got[..., 3], accum[0, ..., 3] / n, atol=1e-3)


def test_persistent_convolutions_equal_per_tile(ctx, monkeypatch):
    rng = np.random.default_rng(5)
    net = Unet.pretrained(ctx)
    accum, albedo, normal, n, _ = synthetic(rng, 1, 264, 1000)  # partial tiles on both axes
    monkeypatch.setenv("VVR_UNET_STREAM", "1")
    streamed = net.denoise(accum[0], albedo[0], normal[0], n)
    monkeypatch.setenv("VVR_UNET_STREAM", "0")
    per_tile = net.denoise(accum[0], albedo[0], normal[0], n)
    assert np.array_equal(streamed, per_tile)
    assert np.isfinite(streamed).all() and streamed[..., :3].max() > 0


def test_training_fits_a_batch(ctx):
    dims = (40, 40, 40)
    hu = phantom.ct(dims, 9)
    labels, sg, _, nl = phantom.organs(dims, 9, 8)
    r = Renderer(ctx, hu, TransferFunction.preset(0))
    r.set_labels(labels, sg)
    w = h = 32
    r.resize(w, h)
    acc, alb, nrm, tgt = [], [], [], []
    for cam in range(4):
        r.record(1, workloads.pt_params(r, nl, 128, bounces=4, camera_index=cam))
        tgt.append(r.read_hdr())
        r.record(2, workloads.pt_params(r, nl, 4, bounces=4, camera_index=cam))
        a, b, c = r.read_accum()
        acc.append(a), alb.append(b), nrm.append(c)
    batch = [np.stack(v) for v in (acc, alb, nrm)]
    target = np.stack(tgt)
    net = Unet(ctx, (16, 32, 48), seed=3)
    losses = [net.train_step(*batch, 4.0, target, lr=2e-3) for _ in range(150)]
    print(f"loss {losses[0]:.4f} -> {losses[-1]:.4f}")
    assert losses[-1] < 0.67 * losses[0]
