"""The benchmark configurations themselves on the GPU, through the C ABI.

- Batched-32 (32 seeded phantoms of edge 128-512, images 256^2-1024^2, 16
  ray-marched and 16 path-traced requests at 1-64 paths): every request served
  in one submission equals, bit for bit, the same request rendered alone by its
  renderer.
- The shipped learned denoiser (5 levels, 32-128 channels, the trained weights)
  equals the float64 oracle's forward pass within bf16 rounding on a small
  frame, and on a 1024^2 frame (the configuration's size) its output is finite
  and the persistent convolutions equal the per-tile ones bit for bit.
"""
import numpy as np
import pytest

import unet_ref as U  # oracle/unet_ref.py
from volview_amd import phantom, workloads
from volview_amd.renderer import Renderer, TransferFunction
from volview_amd.server import Server
from volview_amd.unet import Unet

from test_gpu_unet import oracle_inputs, synthetic

pytestmark = pytest.mark.gpu


def test_batched32_served_equals_alone(ctx):
    vols, reqs = workloads.batched32()
    rs = [Renderer(ctx, phantom.ct(dims, seed), TransferFunction.preset(0)) for dims, seed in vols]
    params = [workloads.request_params(rs[q.volume], q) for q in reqs]
    alone = []
    for k, (q, p) in enumerate(zip(reqs, params)):
        r = rs[q.volume]
        r.resize(q.width, q.height)
        r.record(k + 1, p)
        alone.append(r.read_hdr())
    print(f"{sum(q.path_traced for q in reqs)} path-traced requests, "
          f"{sum(q.width * q.height * q.spp for q in reqs) / 1e6:.0f} M paths")
    srv = Server(ctx)
    ids = [srv.submit(rs[q.volume], p, q.width, q.height, tag=k) for k, (q, p) in enumerate(zip(reqs, params))]
    done = set()
    for _ in range(3000):
        for c in srv.poll(wait_ms=100):
            assert c.status == 0, c.status
            if c.final_frame:
                done.add(c.id)
        if done == set(ids):
            break
    assert done == set(ids)
    for k, (i, q) in enumerate(zip(ids, reqs)):
        got = srv.read_hdr(i)
        assert np.array_equal(got.view(np.uint32), alone[k].view(np.uint32)), (k, q)
        srv.release(i)
    srv.close()
    for r in rs:
        r.close()


def test_shipped_denoiser_equals_oracle_forward(ctx):
    rng = np.random.default_rng(21)
    net = Unet.pretrained(ctx)
    channels = net.channels
    accum, albedo, normal, n, target = synthetic(rng, 1, 64, 96)
    got = net.denoise(accum[0], albedo[0], normal[0], n)
    p, x, alb, _ = oracle_inputs(net.params(), channels, accum, albedo, normal, n, target)
    out, _ = U.forward(p, channels, x)
    want = np.expm1(np.maximum(out, 0.0))[0] * alb[0]
    err = np.abs(got[..., :3] - want)
    bad = err > 0.03 * np.abs(want) + 2e-3
    print(f"5-level forward: {bad.mean():.4f} of channels beyond bf16 tolerance, worst {err.max():.4g}")
    assert bad.mean() < 0.01


def test_shipped_denoiser_on_a_1024_frame(ctx, monkeypatch):
    rng = np.random.default_rng(22)
    net = Unet.pretrained(ctx)
    accum, albedo, normal, n, _ = synthetic(rng, 1, 1024, 1024)
    monkeypatch.setenv("VVR_UNET_STREAM", "1")
    streamed = net.denoise(accum[0], albedo[0], normal[0], n)
    monkeypatch.setenv("VVR_UNET_STREAM", "0")
    per_tile = net.denoise(accum[0], albedo[0], normal[0], n)
    assert np.isfinite(streamed).all() and streamed[..., :3].max() > 0
    assert np.array_equal(streamed.view(np.uint32), per_tile.view(np.uint32))


def test_long_frame_accumulates_no_nan(ctx):
    """The serving trace's first path-traced request (1024^2, 40 organs) with 512
    paths per pixel in one frame, 537 M paths: every accumulated value is finite.
    (A Fresnel base 1 - dot(v, h) that rounding leaves just below 0 once made
    pow NaN on a few paths in 10^8.)"""
    vols, reqs = workloads.serving_trace(5.0)
    q = next(q for q in reqs if q.path_traced)
    dims, seed = vols[q.volume]
    r = Renderer(ctx, phantom.ct(dims, seed), TransferFunction.preset(0))
    labels, soft, _, n = phantom.organs(dims, seed, 40)
    r.set_labels(labels, soft)
    p = workloads.request_params(r, q, n)
    p.path_tracing.samples_per_frame = 512
    r.resize(q.width, q.height)
    r.record(1, p)
    acc, alb, nrm = r.read_accum()
    for name, a in (("radiance", acc), ("albedo", alb), ("normal", nrm)):
        assert np.isfinite(a).all(), f"{name}: {int((~np.isfinite(a)).sum())} values not finite"
    r.close()
