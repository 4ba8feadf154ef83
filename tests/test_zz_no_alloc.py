"""The allocation-free denoise (vvr_unet_denoise_noalloc) against vvr_unet_denoise
through the C ABI: the same kernels on the same inputs, so the frames are equal
bit for bit, on two calls in a row (the second on a smaller frame that is not a
multiple of the network's 16-pixel step) and with the allocating entry point
used again afterwards. A frame beyond the static scratch fails with VVR_ELIMIT
before anything runs."""
import numpy as np
import pytest

from volview_amd import _abi
from volview_amd.unet import Unet

pytestmark = pytest.mark.gpu


def frame(rng, h, w, samples):
    accum = np.empty((h, w, 4), np.float32)
    accum[..., :3] = rng.gamma(0.6, 0.5, (h, w, 3)) * samples
    accum[..., 3] = rng.integers(0, samples + 1, (h, w))
    albedo = np.empty((h, w, 4), np.float32)
    albedo[..., :3] = rng.uniform(0.2, 0.9, (h, w, 3)) * samples
    albedo[..., 3] = rng.gamma(0.6, 1.0, (h, w)) * samples
    n = rng.normal(size=(h, w, 3))
    n /= np.linalg.norm(n, axis=-1, keepdims=True)
    normal = np.concatenate([n * samples, rng.uniform(50, 400, (h, w, 1)) * samples], -1).astype(np.float32)
    return accum, albedo, normal


def test_noalloc_denoise_equals_denoise(ctx):
    net = Unet.pretrained(ctx)
    rng = np.random.default_rng(5)
    for h, w in ((192, 256), (80, 96), (37, 53)):
        a, b, c = frame(rng, h, w, 4)
        got = net.denoise_noalloc(a, b, c, 4.0)
        want = net.denoise(a, b, c, 4.0)
        assert got.shape == want.shape and np.isfinite(got).all()
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), f"{h}x{w}: frames differ"
        again = net.denoise_noalloc(a, b, c, 4.0)
        assert np.array_equal(again.view(np.uint32), want.view(np.uint32))
    big = frame(rng, 512, 512, 4)
    with pytest.raises(_abi.VvrError) as e:
        net.denoise_noalloc(*big, 4.0)
    assert e.value.code == _abi.VVR_ELIMIT
    net.close()
