"""What the trained denoiser gains (volview_amd/data/denoiser_unet.npy, trained
by volview_amd.train_denoiser on the path tracer's own 4- and 1024-path pairs).

- On phantoms used neither in training nor in its evaluation, the network's
  denoised 4-path frame has a lower relative MSE against a 512-path reference
  than the raw 4-path frame, the a-trous filter at 4 paths, and the a-trous
  filter at 64 paths (the fallback with more samples).
- A renderer with the network attached denoises VVR_DENOISE_LEARNED frames of
  any size (padded to the network's multiple of 16) exactly as the host entry
  point does; without one attached such frames are rejected.
"""
import numpy as np
import pytest

from volview_amd import _abi, workloads
from volview_amd.train_denoiser import evaluate, render_pairs, scene
from volview_amd.unet import Unet

pytestmark = pytest.mark.gpu


def test_pretrained_network_beats_the_method_without_it(ctx):
    pairs = render_pairs(ctx, range(7000, 7003), low=4, high=512, cameras=(1, 3), atrous_spp=(4, 64))
    net = Unet.pretrained(ctx)
    s = evaluate(net, pairs, 4)
    print({k: round(v, 5) for k, v in s.items()})
    assert s["unet"] < s["raw"]
    assert s["unet"] < s["atrous4"]
    assert s["unet"] < s["atrous64"]


def test_renderer_runs_the_learned_denoiser(ctx):
    r, n = scene(ctx, 7100)
    w, h = 100, 90
    r.resize(w, h)
    p = workloads.pt_params(r, n, 4, bounces=8, camera_index=2)
    p.path_tracing.denoiser = _abi.DENOISE_LEARNED
    with pytest.raises(_abi.VvrError) as e:
        r.record(1, p)
    assert e.value.code == _abi.VVR_EINVAL
    net = Unet.pretrained(ctx)
    r.set_denoiser(net)
    r.record(1, p)
    got = r.read_denoised()
    acc, alb, nrm = r.read_accum()
    want = net.denoise(acc, alb, nrm, 4.0)
    assert np.array_equal(got, want)
    assert np.isfinite(got).all() and got[..., :3].max() > 0
    r.set_denoiser(None)
