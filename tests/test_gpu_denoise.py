"""The a-trous denoiser (denoise.comp) on the GPU, through the C ABI.

- After a path-traced frame with VVR_DENOISE_ATROUS the HDR image equals the
  FP32 oracle's denoise.comp over the GPU's own accumulation and guide images
  within binary16 rounding + 1e-3 on all but 0.1 % of pixels (discrete
  decisions: exp/pow of near-equal values), after one frame and after a
  second that accumulates onto the first.
- Denoising does not restart accumulation, and read_denoised fails after a
  frame that was not denoised.
- The filter helps: at 4 paths per pixel the denoised frame is closer (relative
  RMSE) to a 256-path frame than the raw 4-path frame is.
"""
import numpy as np
import pytest

from volview_amd import _abi, phantom, workloads
from volview_amd.renderer import Renderer, TransferFunction

pytestmark = pytest.mark.gpu


def scene(ctx, labels=True, dims=(48, 44, 40), seed=6):
    hu = phantom.ct(dims, seed)
    r = Renderer(ctx, hu, TransferFunction.preset(0))
    n = 0
    if labels:
        lab, soft, _, n = phantom.organs(dims, seed, 12)
        r.set_labels(lab, soft)
    return r, n


def compare(got, want):
    tol = 1e-3 + np.abs(want) * 2.0 ** -10
    bad = np.any(np.abs(got - want) > tol, axis=-1)
    print(f"denoised: {bad.mean():.5f} of pixels off, worst {np.abs(got - want).max():.3g}")
    assert bad.mean() <= 1e-3


def test_denoised_frames_equal_oracle(ctx, oracle):
    r, n = scene(ctx)
    w, h = 80, 64
    r.resize(w, h)
    p = workloads.pt_params(r, n, 8, bounces=4, camera_index=1)
    p.path_tracing.denoiser = _abi.DENOISE_ATROUS
    for frame in (1, 2):
        r.record(frame, p)
        got = r.read_denoised()
        assert np.array_equal(got, r.read_hdr())
        accum, albedo, normal = r.read_accum()
        assert r.accumulated_samples() == 8 * frame
        want, _ = oracle.denoise(accum, albedo, normal, float(8 * frame))
        compare(got.reshape(-1, 4), want.reshape(-1, 4))


def test_denoiser_off_frame_has_no_denoised_image(ctx):
    r, n = scene(ctx, labels=False)
    r.resize(16, 16)
    p = workloads.pt_params(r, 0, 2, bounces=2)
    p.path_tracing.denoiser = _abi.DENOISE_ATROUS
    r.record(1, p)
    p.path_tracing.denoiser = _abi.DENOISE_OFF
    r.record(2, p)
    assert r.accumulated_samples() == 4  # the denoiser is not part of the accumulation key
    with pytest.raises(_abi.VvrError) as e:
        r.read_denoised()
    assert e.value.code == _abi.VVR_ESTATE


def rel_rmse(a, b):
    return float(np.sqrt(np.mean((a - b) ** 2)) / np.sqrt(np.mean(b ** 2)))


def test_denoiser_reduces_error_at_low_sample_counts(ctx):
    r, n = scene(ctx)
    w, h = 96, 80
    r.resize(w, h)
    r.record(1, workloads.pt_params(r, n, 256, bounces=4, camera_index=2))
    ref = r.read_hdr()[..., :3]
    p = workloads.pt_params(r, n, 4, bounces=4, camera_index=2)  # a new key: restarts at 4 paths
    r.record(2, p)
    raw = r.read_hdr()[..., :3]
    p.path_tracing.denoiser = _abi.DENOISE_ATROUS
    r2, _ = scene(ctx)
    r2.resize(w, h)
    r2.record(1, p)
    den = r2.read_denoised()[..., :3]
    e_raw, e_den = rel_rmse(raw, ref), rel_rmse(den, ref)
    print(f"4 paths: raw {e_raw:.4f}, denoised {e_den:.4f} relative RMSE vs 256 paths")
    assert e_den < e_raw
