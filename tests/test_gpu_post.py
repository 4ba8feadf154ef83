"""Camera finishing and tone mapping (bloom.comp, tonemap.comp) on the GPU,
through the C ABI, against the oracle run on the GPU's own HDR image and
uniform block.

- The RGBA8 display image of a ray-marched frame with the reference's default
  finishing (bloom 0.04, vignette 0.25, grain 0.012, ACES, auto exposure, sRGB)
  is within 1 of the oracle's on every channel and equal on >= 99 % of them
  (the UNORM store rounds; log / exp / pow differ in the last bit), and the
  metered exposure agrees to 1e-5.
- The exposure adapts over frames as the oracle adapts it.
- A path-traced frame (background share in alpha), and AgX without sRGB.
- render_to_rgba equals resize + record + read_rgba.
"""
import numpy as np
import pytest

from volview_amd import _abi, phantom, workloads
from volview_amd.renderer import Renderer, TransferFunction, render_to_rgba

pytestmark = pytest.mark.gpu


def oracle_post(oracle, r, exposure):
    u = r.read_uniforms()
    nx, ny, nz = r.dims
    sc = oracle.Scene(np.zeros((nz, ny, nx), np.uint16))  # camera rays only read the uniforms
    return sc.post(u, r.read_hdr(), exposure)


def compare_rgba(got, want, min_equal=0.99):
    d = np.abs(got.astype(int) - want.astype(int))
    print(f"rgba: max diff {d.max()}, {np.mean(d == 0):.5f} of channels equal")
    assert d.max() <= 1
    assert np.mean(d == 0) >= min_equal


def test_dvr_display_image_and_exposure_adaptation(ctx, oracle):
    hu = phantom.ct((64, 56, 48), 3)
    r = Renderer(ctx, hu, TransferFunction.preset(0))
    w, h = 96, 72
    r.resize(w, h)
    exposure = 0.0
    for frame, cam in ((1, 0), (2, 3)):
        r.record(frame, workloads.dvr_params(r, cam, lighting=True))
        got = r.read_rgba()
        want, exposure, _, _ = oracle_post(oracle, r, exposure)
        compare_rgba(got, want)
        ex = r.read_exposure()
        print(f"frame {frame}: exposure {ex:.6f} (oracle {exposure:.6f})")
        assert abs(ex - exposure) <= 1e-5 * exposure
        exposure = ex  # the next frame adapts from the GPU's value
        assert np.all(got[..., 3] == 255)


def test_path_traced_display_image_agx(ctx, oracle):
    hu = phantom.ct((40, 40, 40), 5)
    r = Renderer(ctx, hu, TransferFunction.preset(0))
    r.resize(64, 48)
    p = workloads.pt_params(r, 0, 4, bounces=3)
    p.lighting.tone_mapper = _abi.TONE_AGX
    p.srgb_encode = 0
    r.record(1, p)
    want, _, _, _ = oracle_post(oracle, r, 0.0)
    compare_rgba(r.read_rgba(), want)


def test_render_to_rgba_is_record_and_read(ctx):
    hu = phantom.ct((32, 32, 32), 2)
    r = Renderer(ctx, hu, TransferFunction.preset(0))
    p = workloads.dvr_params(r, 1)
    a = render_to_rgba(r, p, 40, 30, frames=2)
    r2 = Renderer(ctx, hu, TransferFunction.preset(0))
    r2.resize(40, 30)
    r2.record(1, p)
    r2.record(2, p)
    assert np.array_equal(a, r2.read_rgba())
