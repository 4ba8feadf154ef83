"""Pins of the oracle's lighting volumes (shadow.comp, ao.comp) and their lookup
(lighting_lookup, common.glsl:141-156) by closed forms.

- Empty volume: direct and diffused light 1 everywhere, ambient occlusion 1.
- Homogeneous medium lit along an axis: the direct transmittance of the slice k
  slices from the lit side is exp(-sigma s)^(k+1), s the slice spacing in mm
  (2 voxels); the diffused light in the middle of the slab decays per channel by
  exp(-sigma y (1.2 - chroma) s) a slice (the 0.4 + 4 x 0.15 gather sums to one),
  each step rounded to binary16 as the RGBA16F image stores it.
- The sweep starts on the lit side for either sign and each axis: a slab at one
  end of the volume shadows the far side only when the light is behind it.
- Ambient occlusion of a homogeneous medium is min(2 exp(-15 sigma), 1) at every
  texel (16 directions x 8 steps of 15/8 mm).
- Lookups: a linear ramp stored in the shadow volume is read back exactly by the
  linear and the cubic filter inside the image (B-spline weights reproduce affine
  functions), which pins the texture-to-texel mapping at the volume's own
  resolution; far outside the image the border reads 1.
"""
import numpy as np
import pytest

from helpers import flat_tf, frame_params, oracle_scene, uniforms
from volview_amd.renderer import TransferFunction


def lit_scene(oracle, hu, tf, light, shadows=True, ao=False, smooth=True):
    p = frame_params(hu)
    p.lighting.shadows = int(shadows)
    p.lighting.ambient_occlusion = int(ao)
    p.smooth_sampling = int(smooth)
    u = uniforms(oracle, hu, tf, p, width=16, height=16)
    if light is not None:
        u.light_dir[0], u.light_dir[1], u.light_dir[2] = light
    return oracle_scene(oracle, hu, tf, u), u


def f16(bits):
    return np.asarray(bits, np.uint16).view(np.float16).astype(np.float64)


def test_empty_volume_is_fully_lit(oracle):
    hu = np.full((12, 20, 16), -1000, np.int16)
    tf = TransferFunction.preset(0)
    sc, u = lit_scene(oracle, hu, tf, (0.3, -0.5, 0.81), ao=True)
    assert np.all(f16(sc.shadow(u)) == 1.0)
    assert np.all(f16(sc.ao(u)) == 1.0)


@pytest.mark.parametrize("precision", [32, 64])
def test_homogeneous_slab_closed_form(oracle, precision):
    sigma, albedo = 0.15, (0.8, 0.4, 0.2)
    hu = np.zeros((24, 64, 64), np.int16)  # (z, y, x): shadow volume 32 x 32 x 12
    tf = flat_tf(sigma, albedo)
    sc, u = lit_scene(oracle, hu, tf, (0.0, 0.0, 1.0))
    lut = oracle.tf_lut_f16(tf).view(np.float16).astype(np.float64)[0]
    sig, alb = lut[3], lut[:3]
    s = 2.0  # mm between shadow slices along z (1 mm voxels, 2x downsampled)
    chroma = alb / alb.max()
    k_ind = sig * u.subsurface[1] * (1.2 - chroma)
    out = f16(sc.shadow(u, precision))
    nz = out.shape[0]
    direct, ind = 1.0, np.ones(3)
    for k in range(nz):
        z = nz - 1 - k  # lit side first
        direct = float(np.float16(direct * np.exp(-sig * s)))
        ind = np.float16(ind * np.exp(-k_ind * s)).astype(np.float64)
        got = out[z, 16, 16]
        assert abs(got[3] - direct) <= 2 * 2.0 ** -11 * direct, (k, got[3], direct)
        assert np.all(np.abs(got[:3] - ind) <= 2 * 2.0 ** -11 * ind), (k, got[:3], ind)
        # Direct light does not spread sideways: the whole slice agrees.
        assert np.all(out[z, :, :, 3] == got[3])


@pytest.mark.parametrize("axis,sign", [(2, 1), (2, -1), (0, 1), (0, -1), (1, 1)])
def test_sweep_starts_on_the_lit_side(oracle, axis, sign):
    n = 32
    hu = np.full((n, n, n), -1000, np.int16)
    sl = [slice(None)] * 3
    sl[2 - axis] = slice(n - 6, n)  # a slab at the high end of `axis`
    hu[tuple(sl)] = 1000
    alb = (0.7, 0.7, 0.7)
    tf = TransferFunction([(-1024.0, alb, 0.0), (200.0, alb, 0.0), (300.0, alb, 4.0), (3071.0, alb, 4.0)])
    light = [0.0, 0.0, 0.0]
    light[axis] = float(sign)
    sc, u = lit_scene(oracle, hu, tf, light)
    out = f16(sc.shadow(u))[..., 3]
    low = [slice(None)] * 3
    low[2 - axis] = 0  # the slice at the low end, across the volume from the slab
    far = out[tuple(low)]
    if sign > 0:  # light beyond the slab: the low end is in its shadow
        assert far.max() < 1e-3
    else:  # light on the low side: nothing between it and the low end
        assert np.all(far == 1.0)


def test_ambient_occlusion_homogeneous_closed_form(oracle):
    sigma = 0.05
    hu = np.zeros((16, 24, 20), np.int16)
    tf = flat_tf(sigma)
    sc, u = lit_scene(oracle, hu, tf, None, shadows=False, ao=True)
    sig = oracle.tf_lut_f16(tf).view(np.float16).astype(np.float64)[0, 3]
    want = min(2.0 * np.exp(-15.0 * sig), 1.0)
    for precision in (32, 64):
        got = f16(sc.ao(u, precision))
        assert got.shape == (4, 6, 5)
        assert np.all(np.abs(got - want) <= 2.0 ** -11 * want), (got.min(), got.max(), want)


@pytest.mark.parametrize("smooth", [False, True])
def test_lookup_reproduces_a_ramp_and_reads_the_border(oracle, smooth):
    hu = np.zeros((20, 24, 28), np.int16)  # shadow volume 14 x 12 x 10
    tf = flat_tf(0.1)
    sc, u = lit_scene(oracle, hu, tf, None, smooth=smooth)
    sx, sy, sz = sc.shadow_dims()
    zz, yy, xx = np.mgrid[0:sz, 0:sy, 0:sx]
    vol = np.zeros((sz, sy, sx, 4), np.float16)
    vol[..., 0] = xx / 64.0
    vol[..., 1] = yy / 64.0
    vol[..., 2] = zz / 64.0
    vol[..., 3] = 0.25
    sc.set("shadow", vol.view(np.uint16))
    rng = np.random.default_rng(3)
    for _ in range(50):
        t = rng.uniform([2, 2, 2], [sx - 3, sy - 3, sz - 3])  # texel-space position inside the image
        tc = (t + 0.5) / np.array([sx, sy, sz])
        got = sc.lighting_lookup(u, 0, tc)
        assert np.allclose(got[:3], t / 64.0, atol=2e-6), (t, got)
        assert abs(got[3] - 0.25) < 1e-6
    assert np.all(sc.lighting_lookup(u, 0, (-2.0, 3.0, 0.5)) == 1.0)
