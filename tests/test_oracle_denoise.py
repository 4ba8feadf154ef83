"""Pins of the oracle's a-trous denoiser (denoise.comp) by properties that the
filter's definition fixes.

- A constant image passes unchanged (the B3 taps and every edge weight are
  normalised), background share included.
- With no per-pixel variance and many samples (the neighbourhood term fades
  as 4 / n) the luminance weight stops every neighbour of another luminance:
  the filter returns the path-traced image (the convergence the shader's
  header promises).
- Normals at right angles stop the filter exactly (max(n.m, 0)^64 = 0): a flat
  region beside a noisy one stays flat up to the edge.
- One level on a flat image with per-pixel variance v leaves v (sum k^2)^2 in
  the interior, k = (1/16, 1/4, 3/8, 1/4, 1/16), and v (53/121)^2 in a corner,
  where taps outside the image are skipped rather than clamped.
"""
import numpy as np
import pytest

KERNEL = np.array([1 / 16, 1 / 4, 3 / 8, 1 / 4, 1 / 16])
LUM = np.array([0.2126, 0.7152, 0.0722])


def images(irradiance, albedo, n, normal=(0.0, 0.0, 1.0), depth=100.0, extra_var=0.0, background=0.25):
    """Accumulation and guide images whose means are the given irradiance x albedo."""
    h, w = irradiance.shape[:2]
    radiance = irradiance * albedo
    accum = np.zeros((h, w, 4), np.float32)
    accum[..., :3] = radiance * n
    accum[..., 3] = background * n
    lum = radiance @ LUM
    alb = np.zeros((h, w, 4), np.float32)
    alb[..., :3] = albedo * n
    alb[..., 3] = n * (lum * lum + extra_var)  # sum of squared luminances: temporal variance extra_var / n
    nrm = np.zeros((h, w, 4), np.float32)
    nrm[..., :3] = np.broadcast_to(np.asarray(normal, np.float32), (h, w, 3)) * n
    nrm[..., 3] = depth * n
    return accum, alb, nrm


@pytest.mark.parametrize("precision", [32, 64])
def test_constant_image_passes_unchanged(oracle, precision):
    h, w, n = 24, 20, 16.0
    irr = np.full((h, w, 3), (0.8, 0.6, 0.3), np.float32)
    alb = np.full((h, w, 3), (0.5, 0.7, 0.9), np.float32)
    hdr, _ = oracle.denoise(*images(irr, alb, n, extra_var=0.3), n, precision=precision)
    assert np.allclose(hdr[..., :3], irr * alb, rtol=1e-5)
    assert np.allclose(hdr[..., 3], 0.25)


def test_no_variance_returns_the_image(oracle):
    # No per-pixel variance, and the neighbourhood spread fades as 4 / n: at
    # 10^7 samples the luminance weight of a neighbour 0.1 away is ~exp(-20).
    h, w, n = 10, 12, 1.0e7
    k = np.arange(h * w, dtype=np.float64).reshape(h, w)
    irr = np.stack([0.1 * k + 0.1] * 3, -1).astype(np.float32)  # luminances 0.1 apart
    alb = np.full((h, w, 3), 0.5, np.float32)
    hdr, _ = oracle.denoise(*images(irr, alb, n), n)
    assert np.allclose(hdr[..., :3], irr * alb, rtol=1e-5, atol=1e-7)


def test_normals_at_right_angles_stop_the_filter(oracle):
    h, w, n = 16, 16, 8.0
    rng = np.random.default_rng(5)
    irr = np.ones((h, w, 3), np.float32)
    irr[:, 8:] = rng.uniform(0.5, 1.5, (h, 8, 1))  # noisy right half
    alb = np.full((h, w, 3), 0.6, np.float32)
    acc, ag, ng = images(irr, alb, n, extra_var=2.0)
    ng[:, 8:, :3] = np.array([1.0, 0.0, 0.0]) * n  # the right half faces another way
    hdr, _ = oracle.denoise(acc, ag, ng, n)
    assert np.allclose(hdr[:, :8, :3], 0.6, rtol=1e-6)
    # Without the edge the flat half takes up the noise.
    acc, ag, ng = images(irr, alb, n, extra_var=2.0)
    hdr2, _ = oracle.denoise(acc, ag, ng, n)
    assert np.abs(hdr2[:, 6:8, :3] - 0.6).max() > 1e-3


def test_variance_propagation_closed_form(oracle):
    h, w, n = 12, 14, 4.0
    irr = np.full((h, w, 3), 1.0, np.float32)
    alb = np.full((h, w, 3), 0.5, np.float32)
    t = 0.08
    _, last = oracle.denoise(*images(irr, alb, n, extra_var=t), n, levels=1, precision=64)
    a = 0.5  # luminance of the mean albedo
    v = (t / n) / (a * a)
    s2 = float((KERNEL ** 2).sum())
    assert np.allclose(last[2:-2, 2:-2, 3], v * s2 * s2, rtol=1e-6)
    assert np.isclose(last[0, 0, 3], v * (53 / 121) ** 2, rtol=1e-6)
    assert np.allclose(last[..., :3], 1.0, rtol=1e-6)
