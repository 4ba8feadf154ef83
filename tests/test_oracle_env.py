"""Pins of the oracle's light map (environment.rs:56-180, pt.comp:192-233,
common.glsl:457-485) by properties of importance sampling and the sphere.

- The sampling pdf integrates to 1 over the sphere (the table's probabilities
  are normalised and the solid-angle Jacobian 2 pi^2 sin(theta) is right).
- Drawn directions have, as their pdf, env_pdf of the direction (sample_env
  and env_uv / env_direction invert each other, at any map rotation), and
  E[1 / pdf] = 4 pi and E[g / pdf] = integral of g for g = 1, z, z^2.
- A uniform map normalises to radiance 1 and its cosine-convolved diffuse map
  is 1 (to the discretisation); a single bright texel takes the peak column
  and almost every sample.
- The uniform block selects the map only when one is set, rotated so the
  brightest column faces the key light: (peak - 1/2) 2 pi - atan2(sin a, -cos a)
  + rotation.
"""
import numpy as np
import pytest

from helpers import flat_tf, frame_params, oracle_scene
from volview_amd import _abi
from volview_amd.renderer import make_volume


def scene_with_map(oracle, rgb, rotation=0.0):
    hu = np.zeros((8, 8, 8), np.int16)
    tf = flat_tf(0.1)
    p = frame_params(hu, path_tracing=True)
    p.lighting.environment = _abi.ENV_MAP
    p.lighting.env_rotation = rotation
    rad, table, dif, peak = oracle.env_prepare(rgb)
    u = oracle.pack_uniforms(make_volume(hu), p, oracle.bucket_max(tf), False, 0, 16, 16, 1, env_peak_u=peak)
    sc = oracle_scene(oracle, hu, tf, u)
    sc.set_env(rad, table, dif)
    return sc, u, (rad, table, dif, peak), p, hu, tf


def test_pdf_integrates_to_one(oracle):
    rng = np.random.default_rng(2)
    rgb = rng.uniform(0.05, 2.0, (8, 16, 3)) ** 2
    sc, u, *_ = scene_with_map(oracle, rgb, rotation=0.7)
    # Integrate over the map's own (u, v) grid, 4 x 4 points per texel.
    nu, nv = 16 * 4, 8 * 4
    total = 0.0
    for j in range(nv):
        for i in range(nu):
            uu, vv = (i + 0.5) / nu, (j + 0.5) / nv
            phi = (uu - 0.5) * 2 * np.pi - u.env[0]
            theta = vv * np.pi
            f, r, up = (np.array(v[:3]) / np.linalg.norm(v[:3]) for v in (u.forward, u.right, u.up))
            d = np.sin(theta) * (np.sin(phi) * r + np.cos(phi) * f) + np.cos(theta) * up
            total += sc.env_pdf(u, d) * 2 * np.pi ** 2 * np.sin(theta) / (nu * nv)
    assert abs(total - 1.0) < 2e-3


def test_samples_carry_their_pdf_and_integrate_test_functions(oracle):
    rng = np.random.default_rng(3)
    rgb = rng.uniform(0.2, 3.0, (16, 32, 3))
    sc, u, *_ = scene_with_map(oracle, rgb, rotation=-1.3)
    dirs, pdfs = sc.sample_env(u, 12345, 60000)
    assert np.allclose(np.linalg.norm(dirs, axis=1), 1.0, atol=1e-5)
    for k in range(0, 60000, 997):
        assert abs(sc.env_pdf(u, dirs[k]) - pdfs[k]) <= 1e-4 * pdfs[k]
    z = dirs[:, 2]
    assert abs(np.mean(1 / pdfs) / (4 * np.pi) - 1) < 0.02
    assert abs(np.mean(z / pdfs)) < 0.05 * 4 * np.pi
    assert abs(np.mean(z * z / pdfs) / (4 * np.pi / 3) - 1) < 0.03


def test_uniform_map_and_single_bright_texel(oracle):
    rad, table, dif, _ = oracle.env_prepare(np.full((16, 32, 3), 2.5, np.float32))
    f16 = lambda a: a.view(np.float16).astype(np.float64)
    assert np.allclose(f16(rad)[..., :3], 1.0, atol=1e-3)
    assert np.allclose(f16(dif)[..., :3], 1.0, atol=0.02)
    rgb = np.full((16, 32, 3), 1e-3, np.float32)
    rgb[5, 21] = 5000.0
    sc, u, (rad, table, dif, peak), *_ = scene_with_map(oracle, rgb)
    assert abs(peak - (21 + 0.5) / 32) < 1e-6
    dirs, pdfs = sc.sample_env(u, 777, 2000)
    assert np.mean(pdfs > 0.5 * pdfs.max()) > 0.98  # nearly all land in the bright texel


@pytest.mark.parametrize("azimuth,rotation", [(-60.0, 0.0), (30.0, 0.4)])
def test_uniforms_select_and_rotate_the_map(oracle, azimuth, rotation):
    rgb = np.random.default_rng(9).uniform(0.1, 1.0, (8, 16, 3))
    hu = np.zeros((8, 8, 8), np.int16)
    tf = flat_tf(0.1)
    p = frame_params(hu, path_tracing=True)
    p.lighting.environment = _abi.ENV_MAP
    p.lighting.azimuth = np.radians(azimuth)
    p.lighting.env_rotation = rotation
    vol, bmax = make_volume(hu), oracle.bucket_max(tf)
    no_map = oracle.pack_uniforms(vol, p, bmax, False, 0, 16, 16, 1)
    assert no_map.lighting[3] == _abi.ENV_STUDIO and no_map.env[0] == 0.0
    peak = oracle.env_prepare(rgb)[3]
    u = oracle.pack_uniforms(vol, p, bmax, False, 0, 16, 16, 1, env_peak_u=peak)
    a = np.float32(p.lighting.azimuth)
    want = (peak - 0.5) * 2 * np.pi - np.arctan2(np.sin(a), -np.cos(a)) + rotation
    assert u.lighting[3] == _abi.ENV_MAP
    assert abs(u.env[0] - want) < 1e-5
