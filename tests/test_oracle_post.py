"""Pins of the oracle's camera finishing (bloom.comp, tonemap.comp).

- linear_to_srgb is the IEC 61966-2-1 encoding: 12.92 c below 0.0031308, else
  1.055 c^(1/2.4) - 0.055 (values from the standard's float64 form).
- Narkowicz's ACES fit: 0 at 0, (2.51 + 0.03) / (2.43 + 0.59 + 0.14) at 1,
  increasing, and clamped to 1 above the root of 0.08 x^2 - 0.56 x - 0.14.
- AgX: black stays black; a very bright grey reaches the sigmoid's value at 1,
  (15.5 - 40.14 + 31.96 - 6.868 + 0.4298 + 0.1191 - 0.00232), through the
  outset matrix's column sums, then gamma 2.2.
- The grain hash is uniform on [0, 1).
- A constant covered HDR image: the Karis average keeps it, each 9-tap blur
  scales it by the taps' sum 0.9996, and the first auto exposure is
  clamp(0.24 / luminance, 0.25, 4); the next frame adapts 15 % of the way in
  log space.
- With bloom, vignette and grain off and no tone curve, a pixel is the sRGB
  encoding of its exposed value; the vignette darkens a corner by
  1 - v (sqrt(2) 0.7071)^2.5.
"""
import numpy as np
import pytest

from helpers import flat_tf, frame_params, oracle_scene, uniforms

LUM = np.array([0.2126, 0.7152, 0.0722])


def srgb64(c):
    c = np.clip(c, 0.0, 1.0)
    return np.where(c < 0.0031308, 12.92 * c, 1.055 * np.power(c, 1 / 2.4) - 0.055)


def test_linear_to_srgb_standard(oracle):
    assert oracle.linear_to_srgb(0.0) == 0.0
    assert abs(oracle.linear_to_srgb(1.0) - 1.0) < 1e-6
    assert abs(oracle.linear_to_srgb(0.5) - 0.7353569830524495) < 2e-7
    for c in np.random.default_rng(1).uniform(0, 1.2, 200):
        assert abs(oracle.linear_to_srgb(float(c)) - srgb64(np.float32(c))) < 3e-7


def test_aces_fit(oracle):
    assert oracle.aces(0.0) == 0.0
    assert abs(oracle.aces(1.0) - 2.54 / 3.16) < 1e-6
    xs = np.linspace(0, 7.2, 400)
    v = [oracle.aces(float(x)) for x in xs]
    assert np.all(np.diff(v) > 0)
    root = (0.56 + np.sqrt(0.56 ** 2 + 4 * 0.08 * 0.14)) / 0.16
    assert oracle.aces(root + 0.05) == 1.0 and oracle.aces(root - 0.05) < 1.0


def test_agx_black_and_white(oracle):
    assert np.all(oracle.agx([0.0, 0.0, 0.0]) == 0.0)
    top = 15.5 - 40.14 + 31.96 - 6.868 + 0.4298 + 0.1191 - 0.00232
    outset_cols = np.array([1.19687900512017 - 0.0980208811401368 - 0.0990297440797205,
                            -0.0528968517574562 + 1.15190312990417 - 0.0989611768448433,
                            -0.0529716355144438 - 0.0980434501171241 + 1.15107367264116])
    want = np.clip(top * outset_cols, 0, 1) ** 2.2
    assert np.allclose(oracle.agx([1e6, 1e6, 1e6]), want, atol=1e-5)
    g = [oracle.agx([x, x, x])[1] for x in np.geomspace(1e-3, 10, 60)]
    assert np.all(np.diff(g) > 0)


def test_grain_hash_uniform(oracle):
    v = np.array([oracle.hash01(i * 7919 + 3) for i in range(20000)])
    assert v.min() >= 0.0 and v.max() < 1.0
    assert abs(v.mean() - 0.5) < 0.01
    hist, _ = np.histogram(v, bins=10, range=(0, 1))
    assert hist.min() > 1800 and hist.max() < 2200


def post_scene(oracle, w, h, finish=(0.0, 0.0, 0.0, 0.0), tone=0, exposure_stops=0.0, srgb=1):
    hu = np.zeros((8, 8, 8), np.int16)
    tf = flat_tf(0.1)
    p = frame_params(hu)
    p.lighting.bloom, p.lighting.vignette, p.lighting.grain = finish[:3]
    p.lighting.auto_exposure = int(finish[3])
    p.lighting.tone_mapper = tone
    p.lighting.exposure = exposure_stops
    p.srgb_encode = srgb
    u = uniforms(oracle, hu, tf, p, width=w, height=h)
    return oracle_scene(oracle, hu, tf, u), u


def test_bloom_and_exposure_of_a_constant_image(oracle):
    w, h = 40, 28
    sc, u = post_scene(oracle, w, h, finish=(0.04, 0.0, 0.0, 1.0))
    c = np.array([0.3, 0.5, 0.2], np.float32)
    hdr = np.zeros((h, w, 4), np.float32)
    hdr[..., :3] = c  # alpha 0: covered
    _, ex, a, cc = sc.post(u, hdr)
    taps = 0.2270 + 2 * (0.1945 + 0.1216 + 0.0540 + 0.0162)
    assert abs(taps - 0.9996) < 1e-12
    assert np.allclose(a[..., :3], c * taps ** 2, rtol=2e-3)  # binary16 stores
    assert np.allclose(a[..., 3], taps ** 2, rtol=2e-3)
    assert np.allclose(cc[..., :3], c * taps ** 4, rtol=3e-3)
    lum = float(np.float16(c @ LUM))
    first = np.clip(0.24 / lum, 0.25, 4.0)
    assert abs(ex - first) < 2e-3 * first
    # A brighter image next frame: 15 % of the way in log space.
    _, ex2, _, _ = sc.post(u, hdr * 4, exposure=ex)
    target = np.clip(0.24 / (4 * lum), 0.25, 4.0)
    assert abs(ex2 - np.exp(np.log(ex) + 0.15 * (np.log(target) - np.log(ex)))) < 2e-3 * ex2


@pytest.mark.parametrize("stops", [0.0, 1.0])
def test_plain_encoding_and_vignette(oracle, stops):
    w, h = 16, 12
    rng = np.random.default_rng(4)
    hdr = np.zeros((h, w, 4), np.float32)
    hdr[..., :3] = rng.uniform(0, 0.9, (h, w, 3)).astype(np.float16)
    sc, u = post_scene(oracle, w, h, exposure_stops=stops)
    out, _, _, _ = sc.post(u, hdr)
    want = np.floor(srgb64(hdr[..., :3].astype(np.float64) * 2.0 ** stops) * 255 + 0.5)
    assert np.abs(out[..., :3].astype(int) - want).max() <= 1
    assert np.all(out[..., 3] == 255)
    v = 0.25
    sc, u = post_scene(oracle, w, h, finish=(0.0, v, 0.0, 0.0), srgb=0)
    flat = np.zeros((h, w, 4), np.float32)
    flat[..., :3] = 0.5
    out, _, _, _ = sc.post(u, flat)
    uv = (np.array([0.5, 0.5]) / np.array([w, h])) * 2 - 1  # pixel (0, 0)
    r = np.sqrt((uv ** 2).sum()) * 0.7071
    assert abs(int(out[0, 0, 0]) - np.floor(0.5 * (1 - v * r ** 2.5) * 255 + 0.5)) <= 1
