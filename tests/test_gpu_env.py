"""The light map (Environment::Map) on the GPU, through the C ABI.

- set_environment_map builds the sampling table and the brightest column
  equal, bit for bit, to the oracle's restatement of EnvironmentMap::prepare.
- Path tracing with the map (importance-sampled light-map rays at every
  event, MIS on escaped paths) traces the oracle's paths: same random
  streams, nearly every pixel's sums agree.
- The ray marcher's ambient term takes the map's diffuse response, within the
  DVR tolerance of the oracle.
- Removing the map falls back to the studio.
"""
import numpy as np
import pytest

from helpers import all_pixels
from test_gpu_render import compare_dvr, oracle_for
from volview_amd import _abi, phantom, workloads
from volview_amd.renderer import Renderer, TransferFunction, make_volume

pytestmark = pytest.mark.gpu


def light_map(seed=4, w=64, h=32):
    """A dim sky with a bright sun and a warm floor."""
    rng = np.random.default_rng(seed)
    rgb = np.full((h, w, 3), 0.3, np.float32) * rng.uniform(0.8, 1.2, (h, w, 1)).astype(np.float32)
    rgb[h // 2:] *= np.array([1.2, 0.9, 0.6], np.float32)
    rgb[6:9, 40:44] = 60.0
    return rgb


def with_map(oracle, r, hu, tf, params, w, h, rgb, has_labels=False, accumulated=0, labels=None):
    rad, table, dif, peak = oracle.env_prepare(rgb)
    u = oracle.pack_uniforms(make_volume(hu), params, oracle.bucket_max(tf), has_labels, accumulated, w, h, 1,
                             env_peak_u=peak)
    _, sc = oracle_for(oracle, r, hu, tf, params, w, h, has_labels=has_labels, accumulated=accumulated, labels=labels)
    sc.set_env(rad, table, dif)
    return u, sc


def test_tables_equal_oracle(ctx, oracle):
    hu = phantom.ct((16, 16, 16), 1)
    r = Renderer(ctx, hu, TransferFunction.preset(0))
    rgb = light_map()
    r.set_environment_map(rgb)
    peak, table = r.read_environment()
    _, want_table, _, want_peak = oracle.env_prepare(rgb)
    assert peak == want_peak
    assert np.array_equal(table, want_table)


def test_pt_with_light_map_traces_the_oracles_paths(ctx, oracle):
    dims = (40, 36, 32)
    hu = phantom.ct(dims, 7)
    tf = TransferFunction.preset(0)
    r = Renderer(ctx, hu, tf)
    labels, sg, _, n = phantom.organs(dims, 7, 8)
    r.set_labels(labels, sg)
    rgb = light_map()
    r.set_environment_map(rgb)
    w, h, spp = 32, 24, 2
    r.resize(w, h)
    p = workloads.pt_params(r, n, spp, bounces=4, camera_index=1)
    p.lighting.environment = _abi.ENV_MAP
    r.record(1, p)
    acc, alb, nrm = r.read_accum()
    want = [np.zeros((w * h, 4)) for _ in range(3)]
    for k in range(spp):
        q = workloads.pt_params(r, n, 1, bounces=4, camera_index=1)
        q.lighting.environment = _abi.ENV_MAP
        u, sc = with_map(oracle, r, hu, tf, q, w, h, rgb, has_labels=True, accumulated=k, labels=labels)
        assert u.lighting[3] == _abi.ENV_MAP
        for a, part in zip(want, sc.pt(u, all_pixels(w, h))):
            a += part
    err = np.abs(acc.reshape(-1, 4) - want[0])
    off = np.any(err > 1e-3 * (1 + np.abs(want[0])), axis=-1)
    print(f"radiance with a light map: {off.mean():.4f} of pixels differ")
    assert off.mean() < 0.10
    assert np.abs(acc[..., :3]).max() > 0


def test_dvr_ambient_from_light_map(ctx, oracle):
    hu = phantom.ct((48, 48, 40), 2)
    tf = TransferFunction.preset(0)
    r = Renderer(ctx, hu, tf)
    rgb = light_map(seed=5)
    r.set_environment_map(rgb)
    w, h = 64, 48
    r.resize(w, h)
    p = workloads.dvr_params(r, 2)
    p.lighting.environment = _abi.ENV_MAP
    r.record(1, p)
    got = r.read_hdr().reshape(-1, 4)
    u, sc = with_map(oracle, r, hu, tf, p, w, h, rgb)
    compare_dvr(got, sc.dvr(u, all_pixels(w, h), 32))
    r.set_environment_map(None)
    r.record(2, p)
    assert r.read_uniforms().lighting[3] == _abi.ENV_STUDIO
