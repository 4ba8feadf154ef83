"""GPU parity of the lighting volumes (shadow.comp, ao.comp) and of the ray
marcher reading them (dvr.comp:107-119), through the C ABI.

- The shadow volume swept on the GPU equals the FP32 oracle's sweep but for
  binary16 differences of a few spacings on a few texels: both round every slice through
  RGBA16F, and expf and fused multiply-adds differ in the last f32 bit. Cameras 0-4 put the light along
  different dominant axes and signs; one case renders with tinted labels,
  whose colours feed the diffused light.
- The ambient-occlusion volume likewise.
- DVR with shadows and AO on is within the DVR tolerance of the FP32 oracle
  reading the GPU's own lighting volumes (the lookups and shading alone), and
  at DVR-1 size with the oracle's own volumes.
- The volumes are rebuilt only when their key changes: the same frame again
  launches no lighting pass; a new camera moves the light and re-sweeps the
  shadow volume only.
"""
import numpy as np
import pytest

from test_gpu_render import compare_dvr, oracle_for
from helpers import all_pixels
from volview_amd import _abi, phantom, workloads
from volview_amd.renderer import Renderer, TransferFunction

pytestmark = pytest.mark.gpu


def f16(bits):
    return np.asarray(bits, np.uint16).view(np.float16).astype(np.float64)


def compare_f16(got, want, name):
    """At most 0.5 % of texels differ, by at most 8 binary16 spacings of the
    larger value. A one-spacing difference in one slice of the sweep carries
    into the slices after it (measured on MI355X: 0.04 % of shadow texels
    differ, by up to 4 spacings at DVR-1; occlusion is bit-exact)."""
    big = np.maximum(np.abs(got), np.abs(want)).astype(np.float16)
    spacing = np.spacing(big).astype(np.float64)
    err = np.abs(got - want)
    n_sp = err / spacing
    print(f"{name}: max |d| {err.max():.3g}, {np.mean(err > 0):.5f} of texels differ, "
          f"max {n_sp.max():.0f} spacings, counts by spacings {np.bincount(np.minimum(n_sp, 9).astype(int).ravel())}")
    assert np.all(n_sp <= 8), (name, n_sp.max())
    assert np.mean(err > 0) <= 5e-3, name


def lit_renderer(ctx, seed=4, dims=(40, 48, 36), labels=False):
    hu = phantom.ct(dims, seed)
    tf = TransferFunction.preset(0)
    r = Renderer(ctx, hu, tf)
    lab, n = None, 0
    if labels:
        lab, soft, _, n = phantom.organs(dims, seed, 10)
        r.set_labels(lab, soft)
    return r, hu, tf, lab, n


@pytest.mark.parametrize("camera,labels", [(0, False), (1, False), (2, True), (3, False), (4, False)])
def test_lighting_volumes_equal_oracle(ctx, oracle, camera, labels):
    r, hu, tf, lab, n = lit_renderer(ctx, labels=labels)
    p = workloads.dvr_params(r, camera, lighting=True)
    if labels:
        workloads.organ_label_params(p, n, 3, mode=_abi.LABELS_TINT)
    w, h = 48, 40
    r.resize(w, h)
    r.record(1, p)
    u, sc = oracle_for(oracle, r, hu, tf, p, w, h, has_labels=labels, labels=lab)
    print("light", list(u.light_dir)[:3])
    compare_f16(f16(r.read_volume(_abi.VOL_SHADOW)), f16(sc.shadow(u)), "shadow")
    compare_f16(f16(r.read_volume(_abi.VOL_AO)), f16(sc.ao(u)), "ao")
    # The frame, against the oracle reading the GPU's lighting volumes.
    sc.set("shadow", r.read_volume(_abi.VOL_SHADOW))
    sc.set("ao", r.read_volume(_abi.VOL_AO))
    got = r.read_hdr().reshape(-1, 4)
    px = all_pixels(w, h)
    compare_dvr(got, sc.dvr(u, px, 32))
    assert (got[:, 3] < 0.5).mean() > 0.01


def test_dvr_config_lit_sampled_pixels(ctx, oracle):
    """DVR-1 with the reference's default lighting: shadows and ambient occlusion on."""
    s = workloads.dvr_scene(ctx, lighting=True)
    s.renderer.set_timing(True)
    s.renderer.record(1, s.params)
    got = s.renderer.read_hdr()
    ms, n = s.renderer.kernel_time(_abi.KERNEL_LIGHTING)
    print(f"lighting volumes: {ms:.2f} ms in {n} passes; DVR {s.renderer.kernel_time(_abi.KERNEL_DVR)[0]:.2f} ms")
    s.renderer.set_timing(False)
    tf = TransferFunction.preset(0)
    u, sc = oracle_for(oracle, s.renderer, s.hu, tf, s.params, 512, 512)
    want_shadow, want_ao = sc.shadow(u), sc.ao(u)
    compare_f16(f16(s.renderer.read_volume(_abi.VOL_SHADOW)), f16(want_shadow), "shadow DVR-1")
    compare_f16(f16(s.renderer.read_volume(_abi.VOL_AO)), f16(want_ao), "ao DVR-1")
    sc.set("shadow", want_shadow)
    sc.set("ao", want_ao)
    rng = np.random.default_rng(12)
    px = np.stack([rng.integers(0, 512, 3000), rng.integers(0, 512, 3000)], 1).astype(np.uint32)
    compare_dvr(got[px[:, 1], px[:, 0]], sc.dvr(u, px, 32))


def test_lighting_rebuilt_only_when_its_key_changes(ctx):
    r, hu, tf, _, _ = lit_renderer(ctx)
    r.resize(16, 16)
    p = workloads.dvr_params(r, 0, lighting=True)
    r.set_timing(True)
    r.record(1, p)
    r.synchronize()
    assert r.kernel_time(_abi.KERNEL_LIGHTING)[1] == 2  # AO, then the shadow sweep
    r.record(2, p)
    r.synchronize()
    assert r.kernel_time(_abi.KERNEL_LIGHTING)[1] == 2
    q = workloads.dvr_params(r, 1, lighting=True)  # another camera: the light moves with it
    r.record(3, q)
    r.synchronize()
    assert r.kernel_time(_abi.KERNEL_LIGHTING)[1] == 3
    r.set_timing(False)
