"""The C ABI and the host-side logic, without a device.

- libvvr.so loads and exports every entry point include/vvr.h declares; the
  ctypes layouts match the C structs.
- Without a device, creating a context fails with an error, not a crash.
- The uniform block packing (renderer.rs:1516-1660) equals the oracle's
  restatement bit for bit, and hand-derived values for the orbit camera, the
  key light and the per-label tables.
- The phantom generator equals its numpy twin bit for bit.
"""
import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

from helpers import frame_params
from volview_amd import _abi, phantom
from volview_amd.renderer import TransferFunction, frame_uniforms, make_volume

ROOT = Path(__file__).resolve().parents[1]


def test_every_declared_entry_point_is_exported():
    lib = _abi.load()
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        names |= set(re.findall(r"\b(vvr_[a-z0-9_]+)\s*\(", h.read_text()))
    assert len(names) > 30
    for n in sorted(names):
        assert hasattr(lib, n), n
    assert names == {s[0] for s in _abi.SIGNATURES}


def test_struct_layouts_match():
    lib = _abi.load()
    for i, name in enumerate(_abi.ABI_STRUCTS):
        assert lib.vvr_abi_sizeof(i) == ctypes.sizeof(getattr(_abi, name)), name


def test_context_without_device_fails_loudly():
    lib = _abi.load()
    h = ctypes.c_void_p()
    rc = lib.vvr_ctx_create(0, ctypes.byref(h))
    if rc != 0:
        assert rc == _abi.VVR_EDEVICE
        assert lib.vvr_last_error()
    else:  # on a GPU host
        lib.vvr_ctx_destroy(h)
    assert lib.vvr_ctx_create(-1, ctypes.byref(h)) in (_abi.VVR_EINVAL, _abi.VVR_EDEVICE)


def random_params(rng, hu):
    p = frame_params(hu, path_tracing=bool(rng.integers(0, 2)), spp=int(rng.integers(1, 5)),
                     g=float(rng.uniform(0, 0.9)), yaw=float(rng.uniform(-3, 3)), pitch=float(rng.uniform(-1.5, 1.5)))
    p.labels.mode = int(rng.integers(0, 4))
    p.labels.tint = float(rng.uniform())
    p.labels.extinction = float(rng.uniform(0.1, 2))
    for k in rng.integers(1, 256, 20):
        p.labels.colors[k][:] = (*rng.uniform(0, 1, 3), float(rng.integers(0, 2)))
        p.labels.materials[k][:] = rng.uniform(0, 1, 4)
        p.labels.optics[k][:] = (*rng.uniform(0, 3, 3), 1.0)
        p.labels.surface[k][:] = (float(rng.integers(0, 3)), *rng.uniform(0, 1, 3))
        p.labels.ghost[k] = int(rng.integers(0, 2))
    p.lighting.environment = int(rng.integers(0, 3))
    p.lighting.azimuth = float(rng.uniform(-3, 3))
    p.lighting.elevation = float(rng.uniform(-1.5, 1.5))
    p.lighting.exposure = float(rng.uniform(-2, 2))
    p.lighting.subsurface_strength = float(rng.uniform(0, 2))
    p.lighting.surface_detail = float(rng.uniform(0, 2))
    for i in range(3):
        p.clipping.planes[i].enabled = int(rng.integers(0, 2))
        p.clipping.planes[i].cut_positive = int(rng.integers(0, 2))
        p.clipping.planes[i].position = float(rng.uniform(-20, 20))
    p.clipping.all = int(rng.integers(0, 2))
    p.path_tracing.aperture = float(rng.uniform(0, 2))
    p.path_tracing.has_focus = int(rng.integers(0, 2))
    p.path_tracing.focus[:] = rng.uniform(-10, 10, 3)
    return p


def test_server_argument_checks():
    lib = _abi.load()
    out = ctypes.c_void_p()
    assert lib.vvr_server_create(None, None, ctypes.byref(out)) == _abi.VVR_EINVAL
    assert lib.vvr_server_submit(None, None, None) == _abi.VVR_EINVAL
    assert lib.vvr_server_poll(None, None, 0, None, 0) == _abi.VVR_EINVAL
    assert lib.vvr_server_release(None, 1) == _abi.VVR_EINVAL
    lib.vvr_server_destroy(None)  # no-op


def test_uniform_packing_equals_oracle(oracle):
    rng = np.random.default_rng(0)
    tf = TransferFunction.preset(0)
    for trial in range(40):
        hu = np.zeros(tuple(rng.integers(4, 40, 3)), np.int16)
        spacing = tuple(rng.uniform(0.3, 2.5, 3))
        vol = make_volume(hu, spacing)
        p = random_params(rng, hu)
        args = (bool(rng.integers(0, 2)), int(rng.integers(0, 100)), int(rng.integers(1, 2000)),
                int(rng.integers(1, 2000)), int(rng.integers(0, 10**6)))
        prod = frame_uniforms(vol, tf, p, *args)
        orc = oracle.pack_uniforms(vol, p, oracle.bucket_max(tf), *args)
        assert bytes(prod) == bytes(orc), trial


def test_uniform_hand_values(oracle):
    hu = np.zeros((32, 32, 32), np.int16)
    tf = TransferFunction.preset(0)
    p = frame_params(hu)
    p.labels.mode = _abi.LABELS_ISOLATE
    p.labels.materials[3][:] = (0.1, 0.2, 0.3, 0.4)  # roughness, metalness, subsurface, coat
    p.labels.ghost[4] = 1
    p.labels.optics[5][:] = (1.0, 3.0, 2.0, 1.0)
    p.lighting.anisotropy = 0.5
    u = oracle.pack_uniforms(make_volume(hu), p, oracle.bucket_max(tf), True, 0, 200, 100, 7)
    # yaw = pitch = 0: looking along +y, right = +x, up = +z (camera.rs:35-50).
    assert list(u.forward[:3]) == [0.0, 1.0, -0.0]
    th = np.tan(np.float32(np.radians(30.0)) / 2)
    assert np.allclose(u.right[:3], [th * 2.0, 0, 0], rtol=1e-6)
    assert np.allclose(u.up[:3], [0, 0, th], rtol=1e-6)
    # Key light at azimuth -60 deg, elevation 40 deg from the view axis (renderer.rs:299-307).
    el, az = np.radians(40.0), np.radians(-60.0)
    want = np.array([np.cos(el) * np.sin(az), -np.cos(el) * np.cos(az), np.sin(el)])
    assert np.allclose(u.light_dir[:3], want, atol=1e-6)
    assert u.ambient[0] == pytest.approx(0.15)  # studio: half the ambient
    assert list(u.label_materials[3]) == pytest.approx([0.1, 0.2, 0.4, 0.3])  # coat and subsurface swap
    assert u.label_colors[4][3] == 2.0  # ghosted
    assert u.tissue[0] == pytest.approx(min(3.0 / 0.5, 0.5))  # capped at 1 / (2 mm)
    assert u.tissue[1] == pytest.approx(0.5)
    assert list(u.flags) == [7, 1, 200, 100]


def test_world_to_index_is_the_closed_form_inverse():
    """volume.rs:23-35: index_to_world = [D diag(s) | o] with D orthonormal, so
    world_to_index = [diag(1/s) D^T | -diag(1/s) D^T o]. The library inverts the
    f64 matrix and rounds to f32: equal to the closed form within 1 f32 ulp."""
    rng = np.random.default_rng(4)
    hu = np.zeros((5, 6, 7), np.int16)
    tf = TransferFunction.preset(0)
    for trial in range(30):
        q, r = np.linalg.qr(rng.normal(size=(3, 3)))
        d = q * np.sign(np.diag(r))  # a random rotation or reflection
        s = rng.uniform(0.2, 3.0, 3)
        o = rng.uniform(-300.0, 300.0, 3)
        vol = make_volume(hu, tuple(s), tuple(o), tuple(d.T.reshape(-1)))  # columns = axis directions
        u = frame_uniforms(vol, tf, frame_params(hu), False, 0, 8, 8, 1)
        lin = np.diag(1.0 / s) @ d.T
        want = np.eye(4)
        want[:3, :3] = lin
        want[:3, 3] = -lin @ o
        got = np.array(u.world_to_index, np.float32).reshape(4, 4).T  # stored column-major
        ulp = np.spacing(np.abs(want.astype(np.float32)))
        assert np.all(np.abs(got - want) <= np.maximum(ulp, 1e-12)), trial


@pytest.mark.parametrize("dims,seed", [((40, 36, 30), 1), ((33, 47, 29), 99)])
def test_phantom_equals_numpy_twin(dims, seed):
    assert np.array_equal(phantom.ct(dims, seed), phantom.ct_twin(dims, seed))
    labels, soft, margin, n = phantom.organs(dims, seed, 40)
    tl, ts, tm, tn = phantom.organs_twin(dims, seed, 40)
    assert n == tn and n > 0
    assert np.array_equal(labels, tl)
    assert np.array_equal(soft, ts)
    assert np.array_equal(margin.view(np.uint16), tm.view(np.uint16))


def test_phantom_structures():
    hu = phantom.ct((128, 128, 128), 3)
    assert hu[0, 0, 0] == -1000  # air, no noise
    body = hu[30, 60:68, 60:70]  # below the lungs, off the vessel
    assert abs(body.mean() - 40) < 10 and 10 < body.std() < 30
    assert (hu > 700).sum() > 1000  # bone
    assert (hu < -600).sum() > (hu == -1000).sum() * 0 + 10000  # lungs and air
    labels, soft, margin, n = phantom.organs((128, 128, 128), 3, 40)
    assert n >= 10
    present = np.unique(labels)
    assert present[0] == 0 and len(present) == n + 1
    inside = labels > 0
    assert np.all(margin[inside].astype(np.float32) <= 0)
    assert np.all(soft[inside][:, 0] == labels[inside])
