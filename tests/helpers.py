"""Shared test helpers: volumes, uniform blocks and scenes for the oracle and the GPU."""
from __future__ import annotations

import numpy as np

from volview_amd import _abi
from volview_amd.renderer import TransferFunction, camera_framing, default_frame_params, make_volume


def flat_tf(extinction: float, albedo=(0.6, 0.6, 0.6)) -> TransferFunction:
    """Constant albedo and extinction at every HU."""
    return TransferFunction([(-1024.0, albedo, extinction), (3071.0, albedo, extinction)])


def geometry_camera(hu: np.ndarray, spacing=(1.0, 1.0, 1.0)):
    nz, ny, nx = hu.shape
    center = [0.0, 0.0, 0.0]
    diag = float(np.sqrt(sum(((n - 1) * s) ** 2 for n, s in zip((nx, ny, nz), spacing))))
    return camera_framing(center, diag), min(spacing)


def frame_params(hu: np.ndarray, *, path_tracing=False, spp=1, bounces=4, g=0.0, yaw=0.0, pitch=0.0,
                 spacing=(1.0, 1.0, 1.0)) -> _abi.FrameParams:
    cam, ms = geometry_camera(hu, spacing)
    cam.yaw, cam.pitch = yaw, pitch
    p = default_frame_params(cam, ms)
    p.lighting.shadows = 0
    p.lighting.ambient_occlusion = 0
    p.lighting.anisotropy = g
    if path_tracing:
        p.path_tracing_enabled = 1
        p.path_tracing.samples_per_frame = spp
        p.path_tracing.bounces = bounces
        p.path_tracing.denoiser = _abi.DENOISE_OFF
    return p


def uniforms(oracle, hu, tf, params, *, has_labels=False, accumulated=0, width=64, height=64, frame=1,
             spacing=(1.0, 1.0, 1.0)):
    vol = make_volume(hu, spacing)
    return oracle.pack_uniforms(vol, params, oracle.bucket_max(tf), has_labels, accumulated, width, height, frame)


def oracle_scene(oracle, hu, tf, u, labels=None, soft=None, mask=None):
    """Oracle textures for `hu` with the oracle's own normals, mask and majorant."""
    density = oracle.hu_to_unorm16(hu)
    sc = oracle.Scene(density, tf=tf, labels=labels)
    sc.set("normals", sc.normals(u))
    if labels is not None and mask is None:
        push = (0, 1, 2, 8) if soft is not None else (0, 0, 0, 0)
        dims = hu.shape[::-1] if soft is not None else None
        mask = sc.mask(u, soft, dims, push, margin=0, radius=1)
    if mask is not None:
        sc.set("mask", mask)
    sc.set("majorant", sc.majorant(u))
    return sc


def all_pixels(width, height):
    y, x = np.mgrid[0:height, 0:width]
    return np.stack([x.ravel(), y.ravel()], 1).astype(np.uint32)
