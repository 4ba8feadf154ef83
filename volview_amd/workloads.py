"""The benchmark and test workloads: seeded phantoms plus the render settings of
each configuration (volume, labels, transfer function, camera, lighting).

  DVR-1  256^3 int16 phantom, bone preset, 0.5-voxel step, brick skipping,
         cubic sampling, no labels, shadows and ambient occlusion off, 512^2.
  PT-1   512 x 512 x 400 phantom + 40 organ labels with soft margins, isolate
         mode, measured soft-tissue optics, HG g = 0.8, 8 bounces, GGX + clear
         coat, 1024^2 at 64 paths per pixel.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import _abi, phantom
from .renderer import Context, Renderer, TransferFunction, camera_framing, default_frame_params

# Group::optics of "other soft tissues" (crates/app/src/labels.rs:213-237):
# mu_s'(lambda) = a (lambda / 500 nm)^-b at 650, 550, 450 nm.
SOFT_TISSUE = (1.89, 1.286)


def measured_optics(a: float, b: float) -> tuple:
    nm = np.array([650.0, 550.0, 450.0], np.float32)
    mu = np.float32(a) * np.power(nm / np.float32(500.0), np.float32(-b))
    return (float(mu[0]), float(mu[1]), float(mu[2]), 1.0)


def organ_label_params(params: _abi.FrameParams, n_labels: int, seed: int, mode: int = _abi.LABELS_ISOLATE):
    """Seeded per-label materials: albedo in [0.3, 0.9]^3, roughness [0.2, 0.6],
    clear coat [0, 0.8], subsurface [0.2, 0.6], soft-tissue optics."""
    rng = np.random.default_rng(seed)
    lp = params.labels
    lp.mode = mode
    for k in range(1, n_labels + 1):
        albedo = rng.uniform(0.3, 0.9, 3)
        rough = rng.uniform(0.2, 0.6)
        coat = rng.uniform(0.0, 0.8)
        sss = rng.uniform(0.2, 0.6)
        lp.colors[k][:] = (*albedo, 1.0)
        lp.materials[k][:] = (rough, 0.0, sss, coat)
        lp.optics[k][:] = measured_optics(*SOFT_TISSUE)
    return params


@dataclass
class Scene:
    renderer: Renderer
    params: _abi.FrameParams
    width: int
    height: int
    n_labels: int = 0
    hu: np.ndarray | None = None
    labels: np.ndarray | None = None
    soft: np.ndarray | None = None


def seeded_camera(renderer: Renderer, index: int):
    """Camera `index` of the seeded list: 0 is OrbitCamera::framing (yaw = pitch = 0)."""
    cam = renderer.default_camera()
    if index:
        rng = np.random.default_rng(1000 + index)
        cam.yaw = float(rng.uniform(-np.pi, np.pi))
        cam.pitch = float(rng.uniform(-0.8, 0.8))
    return cam


def dvr_params(renderer: Renderer, camera_index: int = 0, lighting: bool = False) -> _abi.FrameParams:
    """DVR-1 settings on a seeded camera; `lighting` turns the shadow and
    ambient-occlusion volumes on (the reference's default Lighting)."""
    g = renderer.geometry()
    p = default_frame_params(seeded_camera(renderer, camera_index), g["min_spacing"])
    p.lighting.shadows = int(lighting)
    p.lighting.ambient_occlusion = int(lighting)
    p.lighting.environment = _abi.ENV_STUDIO
    return p


def pt_params(renderer: Renderer, n_labels: int, spp: int, bounces: int = 8, g: float = 0.8, camera_index: int = 0,
              label_seed: int = 7) -> _abi.FrameParams:
    geo = renderer.geometry()
    p = default_frame_params(seeded_camera(renderer, camera_index), geo["min_spacing"])
    p.lighting.anisotropy = g
    p.lighting.chromatic = 1.0
    p.lighting.environment = _abi.ENV_STUDIO
    p.path_tracing_enabled = 1
    p.path_tracing.bounces = bounces
    p.path_tracing.samples_per_frame = spp
    p.path_tracing.denoiser = _abi.DENOISE_OFF
    if n_labels:
        organ_label_params(p, n_labels, label_seed)
    return p


def dvr_scene(ctx: Context, dims=(256, 256, 256), seed: int = 1, size=(512, 512), camera_index: int = 0,
              lighting: bool = False) -> Scene:
    hu = phantom.ct(dims, seed)
    r = Renderer(ctx, hu, TransferFunction.preset(0))
    r.resize(*size)
    return Scene(r, dvr_params(r, camera_index, lighting), size[0], size[1], hu=hu)


def pt_scene(ctx: Context, dims=(512, 512, 400), seed: int = 1, size=(1024, 1024), spp: int = 64, n_organs: int = 40,
             camera_index: int = 0, bounces: int = 8) -> Scene:
    hu = phantom.ct(dims, seed)
    labels, soft, _, n = phantom.organs(dims, seed, n_organs)
    r = Renderer(ctx, hu, TransferFunction.preset(0))
    r.set_labels(labels, soft)
    r.resize(*size)
    p = pt_params(r, n, spp, bounces=bounces, camera_index=camera_index)
    return Scene(r, p, size[0], size[1], n, hu=hu, labels=labels, soft=soft)


# ---- Batched-32 and the serving trace ------------------------------------------------------


@dataclass
class RequestSpec:
    volume: int          # index into the workload's volumes
    path_traced: bool
    width: int
    height: int
    spp: int = 0
    camera_index: int = 0
    arrival: float = 0.0  # seconds from the start of a trace


def batched32(seed0: int = 100, n: int = 32):
    """Batched-32: n requests with seeds seed0 .. seed0 + n - 1, each on its own
    phantom: cubic volumes of edge U{128..512}, square images of edge
    U{256..1024}, half ray-marched (bone preset, DVR-1 settings), half path-traced
    (PT-1 settings without labels) with U{1..64} paths per pixel.
    Returns (volume specs [(dims, seed)], request specs)."""
    rng = np.random.default_rng(seed0)
    kinds = rng.permutation(np.arange(n) % 2 == 1)
    vols, reqs = [], []
    for k in range(n):
        edge = int(rng.integers(128, 513))
        size = int(rng.integers(256, 1025))
        spp = int(rng.integers(1, 65))
        vols.append(((edge, edge, edge), seed0 + k))
        reqs.append(RequestSpec(k, bool(kinds[k]), size, size, spp if kinds[k] else 0, camera_index=k % 5))
    return vols, reqs


def serving_trace(seconds: float, rate: float = 20.0, seed: int = 7, n_volumes: int = 8):
    """The serving workload: Poisson arrivals at `rate` per second for `seconds`;
    70 % ray-marched 1280 x 720 frames with a new camera each, 30 % progressive
    path tracing at 1024^2 with a 256-path budget; 8 resident phantoms of edge
    U{128..512} with 40 organ labels (isolate mode).
    Returns (volume specs [(dims, seed)], request specs by arrival)."""
    rng = np.random.default_rng(seed)
    vols = [((e, e, e), 200 + k) for k, e in enumerate(rng.integers(128, 513, n_volumes))]
    reqs, t = [], 0.0
    while True:
        t += float(rng.exponential(1.0 / rate))
        if t >= seconds:
            break
        pt = bool(rng.random() < 0.3)
        v = int(rng.integers(0, n_volumes))
        cam = int(rng.integers(1, 10**6))
        if pt:
            reqs.append(RequestSpec(v, True, 1024, 1024, 256, cam, t))
        else:
            reqs.append(RequestSpec(v, False, 1280, 720, 0, cam, t))
    return vols, reqs


def request_params(renderer: Renderer, spec: RequestSpec, n_labels: int = 0) -> _abi.FrameParams:
    """Frame parameters of a request: DVR-1 or PT-1 settings on its camera, with
    the volume's organ labels (isolate mode) when it has them."""
    if spec.path_traced:
        return pt_params(renderer, n_labels, spec.spp, camera_index=spec.camera_index)
    p = dvr_params(renderer, spec.camera_index)
    if n_labels:
        organ_label_params(p, n_labels, 7)
    return p
