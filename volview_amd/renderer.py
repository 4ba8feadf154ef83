"""Python host of the MI355X renderer, mirroring volview's `render::Renderer`
(crates/render/src/renderer.rs) and `render_to_rgba` (crates/render/src/offscreen.rs)
over the C ABI of libvvr.so.
"""
from __future__ import annotations

import ctypes
from ctypes import POINTER, byref, c_double, c_float, c_int16, c_uint8, c_uint16, c_uint32, c_uint64, c_void_p

import numpy as np

from . import _abi
from ._abi import (Camera, ControlPoint, EnvMap, FrameParams, LabelParams, Lighting, PathTracing, PtShape, SoftGrid, TF,
                   Uniforms, Volume, check, load)


def _fptr(a: np.ndarray):
    return a.ctypes.data_as(POINTER(c_float))


class TransferFunction:
    """transfer::TransferFunction (transfer.rs:26-30): control points (hu, rgb, extinction)."""

    def __init__(self, points):
        self.points = [(float(h), tuple(float(c) for c in col), float(e)) for h, col, e in points]
        arr = (ControlPoint * len(self.points))()
        for i, (h, col, e) in enumerate(self.points):
            arr[i].hu = h
            arr[i].color[:] = col
            arr[i].extinction = e
        self._arr = arr
        self.c = TF(len(self.points), arr)

    @staticmethod
    def preset(index: int = 0) -> "TransferFunction":
        """TransferFunction::presets (transfer.rs:32-69): 0 bone, 1 angio, 2 soft tissue."""
        lib = load()
        buf = (ControlPoint * 16)()
        n = lib.vvr_tf_preset(index, buf, 16)
        if n < 0:
            raise ValueError(f"no preset {index}")
        return TransferFunction([(buf[i].hu, tuple(buf[i].color), buf[i].extinction) for i in range(n)])

    def lut(self) -> np.ndarray:
        out = np.zeros((4096, 4), np.float32)
        check(load().vvr_tf_lut(byref(self.c), _fptr(out)))
        return out

    def lut_f16(self) -> np.ndarray:
        out = np.zeros((4096, 4), np.uint16)
        check(load().vvr_tf_lut_f16(byref(self.c), out.ctypes.data_as(POINTER(c_uint16))))
        return out

    def bucket_max(self) -> np.ndarray:
        out = np.zeros(64, np.float32)
        check(load().vvr_tf_bucket_max(byref(self.c), _fptr(out)))
        return out

    def cumulative_integrals(self) -> np.ndarray:
        out = np.zeros((4096, 4), np.float32)
        check(load().vvr_tf_cumulative_integrals(byref(self.c), _fptr(out)))
        return out

    def evaluate(self, hu: float) -> np.ndarray:
        out = np.zeros(4, np.float32)
        load().vvr_tf_evaluate(byref(self.c), hu, _fptr(out))
        return out


def hu_to_unorm16(hu: int) -> int:
    return int(load().vvr_hu_to_unorm16(int(hu)))


def make_volume(hu: np.ndarray, spacing=(1.0, 1.0, 1.0), origin=None, direction=None) -> Volume:
    """volio::Volume for an array indexed [k, j, i] (x fastest in memory)."""
    hu = np.ascontiguousarray(hu, dtype=np.int16)
    nz, ny, nx = hu.shape
    v = Volume()
    v.dims[:] = (nx, ny, nz)
    v.spacing[:] = spacing
    if origin is None:  # centred on the origin
        origin = [-(n - 1) * 0.5 * s for n, s in zip((nx, ny, nz), spacing)]
    v.origin[:] = origin
    v.direction[:] = direction if direction is not None else (1, 0, 0, 0, 1, 0, 0, 0, 1)
    v.hu = hu.ctypes.data_as(POINTER(c_int16))
    v._keep = hu
    return v


def frame_uniforms(volume: Volume, tf: TransferFunction, params: FrameParams, has_labels: bool, accumulated: int,
                   width: int, height: int, frame: int) -> Uniforms:
    """The uniform block record() would write (host-only, no device needed)."""
    u = Uniforms()
    check(load().vvr_frame_uniforms(byref(volume), byref(tf.c), byref(params), 1 if has_labels else 0, accumulated,
                                    width, height, frame, byref(u), ctypes.sizeof(u)))
    return u


def camera_framing(center, diagonal: float) -> Camera:
    cam = Camera()
    c = (c_float * 3)(*center)
    load().vvr_camera_framing(c, diagonal, byref(cam))
    return cam


def default_frame_params(camera: Camera, min_spacing: float) -> FrameParams:
    p = FrameParams()
    load().vvr_frame_params_default(byref(p), byref(camera), min_spacing)
    return p


class Context:
    """Context::new (crates/render/src/context.rs:37): one HIP device."""

    def __init__(self, device: int = 0):
        self._lib = load()
        h = c_void_p()
        check(self._lib.vvr_ctx_create(device, byref(h)))
        self.handle = h

    def close(self):
        if self.handle:
            self._lib.vvr_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Renderer:
    """render::Renderer over libvvr: new / set_labels / set_transfer_function /
    resize / record, plus read-backs that synchronise."""

    def __init__(self, ctx: Context, hu: np.ndarray, tf: TransferFunction, spacing=(1.0, 1.0, 1.0), origin=None,
                 direction=None):
        self._lib = load()
        self.ctx = ctx
        self.dims = tuple(int(d) for d in hu.shape[::-1])  # (nx, ny, nz)
        vol = make_volume(hu, spacing, origin, direction)
        h = c_void_p()
        check(self._lib.vvr_renderer_new(ctx.handle, byref(vol), byref(tf.c), byref(h)))
        self.handle = h
        self.width = self.height = 0

    def close(self):
        if getattr(self, "handle", None):
            self._lib.vvr_renderer_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_labels(self, labels: np.ndarray, soft: np.ndarray | None = None, soft_dims=None, src_axis=(0, 1, 2),
                   flip=(0, 0, 0)):
        labels = np.ascontiguousarray(labels, dtype=np.uint8)
        sg = None
        if soft is not None:
            soft = np.ascontiguousarray(soft, dtype=np.uint8)
            sg = SoftGrid()
            sg.dims[:] = soft_dims if soft_dims is not None else self.dims
            sg.src_axis[:] = src_axis
            sg.flip[:] = flip
            sg.data = soft.ctypes.data_as(POINTER(c_uint8))
        check(self._lib.vvr_renderer_set_labels(self.handle, labels.ctypes.data_as(POINTER(c_uint8)), labels.size,
                                                byref(sg) if sg is not None else None))

    def set_transfer_function(self, tf: TransferFunction):
        check(self._lib.vvr_renderer_set_transfer_function(self.handle, byref(tf.c)))

    def resize(self, width: int, height: int):
        check(self._lib.vvr_renderer_resize(self.handle, width, height))
        self.width, self.height = width, height

    def record(self, frame: int, params: FrameParams):
        check(self._lib.vvr_renderer_record(self.handle, frame, byref(params)))

    def synchronize(self):
        check(self._lib.vvr_renderer_synchronize(self.handle))

    def accumulated_samples(self) -> int:
        return int(self._lib.vvr_renderer_accumulated_samples(self.handle))

    def default_camera(self) -> Camera:
        cam = Camera()
        check(self._lib.vvr_renderer_default_camera(self.handle, byref(cam)))
        return cam

    def geometry(self) -> dict:
        lo, hi, c = (c_float * 3)(), (c_float * 3)(), (c_float * 3)()
        ms, dg = c_float(), c_float()
        check(self._lib.vvr_renderer_geometry(self.handle, lo, hi, byref(ms), c, byref(dg)))
        return {"bounds": (tuple(lo), tuple(hi)), "min_spacing": ms.value, "center": tuple(c), "diagonal": dg.value}

    def read_hdr(self) -> np.ndarray:
        out = np.zeros((self.height, self.width, 4), np.float32)
        check(self._lib.vvr_renderer_read_hdr(self.handle, _fptr(out)))
        return out

    def set_environment_map(self, rgb):
        """Renderer::set_environment_map: an (h, w, 3) linear RGB light map, rows
        from the zenith, or None to remove it."""
        if rgb is None:
            check(self._lib.vvr_renderer_set_environment_map(self.handle, None))
            self._env = None
            return
        a = np.ascontiguousarray(rgb, np.float32)
        m = EnvMap(a.shape[1], a.shape[0], a.ctypes.data_as(POINTER(c_float)))
        check(self._lib.vvr_renderer_set_environment_map(self.handle, byref(m)))
        self._env = a.shape[:2]

    def read_environment(self):
        """(peak_u, sampling table (h, w + 1, 2)) of the set light map."""
        h, w = self._env
        table = np.zeros((h, w + 1, 2), np.float32)
        peak = c_float()
        check(self._lib.vvr_renderer_read_environment(self.handle, byref(peak), table.ctypes.data_as(POINTER(c_float)),
                                                       table.size))
        return peak.value, table

    def set_denoiser(self, net):
        """Frames with VVR_DENOISE_LEARNED run the learned denoiser `net` (a
        volview_amd.unet.Unet, kept alive by the caller; None detaches)."""
        check(self._lib.vvr_renderer_set_denoiser(self.handle, net.handle if net is not None else None))
        self._denoiser = net

    def read_rgba(self) -> np.ndarray:
        """The tone-mapped RGBA8 display image of the last record()."""
        out = np.zeros((self.height, self.width, 4), np.uint8)
        check(self._lib.vvr_renderer_read_rgba(self.handle, out.ctypes.data_as(POINTER(c_uint8))))
        return out

    def read_exposure(self) -> float:
        v = c_float()
        check(self._lib.vvr_renderer_read_exposure(self.handle, byref(v)))
        return v.value

    def read_denoised(self) -> np.ndarray:
        out = np.zeros((self.height, self.width, 4), np.float32)
        check(self._lib.vvr_renderer_read_denoised(self.handle, _fptr(out)))
        return out

    def read_accum(self):
        shape = (self.height, self.width, 4)
        a, b, c = np.zeros(shape, np.float32), np.zeros(shape, np.float32), np.zeros(shape, np.float32)
        check(self._lib.vvr_renderer_read_accum(self.handle, _fptr(a), _fptr(b), _fptr(c)))
        return a, b, c

    def read_volume(self, which: int) -> np.ndarray:
        nx, ny, nz = self.dims
        if which == _abi.VOL_MAJORANT:
            out = np.zeros(((nz + 7) // 8, (ny + 7) // 8, (nx + 7) // 8), np.float32)
        elif which == _abi.VOL_DENSITY:
            out = np.zeros((nz, ny, nx), np.uint16)
        elif which == _abi.VOL_LABELS:
            out = np.zeros((nz, ny, nx), np.uint8)
        elif which == _abi.VOL_NORMALS:
            out = np.zeros((nz, ny, nx, 4), np.int8)
        elif which == _abi.VOL_MASK:
            out = np.zeros((nz, ny, nx, 2), np.uint16)
        elif which == _abi.VOL_SHADOW:  # RGBA16F bits
            out = np.zeros(((nz + 1) // 2, (ny + 1) // 2, (nx + 1) // 2, 4), np.uint16)
        elif which == _abi.VOL_AO:  # R16F bits
            out = np.zeros(((nz + 3) // 4, (ny + 3) // 4, (nx + 3) // 4), np.uint16)
        else:
            raise ValueError(which)
        check(self._lib.vvr_renderer_read_volume(self.handle, which, out.ctypes.data_as(c_void_p), out.nbytes))
        return out

    def read_majorant(self) -> np.ndarray:
        return self.read_volume(_abi.VOL_MAJORANT)

    def read_uniforms(self) -> Uniforms:
        u = Uniforms()
        check(self._lib.vvr_renderer_read_uniforms(self.handle, byref(u), ctypes.sizeof(u)))
        return u

    TAIL_NEVER = 0xFFFFFFFF

    @staticmethod
    def _shape(sh) -> dict:
        return {"slots": sh.slots, "flight_steps": sh.flight_steps, "shadow_steps": sh.shadow_steps,
                "tail_paths": sh.tail_paths}

    def pt_shape(self) -> dict:
        """The wavefront path tracer's shape: slots, flight_steps, shadow_steps,
        tail_paths (TAIL_NEVER: no tail launch)."""
        sh = PtShape()
        check(self._lib.vvr_renderer_get_pt_shape(self.handle, byref(sh)))
        return self._shape(sh)

    def set_pt_shape(self, slots: int = 0, flight_steps: int = 0, shadow_steps: int = 0, tail_paths: int = 0):
        """Sets the shape's non-zero fields (frames are the same with any shape)."""
        check(self._lib.vvr_renderer_set_pt_shape(self.handle,
                                                  byref(PtShape(slots, flight_steps, shadow_steps, tail_paths))))

    def autotune(self, params: FrameParams, trials: int = 1) -> dict:
        """vvr_renderer_autotune: times `params`' frames over candidate shapes and
        keeps the fastest; returns it. The trial frames accumulate."""
        sh = PtShape()
        check(self._lib.vvr_renderer_autotune(self.handle, byref(params), trials, byref(sh)))
        return self._shape(sh)

    def set_timing(self, enabled: bool):
        check(self._lib.vvr_renderer_set_timing(self.handle, 1 if enabled else 0))

    def kernel_time(self, which: int):
        ms, n = c_double(), c_uint64()
        check(self._lib.vvr_renderer_kernel_time(self.handle, which, byref(ms), byref(n)))
        return ms.value, n.value


def render_hdr(renderer: Renderer, params: FrameParams, width: int, height: int, frames: int = 1) -> np.ndarray:
    """render_to_rgba (offscreen.rs:11-78) up to the linear HDR image: resize, then
    `frames` frames, each recorded and waited on (the plain loop)."""
    renderer.resize(width, height)
    for frame in range(1, frames + 1):
        renderer.record(frame, params)
        renderer.synchronize()
    return renderer.read_hdr()


def render_to_rgba(renderer: Renderer, params: FrameParams, width: int, height: int, frames: int = 1) -> np.ndarray:
    """render_to_rgba (offscreen.rs:11-78): resize, `frames` frames (path tracing
    accumulates them), then the tone-mapped RGBA8 image, (height, width, 4)."""
    out = np.zeros((height, width, 4), np.uint8)
    check(renderer._lib.vvr_render_to_rgba(renderer.handle, byref(params), width, height, frames,
                                           out.ctypes.data_as(POINTER(c_uint8))))
    renderer.width, renderer.height = width, height
    return out
