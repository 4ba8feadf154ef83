"""ctypes mirror of include/vvr.h and include/vvr_uniforms.h.

The structures here must match the C declarations field for field; tests check
their sizes against the library.
"""
from __future__ import annotations

import ctypes
from ctypes import POINTER, Structure, c_char_p, c_double, c_float, c_int, c_int16, c_size_t, c_uint8, c_uint16
from ctypes import c_int32, c_uint32, c_uint64, c_void_p
from pathlib import Path

F4x256 = (c_float * 4) * 256

VVR_OK, VVR_EINVAL, VVR_ELIMIT, VVR_ENOMEM, VVR_EDEVICE, VVR_ESTATE, VVR_ECANCELED = 0, -1, -2, -3, -4, -5, -6
LABELS_OFF, LABELS_TINT, LABELS_ISOLATE, LABELS_HYBRID = 0, 1, 2, 3
ENV_DOME, ENV_STUDIO, ENV_MAP = 0, 1, 2
DENOISE_OFF, DENOISE_ATROUS, DENOISE_LEARNED = 0, 1, 2
TONE_NONE, TONE_ACES, TONE_AGX = 0, 1, 2
VOL_DENSITY, VOL_LABELS, VOL_NORMALS, VOL_MASK, VOL_MAJORANT, VOL_SHADOW, VOL_AO = 0, 1, 2, 3, 4, 5, 6
KERNEL_MAJORANT, KERNEL_NORMALS, KERNEL_MASK, KERNEL_DVR, KERNEL_PT, KERNEL_LIGHTING, KERNEL_DENOISE, KERNEL_POST = (
    0, 1, 2, 3, 4, 5, 6, 7)
ABI_STRUCTS = ['Volume', 'ControlPoint', 'TF', 'SoftGrid', 'Camera', 'LabelParams', 'Lighting', 'Clipping',
               'PathTracing', 'FrameParams', 'Uniforms', 'ServerConfig', 'Request', 'Completion', 'ServerStats',
               'PtShape', 'EnvMap', 'UnetConfig']


class Volume(Structure):
    _fields_ = [("dims", c_uint32 * 3), ("spacing", c_double * 3), ("origin", c_double * 3),
                ("direction", c_double * 9), ("hu", POINTER(c_int16))]


class ControlPoint(Structure):
    _fields_ = [("hu", c_float), ("color", c_float * 3), ("extinction", c_float)]


class TF(Structure):
    _fields_ = [("n_points", c_uint32), ("points", POINTER(ControlPoint))]


class SoftGrid(Structure):
    _fields_ = [("dims", c_uint32 * 3), ("src_axis", c_uint32 * 3), ("flip", c_uint8 * 3), ("data", POINTER(c_uint8))]


class Camera(Structure):
    _fields_ = [("target", c_float * 3), ("distance", c_float), ("yaw", c_float), ("pitch", c_float),
                ("fov_y", c_float)]


class LabelParams(Structure):
    _fields_ = [("mode", c_uint32), ("tint", c_float), ("extinction", c_float), ("colors", F4x256),
                ("materials", F4x256), ("ghost", c_uint8 * 256), ("ghost_opacity", c_float),
                ("cut_margin_mm", c_float), ("smoothing", c_uint32), ("density", F4x256), ("surface", F4x256),
                ("optics", F4x256)]


class Lighting(Structure):
    _fields_ = [("azimuth", c_float), ("elevation", c_float), ("intensity", c_float), ("ambient", c_float),
                ("environment", c_uint32), ("light_size", c_float), ("env_rotation", c_float),
                ("anisotropy", c_float), ("chromatic", c_float), ("shadows", c_uint32),
                ("ambient_occlusion", c_uint32), ("exposure", c_float), ("auto_exposure", c_uint32),
                ("bloom", c_float), ("vignette", c_float), ("grain", c_float), ("tone_mapper", c_uint32),
                ("roughness", c_float), ("metalness", c_float), ("subsurface", c_float),
                ("scattering_distance", c_float), ("subsurface_strength", c_float), ("surface_detail", c_float)]


class ClipPlane(Structure):
    _fields_ = [("enabled", c_uint32), ("plane", c_uint32), ("cut_positive", c_uint32), ("position", c_float)]


class Clipping(Structure):
    _fields_ = [("planes", ClipPlane * 3), ("all", c_uint32)]


class PathTracing(Structure):
    _fields_ = [("bounces", c_uint32), ("samples_per_frame", c_uint32), ("denoiser", c_uint32),
                ("aperture", c_float), ("has_focus", c_uint32), ("focus", c_float * 3)]


class FrameParams(Structure):
    _fields_ = [("camera", Camera), ("step_mm", c_float), ("srgb_encode", c_uint32), ("labels", LabelParams),
                ("lighting", Lighting), ("preintegrate", c_uint32), ("skip_empty_space", c_uint32),
                ("clipping", Clipping), ("path_tracing_enabled", c_uint32), ("path_tracing", PathTracing),
                ("smooth_sampling", c_uint32)]


class Uniforms(Structure):
    """The 21,200-byte per-frame block (renderer.rs:42-74)."""
    _fields_ = [("world_to_index", c_float * 16), ("eye", c_float * 4), ("forward", c_float * 4),
                ("right", c_float * 4), ("up", c_float * 4), ("dims", c_float * 4), ("flags", c_uint32 * 4),
                ("labels", c_float * 4), ("light_dir", c_float * 4), ("ambient", c_float * 4),
                ("material", c_float * 4), ("lighting", c_uint32 * 4), ("light_cells", c_float * 4),
                ("clip", c_uint32 * 4), ("clip_planes", (c_float * 4) * 3), ("clip_normals", (c_float * 4) * 3),
                ("subsurface", c_float * 4), ("path", c_uint32 * 4), ("env", c_float * 4), ("lens", c_float * 4),
                ("finish", c_float * 4), ("tissue", c_float * 4), ("tf_bucket_max", (c_float * 4) * 16),
                ("label_colors", F4x256), ("label_materials", F4x256), ("label_density", F4x256),
                ("label_surface", F4x256), ("label_optics", F4x256)]


assert ctypes.sizeof(Uniforms) == 21200

# (name, restype, argtypes) of every exported entry point of include/vvr.h.
class ServerConfig(Structure):
    _fields_ = [("hbm_budget", c_uint64), ("slots", c_uint32), ("pt_increment", c_uint32), ("max_requests", c_uint32),
                ("flight_steps", c_uint32), ("shadow_steps", c_uint32), ("iterations", c_uint32)]


class PtShape(Structure):
    _fields_ = [("slots", c_uint32), ("flight_steps", c_uint32), ("shadow_steps", c_uint32), ("tail_paths", c_uint32)]


class Request(Structure):
    _fields_ = [("volume", c_void_p), ("params", POINTER(FrameParams)), ("width", c_uint32), ("height", c_uint32),
                ("tag", c_uint64)]


class Completion(Structure):
    _fields_ = [("id", c_uint64), ("tag", c_uint64), ("status", c_int32), ("spp", c_uint32),
                ("final_frame", c_uint32), ("pad", c_uint32), ("submit_ns", c_uint64), ("admit_ns", c_uint64),
                ("frame_ns", c_uint64)]


class ServerStats(Structure):
    _fields_ = [("submitted", c_uint64), ("completed", c_uint64), ("frames", c_uint64), ("pt_paths", c_uint64),
                ("iterations", c_uint64), ("hbm_in_use", c_uint64), ("hbm_peak", c_uint64),
                ("pt_paths_abandoned", c_uint64)]


class UnetConfig(Structure):
    _fields_ = [("levels", c_uint32), ("channels", c_uint32 * 5), ("seed", c_uint32)]


class EnvMap(Structure):
    _fields_ = [("width", c_uint32), ("height", c_uint32), ("rgb", POINTER(c_float))]


SIGNATURES = [
    ("vvr_label_params_default", None, [POINTER(LabelParams)]),
    ("vvr_lighting_default", None, [POINTER(Lighting)]),
    ("vvr_path_tracing_default", None, [POINTER(PathTracing)]),
    ("vvr_frame_params_default", None, [POINTER(FrameParams), POINTER(Camera), c_float]),
    ("vvr_camera_framing", None, [POINTER(c_float), c_float, POINTER(Camera)]),
    ("vvr_hu_to_unorm16", c_uint16, [c_int16]),
    ("vvr_tf_evaluate", None, [POINTER(TF), c_float, POINTER(c_float)]),
    ("vvr_tf_lut", c_int, [POINTER(TF), POINTER(c_float)]),
    ("vvr_tf_lut_f16", c_int, [POINTER(TF), POINTER(c_uint16)]),
    ("vvr_tf_bucket_max", c_int, [POINTER(TF), POINTER(c_float)]),
    ("vvr_tf_cumulative_integrals", c_int, [POINTER(TF), POINTER(c_float)]),
    ("vvr_tf_preset", c_int, [c_uint32, POINTER(ControlPoint), c_uint32]),
    ("vvr_ctx_create", c_int, [c_int, POINTER(c_void_p)]),
    ("vvr_ctx_destroy", None, [c_void_p]),
    ("vvr_renderer_new", c_int, [c_void_p, POINTER(Volume), POINTER(TF), POINTER(c_void_p)]),
    ("vvr_renderer_set_labels", c_int, [c_void_p, POINTER(c_uint8), c_size_t, POINTER(SoftGrid)]),
    ("vvr_renderer_set_transfer_function", c_int, [c_void_p, POINTER(TF)]),
    ("vvr_renderer_resize", c_int, [c_void_p, c_uint32, c_uint32]),
    ("vvr_renderer_record", c_int, [c_void_p, c_uint64, POINTER(FrameParams)]),
    ("vvr_renderer_synchronize", c_int, [c_void_p]),
    ("vvr_renderer_accumulated_samples", c_uint32, [c_void_p]),
    ("vvr_renderer_default_camera", c_int, [c_void_p, POINTER(Camera)]),
    ("vvr_renderer_geometry", c_int, [c_void_p, POINTER(c_float), POINTER(c_float), POINTER(c_float),
                                      POINTER(c_float), POINTER(c_float)]),
    ("vvr_renderer_destroy", None, [c_void_p]),
    ("vvr_renderer_read_hdr", c_int, [c_void_p, POINTER(c_float)]),
    ("vvr_renderer_read_accum", c_int, [c_void_p, POINTER(c_float), POINTER(c_float), POINTER(c_float)]),
    ("vvr_renderer_read_denoised", c_int, [c_void_p, POINTER(c_float)]),
    ("vvr_renderer_read_rgba", c_int, [c_void_p, POINTER(c_uint8)]),
    ("vvr_renderer_set_environment_map", c_int, [c_void_p, POINTER(EnvMap)]),
    ("vvr_unet_create", c_int, [c_void_p, POINTER(UnetConfig), POINTER(c_void_p)]),
    ("vvr_unet_destroy", None, [c_void_p]),
    ("vvr_unet_num_params", c_size_t, [c_void_p]),
    ("vvr_unet_get_params", c_int, [c_void_p, POINTER(c_float), c_size_t]),
    ("vvr_unet_set_params", c_int, [c_void_p, POINTER(c_float), c_size_t]),
    ("vvr_unet_get_grads", c_int, [c_void_p, POINTER(c_float), c_size_t]),
    ("vvr_unet_train_step", c_int, [c_void_p, c_uint32, c_uint32, c_uint32, POINTER(c_float), POINTER(c_float),
                                    POINTER(c_float), c_float, POINTER(c_float), c_float, POINTER(c_float)]),
    ("vvr_unet_denoise", c_int, [c_void_p, c_uint32, c_uint32, POINTER(c_float), POINTER(c_float), POINTER(c_float),
                                 c_float, POINTER(c_float)]),
    ("vvr_unet_denoise_noalloc", c_int, [c_void_p, c_uint32, c_uint32, POINTER(c_float), POINTER(c_float),
                                         POINTER(c_float), c_float, POINTER(c_float)]),
    ("vvr_renderer_set_denoiser", c_int, [c_void_p, c_void_p]),
    ("vvr_renderer_read_environment", c_int, [c_void_p, POINTER(c_float), POINTER(c_float), c_size_t]),
    ("vvr_renderer_read_exposure", c_int, [c_void_p, POINTER(c_float)]),
    ("vvr_render_to_rgba", c_int, [c_void_p, POINTER(FrameParams), c_uint32, c_uint32, c_uint32, POINTER(c_uint8)]),
    ("vvr_renderer_read_volume", c_int, [c_void_p, c_int, c_void_p, c_size_t]),
    ("vvr_renderer_read_majorant", c_int, [c_void_p, POINTER(c_float), c_size_t]),
    ("vvr_renderer_read_uniforms", c_int, [c_void_p, c_void_p, c_size_t]),
    ("vvr_renderer_set_timing", c_int, [c_void_p, c_int]),
    ("vvr_renderer_kernel_time", c_int, [c_void_p, c_int, POINTER(c_double), POINTER(c_uint64)]),
    ("vvr_renderer_get_pt_shape", c_int, [c_void_p, POINTER(PtShape)]),
    ("vvr_renderer_set_pt_shape", c_int, [c_void_p, POINTER(PtShape)]),
    ("vvr_renderer_autotune", c_int, [c_void_p, POINTER(FrameParams), c_uint32, POINTER(PtShape)]),
    ("vvr_phantom_ct", c_int, [POINTER(c_uint32), c_uint64, POINTER(c_int16)]),
    ("vvr_phantom_organs", c_int, [POINTER(c_uint32), c_uint64, c_uint32, POINTER(c_uint8), POINTER(c_uint8),
                                   POINTER(c_uint16)]),
    ("vvr_frame_uniforms", c_int, [POINTER(Volume), POINTER(TF), POINTER(FrameParams), c_int, c_uint32, c_uint32,
                                   c_uint32, c_uint64, c_void_p, c_size_t]),
    ("vvr_abi_sizeof", c_size_t, [c_int]),
    ("vvr_server_create", c_int, [c_void_p, POINTER(ServerConfig), POINTER(c_void_p)]),
    ("vvr_server_submit", c_int, [c_void_p, POINTER(Request), POINTER(c_uint64)]),
    ("vvr_server_poll", c_int, [c_void_p, POINTER(Completion), c_uint32, POINTER(c_uint32), c_uint32]),
    ("vvr_server_read_hdr", c_int, [c_void_p, c_uint64, POINTER(c_float), c_size_t]),
    ("vvr_server_cancel", c_int, [c_void_p, c_uint64]),
    ("vvr_server_release", c_int, [c_void_p, c_uint64]),
    ("vvr_server_get_stats", c_int, [c_void_p, POINTER(ServerStats)]),
    ("vvr_server_autotune", c_int, [c_void_p, POINTER(Request), c_uint32, POINTER(ServerConfig)]),
    ("vvr_server_destroy", None, [c_void_p]),
    ("vvr_last_error", c_char_p, []),
    ("vvr_version", c_char_p, []),
]

LIB_PATH = Path(__file__).resolve().parent / "libvvr.so"
_lib = None


def load() -> ctypes.CDLL:
    """The built libvvr.so. Raises if it has not been built: there is no fallback."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m volview_amd.build`")
        lib = ctypes.CDLL(str(LIB_PATH))
        for name, res, args in SIGNATURES:
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


class VvrError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"vvr error {code}: {message}")
        self.code = code


def check(code: int) -> None:
    if code != VVR_OK:
        raise VvrError(code, load().vvr_last_error().decode())
