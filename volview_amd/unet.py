"""The learned denoiser over libvvr (vvr_unet_*): a convolutional U-Net over a
path-traced frame's accumulation and guide images, trained on the GPU by the
library's own bf16 MFMA kernels.

    net = Unet(ctx, channels=(32, 48, 64, 96, 128))
    loss = net.train_step(accum, albedo, normal, samples, target, lr=1e-3)  # (b, h, w, 4) float32 each
    rgba = net.denoise(accum, albedo, normal, samples)                     # one (h, w, 4) frame
    renderer.set_denoiser(net)   # frames with VVR_DENOISE_LEARNED run it
"""
from __future__ import annotations

from ctypes import POINTER, byref, c_float, c_void_p
from pathlib import Path

import numpy as np

from ._abi import UnetConfig, check, load
from .renderer import Context


def _f(a):
    return a.ctypes.data_as(POINTER(c_float))


class Unet:
    def __init__(self, ctx: Context, channels=(32, 48, 64, 96, 128), seed: int = 1):
        self._lib = load()
        self.ctx = ctx
        self.channels = tuple(int(c) for c in channels)
        cfg = UnetConfig(len(self.channels), tuple(self.channels) + (0,) * (5 - len(self.channels)), seed)
        h = c_void_p()
        check(self._lib.vvr_unet_create(ctx.handle, byref(cfg), byref(h)))
        self.handle = h
        self.n_params = int(self._lib.vvr_unet_num_params(h))

    def params(self) -> np.ndarray:
        out = np.zeros(self.n_params, np.float32)
        check(self._lib.vvr_unet_get_params(self.handle, _f(out), out.size))
        return out

    def set_params(self, p: np.ndarray):
        a = np.ascontiguousarray(p, np.float32)
        check(self._lib.vvr_unet_set_params(self.handle, _f(a), a.size))

    def grads(self) -> np.ndarray:
        out = np.zeros(self.n_params, np.float32)
        check(self._lib.vvr_unet_get_grads(self.handle, _f(out), out.size))
        return out

    def train_step(self, accum, albedo, normal, samples: float, target, lr: float) -> float:
        a, b, c, t = (np.ascontiguousarray(x, np.float32) for x in (accum, albedo, normal, target))
        n, h, w = a.shape[:3]
        loss = c_float()
        check(self._lib.vvr_unet_train_step(self.handle, n, w, h, _f(a), _f(b), _f(c), samples, _f(t), lr,
                                            byref(loss)))
        return loss.value

    def denoise(self, accum, albedo, normal, samples: float) -> np.ndarray:
        a, b, c = (np.ascontiguousarray(x, np.float32) for x in (accum, albedo, normal))
        h, w = a.shape[:2]
        out = np.zeros((h, w, 4), np.float32)
        check(self._lib.vvr_unet_denoise(self.handle, w, h, _f(a), _f(b), _f(c), samples, _f(out)))
        return out

    def denoise_noalloc(self, accum, albedo, normal, samples: float) -> np.ndarray:
        """denoise() through the allocation-free entry point (the library's static
        per-device scratch; frames up to about 96 Ki pixels)."""
        a, b, c = (np.ascontiguousarray(x, np.float32) for x in (accum, albedo, normal))
        h, w = a.shape[:2]
        out = np.zeros((h, w, 4), np.float32)
        check(self._lib.vvr_unet_denoise_noalloc(self.handle, w, h, _f(a), _f(b), _f(c), samples, _f(out)))
        return out

    @classmethod
    def pretrained(cls, ctx: Context) -> "Unet":
        """The network trained by volview_amd.train_denoiser (5 levels, 32-128
        channels; weights shipped in volview_amd/data/denoiser_unet.npy)."""
        net = cls(ctx, (32, 48, 64, 96, 128))
        net.set_params(np.load(Path(__file__).parent / "data" / "denoiser_unet.npy").astype(np.float32))
        return net

    def save(self, path):
        np.save(path, self.params())

    def load(self, path):
        self.set_params(np.load(path))

    def close(self):
        if getattr(self, "handle", None):
            self._lib.vvr_unet_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
