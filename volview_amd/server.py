"""The serving loop over libvvr (vvr_server_*): requests of many resident
volumes rendered side by side on one device.

    srv = Server(ctx, pt_increment=16)
    rid = srv.submit(renderer, params, 1024, 1024, tag=7)
    for c in srv.poll(wait_ms=100): ...            # frame completions
    img = srv.read_hdr(rid)                        # float32 [1024, 1024, 4]
    srv.release(rid)

A request's frames equal those of its renderer recording the same parameters
with samples_per_frame = pt_increment, frame after frame (ray-marched frames
equal one record()). The renderers stay owned by the caller; while a request
holds one (from submission to its final frame), record, resize and the set_*
calls on it are refused with VVR_ESTATE.
"""
from __future__ import annotations

from ctypes import POINTER, byref, c_float, c_uint32, c_uint64, c_void_p, pointer

import numpy as np

from ._abi import Completion, FrameParams, Request, ServerConfig, ServerStats, check, load
from .renderer import Context, Renderer


class Server:
    def __init__(self, ctx: Context, hbm_budget: int = 0, slots: int = 0, pt_increment: int = 0,
                 max_requests: int = 0, flight_steps: int = 0, shadow_steps: int = 0, iterations: int = 0):
        self._lib = load()
        self.ctx = ctx
        cfg = ServerConfig(hbm_budget, slots, pt_increment, max_requests, flight_steps, shadow_steps, iterations)
        self._held = {}  # request id -> its renderer, kept alive while the server may read it
        self._dims = {}  # request id -> (width, height) of its frames
        h = c_void_p()
        check(self._lib.vvr_server_create(ctx.handle, byref(cfg), byref(h)))
        self.handle = h

    @staticmethod
    def autotune(ctx: Context, requests, **config) -> dict:
        """vvr_server_autotune: the engine shape (slots, iterations per pass) under
        which `requests` ((renderer, params, width, height) tuples) finish fastest,
        as keyword arguments for Server(ctx, ...) together with `config`."""
        lib = load()
        keep = [r[1] for r in requests]
        arr = (Request * len(requests))(*[Request(r.handle, pointer(p), w, h, 0) for r, p, w, h in requests])
        cfg = ServerConfig(config.get("hbm_budget", 0), config.get("slots", 0), config.get("pt_increment", 0),
                           config.get("max_requests", 0), config.get("flight_steps", 0),
                           config.get("shadow_steps", 0), config.get("iterations", 0))
        check(lib.vvr_server_autotune(ctx.handle, arr, len(requests), byref(cfg)))
        del keep
        return dict(config, slots=int(cfg.slots), iterations=int(cfg.iterations))

    def submit(self, renderer: Renderer, params: FrameParams, width: int, height: int, tag: int = 0) -> int:
        rq = Request(renderer.handle, pointer(params), width, height, tag)  # copied by the call
        rid = c_uint64()
        check(self._lib.vvr_server_submit(self.handle, byref(rq), byref(rid)))
        self._held[int(rid.value)] = renderer
        self._dims[int(rid.value)] = (int(width), int(height))
        return int(rid.value)

    def poll(self, cap: int = 256, wait_ms: int = 0) -> list[Completion]:
        out = (Completion * cap)()
        n = c_uint32()
        check(self._lib.vvr_server_poll(self.handle, out, cap, byref(n), wait_ms))
        return [out[i] for i in range(n.value)]

    def read_hdr(self, rid: int, width: int | None = None, height: int | None = None) -> np.ndarray:
        """The request's newest delivered frame, sized from its submission;
        width and height, when given, must match it."""
        w, h = self._dims.get(rid, (width or 0, height or 0))
        if (width is not None and width != w) or (height is not None and height != h):
            raise ValueError(f"request {rid} renders {w} x {h} frames, not {width} x {height}")
        img = np.zeros((h, w, 4), np.float32)
        check(self._lib.vvr_server_read_hdr(self.handle, rid, img.ctypes.data_as(POINTER(c_float)), img.size))
        return img

    def cancel(self, rid: int):
        check(self._lib.vvr_server_cancel(self.handle, rid))

    def release(self, rid: int):
        check(self._lib.vvr_server_release(self.handle, rid))
        self._held.pop(rid, None)
        self._dims.pop(rid, None)

    def stats(self) -> dict:
        s = ServerStats()
        check(self._lib.vvr_server_get_stats(self.handle, byref(s)))
        return {k: int(getattr(s, k)) for k, _ in ServerStats._fields_}

    def close(self):
        if getattr(self, "handle", None):
            self._lib.vvr_server_destroy(self.handle)
            self.handle = None
        self._held = {}
        self._dims = {}

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
