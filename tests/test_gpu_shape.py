"""The wavefront path tracer's shape (vvr_pt_shape) and its autotuner, on the GPU.

- Frames are bit-identical whatever the shape: tracking loops suspended after
  1-7 iterations, or never, a few hundred or a million slots, and the frame's
  tail run in one launch from the first iteration on (every path of a small
  frame), late, or never, in the renderer and in the server.
- The autotuner picks one of its candidates, leaves it set, and the renderer's
  frames afterwards equal a default renderer's.
- Shapes out of range are rejected.
"""
import os

import numpy as np
import pytest

from volview_amd import _abi, phantom, workloads
from volview_amd.renderer import Renderer, TransferFunction
from volview_amd.server import Server

pytestmark = pytest.mark.gpu


def scene(ctx):
    dims = (40, 36, 32)
    hu = phantom.ct(dims, 11)
    labels, soft, _, n = phantom.organs(dims, 11, 8)
    r = Renderer(ctx, hu, TransferFunction.preset(0))
    r.set_labels(labels, soft)
    return r, n


def frame(r, p, w, h):
    r.resize(w, h)  # restarts accumulation
    r.record(1, p)
    acc, alb, nrm = r.read_accum()
    return r.read_hdr(), acc, alb, nrm


def test_frames_equal_for_every_shape(ctx):
    r, n = scene(ctx)
    p = workloads.pt_params(r, n, 8, bounces=8, camera_index=2)
    w, h = 48, 40
    if not any(k in os.environ for k in ("VVR_PT_SLOTS", "VVR_PT_FLIGHT_STEPS", "VVR_PT_SHADOW_STEPS")):
        assert r.pt_shape() == {"slots": 1 << 23, "flight_steps": 24, "shadow_steps": 32, "tail_paths": 1 << 17}
    want = frame(r, p, w, h)
    never = Renderer.TAIL_NEVER
    for shape in ({"slots": 192, "flight_steps": 1, "shadow_steps": 1, "tail_paths": never},
                  {"flight_steps": 3, "shadow_steps": 2},
                  {"slots": 1 << 20, "flight_steps": 4096, "shadow_steps": 4096, "tail_paths": 64},
                  {"flight_steps": 5, "shadow_steps": 7, "tail_paths": 1 << 30}):
        r.set_pt_shape(**shape)
        got = frame(r, p, w, h)
        for a, b in zip(got, want):
            assert np.array_equal(a, b), shape
    assert r.pt_shape() == {"slots": 1 << 20, "flight_steps": 5, "shadow_steps": 7, "tail_paths": 1 << 30}
    assert np.abs(want[0][..., :3]).max() > 0


def test_autotune_keeps_a_candidate_and_frames(ctx):
    r, n = scene(ctx)
    p = workloads.pt_params(r, n, 4, bounces=8, camera_index=1)
    w, h = 64, 48
    r.resize(w, h)
    chosen = r.autotune(p, trials=1)
    assert (chosen["flight_steps"], chosen["shadow_steps"]) in {(8, 16), (12, 24), (16, 32), (24, 32), (24, 48), (32, 64)}
    # (the default shape is the incumbent, and is one of the candidates)
    assert chosen["slots"] in {1 << 20, 1 << 21, 1 << 22, 1 << 23}
    assert chosen["tail_paths"] in {1 << 15, 1 << 16, 1 << 17, 1 << 18, Renderer.TAIL_NEVER}
    assert r.pt_shape() == chosen
    assert r.accumulated_samples() > 4  # the trial frames accumulated
    r2, _ = scene(ctx)
    for a, b in zip(frame(r, p, w, h), frame(r2, p, w, h)):
        assert np.array_equal(a, b)


def test_shapes_out_of_range_rejected(ctx):
    r, n = scene(ctx)
    for bad in ({"slots": 8}, {"slots": (1 << 26) + 64}, {"flight_steps": 5000}, {"shadow_steps": 4097}):
        with pytest.raises(_abi.VvrError) as e:
            r.set_pt_shape(**bad)
        assert e.value.code == _abi.VVR_EINVAL
    with pytest.raises(_abi.VvrError):
        r.autotune(workloads.dvr_params(r, 0))  # not path-traced
    assert 1 <= r.pt_shape()["flight_steps"] <= 4096


def test_server_step_bounds_keep_frames(ctx):
    r, n = scene(ctx)
    p = workloads.pt_params(r, n, 8, bounces=8, camera_index=2)
    w, h = 40, 24
    srv = Server(ctx, pt_increment=8, flight_steps=2, shadow_steps=3)
    rid = srv.submit(r, p, w, h)
    final = None
    for _ in range(600):
        final = next((c for c in srv.poll(wait_ms=100) if c.id == rid and c.final_frame), None)
        if final is not None:
            break
    assert final is not None and final.status == 0 and final.spp == 8
    served = srv.read_hdr(rid, w, h)
    srv.release(rid)
    srv.close()
    assert np.array_equal(served, frame(r, p, w, h)[0])
