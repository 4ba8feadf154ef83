#!/usr/bin/env python3
"""Benchmark of the path tracer on one MI355X.

Workload PT-1: seeded 512 x 512 x 400 int16 CT phantom with 40 labelled organs
and soft margins, isolate mode with measured soft-tissue optics, Henyey-
Greenstein g = 0.8, GGX + clear coat, 8 bounces, 1024 x 1024 at 64 paths per
pixel. One step is one request of the plain loop (crates/render/src/offscreen.rs:32-36):
record one frame of 64 paths per pixel, then wait for it.

Prints one JSON line: PT Msamples/s (1 sample = 1 pixel x 1 path), per-request
latency, the roofline of the path-tracing pass (HIP events on the renderer's
stream around the wavefront kernels of each frame: regen, extend, shade, shadow,
scatter iterations and the resolve) and the CPU oracle's rate on a bounded
sample of the same frame.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

from volview_amd import _abi, workloads  # noqa: E402
from volview_amd.renderer import Context, Renderer, TransferFunction, make_volume  # noqa: E402

METRIC = "PT Msamples/s + p50/p99 frame latency at fixed QPS, served vs plain, 1 GPU"
HBM_PEAK_GBS = 8000.0


def pt_frame_bytes(dims, width, height) -> int:
    """Algorithmic HBM bytes of one path-tracing frame: density (2), normals (4),
    labels (1) and mask (4) bytes per voxel read once, 4 bytes per brick, and the
    accumulation and two guide images (16 bytes per pixel each) read and written."""
    nx, ny, nz = dims
    nvox = nx * ny * nz
    nbricks = -(-nx // 8) * -(-ny // 8) * -(-nz // 8)
    return (2 + 4 + 1 + 4) * nvox + 4 * nbricks + 3 * 2 * 16 * width * height


def cpu_model() -> str:
    """The host CPU's model name (/proc/cpuinfo)."""
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_threads() -> int:
    return int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))


def cpu_baseline(scene, params, seconds: float) -> dict:
    """The FP32 CPU oracle (a port of pt.comp, OpenMP over all host cores) on
    random pixels of the same frame, one path each, for about `seconds`."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle  # noqa: E402
    oracle.load()
    r = scene.renderer
    tf = workloads.TransferFunction.preset(0)
    p = workloads.pt_params(r, scene.n_labels, 1)
    u = oracle.pack_uniforms(make_volume(scene.hu), p, oracle.bucket_max(tf), True, 0, scene.width, scene.height, 1)
    sc = oracle.Scene(oracle.hu_to_unorm16(scene.hu), tf=tf, labels=scene.labels)
    sc.set("normals", r.read_volume(_abi.VOL_NORMALS))
    sc.set("mask", r.read_volume(_abi.VOL_MASK))
    sc.set("majorant", r.read_majorant())
    rng = np.random.default_rng(5)
    done, t_total, batch = 0, 0.0, 256
    while t_total < seconds:
        px = np.stack([rng.integers(0, scene.width, batch), rng.integers(0, scene.height, batch)], 1)
        t0 = time.perf_counter()
        sc.pt(u, px.astype(np.uint32), 32)
        dt = time.perf_counter() - t0
        t_total += dt
        done += batch
        if dt < seconds / 20:
            batch *= 2
    return {"value": done / t_total / 1e6, "unit": "Msamples/s", "cores": cpu_threads(), "kind": "port",
            "cpu": cpu_model(),
            "sample": f"{done} paths (1 per random pixel) of the PT-1 frame, {t_total:.1f} s, FP32 C++ oracle"}


def dvr_cpu_baseline(scene, seconds: float) -> dict:
    """The FP32 CPU oracle of dvr.comp (OpenMP over the host cores) on random
    pixels of the DVR-1 frame for about `seconds`, as ms per 512^2 frame."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle  # noqa: E402
    oracle.load()
    r = scene.renderer
    tf = TransferFunction.preset(0)
    u = oracle.pack_uniforms(make_volume(scene.hu), scene.params, oracle.bucket_max(tf), False, 0, scene.width,
                             scene.height, 1)
    sc = oracle.Scene(oracle.hu_to_unorm16(scene.hu), tf=tf)
    sc.set("normals", r.read_volume(_abi.VOL_NORMALS))
    sc.set("majorant", r.read_majorant())
    rng = np.random.default_rng(6)
    done, t_total, batch = 0, 0.0, 1024
    while t_total < seconds:
        px = np.stack([rng.integers(0, scene.width, batch), rng.integers(0, scene.height, batch)], 1)
        t0 = time.perf_counter()
        sc.dvr(u, px.astype(np.uint32), 32)
        dt = time.perf_counter() - t0
        t_total += dt
        done += batch
        if dt < seconds / 20:
            batch *= 2
    per_frame = t_total / done * scene.width * scene.height * 1e3
    return {"value": round(per_frame, 1), "unit": "ms per 512^2 frame", "cores": cpu_threads(), "kind": "port",
            "cpu": cpu_model(),
            "sample": f"{done} random pixels of the DVR-1 frame, {t_total:.1f} s, FP32 C++ oracle, extrapolated"}


def copy_probe(nbytes: int) -> None:
    """A fill (hipMemset) and a plain device-to-device copy (hipMemcpy) of
    `nbytes`: kernels of known HBM traffic whose FETCH_SIZE / WRITE_SIZE readings
    scale the path tracer's under rocprofv3 --pmc."""
    import ctypes
    hip = ctypes.CDLL(os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib", "libamdhip64.so"))
    src, dst = ctypes.c_void_p(), ctypes.c_void_p()
    n = ctypes.c_size_t(nbytes)
    steps = (("hipMalloc", lambda: hip.hipMalloc(ctypes.byref(src), n)),
             ("hipMalloc", lambda: hip.hipMalloc(ctypes.byref(dst), n)),
             ("hipMemset", lambda: hip.hipMemset(src, 1, n)),
             ("hipMemcpy", lambda: hip.hipMemcpy(dst, src, n, 3)),  # hipMemcpyDeviceToDevice
             ("hipDeviceSynchronize", lambda: hip.hipDeviceSynchronize()),
             ("hipFree", lambda: hip.hipFree(src)),
             ("hipFree", lambda: hip.hipFree(dst)))
    for what, call in steps:
        rc = call()
        if rc != 0:
            raise RuntimeError(f"copy probe: {what} failed ({rc})")


def measured_traffic(args, shape: dict) -> int | None:
    """HBM bytes of one path-tracing pass at the default PT-1 configuration, as
    measured by rocprofv3 FETCH_SIZE / WRITE_SIZE passes scaled by a copy of
    known size (profiles/pt1_wavefront_hbm_traffic.md); None for other sizes or
    another engine shape than the profiled one."""
    if (args.spp, args.size, tuple(args.dims)) != (64, 1024, (512, 512, 400)):
        return None
    prof = ROOT / "profiles" / "pt1_wavefront_hbm_traffic.md"
    if not prof.exists():
        return None
    traffic, profiled = None, None
    for line in prof.read_text().splitlines():
        if line.startswith("traffic_bytes_per_pass:"):
            traffic = int(line.split(":")[1])
        if line.startswith("shape:"):
            profiled = [int(x) for x in line.split(":")[1].split()]
    if profiled != [shape["slots"], shape["flight_steps"], shape["shadow_steps"], shape["tail_paths"]]:
        return None
    return traffic


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def percentiles(xs) -> dict:
    if not xs:
        return {"p50": None, "p99": None, "n": 0}
    a = np.asarray(xs) * 1e3
    return {"p50": round(float(np.percentile(a, 50)), 2), "p99": round(float(np.percentile(a, 99)), 2), "n": len(xs)}


def drain(srv, ids, limit_s: float) -> dict:
    """Polls until every request in `ids` delivered its final frame; returns the completions per id."""
    want, got, t_end = set(ids), {}, time.perf_counter() + limit_s
    while want and time.perf_counter() < t_end:
        for c in srv.poll(wait_ms=50):
            got.setdefault(c.id, []).append((c.spp, c.final_frame, c.submit_ns, c.frame_ns, c.status))
            if c.final_frame:
                want.discard(c.id)
    if want:
        raise RuntimeError(f"{len(want)} requests did not finish within {limit_s} s")
    return got


def run_batched(ctx) -> dict:
    """Batched-32 rendered one request at a time (record + synchronize each, as
    the reference's plain loop) and by the server in one submission."""
    from volview_amd import phantom
    from volview_amd.renderer import Renderer, TransferFunction
    from volview_amd.server import Server
    t = time.perf_counter()
    vols, reqs = workloads.batched32()
    rs = [Renderer(ctx, phantom.ct(dims, seed), TransferFunction.preset(0)) for dims, seed in vols]
    params = [workloads.request_params(rs[q.volume], q) for q in reqs]
    for q in reqs:
        rs[q.volume].resize(q.width, q.height)
    pt_paths = sum(q.width * q.height * q.spp for q in reqs if q.path_traced)
    log(f"batched-32 setup {time.perf_counter() - t:.1f} s: {sum(q.path_traced for q in reqs)} path-traced, "
        f"{pt_paths / 1e6:.0f} M paths")

    def plain() -> float:
        t0 = time.perf_counter()
        for k, (q, p) in enumerate(zip(reqs, params)):
            rs[q.volume].record(k + 1, p)
            rs[q.volume].synchronize()
        return time.perf_counter() - t0

    def plain_one_path_frames() -> float:
        """The reference's own loop (render_to_rgba, offscreen.rs:32-36) at its
        default frame of one path per pixel (samples_per_frame = 1,
        renderer.rs:363): a path-traced request of n paths is n frames, each
        submitted and waited for. Its image equals the one-frame request's."""
        t0 = time.perf_counter()
        for k, (q, p) in enumerate(zip(reqs, params)):
            if not q.path_traced:
                rs[q.volume].record(k + 1, p)
                rs[q.volume].synchronize()
                continue
            p1 = type(p).from_buffer_copy(p)
            p1.path_tracing.samples_per_frame = 1
            for f in range(q.spp):
                rs[q.volume].record(f + 1, p1)
                rs[q.volume].synchronize()
        return time.perf_counter() - t0

    plain()  # warm-up: majorants, engines
    t_plain = plain()
    log(f"batched-32 plain {t_plain:.2f} s")
    t_plain_ref = plain_one_path_frames()
    log(f"batched-32 plain, one-path frames {t_plain_ref:.2f} s")
    t = time.perf_counter()  # untimed: the server's shape for this batch (vvr_server_autotune)
    shape = Server.autotune(ctx, [(rs[q.volume], p, q.width, q.height) for q, p in zip(reqs, params)])
    log(f"batched-32 server autotune {time.perf_counter() - t:.1f} s: {shape}")
    srv = Server(ctx, **shape)

    def served() -> float:
        t0 = time.perf_counter()
        ids = [srv.submit(rs[q.volume], p, q.width, q.height) for q, p in zip(reqs, params)]
        drain(srv, ids, 300.0)
        dt = time.perf_counter() - t0
        for i in ids:
            srv.release(i)
        return dt

    served()
    t_served = served()
    srv.close()
    log(f"batched-32 served {t_served:.2f} s")
    for r in rs:
        r.close()
    side = lambda sec: {"seconds": round(sec, 3), "pt_msamples_s": round(pt_paths / sec / 1e6, 2),
                        "requests_per_s": round(len(reqs) / sec, 2)}
    return {"workload": "Batched-32: 32 seeded phantoms of edge U{128..512}, images U{256..1024}^2, "
                        "16 DVR (bone preset) + 16 PT (U{1..64} spp, 8 bounces, g=0.8)",
            "pt_paths": pt_paths, "plain": side(t_plain), "served": side(t_served),
            "plain_one_path_frames": side(t_plain_ref),
            "server_shape": shape, "server_shape_from": "autotune",
            "speedup": round(t_plain / t_served, 2),
            "speedup_vs_one_path_frames": round(t_plain_ref / t_served, 2)}


def run_serving(ctx, seconds: float, grace: float = 5.0, increment: int = 16, slots: int = 2 << 20,
                iterations: int = 0, plain_side: bool = True) -> dict:
    """The serving trace in real time, twice: by a plain loop (one frame at a
    time, record + synchronize: ray-marched requests first, then one increment
    of the oldest progressive request) and by the server (all at once). A new
    path-traced request on a volume replaces the one in progress there (the
    viewer moved). Latency: arrival to frame complete (plain: on the device;
    served: read back to the host). A frame never delivered counts with the
    time it was waited for: until its request was replaced, or the end.
    Useful samples: per progressive request, the paths per pixel of the newest
    frame it delivered before it was replaced (or the end), x 1024^2."""
    from volview_amd import phantom
    from volview_amd.renderer import Renderer, TransferFunction
    from volview_amd.server import Server
    t = time.perf_counter()
    vols, reqs = workloads.serving_trace(seconds)
    pt_r, dvr_r, n_lab = [], [], []
    for dims, seed in vols:
        hu = phantom.ct(dims, seed)
        labels, soft, _, n = phantom.organs(dims, seed, 40)
        pair = []
        for _ in range(2):
            r = Renderer(ctx, hu, TransferFunction.preset(0))
            r.set_labels(labels, soft)
            pair.append(r)
        pt_r.append(pair[0])
        dvr_r.append(pair[1])
        n_lab.append(n)
    params, plain_params = [], []  # served: the whole path budget; plain: one increment per record
    for q in reqs:
        p = workloads.request_params(pt_r[q.volume], q, n_lab[q.volume])
        params.append(p)
        p = _abi.FrameParams.from_buffer_copy(p)
        if q.path_traced:
            p.path_tracing.samples_per_frame = increment
        plain_params.append(p)
    for v in range(len(vols)):
        pt_r[v].resize(1024, 1024)
        dvr_r[v].resize(1280, 720)
    # warm-up: masks, majorants, engines
    for v in range(len(vols)):
        for r, q in ((pt_r[v], next(q for q in reqs if q.path_traced)), (dvr_r[v], next(q for q in reqs
                                                                                           if not q.path_traced))):
            r.record(1, workloads.request_params(r, q, n_lab[v]))
            r.synchronize()
    log(f"serving setup {time.perf_counter() - t:.1f} s: {len(reqs)} requests over {seconds} s, "
        f"{sum(q.path_traced for q in reqs)} path-traced")

    # ---- plain ----
    plain, window = None, seconds + grace
    if plain_side:
        from collections import deque
        dvr_lat, pt_first, pt_paths, frame, useful = [], [], 0, 0, 0
        dvr_q, pt_fifo, active = deque(), deque(), {}
        done_at = {}  # request -> [(time its frame completed, paths per pixel)]
        i, t0 = 0, time.perf_counter()
        while True:
            now = time.perf_counter() - t0
            if now >= seconds + grace:
                break
            while i < len(reqs) and reqs[i].arrival <= now:
                q = reqs[i]
                if q.path_traced:
                    if q.volume not in active:
                        pt_fifo.append(q.volume)
                    else:
                        old, old_done = active[q.volume]
                        # the newest frame completed before the replacement arrived
                        useful += max([d for t, d in done_at.get(old, []) if t <= q.arrival], default=0)
                        if old_done == 0:  # replaced before its first frame
                            pt_first.append(q.arrival - reqs[old].arrival)
                    active[q.volume] = [i, 0]
                else:
                    dvr_q.append(i)
                i += 1
            frame += 1
            if dvr_q:
                k = dvr_q.popleft()
                r = dvr_r[reqs[k].volume]
                r.record(frame, plain_params[k])
                r.synchronize()
                dvr_lat.append(time.perf_counter() - t0 - reqs[k].arrival)
            elif pt_fifo:
                v = pt_fifo.popleft()
                k, done = active[v]
                pt_r[v].record(frame, plain_params[k])
                pt_r[v].synchronize()
                done += increment
                done_at.setdefault(k, []).append((time.perf_counter() - t0, done))
                pt_paths += 1024 * 1024 * increment
                if done == increment:
                    pt_first.append(time.perf_counter() - t0 - reqs[k].arrival)
                if done < reqs[k].spp:
                    active[v][1] = done
                    pt_fifo.append(v)
                else:
                    useful += done
                    del active[v]
            elif i < len(reqs):
                time.sleep(max(0.0, reqs[i].arrival - now))
            else:
                break
        span = time.perf_counter() - t0
        dvr_lat += [span - reqs[k].arrival for k in dvr_q]  # still waiting
        pt_first += [span - reqs[k].arrival for k, done in active.values() if done == 0]
        useful += sum(done for _, done in active.values())
        plain = {"dvr_latency_ms": percentiles(dvr_lat), "pt_first_frame_ms": percentiles(pt_first),
                 "pt_msamples_s": round(pt_paths / window / 1e6, 2), "pt_paths": pt_paths,
                 "pt_useful_msamples_s": round(useful * 1024 * 1024 / window / 1e6, 2), "busy_seconds": round(span, 2)}
        log(f"serving plain: {plain}")

    # ---- served ----
    srv = Server(ctx, pt_increment=increment, slots=slots, iterations=iterations)
    ids, kind, current, finals = {}, {}, {}, set()
    first, dvr_done, pt_paths = {}, {}, 0
    i, t0, t0_ns = 0, time.perf_counter(), time.monotonic_ns()  # the library stamps frames on the same clock
    arrival, replaced, frames, history, replaced_at, admitted = {}, {}, {}, {}, {}, {}

    def take(cs):
        nonlocal pt_paths
        for c in cs:
            if c.final_frame:
                finals.add(c.id)
            if c.status != 0:
                continue
            at = (c.frame_ns - t0_ns) / 1e9 - arrival[c.id]  # from the scheduled arrival, as for plain
            if kind[c.id]:
                admitted.setdefault(c.id, (c.admit_ns - c.submit_ns) / 1e9)
                if c.spp > frames.get(c.id, (0, 0))[1]:
                    pt_paths += 1024 * 1024 * (c.spp - frames.get(c.id, (0, 0))[1])
                    frames[c.id] = ((c.frame_ns - t0_ns) / 1e9, c.spp)
                    history.setdefault(c.id, []).append(frames[c.id])
                first.setdefault(c.id, at)
            else:
                dvr_done[c.id] = at

    while True:
        now = time.perf_counter() - t0
        if now >= seconds + grace or (i == len(reqs) and len(finals) == len(kind)):
            break
        while i < len(reqs) and reqs[i].arrival <= now:
            q = reqs[i]
            rid = srv.submit(pt_r[q.volume], params[i], q.width, q.height, tag=i)
            ids[i], kind[rid], arrival[rid] = rid, q.path_traced, q.arrival
            if q.path_traced:
                if q.volume in current:
                    prev = current[q.volume]
                    srv.cancel(prev)
                    replaced_at[prev] = q.arrival
                    if prev not in first:  # replaced before its first frame
                        replaced[prev] = q.arrival - arrival[prev]
                current[q.volume] = rid
            i += 1
        wait = reqs[i].arrival - (time.perf_counter() - t0) if i < len(reqs) else 0.02
        take(srv.poll(wait_ms=max(0, min(int(wait * 1e3), 20))))
    span = time.perf_counter() - t0
    for rid in kind:
        if rid not in finals:
            srv.cancel(rid)
    stats = srv.stats()
    srv_slots = slots
    srv.close()
    dvr_lat = [dvr_done.get(ids[k], span - q.arrival) for k, q in enumerate(reqs) if not q.path_traced and k in ids]
    pt_first = [first[r] if r in first else replaced.get(r, span - arrival[r]) for r in kind if kind[r]]
    # The newest frame each request delivered before it was replaced (or the end).
    useful = sum(max([spp for t, spp in history.get(r, []) if t <= replaced_at.get(r, window)], default=0)
                 for r in kind if kind[r])
    served = {"dvr_latency_ms": percentiles(dvr_lat), "pt_first_frame_ms": percentiles(pt_first),
              "pt_msamples_s": round(pt_paths / window / 1e6, 2), "pt_paths": pt_paths,
              "pt_useful_msamples_s": round(useful * 1024 * 1024 / window / 1e6, 2),
              "busy_seconds": round(span, 2), "slots": srv_slots, "hbm_peak_bytes": stats["hbm_peak"],
              "pt_paths_abandoned": stats["pt_paths_abandoned"],
              "pt_admission_ms": percentiles(list(admitted.values()))}
    log(f"serving served: {served}")
    for r in pt_r + dvr_r:
        r.close()
    return {"workload": f"Poisson {len(reqs) / seconds:.0f} req/s for {seconds:.0f} s (seed 7): 70% DVR 1280x720, "
                        f"30% progressive PT 1024^2 in {increment}-spp frames up to 256 spp, a newer PT request "
                        "on a volume replacing the older; 8 phantoms of edge U{128..512} with 40 organs",
            "window_seconds": window, "plain": plain, "served": served}


def served_quality(ctx, reference_spp: int = 4096, budget: int = 256, increment: int = 16) -> dict:
    """The serving trace's progressive requests against a converged frame: the
    first path-traced request of the trace served to its whole budget (256
    paths, in 16-path frames) and the same request rendered with 4096 paths by
    its renderer; relative RMSE of the served frame (linear RGB and luminance)."""
    from volview_amd import phantom
    from volview_amd.renderer import Renderer
    from volview_amd.server import Server
    vols, reqs = workloads.serving_trace(5.0)
    q = next(q for q in reqs if q.path_traced)
    dims, seed = vols[q.volume]
    hu = phantom.ct(dims, seed)
    labels, soft, _, n = phantom.organs(dims, seed, 40)
    r = Renderer(ctx, hu, TransferFunction.preset(0))
    r.set_labels(labels, soft)
    p = workloads.request_params(r, q, n)
    p.path_tracing.samples_per_frame = budget
    srv = Server(ctx, pt_increment=increment)
    rid = srv.submit(r, p, q.width, q.height)
    drain(srv, [rid], 300.0)
    served = srv.read_hdr(rid).astype(np.float64)
    srv.release(rid)
    srv.close()
    ref = _abi.FrameParams.from_buffer_copy(p)
    ref.path_tracing.samples_per_frame = 512  # frames of 512 paths accumulate to the reference
    r.resize(q.width, q.height)
    t = time.perf_counter()
    for f in range(reference_spp // 512):
        r.record(1 + f, ref)
    conv = r.read_hdr().astype(np.float64)
    t_ref = time.perf_counter() - t
    r.close()
    lum = np.array([0.2126, 0.7152, 0.0722])
    rel = lambda a, b: float(np.sqrt(np.mean((a - b) ** 2) / np.mean(b ** 2)))
    out = {"request": f"{q.width}x{q.height}, volume {dims[0]}^3 with {n} organs, {budget} paths served in "
                      f"{increment}-path frames", "reference_paths": reference_spp,
           "rel_rmse_rgb": round(rel(served[..., :3], conv[..., :3]), 4),
           "rel_rmse_luminance": round(rel(served[..., :3] @ lum, conv[..., :3] @ lum), 4),
           "reference_render_s": round(t_ref, 2)}
    log(f"served quality: {out}")
    return out


def dvr_frame_bytes(dims, width, height) -> int:
    """Algorithmic bytes of one DVR-1 frame with lighting off: density (2 B) and
    normals (4 B) per voxel, the majorant grid, the RGBA16F image."""
    n = dims[0] * dims[1] * dims[2]
    bricks = -(-dims[0] // 8) * -(-dims[1] // 8) * -(-dims[2] // 8)
    return 6 * n + 4 * bricks + 8 * width * height


def majorant_bytes(dims, mode: int) -> int:
    """Algorithmic bytes of the majorant build: per voxel the inputs its bound
    depends on, read once (the density, 2 B, without labels and in tint mode;
    the mask, 4 B, in the label modes: isolate and hybrid bounds are the label
    opacity wherever the mask is non-zero), and one f32 per 8^3 brick."""
    n = dims[0] * dims[1] * dims[2]
    bricks = -(-dims[0] // 8) * -(-dims[1] // 8) * -(-dims[2] // 8)
    per_voxel = {_abi.LABELS_OFF: 2, _abi.LABELS_TINT: 6}.get(mode, 4)
    return per_voxel * n + 4 * bricks


def roofline_gbs(nbytes: int, ms: float) -> dict:
    achieved = nbytes / (ms * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "algorithmic_bytes": nbytes}


def run_dvr(ctx, frames: int = 40, cpu_seconds: float = 0.0) -> dict:
    """DVR-1 (256^3 phantom, bone preset, 512^2, cubic sampling, lighting off):
    frames/s, the ray marcher's HIP-event time per frame and its roofline, wall
    latency of record + read-back; the majorant's time: its first build on a new
    renderer (the density's cell records, then the bricks; median of 5 renderers)
    and its rebuild after a transfer-function change, which reduces the kept cell
    records only (median of 9 changes)."""
    dims, size = (256, 256, 256), (512, 512)
    sc = workloads.dvr_scene(ctx, dims=dims, size=size)
    first_builds = []
    for _ in range(5):
        rf = Renderer(ctx, sc.hu, TransferFunction.preset(0))
        rf.resize(*size)
        rf.set_timing(True)
        rf.record(1, sc.params)
        rf.synchronize()
        ms, n = rf.kernel_time(_abi.KERNEL_MAJORANT)
        assert n == 1, n
        first_builds.append(ms)
        rf.close()
    maj_ms = float(np.median(first_builds))
    r = sc.renderer
    r.record(1, sc.params)  # builds normals and the majorant
    builds = []
    for _ in range(9):
        r.set_transfer_function(TransferFunction.preset(0))
        r.set_timing(True)
        r.record(1, sc.params)
        r.synchronize()
        ms, n = r.kernel_time(_abi.KERNEL_MAJORANT)
        assert n == 1, n
        builds.append(ms)
    maj_tf_ms = float(np.median(builds))
    r.set_timing(True)
    lat = []
    for f in range(frames):
        t = time.perf_counter()
        r.record(2 + f, sc.params)
        r.read_hdr()
        lat.append(time.perf_counter() - t)
    dvr_ms, n = r.kernel_time(_abi.KERNEL_DVR)
    r.set_timing(False)
    cpu = dvr_cpu_baseline(sc, cpu_seconds) if cpu_seconds > 0 else None
    r.close()
    per = dvr_ms / max(n, 1)
    # With the shadow and occlusion volumes (built once, then read by every frame).
    lit = workloads.dvr_scene(ctx, dims=dims, size=size, lighting=True)
    rl = lit.renderer
    rl.set_timing(True)
    rl.record(1, lit.params)
    rl.synchronize()
    light_ms, _ = rl.kernel_time(_abi.KERNEL_LIGHTING)
    rl.set_timing(True)
    for f in range(frames):
        rl.record(2 + f, lit.params)
    rl.synchronize()
    lit_ms, n_lit = rl.kernel_time(_abi.KERNEL_DVR)
    rl.set_timing(False)
    rl.close()
    log(f"dvr-1: {per:.3f} ms per frame ({lit_ms / max(n_lit, 1):.3f} lit), majorant {maj_ms:.4f} ms "
        f"(after a transfer-function change {maj_tf_ms:.4f} ms)")
    return {"workload": "DVR-1: 256^3 int16 phantom, bone preset, 512^2, step 0.5 mm, cubic sampling, lighting off",
            "ms_per_frame": round(per, 3), "frames_per_s": round(1e3 / per, 1),
            "latency_ms": {"p50": round(float(np.percentile(lat, 50)) * 1e3, 3),
                           "p99": round(float(np.percentile(lat, 99)) * 1e3, 3)},
            "roofline": roofline_gbs(dvr_frame_bytes(dims, *size), per),
            "lit": {"ms_per_frame": round(lit_ms / max(n_lit, 1), 3), "lighting_volumes_ms": round(light_ms, 3)},
            "majorant": {"ms": round(maj_ms, 4),
                         "roofline": roofline_gbs(majorant_bytes(dims, _abi.LABELS_OFF), maj_ms),
                         "transfer_function_change_ms": round(maj_tf_ms, 4)},
            "cpu_baseline": cpu}


def unet_flops(channels, width, height) -> float:
    """Algorithmic FLOPs of one U-Net pass over a width x height frame: 2 x 9 x
    c_in x c_out per pixel of each 3x3 layer at its level's resolution (the
    output layer with its 3 real channels)."""
    L = len(channels)
    per_level = [0.0] * L
    for l in range(L):
        per_level[l] += 9 * (16 if l == 0 else channels[l - 1]) * channels[l] + 9 * channels[l] * channels[l]
    for l in range(L - 1):
        per_level[l] += 9 * (channels[l + 1] + channels[l]) * channels[l] + 9 * channels[l] * channels[l]
    per_level[0] += 9 * channels[0] * 3
    return 2.0 * sum(m * width * height / 4 ** l for l, m in enumerate(per_level))


def run_denoisers(r, params, size: int, frames: int = 5) -> dict:
    """Per-frame cost of the a-trous and the learned denoiser on a 4-path frame
    of the PT-1 scene (HIP events around each denoiser's launches)."""
    from volview_amd.unet import Unet
    net = Unet.pretrained(r.ctx)
    r.set_denoiser(net)
    out = {"frame": f"{size}^2, 4 paths per frame"}
    p = _abi.FrameParams.from_buffer_copy(params)
    p.path_tracing.samples_per_frame = 4
    for name, mode in (("atrous", _abi.DENOISE_ATROUS), ("learned", _abi.DENOISE_LEARNED)):
        p.path_tracing.denoiser = mode
        r.record(1000, p)  # warm-up
        r.synchronize()
        r.set_timing(True)
        for f in range(frames):
            r.record(1001 + f, p)
        r.synchronize()
        ms, n = r.kernel_time(_abi.KERNEL_DENOISE)
        pt_ms, pt_n = r.kernel_time(_abi.KERNEL_PT)
        r.set_timing(False)
        out[f"{name}_ms"] = round(ms / max(n, 1), 3)
        out["pt_4spp_ms"] = round(pt_ms / max(pt_n, 1), 3)
        log(f"denoiser {name}: {ms / max(n, 1):.3f} ms per frame")
    flops = unet_flops(net.channels, size, size)
    tflops = flops / (out["learned_ms"] * 1e-3) / 1e12
    out["learned_roofline"] = {"bound": "mfma", "achieved": round(tflops, 2), "peak": 2500.0, "unit": "TFLOP/s",
                               "frac": round(tflops / 2500.0, 5), "flops": flops}
    r.set_denoiser(None)
    net.close()
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--spp", type=int, default=64)
    ap.add_argument("--size", type=int, default=1024)
    ap.add_argument("--dims", type=int, nargs=3, default=[512, 512, 400])
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batched", action="store_true", help="skip Batched-32 (plain vs served)")
    ap.add_argument("--no-denoisers", action="store_true", help="skip the per-frame denoiser timings")
    ap.add_argument("--no-dvr", action="store_true", help="skip the DVR-1 frame timings")
    ap.add_argument("--no-autotune", dest="autotune", action="store_false",
                    help="keep the default shape instead of the autotuner's (vvr_renderer_autotune)")
    ap.add_argument("--serve-seconds", type=float, default=60.0,
                    help="length of the Poisson serving trace at 20 requests/s (0: skip it)")
    ap.add_argument("--no-quality", action="store_true",
                    help="skip the served frame's relative RMSE against a 4096-path frame")
    ap.add_argument("--only", choices=["dvr", "denoisers"], default=None,
                    help="run only this part (for profiling): no PT-1 steps and no JSON line")
    ap.add_argument("--serve-slots", type=int, default=2 << 20, help="path slots of the server in the trace")
    ap.add_argument("--serve-iterations", type=int, default=0,
                    help="wavefront iterations per pass of the server in the trace (0: its default)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="first copy MIB mebibytes device to device (a plain copy kernel whose HBM counters "
                         "scale the path tracer's under rocprofv3 --pmc)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU: the renderer has no multi-GPU path to measure "
              f"(asked for {args.gpus}).", file=sys.stderr)
        return 2

    t_setup = time.perf_counter()
    ctx = Context(0)
    if args.copy_probe:
        copy_probe(args.copy_probe << 20)
        print(f"copy probe: {args.copy_probe} MiB", file=sys.stderr, flush=True)
    if args.only == "dvr":
        print(json.dumps({"dvr": run_dvr(ctx)}), flush=True)
        return 0
    scene = workloads.pt_scene(ctx, dims=tuple(args.dims), size=(args.size, args.size), spp=args.spp)
    r = scene.renderer
    print(f"setup {time.perf_counter() - t_setup:.1f} s: {scene.n_labels} organs", file=sys.stderr, flush=True)

    r.set_timing(True)  # the first record builds the masks and the majorant (autotune's, warm-up's or the first step's)
    shape = r.pt_shape()
    if args.autotune:  # untimed: the autotuner's trial frames of the workload itself
        t_tune = time.perf_counter()
        shape = r.autotune(scene.params, trials=1)
        print(f"autotune {time.perf_counter() - t_tune:.1f} s: {shape}", file=sys.stderr, flush=True)

    frame = 0
    for _ in range(args.warmup):
        frame += 1
        r.record(frame, scene.params)
        r.synchronize()
        print(f"warmup frame {frame} done", file=sys.stderr, flush=True)
    pt_majorant_ms, _ = r.kernel_time(_abi.KERNEL_MAJORANT)
    r.set_timing(True)
    lat = []
    r.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        frame += 1
        ts = time.perf_counter()
        r.record(frame, scene.params)
        r.synchronize()
        lat.append(time.perf_counter() - ts)
        print(f"step {len(lat)}: {lat[-1] * 1e3:.1f} ms", file=sys.stderr, flush=True)
    r.synchronize()
    elapsed = time.perf_counter() - t0
    pt_ms, pt_launches = r.kernel_time(_abi.KERNEL_PT)
    pt_majorant_ms = max(pt_majorant_ms, r.kernel_time(_abi.KERNEL_MAJORANT)[0])
    r.set_timing(False)

    samples = args.size * args.size * args.spp * args.steps
    value = samples / elapsed / 1e6
    avg_kernel_s = pt_ms / 1e3 / max(pt_launches, 1)
    algo = pt_frame_bytes(args.dims, args.size, args.size)
    achieved = algo / avg_kernel_s / 1e9
    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Msamples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"PT-1 plain: {args.dims[0]}x{args.dims[1]}x{args.dims[2]} int16 phantom, "
                               f"{scene.n_labels} organs (isolate, soft margins), {args.size}^2, {args.spp} spp, "
                               "8 bounces, HG g=0.8, one request at a time",
                   "mode": "plain", "shape": shape, "shape_from": "autotune" if args.autotune else "defaults"},
        "latency_ms": {"p50": round(float(np.percentile(lat, 50)) * 1e3, 3),
                       "p99": round(float(np.percentile(lat, 99)) * 1e3, 3)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 3), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 6), "traffic": measured_traffic(args, shape),
                     "kernel": "PT pass (k_wf_* iterations + k_wf_resolve)", "kernel_ms": round(avg_kernel_s * 1e3, 3), "algorithmic_bytes": algo},
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(scene, scene.params, args.cpu_seconds)
    out["majorant"] = {"workload": f"PT-1 volume {args.dims[0]}x{args.dims[1]}x{args.dims[2]}, isolate label mode",
                       "ms": round(pt_majorant_ms, 4),
                       "roofline": roofline_gbs(majorant_bytes(args.dims, _abi.LABELS_ISOLATE), pt_majorant_ms)}
    if not args.no_denoisers or args.only == "denoisers":
        out["denoisers"] = run_denoisers(r, scene.params, args.size)
    r.close()
    if args.only == "denoisers":
        print(json.dumps({"denoisers": out["denoisers"]}), flush=True)
        return 0
    if not args.no_dvr:
        out["dvr"] = run_dvr(ctx, cpu_seconds=0.0 if args.no_cpu else 8.0)
    if not args.no_batched:
        out["batched"] = run_batched(ctx)
    if args.serve_seconds > 0:
        out["serving"] = run_serving(ctx, args.serve_seconds, slots=args.serve_slots,
                                     iterations=args.serve_iterations)
    if not args.no_quality:
        out["served_quality"] = served_quality(ctx)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
