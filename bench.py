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
from volview_amd.renderer import Context, make_volume  # noqa: E402

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
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": done / t_total / 1e6, "unit": "Msamples/s", "cores": cores, "kind": "port",
            "sample": f"{done} paths (1 per random pixel) of the PT-1 frame, {t_total:.1f} s, FP32 C++ oracle"}


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


def measured_traffic(args) -> int | None:
    """HBM bytes of one path-tracing pass at the default PT-1 configuration, as
    measured by rocprofv3 FETCH_SIZE / WRITE_SIZE passes scaled by a copy of
    known size (profiles/pt1_wavefront_hbm_traffic.md); None for other sizes."""
    if (args.spp, args.size, tuple(args.dims)) != (64, 1024, (512, 512, 400)):
        return None
    prof = ROOT / "profiles" / "pt1_wavefront_hbm_traffic.md"
    if not prof.exists():
        return None
    for line in prof.read_text().splitlines():
        if line.startswith("traffic_bytes_per_pass:"):
            return int(line.split(":")[1])
    return None


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
    scene = workloads.pt_scene(ctx, dims=tuple(args.dims), size=(args.size, args.size), spp=args.spp)
    r = scene.renderer
    print(f"setup {time.perf_counter() - t_setup:.1f} s: {scene.n_labels} organs", file=sys.stderr, flush=True)

    if args.copy_probe:
        copy_probe(args.copy_probe << 20)
        print(f"copy probe: {args.copy_probe} MiB", file=sys.stderr, flush=True)

    frame = 0
    for _ in range(args.warmup):
        frame += 1
        r.record(frame, scene.params)
        r.synchronize()
        print(f"warmup frame {frame} done", file=sys.stderr, flush=True)
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
                   "mode": "plain"},
        "latency_ms": {"p50": round(float(np.percentile(lat, 50)) * 1e3, 3),
                       "p99": round(float(np.percentile(lat, 99)) * 1e3, 3)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 3), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 6), "traffic": measured_traffic(args),
                     "kernel": "PT pass (k_wf_* iterations + k_wf_resolve)", "kernel_ms": round(avg_kernel_s * 1e3, 3), "algorithmic_bytes": algo},
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(scene, scene.params, args.cpu_seconds)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
