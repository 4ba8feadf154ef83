"""Training of the learned denoiser on pairs the path tracer renders itself.

    python -m volview_amd.train_denoiser --out weights.npy
    python -m volview_amd.train_denoiser --evaluate volview_amd/data/denoiser_unet.npy

Pairs: seeded CT phantoms with labelled organs (edge U{96..128}), `--cameras`
seeded cameras each, 128 x 128 frames traced at `--low` paths per pixel (the
input: the accumulation and guide images) and at `--high` paths (the reference). The
network (5 levels, 32-128 channels) trains on random 64 x 64 crops with Adam on
the GPU through libvvr's own kernels; held-out phantoms then score it by
relative MSE against the reference, beside the raw frame and the a-trous
filter. Prints one line per evaluation and progress every 100 steps.

--evaluate scores given weights only: held-out phantoms (seeds 9000 on), two
cameras each, `--size` frames, the network on `--low`-path frames and the
a-trous filter on 64-path frames against `--high`-path references.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from . import _abi, phantom, workloads
from .renderer import Context, Renderer, TransferFunction
from .unet import Unet

CHANNELS = (32, 48, 64, 96, 128)
SIZE = 128


def rel_mse(a: np.ndarray, ref: np.ndarray) -> float:
    """Mean of (a - ref)^2 / (ref^2 + 0.01) over pixels and channels."""
    return float(np.mean((a - ref) ** 2 / (ref ** 2 + 1e-2)))


def rel_rmse(a: np.ndarray, ref: np.ndarray) -> float:
    """sqrt(sum (a - ref)^2 / sum ref^2) over pixels and channels."""
    return float(np.sqrt(np.sum((a - ref) ** 2) / np.sum(ref ** 2)))


def scene(ctx: Context, seed: int, size: int = SIZE):
    rng = np.random.default_rng(seed)
    e = int(rng.integers(96, 129))
    dims = (e, e, e)
    hu = phantom.ct(dims, seed)
    labels, soft, _, n = phantom.organs(dims, seed, 20)
    r = Renderer(ctx, hu, TransferFunction.preset(0))
    r.set_labels(labels, soft)
    r.resize(size, size)
    return r, n


def render_pairs(ctx: Context, seeds, low: int, high: int, cameras=range(4), atrous_spp=(), size: int = SIZE):
    """Per (seed, camera): the low frame's images, the reference radiance, and
    the a-trous filtered frames at the given path counts."""
    out = []
    t0 = time.perf_counter()
    for j, s in enumerate(seeds):
        if j and j % 20 == 0:
            print(f"rendered {j} phantoms ({time.perf_counter() - t0:.1f} s)", flush=True)
        r, n = scene(ctx, s, size)
        for cam in cameras:
            p = workloads.pt_params(r, n, high, bounces=8, camera_index=cam)
            r.record(1, p)
            ref = r.read_hdr()
            p = workloads.pt_params(r, n, low, bounces=8, camera_index=cam)
            r.record(2, p)
            acc, alb, nrm = r.read_accum()
            item = {"accum": acc, "albedo": alb, "normal": nrm, "ref": ref, "raw": r.read_hdr()}
            for spp in atrous_spp:
                q = workloads.pt_params(r, n, spp, bounces=8, camera_index=cam)
                q.path_tracing.denoiser = _abi.DENOISE_ATROUS
                r.resize(size, size)  # restarts accumulation: a frame of `spp` paths
                r.record(3, q)
                item[f"atrous{spp}"] = r.read_denoised()
                q.path_tracing.denoiser = _abi.DENOISE_OFF
                r.resize(size, size)
                r.record(4, q)
                item[f"raw{spp}"] = r.read_hdr()
            out.append(item)
        r.close()
    return out


def evaluate(net: Unet, pairs, low: int, metric=rel_mse) -> dict:
    scores = {}
    for it in pairs:
        den = net.denoise(it["accum"], it["albedo"], it["normal"], float(low))
        ref = it["ref"][..., :3]
        cand = {"raw": it["raw"], "unet": den}
        cand.update({k: v for k, v in it.items() if k.startswith("atrous") or (k.startswith("raw") and k != "raw")})
        for k, v in cand.items():
            scores.setdefault(k, []).append(metric(v[..., :3], ref))
    return {k: float(np.mean(v)) for k, v in scores.items()}


def score(ctx: Context, weights: str, phantoms: int, low: int, high: int, size: int) -> int:
    """Held-out scoring of trained weights; one JSON line per phantom and a summary."""
    net = Unet(ctx, CHANNELS)
    net.set_params(np.load(weights).astype(np.float32))
    t0 = time.perf_counter()
    per = {"rel_mse": {}, "rel_rmse": {}}
    for s in range(9000, 9000 + phantoms):
        pairs = render_pairs(ctx, [s], low, high, cameras=(0, 2), atrous_spp=(64,), size=size)
        for name, metric in (("rel_mse", rel_mse), ("rel_rmse", rel_rmse)):
            for k, v in evaluate(net, pairs, low, metric).items():
                per[name].setdefault(k, []).append(v)
        print(json.dumps({"seed": s, "rel_rmse": {k: round(v[-1], 5) for k, v in per["rel_rmse"].items()},
                          "seconds": round(time.perf_counter() - t0, 1)}), flush=True)
    summary = {name: {k: round(float(np.mean(v)), 5) for k, v in d.items()} for name, d in per.items()}
    summary.update({"phantoms": phantoms, "cameras": 2, "size": size, "low": low, "high": high,
                    "seconds": round(time.perf_counter() - t0, 1)})
    print(json.dumps(summary), flush=True)
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--out")
    ap.add_argument("--evaluate", metavar="WEIGHTS", help="score these weights on held-out phantoms")
    ap.add_argument("--phantoms", type=int, default=20, help="held-out phantoms for --evaluate")
    ap.add_argument("--size", type=int, default=256, help="frame size for --evaluate")
    ap.add_argument("--train-seeds", type=int, default=200)
    ap.add_argument("--cameras", type=int, default=8, help="seeded cameras per training phantom")
    ap.add_argument("--test-seeds", type=int, default=6)
    ap.add_argument("--low", type=int, default=4)
    ap.add_argument("--high", type=int, default=4096)
    ap.add_argument("--steps", type=int, default=30000)
    ap.add_argument("--batch", type=int, default=16)
    ap.add_argument("--lr", type=float, default=1e-3)
    args = ap.parse_args()
    ctx = Context(0)
    if args.evaluate:
        return score(ctx, args.evaluate, args.phantoms, args.low, args.high, args.size)
    if not args.out:
        ap.error("--out is required for training")
    t0 = time.perf_counter()
    train = render_pairs(ctx, range(1000, 1000 + args.train_seeds), args.low, args.high, cameras=range(args.cameras))
    print(f"rendered {len(train)} training pairs in {time.perf_counter() - t0:.1f} s", flush=True)
    test = render_pairs(ctx, range(5000, 5000 + args.test_seeds), args.low, args.high, cameras=(0, 2),
                        atrous_spp=(args.low, 64))
    print(f"rendered {len(test)} test pairs ({time.perf_counter() - t0:.1f} s)", flush=True)
    imgs = {k: np.stack([it[k] for it in train]) for k in ("accum", "albedo", "normal", "ref")}
    net = Unet(ctx, CHANNELS, seed=1)
    rng = np.random.default_rng(0)
    crop = 64
    t1 = time.perf_counter()
    losses = []
    for step in range(1, args.steps + 1):
        idx = rng.integers(0, len(train), args.batch)
        ys = rng.integers(0, SIZE - crop + 1, args.batch)
        xs = rng.integers(0, SIZE - crop + 1, args.batch)
        b = {k: np.stack([v[i, y:y + crop, x:x + crop] for i, y, x in zip(idx, ys, xs)]) for k, v in imgs.items()}
        lr = args.lr * (0.1 ** (step / args.steps))  # exponential decay to a tenth
        losses.append(net.train_step(b["accum"], b["albedo"], b["normal"], float(args.low), b["ref"], lr))
        if step % 100 == 0:
            print(f"step {step}: loss {np.mean(losses[-100:]):.4f} ({(time.perf_counter() - t1) / step * 1e3:.1f} ms/step)",
                  flush=True)
    scores = evaluate(net, test, args.low)
    print(json.dumps({"rel_mse": scores, "steps": args.steps, "train_pairs": len(train), "low": args.low,
                      "high": args.high, "seconds": round(time.perf_counter() - t0, 1)}), flush=True)
    np.save(args.out, net.params().astype(np.float16))
    return 0


if __name__ == "__main__":
    sys.exit(main())
