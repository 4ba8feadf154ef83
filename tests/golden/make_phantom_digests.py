"""Writes the SHA-256 digests of the benchmark phantoms and small crops of them
(phantom_digests.json, phantom_crops.npz) that test_phantom_digests.py pins the
generator (csrc/phantom.cpp) against. Run from the repository root after a
build: python tests/golden/make_phantom_digests.py"""
from __future__ import annotations

import hashlib
import json
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from volview_amd import phantom, workloads  # noqa: E402

HERE = Path(__file__).resolve().parent

# Crops: (name, volume key, z, y, x) of 16^3 corners: through the DVR-1 phantom's
# body, a lung, the vessel and a bone shell, and PT-1's density and labels at an
# organ.
CROPS = [
    ("dvr1_body", "dvr1_hu", 120, 120, 120),
    ("dvr1_lung", "dvr1_hu", 150, 120, 80),
    ("dvr1_vessel", "dvr1_hu", 120, 140, 105),
    ("dvr1_bone", "dvr1_hu", 157, 99, 128),
    ("pt1_hu", "pt1_hu", 184, 200, 384),
    ("pt1_labels", "pt1_labels", 184, 200, 384),
]


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def volumes():
    """The benchmark phantoms by name (arrays indexed [z, y, x])."""
    out = {"dvr1_hu": phantom.ct((256, 256, 256), 1), "pt1_hu": phantom.ct((512, 512, 400), 1)}
    labels, soft, margin, n = phantom.organs((512, 512, 400), 1, 40)
    out.update(pt1_labels=labels, pt1_soft=soft, pt1_margin=margin.view(np.uint16))
    return out, n


def batched_specs():
    vols, reqs = workloads.batched32()
    return {"volumes": [[list(d), s] for d, s in vols],
            "requests": [[r.volume, r.path_traced, r.width, r.height, r.spp, r.camera_index] for r in reqs]}


def digests():
    vols, n = volumes()
    d = {k: sha(v) for k, v in vols.items()}
    d["pt1_organs_placed"] = n
    specs = batched_specs()
    d["batched32_specs"] = hashlib.sha256(json.dumps(specs, sort_keys=True).encode()).hexdigest()
    d["batched32_volumes"] = hashlib.sha256(
        "".join(sha(phantom.ct(tuple(dims), seed)) for dims, seed in specs["volumes"]).encode()).hexdigest()
    crops = {name: vols[key][z:z + 16, y:y + 16, x:x + 16] for name, key, z, y, x in CROPS}
    return d, crops


if __name__ == "__main__":
    d, crops = digests()
    (HERE / "phantom_digests.json").write_text(json.dumps(d, indent=1, sort_keys=True) + "\n")
    np.savez_compressed(HERE / "phantom_crops.npz", **crops)
    print(json.dumps(d, indent=1, sort_keys=True))
