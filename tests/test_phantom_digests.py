"""The benchmark phantoms are pinned: the generator (csrc/phantom.cpp) must
reproduce the SHA-256 digests and 16^3 crops in tests/golden/ of the DVR-1
volume (256^3, seed 1), the PT-1 volume with its 40 organs (labels, soft
grid, margin; 512x512x400, seed 1) and the 32 Batched-32 volumes with their
request specs. The fixtures come from tests/golden/make_phantom_digests.py."""
import json
from pathlib import Path

import numpy as np
import pytest

GOLDEN = Path(__file__).resolve().parent / "golden"


@pytest.fixture(scope="module")
def made():
    import sys
    sys.path.insert(0, str(GOLDEN))
    import make_phantom_digests as m
    return m.digests()


def test_phantom_digests(made):
    want = json.loads((GOLDEN / "phantom_digests.json").read_text())
    got, _ = made
    assert got == want


def test_phantom_crops(made):
    _, crops = made
    with np.load(GOLDEN / "phantom_crops.npz") as want:
        assert sorted(want.files) == sorted(crops)
        for name in want.files:
            assert np.array_equal(crops[name], want[name]), name
        # The crops hold what their names say: soft tissue, lung, the 300 HU
        # vessel, bone, and an organ label.
        assert -100 < want["dvr1_body"].mean() < 150
        assert want["dvr1_lung"].max() < -500
        assert ((want["dvr1_vessel"] > 200) & (want["dvr1_vessel"] < 400)).any()
        assert want["dvr1_bone"].max() > 700
        assert (want["pt1_labels"] > 0).any()
