"""Pins of the oracle's volume passes: majorant.comp, normals.comp, mask.comp.

Majorant: brute force of the shader's definition on tiny volumes (windowed
min/max with integer bucket arithmetic), the bound invariant over samples in the
brick, and the integer bucket formula against the f32 one for every density.
Normals: Sobel of linear ramps has a closed form. Masks: binomial weights sum to
one, the binary mask of one label is its indicator, and uniform soft grids give
the closed-form score.
"""
import numpy as np
import pytest

from helpers import flat_tf, frame_params, uniforms
from volview_amd import _abi
from volview_amd.renderer import TransferFunction


def test_bucket_index_integer_formula_equals_f32_for_every_density():
    c = np.arange(65536, dtype=np.int64)
    f32 = ((c.astype(np.float32) / np.float32(65535.0)) * np.float32(64.0)).astype(np.int64)
    assert np.array_equal(f32, (64 * c) // 65535)


def brute_force_majorant(density, mask, bmax, mode, label_ext):
    nz, ny, nx = density.shape
    out = np.zeros(((nz + 7) // 8, (ny + 7) // 8, (nx + 7) // 8), np.float32)
    for bz in range(out.shape[0]):
        for by in range(out.shape[1]):
            for bx in range(out.shape[2]):
                sl = tuple(slice(max(8 * b - 2, 0), min(8 * b + 9, n - 1) + 1)
                           for b, n in ((bz, nz), (by, ny), (bx, nx)))
                w = density[sl].astype(np.int64)
                lo, hi = (64 * w.min()) // 65535, (64 * w.max()) // 65535
                tfm = np.float32(bmax[lo:hi + 1].max())
                if mode == 0:
                    out[bz, by, bx] = tfm
                    continue
                mm = mask[sl].max()
                if mm == 0:
                    out[bz, by, bx] = 0.0
                elif mode == 1:
                    out[bz, by, bx] = tfm
                else:
                    out[bz, by, bx] = label_ext
    return out


@pytest.mark.parametrize("shape", [(17, 20, 23), (8, 8, 8), (3, 9, 30)])
@pytest.mark.parametrize("mode", [0, 1, 2])
def test_majorant_brute_force(oracle, shape, mode):
    rng = np.random.default_rng(sum(shape) + mode)
    hu = rng.integers(-1100, 1500, shape).astype(np.int16)
    hu[:, : shape[1] // 2] = -1000  # some empty bricks
    tf = TransferFunction.preset(0)
    p = frame_params(hu)
    p.labels.mode = mode
    u = uniforms(oracle, hu, tf, p, has_labels=mode != 0)
    density = oracle.hu_to_unorm16(hu)
    mask = None
    if mode:
        mask = rng.integers(0, 65536, shape + (2,)).astype(np.uint16)
        mask[rng.random(shape) < 0.7] = 0
        mask[: shape[0] // 3] = 0
    sc = oracle.Scene(density, tf=tf, mask=mask)
    got = sc.majorant(u)
    want = brute_force_majorant(density, None if mask is None else mask.max(-1), oracle.bucket_max(tf), mode,
                                np.float32(p.labels.extinction))
    assert np.array_equal(got, want)


def test_majorant_bounds_every_sample_in_the_brick(oracle):
    # Invariant for a non-decreasing transfer function (the bone preset): the bound
    # is >= the extinction of every trilinear or cubic sample whose support lies in
    # the brick's window, since B-spline and trilinear weights are convex.
    rng = np.random.default_rng(9)
    hu = rng.integers(-1000, 2000, (24, 24, 24)).astype(np.int16)
    tf = TransferFunction.preset(0)
    u = uniforms(oracle, hu, tf, frame_params(hu))
    sc = oracle.Scene(oracle.hu_to_unorm16(hu), tf=tf)
    maj = sc.majorant(u)
    lut = oracle.tf_lut(tf)
    for _ in range(400):
        b = rng.integers(0, 3, 3)
        idx = 8 * b - 0.5 + rng.uniform(0, 8, 3)  # inside brick b: floor((idx + 0.5) / 8) == b
        for d in (sc.density_linear(u, idx), sc.density_cubic(u, idx)):
            t = d * 4095
            i0 = int(np.clip(np.floor(t), 0, 4095))
            ext = lut[i0, 3] * (1 - (t - np.floor(t))) + lut[min(i0 + 1, 4095), 3] * (t - np.floor(t))
            assert maj[b[2], b[1], b[0]] >= ext - 1e-6


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_normals_of_a_ramp(oracle, axis):
    n = 10
    ramp = 20000 + 100 * np.arange(n)
    shape = [n, n, n]
    hu = np.zeros(shape, np.int16)
    density = np.ascontiguousarray(np.broadcast_to(ramp.reshape([-1 if a == 2 - axis else 1 for a in range(3)]),
                                                   shape).astype(np.uint16))
    u = uniforms(oracle, hu, flat_tf(0.0), frame_params(hu))
    got = oracle.Scene(density).normals(u)
    g = 100 / 65535  # per voxel and per mm (1 mm spacing)
    want = np.zeros(4, np.int8)
    want[axis] = -127
    want[3] = int(np.floor(min(g * 400, 1.0) * 127 + 0.5))
    interior = got[1:-1, 1:-1, 1:-1].reshape(-1, 4)
    assert np.all(interior == want)


def mask_uniforms(oracle, hu, mode, shown=(), hidden=(), ghost=()):
    p = frame_params(hu)
    p.labels.mode = mode
    for k in hidden:
        p.labels.colors[k][3] = 0.0
    for k in ghost:
        p.labels.ghost[k] = 1
    return uniforms(oracle, hu, flat_tf(0.0), p, has_labels=True)


def test_binary_mask_of_one_label_is_its_indicator(oracle):
    rng = np.random.default_rng(5)
    labels = (rng.random((9, 10, 11)) < 0.3).astype(np.uint8) * 7
    hu = np.zeros(labels.shape, np.int16)
    u = mask_uniforms(oracle, hu, _abi.LABELS_ISOLATE)
    sc = oracle.Scene(oracle.hu_to_unorm16(hu), labels=labels)
    got = sc.mask(u, radius=0)
    assert np.array_equal(got[..., 0], np.where(labels == 7, 65535, 0))
    assert np.all(got[..., 1] == 0)
    # Ghosted: the second channel instead.
    u = mask_uniforms(oracle, hu, _abi.LABELS_ISOLATE, ghost=(7,))
    got = sc.mask(u, radius=0)
    assert np.all(got[..., 0] == 0)
    assert np.array_equal(got[..., 1], np.where(labels == 7, 65535, 0))


@pytest.mark.parametrize("radius", [1, 2, 3])
def test_blur_weights_sum_to_one_and_match_binomials(oracle, radius):
    shape = (12, 12, 16)
    labels = np.zeros(shape, np.uint8)
    labels[:, :, 8] = 3  # a plane x = 8
    hu = np.zeros(shape, np.int16)
    u = mask_uniforms(oracle, hu, _abi.LABELS_ISOLATE)
    sc = oracle.Scene(oracle.hu_to_unorm16(hu), labels=labels)
    got = sc.mask(u, radius=radius)[..., 0]
    from math import comb
    want_x = np.zeros(16)
    for k in range(-radius, radius + 1):
        want_x[8 + k] = np.floor(65535 * comb(2 * radius, radius + k) / 4 ** radius + 0.5)
    assert np.array_equal(got, np.broadcast_to(want_x, shape))
    labels[:] = 3  # everything: stays exactly 1 at every radius
    sc = oracle.Scene(oracle.hu_to_unorm16(hu), labels=labels)
    assert np.all(sc.mask(u, radius=radius)[..., 0] == 65535)


@pytest.mark.parametrize("margin16", [0, 8, 40, 64, 200])
def test_uniform_soft_grid_gives_closed_form_score(oracle, margin16):
    # Every node (best 5, runner-up 0, margin m): every voxel's 27 nodes agree, so
    # the weighted mean score is min(m, 4) exactly (mask.comp:32-37, 99-137).
    shape = (6, 7, 8)
    labels = np.full(shape, 5, np.uint8)
    hu = np.random.default_rng(1).integers(-200, 200, shape).astype(np.int16)
    soft = np.zeros(shape + (4,), np.uint8)
    soft[..., 0] = 5
    soft[..., 2] = margin16
    u = mask_uniforms(oracle, hu, _abi.LABELS_ISOLATE)
    sc = oracle.Scene(oracle.hu_to_unorm16(hu), labels=labels)
    got = sc.mask_binary(u, soft, shape[::-1], (0, 1, 2, 8))
    want = 0.5 + min(margin16 / 16, 4.0) / 8
    assert np.allclose(got[..., 0], want, atol=1e-6)
    assert np.allclose(got[..., 1], 0.0, atol=1e-6)  # nothing ghosted: score -4
    # Runner-up drawn and best not: the negative score.
    soft[..., 0], soft[..., 1] = 0, 5
    got = sc.mask_binary(u, soft, shape[::-1], (0, 1, 2, 8))
    assert np.allclose(got[..., 0], 0.5 - min(margin16 / 16, 4.0) / 8, atol=1e-6)
    # Overridden nodes keep the binary mask.
    soft[..., 3] = 1
    got = sc.mask_binary(u, soft, shape[::-1], (0, 1, 2, 8))
    assert np.all(got[..., 0] == 1.0)


def test_tint_cut_margin_removes_unlabelled_neighbours(oracle):
    shape = (9, 9, 9)
    labels = np.zeros(shape, np.uint8)
    labels[4, 4, 4] = 9
    hu = np.zeros(shape, np.int16)
    u = mask_uniforms(oracle, hu, _abi.LABELS_TINT, hidden=(9,))
    sc = oracle.Scene(oracle.hu_to_unorm16(hu), labels=labels)
    got = sc.mask_binary(u, margin=2)[..., 0]
    z, y, x = np.mgrid[0:9, 0:9, 0:9]
    near = (np.abs(x - 4) <= 2) & (np.abs(y - 4) <= 2) & (np.abs(z - 4) <= 2)
    assert np.all(got[near] == 0.0)       # hidden label and unlabelled voxels within 2
    assert np.all(got[~near] == 1.0)      # unlabelled tissue is solid in tint mode
