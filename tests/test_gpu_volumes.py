"""GPU parity of the volume passes against the oracle: density upload
(transfer.rs:121-124, all 65,536 HU values), normals.comp, mask.comp and
majorant.comp. Quantised outputs must match bit for bit; the soft-mask pass
uses exp(), whose last-bit rounding may differ between the GPU and the host
libm, so it may differ by one unorm16 step on a small share of voxels.
"""
import numpy as np
import pytest

from helpers import frame_params
from volview_amd import _abi, phantom
from volview_amd.renderer import Renderer, TransferFunction

pytestmark = pytest.mark.gpu


def oracle_uniforms(oracle, r, hu, params, tf, has_labels=False):
    from volview_amd.renderer import make_volume
    g = r.geometry()
    return oracle.pack_uniforms(make_volume(hu), params, oracle.bucket_max(tf), has_labels, 0, 8, 8, 1)


def test_density_upload_every_hu_value(ctx, oracle):
    hu = np.arange(-32768, 32768, dtype=np.int32).astype(np.int16).reshape(32, 32, 64)
    r = Renderer(ctx, hu, TransferFunction.preset(0))
    got = r.read_volume(_abi.VOL_DENSITY)
    assert np.array_equal(got, oracle.hu_to_unorm16(hu))


@pytest.mark.parametrize("case", ["phantom", "anisotropic", (5, 3, 1), (9, 1, 17), (1, 1, 1), (13, 6, 9)])
def test_normals_bit_exact(ctx, oracle, case):
    """Sobel normals equal the oracle's bit for bit, also on volumes smaller
    than one 8^3 tile or not a whole number of micro-bricks (clamped halos)."""
    if case == "phantom":
        hu, spacing = phantom.ct((96, 80, 72), 5), (1.0, 1.0, 1.0)
    elif isinstance(case, tuple):
        hu = np.random.default_rng(sum(case)).integers(-1000, 2000, case[::-1]).astype(np.int16)
        spacing = (0.8, 1.1, 0.9)
    else:
        hu = np.random.default_rng(2).integers(-1000, 2000, (29, 41, 37)).astype(np.int16)
        spacing = (0.7, 0.9, 1.3)
    tf = TransferFunction.preset(0)
    r = Renderer(ctx, hu, tf, spacing=spacing)
    got = r.read_volume(_abi.VOL_NORMALS)
    from volview_amd.renderer import make_volume
    p = frame_params(hu, spacing=spacing)
    u = oracle.pack_uniforms(make_volume(hu, spacing), p, oracle.bucket_max(tf), False, 0, 8, 8, 1)
    want = oracle.Scene(oracle.hu_to_unorm16(hu)).normals(u)
    mism = np.argwhere(np.any(got != want, axis=-1))
    assert len(mism) == 0, f"{len(mism)} voxels differ, first {mism[:5].tolist()}"


def render_once(r, params, w=8, h=8):
    r.resize(w, h)
    r.record(1, params)
    r.synchronize()


@pytest.mark.parametrize("dims", [(512, 256, 256), (256, 256, 256), (77, 50, 43), (5, 3, 1), (9, 1, 17), (1, 1, 1)])
def test_majorant_mode_off_bit_exact(ctx, oracle, dims):
    hu = phantom.ct(dims, 1)
    tf = TransferFunction.preset(0)
    r = Renderer(ctx, hu, tf)
    p = frame_params(hu)
    render_once(r, p)
    got = r.read_majorant()
    u = oracle_uniforms(oracle, r, hu, p, tf)
    want = oracle.Scene(oracle.hu_to_unorm16(hu), tf=tf).majorant(u)
    assert np.array_equal(got, want)
    if min(dims) > 8:
        assert (got > 0).mean() < 0.9 and (got == 0).any()  # bricks are skipped


def test_majorant_label_mode_large_tiles(ctx, oracle):
    """Label mode on a volume of many majorant workgroups: the majorant over the
    GPU's own mask equals the oracle's bit for bit."""
    hu, labels, sg, tf, r, p = organ_scene(ctx, (512, 256, 256), 3, _abi.LABELS_ISOLATE)
    render_once(r, p)
    u = oracle_uniforms(oracle, r, hu, p, tf, has_labels=True)
    sc = oracle.Scene(oracle.hu_to_unorm16(hu), tf=tf, labels=labels)
    sc.set("mask", r.read_volume(_abi.VOL_MASK))
    got = r.read_majorant()
    assert np.array_equal(got, sc.majorant(u))
    assert (got == 0).any() and (got > 0).any()


@pytest.mark.parametrize("dims", [(64, 48, 40), (66, 45, 38)])  # whole and partial 4^3 micro-bricks
@pytest.mark.parametrize("mode", [_abi.LABELS_OFF, _abi.LABELS_TINT])
def test_majorant_after_transfer_function_change(ctx, oracle, mode, dims):
    """The density's cell records are built once and kept: after a change of
    transfer function (and, in tint mode, of the mask) the rebuilt majorant still
    equals the oracle's bit for bit."""
    hu, labels, sg, tf, r, p = organ_scene(ctx, dims, 5, mode, soft=False)
    render_once(r, p)
    first = r.read_majorant()
    for preset in (1, 2):
        tf2 = TransferFunction.preset(preset)
        r.set_transfer_function(tf2)
        if mode == _abi.LABELS_TINT:
            p.labels.colors[preset][3] = 0.0  # a hidden label: a different mask
        render_once(r, p)
        u = oracle_uniforms(oracle, r, hu, p, tf2, has_labels=True)
        sc = oracle.Scene(oracle.hu_to_unorm16(hu), tf=tf2, labels=labels)
        if mode != _abi.LABELS_OFF:
            sc.set("mask", r.read_volume(_abi.VOL_MASK))
        got = r.read_majorant()
        assert np.array_equal(got, sc.majorant(u))
        assert not np.array_equal(got, first)


def organ_scene(ctx, dims, seed, mode, soft=True, **label_kw):
    hu = phantom.ct(dims, seed)
    labels, sg, _, n = phantom.organs(dims, seed, 40)
    tf = TransferFunction.preset(0)
    r = Renderer(ctx, hu, tf)
    r.set_labels(labels, sg if soft else None)
    p = frame_params(hu)
    p.labels.mode = mode
    for k, v in label_kw.items():
        setattr(p.labels, k, v)
    return hu, labels, sg, tf, r, p


@pytest.mark.parametrize("soft_dims,src_axis,flip", [((23, 17, 29), (2, 0, 1), (1, 0, 1)),
                                                     ((70, 41, 30), (1, 2, 0), (0, 1, 0))])
def test_mask_soft_grid_other_shape(ctx, oracle, soft_dims, src_axis, flip):
    """A soft grid coarser than the volume (its node densities staged once)
    and one with more nodes than the volume has voxels (fetched per window),
    each following permuted and flipped volume axes, match the oracle."""
    dims = (36, 30, 26)
    hu = phantom.ct(dims, 9)
    labels, _, _, n = phantom.organs(dims, 9, 12)
    rng = np.random.default_rng(sum(soft_dims))
    sd0, sd1, sd2 = soft_dims
    sg = np.zeros((sd2, sd1, sd0, 4), np.uint8)
    sg[..., 0] = rng.integers(0, n + 1, sg.shape[:3])
    sg[..., 1] = rng.integers(0, n + 1, sg.shape[:3])
    sg[..., 2] = rng.integers(0, 80, sg.shape[:3])
    sg[..., 3] = rng.random(sg.shape[:3]) < 0.05
    tf = TransferFunction.preset(0)
    r = Renderer(ctx, hu, tf)
    r.set_labels(labels, sg, soft_dims=soft_dims, src_axis=src_axis, flip=flip)
    p = frame_params(hu)
    p.labels.mode = _abi.LABELS_ISOLATE
    p.labels.colors[2][3] = 0.0  # a hidden label
    p.labels.ghost[4] = 1
    render_once(r, p)
    got = r.read_volume(_abi.VOL_MASK)
    u = oracle_uniforms(oracle, r, hu, p, tf, has_labels=True)
    sc = oracle.Scene(oracle.hu_to_unorm16(hu), tf=tf, labels=labels)
    flags = flip[0] | (flip[1] << 1) | (flip[2] << 2) | 8
    want = sc.mask(u, sg, soft_dims, (*src_axis, flags), margin=0, radius=1)
    diff = np.abs(got.astype(np.int64) - want.astype(np.int64))
    assert diff.max() <= 1, diff.max()
    assert (diff > 0).mean() < 5e-3
    assert got[..., 0].min() < got[..., 0].max() and got[..., 1].max() > 0  # neither constant nor unghosted


@pytest.mark.parametrize("dims", [(70, 64, 58), (64, 48, 32)])  # partial and whole 4^3 micro-bricks
@pytest.mark.parametrize("mode", [_abi.LABELS_TINT, _abi.LABELS_ISOLATE, _abi.LABELS_HYBRID])
def test_mask_binary_path_bit_exact(ctx, oracle, mode, dims):
    hu, labels, sg, tf, r, p = organ_scene(ctx, dims, 4, mode, soft=False)
    p.labels.colors[3][3] = 0.0  # a hidden label (cut in tint mode)
    p.labels.ghost[5] = 1
    render_once(r, p)
    got = r.read_volume(_abi.VOL_MASK)
    u = oracle_uniforms(oracle, r, hu, p, tf, has_labels=True)
    sc = oracle.Scene(oracle.hu_to_unorm16(hu), tf=tf, labels=labels)
    want = sc.mask(u, None, None, (0, 0, 0, 0), margin=2, radius=1)
    assert np.array_equal(got, want)
    sc.set("mask", got)
    assert np.array_equal(r.read_majorant(), sc.majorant(u))


def test_mask_soft_grid_and_majorant(ctx, oracle):
    hu, labels, sg, tf, r, p = organ_scene(ctx, (80, 72, 64), 6, _abi.LABELS_ISOLATE)
    render_once(r, p)
    got = r.read_volume(_abi.VOL_MASK)
    u = oracle_uniforms(oracle, r, hu, p, tf, has_labels=True)
    sc = oracle.Scene(oracle.hu_to_unorm16(hu), tf=tf, labels=labels)
    want = sc.mask(u, sg, hu.shape[::-1], (0, 1, 2, 8), margin=0, radius=1)
    diff = np.abs(got.astype(np.int64) - want.astype(np.int64))
    assert diff.max() <= 1, diff.max()
    assert (diff > 0).mean() < 5e-3
    # The majorant over the GPU's own mask is bit-exact.
    sc.set("mask", got)
    assert np.array_equal(r.read_majorant(), sc.majorant(u))
