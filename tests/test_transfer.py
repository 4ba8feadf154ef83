"""Transfer functions, density normalisation and the fp16 LUT.

Pins: the reference's own known answers (crates/render/src/transfer.rs:126-150),
IEEE binary16 rounding checked against numpy, and closed forms for the
cumulative integrals and bucket maxima. Both the oracle and the product's host
code are checked against the pins and against each other.
"""
import numpy as np
import pytest

from volview_amd.renderer import TransferFunction, hu_to_unorm16

TWO_POINT = [(0.0, (0.0, 0.0, 0.0), 0.0), (100.0, (1.0, 1.0, 1.0), 2.0)]  # transfer.rs:133-136


def test_reference_interpolates_between_points(oracle):
    # transfer.rs:131-139
    tf = TransferFunction(TWO_POINT)
    for impl in (tf.evaluate, lambda h: oracle.tf_evaluate(TWO_POINT, h)):
        assert list(impl(-50.0)) == [0.0, 0.0, 0.0, 0.0]
        assert list(impl(50.0)) == [0.5, 0.5, 0.5, 1.0]
        assert list(impl(500.0)) == [1.0, 1.0, 1.0, 2.0]


def test_reference_lut_matches_density_normalisation(oracle):
    # transfer.rs:141-150
    lib = oracle.load()
    for impl in (hu_to_unorm16, lib.oracle_hu_to_unorm16):
        assert impl(-1024) == 0
        assert impl(3071) == 65535
        assert impl(-2000) == 0
    bone = TransferFunction.preset(0)
    lut = bone.lut()
    assert np.array_equal(lut[1024 + 300], bone.evaluate(300.0))
    assert np.array_equal(oracle.tf_lut(bone)[1024 + 300], oracle.tf_evaluate(bone, 300.0))


def test_hu_to_unorm16_all_values_agree_and_follow_closed_form(oracle):
    hu = np.arange(-32768, 32768, dtype=np.int32)
    prod = np.array([hu_to_unorm16(int(h)) for h in hu[::7]], np.int64)
    orc = oracle.hu_to_unorm16(hu[::7].astype(np.int16)).astype(np.int64)
    assert np.array_equal(prod, orc)
    # Closed form in exact arithmetic: round(65535 * clamp((hu + 1024) / 4095, 0, 1)).
    # The f32 evaluation can differ only where the exact value is within f32
    # rounding of a .5 tie; there are none in the HU range.
    t = np.clip((hu[::7] + 1024) / 4095.0, 0.0, 1.0) * 65535.0
    assert np.array_equal(orc, np.floor(t + 0.5).astype(np.int64))


def test_presets_match_reference_table():
    # transfer.rs:37-41 (bone): extinction 0 below 180 HU, 0.6 at 300, 3 at 1200, 4 at 3071.
    bone = TransferFunction.preset(0)
    assert [p[0] for p in bone.points] == [-1024.0, 180.0, 300.0, 1200.0, 3071.0]
    assert [p[2] for p in bone.points] == pytest.approx([0.0, 0.0, 0.6, 3.0, 4.0])
    assert len(TransferFunction.preset(1).points) == 6
    assert len(TransferFunction.preset(2).points) == 9
    with pytest.raises(ValueError):
        TransferFunction.preset(3)


def test_f16_rounding_matches_ieee(oracle):
    rng = np.random.default_rng(0)
    bits = np.concatenate([rng.integers(0, 2**32, 20000, dtype=np.uint64).astype(np.uint32),
                           np.array([0x38000000, 0x33000000, 0x33000001, 0x477FE000, 0x477FF000, 0x7F800000,
                                     0x00000001, 0x387FC000, 0x3F800000, 0xBF801000], np.uint32)])
    f = bits.view(np.float32)
    f = f[np.isfinite(f)]
    with np.errstate(over="ignore"):
        ref = f.astype(np.float16).view(np.uint16)
    got = np.array([oracle.f32_to_f16(float(x)) for x in f], np.uint16)
    assert np.array_equal(got, ref)
    back = np.array([oracle.f16_to_f32(int(h)) for h in range(0, 65536, 3)], np.float32)
    ref_back = np.arange(0, 65536, 3, dtype=np.uint16).view(np.float16).astype(np.float32)
    finite = np.isfinite(ref_back)
    assert np.array_equal(back[finite], ref_back[finite])


@pytest.mark.parametrize("preset", [0, 1, 2])
def test_lut_f16_product_equals_oracle_and_numpy(oracle, preset):
    tf = TransferFunction.preset(preset)
    prod = tf.lut_f16()
    orc = oracle.tf_lut_f16(tf)
    assert np.array_equal(prod, orc)
    assert np.array_equal(orc, oracle.tf_lut(tf).astype(np.float16).view(np.uint16))


@pytest.mark.parametrize("preset", [0, 1, 2])
def test_bucket_max_definition(oracle, preset):
    # renderer.rs:2164-2174: max extinction over LUT entries 64b .. 64b + 64 (inclusive of the next first).
    tf = TransferFunction.preset(preset)
    lut = oracle.tf_lut(tf)
    want = np.array([lut[64 * b: min(64 * b + 65, 4096), 3].max() for b in range(64)], np.float32)
    assert np.array_equal(oracle.bucket_max(tf), want)
    assert np.array_equal(tf.bucket_max(), want)


def test_cumulative_integrals_closed_form(oracle):
    # A linear extinction ramp e(d) = 2 d over density d in [0, 1], constant albedo 0.5:
    # the trapezoid rule is exact for linear integrands: cumulative e = d^2, albedo*e = 0.5 d^2.
    tf = [(-1024.0, (0.5, 0.5, 0.5), 0.0), (3071.0, (0.5, 0.5, 0.5), 2.0)]
    got = oracle.tf_cumulative(tf)
    d = np.arange(4096) / 4095.0
    lut = oracle.tf_lut(tf)  # the f32 LUT entries being integrated (near 2 d)
    ext = lut[:, 3].astype(np.float64)
    trap = np.concatenate([[0.0], np.cumsum(0.5 * (ext[1:] + ext[:-1]) / 4095.0)])
    assert np.allclose(got[:, 3], trap, rtol=0, atol=1e-6)
    assert np.allclose(got[:, 3], d * d, atol=1e-6)
    assert np.allclose(got[:, 0], 0.5 * d * d, atol=1e-6)
    prod = TransferFunction(tf).cumulative_integrals()
    assert np.array_equal(prod, got)
