"""Pins of the oracle's texture filtering (common.glsl:68-136).

Cubic B-spline (Sigg-Hadwiger fold): partition of unity, exact reproduction of
affine fields, and equality with the direct 64-tap B-spline sum (brute force).
Trilinear: closed form on random texels, clamp-to-edge at the border.
"""
import numpy as np
import pytest

from helpers import flat_tf, frame_params, uniforms


def bspline_weights(f):
    return np.array([(1 - f) ** 3 / 6, (4 - 6 * f * f + 3 * f ** 3) / 6, (1 + 3 * f + 3 * f * f - 3 * f ** 3) / 6,
                     f ** 3 / 6])


def bspline_derivs(f):
    return np.array([-0.5 * (1 - f) ** 2, 0.5 * f * (3 * f - 4), 0.5 * (1 - f) * (3 * f + 1), 0.5 * f * f])


def setup(oracle, density):
    hu = np.zeros(density.shape, np.int16)
    tf = flat_tf(0.0)
    u = uniforms(oracle, hu, tf, frame_params(hu))
    return oracle.Scene(density, tf=tf), u


def test_cubic_partition_of_unity(oracle):
    d = np.full((12, 12, 12), 40000, np.uint16)
    sc, u = setup(oracle, d)
    rng = np.random.default_rng(1)
    for idx in rng.uniform(-0.5, 11.5, (200, 3)):
        assert sc.density_cubic(u, idx) == pytest.approx(40000 / 65535, rel=2e-6)
        assert np.all(np.abs(sc.density_gradient_cubic(u, idx)) < 1e-5)


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_cubic_reproduces_affine_field(oracle, axis):
    n = 16
    i = np.arange(n)
    ramp = 1000 + 3000 * i  # counts per index unit along `axis`
    shape = [n, n, n]
    d = np.broadcast_to(ramp.reshape([-1 if a == 2 - axis else 1 for a in range(3)]), shape).astype(np.uint16)
    sc, u = setup(oracle, np.ascontiguousarray(d))
    rng = np.random.default_rng(2)
    # The entry points take f32 positions and return f32 values.
    for idx in rng.uniform(2.0, n - 3.0, (100, 3)).astype(np.float32).astype(np.float64):
        want = (1000 + 3000 * idx[axis]) / 65535
        for prec, tol in ((32, 3e-6), (64, 2e-7)):
            assert sc.density_cubic(u, idx, prec) == pytest.approx(want, rel=tol, abs=tol)
            g = sc.density_gradient_cubic(u, idx, prec)
            assert g[axis] == pytest.approx(3000 / 65535, rel=max(tol, 1e-6) * 10)
            assert np.allclose(np.delete(g, axis), 0.0, atol=1e-5)


def test_cubic_fold_equals_direct_64_tap_sum(oracle):
    rng = np.random.default_rng(3)
    d = rng.integers(0, 65536, (10, 10, 10)).astype(np.uint16)
    sc, u = setup(oracle, d)
    v = d.astype(np.float64) / 65535
    for idx in rng.uniform(2.0, 6.99, (300, 3)).astype(np.float32).astype(np.float64):
        i = np.floor(idx).astype(int)
        f = idx - i
        w = [bspline_weights(f[a]) for a in range(3)]
        dw = [bspline_derivs(f[a]) for a in range(3)]
        block = v[i[2] - 1:i[2] + 3, i[1] - 1:i[1] + 3, i[0] - 1:i[0] + 3]  # [z, y, x]
        direct = np.einsum("zyx,x,y,z->", block, w[0], w[1], w[2])
        assert sc.density_cubic(u, idx, 64) == pytest.approx(direct, abs=1e-7)
        assert sc.density_cubic(u, idx, 32) == pytest.approx(direct, abs=2e-6)
        grad = [np.einsum("zyx,x,y,z->", block, dw[0], w[1], w[2]), np.einsum("zyx,x,y,z->", block, w[0], dw[1], w[2]),
                np.einsum("zyx,x,y,z->", block, w[0], w[1], dw[2])]
        assert np.allclose(sc.density_gradient_cubic(u, idx, 64), grad, atol=1e-7)
        assert np.allclose(sc.density_gradient_cubic(u, idx, 32), grad, atol=2e-5)


def test_trilinear_closed_form_and_clamp_to_edge(oracle):
    rng = np.random.default_rng(4)
    d = rng.integers(0, 65536, (6, 7, 8)).astype(np.uint16)
    sc, u = setup(oracle, d)
    v = d.astype(np.float64) / 65535
    for idx in rng.uniform(-1.5, 8.5, (300, 3)).astype(np.float32).astype(np.float64):
        fl = np.floor(idx).astype(int)
        a = idx - fl
        acc = 0.0
        for dz in (0, 1):
            for dy in (0, 1):
                for dx in (0, 1):
                    x = np.clip(fl[0] + dx, 0, 7)
                    y = np.clip(fl[1] + dy, 0, 6)
                    z = np.clip(fl[2] + dz, 0, 5)
                    w = (a[0] if dx else 1 - a[0]) * (a[1] if dy else 1 - a[1]) * (a[2] if dz else 1 - a[2])
                    acc += w * v[z, y, x]
        assert sc.density_linear(u, idx, 64) == pytest.approx(acc, abs=1e-7)
        assert sc.density_linear(u, idx, 32) == pytest.approx(acc, abs=2e-7)


def test_kernels_division_free_cubic_weights(oracle):
    """The kernels compute the cubic weights' x / 6 as (x / 2) RN(1/3) plus one
    FMA-exact correction (a division below 2^-125). That form equals the
    division for every dividend tried and for the weights of every fraction f
    tried: all floats 101 apart of either sign, all subnormals up to 2^-140,
    every 23rd f in [0, 1) and every 3rd f around 2^-42, where f^3 crosses
    2^-125 (on the CPU all 2^31 positive dividends and all 2^30 fractions were
    checked). The product with RN(1/6) alone differs, so the check can fail."""
    assert oracle.div6_mismatches(0, 0x7F7FFFFF, 101, 0) == 0
    assert oracle.div6_mismatches(0x80000000, 0xFF7FFFFF, 101, 0) == 0
    assert oracle.div6_mismatches(1, 0x00080000, 1, 0) == 0
    assert oracle.div6_mismatches(0, 0x3F7FFFFF, 23, 1) == 0
    assert oracle.div6_mismatches(0x2A000000, 0x2B800000, 3, 1) == 0
    assert oracle.div6_mismatches(0x3F800000, 0x40800000, 101, 2) > 1000
