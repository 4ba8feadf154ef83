"""Pins of the oracle's U-Net (oracle/unet_ref.py): its hand-written backward
equals central finite differences of its own loss (float64), for parameters
of every layer, on a two- and a three-level network; pooling and upsampling
are adjoint (<pool x, y> = <x, pool_back y>, <up x, y> = <x, up_back y>); a
convolution's input gradient is adjoint to the convolution."""
import numpy as np
import pytest

import unet_ref as U  # oracle/unet_ref.py


def test_pool_up_and_conv_adjoints():
    rng = np.random.default_rng(0)
    x = rng.normal(size=(2, 8, 6, 3))
    y = rng.normal(size=(2, 4, 3, 3))
    assert np.isclose((U.pool(x) * y).sum(), (x * U.pool_back(y)).sum())
    z = rng.normal(size=(2, 16, 12, 3))
    assert np.isclose((U.up(x) * z).sum(), (x * U.up_back(z)).sum())
    w = rng.normal(size=(5, 9, 3))
    gy = rng.normal(size=(2, 8, 6, 5))
    gx, _, _ = U.conv_back(x, w, gy)
    assert np.isclose((U.conv(x, w, 0.0) * gy).sum(), (x * gx).sum())


@pytest.mark.parametrize("channels", [(16, 32), (16, 16, 32)])
def test_backward_equals_finite_differences(channels):
    rng = np.random.default_rng(4)
    params = U.init_params(channels, seed=3)
    x = rng.normal(size=(2, 8, 8, 16)) * 0.5
    target = rng.normal(size=(2, 8, 8, 3))
    loss, grads, _ = U.loss_and_grads(params, channels, x, target)
    # Parameters of every layer: offsets of each layer's block.
    offs, k = [], 0
    for cin, cout in U.layer_shapes(channels):
        offs.append(k)
        k += cout * 9 * cin + cout
    picks = [o + int(rng.integers(0, 9 * 16)) for o in offs] + [offs[-1] + 3 * 9 * channels[0]]  # + an output bias
    for i in picks:
        h = 1e-6
        p1, p2 = params.copy(), params.copy()
        p1[i] += h
        p2[i] -= h
        fd = (U.loss_and_grads(p1, channels, x, target)[0] - U.loss_and_grads(p2, channels, x, target)[0]) / (2 * h)
        assert abs(fd - grads[i]) <= 1e-6 + 1e-4 * abs(grads[i]), (i, fd, grads[i])
