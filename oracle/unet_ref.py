"""Test infrastructure only: the learned denoiser's U-Net in float64 numpy,
forward and backward by hand, as the checker of libvvr's bf16 MFMA kernels.

The network (no counterpart in the reference; libvvr's unet.cpp builds the
same one): levels l = 0 .. L-1 at resolution / 2^l with c_l channels,
  encoder  e_l = relu(conv(relu(conv(l == 0 ? x : avgpool2(e_{l-1})))))
  decoder  d_l = relu(conv(relu(conv(concat(up2(d_{l+1}), e_l))))), d_{L-1} = e_{L-1}
  output   conv(d_0): 3 channels, log1p of the demodulated irradiance.
3x3 convolutions with zero padding; weights W[cout][tap][cin] with tap =
3 (dy + 1) + (dx + 1); parameters flat, per layer W then b, layers in the order
encoder (2 per level), decoder (2 per level, deepest first), output.
Input: 16 channels from a path-traced frame (log1p irradiance demodulated by
the mean albedo floored at 0.02 as denoise.comp:37-39 does, the mean albedo,
the unit mean normal, the background share, zeros). Loss: mean L1 between the
output and log1p of the reference radiance demodulated by the same albedo.
"""
from __future__ import annotations

import numpy as np


def layer_shapes(channels):
    L = len(channels)
    shapes = []
    for l in range(L):
        shapes.append((16 if l == 0 else channels[l - 1], channels[l]))
        shapes.append((channels[l], channels[l]))
    for l in range(L - 2, -1, -1):
        shapes.append((channels[l + 1] + channels[l], channels[l]))
        shapes.append((channels[l], channels[l]))
    shapes.append((channels[0], 3))
    return shapes


def unflatten(params, channels):
    out, k = [], 0
    for cin, cout in layer_shapes(channels):
        w = params[k:k + cout * 9 * cin].reshape(cout, 9, cin)
        k += cout * 9 * cin
        b = params[k:k + cout]
        k += cout
        out.append((w, b))
    assert k == len(params)
    return out


def num_params(channels):
    return sum(co * 9 * ci + co for ci, co in layer_shapes(channels))


def conv(x, w, b):
    """x (n, h, w, cin), w (cout, 9, cin): zero-padded 3x3 convolution."""
    n, h, wd, _ = x.shape
    xp = np.pad(x, ((0, 0), (1, 1), (1, 1), (0, 0)))
    y = np.zeros((n, h, wd, w.shape[0]))
    for t in range(9):
        dy, dx = t // 3 - 1, t % 3 - 1
        y += xp[:, 1 + dy:1 + dy + h, 1 + dx:1 + dx + wd, :] @ w[:, t, :].T
    return y + b


def conv_back(x, w, gy):
    """Gradients of conv: (dx, dw, db)."""
    n, h, wd, _ = x.shape
    xp = np.pad(x, ((0, 0), (1, 1), (1, 1), (0, 0)))
    gxp = np.zeros_like(xp)
    dw = np.zeros_like(w)
    for t in range(9):
        dy, dx = t // 3 - 1, t % 3 - 1
        xs = xp[:, 1 + dy:1 + dy + h, 1 + dx:1 + dx + wd, :]
        dw[:, t, :] = np.einsum('nhwo,nhwi->oi', gy, xs)
        gxp[:, 1 + dy:1 + dy + h, 1 + dx:1 + dx + wd, :] += gy @ w[:, t, :]
    return gxp[:, 1:-1, 1:-1, :], dw, gy.sum(axis=(0, 1, 2))


def pool(x):
    n, h, w, c = x.shape
    return x.reshape(n, h // 2, 2, w // 2, 2, c).mean(axis=(2, 4))


def pool_back(g):
    return np.repeat(np.repeat(g, 2, axis=1), 2, axis=2) * 0.25


def up(x):
    return np.repeat(np.repeat(x, 2, axis=1), 2, axis=2)


def up_back(g):
    n, h, w, c = g.shape
    return g.reshape(n, h // 2, 2, w // 2, 2, c).sum(axis=(2, 4))


def prepare(accum, albedo, normal, samples):
    """(n, h, w, 4) images -> the 16-channel input and the demodulating albedo."""
    n = float(samples)
    alb = np.maximum(albedo[..., :3] / n, 0.02)
    x = np.zeros(accum.shape[:3] + (16,))
    x[..., 0:3] = np.log1p(np.maximum(accum[..., :3] / n, 0.0) / alb)
    x[..., 3:6] = alb
    g = normal[..., :3] / n
    ln = np.linalg.norm(g, axis=-1, keepdims=True)
    x[..., 6:9] = np.where(ln > 1e-3, g / np.where(ln > 1e-3, ln, 1.0), 0.0)
    x[..., 9] = accum[..., 3] / n
    return x, alb


def forward(params, channels, x):
    layers = unflatten(params, channels)
    L = len(channels)
    cache = {"x": x}
    e = []
    for l in range(L):
        inp = x if l == 0 else pool(e[l - 1])
        a = np.maximum(conv(inp, *layers[2 * l]), 0.0)
        o = np.maximum(conv(a, *layers[2 * l + 1]), 0.0)
        cache[("in", l)], cache[("ea", l)] = inp, a
        e.append(o)
    d = {L - 1: e[L - 1]}
    for i, l in enumerate(range(L - 2, -1, -1)):
        cat = np.concatenate([up(d[l + 1]), e[l]], axis=-1)
        a = np.maximum(conv(cat, *layers[2 * L + 2 * i]), 0.0)
        d[l] = np.maximum(conv(a, *layers[2 * L + 2 * i + 1]), 0.0)
        cache[("cat", l)], cache[("da", l)] = cat, a
    out = conv(d[0], *layers[-1])
    cache["e"], cache["d"] = e, d
    return out, cache


def loss_and_grads(params, channels, x, target_log):
    """Mean L1 loss and its gradient with respect to the flat parameters."""
    layers = unflatten(params, channels)
    L = len(channels)
    out, c = forward(params, channels, x)
    diff = out - target_log
    loss = np.abs(diff).mean()
    g_out = np.sign(diff) / diff.size
    grads = [None] * len(layers)
    gx, dw, db = conv_back(c["d"][0], layers[-1][0], g_out)
    grads[-1] = (dw, db)
    g_d = {0: gx * (c["d"][0] > 0)}
    g_skip, g_deep = {}, None
    for l in range(0, L - 1):
        i = L - 2 - l
        la, lb = 2 * L + 2 * i, 2 * L + 2 * i + 1
        gx, dw, db = conv_back(c[("da", l)], layers[lb][0], g_d[l])
        grads[lb] = (dw, db)
        g_da = gx * (c[("da", l)] > 0)
        gcat, dw, db = conv_back(c[("cat", l)], layers[la][0], g_da)
        grads[la] = (dw, db)
        cu = channels[l + 1]
        g_up, g_skip[l] = gcat[..., :cu], gcat[..., cu:]
        if l + 1 <= L - 2:
            g_d[l + 1] = up_back(g_up) * (c["d"][l + 1] > 0)
        else:
            g_deep = up_back(g_up)
    g_pool = None
    for l in range(L - 1, -1, -1):
        g = np.zeros_like(c["e"][l])
        if l <= L - 2:
            g += g_skip[l]
        if l == L - 1:
            g += g_deep
        if g_pool is not None:
            g += pool_back(g_pool)
        g = g * (c["e"][l] > 0)
        gx, dw, db = conv_back(c[("ea", l)], layers[2 * l + 1][0], g)
        grads[2 * l + 1] = (dw, db)
        g_a = gx * (c[("ea", l)] > 0)
        gx, dw, db = conv_back(c[("in", l)], layers[2 * l][0], g_a)
        grads[2 * l] = (dw, db)
        g_pool = gx
    flat = np.concatenate([np.concatenate([w.ravel(), b]) for w, b in grads])
    return loss, flat, out


def init_params(channels, seed=1):
    rng = np.random.default_rng(seed)
    parts = []
    for cin, cout in layer_shapes(channels):
        parts.append(rng.normal(0, np.sqrt(2.0 / (9 * cin)), cout * 9 * cin))
        parts.append(rng.normal(0, 0.05, cout))
    return np.concatenate(parts)
