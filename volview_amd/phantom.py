"""Seeded synthetic phantoms: wrappers of the C generator in libvvr.so
(csrc/phantom.cpp) and a numpy twin of it that reproduces its output bit for
bit (used by the tests to pin the generator).
"""
from __future__ import annotations

from ctypes import POINTER, c_int16, c_uint8, c_uint16, c_uint32

import numpy as np

from ._abi import check, load


def ct(dims, seed: int) -> np.ndarray:
    """int16 HU volume indexed [k, j, i]; dims = (nx, ny, nz)."""
    nx, ny, nz = (int(d) for d in dims)
    out = np.empty((nz, ny, nx), np.int16)
    d = (c_uint32 * 3)(nx, ny, nz)
    check(load().vvr_phantom_ct(d, seed, out.ctypes.data_as(POINTER(c_int16))))
    return out


def organs(dims, seed: int, n_organs: int = 40):
    """(labels uint8 [k,j,i], soft grid uint8 [k,j,i,4], fp16 margin [k,j,i], organs placed)."""
    nx, ny, nz = (int(d) for d in dims)
    labels = np.empty((nz, ny, nx), np.uint8)
    soft = np.empty((nz, ny, nx, 4), np.uint8)
    margin = np.empty((nz, ny, nx), np.uint16)
    d = (c_uint32 * 3)(nx, ny, nz)
    n = load().vvr_phantom_organs(d, seed, n_organs, labels.ctypes.data_as(POINTER(c_uint8)),
                                  soft.ctypes.data_as(POINTER(c_uint8)), margin.ctypes.data_as(POINTER(c_uint16)))
    if n < 0:
        check(n)
    return labels, soft, margin.view(np.float16), n


# ---- numpy twin ---------------------------------------------------------------------

M32 = np.uint32(0xFFFFFFFF)


def _lowbias32(x: np.ndarray) -> np.ndarray:
    x = np.atleast_1d(np.asarray(x, dtype=np.uint32)).copy()
    with np.errstate(over="ignore"):  # multiplication modulo 2^32
        x ^= x >> np.uint32(16)
        x *= np.uint32(0x7FEB352D)
        x ^= x >> np.uint32(15)
        x *= np.uint32(0x846CA68B)
        x ^= x >> np.uint32(16)
    return x


def _voxel_hash(seed: int, n: np.ndarray, k: int) -> np.ndarray:
    h = _lowbias32(np.uint32((seed >> 32) & 0xFFFFFFFF) ^ np.uint32(k))
    h = _lowbias32(np.uint32(seed & 0xFFFFFFFF) ^ h)
    h = _lowbias32((n >> np.uint64(32)).astype(np.uint32) ^ h)
    return _lowbias32((n & np.uint64(0xFFFFFFFF)).astype(np.uint32) ^ h)


def _noise_hu(seed: int, n: np.ndarray, sigma: int) -> np.ndarray:
    a = _voxel_hash(seed, n, 0).astype(np.int64)
    b = _voxel_hash(seed, n, 1).astype(np.int64)
    c = (a & 0xFFFF) + (a >> 16) + (b & 0xFFFF) + (b >> 16) - 131070
    mag = (2 * sigma * np.abs(c) + 37837) // 75674
    return np.where(c < 0, -mag, mag)


def ct_twin(dims, seed: int) -> np.ndarray:
    nx, ny, nz = (int(d) for d in dims)
    s = [(d - 1.0) / 255.0 for d in (nx, ny, nz)]
    smin = min(s)
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    n = (i.astype(np.uint64) + np.uint64(nx) * (j.astype(np.uint64) + np.uint64(ny) * k.astype(np.uint64)))
    p = [i - (nx - 1.0) * 0.5, j - (ny - 1.0) * 0.5, k - (nz - 1.0) * 0.5]

    def q(c, r):
        x = (p[0] - c[0]) / r[0]
        y = (p[1] - c[1]) / r[1]
        z = (p[2] - c[2]) / r[2]
        return x * x + y * y + z * z

    body = (100 * s[0], 75 * s[1], 115 * s[2])
    lung_r = (35 * s[0], 45 * s[1], 60 * s[2])
    rib_o = (92 * s[0], 69 * s[1], 106 * s[2])
    rib_i = (rib_o[0] - 4.0, rib_o[1] - 4.0, rib_o[2] - 4.0)
    v = np.full((nz, ny, nx), -1000, np.int64)
    m = q((0.0, 0.0, 0.0), body) <= 1.0
    v = np.where(m, 40 + _noise_hu(seed, n, 20), v)
    m = (q((-45 * s[0], 0.0, 25 * s[2]), lung_r) <= 1.0) | (q((45 * s[0], 0.0, 25 * s[2]), lung_r) <= 1.0)
    v = np.where(m, -800 + _noise_hu(seed, n, 50), v)
    dx, dy = p[0] - (-15 * s[0]), p[1] - 20 * s[1]
    m = (dx * dx + dy * dy <= 9.0) & (np.abs(p[2]) < 90 * s[2])
    v = np.where(m, 300 + _noise_hu(seed, n, 20), v)
    zr = p[2] + 60 * s[2]
    m = ((q((0.0, 0.0, 0.0), rib_o) <= 1.0) & (q((0.0, 0.0, 0.0), rib_i) > 1.0) & (p[2] >= -60 * s[2])
         & (p[2] <= 50 * s[2]) & (np.fmod(zr, 16.0) < 6.0))
    v = np.where(m, 1000 + _noise_hu(seed, n, 150), v)
    sz = p[2] - 70 * s[2]
    rr = np.sqrt(p[0] * p[0] + p[1] * p[1] + sz * sz)
    R = 40 * smin
    m = (rr <= R) & (rr > R - 4.0)
    v = np.where(m, 1000 + _noise_hu(seed, n, 150), v)
    return np.clip(v, -32768, 32767).astype(np.int16)


def _splitmix64(state):
    state = (state + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
    z = state
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
    return state, z ^ (z >> 31)


def organ_list(dims, seed: int, n_organs: int = 40):
    """The organ ellipsoids (centre, semi-axes) the generator places."""
    s = [(d - 1.0) / 255.0 for d in dims]
    smin = min(s)
    body = (100 * s[0], 75 * s[1], 115 * s[2])
    size = min(1.0, smin / 1.5)
    n_organs = min(n_organs, 249)
    st = (seed ^ 0x6F7267616E73) & 0xFFFFFFFFFFFFFFFF
    organs = []

    def uniform():
        nonlocal st
        st, z = _splitmix64(st)
        return (z >> 11) * (1.0 / 9007199254740992.0)

    for _ in range(4000):
        if len(organs) >= n_organs:
            break
        r = [(10.0 + 30.0 * uniform()) * size for _ in range(3)]
        c = [(2.0 * uniform() - 1.0) * 0.8 * body[a] for a in range(3)]
        R = max(r)
        fits, qb = True, 0.0
        for a in range(3):
            room = body[a] - R
            if room <= 0.0:
                fits = False
            else:
                qb += (c[a] / room) * (c[a] / room)
        if not fits or qb > 1.0:
            continue
        for (pc, pr) in organs:
            d2 = sum((c[a] - pc[a]) * (c[a] - pc[a]) for a in range(3))
            if np.sqrt(d2) <= R + max(pr) + 2.0:
                fits = False
                break
        if fits:
            organs.append((c, r))
    return organs


def organs_twin(dims, seed: int, n_organs: int = 40):
    nx, ny, nz = (int(d) for d in dims)
    organs = organ_list((nx, ny, nz), seed, n_organs)
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    p = [i - (nx - 1.0) * 0.5, j - (ny - 1.0) * 0.5, k - (nz - 1.0) * 0.5]
    far = 2.0 * 15.9
    best = np.zeros((nz, ny, nx), np.int64)
    second = np.zeros((nz, ny, nx), np.int64)
    sd = np.full((nz, ny, nx), far)
    done = np.zeros((nz, ny, nx), bool)
    for o, (c, r) in enumerate(organs):
        near = ~done
        for a in range(3):
            near &= np.abs(p[a] - c[a]) <= r[a] + far
        t0 = (p[0] - c[0]) / r[0]
        t1 = (p[1] - c[1]) / r[1]
        t2 = (p[2] - c[2]) / r[2]
        qq = t0 * t0 + t1 * t1 + t2 * t2
        d = (np.sqrt(qq) - 1.0) * min(r[0], min(r[1], r[2]))
        inside = near & (qq <= 1.0)
        closer = near & ~inside & (d < sd)
        best = np.where(inside, o + 1, best)
        second = np.where(inside, 0, np.where(closer, o + 1, second))
        sd = np.where(inside | closer, d, sd)
        done |= inside
    m = np.minimum(np.abs(sd) / 2.0, 15.9)
    soft = np.stack([best, second, np.floor(16.0 * m + 0.5), np.zeros_like(best)], -1).astype(np.uint8)
    margin = (sd / 2.0).astype(np.float32).astype(np.float16)
    return best.astype(np.uint8), soft, margin, len(organs)
