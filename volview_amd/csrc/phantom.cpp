// Seeded synthetic CT phantoms and organ label volumes: the workload
// generator of the benchmarks and tests. Every output is reproducible bit for
// bit from (dims, seed): shapes are tested in f64 and the noise is integer
// (sum of four 16-bit hash values, an Irwin-Hall approximation of a Gaussian),
// so the numpy twin in volview_amd/phantom.py produces identical volumes.
//
// Geometry, with s_a = (dims_a - 1) / 255 so a 256^3 volume has the sizes
// below in mm (1 mm voxels, centred on the origin, identity direction):
//   air -1000 HU; body ellipsoid (100, 75, 115) s, 40 +- 20 HU; lungs centred
//   (+-45 s_x, 0, 25 s_z), semi-axes (35, 45, 60) s, -800 +- 50 HU; a 6 mm
//   contrast vessel along z at (-15 s_x, 20 s_y), |z| < 90 s_z, 300 +- 20 HU;
//   4 mm rib bands on a (92, 69, 106) s shell, z in [-60, 50] s_z, 6 mm of
//   every 16 mm; a 4 mm skull shell, sphere of radius 40 min(s) at (0, 0, 70 s_z);
//   bone 1000 +- 150 HU. Later structures override earlier ones.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

#include "host.hpp"

namespace vvr {

namespace {

inline uint32_t lowbias32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

inline uint32_t voxel_hash(uint64_t seed, uint64_t n, uint32_t k) {
    uint32_t h = lowbias32(uint32_t(seed >> 32) ^ k);
    h = lowbias32(uint32_t(seed) ^ h);
    h = lowbias32(uint32_t(n >> 32) ^ h);
    return lowbias32(uint32_t(n) ^ h);
}

// Integer noise of standard deviation `sigma` HU, rounded half away from zero.
inline int32_t noise_hu(uint64_t seed, uint64_t n, int32_t sigma) {
    const uint32_t a = voxel_hash(seed, n, 0u), b = voxel_hash(seed, n, 1u);
    const int64_t c = int64_t(a & 0xffffu) + int64_t(a >> 16) + int64_t(b & 0xffffu) + int64_t(b >> 16) - 131070;
    const int64_t mag = (2 * int64_t(sigma) * (c < 0 ? -c : c) + 37837) / 75674;  // 37837 ~ sd of c
    return int32_t(c < 0 ? -mag : mag);
}

inline uint64_t splitmix64(uint64_t& s) {
    uint64_t z = (s += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}
// Uniform double in [0, 1) from the top 53 bits.
inline double uniform01(uint64_t& s) { return double(splitmix64(s) >> 11) * (1.0 / 9007199254740992.0); }

template <class F> void parallel_slices(uint32_t nz, F f) {
    unsigned nt = std::max(1u, std::min(32u, std::thread::hardware_concurrency()));
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; t++)
        th.emplace_back([=] {
            for (uint32_t k = t; k < nz; k += nt) f(k);
        });
    for (auto& x : th) x.join();
}

struct Organ {
    double c[3], r[3];
};

}  // namespace

void phantom_ct(const uint32_t dims[3], uint64_t seed, int16_t* hu) {
    const uint32_t nx = dims[0], ny = dims[1], nz = dims[2];
    double s[3];
    for (int a = 0; a < 3; a++) s[a] = (double(dims[a]) - 1.0) / 255.0;
    const double smin = std::min(s[0], std::min(s[1], s[2]));
    const double body[3] = {100 * s[0], 75 * s[1], 115 * s[2]};
    const double lung_c[2][3] = {{-45 * s[0], 0.0, 25 * s[2]}, {45 * s[0], 0.0, 25 * s[2]}};
    const double lung_r[3] = {35 * s[0], 45 * s[1], 60 * s[2]};
    const double rib_o[3] = {92 * s[0], 69 * s[1], 106 * s[2]};
    const double rib_i[3] = {rib_o[0] - 4.0, rib_o[1] - 4.0, rib_o[2] - 4.0};
    const double skull_c = 70 * s[2], skull_r = 40 * smin;
    const double vessel[2] = {-15 * s[0], 20 * s[1]};
    auto q = [](const double p[3], const double c[3], const double r[3]) {
        const double x = (p[0] - c[0]) / r[0], y = (p[1] - c[1]) / r[1], z = (p[2] - c[2]) / r[2];
        return x * x + y * y + z * z;
    };
    const double zero[3] = {0.0, 0.0, 0.0};
    parallel_slices(nz, [&](uint32_t k) {
        for (uint32_t j = 0; j < ny; j++)
            for (uint32_t i = 0; i < nx; i++) {
                const uint64_t n = uint64_t(i) + uint64_t(nx) * (uint64_t(j) + uint64_t(ny) * k);
                const double p[3] = {double(i) - (double(nx) - 1.0) * 0.5, double(j) - (double(ny) - 1.0) * 0.5,
                                     double(k) - (double(nz) - 1.0) * 0.5};
                int32_t v = -1000;
                if (q(p, zero, body) <= 1.0) v = 40 + noise_hu(seed, n, 20);
                if (q(p, lung_c[0], lung_r) <= 1.0 || q(p, lung_c[1], lung_r) <= 1.0) v = -800 + noise_hu(seed, n, 50);
                const double dx = p[0] - vessel[0], dy = p[1] - vessel[1];
                if (dx * dx + dy * dy <= 9.0 && std::fabs(p[2]) < 90 * s[2]) v = 300 + noise_hu(seed, n, 20);
                const double zr = p[2] + 60 * s[2];
                if (q(p, zero, rib_o) <= 1.0 && q(p, zero, rib_i) > 1.0 && p[2] >= -60 * s[2] && p[2] <= 50 * s[2] &&
                    std::fmod(zr, 16.0) < 6.0)
                    v = 1000 + noise_hu(seed, n, 150);
                const double sz = p[2] - skull_c;
                const double rr = std::sqrt(p[0] * p[0] + p[1] * p[1] + sz * sz);
                if (rr <= skull_r && rr > skull_r - 4.0) v = 1000 + noise_hu(seed, n, 150);
                hu[n] = int16_t(std::min(std::max(v, -32768), 32767));
            }
    });
}

// Up to n seeded axis-aligned organ ellipsoids inside the body, at least 2 mm
// apart. Labels 1..n; soft grid on the volume grid (best label, label across
// the nearest boundary, round(16 min(|sd| / 2 mm, 15.9)), 0); fp16 margin
// sd / 2 mm (negative inside). The signed distance of an ellipsoid is taken as
// (sqrt(q) - 1) x its smallest semi-axis, q its quadratic form.
int phantom_organs(const uint32_t dims[3], uint64_t seed, uint32_t n_organs, uint8_t* labels, uint8_t* soft,
                   uint16_t* margin_f16) {
    const uint32_t nx = dims[0], ny = dims[1], nz = dims[2];
    double s[3];
    for (int a = 0; a < 3; a++) s[a] = (double(dims[a]) - 1.0) / 255.0;
    const double smin = std::min(s[0], std::min(s[1], s[2]));
    const double body[3] = {100 * s[0], 75 * s[1], 115 * s[2]};
    const double size = std::min(1.0, smin / 1.5);
    n_organs = std::min<uint32_t>(n_organs, 249);  // labels from 250 up are CT-derived in the reference

    uint64_t st = seed ^ 0x6f7267616e73ull;  // "organs"
    std::vector<Organ> organs;
    for (int attempt = 0; attempt < 4000 && organs.size() < n_organs; attempt++) {
        Organ o;
        for (int a = 0; a < 3; a++) o.r[a] = (10.0 + 30.0 * uniform01(st)) * size;
        for (int a = 0; a < 3; a++) o.c[a] = (2.0 * uniform01(st) - 1.0) * 0.8 * body[a];
        const double R = std::max(o.r[0], std::max(o.r[1], o.r[2]));
        double qb = 0.0;
        bool fits = true;
        for (int a = 0; a < 3; a++) {
            const double room = body[a] - R;
            if (room <= 0.0) fits = false;
            else qb += (o.c[a] / room) * (o.c[a] / room);
        }
        if (!fits || qb > 1.0) continue;
        for (const Organ& p : organs) {
            const double Rp = std::max(p.r[0], std::max(p.r[1], p.r[2]));
            double d2 = 0.0;
            for (int a = 0; a < 3; a++) d2 += (o.c[a] - p.c[a]) * (o.c[a] - p.c[a]);
            if (std::sqrt(d2) <= R + Rp + 2.0) {
                fits = false;
                break;
            }
        }
        if (fits) organs.push_back(o);
    }

    const double kFar = 2.0 * 15.9;  // margins saturate at 15.9 x 2 mm
    parallel_slices(nz, [&](uint32_t k) {
        for (uint32_t j = 0; j < ny; j++)
            for (uint32_t i = 0; i < nx; i++) {
                const uint64_t n = uint64_t(i) + uint64_t(nx) * (uint64_t(j) + uint64_t(ny) * k);
                const double p[3] = {double(i) - (double(nx) - 1.0) * 0.5, double(j) - (double(ny) - 1.0) * 0.5,
                                     double(k) - (double(nz) - 1.0) * 0.5};
                uint32_t best = 0, second = 0;
                double sd = kFar;  // signed distance to the nearest boundary (negative inside)
                for (size_t o = 0; o < organs.size(); o++) {
                    const Organ& g = organs[o];
                    bool near = true;
                    for (int a = 0; a < 3; a++)
                        if (std::fabs(p[a] - g.c[a]) > g.r[a] + kFar) near = false;
                    if (!near) continue;
                    double qq = 0.0;
                    for (int a = 0; a < 3; a++) {
                        const double t = (p[a] - g.c[a]) / g.r[a];
                        qq += t * t;
                    }
                    const double rmin = std::min(g.r[0], std::min(g.r[1], g.r[2]));
                    const double d = (std::sqrt(qq) - 1.0) * rmin;
                    if (qq <= 1.0) {
                        best = uint32_t(o + 1);
                        second = 0;
                        sd = d;
                        break;  // organs do not overlap
                    }
                    if (d < sd) {
                        sd = d;
                        second = uint32_t(o + 1);
                    }
                }
                if (labels) labels[n] = uint8_t(best);
                if (soft) {
                    const double m = std::min(std::fabs(sd) / 2.0, 15.9);
                    soft[4 * n + 0] = uint8_t(best);
                    soft[4 * n + 1] = uint8_t(second);
                    soft[4 * n + 2] = uint8_t(std::floor(16.0 * m + 0.5));
                    soft[4 * n + 3] = 0;
                }
                if (margin_f16) margin_f16[n] = f32_to_f16_bits(float(sd / 2.0));
            }
    });
    return int(organs.size());
}

}  // namespace vvr

extern "C" int vvr_phantom_ct(const uint32_t dims[3], uint64_t seed, int16_t* hu) {
    if (!dims || !hu || !dims[0] || !dims[1] || !dims[2]) return VVR_EINVAL;
    vvr::phantom_ct(dims, seed, hu);
    return VVR_OK;
}

extern "C" int vvr_phantom_organs(const uint32_t dims[3], uint64_t seed, uint32_t n_organs, uint8_t* labels,
                                  uint8_t* soft, uint16_t* margin_f16) {
    if (!dims || !dims[0] || !dims[1] || !dims[2]) return VVR_EINVAL;
    return vvr::phantom_organs(dims, seed, n_organs, labels, soft, margin_f16);
}
