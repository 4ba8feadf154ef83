// The light map of Environment::Map on the host: EnvironmentMap::prepare
// (environment.rs:56-180) and the binary16 conversion of upload_environment
// (renderer.rs:2224-2290). Sums run in double where the reference's do, in the
// reference's order, so the tables equal its tables.
#include <algorithm>
#include <cmath>

#include "host.hpp"

namespace vvr {

namespace {

constexpr float kPi32 = 3.14159265358979323846f;

float lum(float r, float g, float b) { return (0.2126f * r + 0.7152f * g) + 0.0722f * b; }

// environment.rs:44-54: unit direction (x right, y up, z ahead) of image
// coordinates (u, v).
void direction(float u, float v, float d[3]) {
    const float phi = (u - 0.5f) * 2.0f * kPi32;
    const float theta = v * kPi32;
    d[0] = std::sin(theta) * std::sin(phi);
    d[1] = std::cos(theta);
    d[2] = std::sin(theta) * std::cos(phi);
}

uint16_t half_clamped(float v) { return f32_to_f16_bits(std::min(v, 60000.0f)); }

}  // namespace

EnvTables prepare_environment(uint32_t width, uint32_t height, const float* rgb) {
    const size_t w = width, h = height, n = w * h;
    EnvTables t;
    t.width = width, t.height = height;
    std::vector<float> sin_theta(h);
    for (size_t y = 0; y < h; y++) sin_theta[y] = std::sin(kPi32 * (float(y) + 0.5f) / float(h));

    // Mean radiance over the sphere, texels weighted by their solid angle.
    double sum[3] = {0.0, 0.0, 0.0}, weight = 0.0;
    for (size_t y = 0; y < h; y++) {
        double row[3] = {0.0, 0.0, 0.0};
        for (size_t x = 0; x < w; x++)
            for (int c = 0; c < 3; c++) row[c] += double(rgb[3 * (y * w + x) + c]);
        for (int c = 0; c < 3; c++) sum[c] += row[c] * double(sin_theta[y]);
        weight += double(sin_theta[y]) * double(w);
    }
    const float mean = std::max(lum(float(sum[0] / weight), float(sum[1] / weight), float(sum[2] / weight)), 1e-6f);
    std::vector<float> radiance(4 * n);
    for (size_t i = 0; i < n; i++) {
        for (int c = 0; c < 3; c++) radiance[4 * i + c] = rgb[3 * i + c] / mean;
        radiance[4 * i + 3] = 1.0f;
    }

    // Importance sampling proportional to luminance x solid angle: per row the
    // conditional CDF and each texel's probability, in the extra column the
    // marginal CDF and the row's probability.
    std::vector<float> f(n);
    for (size_t i = 0; i < n; i++) f[i] = lum(rgb[3 * i], rgb[3 * i + 1], rgb[3 * i + 2]) * sin_theta[i / w];
    std::vector<double> row_sum(h, 0.0);
    double total = 0.0;
    for (size_t y = 0; y < h; y++) {
        for (size_t x = 0; x < w; x++) row_sum[y] += double(f[y * w + x]);
        total += row_sum[y];
    }
    total = std::max(total, 1e-12);
    t.table.assign(2 * (w + 1) * h, 0.0f);
    double marginal = 0.0;
    for (size_t y = 0; y < h; y++) {
        float* row = &t.table[2 * y * (w + 1)];
        double acc = 0.0;
        for (size_t x = 0; x < w; x++) {
            acc += double(f[y * w + x]);
            row[2 * x] = row_sum[y] > 0.0 ? float(acc / row_sum[y]) : float(x + 1) / float(w);
            row[2 * x + 1] = float(double(f[y * w + x]) / total);
        }
        row[2 * (w - 1)] = 1.0f;
        marginal += row_sum[y];
        row[2 * w] = float(marginal / total);
        row[2 * w + 1] = float(row_sum[y] / total);
    }
    t.table[2 * ((h - 1) * (w + 1) + w)] = 1.0f;

    // The brightest region's azimuth: the power-weighted mean direction of the
    // texels, squared luminance favouring the light sources.
    double px = 0.0, pz = 0.0;
    for (size_t y = 0; y < h; y++)
        for (size_t x = 0; x < w; x++) {
            const double l = double(f[y * w + x]);
            float d[3];
            direction((float(x) + 0.5f) / float(w), (float(y) + 0.5f) / float(h), d);
            px += l * l * double(d[0]);
            pz += l * l * double(d[2]);
        }
    t.peak_u = 0.5f + float(std::atan2(px, pz) / (2.0 * 3.14159265358979323846));

    // Diffuse response: the normalised radiance box-filtered to at most 128 x 64,
    // then convolved with a clamped cosine lobe at 32 x 16 directions.
    const size_t sw = std::min<size_t>(128, w), sh = std::max<size_t>(sw / 2, 1);
    std::vector<float> light(3 * sw * sh), dirs(3 * sw * sh);
    for (size_t i = 0; i < sw * sh; i++) {
        const size_t sx = i % sw, sy = i / sw;
        const size_t x0 = sx * w / sw, x1 = std::max((sx + 1) * w / sw, x0 + 1);
        const size_t y0 = sy * h / sh, y1 = std::max((sy + 1) * h / sh, y0 + 1);
        float c[3] = {0.0f, 0.0f, 0.0f};
        for (size_t y = y0; y < y1; y++)
            for (size_t x = x0; x < x1; x++)
                for (int k = 0; k < 3; k++) c[k] = c[k] + radiance[4 * (y * w + x) + k];
        const float cnt = float((y1 - y0) * (x1 - x0));
        const float u = (float(sx) + 0.5f) / float(sw), v = (float(sy) + 0.5f) / float(sh);
        const float solid_angle = (2.0f * kPi32 / float(sw)) * (kPi32 / float(sh)) * std::sin(kPi32 * v);
        for (int k = 0; k < 3; k++) light[3 * i + k] = c[k] / cnt * solid_angle;
        direction(u, v, &dirs[3 * i]);
    }
    constexpr size_t kDw = 32, kDh = 16;
    t.diffuse.resize(4 * kDw * kDh);
    for (size_t i = 0; i < kDw * kDh; i++) {
        float nrm[3];
        direction((float(i % kDw) + 0.5f) / float(kDw), (float(i / kDw) + 0.5f) / float(kDh), nrm);
        float c[3] = {0.0f, 0.0f, 0.0f};
        for (size_t j = 0; j < sw * sh; j++) {
            const float* d = &dirs[3 * j];
            const float cs = (nrm[0] * d[0] + nrm[1] * d[1]) + nrm[2] * d[2];
            if (cs > 0.0f)
                for (int k = 0; k < 3; k++) c[k] = c[k] + light[3 * j + k] * cs;
        }
        for (int k = 0; k < 3; k++) t.diffuse[4 * i + k] = half_clamped(c[k] / kPi32);
        t.diffuse[4 * i + 3] = half_clamped(1.0f);
    }
    t.radiance.resize(4 * n);
    for (size_t i = 0; i < 4 * n; i++) t.radiance[i] = half_clamped(radiance[i]);
    return t;
}

}  // namespace vvr
