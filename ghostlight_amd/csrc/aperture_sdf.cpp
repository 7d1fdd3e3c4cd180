// aperture_sdf.cpp — signed distance field of an IMAGE aperture stop for the
// HURB edge kick (bake_aperture_sdf, aperture_sdf.cpp:110-189 of the
// reference).  Host-only, built by the system C++ compiler.
//
// The bitmap is thresholded at 0.5 (the trace's pass test); the exact
// Euclidean distance of every texel centre to the nearest texel of the other
// class is the lower envelope of parabolas (Felzenszwalb & Huttenlocher,
// "Distance Transforms of Sampled Functions", 2012), one pass down the
// columns and one along the rows with the physical texel pitch of each axis.
// The signed distance (positive inside) is pulled in by half a texel along
// the central-difference normal so its zero level sits on the matte edge, and
// is stored with that unit normal as [d_mm, nx, ny, 0] per texel.
#include "../../include/ghostlight_amd.h"

#include <cmath>
#include <vector>

namespace {

// Squared distance standing in for "no seed on this line": large enough to be
// unreachable, small enough that a q^2 parabola term still survives float
// addition.
const float kFar = 1.0e10f;

// Lower envelope of the parabolas a (q - p)^2 + f[p] over p, sampled at every
// q; breakpoints in double.
void envelope_1d(const float* f, float* d, int n, float a, std::vector<int>& v, std::vector<double>& z) {
    int k = 0;
    v[0] = 0;
    z[0] = -1e300;
    z[1] = 1e300;
    for (int q = 1; q < n; ++q) {
        double s;
        for (;;) {
            const int p = v[k];
            s = (((double)f[q] + (double)a * (double)q * q) - ((double)f[p] + (double)a * (double)p * p)) /
                (2.0 * (double)a * (double)(q - p));
            if (s <= z[k]) --k;
            else break;
        }
        ++k;
        v[k] = q;
        z[k] = s;
        z[k + 1] = 1e300;
    }
    k = 0;
    for (int q = 0; q < n; ++q) {
        while (z[k + 1] < (double)q) ++k;
        const int p = v[k];
        const float dx = (float)(q - p);
        d[q] = a * dx * dx + f[p];
    }
}

// Squared physical distance transform of `g` (0 on seeds, kFar elsewhere) in
// place: columns (pitch sy), then rows (pitch sx).
void edt(std::vector<float>& g, int W, int H, float sx, float sy) {
    const int m = W > H ? W : H;
    std::vector<int> v(m);
    std::vector<double> z(m + 1);
    std::vector<float> col(H), out(m);
    for (int i = 0; i < W; ++i) {
        for (int j = 0; j < H; ++j) col[j] = g[(size_t)j * W + i];
        envelope_1d(col.data(), out.data(), H, sy * sy, v, z);
        for (int j = 0; j < H; ++j) g[(size_t)j * W + i] = out[j];
    }
    for (int j = 0; j < H; ++j) {
        float* row = &g[(size_t)j * W];
        envelope_1d(row, out.data(), W, sx * sx, v, z);
        for (int i = 0; i < W; ++i) row[i] = out[i];
    }
}

}  // namespace

extern "C" {

int gl_bake_aperture_sdf(const float* pixels, int32_t W, int32_t H, float semi_diameter_mm, float aspect,
                         float* out4) {
    if (!pixels || !out4 || W <= 0 || H <= 0 || !(semi_diameter_mm > 0.0f)) return GL_ERR_INVALID;
    if (!(aspect > 0.0f)) aspect = 1.0f;
    // mm per texel under the trace's UV map: x spans 2 aspect sd over W, y 2 sd over H
    const float sx = 2.0f * aspect * semi_diameter_mm / (float)W;
    const float sy = 2.0f * semi_diameter_mm / (float)H;
    const size_t n = (size_t)W * H;
    std::vector<float> to_blocked(n), to_pass(n);
    for (size_t p = 0; p < n; ++p) {
        const bool in = pixels[p] > 0.5f;
        to_blocked[p] = in ? kFar : 0.0f;
        to_pass[p] = in ? 0.0f : kFar;
    }
    edt(to_blocked, W, H, sx, sy);
    edt(to_pass, W, H, sx, sy);
    std::vector<float> sd(n);
    for (size_t p = 0; p < n; ++p) sd[p] = pixels[p] > 0.5f ? std::sqrt(to_blocked[p]) : -std::sqrt(to_pass[p]);
    auto at = [&](int i, int j) {
        i = i < 0 ? 0 : (i >= W ? W - 1 : i);
        j = j < 0 ? 0 : (j >= H ? H - 1 : j);
        return sd[(size_t)j * W + i];
    };
    for (int j = 0; j < H; ++j)
        for (int i = 0; i < W; ++i) {
            const float gx = (at(i + 1, j) - at(i - 1, j)) / (2.0f * sx);
            const float gy = (at(i, j + 1) - at(i, j - 1)) / (2.0f * sy);
            const float gl = std::sqrt(gx * gx + gy * gy);
            float nx = 1.0f, ny = 0.0f;
            if (!(gl < 1.0e-12f)) { nx = gx / gl; ny = gy / gl; }
            // half a texel along the normal: the matte edge lies between centres
            const float c = 0.5f * std::sqrt((nx * sx) * (nx * sx) + (ny * sy) * (ny * sy));
            const float v = sd[(size_t)j * W + i];
            float mag = std::fabs(v) - c;
            if (mag < 0.0f) mag = 0.0f;
            float* t = out4 + ((size_t)j * W + i) * 4;
            t[0] = v >= 0.0f ? mag : -mag;
            t[1] = nx;
            t[2] = ny;
            t[3] = 0.0f;
        }
    return GL_OK;
}

}  // extern "C"
