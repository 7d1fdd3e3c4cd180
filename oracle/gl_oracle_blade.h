// ORACLE — test infrastructure only (see gl_oracle.cpp).
//
// gl_oracle_blade.h — CPU restatement of the deformed blade silhouette of a
// bladed stop (ApertureProfile, ghostlight/src/aperture_profile.h):
//   orc_blade_build   make_aperture_profile  aperture_profile.h:285-363
//                     ap_area_fraction       aperture_profile.h:271-282
//   blade_radius      radius_at / u_raw / fold / sector_angle / edge_at
//                                            aperture_profile.h:116-181
// Float expressions keep the reference's operation order (the C library's
// sinf / cosf / atan2f / fmodf), so with S = float the silhouette is the
// reference's bit for bit; tests pin it against the reference's compiled CPU
// code and the blade_shape calibration manifest (pupil_area_frac).
#pragma once

#include "../include/ghostlight_amd.h"

#include <cmath>
#include <utility>

namespace orc {

struct Blade {
    bool deformed = false;       // blades >= 3 with curvature, twist or notch authored
    int blades = 0, facets = 0;
    float rotation = 0.f, sigma = 0.f, half = 0.f, r_w = 1.f, phi_w = 0.f, u_w = 1.f, u_rise = 0.f;
    float qR = 0.f, qL = 0.f, p = 0.f, beta = 0.f, eA = 0.f, eB = 0.f, area_frac = 1.f;
};

inline float blade_clamp(float v, float lo, float hi) { return v < lo ? lo : (v > hi ? hi : v); }

// 1/r on the conic half edge at sector angle phi.
template <class S> inline S blade_edge_u(const Blade& b, S phi) {
    const S psi = phi - S(b.phi_w);
    const S sh = std::sin(S(0.5f) * psi);
    const S q = (psi >= S(0)) ? S(b.qR) : S(b.qL);
    return S(b.u_w) - S(b.u_rise) * (sh * sh) * q;
}

template <class S> inline S blade_u(const Blade& b, S phi) {
    S u = blade_edge_u<S>(b, phi);
    if (b.facets) {
        S w = std::fmod((phi - S(b.half)) + S(b.half), S(b.sigma));
        if (w < S(0)) w += S(b.sigma);
        w -= S(b.half);
        if (w >= -S(b.eA) && w <= S(b.eB)) {
            const S uf = std::cos(w + S(b.half) - S(b.beta)) / S(b.p);
            if (uf > u) u = uf;
        }
    }
    return u;
}

// Normalised boundary radius at absolute angle theta.
template <class S> inline S blade_radius(const Blade& b, S theta) {
    if (b.blades < 3) return S(1);
    S phi = std::fmod(theta - S(b.rotation), S(b.sigma));
    if (phi < S(0)) phi += S(b.sigma);
    return S(1) / blade_u<S>(b, phi - S(b.half));
}

// ApertureProfile::edge_gap (aperture_profile.h:205-230): first-order
// distance from normalised (r, theta) to the conic edge, exact to a tip facet,
// the nearer wins, with the boundary normal.
template <class S> inline S blade_edge_gap(const Blade& b, S r, S theta, S& nx, S& ny) {
    S phi = std::fmod(theta - S(b.rotation), S(b.sigma));
    if (phi < S(0)) phi += S(b.sigma);
    phi -= S(b.half);
    const S ct = std::cos(theta), st = std::sin(theta);
    const S psi = phi - S(b.phi_w);
    const S sh = std::sin(S(0.5f) * psi);
    const S q = (psi >= S(0)) ? S(b.qR) : S(b.qL);
    const S u = S(b.u_w) - S(b.u_rise) * (sh * sh) * q;
    const S du = S(-0.5f) * S(b.u_rise) * q * std::sin(psi);
    const S k = (-du / (u * u)) / r;
    const S inv = S(1) / std::sqrt(S(1) + k * k);
    S best = (S(1) / u - r) * inv;
    nx = (ct + k * st) * inv;
    ny = (st - k * ct) * inv;
    if (b.facets) {
        S w = std::fmod((phi - S(b.half)) + S(b.half), S(b.sigma));
        if (w < S(0)) w += S(b.sigma);
        w -= S(b.half);
        const S ang = w + S(b.half) - S(b.beta);
        const S d = S(b.p) - r * std::cos(ang);
        if (d < best) { best = d; nx = std::cos(theta - ang); ny = std::sin(theta - ang); }
    }
    return best;
}

inline Blade orc_blade_build(int blades, float rot, float curvature_in, float twist_in, float notch_in,
                             float rake_in) {
    Blade b;
    if (blades < 3) return b;
    const float neg_depth = 0.55f, min_span = 4.0f * (float)(M_PI / 180.0);
    const float notch_full = 45.0f * (float)(M_PI / 180.0), notch_max = 0.45f, notch_sector = 0.90f;
    const float curvature = blade_clamp(curvature_in, -1.0f, 1.0f);
    const float twist = blade_clamp(twist_in, -1.0f, 1.0f);
    const float notch = blade_clamp(notch_in, -(float)(M_PI / 4.0), (float)(M_PI / 4.0));
    const float rake = blade_clamp(rake_in, 0.0f, (float)(M_PI / 4.0));
    b.blades = blades;
    b.rotation = rot;
    b.sigma = 2.0f * (float)M_PI / (float)blades;
    b.half = 0.5f * b.sigma;
    b.deformed = !(curvature == 0.0f && twist == 0.0f && notch == 0.0f);
    const float a = std::cos(b.half);
    b.r_w = curvature >= 0.0f ? a + curvature * (1.0f - a) : a * (1.0f + curvature * neg_depth);
    const float lim = std::fmax(0.0f, b.half - min_span);
    b.phi_w = blade_clamp(twist * b.half, -lim, lim);
    b.u_w = 1.0f / b.r_w;
    b.u_rise = b.u_w - 1.0f;
    const float sr = std::sin(0.5f * (b.half - b.phi_w)), sl = std::sin(0.5f * (b.half + b.phi_w));
    b.qR = 1.0f / (sr * sr);
    b.qL = 1.0f / (sl * sl);
    const float e = (std::fabs(notch) / notch_full) * notch_max * b.half;
    const float sk = blade_clamp(rake / (float)(M_PI / 4.0), 0.0f, 1.0f);
    float ea = std::fmin(e * (1.0f + sk), notch_sector * b.half), eb = e * (1.0f - sk);
    if (notch < 0.0f) std::swap(ea, eb);
    if (e > 1e-9f && b.r_w < 1.0f && (ea + eb) > 1e-9f) {
        const float pa = b.half - ea, pb = b.half + eb;
        const float ra = 1.0f / blade_edge_u<float>(b, pa);
        const float rb = 1.0f / blade_edge_u<float>(b, -b.half + eb);
        const float ax = ra * std::cos(pa), ay = ra * std::sin(pa);
        const float bx = rb * std::cos(pb), by = rb * std::sin(pb);
        const float dx = bx - ax, dy = by - ay, len = std::sqrt(dx * dx + dy * dy);
        if (len >= 1e-9f) {
            float nx = dy / len, ny = -dx / len;
            if (ax * nx + ay * ny < 0.0f) { nx = -nx; ny = -ny; }
            const float p = ax * nx + ay * ny;
            if (p > 1e-6f) { b.facets = 1; b.p = p; b.beta = std::atan2(ny, nx); b.eA = ea; b.eB = eb; }
        }
    }
    // Simpson over one sector (512 panels), in double.
    const int n = 512;
    const double h = (double)b.sigma / n;
    double tot = 0.0;
    for (int i = 0; i <= n; ++i) {
        const float phi = (float)(-(double)b.half + i * h);
        const double r = 1.0 / (double)blade_u<float>(b, phi);
        tot += ((i == 0 || i == n) ? 1.0 : ((i % 2) ? 4.0 : 2.0)) * r * r;
    }
    b.area_frac = (float)(b.blades * (tot * h / 3.0) / (2.0 * M_PI));
    return b;
}

inline Blade orc_blade_of(const gl_surface& g) {
    if (g.aperture_shape != GL_APERTURE_POLYGON) return Blade{};
    return orc_blade_build(g.aperture_blades, g.aperture_rotation_rad, g.aperture_curvature, g.aperture_twist,
                           g.aperture_notch_rad, g.aperture_notch_angle_rad);
}

}  // namespace orc
