// blade_profile.cpp — the silhouette of a bladed stop with authored blade
// shape (ApertureProfile, aperture_profile.h:71-363).
//
// Host-only: built by the system C++ compiler so sinf / cosf / atan2f / fmodf
// are the C library's, as in the reference's CPU build, and every field of
// the profile is bit-identical to the one the reference derives (pinned by
// tests against its compiled CPU code).  The device evaluates the same
// r(theta) from these fields in gl_device.h.
//
// Geometry.  In the frame normalised to the blade-tip radius, each half of a
// blade edge is a conic arc with its focus at the centre, i.e. linear in
// u = 1/r:  u(phi) = u_w - (u_w - 1) sin^2((phi - phi_w)/2) / sin^2(d/2),
// with the waist (radius r_w) at phi_w inside the sector and d the angular
// span from the waist to the tip on that side.  Curvature 0 gives the regular
// polygon, curvature +1 the circle.  A notch replaces the tip by a straight
// chord whose end points stay on the two adjacent edges.
#include "../../include/ghostlight_amd.h"

#include <cmath>

namespace {

// Authored ranges and deformation constants (aperture_profile.h:47-61), each
// formed with the same float roundings as the reference's constants.
const float kNegDepth = 0.55f;                                // waist scoop, fraction of the apothem
const float kMinSpan = 4.0f * (float)(M_PI / 180.0);          // narrowest half edge left by twist
const float kNotchFull = 45.0f * (float)(M_PI / 180.0);       // notch that opens the full bite
const float kNotchMax = 0.45f;                                // bite as a fraction of the half sector
const float kNotchSector = 0.90f;                             // facet span cap, fraction of half sector
const int kAreaPanels = 512;                                  // Simpson panels over one sector
const float kNotchLimit = (float)(M_PI / 4.0);
const float kRakeLimit = (float)(M_PI / 4.0);

float clampf(float v, float lo, float hi) { return v < lo ? lo : (v > hi ? hi : v); }

// u = 1/r of the conic edge at sector-local angle phi.
float edge_u(const gl_blade_profile& b, float phi) {
    const float psi = phi - b.phi_w;
    const float sh = std::sin(0.5f * psi);
    const float q = (psi >= 0.0f) ? b.qR : b.qL;
    return b.u_w - b.u_rise * (sh * sh) * q;
}

// Angle folded into [-sigma/2, sigma/2).
float fold(const gl_blade_profile& b, float ang) {
    float a = std::fmod(ang + b.half, b.sigma);
    if (a < 0.0f) a += b.sigma;
    return a - b.half;
}

// u at a sector-local angle: the edge, or the tip facet across its own span
// when it cuts deeper (larger u = smaller r).
float u_at(const gl_blade_profile& b, float phi) {
    float u = edge_u(b, phi);
    if (b.facets) {
        const float w = fold(b, phi - b.half);
        if (w >= -b.eA && w <= b.eB) {
            const float uf = std::cos(w + b.half - b.beta) / b.p;
            if (uf > u) u = uf;
        }
    }
    return u;
}

float sector_angle(const gl_blade_profile& b, float theta) {
    float phi = std::fmod(theta - b.rotation, b.sigma);
    if (phi < 0.0f) phi += b.sigma;
    return phi - b.half;
}

float inv_sin2_half(float d) {
    const float s = std::sin(0.5f * d);
    return 1.0f / (s * s);
}

// Open area over the circumscribed disk: Simpson over one sector, summed in
// double (it scales brightness).
float area_fraction(const gl_blade_profile& b) {
    if (b.blades < 3) return 1.0f;
    const int n = (kAreaPanels % 2 == 0) ? kAreaPanels : kAreaPanels + 1;
    const double h = (double)b.sigma / n;
    double total = 0.0;
    for (int i = 0; i <= n; ++i) {
        const float phi = (float)(-(double)b.half + i * h);
        const double r = 1.0 / (double)u_at(b, phi);
        const double w = (i == 0 || i == n) ? 1.0 : ((i % 2) ? 4.0 : 2.0);
        total += w * r * r;
    }
    return (float)(b.blades * (total * h / 3.0) / (2.0 * M_PI));
}

gl_blade_profile unused_profile() {
    gl_blade_profile b{};
    b.plain = 1;
    b.r_w = 1.0f;
    b.u_w = 1.0f;
    b.area_frac = 1.0f;
    return b;
}

}  // namespace

extern "C" {

int gl_blade_profile_build(int32_t blades, float rotation_rad, float curvature_in, float twist_in, float notch_in,
                           float rake_in, gl_blade_profile* out) {
    if (!out) return GL_ERR_INVALID;
    gl_blade_profile b = unused_profile();
    if (blades < 3) { *out = b; return GL_OK; }
    const float curvature = clampf(curvature_in, -1.0f, 1.0f);
    const float twist = clampf(twist_in, -1.0f, 1.0f);
    const float notch = clampf(notch_in, -kNotchLimit, kNotchLimit);
    const float rake = clampf(rake_in, 0.0f, kRakeLimit);

    b.blades = blades;
    b.rotation = rotation_rad;
    b.sigma = 2.0f * (float)M_PI / (float)blades;
    b.half = 0.5f * b.sigma;
    // A skew without a notch changes nothing, so it does not make the blade deformed.
    b.plain = (curvature == 0.0f && twist == 0.0f && notch == 0.0f) ? 1 : 0;

    const float apothem = std::cos(b.half);
    b.r_w = (curvature >= 0.0f) ? (apothem + curvature * (1.0f - apothem))
                                : (apothem * (1.0f + curvature * kNegDepth));
    const float lim = std::fmax(0.0f, b.half - kMinSpan);
    b.phi_w = clampf(twist * b.half, -lim, lim);
    const float span_r = b.half - b.phi_w, span_l = b.half + b.phi_w;
    b.u_w = 1.0f / b.r_w;
    b.u_rise = b.u_w - 1.0f;
    b.qR = inv_sin2_half(span_r);
    b.qL = inv_sin2_half(span_l);

    // Tip facet: the chord between a point on this sector's edge (eA before
    // the tip) and one on the neighbour's edge (eB after it); the skew moves
    // the two ends in opposite directions.
    const float e = (std::fabs(notch) / kNotchFull) * kNotchMax * b.half;
    const float skew = clampf(rake / (float)(M_PI / 4.0), 0.0f, 1.0f);
    float eA = std::fmin(e * (1.0f + skew), kNotchSector * b.half);
    float eB = e * (1.0f - skew);
    if (notch < 0.0f) { const float t = eA; eA = eB; eB = t; }
    if (e > 1e-9f && b.r_w < 1.0f && (eA + eB) > 1e-9f) {
        const float pa = b.half - eA;
        const float ra = 1.0f / edge_u(b, pa);
        const float pb = b.half + eB;
        const float rb = 1.0f / edge_u(b, -b.half + eB);      // the neighbour's edge, in its own frame
        const float ax = ra * std::cos(pa), ay = ra * std::sin(pa);
        const float bx = rb * std::cos(pb), by = rb * std::sin(pb);
        const float dx = bx - ax, dy = by - ay;
        const float dlen = std::sqrt(dx * dx + dy * dy);
        if (dlen >= 1e-9f) {
            float nx = dy / dlen, ny = -dx / dlen;
            if (ax * nx + ay * ny < 0.0f) { nx = -nx; ny = -ny; }
            const float p = ax * nx + ay * ny;
            if (p > 1e-6f) {
                b.facets = 1;
                b.p = p;
                b.beta = std::atan2(ny, nx);
                b.eA = eA;
                b.eB = eB;
            }
        }
    }
    b.area_frac = area_fraction(b);
    *out = b;
    return GL_OK;
}

float gl_blade_radius_at(const gl_blade_profile* b, float theta) {
    if (!b || b->blades < 3) return 1.0f;
    return 1.0f / u_at(*b, sector_angle(*b, theta));
}

// dr/dtheta = -(du/dphi) / u^2 of whichever boundary u_at picked: the conic
// edge (du/dphi = -u_rise q sin(psi) / 2 from u = u_w - u_rise q sin^2(psi/2))
// or the tip facet (u = cos(w + half - beta) / p).
float gl_blade_dr_dtheta(const gl_blade_profile* b, float theta) {
    if (!b || b->blades < 3) return 0.0f;
    const float phi = sector_angle(*b, theta);
    const float psi = phi - b->phi_w;
    const float q = (psi >= 0.0f) ? b->qR : b->qL;
    float u = edge_u(*b, phi);
    float du = -0.5f * b->u_rise * q * std::sin(psi);
    if (b->facets) {
        const float w = fold(*b, phi - b->half);
        if (w >= -b->eA && w <= b->eB) {
            const float ang = w + b->half - b->beta;
            const float uf = std::cos(ang) / b->p;
            if (uf > u) {
                u = uf;
                du = -std::sin(ang) / b->p;
            }
        }
    }
    return -du / (u * u);
}

}  // extern "C"
