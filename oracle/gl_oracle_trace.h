// ORACLE — test infrastructure only.  Imported by tests/, __graft_entry__.smoke()
// and bench.py's cpu_baseline leg as the CHECKER; never linked into, called by
// or shipped with the ghostlight_amd product library.
//
// gl_oracle_trace.h — CPU restatement of the reference's sequential ghost
// tracer, templated on the scalar type:
//   S = float  : the bit-exact oracle for the device tracer (masks, statuses,
//                landing positions, weights up to libm-vs-device cos/asin/exp)
//   S = double : the FP64 image oracle.
//
// Followed line by line (reference paths relative to its repo root):
//   intersect_surface      ghostlight/src/trace_core.h:299-525
//   check_aperture         ghostlight/src/trace_core.h:133-170  (polygon test
//                          restated in half-plane form, see check_aperture)
//   refract_ray/reflect_ray ghostlight/src/trace_core.h:533-564
//   surface_weight         ghostlight/src/trace_core.h:635-783
//   surface_attenuator     ghostlight/src/trace_core.h:791-809
//   fresnel / coatings     ghostlight/src/fresnel.h:27-178
//   d_trace_ghost_ray      ghostlight/src/trace_cuda.h:195-382 (GPU semantics:
//                          per-lambda IOR slices, t > 0 at the sensor)
//   trace_ghost_ray (CPU)  ghostlight/src/trace.cpp:171-240 (t >= 0 via t_ge0)
//   d_trace_primary_ray    ghostlight/src/trace_cuda.h:111-172
#pragma once

#include "../include/ghostlight_amd.h"
#include "gl_oracle_blade.h"

#include <cmath>
#include <cstdint>
#include <cstring>

namespace orc {

template <class S> struct V3 {
    S x, y, z;
    V3() : x(0), y(0), z(0) {}
    V3(S a, S b, S c) : x(a), y(b), z(c) {}
    V3 operator+(const V3& o) const { return V3(x + o.x, y + o.y, z + o.z); }
    V3 operator-(const V3& o) const { return V3(x - o.x, y - o.y, z - o.z); }
    V3 operator*(S s) const { return V3(x * s, y * s, z * s); }
    V3 operator-() const { return V3(-x, -y, -z); }
};
template <class S> inline S dot(const V3<S>& a, const V3<S>& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class S> inline V3<S> normalized(const V3<S>& v) {
    S l = std::sqrt(v.x * v.x + v.y * v.y + v.z * v.z);
    return (l > S(1e-12f)) ? v * (S(1) / l) : V3<S>(0, 0, 0);
}

// Polygon data precomputed per surface in FP64 and rounded once: side-normal
// directions and the apothem (mm).  The product precomputes the same values
// with the same expressions.
struct PolyPre {
    int   n = 0;
    float c[GL_MAX_BLADES], s[GL_MAX_BLADES];
    float apothem = 0.f;
    double cd[GL_MAX_BLADES], sd[GL_MAX_BLADES];
    double apothem_d = 0.0;
    Blade blade;                 // authored blade shape (deformed polygon)
};

inline void poly_precompute(int n_blades, float rot_rad, float scale, PolyPre& p) {
    p.n = (n_blades >= 3 && n_blades <= GL_MAX_BLADES) ? n_blades : 0;
    if (!p.n) return;
    const double sector = 2.0 * M_PI / (double)p.n;
    for (int k = 0; k < p.n; ++k) {
        const double phi = (double)rot_rad + sector * ((double)k + 0.5);
        p.cd[k] = std::cos(phi); p.sd[k] = std::sin(phi);
        p.c[k] = (float)p.cd[k]; p.s[k] = (float)p.sd[k];
    }
    p.apothem_d = (double)scale * std::cos(M_PI / (double)p.n);
    p.apothem = (float)p.apothem_d;
}

template <class S> inline S poly_get_c(const PolyPre& p, int k) { return S(p.c[k]); }
template <> inline double poly_get_c<double>(const PolyPre& p, int k) { return p.cd[k]; }
template <class S> inline S poly_get_s(const PolyPre& p, int k) { return S(p.s[k]); }
template <> inline double poly_get_s<double>(const PolyPre& p, int k) { return p.sd[k]; }
template <class S> inline S poly_apothem(const PolyPre& p) { return S(p.apothem); }
template <> inline double poly_apothem<double>(const PolyPre& p) { return p.apothem_d; }

struct Lens {
    const gl_surface* s = nullptr;
    int n = 0;
    const float* arena = nullptr;
    PolyPre poly[GL_MAX_SURFACES];
    bool transformed[GL_MAX_SURFACES];
    float fp64_rmin = 0.f;      // <= 0: FP64 sphere path off
};

inline bool rot_is_identity(const float m[9]) {
    const float eps = 1e-6f;
    return (std::fabs(m[0] - 1.f) < eps && std::fabs(m[1]) < eps && std::fabs(m[2]) < eps &&
            std::fabs(m[3]) < eps && std::fabs(m[4] - 1.f) < eps && std::fabs(m[5]) < eps &&
            std::fabs(m[6]) < eps && std::fabs(m[7]) < eps && std::fabs(m[8] - 1.f) < eps);
}

inline void lens_init(Lens& L, const gl_surface* s, int n, const float* arena, float fp64_rmin) {
    L.s = s; L.n = n; L.arena = arena; L.fp64_rmin = fp64_rmin;
    for (int i = 0; i < n && i < GL_MAX_SURFACES; ++i) {
        L.transformed[i] = (s[i].decenter_x != 0.0f || s[i].decenter_y != 0.0f || !rot_is_identity(s[i].rot));
        L.poly[i] = PolyPre{};
        if (s[i].aperture_shape == GL_APERTURE_POLYGON) {
            poly_precompute(s[i].aperture_blades, s[i].aperture_rotation_rad, s[i].semi_aperture, L.poly[i]);
            L.poly[i].blade = orc_blade_of(s[i]);
        }
    }
}

// ---- dispersion (fresnel.h:27-63, optical_system.h:230-235) ---------------
inline float dispersion_ior(float n_d, float V_d, float lambda_nm) {
    if (V_d < 0.1f || n_d <= 1.0001f) return n_d;
    const float lF = 486.13f, lC = 656.27f, ld = 587.56f;
    float dn = (n_d - 1.0f) / V_d;
    float inv_lF2 = 1.0f / (lF * lF);
    float inv_lC2 = 1.0f / (lC * lC);
    float inv_ld2 = 1.0f / (ld * ld);
    float B = dn / (inv_lF2 - inv_lC2);
    float A = n_d - B * inv_ld2;
    return A + B / (lambda_nm * lambda_nm);
}
inline float sellmeier_n(const float B[3], const float C[3], float lambda_nm) {
    float l = lambda_nm * 1e-3f;
    float l2 = l * l;
    float n2 = 1.0f;
    n2 += B[0] * l2 / (l2 - C[0]);
    n2 += B[1] * l2 / (l2 - C[1]);
    n2 += B[2] * l2 / (l2 - C[2]);
    return (n2 > 0.0f) ? std::sqrt(n2) : 1.0f;
}
inline float ior_at(const gl_surface& s, float lambda_nm) {
    if (s.disp_model == GL_DISP_SELLMEIER) return sellmeier_n(s.sellmeier_B, s.sellmeier_C, lambda_nm);
    return dispersion_ior(s.ior, s.abbe_v, lambda_nm);
}
inline void ior_tables(const gl_surface* s, int n, const gl_spec* spec, int ns, float* at, float* bf) {
    for (int si = 0; si < ns; ++si) {
        float last = 1.0f;
        for (int k = 0; k < n; ++k) {
            at[si * n + k] = ior_at(s[k], spec[si].lambda);
            bf[si * n + k] = last;
            if (s[k].is_active) last = at[si * n + k];
        }
    }
}

// ---- fresnel / coatings (fresnel.h:69-178) --------------------------------
template <class S> inline S fresnel_reflectance(S cos_i, S n1, S n2) {
    cos_i = std::fabs(cos_i);
    S eta = n1 / n2;
    S sin2_t = eta * eta * (S(1) - cos_i * cos_i);
    if (sin2_t >= S(1)) return S(1);
    S cos_t = std::sqrt(S(1) - sin2_t);
    S rs = (n1 * cos_i - n2 * cos_t) / (n1 * cos_i + n2 * cos_t);
    S rp = (n2 * cos_i - n1 * cos_t) / (n2 * cos_i + n1 * cos_t);
    return S(0.5f) * (rs * rs + rp * rp);
}
template <class S> inline S clamp01(S R) { return std::fmin(std::fmax(R, S(0)), S(1)); }
template <class S> inline S coating_reflectance(S cos_i, S n1, S n2, S cn, S d_nm, S lambda_nm) {
    S sin2_c = (n1 / cn) * (n1 / cn) * (S(1) - cos_i * cos_i);
    if (sin2_c >= S(1)) return fresnel_reflectance(cos_i, n1, n2);
    S cos_c = std::sqrt(S(1) - sin2_c);
    S delta = S(2) * S((float)M_PI) * cn * d_nm * cos_c / lambda_nm;
    S r01 = (n1 * cos_i - cn * cos_c) / (n1 * cos_i + cn * cos_c);
    S sin2_2 = (cn / n2) * (cn / n2) * (S(1) - cos_c * cos_c);
    if (sin2_2 >= S(1)) return fresnel_reflectance(cos_i, n1, n2);
    S cos_2 = std::sqrt(S(1) - sin2_2);
    S r12 = (cn * cos_c - n2 * cos_2) / (cn * cos_c + n2 * cos_2);
    S cos_2delta = std::cos(S(2) * delta);
    S num = r01 * r01 + r12 * r12 + S(2) * r01 * r12 * cos_2delta;
    S den = S(1) + r01 * r01 * r12 * r12 + S(2) * r01 * r12 * cos_2delta;
    return clamp01(num / den);
}
template <class S> inline S surface_reflectance(S cos_i, S n1, S n2, int layers, S lambda_nm) {
    if (layers <= 0) return fresnel_reflectance(cos_i, n1, n2);
    const S mgf2_n = S(1.38f);
    const S qw = S(550.0f) / (S(4) * mgf2_n);
    S R = coating_reflectance(cos_i, n1, n2, mgf2_n, qw, lambda_nm);
    for (int i = 1; i < layers; ++i) R *= S(0.25f);
    return clamp01(R);
}
template <class S> inline S artist_reflectance(S tr, S tg, S tb, S strength, S lambda_nm) {
    const S inv_2s2 = S(1.0f / (2.0f * 65.0f * 65.0f));
    S dr = lambda_nm - S(650), dg = lambda_nm - S(550), db = lambda_nm - S(450);
    S br = std::exp(-dr * dr * inv_2s2), bg = std::exp(-dg * dg * inv_2s2), bb = std::exp(-db * db * inv_2s2);
    S sum = br + bg + bb;
    return clamp01(strength * (tr * br + tg * bg + tb * bb) / sum);
}
template <class S> inline S coating_lerp_1d(const float* t, int n, S key, bool discard, bool* discarded) {
    if (!t || n <= 0) return S(0);
    if (key <= S(t[0])) { if (discard && key < S(t[0])) *discarded = true; return S(t[1]); }
    if (key >= S(t[2 * (n - 1)])) { if (discard && key > S(t[2 * (n - 1)])) *discarded = true; return S(t[2 * (n - 1) + 1]); }
    for (int i = 1; i < n; ++i) {
        if (key <= S(t[2 * i])) {
            S span = S(t[2 * i]) - S(t[2 * (i - 1)]);
            S f = (span > S(0)) ? (key - S(t[2 * (i - 1)])) / span : S(0);
            return S(t[2 * (i - 1) + 1]) + f * (S(t[2 * i + 1]) - S(t[2 * (i - 1) + 1]));
        }
    }
    return S(t[2 * (n - 1) + 1]);
}
template <class S> inline S coating_ref_angle_deg(S cos_i, S n1, S angle_ref_ior) {
    S sin_i = std::sqrt(std::fmax(S(0), S(1) - cos_i * cos_i));
    S s_ref = (angle_ref_ior > S(0)) ? (n1 * sin_i / angle_ref_ior) : sin_i;
    if (s_ref >= S(1)) return S(90);
    return std::asin(s_ref) * (S(180.0f) / S((float)M_PI));
}
template <class S> inline S surface_weight(const V3<S>& dir, const V3<S>& normal, S n1, S n2,
                                           const gl_surface& s, const float* arena, S lambda_nm, bool reflecting) {
    S cos_i = std::fabs(dot(normal, dir));
    S R;
    switch (s.coating_model) {
    case GL_COAT_SIMPLE: R = surface_reflectance(cos_i, n1, n2, s.ar_layers, lambda_nm); break;
    case GL_COAT_ARTIST: R = artist_reflectance(S(s.tint_r), S(s.tint_g), S(s.tint_b), S(s.tint_strength), lambda_nm); break;
    case GL_COAT_SPECTRAL: {
        if (s.table_count <= 0) { R = fresnel_reflectance(cos_i, n1, n2); break; }
        bool disc = false;
        R = coating_lerp_1d<S>(arena + s.table_offset, s.table_count, lambda_nm, s.out_of_range_discard != 0, &disc);
        if (disc) return S(0);
        break;
    }
    case GL_COAT_ANGULAR: {
        if (s.table_count <= 0) { R = fresnel_reflectance(cos_i, n1, n2); break; }
        S ang = coating_ref_angle_deg(cos_i, n1, S(s.angle_ref_ior));
        bool disc = false;
        R = coating_lerp_1d<S>(arena + s.table_offset, s.table_count, ang, s.out_of_range_discard != 0, &disc);
        if (disc) return S(0);
        break;
    }
    case GL_COAT_SPECTRAL_ANGULAR: {
        const int nw = s.sa_n_wavelengths, na = s.sa_n_angles;
        if (nw <= 0 || na <= 0) { R = fresnel_reflectance(cos_i, n1, n2); break; }
        const float* wl = arena + s.sa_offset;
        const float* an = wl + nw;
        const float* rr = an + na;
        S ang = coating_ref_angle_deg(cos_i, n1, S(s.angle_ref_ior));
        bool disc = false;
        const bool oob = s.out_of_range_discard != 0;
        int iw = 0; S fw = S(0);
        if (lambda_nm <= S(wl[0])) { if (oob && lambda_nm < S(wl[0])) disc = true; }
        else if (lambda_nm >= S(wl[nw - 1])) { iw = nw > 1 ? nw - 2 : 0; fw = nw > 1 ? S(1) : S(0); if (oob && lambda_nm > S(wl[nw - 1])) disc = true; }
        else for (int i = 1; i < nw; ++i) if (lambda_nm <= S(wl[i])) {
            iw = i - 1; S span = S(wl[i]) - S(wl[i - 1]);
            fw = span > S(0) ? (lambda_nm - S(wl[i - 1])) / span : S(0); break; }
        int ia = 0; S fa = S(0);
        if (ang <= S(an[0])) { if (oob && ang < S(an[0])) disc = true; }
        else if (ang >= S(an[na - 1])) { ia = na > 1 ? na - 2 : 0; fa = na > 1 ? S(1) : S(0); if (oob && ang > S(an[na - 1])) disc = true; }
        else for (int i = 1; i < na; ++i) if (ang <= S(an[i])) {
            ia = i - 1; S span = S(an[i]) - S(an[i - 1]);
            fa = span > S(0) ? (ang - S(an[i - 1])) / span : S(0); break; }
        if (disc) return S(0);
        const int iw1 = iw + 1 < nw ? iw + 1 : iw, ia1 = ia + 1 < na ? ia + 1 : ia;
        S r00 = rr[iw * na + ia], r01 = rr[iw * na + ia1], r10 = rr[iw1 * na + ia], r11 = rr[iw1 * na + ia1];
        R = (S(1) - fw) * ((S(1) - fa) * r00 + fa * r01) + fw * ((S(1) - fa) * r10 + fa * r11);
        break;
    }
    default: R = fresnel_reflectance(cos_i, n1, n2); break;
    }
    R = clamp01(R);
    return reflecting ? R : (S(1) - R);
}

// ---- frames (trace_core.h:33-65) ------------------------------------------
template <class S> inline V3<S> m_mul(const float m[9], const V3<S>& v) {
    return V3<S>(S(m[0]) * v.x + S(m[1]) * v.y + S(m[2]) * v.z,
                 S(m[3]) * v.x + S(m[4]) * v.y + S(m[5]) * v.z,
                 S(m[6]) * v.x + S(m[7]) * v.y + S(m[8]) * v.z);
}
template <class S> inline V3<S> m_tmul(const float m[9], const V3<S>& v) {
    return V3<S>(S(m[0]) * v.x + S(m[3]) * v.y + S(m[6]) * v.z,
                 S(m[1]) * v.x + S(m[4]) * v.y + S(m[7]) * v.z,
                 S(m[2]) * v.x + S(m[5]) * v.y + S(m[8]) * v.z);
}
template <class S> inline S surface_attenuator(const V3<S>& hit, const gl_surface& s, bool transformed) {
    if (s.coating_model != GL_COAT_ATTENUATOR_GAUSS) return S(1);
    if (!(s.gauss_sigma > 0.0f)) return S(s.gauss_background);
    V3<S> local = hit;
    if (transformed) {
        V3<S> vtx(S(s.decenter_x), S(s.decenter_y), S(s.z));
        local = m_tmul<S>(s.rot, hit - vtx) + V3<S>(0, 0, S(s.z));
    }
    S dx = local.x - S(s.gauss_decenter_x), dy = local.y - S(s.gauss_decenter_y);
    S r2 = dx * dx + dy * dy;
    S s2 = S(s.gauss_sigma) * S(s.gauss_sigma);
    return S(s.gauss_background) + S(s.gauss_peak) * std::exp(S(-0.5f) * r2 / s2);
}

// ---- aperture (trace_core.h:133-170) --------------------------------------
// Circle / ellipse exactly as the reference.  The regular polygon is tested as
// "inside every side half-plane": max_k (h . n_k) <= apothem, with the side
// normals n_k at rot + sector*(k + 1/2).  This equals the reference's
// r <= apothem / cos(offset) (offset = angle to the nearest side midpoint),
// since the sector holding the point maximises h . n_k; it needs no
// atan2/fmod/cos, so device and host evaluate the same IEEE operations.
template <class S> inline bool check_aperture(const V3<S>& hit, const gl_surface& s, const PolyPre& pp) {
    S hx = hit.x / S(s.aperture_aspect);
    S hy = hit.y;
    S h2 = hx * hx + hy * hy;
    S sa = S(s.semi_aperture);
    if (!(h2 <= sa * sa)) return false;
    if (s.aperture_shape == GL_APERTURE_POLYGON && pp.blade.deformed) {
        // authored blade shape: r <= sa * r(theta), as the reference
        const S r = std::sqrt(h2);
        if (r < S(1e-9f)) return true;
        return r <= sa * blade_radius<S>(pp.blade, std::atan2(hy, hx));
    }
    if (s.aperture_shape == GL_APERTURE_POLYGON && pp.n >= 3) {
        S m = hx * poly_get_c<S>(pp, 0) + hy * poly_get_s<S>(pp, 0);
        for (int k = 1; k < pp.n; ++k) {
            S d = hx * poly_get_c<S>(pp, k) + hy * poly_get_s<S>(pp, k);
            m = d > m ? d : m;
        }
        return m <= poly_apothem<S>(pp);
    }
    return true;
}

// ---- intersection (trace_core.h:299-525) ----------------------------------
template <class S> struct Ray { V3<S> o, d; };

// Ray-sphere in FP64 for large |R| (sag cancellation in c = |oc|^2 - R^2);
// returns t (rounded to S) or a negative sentinel when the ray misses.
template <class S> inline bool sphere_t_fp64(const Ray<S>& rc, double R, double zv, S& t_out) {
    const double ox = (double)rc.o.x, oy = (double)rc.o.y, oz = (double)rc.o.z - (zv + R);
    const double dx = (double)rc.d.x, dy = (double)rc.d.y, dz = (double)rc.d.z;
    const double a = dx * dx + dy * dy + dz * dz;
    const double b = 2.0 * (ox * dx + oy * dy + oz * dz);
    const double c = ox * ox + oy * oy + oz * oz - R * R;
    const double disc = b * b - 4.0 * a * c;
    if (disc < 0.0) return false;
    const double sd = std::sqrt(disc);
    const double inv_2a = 0.5 / a;
    const double t1 = (-b - sd) * inv_2a, t2 = (-b + sd) * inv_2a;
    double t;
    if (t1 > 1e-6 && t2 > 1e-6) {
        const double z1 = (double)rc.o.z + t1 * dz, z2 = (double)rc.o.z + t2 * dz;
        t = (std::fabs(z1 - zv) < std::fabs(z2 - zv)) ? t1 : t2;
    } else if (t1 > 1e-6) t = t1;
    else if (t2 > 1e-6) t = t2;
    else return false;
    t_out = (S)t;
    return true;
}

template <class S> inline S asphere_sag(S r, S inv_R, S K, const float* terms, int n_terms) {
    if (n_terms < 0) n_terms = 0;
    if (n_terms > GL_MAX_ASPHERE_TERMS) n_terms = GL_MAX_ASPHERE_TERMS;
    S r2 = r * r;
    S D = S(1) - (S(1) + K) * inv_R * inv_R * r2;
    if (D <= S(0)) return S(1e30f);
    S z = inv_R * r2 / (S(1) + std::sqrt(D));
    S rp = r2 * r2;
    for (int i = 0; i < n_terms; ++i) { z += S(terms[i]) * rp; rp *= r2; }
    return z;
}
template <class S> inline S asphere_dsag_dr(S r, S inv_R, S K, const float* terms, int n_terms) {
    if (n_terms < 0) n_terms = 0;
    if (n_terms > GL_MAX_ASPHERE_TERMS) n_terms = GL_MAX_ASPHERE_TERMS;
    S r2 = r * r;
    S D = S(1) - (S(1) + K) * inv_R * inv_R * r2;
    if (D <= S(0)) return S(0);
    S dz = inv_R * r / std::sqrt(D);
    S rp = r2 * r;
    for (int i = 0; i < n_terms; ++i) { dz += (S(4) + S(2) * S(i)) * S(terms[i]) * rp; rp *= r2; }
    return dz;
}

template <class S>
inline bool intersect_surface(const Ray<S>& r, const gl_surface& s, const PolyPre& pp, bool transformed,
                              float fp64_rmin, S& t_out, V3<S>& n_out) {
    Ray<S> rc = r;
    const S zv = S(s.z);
    if (transformed) {
        V3<S> V(S(s.decenter_x), S(s.decenter_y), zv);
        rc.o = m_tmul<S>(s.rot, r.o - V) + V3<S>(0, 0, zv);
        rc.d = m_tmul<S>(s.rot, r.d);
    }
    if (std::fabs(s.radius) < 1e-6f) {
        if (std::fabs(rc.d.z) < S(1e-12f)) return false;
        S t = (zv - rc.o.z) / rc.d.z;
        if (!(t > S(1e-6f))) return false;
        V3<S> hit = rc.o + rc.d * t;
        if (!check_aperture(hit, s, pp)) return false;
        t_out = t;
        n_out = V3<S>(0, 0, (rc.d.z > S(0)) ? S(-1) : S(1));
        if (transformed) n_out = m_mul<S>(s.rot, n_out);
        return true;
    } else if (s.form == GL_FORM_ASPHERE) {
        S R = S(s.radius), inv_R = S(1) / R, K = S(s.conic_k);
        V3<S> ctr(0, 0, zv + R);
        V3<S> oc = rc.o - ctr;
        S qa = dot(rc.d, rc.d), qb = S(2) * dot(oc, rc.d), qc = dot(oc, oc) - R * R;
        S disc = qb * qb - S(4) * qa * qc;
        S t;
        if (disc >= S(0)) {
            S sd = std::sqrt(disc), inv_2a = S(0.5f) / qa;
            S t1 = (-qb - sd) * inv_2a, t2 = (-qb + sd) * inv_2a;
            if (t1 > S(1e-6f) && t2 > S(1e-6f)) {
                S z1 = rc.o.z + t1 * rc.d.z, z2 = rc.o.z + t2 * rc.d.z;
                t = (std::fabs(z1 - zv) < std::fabs(z2 - zv)) ? t1 : t2;
            } else if (t1 > S(1e-6f)) t = t1;
            else if (t2 > S(1e-6f)) t = t2;
            else return false;
        } else {
            if (std::fabs(rc.d.z) < S(1e-12f)) return false;
            t = (zv - rc.o.z) / rc.d.z;
            if (!(t > S(1e-6f))) return false;
        }
        for (int iter = 0; iter < 15; ++iter) {
            V3<S> hi = rc.o + rc.d * t;
            S rx = hi.x, ry = hi.y, rad = std::sqrt(rx * rx + ry * ry);
            S sag = asphere_sag<S>(rad, inv_R, K, s.asphere_terms, s.n_asphere_terms);
            if (sag > S(1e29f)) return false;
            S f = hi.z - zv - sag;
            if (std::fabs(f) < S(1e-5f)) break;
            S dsag = asphere_dsag_dr<S>(rad, inv_R, K, s.asphere_terms, s.n_asphere_terms);
            S r_safe = (rad > S(1e-9f)) ? rad : S(1e-9f);
            S df_dt = rc.d.z - dsag * (rx * rc.d.x + ry * rc.d.y) / r_safe;
            if (std::fabs(df_dt) < S(1e-12f)) break;
            t -= f / df_dt;
        }
        if (!(t > S(1e-6f))) return false;
        V3<S> hit = rc.o + rc.d * t;
        S rad = std::sqrt(hit.x * hit.x + hit.y * hit.y);
        if (!check_aperture(hit, s, pp)) return false;
        S r_safe = (rad > S(1e-9f)) ? rad : S(1e-9f);
        S dsag = asphere_dsag_dr<S>(rad, inv_R, K, s.asphere_terms, s.n_asphere_terms);
        V3<S> nrm(-dsag * hit.x / r_safe, -dsag * hit.y / r_safe, S(1));
        S nlen = std::sqrt(nrm.x * nrm.x + nrm.y * nrm.y + nrm.z * nrm.z);
        nrm = nrm * (S(1) / nlen);
        if (dot(nrm, rc.d) > S(0)) nrm = -nrm;
        if (transformed) nrm = m_mul<S>(s.rot, nrm);
        t_out = t; n_out = nrm;
        return true;
    } else if (s.form == GL_FORM_CYLINDRICAL) {
        S vx, vy;
        if (s.cyl_axis == GL_CYL_AXIS_Y) { vx = S(1); vy = S(0); } else { vx = S(0); vy = S(1); }
        S R = S(s.radius);
        S oc_v = rc.o.x * vx + rc.o.y * vy;
        S oc_z = rc.o.z - zv - R;
        S dv = rc.d.x * vx + rc.d.y * vy;
        S qa = dv * dv + rc.d.z * rc.d.z;
        S qb = S(2) * (oc_v * dv + oc_z * rc.d.z);
        S qc = oc_v * oc_v + oc_z * oc_z - R * R;
        S disc = qb * qb - S(4) * qa * qc;
        if (disc < S(0)) return false;
        S sd = std::sqrt(disc), inv_2a = S(0.5f) / qa;
        S t1 = (-qb - sd) * inv_2a, t2 = (-qb + sd) * inv_2a, t;
        if (t1 > S(1e-6f) && t2 > S(1e-6f)) {
            S z1 = rc.o.z + t1 * rc.d.z, z2 = rc.o.z + t2 * rc.d.z;
            t = (std::fabs(z1 - zv) < std::fabs(z2 - zv)) ? t1 : t2;
        } else if (t1 > S(1e-6f)) t = t1;
        else if (t2 > S(1e-6f)) t = t2;
        else return false;
        V3<S> hit = rc.o + rc.d * t;
        if (!check_aperture(hit, s, pp)) return false;
        S hv = hit.x * vx + hit.y * vy;
        V3<S> nrm(hv * vx, hv * vy, hit.z - zv - R);
        S nlen = std::sqrt(nrm.x * nrm.x + nrm.y * nrm.y + nrm.z * nrm.z);
        if (nlen < S(1e-12f)) return false;
        nrm = nrm * (S(1) / nlen);
        if (dot(nrm, rc.d) > S(0)) nrm = -nrm;
        if (transformed) nrm = m_mul<S>(s.rot, nrm);
        t_out = t; n_out = nrm;
        return true;
    } else {
        S R = S(s.radius);
        V3<S> ctr(0, 0, zv + R);
        S t;
        if (sizeof(S) == 4 && fp64_rmin > 0.0f && std::fabs(s.radius) >= fp64_rmin) {
            if (!sphere_t_fp64<S>(rc, (double)s.radius, (double)s.z, t)) return false;
        } else {
            V3<S> oc = rc.o - ctr;
            S a = dot(rc.d, rc.d), b = S(2) * dot(oc, rc.d), c = dot(oc, oc) - R * R;
            S disc = b * b - S(4) * a * c;
            if (disc < S(0)) return false;
            S sd = std::sqrt(disc), inv_2a = S(0.5f) / a;
            S t1 = (-b - sd) * inv_2a, t2 = (-b + sd) * inv_2a;
            if (t1 > S(1e-6f) && t2 > S(1e-6f)) {
                S z1 = rc.o.z + t1 * rc.d.z, z2 = rc.o.z + t2 * rc.d.z;
                t = (std::fabs(z1 - zv) < std::fabs(z2 - zv)) ? t1 : t2;
            } else if (t1 > S(1e-6f)) t = t1;
            else if (t2 > S(1e-6f)) t = t2;
            else return false;
        }
        V3<S> hit = rc.o + rc.d * t;
        if (!check_aperture(hit, s, pp)) return false;
        V3<S> nrm = (hit - ctr) * (S(1) / std::fabs(R));
        if (dot(nrm, rc.d) > S(0)) nrm = -nrm;
        if (transformed) nrm = m_mul<S>(s.rot, nrm);
        t_out = t; n_out = nrm;
        return true;
    }
}

template <class S> inline bool refract_ray(const V3<S>& dir, const V3<S>& nrm, S n1, S n2, V3<S>& out) {
    S n_ratio = n1 / n2;
    S cos_i = -dot(nrm, dir);
    S sin2_t = n_ratio * n_ratio * (S(1) - cos_i * cos_i);
    if (sin2_t >= S(1)) return false;
    S cos_t = std::sqrt(S(1) - sin2_t);
    V3<S> tmp = dir * n_ratio + nrm * (n_ratio * cos_i - cos_t);
    S sq = tmp.x * tmp.x + tmp.y * tmp.y + tmp.z * tmp.z;
    if (sq < S(1e-18f)) return false;
    S inv = S(1) / std::sqrt(sq);
    out = tmp * inv;
    return true;
}
template <class S> inline V3<S> reflect_ray(const V3<S>& dir, const V3<S>& nrm) {
    V3<S> tmp = dir - nrm * (S(2) * dot(dir, nrm));
    return normalized(tmp);
}

template <class S> struct TraceOut { V3<S> pos; S weight; int status; int segments; };

// Bilinear bitmap sample between texel centres, clamped edges (tex2D, linear
// filter, normalised coordinates, clamp); stride > 1 reads one channel of
// interleaved texels.
inline float image_bilinear(const float* img, int W, int H, float u, float v, int stride = 1) {
    const float x = u * (float)W - 0.5f, y = v * (float)H - 0.5f;
    const float fx = std::floor(x), fy = std::floor(y);
    const float a = x - fx, b = y - fy;
    auto cl = [](int i, int n) { return i < 0 ? 0 : (i > n - 1 ? n - 1 : i); };
    const int i0 = cl((int)fx, W), i1 = cl((int)fx + 1, W), j0 = cl((int)fy, H), j1 = cl((int)fy + 1, H);
    const float top = (1.0f - a) * img[(j0 * W + i0) * stride] + a * img[(j0 * W + i1) * stride];
    const float bot = (1.0f - a) * img[(j1 * W + i0) * stride] + a * img[(j1 * W + i1) * stride];
    return (1.0f - b) * top + b * bot;
}

// ---- HURB edge diffraction (hurb.h:40-150) ---------------------------------
struct Hurb { int on = 0, dist = GL_HURB_LORENTZIAN; float min_sigma = 1.0e-5f, max_kick = 0.35f; };
inline uint32_t hurb_wang(uint32_t s) {
    s = (s ^ 61u) ^ (s >> 16); s *= 9u; s ^= s >> 4; s *= 0x27d4eb2du; s ^= s >> 15; return s;
}
inline float hurb_u(uint32_t& st) {
    st ^= st << 13; st ^= st >> 17; st ^= st << 5;
    return (float)(st & 0x00FFFFFFu) * (1.0f / 16777216.0f);
}
template <class S> inline S hurb_sigma(S lam_nm, S d, int dist) {
    if (d <= S(0)) return S(0);
    const S lam_mm = lam_nm * S(1.0e-6f);
    return lam_mm * S(dist == GL_HURB_LORENTZIAN ? 0.159154943f : 0.112539539f) / d;
}
template <class S> inline S hurb_kick(int dist, S sigma, S max_kick, uint32_t& st) {
    S k;
    const S pi = S(3.14159265358979323846f);
    if (dist == GL_HURB_LORENTZIAN) {
        const S u = S(hurb_u(st));
        k = sigma * std::tan(pi * (u - S(0.5f)));
    } else {
        const S u1 = S(std::fmax(hurb_u(st), 1.0e-7f));
        const S u2 = S(hurb_u(st));
        k = sigma * std::sqrt(S(-2.0f) * std::log(u1)) * std::cos(S(2.0f) * pi * u2);
    }
    if (k > max_kick) k = max_kick;
    if (k < -max_kick) k = -max_kick;
    return k;
}
template <class S> inline void hurb_apply(const Hurb& h, V3<S>& dir, const V3<S>& n, S d, S lam, uint32_t& st) {
    const S sigma = hurb_sigma<S>(lam, d, h.dist);
    if (sigma <= S(h.min_sigma)) return;
    const S theta = hurb_kick<S>(h.dist, sigma, S(h.max_kick), st);
    V3<S> t = n - dir * dot(n, dir);
    const S tl = std::sqrt(dot(t, t));
    if (tl < S(1.0e-8f)) return;
    t = t * (S(1) / tl);
    V3<S> nd = dir + t * std::tan(theta);
    const S nl = std::sqrt(dot(nd, nd));
    if (nl < S(1.0e-12f)) return;
    dir = nd * (S(1) / nl);
}

// aperture_edge_distance (trace_core.h:206-290).
//
// An image stop samples its baked distance field (arena[image_sdf_offset])
// where the trace samples the bitmap; other image surfaces never kick
// (d_aperture_edge_distance, trace_cuda.h:72-97).
template <class S> inline S edge_distance(const V3<S>& hit, const gl_surface& s, const PolyPre& pp, bool transformed,
                                          V3<S>& n_out, const float* arena = nullptr) {
    V3<S> h = hit;
    if (transformed) h = m_tmul<S>(s.rot, hit - V3<S>(S(s.decenter_x), S(s.decenter_y), S(s.z))) + V3<S>(0, 0, S(s.z));
    if (s.aperture_shape == GL_APERTURE_IMAGE) {
        if (!arena || !s.image_sdf || s.image_width <= 0 || s.image_height <= 0) {
            n_out = V3<S>(0, 0, 0);
            return S(1e30f);
        }
        const float hx = (float)h.x / s.aperture_aspect;
        const float u = 0.5f + hx / (2.0f * s.aperture_semi_diameter);
        const float v = 0.5f + (float)h.y / (2.0f * s.aperture_semi_diameter);
        const float* f = arena + s.image_sdf_offset;
        const float d = image_bilinear(f, s.image_width, s.image_height, u, v, 4);
        const V3<S> nl(S(image_bilinear(f + 1, s.image_width, s.image_height, u, v, 4)),
                       S(image_bilinear(f + 2, s.image_width, s.image_height, u, v, 4)), S(0));
        n_out = transformed ? m_mul<S>(s.rot, nl) : nl;
        return S(d > 0.0f ? d : 0.0f);
    }
    const S sa = S(s.semi_aperture);
    const S hx = h.x / S(s.aperture_aspect), hy = h.y;
    const S r_s = std::sqrt(hx * hx + hy * hy);
    S d;
    V3<S> n(0, 0, 0);
    if (s.aperture_shape == GL_APERTURE_POLYGON && s.aperture_blades >= 3) {
        const S sector = S(2.0f * (float)M_PI / (float)s.aperture_blades);
        const S apothem = sa * std::cos(S(0.5f) * sector);
        if (r_s < S(1e-9f)) {
            d = pp.blade.deformed ? sa * S(pp.blade.r_w) : apothem;
            n = V3<S>(1, 0, 0);
        } else if (pp.blade.deformed) {
            const S theta = std::atan2(hy, hx);
            S nx, ny;
            d = sa * blade_edge_gap<S>(pp.blade, r_s / sa, theta, nx, ny);
            n = V3<S>(nx, ny, 0);
        } else {
            const S theta = std::atan2(hy, hx);
            S off = std::fmod(theta - S(s.aperture_rotation_rad), sector);
            if (off < S(0)) off += sector;
            off -= S(0.5f) * sector;
            d = apothem - r_s * std::cos(off);
            const S phi = theta - off;
            n = V3<S>(std::cos(phi), std::sin(phi), 0);
        }
    } else {
        const S asp2 = S(s.aperture_aspect) * S(s.aperture_aspect);
        const S gx = h.x / asp2, gy = h.y;
        const S glen = std::sqrt(gx * gx + gy * gy);
        if (glen < S(1e-12f)) { d = sa; n = V3<S>(1, 0, 0); }
        else { d = (sa - r_s) * r_s / glen; n = V3<S>(gx / glen, gy / glen, 0); }
    }
    if (d < S(0)) d = S(0);
    n_out = transformed ? m_mul<S>(s.rot, n) : n;
    return d;
}

// ---- front-of-lens baffles (baffle.h:49-113) -------------------------------
inline bool baffles_block(const gl_baffle* b, int nb, float px, float py, float tx, float ty) {
    for (int k = 0; k < nb; ++k) {
        const float mx = px + tx * b[k].dz, my = py + ty * b[k].dz;
        if (b[k].shape == GL_BAFFLE_RECT) {
            if (mx > b[k].right || mx < -b[k].left || my > b[k].top || my < -b[k].bottom) return true;
        } else {
            const float dx = (mx - b[k].cx) / b[k].radius, dy = (my - b[k].cy) / (b[k].radius * b[k].aspect);
            if (dx * dx + dy * dy > 1.0f) return true;
        }
    }
    return false;
}
inline float baffles_edge(const gl_baffle* b, int nb, float px, float py, float tx, float ty, V3<float>& n_out) {
    float best = 1e30f;
    V3<float> bn(0, 0, 0);
    for (int k = 0; k < nb; ++k) {
        const float mx = px + tx * b[k].dz, my = py + ty * b[k].dz;
        float d; V3<float> n;
        if (b[k].shape == GL_BAFFLE_RECT) {
            d = b[k].right - mx; n = V3<float>(1, 0, 0);
            const float dl = mx + b[k].left, dt = b[k].top - my, db = my + b[k].bottom;
            if (dl < d) { d = dl; n = V3<float>(-1, 0, 0); }
            if (dt < d) { d = dt; n = V3<float>(0, 1, 0); }
            if (db < d) { d = db; n = V3<float>(0, -1, 0); }
        } else {
            const float rr = b[k].radius > 1e-12f ? b[k].radius : 1e-12f;
            const float ra = rr * (b[k].aspect > 0.0f ? b[k].aspect : 1.0f);
            const float ex = mx - b[k].cx, ey = my - b[k].cy;
            const float g = std::sqrt((ex / rr) * (ex / rr) + (ey / ra) * (ey / ra));
            const float gx = ex / (rr * rr), gy = ey / (ra * ra);
            const float gl = std::sqrt(gx * gx + gy * gy);
            if (gl < 1e-20f) { d = rr; n = V3<float>(1, 0, 0); }
            else { d = (1.0f - g) / gl; n = V3<float>(gx / gl, gy / gl, 0); }
        }
        if (d < best) { best = d; bn = n; }
    }
    if (best < 0.0f) best = 0.0f;
    n_out = bn;
    return best;
}

// d_image_aperture_passes (trace_cuda.h:37-54): bitmap transmission at the
// surface-frame hit, bilinear between texel centres with clamped edges (the
// reference's tex2D, linear filter, normalised coordinates, clamp), > 0.5.
// No bitmap loaded: every ray passes.
template <class S> inline bool image_aperture_passes(const Lens& L, int k, const V3<S>& hit) {
    const gl_surface& s = L.s[k];
    if (s.aperture_shape != GL_APERTURE_IMAGE || s.image_width <= 0 || s.image_height <= 0 || !L.arena) return true;
    if (!(s.aperture_aspect > 0.0f) || !(s.aperture_semi_diameter > 0.0f)) return false;
    V3<float> local((float)hit.x, (float)hit.y, (float)hit.z);
    if (L.transformed[k])
        local = m_tmul<float>(s.rot, local - V3<float>(s.decenter_x, s.decenter_y, s.z)) + V3<float>(0, 0, s.z);
    const float hx = local.x / s.aperture_aspect;
    const float u = 0.5f + hx / (2.0f * s.aperture_semi_diameter);
    const float v = 0.5f + local.y / (2.0f * s.aperture_semi_diameter);
    return image_bilinear(L.arena + s.image_offset, s.image_width, s.image_height, u, v) > 0.5f;
}

// One surface interaction; returns false when the ray died (status set).
template <class S>
inline bool step(const Lens& L, int k, Ray<S>& ray, S lambda, S n2, bool reflecting, S reflect_ior,
                 S& cur_ior, TraceOut<S>& res) {
    const gl_surface& s = L.s[k];
    S t; V3<S> nrm;
    ++res.segments;
    if (!intersect_surface<S>(ray, s, L.poly[k], L.transformed[k], L.fp64_rmin, t, nrm)) return false;
    V3<S> hit = ray.o + ray.d * t;
    ray.o = hit;
    if (!image_aperture_passes<S>(L, k, hit)) return false;
    const S n1 = cur_ior;
    if (reflecting) {
        res.weight *= surface_weight<S>(ray.d, nrm, n1, n2, s, L.arena, lambda, true);
        res.weight *= surface_attenuator<S>(hit, s, L.transformed[k]);
        ray.d = reflect_ray(ray.d, nrm);
        cur_ior = reflect_ior;
    } else {
        V3<S> nd;
        if (!refract_ray<S>(ray.d, nrm, n1, n2, nd)) { res.status = GL_TRACE_TIR; return false; }
        res.weight *= surface_weight<S>(ray.d, nrm, n1, n2, s, L.arena, lambda, false);
        res.weight *= surface_attenuator<S>(hit, s, L.transformed[k]);
        ray.d = nd;
        cur_ior = n2;
    }
    return true;
}

template <class S> inline void to_sensor(const Ray<S>& ray, bool t_ge0, TraceOut<S>& res) {
    if (std::fabs(ray.d.z) < S(1e-12f)) { res.status = GL_TRACE_MISSED_SURFACE; return; }
    S ts = -ray.o.z / ray.d.z;
    bool ok = t_ge0 ? (ts >= S(0)) : (ts > S(0));
    if (!ok) { res.status = GL_TRACE_MISSED_SURFACE; return; }
    res.pos = ray.o + ray.d * ts;
    res.status = GL_TRACE_OK;
}

// Three-leg ghost trace.  ior_at / ior_before: this lambda's slices.
// hurb (may be null): kick the new direction after every surface on the
// stream *hs (d_trace_ghost_ray<HURB>, trace_cuda.h:248-361).
template <class S>
inline TraceOut<S> trace_ghost(const Lens& L, int a, int b, Ray<S> ray, S lambda,
                               const float* ior_at, const float* ior_before, bool t_ge0,
                               const Hurb* hurb = nullptr, uint32_t* hs = nullptr) {
    TraceOut<S> res; res.status = GL_TRACE_VIGNETTED; res.weight = S(1); res.segments = 0;
    if (a < 0 || b <= a || b >= L.n) { res.status = GL_TRACE_INVALID_INPUT; res.weight = S(0); return res; }
    auto kick = [&](int k) {
        if (!hurb || !hurb->on) return;
        V3<S> en;
        const S ed = edge_distance<S>(ray.o, L.s[k], L.poly[k], L.transformed[k], en, L.arena);
        hurb_apply<S>(*hurb, ray.d, en, ed, lambda, *hs);
    };
    S cur = S(1);
    for (int k = 0; k <= b; ++k) {
        if (!L.s[k].is_active) continue;
        if (!step<S>(L, k, ray, lambda, S(ior_at[k]), k == b, cur, cur, res)) return res;
        kick(k);
    }
    for (int k = b - 1; k >= a; --k) {
        if (!L.s[k].is_active) continue;
        if (!step<S>(L, k, ray, lambda, S(ior_before[k]), k == a, S(ior_at[a]), cur, res)) return res;
        kick(k);
    }
    for (int k = a + 1; k < L.n; ++k) {
        if (!L.s[k].is_active) continue;
        if (!step<S>(L, k, ray, lambda, S(ior_at[k]), false, S(0), cur, res)) return res;
        kick(k);
    }
    to_sensor(ray, t_ge0, res);
    return res;
}

template <class S>
inline TraceOut<S> trace_primary(const Lens& L, Ray<S> ray, S lambda, const float* ior_at, bool t_ge0,
                                 V3<S>* out_dir = nullptr) {
    TraceOut<S> res; res.status = GL_TRACE_VIGNETTED; res.weight = S(1); res.segments = 0;
    S cur = S(1);
    for (int k = 0; k < L.n; ++k) {
        if (!L.s[k].is_active) continue;
        if (!step<S>(L, k, ray, lambda, S(ior_at[k]), false, S(0), cur, res)) return res;
    }
    to_sensor(ray, t_ge0, res);
    if (out_dir) *out_dir = ray.d;   // final direction (trace_cuda.h:170-171)
    return res;
}

}  // namespace orc
