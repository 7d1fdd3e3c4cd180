// gl_device.h — device-side prescription layout and the FP32 sequential
// ghost tracer for gfx950.
//
// Layout.  The prescription is wave-uniform: every lane of a workgroup traces
// the same ghost pair through the same surfaces, and the surface loop counter
// is uniform.  DevSurf records therefore sit in HBM behind a __restrict__ const
// pointer and are read with uniform indices, which the AMDGPU backend turns
// into scalar (SMEM) loads into SGPRs: no LDS traffic and no VGPRs spent on
// the prescription.  Per-lane state (origin, direction, weight, medium) is
// the only vector data, kept in registers.
//
// Arithmetic.  Every expression follows the reference's operation order
// (trace_core.h, trace_cuda.h, fresnel.h; see citations per function) and
// the whole file is compiled with -ffp-contract=off and IEEE division and
// square root, so statuses, hit/vignette masks and landing positions are
// bit-identical to the FP32 CPU oracle.  Arithmetic that feeds only a weight
// (Fresnel and coating amplitudes, the coating-table angle, cos/asin/exp)
// uses hardware reciprocal and square root estimates instead: weights move
// by a few ulp, never a mask.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/ghostlight_amd.h"

// Constant address space: loads through it with a wave-uniform address are
// selected as scalar (SMEM) loads.  Only read-only, wave-uniform data uses it.
#define GLC __attribute__((address_space(4)))

namespace gl {

template <class T> __host__ __device__ inline const T GLC* as_const(const T* p) { return (const T GLC*)p; }

enum : int32_t {
    SF_ACTIVE = 1, SF_TRANSFORMED = 2, SF_FLAT = 4, SF_POLYGON = 8, SF_FP64 = 16,
    SF_ASPECT = 32, SF_CYL_Y = 64,
    SF_TAB_SORTED = 128,      // 1-D coating table keys non-decreasing (binary search allowed)
    SF_SA_SORTED = 256,       // spectral-angular wavelength and angle axes non-decreasing
    SF_DEFORMED = 512,        // polygon with authored blade shape: test r <= sa * r(theta)
    SF_IMAGE = 1024,          // IMAGE aperture with a loaded bitmap (img_* below)
    SF_SA_UNIFORM = 2048,     // spectral-angular angle axis on an exact uniform grid (sa_a0, sa_inv_da)
};

// Blade silhouette r(theta) of a deformed polygon (gl_blade_profile as the
// device evaluates it, aperture_profile.h:116-181).
struct DevBlade {
    float rotation, sigma, half, phi_w, u_w, u_rise, qR, qL, p, beta, eA, eB, r_w;
    int32_t facets;
};

// HURB edge diffraction switches (GpuHurb, hurb.h:103-108).
struct HurbDev {
    int32_t on, dist;
    float min_sigma, max_kick;
};

// Front-of-lens baffle stack (GpuBaffleStack, baffle.h:43-46).
struct DevBaffles {
    int32_t n;
    gl_baffle b[GL_MAX_BAFFLES];
};

// One surface as the tracer reads it (built on the host from gl_surface).
struct DevSurf {
    float z, R, sa, aspect;          // vertex z, radius, semi-aperture, X stretch
    float dx, dy, inv_absR, conic;   // decenter, 1/|R| (float division), conic K
    float cz, R2, sa2, pad0;         // z + R, R * R, sa * sa: the same float operations, done once on the host
    float rot[9];
    int32_t flags, form, nblades, poly_off;   // poly_off: index into the normals array
    float apothem;                   // polygon apothem (mm), rounded from FP64
    int32_t coat_model, ar_layers, discard;
    float angle_ref_ior;
    int32_t tab_off, tab_n, sa_off, sa_nw;
    int32_t sa_na, n_asph;
    float sa_a0, sa_inv_da, sa_da;   // SF_SA_UNIFORM: first angle, 1 / spacing, spacing
    float tint_r, tint_g, tint_b, tint_strength;
    float g_sigma, g_bg, g_peak, g_dx, g_dy;
    float asph[GL_MAX_ASPHERE_TERMS];
    DevBlade blade;                  // SF_DEFORMED only
    float ap_rot;                    // polygon rotation (radians): HURB edge distance
    int32_t ap_blades;               // polygon blade count (0: round, -1: image aperture)
    int32_t img_off, img_w, img_h;   // SF_IMAGE: bitmap at arena[img_off], img_w x img_h
    float img_semi;                  // SF_IMAGE: world radius of the bitmap's edge (mm)
    int32_t sdf_off;                 // image stop: [d, nx, ny, 0] texels at arena[sdf_off]; -1 none
};

// One wavelength's tables for one surface, interleaved so that a trace with
// a lane-varying wavelength reaches all of them from one address per surface:
// the indices after (iat) and before (ibf) the surface, the refraction ratios
// forward (ibf / iat) and backward (iat / ibf), and the coating-row entry of
// lambda_rows, whose self-relative row offset counts from this lrow field.
struct __attribute__((aligned(16))) LamRec {
    float iat, ibf, rfw, rbw;
    int2 lrow;
    int2 pad;
};
static_assert(sizeof(LamRec) == 32, "LamRec is two 16-byte loads");

struct TraceRes {
    float px, py, pz, w;
    int32_t status, segs;
};

#define GLD __device__ __forceinline__

struct F3 { float x, y, z; };
GLD F3 f3(float x, float y, float z) { F3 r; r.x = x; r.y = y; r.z = z; return r; }
GLD F3 add(F3 a, F3 b) { return f3(a.x + b.x, a.y + b.y, a.z + b.z); }
GLD F3 sub(F3 a, F3 b) { return f3(a.x - b.x, a.y - b.y, a.z - b.z); }
GLD F3 mul(F3 a, float s) { return f3(a.x * s, a.y * s, a.z * s); }
GLD float dot3(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
GLD F3 neg(F3 a) { return f3(-a.x, -a.y, -a.z); }
GLD F3 normalized(F3 v) {  // vec3.h:67-71
    float l = sqrtf(v.x * v.x + v.y * v.y + v.z * v.z);
    return (l > 1e-12f) ? mul(v, 1.0f / l) : f3(0.f, 0.f, 0.f);
}
GLD float sqrt_rn_big(float x);
// normalized() for a vector of length near 1 (a reflected unit direction):
// |v|^2 stays far above the small-subnormal range, so sqrt_rn_big gives
// sqrtf's bits, and rcp_rn the IEEE reciprocal of the length.
GLD float rcp_rn(float x);
GLD F3 normalized_unit(F3 v) {
    float l = sqrt_rn_big(v.x * v.x + v.y * v.y + v.z * v.z);
    return (l > 1e-12f) ? mul(v, rcp_rn(l)) : f3(0.f, 0.f, 0.f);
}
template <class P> GLD F3 m_mul(P m, F3 v) {   // trace_core.h:33-39
    return f3(m[0] * v.x + m[1] * v.y + m[2] * v.z, m[3] * v.x + m[4] * v.y + m[5] * v.z,
              m[6] * v.x + m[7] * v.y + m[8] * v.z);
}
template <class P> GLD F3 m_tmul(P m, F3 v) {  // trace_core.h:41-47
    return f3(m[0] * v.x + m[3] * v.y + m[6] * v.z, m[1] * v.x + m[4] * v.y + m[7] * v.z,
              m[2] * v.x + m[5] * v.y + m[8] * v.z);
}

// ---- Fresnel and coatings (fresnel.h:69-178) -----------------------------
// Correctly rounded square root for arguments that are 0, at least 2^-96,
// +inf or NaN.  The hardware estimate (v_sqrt_f32, within 1 ulp) is moved to
// the neighbour whose residual x - s * s' has the right sign: the IEEE square
// root's own rounding step without its input scaling, which is the identity
// on this range, so the bits equal sqrtf's.  Used where the argument is
// bounded away from the small-subnormal range by construction.
GLD float sqrt_rn_big(float x) {
    const float s = __builtin_amdgcn_sqrtf(x);
    const float sm = __int_as_float(__float_as_int(s) - 1), sp = __int_as_float(__float_as_int(s) + 1);
    const float rm = __builtin_fmaf(-sm, s, x), rp = __builtin_fmaf(-sp, s, x);
    const float r = rm <= 0.0f ? sm : s;
    return rp > 0.0f ? sp : r;
}
// Correctly rounded square root for x >= 0 (any finite or infinite value,
// subnormals included; NaN gives NaN): arguments below 2^-96 are scaled by
// 2^32 before sqrt_rn_big and the root by 2^-16 after -- exact power-of-two
// steps -- as the IEEE square root's own expansion does.  Equal to sqrtf in
// every bit on [0, +inf] (checked over every float of it by gl_selftest_rn);
// sqrtf's separate handling of +-0 and +inf is not needed, sqrt_rn_big
// returns them.
GLD float sqrt_rn_pos(float x) {
    const bool tiny = x < 0x1p-96f;
    const float r = sqrt_rn_big(tiny ? x * 0x1p32f : x);
    return tiny ? r * 0x1p-16f : r;
}
// Correctly rounded reciprocal for normal x whose reciprocal is normal
// (2^-126 <= |x| <= 2^126): the hardware estimate (v_rcp_f32, within 1 ulp)
// refined by one Newton step whose residual 1 - x * y the fused
// multiply-add forms exactly.  Equal to 1.0f / x in every bit on that range
// (checked over every float of it by gl_selftest_rn).
GLD float rcp_rn(float x) {
    const float y = __builtin_amdgcn_rcpf(x);
    const float e = __builtin_fmaf(-x, y, 1.0f);
    return __builtin_fmaf(e, y, y);
}
// Weight-only quotient: reciprocal estimate (1 ulp) times the numerator.
GLD float wdiv(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }

// eta = n1 / n2, IEEE-rounded (the refraction's own ratio), so the total
// reflection test sees the same bits as a division here.  cos_t is exact
// (1 - sin2_t >= 2^-24), so the amplitudes' numerators carry no extra error
// into the cancellation of nearly matched indices; only the two quotients use
// the reciprocal estimate (a few ulp on the weight).
// Total reflection (sin2_t >= 1) gives 1; the amplitudes are formed for
// every lane and the result selected, so the lanes of a wave do not branch.
GLD float fresnel_R(float cos_i, float n1, float n2, float eta) {
    cos_i = fabsf(cos_i);
    float sin2_t = eta * eta * (1.0f - cos_i * cos_i);
    float cos_t = sqrt_rn_big(1.0f - sin2_t);
    float rs = wdiv(n1 * cos_i - n2 * cos_t, n1 * cos_i + n2 * cos_t);
    float rp = wdiv(n2 * cos_i - n1 * cos_t, n2 * cos_i + n1 * cos_t);
    return (sin2_t >= 1.0f) ? 1.0f : 0.5f * (rs * rs + rp * rp);
}
GLD float clamp01(float r) { return fminf(fmaxf(r, 0.0f), 1.0f); }
// Single-layer (Airy) coating.  The interference sum cancels for an AR
// layer, so the phase and the two amplitudes keep IEEE quotients; the square
// roots are exact (arguments >= 2^-24) and only the final ratio, whose
// numerator is already formed, uses the reciprocal estimate.
GLD float coating_R(float cos_i, float n1, float n2, float cn, float d_nm, float lam, float eta) {
    float sin2_c = (n1 / cn) * (n1 / cn) * (1.0f - cos_i * cos_i);
    if (sin2_c >= 1.0f) return fresnel_R(cos_i, n1, n2, eta);
    float cos_c = sqrt_rn_big(1.0f - sin2_c);
    float delta = 2.0f * 3.14159265358979323846f * cn * d_nm * cos_c / lam;
    float r01 = (n1 * cos_i - cn * cos_c) / (n1 * cos_i + cn * cos_c);
    float sin2_2 = (cn / n2) * (cn / n2) * (1.0f - cos_c * cos_c);
    if (sin2_2 >= 1.0f) return fresnel_R(cos_i, n1, n2, eta);
    float cos_2 = sqrt_rn_big(1.0f - sin2_2);
    float r12 = (cn * cos_c - n2 * cos_2) / (cn * cos_c + n2 * cos_2);
    float c2d = cosf(2.0f * delta);
    float num = r01 * r01 + r12 * r12 + 2.0f * r01 * r12 * c2d;
    float den = 1.0f + r01 * r01 * r12 * r12 + 2.0f * r01 * r12 * c2d;
    return clamp01(wdiv(num, den));
}
// The coating-table angle feeds only the weight (never a hit, vignette or
// segment decision), so it is computed with the hardware square root and
// reciprocal (1 ulp) and asin from Abramowitz & Stegun 4.4.46
// (asin x = pi/2 - sqrt(1 - x) * sum a_k x^k, |error| <= 2e-8 rad on [0, 1]):
// a tenth of the instructions of the correctly rounded forms, with weights
// within ~1e-6 relative of the reference's.
GLD float ref_angle_deg(float cos_i, float n1, float aref) {   // trace_core.h:614-621
    const float sin_i = __builtin_amdgcn_sqrtf(fmaxf(0.0f, 1.0f - cos_i * cos_i));
    const float s = (aref > 0.0f) ? (n1 * sin_i * __builtin_amdgcn_rcpf(aref)) : sin_i;
    float q = __builtin_fmaf(s, -0.0012624911f, 0.0066700901f);
    q = __builtin_fmaf(s, q, -0.0170881256f);
    q = __builtin_fmaf(s, q, 0.0308918810f);
    q = __builtin_fmaf(s, q, -0.0501743046f);
    q = __builtin_fmaf(s, q, 0.0889789874f);
    q = __builtin_fmaf(s, q, -0.2145988016f);
    q = __builtin_fmaf(s, q, 1.5707963050f);
    const float ang = __builtin_fmaf(-__builtin_amdgcn_sqrtf(1.0f - s), q, 1.57079632679489662f) *
                      (180.0f / 3.14159265358979323846f);
    return (s >= 1.0f) ? 90.0f : ang;      // selected, not branched
}
GLD int2 ld_i2(const int2 GLC* p) {
    const int GLC* q = reinterpret_cast<const int GLC*>(p);
    return make_int2(q[0], q[1]);
}

// First i in [lo, hi] with key <= v[i * stride], for v non-decreasing and
// key <= v[hi * stride]: the same cell the reference's forward scan stops at.
GLD int lower_cell(const float* v, int stride, int lo, int hi, float key) {
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (key <= v[mid * stride]) hi = mid; else lo = mid + 1;
    }
    return lo;
}
GLD float lerp1d(const float* t, int n, float key, bool disc, bool* out_disc, bool sorted) {  // trace_core.h:575-602
    if (n <= 0) return 0.0f;
    if (key <= t[0]) { if (disc && key < t[0]) *out_disc = true; return t[1]; }
    if (key >= t[2 * (n - 1)]) { if (disc && key > t[2 * (n - 1)]) *out_disc = true; return t[2 * (n - 1) + 1]; }
    if (sorted) {
        const int i = lower_cell(t, 2, 1, n - 1, key);
        float span = t[2 * i] - t[2 * (i - 1)];
        float f = (span > 0.0f) ? wdiv(key - t[2 * (i - 1)], span) : 0.0f;
        return t[2 * (i - 1) + 1] + f * (t[2 * i + 1] - t[2 * (i - 1) + 1]);
    }
    for (int i = 1; i < n; ++i)
        if (key <= t[2 * i]) {
            float span = t[2 * i] - t[2 * (i - 1)];
            float f = (span > 0.0f) ? (key - t[2 * (i - 1)]) / span : 0.0f;
            return t[2 * (i - 1) + 1] + f * (t[2 * i + 1] - t[2 * (i - 1) + 1]);
        }
    return t[2 * (n - 1) + 1];
}

// Coating classes a lens can use (coating_classes); a kernel instantiated
// for a subset compiles only those branches of surface_weight.
enum CoatClass : int {
    CC_FRESNEL = 1,   // bare interface (also every table-less model)
    CC_AIRY = 2,      // SIMPLE with MgF2 layers
    CC_ROW = 4,       // lambda-only value from the lambda row (SPECTRAL, ARTIST)
    CC_TABLE = 8,     // ANGULAR 1-D table on the angle
    CC_SA = 16,       // SPECTRAL_ANGULAR
    CC_ALL = 31,
};

// surface_weight (trace_core.h:635-783).  cos_i = |dot(nrm, dir)|, the
// caller's (intersect hands it over with the normal).
// lr: this surface's entry of the per-lambda lookup row (the lambda-only part
// of the weight, precomputed on the host with the same float operations), or
// nullptr to evaluate it here.  CM: the coating classes the lens uses.
template <int CM = CC_ALL>
GLD float surface_weight(float cos_i, float n1, float n2, float eta, const DevSurf GLC& s,
                         const float* __restrict__ arena, float lam, bool reflecting, const int2 GLC* lr) {
    float R;
    const int m = s.coat_model;
    if (CM == CC_FRESNEL) {
        R = fresnel_R(cos_i, n1, n2, eta);
    } else if (CM == (CC_SA | CC_FRESNEL) && m != GL_COAT_SPECTRAL_ANGULAR) {
        R = fresnel_R(cos_i, n1, n2, eta);
    } else if (CM == (CC_AIRY | CC_FRESNEL) && m == GL_COAT_SIMPLE && s.ar_layers <= 0) {
        R = fresnel_R(cos_i, n1, n2, eta);
    } else if (CM == (CC_AIRY | CC_FRESNEL)) {
        const float mg = 1.38f;
        const float qw = 550.0f / (4.0f * mg);
        R = coating_R(cos_i, n1, n2, mg, qw, lam, eta);
        for (int i = 1; i < s.ar_layers; ++i) R *= 0.25f;
        R = clamp01(R);
    } else if (m == GL_COAT_SIMPLE) {
        if (s.ar_layers <= 0) R = fresnel_R(cos_i, n1, n2, eta);
        else {
            const float mg = 1.38f;
            const float qw = 550.0f / (4.0f * mg);
            R = coating_R(cos_i, n1, n2, mg, qw, lam, eta);
            for (int i = 1; i < s.ar_layers; ++i) R *= 0.25f;
            R = clamp01(R);
        }
    } else if (m == GL_COAT_ARTIST && lr) {
        R = __int_as_float(ld_i2(lr).y);
    } else if (m == GL_COAT_ARTIST) {
        const float inv_2s2 = 1.0f / (2.0f * 65.0f * 65.0f);
        float dr = lam - 650.0f, dg = lam - 550.0f, db = lam - 450.0f;
        float br = expf(-dr * dr * inv_2s2), bg = expf(-dg * dg * inv_2s2), bb = expf(-db * db * inv_2s2);
        float sum = br + bg + bb;
        R = clamp01(s.tint_strength * (s.tint_r * br + s.tint_g * bg + s.tint_b * bb) / sum);
    } else if (m == GL_COAT_SPECTRAL || m == GL_COAT_ANGULAR) {
        if (s.tab_n <= 0) R = fresnel_R(cos_i, n1, n2, eta);
        else if (m == GL_COAT_SPECTRAL && lr) {
            const int2 e = ld_i2(lr);
            if (e.x == -1) return 0.0f;
            R = __int_as_float(e.y);
        } else {
            float key = (m == GL_COAT_SPECTRAL) ? lam : ref_angle_deg(cos_i, n1, s.angle_ref_ior);
            bool d = false;
            R = lerp1d(arena + s.tab_off, s.tab_n, key, s.discard != 0, &d, (s.flags & SF_TAB_SORTED) != 0);
            if (d) return 0.0f;
        }
    } else if (m == GL_COAT_SPECTRAL_ANGULAR) {
        const int nw = s.sa_nw, na = s.sa_na;
        if (nw <= 0 || na <= 0) R = fresnel_R(cos_i, n1, n2, eta);
        else {
            const float* wl = arena + s.sa_off;
            const float* an = wl + nw;
            const float* rr = an + na;
            const float ang = ref_angle_deg(cos_i, n1, s.angle_ref_ior);
            bool disc = false;
            const bool oob = s.discard != 0;
            const bool sorted = (s.flags & SF_SA_SORTED) != 0;
            // Angle cell: ia, fa.
            int ia = 0; float fa = 0.0f;
            if (s.flags & SF_SA_UNIFORM) {
                // The host checked an[j] == a0 + j * da in float for every j,
                // so the cell comes from arithmetic: ia = floor((ang - a0) /
                // da) clamped to the table, fa measured from the recomputed
                // node (same float operations as the host's check) and
                // clamped to [0, 1].  Where the reference's scan would pick
                // the cell below a node that the floor puts above it, both
                // interpolate to that node's value.
                const float a0 = s.sa_a0, da = s.sa_da;
                const int top = na > 1 ? na - 2 : 0;
                const float t = fmaxf((ang - a0) * s.sa_inv_da, 0.0f);
                ia = min((int)t, top);
                fa = fminf(fmaxf((ang - (a0 + (float)ia * da)) * s.sa_inv_da, 0.0f), na > 1 ? 1.0f : 0.0f);
                if (oob && (ang < a0 || ang > a0 + (float)(na - 1) * da)) disc = true;
            } else if (ang <= an[0]) { if (oob && ang < an[0]) disc = true; }
            else if (ang >= an[na - 1]) { ia = na > 1 ? na - 2 : 0; fa = na > 1 ? 1.0f : 0.0f; if (oob && ang > an[na - 1]) disc = true; }
            else {
                const int i = sorted ? lower_cell(an, 1, 1, na - 1, ang) : 1;
                for (int j = i; j < na; ++j) if (ang <= an[j]) {
                    ia = j - 1; float span = an[j] - an[j - 1];
                    fa = span > 0.0f ? (ang - an[j - 1]) * __builtin_amdgcn_rcpf(span) : 0.0f; break; }
            }
            if (lr) {
                // The lambda row holds this wavelength's table blended along
                // lambda on the host, node by node, as {R_q, R_q+1 - R_q}
                // pairs at a self-relative offset (-1: the wavelength is
                // discarded): one 8-byte gather instead of four table reads.
                const int2 e = ld_i2(lr);
                if (e.x == -1 || disc) return 0.0f;
                // (entries and rows both start at even words: 8-byte aligned)
                const unsigned long long GLC* row =
                    reinterpret_cast<const unsigned long long GLC*>(reinterpret_cast<const float GLC*>(lr) + e.x);
                const unsigned long long q = row[ia];
                R = __builtin_fmaf(fa, __uint_as_float((unsigned)(q >> 32)), __uint_as_float((unsigned)q));
            } else {
                int iw = 0; float fw = 0.0f;
                if (lam <= wl[0]) { if (oob && lam < wl[0]) disc = true; }
                else if (lam >= wl[nw - 1]) { iw = nw > 1 ? nw - 2 : 0; fw = nw > 1 ? 1.0f : 0.0f; if (oob && lam > wl[nw - 1]) disc = true; }
                else {
                    const int i = sorted ? lower_cell(wl, 1, 1, nw - 1, lam) : 1;
                    for (int j = i; j < nw; ++j) if (lam <= wl[j]) {
                        iw = j - 1; float span = wl[j] - wl[j - 1];
                        fw = span > 0.0f ? (lam - wl[j - 1]) / span : 0.0f; break; }
                }
                if (disc) return 0.0f;
                const int iw1 = iw + 1 < nw ? iw + 1 : iw, ia1 = ia + 1 < na ? ia + 1 : ia;
                float r00 = rr[iw * na + ia], r01 = rr[iw * na + ia1], r10 = rr[iw1 * na + ia], r11 = rr[iw1 * na + ia1];
                // bilinear blend (weight only: fused multiply-adds)
                const float lo = __builtin_fmaf(fa, r01 - r00, r00), hi = __builtin_fmaf(fa, r11 - r10, r10);
                R = __builtin_fmaf(fw, hi - lo, lo);
            }
        }
    } else {
        R = fresnel_R(cos_i, n1, n2, eta);
    }
    R = clamp01(R);
    return reflecting ? R : (1.0f - R);
}

GLD float attenuator(F3 hit, const DevSurf GLC& s) {   // trace_core.h:791-809
    if (s.coat_model != GL_COAT_ATTENUATOR_GAUSS) return 1.0f;
    if (!(s.g_sigma > 0.0f)) return s.g_bg;
    F3 local = hit;
    if (s.flags & SF_TRANSFORMED) local = add(m_tmul(s.rot, sub(hit, f3(s.dx, s.dy, s.z))), f3(0.f, 0.f, s.z));
    float ddx = local.x - s.g_dx, ddy = local.y - s.g_dy;
    float r2 = ddx * ddx + ddy * ddy;
    float s2 = s.g_sigma * s.g_sigma;
    return s.g_bg + s.g_peak * expf(-0.5f * r2 / s2);
}

// ---- deformed blades (ApertureProfile::u_edge / fold / u_raw /
// sector_angle / radius_at, aperture_profile.h:116-181) ---------------------
template <class B> __host__ __device__ inline float blade_radius_at(const B& b, float theta) {
    float phi = fmodf(theta - b.rotation, b.sigma);
    if (phi < 0.0f) phi += b.sigma;
    phi -= b.half;
    const float psi = phi - b.phi_w;
    const float sh = sinf(0.5f * psi);
    float u = b.u_w - b.u_rise * (sh * sh) * (psi >= 0.0f ? b.qR : b.qL);
    if (b.facets) {
        float w = fmodf(phi - b.half + b.half, b.sigma);
        if (w < 0.0f) w += b.sigma;
        w -= b.half;
        if (w >= -b.eA && w <= b.eB) {
            const float uf = cosf(w + b.half - b.beta) / b.p;
            if (uf > u) u = uf;
        }
    }
    return 1.0f / u;
}

// Kept out of line: only a deformed stop takes it, and inlined its
// trigonometry would weigh on every surface step's registers.
__device__ __noinline__ bool blade_inside(const DevSurf GLC& s, float hx, float hy, float h2) {
    const float r = sqrtf(h2);
    if (r < 1e-9f) return true;
    return r <= s.sa * blade_radius_at(s.blade, atan2f(hy, hx));
}

// Bitmap transmission at normalised (u, v): bilinear between texel centres
// with clamped edges (tex2D with linear filtering, normalised coordinates and
// clamp addressing, as the reference samples its aperture textures).
template <class SV> __host__ __device__ inline float image_sample(const SV& s, const float* img, float u, float v,
                                                                  int stride = 1) {
    const int W = s.img_w, H = s.img_h;
    const float x = u * (float)W - 0.5f, y = v * (float)H - 0.5f;
    const float fx = floorf(x), fy = floorf(y);
    const float a = x - fx, b = y - fy;
    int i0 = (int)fx, j0 = (int)fy;
    int i1 = i0 + 1, j1 = j0 + 1;
    i0 = i0 < 0 ? 0 : (i0 > W - 1 ? W - 1 : i0);
    i1 = i1 < 0 ? 0 : (i1 > W - 1 ? W - 1 : i1);
    j0 = j0 < 0 ? 0 : (j0 > H - 1 ? H - 1 : j0);
    j1 = j1 < 0 ? 0 : (j1 > H - 1 ? H - 1 : j1);
    const float top = (1.0f - a) * img[(j0 * W + i0) * stride] + a * img[(j0 * W + i1) * stride];
    const float bot = (1.0f - a) * img[(j1 * W + i0) * stride] + a * img[(j1 * W + i1) * stride];
    return (1.0f - b) * top + b * bot;
}

// d_image_aperture_passes (trace_cuda.h:37-54): the bitmap at the surface-
// frame hit must exceed 0.5.
GLD bool image_passes(const DevSurf GLC& s, const float* __restrict__ arena, F3 hit) {
    if (!(s.aspect > 0.0f) || !(s.img_semi > 0.0f)) return false;
    F3 local = hit;
    if (s.flags & SF_TRANSFORMED) local = add(m_tmul(s.rot, sub(hit, f3(s.dx, s.dy, s.z))), f3(0.f, 0.f, s.z));
    const float hx = local.x / s.aspect;
    const float u = 0.5f + hx / (2.0f * s.img_semi);
    const float v = 0.5f + local.y / (2.0f * s.img_semi);
    return image_sample(s, arena + s.img_off, u, v) > 0.5f;
}

// check_aperture (trace_core.h:133-170); the regular polygon is the
// half-plane intersection max_k(h . n_k) <= apothem (same set as the
// reference's r <= apothem / cos(offset)).
// LF ("lean features"): the lens has no stretched and no deformed apertures,
// so those branches are compiled out.  (The tracer's own LF instantiation
// takes step_lean, which handles stretched apertures itself.)
template <bool LF = false>
GLD bool check_aperture(F3 hit, const DevSurf GLC& s, const float GLC* poly) {
    float hx = (!LF && (s.flags & SF_ASPECT)) ? hit.x / s.aspect : hit.x;
    float hy = hit.y;
    float h2 = hx * hx + hy * hy;
    if (!(h2 <= s.sa2)) return false;
    if (!LF && (s.flags & SF_DEFORMED)) return blade_inside(s, hx, hy, h2);
    if (s.flags & SF_POLYGON) {
        const float GLC* pc = poly + s.poly_off;
        const float GLC* ps = pc + GL_MAX_BLADES;
        float m = hx * pc[0] + hy * ps[0];
        for (int k = 1; k < s.nblades; ++k) {
            float d = hx * pc[k] + hy * ps[k];
            m = d > m ? d : m;
        }
        return m <= s.apothem;
    }
    return true;
}

GLD bool pick_root(float t1, float t2, float oz, float dz, float zv, float& t) {
    if (t1 > 1e-6f && t2 > 1e-6f) {
        float z1 = oz + t1 * dz, z2 = oz + t2 * dz;
        t = (fabsf(z1 - zv) < fabsf(z2 - zv)) ? t1 : t2;
    } else if (t1 > 1e-6f) t = t1;
    else if (t2 > 1e-6f) t = t2;
    else return false;
    return true;
}

GLD float asph_sag(float r, float inv_R, float K, const DevSurf GLC& s) {     // trace_core.h:75-88
    float r2 = r * r;
    float D = 1.0f - (1.0f + K) * inv_R * inv_R * r2;
    if (D <= 0.0f) return 1e30f;
    float z = inv_R * r2 / (1.0f + sqrtf(D));
    float rp = r2 * r2;
    for (int i = 0; i < s.n_asph; ++i) { z += s.asph[i] * rp; rp *= r2; }
    return z;
}
GLD float asph_dsag(float r, float inv_R, float K, const DevSurf GLC& s) {    // trace_core.h:96-112
    float r2 = r * r;
    float D = 1.0f - (1.0f + K) * inv_R * inv_R * r2;
    if (D <= 0.0f) return 0.0f;
    float dz = inv_R * r / sqrtf(D);
    float rp = r2 * r;
    for (int i = 0; i < s.n_asph; ++i) { dz += (4.0f + 2.0f * (float)i) * s.asph[i] * rp; rp *= r2; }
    return dz;
}

// intersect_surface (trace_core.h:299-525).  Returns false on miss / clip.
// LF: every surface is flat or spherical, untransformed and solved in FP32
// (the tracer's LF instantiation takes step_lean instead).
// dn_out = dot(n_out, d) in the caller's frame, with the bits the caller's
// own dot product would give: the refraction, reflection and coating weight
// take their incidence cosine from it instead of forming it again.  Where
// the normal was flipped to face the ray, dot(-n, d) = -dot(n, d) exactly
// (negation commutes with every rounding; a zero partial sum changes sign
// only where dot(n, d) itself could not have been > 0).
template <bool LF = false>
GLD bool intersect_dn(F3 o, F3 d, const DevSurf GLC& s, const float GLC* poly, float& t_out, F3& n_out,
                      float& dn_out) {
    const float zv = s.z;
    const bool tr = !LF && (s.flags & SF_TRANSFORMED) != 0;
    const F3 dw = d;
    if (tr) {
        o = add(m_tmul(s.rot, sub(o, f3(s.dx, s.dy, zv))), f3(0.f, 0.f, zv));
        d = m_tmul(s.rot, d);
    }
    if (s.flags & SF_FLAT) {
        if (fabsf(d.z) < 1e-12f) return false;
        float t = (zv - o.z) / d.z;
        if (!(t > 1e-6f)) return false;
        F3 hit = add(o, mul(d, t));
        if (!check_aperture<LF>(hit, s, poly)) return false;
        t_out = t;
        n_out = f3(0.f, 0.f, (d.z > 0.0f) ? -1.0f : 1.0f);
        if (tr) n_out = m_mul(s.rot, n_out);
        dn_out = dot3(n_out, dw);
        return true;
    }
    if (s.form == GL_FORM_SPHERE) {
        const float R = s.R;
        F3 ctr = f3(0.f, 0.f, s.cz);
        float t;
        if (!LF && (s.flags & SF_FP64)) {
            // FP64 quadratic for large |R| (sag cancellation in |oc|^2 - R^2).
            const double Rd = (double)R, zd = (double)zv;
            const double ox = o.x, oy = o.y, oz = (double)o.z - (zd + Rd);
            const double dx = d.x, dy = d.y, dz = d.z;
            const double a = dx * dx + dy * dy + dz * dz;
            const double b = 2.0 * (ox * dx + oy * dy + oz * dz);
            const double c = ox * ox + oy * oy + oz * oz - Rd * Rd;
            const double disc = b * b - 4.0 * a * c;
            if (disc < 0.0) return false;
            const double sd = sqrt(disc), inv_2a = 0.5 / a;
            const double t1 = (-b - sd) * inv_2a, t2 = (-b + sd) * inv_2a;
            double td;
            if (t1 > 1e-6 && t2 > 1e-6) {
                const double z1 = (double)o.z + t1 * dz, z2 = (double)o.z + t2 * dz;
                td = (fabs(z1 - zd) < fabs(z2 - zd)) ? t1 : t2;
            } else if (t1 > 1e-6) td = t1;
            else if (t2 > 1e-6) td = t2;
            else return false;
            t = (float)td;
        } else {
            // a = |d|^2 is within a few ulp of 1 (unit directions), where
            // 0.5 * rcp_rn(a) is the IEEE quotient 0.5 / a.
            F3 oc = sub(o, ctr);
            float a = dot3(d, d), b = 2.0f * dot3(oc, d), c = dot3(oc, oc) - s.R2;
            float disc = b * b - 4.0f * a * c;
            if (disc < 0.0f) return false;
            float sd = sqrtf(disc), inv_2a = 0.5f * rcp_rn(a);
            if (!pick_root((-b - sd) * inv_2a, (-b + sd) * inv_2a, o.z, d.z, zv, t)) return false;
        }
        F3 hit = add(o, mul(d, t));
        if (!check_aperture<LF>(hit, s, poly)) return false;
        F3 nrm = mul(sub(hit, ctr), s.inv_absR);
        float dn = dot3(nrm, d);
        if (dn > 0.0f) { nrm = neg(nrm); dn = -dn; }
        if (tr) { nrm = m_mul(s.rot, nrm); dn = dot3(nrm, dw); }
        t_out = t; n_out = nrm; dn_out = dn;
        return true;
    }
    if (LF) return false;
    if (s.form == GL_FORM_CYLINDRICAL) {
        const float vx = (s.flags & SF_CYL_Y) ? 1.0f : 0.0f, vy = (s.flags & SF_CYL_Y) ? 0.0f : 1.0f;
        const float R = s.R;
        float oc_v = o.x * vx + o.y * vy;
        float oc_z = o.z - zv - R;
        float dv = d.x * vx + d.y * vy;
        float qa = dv * dv + d.z * d.z;
        float qb = 2.0f * (oc_v * dv + oc_z * d.z);
        float qc = oc_v * oc_v + oc_z * oc_z - R * R;
        float disc = qb * qb - 4.0f * qa * qc;
        if (disc < 0.0f) return false;
        float sd = sqrtf(disc), inv_2a = 0.5f / qa, t;
        if (!pick_root((-qb - sd) * inv_2a, (-qb + sd) * inv_2a, o.z, d.z, zv, t)) return false;
        F3 hit = add(o, mul(d, t));
        if (!check_aperture<LF>(hit, s, poly)) return false;
        float hv = hit.x * vx + hit.y * vy;
        F3 nrm = f3(hv * vx, hv * vy, hit.z - zv - R);
        float nlen = sqrtf(nrm.x * nrm.x + nrm.y * nrm.y + nrm.z * nrm.z);
        if (nlen < 1e-12f) return false;
        nrm = mul(nrm, 1.0f / nlen);
        if (dot3(nrm, d) > 0.0f) nrm = neg(nrm);
        if (tr) nrm = m_mul(s.rot, nrm);
        t_out = t; n_out = nrm; dn_out = dot3(nrm, dw);
        return true;
    }
    // FORM_ASPHERE: Newton on the sag residual from the spherical guess.
    {
        const float R = s.R, inv_R = 1.0f / R, K = s.conic;
        F3 ctr = f3(0.f, 0.f, zv + R);
        F3 oc = sub(o, ctr);
        float qa = dot3(d, d), qb = 2.0f * dot3(oc, d), qc = dot3(oc, oc) - R * R;
        float disc = qb * qb - 4.0f * qa * qc;
        float t;
        if (disc >= 0.0f) {
            float sd = sqrtf(disc), inv_2a = 0.5f / qa;
            if (!pick_root((-qb - sd) * inv_2a, (-qb + sd) * inv_2a, o.z, d.z, zv, t)) return false;
        } else {
            if (fabsf(d.z) < 1e-12f) return false;
            t = (zv - o.z) / d.z;
            if (!(t > 1e-6f)) return false;
        }
        for (int it = 0; it < 15; ++it) {
            F3 hi = add(o, mul(d, t));
            float rx = hi.x, ry = hi.y, rad = sqrtf(rx * rx + ry * ry);
            float sag = asph_sag(rad, inv_R, K, s);
            if (sag > 1e29f) return false;
            float f = hi.z - zv - sag;
            if (fabsf(f) < 1e-5f) break;
            float dsag = asph_dsag(rad, inv_R, K, s);
            float r_safe = (rad > 1e-9f) ? rad : 1e-9f;
            float dfdt = d.z - dsag * (rx * d.x + ry * d.y) / r_safe;
            if (fabsf(dfdt) < 1e-12f) break;
            t -= f / dfdt;
        }
        if (!(t > 1e-6f)) return false;
        F3 hit = add(o, mul(d, t));
        float rad = sqrtf(hit.x * hit.x + hit.y * hit.y);
        if (!check_aperture<LF>(hit, s, poly)) return false;
        float r_safe = (rad > 1e-9f) ? rad : 1e-9f;
        float dsag = asph_dsag(rad, inv_R, K, s);
        F3 nrm = f3(-dsag * hit.x / r_safe, -dsag * hit.y / r_safe, 1.0f);
        float nlen = sqrtf(nrm.x * nrm.x + nrm.y * nrm.y + nrm.z * nrm.z);
        nrm = mul(nrm, 1.0f / nlen);
        if (dot3(nrm, d) > 0.0f) nrm = neg(nrm);
        if (tr) nrm = m_mul(s.rot, nrm);
        t_out = t; n_out = nrm; dn_out = dot3(nrm, dw);
        return true;
    }
}
template <bool LF = false>
GLD bool intersect(F3 o, F3 d, const DevSurf GLC& s, const float GLC* poly, float& t_out, F3& n_out) {
    float dn;
    return intersect_dn<LF>(o, d, s, poly, t_out, n_out, dn);
}

// n_ratio = n1 / n2, IEEE-rounded (by the caller, or from the lambda row).
// Both square roots take arguments bounded away from the small-subnormal
// range (1 - sin2_t >= 2^-24 since sin2_t < 1; sq >= 1e-18), so sqrt_rn_big
// gives sqrtf's bits.
// cos_i = -dot(nrm, dir), the caller's.  |tmp| is within rounding of 1 and
// at least 1e-9, so rcp_rn gives the IEEE reciprocal of its length.
GLD bool refract_cos(F3 dir, F3 nrm, float cos_i, float n_ratio, F3& out) {   // trace_core.h:533-552
    float sin2_t = n_ratio * n_ratio * (1.0f - cos_i * cos_i);
    if (sin2_t >= 1.0f) return false;
    float cos_t = sqrt_rn_big(1.0f - sin2_t);
    F3 tmp = add(mul(dir, n_ratio), mul(nrm, n_ratio * cos_i - cos_t));
    float sq = tmp.x * tmp.x + tmp.y * tmp.y + tmp.z * tmp.z;
    if (sq < 1e-18f) return false;
    float inv = rcp_rn(sqrt_rn_big(sq));
    out = mul(tmp, inv);
    return true;
}
GLD bool refract_ratio(F3 dir, F3 nrm, float n_ratio, F3& out) {
    return refract_cos(dir, nrm, -dot3(nrm, dir), n_ratio, out);
}
GLD bool refract(F3 dir, F3 nrm, float n1, float n2, F3& out) { return refract_ratio(dir, nrm, n1 / n2, out); }
// dn = dot(dir, nrm) (the product is commutative, so dot(nrm, dir) too).
GLD F3 reflect_dn(F3 dir, F3 nrm, float dn) { return normalized_unit(sub(dir, mul(nrm, 2.0f * dn))); }
GLD F3 reflect(F3 dir, F3 nrm) { return reflect_dn(dir, nrm, dot3(dir, nrm)); }

// One surface interaction of the lean path (flat, FP32 sphere or cylinder,
// decentred / tilted or not, round, stretched or polygon aperture, no
// bitmap, deformed blade or attenuator): the same float operations as
// intersect_dn / check_aperture / refract_cos below, written so that a lane's
// failure is a flag rather than a branch.  Every lane forms the whole
// interaction and the lanes whose ray died leave at one point at the end
// (status VIGNETTED for a miss or clip, TIR for a failed refraction); what a
// dead lane computed past its failure is discarded.  With one divergent exit
// per surface the compiler keeps the ray state in place instead of merging
// it at each early return.  The surface kind, the transform and the
// stretched aperture are tests of the surface record: uniform over the wave.
// MED = false compiles out the cylinder, the transform and the stretched
// aperture (lenses of flat and spherical surfaces only, lean level 1);
// MED = true keeps them (lean level 2).
template <bool W, int CM, bool RT, bool MED>
GLD bool step_lean(const DevSurf GLC& s, const float GLC* poly, const float* __restrict__ arena, F3& o, F3& d,
                   float lam, float n2, bool reflecting, float reflect_ior, float& cur, TraceRes& res,
                   const int2 GLC* lr, float ratio) {
    ++res.segs;
    const float zv = s.z;
    const int fl = s.flags;
    const bool flat = (fl & SF_FLAT) != 0;                 // uniform
    const bool cyl = MED && !flat && s.form == GL_FORM_CYLINDRICAL;
    const bool tr = MED && (fl & SF_TRANSFORMED) != 0;
    // The interaction is solved in the surface's frame (ol, dl); the hit that
    // moves the ray is formed from the caller's o and d with the same t.
    F3 ol = o, dl = d;
    if (tr) {
        ol = add(m_tmul(s.rot, sub(o, f3(s.dx, s.dy, zv))), f3(0.f, 0.f, zv));
        dl = m_tmul(s.rot, d);
    }
    float t;
    bool ok;
    if (flat) {                                            // intersect_surface, plane
        ok = !(fabsf(dl.z) < 1e-12f);
        t = (zv - ol.z) / dl.z;
        ok = ok && (t > 1e-6f);
    } else if (cyl) {                                      // cylinder about x or y
        const float vx = (fl & SF_CYL_Y) ? 1.0f : 0.0f, vy = (fl & SF_CYL_Y) ? 0.0f : 1.0f;
        const float R = s.R;
        const float oc_v = ol.x * vx + ol.y * vy;
        const float oc_z = ol.z - zv - R;
        const float dv = dl.x * vx + dl.y * vy;
        const float qa = dv * dv + dl.z * dl.z;
        const float qb = 2.0f * (oc_v * dv + oc_z * dl.z);
        const float qc = oc_v * oc_v + oc_z * oc_z - R * R;
        const float disc = qb * qb - 4.0f * qa * qc;
        const float sd = sqrtf(disc), inv_2a = 0.5f / qa;
        const float t1 = (-qb - sd) * inv_2a, t2 = (-qb + sd) * inv_2a;
        const bool c1 = t1 > 1e-6f, c2 = t2 > 1e-6f;
        const float z1 = ol.z + t1 * dl.z, z2 = ol.z + t2 * dl.z;
        const bool near1 = fabsf(z1 - zv) < fabsf(z2 - zv);
        t = c1 ? ((c2 && !near1) ? t2 : t1) : t2;
        ok = !(disc < 0.0f) && (c1 || c2);
    } else {                                               // sphere: nearer root to the vertex plane
        const F3 oc = sub(ol, f3(0.f, 0.f, s.cz));
        const float a = dot3(dl, dl), b = 2.0f * dot3(oc, dl), c = dot3(oc, oc) - s.R2;
        const float disc = b * b - 4.0f * a * c;
        // sqrt_rn_pos: sqrtf's bits wherever disc >= 0 (a lane with disc < 0
        // is dead and its root unused); a = |dl|^2 is within rounding of 1
        const float sd = sqrt_rn_pos(disc), inv_2a = 0.5f * rcp_rn(a);
        const float t1 = (-b - sd) * inv_2a, t2 = (-b + sd) * inv_2a;
        const bool c1 = t1 > 1e-6f, c2 = t2 > 1e-6f;
        const float z1 = ol.z + t1 * dl.z, z2 = ol.z + t2 * dl.z;
        const bool near1 = fabsf(z1 - zv) < fabsf(z2 - zv);
        t = c1 ? ((c2 && !near1) ? t2 : t1) : t2;
        ok = !(disc < 0.0f) && (c1 || c2);
    }
    const F3 hl = add(ol, mul(dl, t));                     // the hit in the surface's frame
    const float hx = (MED && (fl & SF_ASPECT)) ? hl.x / s.aspect : hl.x, hy = hl.y;   // check_aperture
    const float h2 = hx * hx + hy * hy;
    ok = ok && (h2 <= s.sa2);
    if (fl & SF_POLYGON) {
        const float GLC* pc = poly + s.poly_off;
        const float GLC* ps = pc + GL_MAX_BLADES;
        float m = hx * pc[0] + hy * ps[0];
        for (int k = 1; k < s.nblades; ++k) {
            const float e = hx * pc[k] + hy * ps[k];
            m = e > m ? e : m;
        }
        ok = ok && (m <= s.apothem);
    }
    F3 nrm;
    float dn;
    if (flat) {
        nrm = f3(0.f, 0.f, (dl.z > 0.0f) ? -1.0f : 1.0f);
        if (tr) nrm = m_mul(s.rot, nrm);
        dn = dot3(nrm, d);
    } else {
        if (cyl) {
            const float vx = (fl & SF_CYL_Y) ? 1.0f : 0.0f, vy = (fl & SF_CYL_Y) ? 0.0f : 1.0f;
            const float hv = hl.x * vx + hl.y * vy;
            nrm = f3(hv * vx, hv * vy, hl.z - zv - s.R);
            const float nlen = sqrtf(nrm.x * nrm.x + nrm.y * nrm.y + nrm.z * nrm.z);
            ok = ok && !(nlen < 1e-12f);
            nrm = mul(nrm, 1.0f / nlen);
        } else {
            nrm = mul(sub(hl, f3(0.f, 0.f, s.cz)), s.inv_absR);
        }
        dn = dot3(nrm, dl);
        const bool flip = dn > 0.0f;
        nrm = flip ? neg(nrm) : nrm;
        dn = flip ? -dn : dn;
        if (tr) {
            nrm = m_mul(s.rot, nrm);
            dn = dot3(nrm, d);
        }
    }
    o = tr ? add(o, mul(d, t)) : hl;
    const float n1 = cur;
    bool bent = true;
    if (reflecting) {                                      // uniform
        if (W) res.w *= surface_weight<CM>(fabsf(dn), n1, n2, RT ? ratio : n1 / n2, s, arena, lam, true, lr);
        d = reflect_dn(d, nrm, dn);
        cur = reflect_ior;
    } else {                                               // refract_cos
        const float eta = RT ? ratio : n1 / n2;
        const float ci = -dn;
        const float sin2_t = eta * eta * (1.0f - ci * ci);
        const float cos_t = sqrt_rn_big(1.0f - sin2_t);
        const F3 tmp = add(mul(d, eta), mul(nrm, eta * ci - cos_t));
        const float sq = tmp.x * tmp.x + tmp.y * tmp.y + tmp.z * tmp.z;
        bent = !(sin2_t >= 1.0f) && !(sq < 1e-18f);
        if (W) res.w *= surface_weight<CM>(fabsf(dn), n1, n2, eta, s, arena, lam, false, lr);
        d = mul(tmp, rcp_rn(sqrt_rn_big(sq)));
        cur = n2;
    }
    if (!(ok && bent)) {
        res.status = ok ? GL_TRACE_TIR : GL_TRACE_VIGNETTED;
        return false;
    }
    return true;
}

// One surface interaction (trace_cuda.h:226-361 loop bodies).  Returns false
// when the ray died; res.status is set to TIR for total internal reflection.
// W = false skips the Fresnel / coating / attenuator weights (callers that
// need only landings, statuses and segment counts).
// ratio: n1 / n2 from the caller's table when RT (the same IEEE quotient);
// otherwise it is divided here, and only where it is used.  LF: the lens's
// lean level (lean_level); 1 and 2 take the predicated step_lean.
template <bool W, int LF = 0, int CM = CC_ALL, bool RT = false>
GLD bool step(const DevSurf GLC& s, const float GLC* poly, const float* __restrict__ arena, F3& o, F3& d,
              float lam, float n2, bool reflecting, float reflect_ior, float& cur, TraceRes& res,
              const int2 GLC* lr, float ratio = 0.0f) {
    if constexpr (LF != 0) {
        return step_lean<W, CM, RT, LF == 2>(s, poly, arena, o, d, lam, n2, reflecting, reflect_ior, cur, res, lr,
                                            ratio);
    } else {
    float t, dn; F3 nrm;
    ++res.segs;
    if (!intersect_dn<false>(o, d, s, poly, t, nrm, dn)) return false;
    F3 hit = add(o, mul(d, t));
    o = hit;
    if (!LF && (s.flags & SF_IMAGE) && !image_passes(s, arena, hit)) return false;
    const float n1 = cur;
    if (reflecting) {
        if (W) {
            res.w *= surface_weight<CM>(fabsf(dn), n1, n2, RT ? ratio : n1 / n2, s, arena, lam, true, lr);
            if (!LF) res.w *= attenuator(hit, s);
        }
        d = reflect_dn(d, nrm, dn);
        cur = reflect_ior;
    } else {
        F3 nd;
        const float eta = RT ? ratio : n1 / n2;
        if (!refract_cos(d, nrm, -dn, eta, nd)) { res.status = GL_TRACE_TIR; return false; }
        if (W) {
            res.w *= surface_weight<CM>(fabsf(dn), n1, n2, eta, s, arena, lam, false, lr);
            if (!LF) res.w *= attenuator(hit, s);
        }
        d = nd;
        cur = n2;
    }
    return true;
    }
}

// ---- front-of-lens baffles (baffle_stack_blocks / baffle_stack_edge_distance,
// baffle.h:49-113).  A pupil ray at spawn (px, py) heading (tx, ty) meets
// baffle k at (px + tx dz_k, py + ty dz_k). -----------------------------------
template <class BS> __host__ __device__ inline bool baffles_block(const BS& st, float px, float py, float tx,
                                                                  float ty) {
    for (int k = 0; k < st.n; ++k) {
        const gl_baffle& b = st.b[k];
        const float mx = px + tx * b.dz, my = py + ty * b.dz;
        if (b.shape == GL_BAFFLE_RECT) {
            if (mx > b.right || mx < -b.left || my > b.top || my < -b.bottom) return true;
        } else {
            const float ex = (mx - b.cx) / b.radius, ey = (my - b.cy) / (b.radius * b.aspect);
            if (ex * ex + ey * ey > 1.0f) return true;
        }
    }
    return false;
}
// Clearance to the nearest baffle edge inside the shapes, and its normal.
template <class BS> __host__ __device__ inline float baffles_edge(const BS& st, float px, float py, float tx,
                                                                 float ty, F3& nrm) {
    float best = 1e30f;
    F3 bn; bn.x = bn.y = bn.z = 0.0f;
    for (int k = 0; k < st.n; ++k) {
        const gl_baffle& b = st.b[k];
        const float mx = px + tx * b.dz, my = py + ty * b.dz;
        float d; F3 n; n.z = 0.0f;
        if (b.shape == GL_BAFFLE_RECT) {
            d = b.right - mx; n.x = 1.0f; n.y = 0.0f;
            const float dl = mx + b.left, dt = b.top - my, db = my + b.bottom;
            if (dl < d) { d = dl; n.x = -1.0f; n.y = 0.0f; }
            if (dt < d) { d = dt; n.x = 0.0f; n.y = 1.0f; }
            if (db < d) { d = db; n.x = 0.0f; n.y = -1.0f; }
        } else {
            const float rr = (b.radius > 1e-12f) ? b.radius : 1e-12f;
            const float ra = rr * ((b.aspect > 0.0f) ? b.aspect : 1.0f);
            const float ex = mx - b.cx, ey = my - b.cy;
            const float g = sqrtf((ex / rr) * (ex / rr) + (ey / ra) * (ey / ra));
            const float gx = ex / (rr * rr), gy = ey / (ra * ra);
            const float gl = sqrtf(gx * gx + gy * gy);
            if (gl < 1e-20f) { d = rr; n.x = 1.0f; n.y = 0.0f; }
            else { d = (1.0f - g) / gl; n.x = gx / gl; n.y = gy / gl; }
        }
        if (d < best) { best = d; bn = n; }
    }
    if (best < 0.0f) best = 0.0f;
    nrm = bn;
    return best;
}

// ---- HURB (hurb.h:40-150): a ray passing an edge at distance d gets a random
// transverse kick of scale lambda K / d, Cauchy or Gaussian, clamped. ---------
__host__ __device__ inline float hurb_sigma(float lam_nm, float d_mm, int dist) {
    if (d_mm <= 0.0f) return 0.0f;
    const float lam_mm = lam_nm * 1.0e-6f;
    return lam_mm * (dist == GL_HURB_LORENTZIAN ? 0.159154943f : 0.112539539f) / d_mm;
}
__host__ __device__ inline uint32_t hurb_hash(uint32_t s) {
    s = (s ^ 61u) ^ (s >> 16); s *= 9u; s = s ^ (s >> 4); s *= 0x27d4eb2du; s = s ^ (s >> 15); return s;
}
__host__ __device__ inline float hurb_uniform(uint32_t& st) {
    st ^= st << 13; st ^= st >> 17; st ^= st << 5;
    return (float)(st & 0x00FFFFFFu) * (1.0f / 16777216.0f);
}
__host__ __device__ inline float hurb_draw(int dist, float sigma, float max_kick, uint32_t& st) {
    float k;
    if (dist == GL_HURB_LORENTZIAN) {
        const float u = hurb_uniform(st);
        k = sigma * tanf(3.14159265358979323846f * (u - 0.5f));
    } else {
        const float u1 = fmaxf(hurb_uniform(st), 1.0e-7f);
        const float u2 = hurb_uniform(st);
        k = sigma * sqrtf(-2.0f * logf(u1)) * cosf(2.0f * 3.14159265358979323846f * u2);
    }
    if (k > max_kick) k = max_kick;
    if (k < -max_kick) k = -max_kick;
    return k;
}
// Rotate `dir` by the kick inside the plane of dir and the edge normal (|dir|
// stays 1); far edges (sigma <= min_sigma) draw nothing.
template <class H> __host__ __device__ inline void hurb_apply(const H& h, F3& dir, F3 n_edge, float d_mm,
                                                              float lam_nm, uint32_t& st) {
    const float sigma = hurb_sigma(lam_nm, d_mm, h.dist);
    if (sigma <= h.min_sigma) return;
    const float theta = hurb_draw(h.dist, sigma, h.max_kick, st);
    const float dn = n_edge.x * dir.x + n_edge.y * dir.y + n_edge.z * dir.z;
    F3 t; t.x = n_edge.x - dir.x * dn; t.y = n_edge.y - dir.y * dn; t.z = n_edge.z - dir.z * dn;
    const float tl = sqrtf(t.x * t.x + t.y * t.y + t.z * t.z);
    if (tl < 1.0e-8f) return;
    const float it = 1.0f / tl;
    t.x *= it; t.y *= it; t.z *= it;
    const float tk = tanf(theta);
    F3 nd; nd.x = dir.x + t.x * tk; nd.y = dir.y + t.y * tk; nd.z = dir.z + t.z * tk;
    const float nl = sqrtf(nd.x * nd.x + nd.y * nd.y + nd.z * nd.z);
    if (nl < 1.0e-12f) return;
    const float in = 1.0f / nl;
    dir.x = nd.x * in; dir.y = nd.y * in; dir.z = nd.z * in;
}

// Distance from a deformed blade edge (ApertureProfile::edge_gap,
// aperture_profile.h:205-230): first order to the conic edge, exact to a tip
// facet, the nearer wins; r in the normalised frame.
template <class B> __host__ __device__ inline float blade_edge_gap(const B& b, float r, float theta, float& nx,
                                                                   float& ny) {
    float phi = fmodf(theta - b.rotation, b.sigma);
    if (phi < 0.0f) phi += b.sigma;
    phi -= b.half;
    const float ct = cosf(theta), st = sinf(theta);
    const float psi = phi - b.phi_w;
    const float sh = sinf(0.5f * psi);
    const float q = (psi >= 0.0f) ? b.qR : b.qL;
    const float u = b.u_w - b.u_rise * (sh * sh) * q;
    const float du = -0.5f * b.u_rise * q * sinf(psi);
    const float k = (-du / (u * u)) / r;
    const float inv = 1.0f / sqrtf(1.0f + k * k);
    float best = (1.0f / u - r) * inv;
    nx = (ct + k * st) * inv;
    ny = (st - k * ct) * inv;
    if (b.facets) {
        float w = fmodf(phi - b.half + b.half, b.sigma);
        if (w < 0.0f) w += b.sigma;
        w -= b.half;
        const float ang = w + b.half - b.beta;
        const float d = b.p - r * cosf(ang);
        if (d < best) { best = d; nx = cosf(theta - ang); ny = sinf(theta - ang); }
    }
    return best;
}

// aperture_edge_distance (trace_core.h:206-290): distance from a world-space
// hit on surface s to the clear-aperture edge and the world-space edge normal,
// in the same stretched frame and polygon geometry the aperture test uses.
//
// An image aperture has no analytic edge: the stop's baked distance field is
// sampled where the trace samples its bitmap (d_aperture_edge_distance,
// trace_cuda.h:72-97); any other image surface never kicks.
template <class SV> __host__ __device__ inline float surface_edge(const SV& s, F3 hit, F3& n_out,
                                                                  const float* arena) {
    F3 h = hit;
    const bool tr = (s.flags & SF_TRANSFORMED) != 0;
    if (tr) {
        F3 q; q.x = hit.x - s.dx; q.y = hit.y - s.dy; q.z = hit.z - s.z;
        h.x = s.rot[0] * q.x + s.rot[3] * q.y + s.rot[6] * q.z;
        h.y = s.rot[1] * q.x + s.rot[4] * q.y + s.rot[7] * q.z;
        h.z = s.rot[2] * q.x + s.rot[5] * q.y + s.rot[8] * q.z + s.z;
    }
    if (s.ap_blades < 0) {
        if (s.sdf_off < 0 || arena == nullptr) {
            n_out.x = n_out.y = n_out.z = 0.0f;
            return 1e30f;
        }
        const float hxi = h.x / s.aspect;
        const float ui = 0.5f + hxi / (2.0f * s.img_semi);
        const float vi = 0.5f + h.y / (2.0f * s.img_semi);
        const float* f = arena + s.sdf_off;
        const float d = image_sample(s, f, ui, vi, 4);
        F3 nl; nl.x = image_sample(s, f + 1, ui, vi, 4); nl.y = image_sample(s, f + 2, ui, vi, 4); nl.z = 0.0f;
        if (tr) {
            n_out.x = s.rot[0] * nl.x + s.rot[1] * nl.y + s.rot[2] * nl.z;
            n_out.y = s.rot[3] * nl.x + s.rot[4] * nl.y + s.rot[5] * nl.z;
            n_out.z = s.rot[6] * nl.x + s.rot[7] * nl.y + s.rot[8] * nl.z;
        } else {
            n_out = nl;
        }
        return d > 0.0f ? d : 0.0f;
    }
    const float sa = s.sa;
    const float hx = h.x / s.aspect, hy = h.y;
    const float r_s = sqrtf(hx * hx + hy * hy);
    float d;
    F3 n; n.z = 0.0f;
    if (s.ap_blades >= 3) {
        const bool deformed = (s.flags & SF_DEFORMED) != 0;
        const float sector = 2.0f * 3.14159265358979323846f / (float)s.ap_blades;
        const float apothem = sa * cosf(0.5f * sector);
        if (r_s < 1e-9f) {
            d = deformed ? sa * s.blade.r_w : apothem;
            n.x = 1.0f; n.y = 0.0f;
        } else if (deformed) {
            const float theta = atan2f(hy, hx);
            float nx, ny;
            d = sa * blade_edge_gap(s.blade, r_s / sa, theta, nx, ny);
            n.x = nx; n.y = ny;
        } else {
            const float theta = atan2f(hy, hx);
            float off = fmodf(theta - s.ap_rot, sector);
            if (off < 0.0f) off += sector;
            off -= 0.5f * sector;
            d = apothem - r_s * cosf(off);
            const float phi = theta - off;
            n.x = cosf(phi); n.y = sinf(phi);
        }
    } else {
        const float asp2 = s.aspect * s.aspect;
        const float gx = h.x / asp2, gy = h.y;
        const float glen = sqrtf(gx * gx + gy * gy);
        if (glen < 1e-12f) { d = sa; n.x = 1.0f; n.y = 0.0f; }
        else { d = (sa - r_s) * r_s / glen; n.x = gx / glen; n.y = gy / glen; }
    }
    if (d < 0.0f) d = 0.0f;
    if (tr) {
        n_out.x = s.rot[0] * n.x + s.rot[1] * n.y + s.rot[2] * n.z;
        n_out.y = s.rot[3] * n.x + s.rot[4] * n.y + s.rot[5] * n.z;
        n_out.z = s.rot[6] * n.x + s.rot[7] * n.y + s.rot[8] * n.z;
    } else {
        n_out = n;
    }
    return d;
}

GLD void to_sensor(F3 o, F3 d, bool t_ge0, TraceRes& res) {   // trace_cuda.h:366-380
    if (fabsf(d.z) < 1e-12f) { res.status = GL_TRACE_MISSED_SURFACE; return; }
    float ts = -o.z / d.z;
    bool ok = t_ge0 ? (ts >= 0.0f) : (ts > 0.0f);
    if (!ok) { res.status = GL_TRACE_MISSED_SURFACE; return; }
    F3 p = add(o, mul(d, ts));
    res.px = p.x; res.py = p.y; res.pz = p.z;
    res.status = GL_TRACE_OK;
}

// d_trace_ghost_ray (trace_cuda.h:195-382) with per-lambda IOR slices.  The
// three legs -- forward 0..b reflecting at b, backward b-1..a against the
// index before each surface reflecting at a, forward a+1..n-1 -- run as ONE
// loop over a wave-uniform step counter, so the surface interaction is
// instantiated once (a third of the code of three inlined legs, which keeps
// the kernel inside the instruction cache).  The surface record is read
// through the constant address space (scalar loads, uniform index).
//
// H = HURB: after every surface interaction the new direction takes an edge
// kick (hurb_apply) drawn from the per-trace stream `hs`
// (d_trace_ghost_ray<HURB>, trace_cuda.h:248-361).
//
// rat (read when RT; null otherwise): this wavelength's refraction ratios,
// {ibf[k] / iat[k], iat[k] / ibf[k]} per surface, divided on the host in IEEE
// float; they also serve as the Fresnel eta.  The index
// on the incident side of surface k is ibf[k] on the two forward legs and
// iat[k] on the backward one (muted surfaces are skipped, so the medium
// carried between surfaces is exactly that), so the ratio the refraction
// needs is rat[k].x forward and rat[k].y backward.
template <bool W = true, bool H = false, int LF = 0, int CM = CC_ALL, bool RT = false>
GLD TraceRes trace_ghost(const DevSurf GLC* S, int n, const float GLC* poly, const float* __restrict__ arena,
                         int a, int b, F3 o, F3 d, float lam, const float GLC* iat, const float GLC* ibf,
                         bool t_ge0, const int2 GLC* lrow = nullptr, const HurbDev* hb = nullptr,
                         uint32_t* hs = nullptr, const float2 GLC* rat = nullptr) {
    TraceRes res; res.status = GL_TRACE_VIGNETTED; res.w = 1.0f; res.segs = 0; res.px = res.py = res.pz = 0.f;
    float cur = 1.0f;
    const float refl_a = iat[a];
    const int end1 = b + 1, end2 = end1 + (b - a), end3 = end2 + (n - 1 - a);
    for (int i = 0; i < end3; ++i) {
        int k;
        float n2, rior;
        bool refl;
        if (i < end1) {
            k = i; n2 = iat[k]; refl = (k == b); rior = cur;
        } else if (i < end2) {
            k = b - 1 - (i - end1); n2 = ibf[k]; refl = (k == a); rior = refl_a;
        } else {
            k = a + 1 + (i - end2); n2 = iat[k]; refl = false; rior = 0.0f;
        }
        const DevSurf GLC& s = S[k];
        if (!(s.flags & SF_ACTIVE)) continue;
        float ratio = 0.0f;
        if (RT) {
            const float GLC* rk = reinterpret_cast<const float GLC*>(rat + k);
            ratio = (i >= end1 && i < end2) ? rk[1] : rk[0];
        }
        if (!step<W, LF, CM, RT>(s, poly, arena, o, d, lam, n2, refl, rior, cur, res, lrow ? lrow + k : nullptr,
                                 ratio))
            return res;
        if (H) {
            F3 en;
            const float ed = surface_edge(s, o, en, arena);
            hurb_apply(*hb, d, en, ed, lam, *hs);
        }
    }
    to_sensor(o, d, t_ge0, res);
    return res;
}

// The same trace reading this wavelength's tables from interleaved records
// (LamRec, one per surface): the index, ratio and coating row of surface k
// come from one address, which is what a lane-varying wavelength needs.
// Always the refraction-ratio (RT) form.
template <int LF, int CM>
GLD TraceRes trace_ghost_rec(const DevSurf GLC* S, int n, const float GLC* poly, const float* __restrict__ arena,
                             int a, int b, F3 o, F3 d, float lam, const LamRec GLC* rec, bool t_ge0) {
    TraceRes res; res.status = GL_TRACE_VIGNETTED; res.w = 1.0f; res.segs = 0; res.px = res.py = res.pz = 0.f;
    float cur = 1.0f;
    const float refl_a = rec[a].iat;
    const int end1 = b + 1, end2 = end1 + (b - a), end3 = end2 + (n - 1 - a);
    for (int i = 0; i < end3; ++i) {
        int k;
        float rior;
        bool refl, back;
        if (i < end1) {
            k = i; refl = (k == b); rior = cur; back = false;
        } else if (i < end2) {
            k = b - 1 - (i - end1); refl = (k == a); rior = refl_a; back = true;
        } else {
            k = a + 1 + (i - end2); refl = false; rior = 0.0f; back = false;
        }
        const DevSurf GLC& s = S[k];
        if (!(s.flags & SF_ACTIVE)) continue;
        const LamRec GLC& q = rec[k];
        const float GLC* t4 = reinterpret_cast<const float GLC*>(&q);
        const float n2 = back ? t4[1] : t4[0];
        const float ratio = back ? t4[3] : t4[2];
        if (!step<true, LF, CM, true>(s, poly, arena, o, d, lam, n2, refl, rior, cur, res, &q.lrow, ratio))
            return res;
    }
    to_sensor(o, d, t_ge0, res);
    return res;
}

// d_trace_primary_ray (trace_cuda.h:111-172).
// LF: the lean level of the lens (lean_level: 0 generic, 1 or 2 predicated).
template <bool W = true, int LF = 0>
GLD TraceRes trace_primary(const DevSurf GLC* S, int n, const float GLC* poly, const float* __restrict__ arena,
                           F3 o, F3 d, float lam, const float GLC* iat, bool t_ge0, F3* out_dir) {
    TraceRes res; res.status = GL_TRACE_VIGNETTED; res.w = 1.0f; res.segs = 0; res.px = res.py = res.pz = 0.f;
    float cur = 1.0f;
    for (int k = 0; k < n; ++k) {
        const DevSurf GLC& s = S[k];
        if (!(s.flags & SF_ACTIVE)) continue;
        if (!step<W, LF>(s, poly, arena, o, d, lam, iat[k], false, 0.0f, cur, res, nullptr)) return res;
    }
    to_sensor(o, d, t_ge0, res);
    if (out_dir) *out_dir = d;
    return res;
}

// ---- pupil sampling on the device (aperture_sampler.h:136-170) ------------
__host__ __device__ __forceinline__ uint32_t wang_hash(uint32_t s) {
    s = (s ^ 61u) ^ (s >> 16u); s *= 9u; s ^= s >> 4u; s *= 0x27d4eb2du; s ^= s >> 15u; return s;
}
GLD float halton2(uint32_t n) {
    n = (n << 16u) | (n >> 16u);
    n = ((n & 0x00ff00ffu) << 8u) | ((n & 0xff00ff00u) >> 8u);
    n = ((n & 0x0f0f0f0fu) << 4u) | ((n & 0xf0f0f0f0u) >> 4u);
    n = ((n & 0x33333333u) << 2u) | ((n & 0xccccccccu) >> 2u);
    n = ((n & 0x55555555u) << 1u) | ((n & 0xaaaaaaaau) >> 1u);
    return (float)n * (1.0f / 4294967296.0f);
}
GLD float halton3(uint32_t n) {
    float r = 0.0f, f = 1.0f / 3.0f;
    while (n > 0) { r += (float)(n % 3u) * f; n /= 3u; f /= 3.0f; }
    return r;
}

// Pupil silhouette (PupilMask, aperture_sampler.h:49-80) in half-plane form.
struct DevMask {
    float aspect;
    int32_t nb;
    float apothem;
    int32_t deformed;                // the stop's blade shape is authored: test r <= r(theta)
    float c[GL_MAX_BLADES], s[GL_MAX_BLADES];
    DevBlade blade;
};
GLD bool mask_contains(const DevMask& m, float u, float v) {
    const float ux = u / m.aspect;
    const float r2 = ux * ux + v * v;
    if (r2 > 1.0f) return false;
    if (m.deformed) return sqrtf(r2) <= blade_radius_at(m.blade, atan2f(v, ux));
    if (m.nb >= 3) {
        float mx = ux * m.c[0] + v * m.s[0];
        for (int k = 1; k < m.nb; ++k) { float dd = ux * m.c[k] + v * m.s[k]; mx = dd > mx ? dd : mx; }
        if (mx > m.apothem) return false;
    }
    return true;
}

}  // namespace gl
