// hostprep.hip — host preparation for the flare path: spectral tables and
// colour matrices, the entrance-pupil sampler, per-lambda IOR tables, the
// thin-film (TMM) coating bake, parameter defaults and the device surface
// layout.  Float expressions keep the reference's operation order so the
// tables are bit-identical to its CPU code (checked against oracle/_ref).
#include "gl_runtime.h"

#include <chrono>
#include <cmath>

namespace gl {

// CIE 1931 2-degree standard observer, 380-780 nm in 5 nm steps
// (CIE 15:2004, Table 1).  Rows: x-bar, y-bar, z-bar.
static const float kCieX[81] = {
    0.001368f, 0.002236f, 0.004243f, 0.007650f, 0.014310f, 0.023190f, 0.043510f, 0.077630f, 0.134380f,
    0.214770f, 0.283900f, 0.328500f, 0.348280f, 0.348060f, 0.336200f, 0.318700f, 0.290800f, 0.251100f,
    0.195360f, 0.142100f, 0.095640f, 0.057950f, 0.032010f, 0.014700f, 0.004900f, 0.002400f, 0.009300f,
    0.029100f, 0.063270f, 0.109600f, 0.165500f, 0.225750f, 0.290400f, 0.359700f, 0.433450f, 0.512050f,
    0.594500f, 0.678400f, 0.762100f, 0.842500f, 0.916300f, 0.978600f, 1.026300f, 1.056700f, 1.062200f,
    1.045600f, 1.002600f, 0.938400f, 0.854450f, 0.751400f, 0.642400f, 0.541900f, 0.447900f, 0.360800f,
    0.283500f, 0.218700f, 0.164900f, 0.121200f, 0.087400f, 0.063600f, 0.046770f, 0.032900f, 0.022700f,
    0.015840f, 0.011359f, 0.008111f, 0.005790f, 0.004109f, 0.002899f, 0.002049f, 0.001440f, 0.001000f,
    0.000690f, 0.000476f, 0.000332f, 0.000235f, 0.000166f, 0.000117f, 0.000083f, 0.000059f, 0.000042f};
static const float kCieY[81] = {
    0.000039f, 0.000064f, 0.000120f, 0.000217f, 0.000396f, 0.000640f, 0.001210f, 0.002180f, 0.004000f,
    0.007300f, 0.011600f, 0.016840f, 0.023000f, 0.029800f, 0.038000f, 0.048000f, 0.060000f, 0.073900f,
    0.090980f, 0.112600f, 0.139020f, 0.169300f, 0.208020f, 0.258600f, 0.323000f, 0.407300f, 0.503000f,
    0.608200f, 0.710000f, 0.793200f, 0.862000f, 0.914850f, 0.954000f, 0.980300f, 0.994950f, 1.000000f,
    0.995000f, 0.978600f, 0.952000f, 0.915400f, 0.870000f, 0.816300f, 0.757000f, 0.694900f, 0.631000f,
    0.566800f, 0.503000f, 0.441200f, 0.381000f, 0.321000f, 0.265000f, 0.217000f, 0.175000f, 0.138200f,
    0.107000f, 0.081600f, 0.061000f, 0.044580f, 0.032000f, 0.023200f, 0.017000f, 0.011920f, 0.008210f,
    0.005723f, 0.004102f, 0.002929f, 0.002091f, 0.001484f, 0.001047f, 0.000740f, 0.000520f, 0.000361f,
    0.000249f, 0.000172f, 0.000120f, 0.000085f, 0.000060f, 0.000042f, 0.000030f, 0.000021f, 0.000015f};
static const float kCieZ[81] = {
    0.006450f, 0.010550f, 0.020050f, 0.036210f, 0.067850f, 0.110200f, 0.207400f, 0.371300f, 0.645600f,
    1.039050f, 1.385600f, 1.622960f, 1.747060f, 1.782600f, 1.772110f, 1.744100f, 1.669200f, 1.528100f,
    1.287640f, 1.041900f, 0.812950f, 0.616200f, 0.465180f, 0.353300f, 0.272000f, 0.212300f, 0.158200f,
    0.111700f, 0.078250f, 0.057250f, 0.042160f, 0.029840f, 0.020300f, 0.013400f, 0.008750f, 0.005750f,
    0.003900f, 0.002750f, 0.002100f, 0.001800f, 0.001650f, 0.001400f, 0.001100f, 0.001000f, 0.000800f,
    0.000600f, 0.000340f, 0.000240f, 0.000190f, 0.000100f, 0.000050f, 0.000030f, 0.000020f, 0.000010f,
    0.000000f, 0.000000f, 0.000000f, 0.000000f, 0.000000f, 0.000000f, 0.000000f, 0.000000f, 0.000000f,
    0.000000f, 0.000000f, 0.000000f, 0.000000f, 0.000000f, 0.000000f, 0.000000f, 0.000000f, 0.000000f,
    0.000000f, 0.000000f, 0.000000f, 0.000000f, 0.000000f, 0.000000f, 0.000000f, 0.000000f, 0.000000f};

static void cie_at(float lam, float* x, float* y, float* z) {
    // Linear interpolation between 5 nm nodes; zero outside 380-780 nm.
    if (lam < 380.0f || lam > 780.0f) { *x = *y = *z = 0.0f; return; }
    int lo = 0, hi = 79;
    while (lo < hi) {
        int mid = (lo + hi) / 2;
        if (380.0f + 5.0f * (float)(mid + 1) < lam) lo = mid + 1; else hi = mid;
    }
    const float w0 = 380.0f + 5.0f * (float)lo, w1 = 380.0f + 5.0f * (float)(lo + 1);
    const float t = (lam - w0) / (w1 - w0);
    *x = kCieX[lo] + t * (kCieX[lo + 1] - kCieX[lo]);
    *y = kCieY[lo] + t * (kCieY[lo + 1] - kCieY[lo]);
    *z = kCieZ[lo] + t * (kCieZ[lo + 1] - kCieZ[lo]);
}

// Published colour matrices (row-major): ACES AP1 (AMF-0-2008-001), sRGB /
// Rec.709 (IEC 61966-2-1), DCI-P3, Bradford E->D60 adaptation.
static const float kAP1toXYZ[9] = {0.66246f, 0.13400f, 0.15618f, 0.27223f, 0.67408f, 0.05369f,
                                   -0.00557f, 0.00406f, 1.01034f};
static const float kSRGBtoXYZ[9] = {0.4124f, 0.3576f, 0.1805f, 0.2126f, 0.7152f, 0.0722f,
                                    0.0193f, 0.1192f, 0.9505f};
static const float kBradfordEtoD60[9] = {0.95307f, -0.02663f, 0.02375f, -0.03845f, 1.02927f, 0.00937f,
                                         0.00258f, -0.00298f, 1.08912f};
static const float kXYZtoAP1[9] = {1.6410f, -0.3248f, -0.2358f, -0.6636f, 1.6153f, 0.0167f,
                                   0.0117f, -0.0082f, 0.9883f};
static const float kXYZtoSRGB[9] = {3.2406f, -1.5372f, -0.4986f, -0.9689f, 1.8758f, 0.0415f,
                                    0.0557f, -0.2040f, 1.0570f};
static const float kXYZtoP3D65[9] = {2.4935f, -0.9314f, -0.4027f, -0.8295f, 1.7627f, 0.0236f,
                                     0.0358f, -0.0761f, 0.9569f};
static const float kXYZtoP3D60[9] = {2.4932f, -0.9316f, -0.4027f, -0.8320f, 1.7648f, 0.0232f,
                                     0.0363f, -0.0765f, 0.9568f};

static void m3_ident(float* m) { for (int i = 0; i < 9; ++i) m[i] = (i % 4 == 0) ? 1.0f : 0.0f; }
static void m3_mul(const float* A, const float* B, float* out) {
    float t[9];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
            float acc = 0.0f;
            for (int k = 0; k < 3; ++k) acc += A[3 * r + k] * B[3 * k + c];
            t[3 * r + c] = acc;
        }
    std::memcpy(out, t, sizeof t);
}

static void input_matrix(int cs, const float* custom, float* out) {
    if (cs == GL_CS_INPUT_ACESCG) std::memcpy(out, kAP1toXYZ, 36);
    else if (cs == GL_CS_INPUT_SRGB_LINEAR) std::memcpy(out, kSRGBtoXYZ, 36);
    else if (cs == GL_CS_INPUT_CUSTOM && custom) std::memcpy(out, custom, 36);
    else m3_ident(out);
}
static void output_matrix(int cs, const float* custom, float* out) {
    switch (cs) {
    case GL_CS_ACESCG: m3_mul(kXYZtoAP1, kBradfordEtoD60, out); break;
    case GL_CS_SRGB_LINEAR: std::memcpy(out, kXYZtoSRGB, 36); break;
    case GL_CS_P3_D65: std::memcpy(out, kXYZtoP3D65, 36); break;
    case GL_CS_P3_D60: m3_mul(kXYZtoP3D60, kBradfordEtoD60, out); break;
    case GL_CS_CUSTOM: if (custom) std::memcpy(out, custom, 36); else m3_ident(out); break;
    default: m3_ident(out); break;
    }
}

static std::vector<gl_spec> table_of(int n, const float* M) {
    std::vector<gl_spec> t(n);
    const float delta = (700.0f - 400.0f) / (float)n;
    for (int i = 0; i < n; ++i) {
        float lam = 400.0f + ((float)i + 0.5f) * delta;
        float x, y, z;
        cie_at(lam, &x, &y, &z);
        // The CIE observer is its own XYZ basis (identity sensor matrix).
        float xr = 1.0f * x + 0.0f * y + 0.0f * z;
        float xg = 0.0f * x + 1.0f * y + 0.0f * z;
        float xb = 0.0f * x + 0.0f * y + 1.0f * z;
        t[i].lambda = lam;
        t[i].cmf_r = M[0] * xr + M[1] * xg + M[2] * xb;
        t[i].cmf_g = M[3] * xr + M[4] * xg + M[5] * xb;
        t[i].cmf_b = M[6] * xr + M[7] * xg + M[8] * xb;
    }
    float sum_g = 0.0f;
    for (int i = 0; i < n; ++i) sum_g += t[i].cmf_g * delta;
    if (sum_g > 1e-9f) {
        const float inv = 1.0f / sum_g;
        for (auto& s : t) { s.cmf_r *= inv; s.cmf_g *= inv; s.cmf_b *= inv; }
    }
    return t;
}

int spectral_table(int n, int output_cs, const float* custom9, std::vector<gl_spec>& out) {
    if (n <= 0 || n > GL_MAX_SPECTRAL) return set_error(GL_ERR_INVALID, "spectral_samples %d outside [1, %d]", n, GL_MAX_SPECTRAL);
    if (n == 3) {   // fixed RGB preview path (ghost_render.cu:777-785)
        out = {{650.0f, 1.0f, 0.0f, 0.0f}, {550.0f, 0.0f, 1.0f, 0.0f}, {450.0f, 0.0f, 0.0f, 1.0f}};
        return GL_OK;
    }
    float M[9];
    output_matrix(output_cs, custom9, M);
    out = table_of(n, M);
    return GL_OK;
}

int spectral_table_cmf(int n, int output_cs, const float* custom9, std::vector<gl_spec>& out) {
    if (n <= 0 || n > GL_MAX_SPECTRAL) return set_error(GL_ERR_INVALID, "spectral_samples %d outside [1, %d]", n, GL_MAX_SPECTRAL);
    float M[9];
    output_matrix(output_cs, custom9, M);
    out = table_of(n, M);
    return GL_OK;
}

static void gram(const std::vector<gl_spec>& t, float O[9]) {
    for (int i = 0; i < 9; ++i) O[i] = 0.0f;
    for (const auto& s : t) {
        const float c[3] = {s.cmf_r, s.cmf_g, s.cmf_b};
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) O[3 * a + b] += c[a] * c[b];
    }
}
static bool m3_inv(const float* m, float* inv) {
    const float det = m[0] * (m[4] * m[8] - m[5] * m[7]) - m[1] * (m[3] * m[8] - m[5] * m[6]) +
                      m[2] * (m[3] * m[7] - m[4] * m[6]);
    if (std::fabs(det) < 1e-20f) { m3_ident(inv); return false; }
    const float id = 1.0f / det;
    inv[0] = (m[4] * m[8] - m[5] * m[7]) * id;
    inv[1] = -(m[1] * m[8] - m[2] * m[7]) * id;
    inv[2] = (m[1] * m[5] - m[2] * m[4]) * id;
    inv[3] = -(m[3] * m[8] - m[5] * m[6]) * id;
    inv[4] = (m[0] * m[8] - m[2] * m[6]) * id;
    inv[5] = -(m[0] * m[5] - m[2] * m[3]) * id;
    inv[6] = (m[3] * m[7] - m[4] * m[6]) * id;
    inv[7] = -(m[0] * m[7] - m[1] * m[6]) * id;
    inv[8] = (m[0] * m[4] - m[1] * m[3]) * id;
    return true;
}

// C = (b_anchor / b_n) * (O_dense_hat . O_n_hat^-1): maps every sample count
// to the converged (300-sample) colour at the 16-sample brightness.
int spectral_correction(const std::vector<gl_spec>& t, int output_cs, const float* custom9, float C[9]) {
    float M[9];
    output_matrix(output_cs, custom9, M);
    float On[9], Od[9], Oa[9];
    gram(t, On);
    gram(table_of(300, M), Od);
    gram(table_of(16, M), Oa);
    const float bn = On[3] + On[4] + On[5], bd = Od[3] + Od[4] + Od[5], ba = Oa[3] + Oa[4] + Oa[5];
    if (!(bn > 1e-20f) || !(bd > 1e-20f)) { m3_ident(C); return GL_OK; }
    float Onh[9], Odh[9];
    const float ibn = 1.0f / bn, ibd = 1.0f / bd;
    for (int i = 0; i < 9; ++i) { Onh[i] = On[i] * ibn; Odh[i] = Od[i] * ibd; }
    float Ninv[9], Cc[9];
    m3_inv(Onh, Ninv);
    m3_mul(Odh, Ninv, Cc);
    const float s = ba / bn;
    for (int i = 0; i < 9; ++i) C[i] = Cc[i] * s;
    return GL_OK;
}

// ---- dispersion (fresnel.h:27-63) -----------------------------------------
static float dispersion_ior(float n_d, float V_d, float lambda_nm) {
    if (V_d < 0.1f || n_d <= 1.0001f) return n_d;
    const float lF = 486.13f, lC = 656.27f, ld = 587.56f;
    float dn = (n_d - 1.0f) / V_d;
    float iF = 1.0f / (lF * lF), iC = 1.0f / (lC * lC), idl = 1.0f / (ld * ld);
    float B = dn / (iF - iC);
    float A = n_d - B * idl;
    return A + B / (lambda_nm * lambda_nm);
}
float ior_at(const gl_surface& s, float lambda_nm) {
    if (s.disp_model == GL_DISP_SELLMEIER) {
        float l = lambda_nm * 1e-3f, l2 = l * l, n2 = 1.0f;
        n2 += s.sellmeier_B[0] * l2 / (l2 - s.sellmeier_C[0]);
        n2 += s.sellmeier_B[1] * l2 / (l2 - s.sellmeier_C[1]);
        n2 += s.sellmeier_B[2] * l2 / (l2 - s.sellmeier_C[2]);
        return (n2 > 0.0f) ? std::sqrt(n2) : 1.0f;
    }
    return dispersion_ior(s.ior, s.abbe_v, lambda_nm);
}
void ior_tables(const gl_surface* s, int n, const gl_spec* spec, int ns, float* at, float* bf) {
    for (int si = 0; si < ns; ++si) {
        float last = 1.0f;
        for (int k = 0; k < n; ++k) {
            at[(size_t)si * n + k] = ior_at(s[k], spec[si].lambda);
            bf[(size_t)si * n + k] = last;
            if (s[k].is_active) last = at[(size_t)si * n + k];
        }
    }
}

// ---- polygon half-plane data (FP64, rounded once) -------------------------
void poly_normals(int nb, float rot_rad, float scale, float* c, float* s, float* apothem) {
    const double sector = 2.0 * M_PI / (double)nb;
    for (int k = 0; k < nb; ++k) {
        const double phi = (double)rot_rad + sector * ((double)k + 0.5);
        c[k] = (float)std::cos(phi);
        s[k] = (float)std::sin(phi);
    }
    *apothem = (float)((double)scale * std::cos(M_PI / (double)nb));
}

// The device form of a surface's blade profile; false when its polygon is
// plain (or it is not a bladed polygon).
static bool deformed_blade(const gl_surface& g, DevBlade* out) {
    if (g.aperture_shape != GL_APERTURE_POLYGON || g.aperture_blades < 3) return false;
    gl_blade_profile b;
    gl_blade_profile_build(g.aperture_blades, g.aperture_rotation_rad, g.aperture_curvature, g.aperture_twist,
                           g.aperture_notch_rad, g.aperture_notch_angle_rad, &b);
    if (b.blades < 3 || b.plain) return false;
    DevBlade d{};
    d.rotation = b.rotation; d.sigma = b.sigma; d.half = b.half; d.phi_w = b.phi_w;
    d.u_w = b.u_w; d.u_rise = b.u_rise; d.qR = b.qR; d.qL = b.qL;
    d.p = b.p; d.beta = b.beta; d.eA = b.eA; d.eB = b.eB; d.facets = b.facets; d.r_w = b.r_w;
    *out = d;
    return true;
}

DevMask build_mask(const gl_surface* s, int n, int bo, float rdeg) {
    int nb = 0; float rot = 0.f, aspect = 1.f;
    const gl_surface* stop = nullptr;
    for (int i = 0; i < n; ++i)
        if (s[i].is_stop) {
            aspect = s[i].aperture_aspect;
            if (s[i].aperture_shape == GL_APERTURE_POLYGON && s[i].aperture_blades >= 3) {
                nb = s[i].aperture_blades; rot = s[i].aperture_rotation_rad;
                stop = &s[i];
            }
            break;
        }
    // An override describes a plain n-gon: it drops any authored blade shape.
    if (bo > 0) { nb = bo; rot = rdeg * ((float)M_PI / 180.0f); stop = nullptr; }
    DevMask m{};
    m.aspect = aspect;
    m.nb = (nb >= 3 && nb <= GL_MAX_BLADES) ? nb : 0;
    m.apothem = 1.0f;
    if (m.nb) poly_normals(m.nb, rot, 1.0f, m.c, m.s, &m.apothem);
    m.deformed = (m.nb && stop && deformed_blade(*stop, &m.blade)) ? 1 : 0;
    return m;
}

bool mask_contains_host(const DevMask& m, float u, float v) {
    const float ux = u / m.aspect;
    const float r2 = ux * ux + v * v;
    if (r2 > 1.0f) return false;
    if (m.deformed) return std::sqrt(r2) <= blade_radius_at(m.blade, std::atan2(v, ux));
    if (m.nb >= 3) {
        float mx = ux * m.c[0] + v * m.s[0];
        for (int k = 1; k < m.nb; ++k) { float d = ux * m.c[k] + v * m.s[k]; mx = d > mx ? d : mx; }
        if (mx > m.apothem) return false;
    }
    return true;
}

static uint32_t wang(uint32_t s) {
    s = (s ^ 61u) ^ (s >> 16u); s *= 9u; s ^= s >> 4u; s *= 0x27d4eb2du; s ^= s >> 15u; return s;
}
static float h2(uint32_t n) {
    n = (n << 16u) | (n >> 16u);
    n = ((n & 0x00ff00ffu) << 8u) | ((n & 0xff00ff00u) >> 8u);
    n = ((n & 0x0f0f0f0fu) << 4u) | ((n & 0xf0f0f0f0u) >> 4u);
    n = ((n & 0x33333333u) << 2u) | ((n & 0xccccccccu) >> 2u);
    n = ((n & 0x55555555u) << 1u) | ((n & 0xaaaaaaaau) >> 1u);
    return (float)n * (1.0f / 4294967296.0f);
}
static float h3(uint32_t n) {
    float r = 0.0f, f = 1.0f / 3.0f;
    while (n > 0) { r += (float)(n % 3u) * f; n /= 3u; f /= 3.0f; }
    return r;
}

std::vector<gl_sample> build_samples(const gl_surface* s, int n, int N, int jitter, int seed, int bo, float rdeg) {
    std::vector<gl_sample> out;
    if (N <= 0) return out;
    out.reserve((size_t)N * N);
    const DevMask m = build_mask(s, n, bo, rdeg);
    const uint32_t so = (uint32_t)seed * 1000003u;
    for (int k = 0; k < N * N; ++k) {
        const int gx = k % N, gy = k / N;
        float u, v;
        if (jitter == 2) { u = h2((uint32_t)k) * 2.0f - 1.0f; v = h3((uint32_t)k) * 2.0f - 1.0f; }
        else {
            float ju = (jitter == 1) ? (float)wang((uint32_t)k + so) / 4294967296.0f : 0.5f;
            float jv = (jitter == 1) ? (float)wang((uint32_t)k + (uint32_t)(N * N) + so) / 4294967296.0f : 0.5f;
            u = (((float)gx + ju) / (float)N) * 2.0f - 1.0f;
            v = (((float)gy + jv) / (float)N) * 2.0f - 1.0f;
        }
        if (mask_contains_host(m, u, v)) out.push_back({u, v});
    }
    return out;
}

// ---- device surface layout -------------------------------------------------
static bool rot_identity(const float m[9]) {
    const float eps = 1e-6f;
    return (std::fabs(m[0] - 1.f) < eps && std::fabs(m[1]) < eps && std::fabs(m[2]) < eps &&
            std::fabs(m[3]) < eps && std::fabs(m[4] - 1.f) < eps && std::fabs(m[5]) < eps &&
            std::fabs(m[6]) < eps && std::fabs(m[7]) < eps && std::fabs(m[8] - 1.f) < eps);
}

static bool nondecreasing(const float* v, int n, int stride) {
    for (int i = 1; i < n; ++i)
        if (!(v[(size_t)i * stride] >= v[(size_t)(i - 1) * stride])) return false;
    return true;
}

void build_devsurfs(const gl_surface* s, int n, float fp64_rmin, const float* arena, int arena_len,
                    std::vector<DevSurf>& out, std::vector<float>& poly) {
    out.assign(n, DevSurf{});
    poly.clear();
    for (int i = 0; i < n; ++i) {
        const gl_surface& g = s[i];
        DevSurf& d = out[i];
        d.z = g.z; d.R = g.radius; d.sa = g.semi_aperture; d.aspect = g.aperture_aspect;
        d.dx = g.decenter_x; d.dy = g.decenter_y;
        d.inv_absR = (std::fabs(g.radius) > 0.0f) ? 1.0f / std::fabs(g.radius) : 0.0f;
        // The tracer's sphere centre z, squared radius and squared
        // semi-aperture, rounded once here with the same float operations.
        d.cz = g.z + g.radius;
        d.R2 = g.radius * g.radius;
        d.sa2 = g.semi_aperture * g.semi_aperture;
        d.conic = g.conic_k;
        std::memcpy(d.rot, g.rot, sizeof d.rot);
        int f = 0;
        if (g.is_active) f |= SF_ACTIVE;
        if (g.decenter_x != 0.0f || g.decenter_y != 0.0f || !rot_identity(g.rot)) f |= SF_TRANSFORMED;
        if (std::fabs(g.radius) < 1e-6f) f |= SF_FLAT;
        if (g.aperture_aspect != 1.0f) f |= SF_ASPECT;
        if (g.cyl_axis == GL_CYL_AXIS_Y) f |= SF_CYL_Y;
        if (g.form == GL_FORM_SPHERE && fp64_rmin > 0.0f && std::fabs(g.radius) >= fp64_rmin) f |= SF_FP64;
        d.form = g.form;
        d.nblades = 0; d.poly_off = 0; d.apothem = 0.0f;
        if (deformed_blade(g, &d.blade)) f |= SF_DEFORMED;
        d.ap_rot = g.aperture_rotation_rad;
        d.sdf_off = -1;
        if (g.aperture_shape == GL_APERTURE_IMAGE && g.image_width > 0 && g.image_height > 0 && g.image_offset >= 0 &&
            arena != nullptr && (long long)g.image_offset + (long long)g.image_width * g.image_height <= arena_len) {
            f |= SF_IMAGE;
            d.img_off = g.image_offset; d.img_w = g.image_width; d.img_h = g.image_height;
            d.img_semi = g.aperture_semi_diameter;
            if (g.image_sdf && g.image_sdf_offset >= 0 &&
                (long long)g.image_sdf_offset + 4LL * g.image_width * g.image_height <= arena_len)
                d.sdf_off = g.image_sdf_offset;
        }
        d.ap_blades = (g.aperture_shape == GL_APERTURE_IMAGE) ? -1
                      : ((g.aperture_shape == GL_APERTURE_POLYGON && g.aperture_blades >= 3) ? g.aperture_blades : 0);
        if (g.aperture_shape == GL_APERTURE_POLYGON && g.aperture_blades >= 3 && g.aperture_blades <= GL_MAX_BLADES) {
            f |= SF_POLYGON;
            d.nblades = g.aperture_blades;
            d.poly_off = (int)poly.size();
            poly.resize(poly.size() + 2 * GL_MAX_BLADES, 0.0f);
            poly_normals(g.aperture_blades, g.aperture_rotation_rad, g.semi_aperture,
                         &poly[d.poly_off], &poly[d.poly_off + GL_MAX_BLADES], &d.apothem);
        }
        if (g.table_count > 0 && g.table_offset >= 0 && g.table_offset + 2 * g.table_count <= arena_len &&
            nondecreasing(arena + g.table_offset, g.table_count, 2))
            f |= SF_TAB_SORTED;
        if (g.sa_n_wavelengths > 0 && g.sa_n_angles > 0 && g.sa_offset >= 0 &&
            g.sa_offset + g.sa_n_wavelengths + g.sa_n_angles <= arena_len &&
            nondecreasing(arena + g.sa_offset, g.sa_n_wavelengths, 1) &&
            nondecreasing(arena + g.sa_offset + g.sa_n_wavelengths, g.sa_n_angles, 1)) {
            f |= SF_SA_SORTED;
            // An angle axis on an exact uniform grid (the TMM bake's 0, 5, ...,
            // 85 deg) lets the lookup start from an arithmetic guess.
            const float* an = arena + g.sa_offset + g.sa_n_wavelengths;
            const int na_ = g.sa_n_angles;
            if (na_ >= 2 && an[1] > an[0]) {
                const float da = an[1] - an[0];
                bool uniform = true;
                for (int j = 2; j < na_ && uniform; ++j) uniform = an[j] == an[0] + (float)j * da;
                if (uniform) {
                    f |= SF_SA_UNIFORM;
                    d.sa_a0 = an[0];
                    d.sa_inv_da = 1.0f / da;
                    d.sa_da = da;
                }
            }
        }
        d.flags = f;
        d.coat_model = g.coating_model; d.ar_layers = g.ar_layers; d.discard = g.out_of_range_discard;
        d.angle_ref_ior = g.angle_ref_ior;
        d.tab_off = g.table_offset; d.tab_n = g.table_count;
        d.sa_off = g.sa_offset; d.sa_nw = g.sa_n_wavelengths; d.sa_na = g.sa_n_angles;
        int na = g.n_asphere_terms; na = na < 0 ? 0 : (na > GL_MAX_ASPHERE_TERMS ? GL_MAX_ASPHERE_TERMS : na);
        d.n_asph = na;
        std::memcpy(d.asph, g.asphere_terms, sizeof d.asph);
        d.tint_r = g.tint_r; d.tint_g = g.tint_g; d.tint_b = g.tint_b; d.tint_strength = g.tint_strength;
        d.g_sigma = g.gauss_sigma; d.g_bg = g.gauss_background; d.g_peak = g.gauss_peak;
        d.g_dx = g.gauss_decenter_x; d.g_dy = g.gauss_decenter_y;
    }
    if (poly.empty()) poly.assign(2 * GL_MAX_BLADES, 0.0f);
}

// The lambda-only part of surface_weight for every (lambda, surface), with the
// device's float operation order (trace_core.h:575-602, 657-770):
//   SPECTRAL_ANGULAR  x = wavelength cell iw, y = bits of its fraction fw
//   SPECTRAL          x = -2, y = bits of the interpolated reflectance
//   ARTIST            x = -2, y = bits of the tinted reflectance
//   out of range with discard: x = -1 (weight 0);  other models: x = -3.
void lambda_rows(const gl_surface* s, int n, const float* arena, const gl_spec* spec, int ns, std::vector<int32_t>& buf) {
    std::vector<int2> out((size_t)ns * n, make_int2(-3, 0));
    std::vector<float> rows;                 // SPECTRAL_ANGULAR rows, appended after the entries
    const size_t rows_at = 2 * out.size();   // in 4-byte words
    auto bits = [](float f) { int i; std::memcpy(&i, &f, 4); return i; };
    for (int si = 0; si < ns; ++si) {
        const float lam = spec[si].lambda;
        for (int k = 0; k < n; ++k) {
            const gl_surface& g = s[k];
            int2& e = out[(size_t)si * n + k];
            const bool oob = g.out_of_range_discard != 0;
            if (g.coating_model == GL_COAT_SPECTRAL_ANGULAR && g.sa_n_wavelengths > 0 && g.sa_n_angles > 0) {
                const int nw = g.sa_n_wavelengths;
                const float* wl = arena + g.sa_offset;
                int iw = 0; float fw = 0.0f; bool disc = false;
                if (lam <= wl[0]) { if (oob && lam < wl[0]) disc = true; }
                else if (lam >= wl[nw - 1]) { iw = nw > 1 ? nw - 2 : 0; fw = nw > 1 ? 1.0f : 0.0f; if (oob && lam > wl[nw - 1]) disc = true; }
                else for (int i = 1; i < nw; ++i) if (lam <= wl[i]) {
                    iw = i - 1; const float span = wl[i] - wl[i - 1];
                    fw = span > 0.0f ? (lam - wl[i - 1]) / span : 0.0f; break; }
                if (disc) {
                    e = make_int2(-1, 0);
                } else {
                    // This wavelength's table blended along lambda, node by
                    // node, stored as {R_q, R_q+1 - R_q} (the last node's
                    // slope is 0: the device clamps the upper node there).
                    const int na = g.sa_n_angles;
                    const float* rr = arena + g.sa_offset + nw + na;
                    const int iw1 = iw + 1 < nw ? iw + 1 : iw;
                    std::vector<float> bl(na);
                    for (int q = 0; q < na; ++q) {
                        const float r0 = rr[iw * na + q], r1 = rr[iw1 * na + q];
                        bl[q] = std::fma(fw, r1 - r0, r0);
                    }
                    const size_t at = rows_at + rows.size();
                    for (int q = 0; q < na; ++q) {
                        rows.push_back(bl[q]);
                        rows.push_back(q + 1 < na ? bl[q + 1] - bl[q] : 0.0f);
                    }
                    e = make_int2((int)(at - 2 * ((size_t)si * n + k)), 0);
                }
            } else if (g.coating_model == GL_COAT_SPECTRAL && g.table_count > 0) {
                const float* t = arena + g.table_offset;
                const int tn = g.table_count;
                bool disc = false; float R;
                if (lam <= t[0]) { if (oob && lam < t[0]) disc = true; R = t[1]; }
                else if (lam >= t[2 * (tn - 1)]) { if (oob && lam > t[2 * (tn - 1)]) disc = true; R = t[2 * (tn - 1) + 1]; }
                else {
                    R = t[2 * (tn - 1) + 1];
                    for (int i = 1; i < tn; ++i)
                        if (lam <= t[2 * i]) {
                            const float span = t[2 * i] - t[2 * (i - 1)];
                            const float f = (span > 0.0f) ? (lam - t[2 * (i - 1)]) / span : 0.0f;
                            R = t[2 * (i - 1) + 1] + f * (t[2 * i + 1] - t[2 * (i - 1) + 1]);
                            break;
                        }
                }
                e = disc ? make_int2(-1, 0) : make_int2(-2, bits(R));
            } else if (g.coating_model == GL_COAT_ARTIST) {
                const float inv_2s2 = 1.0f / (2.0f * 65.0f * 65.0f);
                const float dr = lam - 650.0f, dg = lam - 550.0f, db = lam - 450.0f;
                const float br = std::exp(-dr * dr * inv_2s2), bg = std::exp(-dg * dg * inv_2s2), bb = std::exp(-db * db * inv_2s2);
                const float sum = br + bg + bb;
                float R = g.tint_strength * (g.tint_r * br + g.tint_g * bg + g.tint_b * bb) / sum;
                R = std::fmin(std::fmax(R, 0.0f), 1.0f);
                e = make_int2(-2, bits(R));
            }
        }
    }
    buf.resize(rows_at + rows.size());
    std::memcpy(buf.data(), out.data(), out.size() * sizeof(int2));
    if (!rows.empty()) std::memcpy(buf.data() + rows_at, rows.data(), rows.size() * sizeof(float));
}

double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}


}  // namespace gl

extern "C" {

void gl_flare_params_init(gl_flare_params* p) {
    std::memset(p, 0, sizeof *p);
    p->ray_grid = 64; p->spectral_samples = 16; p->pupil_jitter = 0; p->jitter_seed = 0;
    p->sensor_model = 0; p->input_cs = GL_CS_INPUT_ACESCG; p->output_cs = GL_CS_ACESCG;
    p->spectral_correction = 1;
    p->aperture_blades = 0; p->aperture_rotation = 0.0f;
    p->flare_gain = 1000.0f; p->min_ghost_intensity = 1e-7f; p->ghost_normalize = 1; p->max_area_boost = 100.0f;
    p->cull_dead_pairs = 1; p->concentrate_samples = 1; p->adaptive_sample_budgets = 1;
    p->adaptive_density_boost = 8.0f; p->adaptive_min_samples = 1024;
    p->aov_mode = GL_AOV_NONE; p->aov_max_pairs = -1;
    p->ghost_filter_mode = GL_FILTER_ALL;
    p->fp64_radius_min = 1000.0f;
    p->batch_mode = GL_BATCH_GRAPH;
    p->schedule = -1;
    p->hurb = 0; p->hurb_kick = GL_HURB_LORENTZIAN;
    p->hurb_min_sigma_rad = 1.0e-5f; p->hurb_max_kick_rad = 0.35f;
    p->n_baffles = 0;
}

int gl_build_spectral_table(int32_t n, int32_t output_cs, const float* custom9, gl_spec* out) {
    std::vector<gl_spec> t;
    int st = gl::spectral_table(n, output_cs, custom9, t);
    if (st) return st;
    std::copy(t.begin(), t.end(), out);
    return GL_OK;
}

int gl_spectral_correction(const gl_spec* table, int32_t n, int32_t output_cs, const float* custom9, float* C) {
    if (!table || n <= 0) return gl::set_error(GL_ERR_INVALID, "empty spectral table");
    std::vector<gl_spec> t(table, table + n);
    return gl::spectral_correction(t, output_cs, custom9, C);
}

int gl_color_matrices(int32_t in_cs, const float* cin, int32_t out_cs, const float* cout, float* Min, float* Mout) {
    gl::input_matrix(in_cs, cin, Min);
    gl::output_matrix(out_cs, cout, Mout);
    return GL_OK;
}

int gl_hurb_sample(float lambda_nm, const float* d_mm, int32_t n, int32_t dist, float max_kick_rad, uint32_t seed,
                   float* kicks, float* sigma) {
    if (n < 0 || (n && (!d_mm || !kicks || !sigma))) return gl::set_error(GL_ERR_INVALID, "null argument");
    for (int i = 0; i < n; ++i) {
        const float sg = gl::hurb_sigma(lambda_nm, d_mm[i], dist);
        uint32_t st = gl::hurb_hash(seed ^ (2654435761u * (uint32_t)i + 1u));
        if (st == 0u) st = 1u;
        kicks[i] = (sg > 0.0f) ? gl::hurb_draw(dist, sg, max_kick_rad, st) : 0.0f;
        sigma[i] = sg;
    }
    return GL_OK;
}

int gl_aperture_edge_distance(const gl_surface* s, const float hit[3], float* d_mm, float normal[3]) {
    if (!s || !hit || !d_mm || !normal) return gl::set_error(GL_ERR_INVALID, "null argument");
    if (s->aperture_shape == GL_APERTURE_IMAGE) {
        *d_mm = 1e30f;
        normal[0] = normal[1] = normal[2] = 0.0f;
        return GL_OK;
    }
    std::vector<gl::DevSurf> ds;
    std::vector<float> poly;
    gl::build_devsurfs(s, 1, 0.0f, nullptr, 0, ds, poly);
    gl::F3 h, nrm;
    h.x = hit[0]; h.y = hit[1]; h.z = hit[2];
    *d_mm = gl::surface_edge(ds[0], h, nrm, nullptr);
    normal[0] = nrm.x; normal[1] = nrm.y; normal[2] = nrm.z;
    return GL_OK;
}

int gl_build_aperture_samples(const gl_surface* s, int32_t n, int32_t ray_grid, int32_t jitter, int32_t seed,
                              int32_t bo, float rdeg, gl_sample* out, int32_t cap, int32_t* n_out) {
    if (!s || n <= 0 || n > GL_MAX_SURFACES) return gl::set_error(GL_ERR_INVALID, "bad surface array");
    std::vector<gl_sample> g = gl::build_samples(s, n, ray_grid, jitter, seed, bo, rdeg);
    *n_out = (int32_t)g.size();
    if ((int)g.size() > cap) return gl::set_error(GL_ERR_INVALID, "sample capacity %d < %zu", cap, g.size());
    std::copy(g.begin(), g.end(), out);
    return GL_OK;
}

int gl_build_ior_tables(const gl_surface* s, int32_t n, const gl_spec* spec, int32_t ns, float* at, float* bf) {
    if (!s || n <= 0 || !spec || ns <= 0) return gl::set_error(GL_ERR_INVALID, "bad arguments");
    gl::ior_tables(s, n, spec, ns, at, bf);
    return GL_OK;
}


}  // extern "C"
