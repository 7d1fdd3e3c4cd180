// tmm_bake.cpp — thin-film coating stack baked to a spectral-angular table
// by the transfer-matrix method (Macleod, "Thin-Film Optical Filters", ch. 2).
//
// Host code only, compiled by the system C++ compiler rather than hipcc:
// complex<float> division and square root are lowered differently by the two
// compilers (promotion and algorithm choice), and the table has to reproduce
// the reference's load-time bake (optical_system.cpp:1483-1570 over
// tmm.h:20-111) to the bit.  The evaluation order below is the one those
// files use; it is an algorithm that must stay alike for identical tables.
#include <cmath>
#include <complex>
#include <vector>

#include "../../include/ghostlight_amd.h"

namespace gl {
int set_error(int code, const char* fmt, ...);
}

namespace {

using cf = std::complex<float>;

// Cosine of the propagation angle inside index n for the transverse invariant
// n0 sin(theta0); the branch with Im(n cos) >= 0 decays into the medium.
cf cos_inside(const cf& n, float invariant) {
    const cf s = cf(invariant, 0.0f) / n;
    cf c = std::sqrt(cf(1.0f, 0.0f) - s * s);
    if ((n * c).imag() < 0.0f) c = -c;
    return c;
}

// Amplitude reflectance of ambient | layers | substrate for one polarisation
// (tilted admittance n cos for s, n / cos for p).
cf amplitude(int n_layers, const cf* idx, const float* thick_nm, cf n_amb, cf n_sub, float lambda_nm,
             float invariant, bool s_pol) {
    auto admittance = [&](const cf& n) {
        const cf c = cos_inside(n, invariant);
        return s_pol ? n * c : n / c;
    };
    const cf eta0 = admittance(n_amb);
    const cf eta_s = admittance(n_sub);
    cf a = cf(1.0f, 0.0f), b(0.0f, 0.0f), c(0.0f, 0.0f), d = cf(1.0f, 0.0f);   // running product [a b; c d]
    for (int l = 0; l < n_layers; ++l) {
        const cf cth = cos_inside(idx[l], invariant);
        const cf eta = s_pol ? idx[l] * cth : idx[l] / cth;
        const cf phase = cf(2.0f * 3.14159265358979323846f * thick_nm[l] / lambda_nm, 0.0f) * idx[l] * cth;
        const cf cp = std::cos(phase);
        const cf sp = std::sin(phase);
        const cf isp(-sp.imag(), sp.real());
        const cf m01 = isp / eta, m10 = isp * eta;
        const cf na = a * cp + b * m10;
        const cf nb = a * m01 + b * cp;
        const cf nc = c * cp + d * m10;
        const cf nd = c * m01 + d * cp;
        a = na; b = nb; c = nc; d = nd;
    }
    const cf B = a + b * eta_s;
    const cf C = c + d * eta_s;
    return (eta0 * B - C) / (eta0 * B + C);
}

}  // namespace

extern "C" {

int gl_bake_tmm_table(int32_t nl, const float* n_re, const float* n_im, const float* thick, const float* n_amb,
                      const float* n_sub, float* out) {
    if (nl <= 0 || nl > 64) return gl::set_error(GL_ERR_INVALID, "layer count %d", nl);
    if (!n_re || !n_im || !thick || !n_amb || !n_sub || !out) return gl::set_error(GL_ERR_INVALID, "null argument");
    const int NW = 31, NA = 18;
    float* wl = out;
    float* ang = out + NW;
    float* R = out + NW + NA;
    for (int w = 0; w < NW; ++w) wl[w] = 400.0f + (float)w * 10.0f;
    for (int k = 0; k < NA; ++k) ang[k] = 0.0f + (float)k * 5.0f;
    std::vector<cf> idx(nl);
    for (int w = 0; w < NW; ++w) {
        for (int l = 0; l < nl; ++l) idx[l] = cf(n_re[l * NW + w], n_im[l * NW + w]);
        const cf amb(n_amb[w], 0.0f), sub(n_sub[w], 0.0f);
        for (int k = 0; k < NA; ++k) {
            const float theta0 = ang[k] * (3.14159265358979323846f / 180.0f);
            const float invariant = amb.real() * std::sin(theta0);
            const cf rs = amplitude(nl, idx.data(), thick, amb, sub, wl[w], invariant, true);
            const cf rp = amplitude(nl, idx.data(), thick, amb, sub, wl[w], invariant, false);
            float r = 0.5f * (std::norm(rs) + std::norm(rp));
            if (r < 0.0f) r = 0.0f;
            if (r > 1.0f) r = 1.0f;
            R[w * NA + k] = r;
        }
    }
    return GL_OK;
}

}  // extern "C"
