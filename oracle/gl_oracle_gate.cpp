// ORACLE — test infrastructure only (see gl_oracle.cpp).
//
// gl_oracle_gate.cpp — CPU restatement of the film-gate scatter pass
// (gate.h:33-211, gate_render.cu:73-280): FP32 primary traces of every
// (source, pupil sample, wavelength) on the oracle tracer, the four-wall band
// test, Schlick reflectance, the hashed scatter lobe and the bilinear splat,
// accumulated in double in a fixed order.  The sample grid and the spectral
// table (already multiplied by d_lambda) come from the caller.
#include "gl_oracle_trace.h"

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

using namespace orc;

namespace {

struct Gate {
    float x_pos, x_neg, y_pos, y_neg, t_mm, zs_mm, sig_wide, sig_tight, cos_g, sin_g, r0, max_kick, inv_scatter, gain;
    int n_scatter, lobe;
};

// GateConfig -> resolved parameters (gate.h:168-211).
Gate resolve(const gl_gate_params& gc, float half_w, float half_h) {
    auto clampf = [](float v, float lo, float hi) { return v < lo ? lo : (v > hi ? hi : v); };
    Gate g{};
    const float hw = half_w > 0.0f ? half_w : 1.0f, hh = half_h > 0.0f ? half_h : 1.0f;
    g.x_pos = std::fmax(hw + gc.offset_right_mm, 0.1f);
    g.x_neg = -std::fmax(hw + gc.offset_left_mm, 0.1f);
    g.y_pos = std::fmax(hh + gc.offset_top_mm, 0.1f);
    g.y_neg = -std::fmax(hh + gc.offset_bottom_mm, 0.1f);
    g.t_mm = std::fmax(gc.thickness_mm, 1.0e-4f);
    g.zs_mm = std::fmax(gc.standoff_mm, 0.0f);
    g.sig_wide = clampf(gc.roughness_rad, 0.0f, 0.5f);
    const float aniso = clampf(gc.groove_aniso, 1.0e-3f, 1.0f);
    g.sig_tight = gc.lobe == GL_GATE_CAUCHY_ISO ? g.sig_wide : g.sig_wide * aniso;
    const float th = gc.groove_angle_deg * (float)(M_PI / 180.0);
    g.cos_g = std::cos(th);
    g.sin_g = std::sin(th);
    g.r0 = clampf(gc.reflectance_r0, 0.0f, 1.0f);
    g.max_kick = clampf(gc.max_kick_rad, 1.0e-4f, 1.5f);
    g.gain = std::fmax(gc.gain, 0.0f);
    int ns = gc.scatter_samples;
    ns = ns < 1 ? 1 : (ns > 64 ? 64 : ns);
    g.n_scatter = ns;
    g.inv_scatter = 1.0f / (float)ns;
    g.lobe = gc.lobe;
    return g;
}

uint32_t hash(uint32_t s) {
    s = (s ^ 61u) ^ (s >> 16); s *= 9u; s ^= s >> 4; s *= 0x27d4eb2du; s ^= s >> 15; return s;
}
float uniform(uint32_t& s) {
    s ^= s << 13; s ^= s >> 17; s ^= s << 5;
    return (float)(s & 0x00FFFFFFu) * (1.0f / 16777216.0f);
}
float cauchy(float sigma, float mk, uint32_t& s) {
    const float u = uniform(s);
    float k = sigma * std::tan(3.14159265358979323846f * (u - 0.5f));
    return k > mk ? mk : (k < -mk ? -mk : k);
}
float gauss(float sigma, float mk, uint32_t& s) {
    const float u1 = std::fmax(uniform(s), 1.0e-7f);
    const float u2 = uniform(s);
    float k = sigma * std::sqrt(-2.0f * std::log(u1)) * std::cos(2.0f * 3.14159265358979323846f * u2);
    return k > mk ? mk : (k < -mk ? -mk : k);
}

uint32_t bits(float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; }

}  // namespace

extern "C" {

int orc_render_gate(const gl_surface* s, int n, const float* arena, const gl_source* src, int n_src,
                    const gl_sample* grid, int n_grid, const gl_spec* spec, int n_spec, const gl_gate_params* gp,
                    int jitter_seed, float fp64_rmin, float half_w, float half_h, int width, int height, int fmt_w,
                    int fmt_h, int x0, int y0, float* out_r, float* out_g, float* out_b, uint64_t* counters,
                    float* dir_r, float* dir_g, float* dir_b, float* reach_out) {
    // dir_* / reach_out (may be null): the pass's telemetry
    // (gate_render.cu:158-172, 209-213): the direct image of the primary
    // traces (each landing's weight in the pixel it falls in) and the largest
    // realised |fold - edge| of a scrape.
    const size_t npx = (size_t)width * height;
    std::vector<double> R(npx, 0.0), G(npx, 0.0), B(npx, 0.0);
    std::vector<double> DR(dir_r ? npx : 0, 0.0), DG(dir_r ? npx : 0, 0.0), DB(dir_r ? npx : 0, 0.0);
    float reach = 0.0f;
    uint64_t traces = 0, scrapes = 0;
    if (n > 0 && n_grid > 0 && n_src > 0) {
        Lens L;
        lens_init(L, s, n, arena, fp64_rmin);
        std::vector<float> at((size_t)n_spec * n), bf((size_t)n_spec * n);
        ior_tables(s, n, spec, n_spec, at.data(), bf.data());
        const Gate g = resolve(*gp, half_w, half_h);
        const float front_R = s[0].semi_aperture, start_z = s[0].z - 20.0f;
        const float ray_weight = 1.0f / (float)n_grid;
        const uint32_t seed_base = (uint32_t)jitter_seed * 1000003u;
        const bool stochastic = g.sig_wide > 0.0f;
        const int n_k = stochastic ? g.n_scatter : 1;
        const float inv_k = stochastic ? g.inv_scatter : 1.0f;
        auto splat = [&](int x, int y, float w, float cr, float cg, float cb) {
            if (x < 0 || x >= width || y < 0 || y >= height) return;
            const size_t p = (size_t)y * width + x;
            if (std::fabs(cr) > 1e-16f) R[p] += (double)(cr * w);
            if (std::fabs(cg) > 1e-16f) G[p] += (double)(cg * w);
            if (std::fabs(cb) > 1e-16f) B[p] += (double)(cb * w);
        };
        for (int si_src = 0; si_src < n_src; ++si_src) {
            const gl_source& sc = src[si_src];
            const float bx = std::tan(sc.angle_x), by = std::tan(sc.angle_y);
            const V3<float> dir0 = normalized(V3<float>(bx, by, 1.0f));
            const float sdx = -20.0f * bx, sdy = -20.0f * by;
            for (int gi = 0; gi < n_grid; ++gi) {
                const uint32_t ray_seed = bits(sc.angle_x) * 0x9E3779B9u + bits(sc.angle_y) * 0x85EBCA6Bu +
                                          (uint32_t)gi * 0xC2B2AE35u + seed_base;
                for (int si = 0; si < n_spec; ++si) {
                    Ray<float> r;
                    r.o = V3<float>(grid[gi].u * front_R + sdx, grid[gi].v * front_R + sdy, start_z);
                    r.d = dir0;
                    V3<float> d;
                    const TraceOut<float> res = trace_primary<float>(L, r, spec[si].lambda, &at[(size_t)si * n],
                                                                     false, &d);
                    ++traces;
                    if (res.status != GL_TRACE_OK) continue;
                    if (!std::isfinite(res.pos.x) || !std::isfinite(res.pos.y)) continue;
                    if (!(std::fabs(d.z) > 1.0e-12f)) continue;
                    const float w_base = res.weight * ray_weight;
                    if (dir_r) {
                        const float dpx = (res.pos.x / (2.0f * half_w) + 0.5f) * (float)fmt_w + (float)x0;
                        const float dpy = (res.pos.y / (2.0f * half_h) + 0.5f) * (float)fmt_h + (float)y0;
                        const int dxi = (int)std::floor(dpx), dyi = (int)std::floor(dpy);
                        if (dxi >= 0 && dxi < width && dyi >= 0 && dyi < height) {
                            const size_t pix = (size_t)dyi * width + dxi;
                            DR[pix] += (double)(sc.r * spec[si].cmf_r * w_base);
                            DG[pix] += (double)(sc.g * spec[si].cmf_g * w_base);
                            DB[pix] += (double)(sc.b * spec[si].cmf_b * w_base);
                        }
                    }
                    const float sx = d.x / d.z, sy = d.y / d.z;
                    for (int wall = 0; wall < 4; ++wall) {
                        const int axis = wall >> 1;
                        const float sgn = (wall & 1) ? -1.0f : 1.0f;
                        float edge, pa, sa, po, so, lo, hi, na;
                        if (axis == 0) {
                            edge = sgn > 0 ? g.x_pos : g.x_neg; pa = res.pos.x; sa = sx; na = d.x;
                            po = res.pos.y; so = sy; lo = g.y_neg; hi = g.y_pos;
                        } else {
                            edge = sgn > 0 ? g.y_pos : g.y_neg; pa = res.pos.y; sa = sy; na = d.y;
                            po = res.pos.x; so = sx; lo = g.x_neg; hi = g.x_pos;
                        }
                        const float out = sgn * sa;
                        if (!(out > 1.0e-12f)) continue;
                        const float over = sgn * (pa - edge);
                        if (!(over > 0.0f)) continue;
                        if (over >= (g.zs_mm + g.t_mm) * out) continue;
                        if (over <= g.zs_mm * out) continue;
                        const float back = over / out;
                        const float hit_other = po - so * back;
                        if (hit_other <= lo || hit_other >= hi) continue;
                        ++scrapes;
                        reach = std::fmax(reach, std::fabs((2.0f * edge - pa) - edge));
                        const float c = 1.0f - std::fmin(std::fmax(std::fabs(na), 0.0f), 1.0f);
                        const float c2 = c * c;
                        const float Rs = g.r0 + (1.0f - g.r0) * c2 * c2 * c;
                        const float w_hit = w_base * Rs * g.gain * inv_k;
                        if (!(w_hit > 0.0f)) continue;
                        const float cr = sc.r * spec[si].cmf_r * w_hit, cg = sc.g * spec[si].cmf_g * w_hit,
                                    cb = sc.b * spec[si].cmf_b * w_hit;
                        for (int k = 0; k < n_k; ++k) {
                            uint32_t st = hash(ray_seed ^ ((uint32_t)si * 0x9E3779B9u) ^
                                               ((uint32_t)wall * 0x85EBCA6Bu) ^ ((uint32_t)k * 0x27D4EB2Fu));
                            float ka = 0.0f, ko = 0.0f, kz = 0.0f;
                            if (stochastic) {
                                if (g.lobe == GL_GATE_CAUCHY_ISO) {
                                    ka = cauchy(g.sig_wide, g.max_kick, st);
                                    ko = cauchy(g.sig_wide, g.max_kick, st);
                                    kz = cauchy(g.sig_wide, g.max_kick, st);
                                } else {
                                    const float wide = cauchy(g.sig_wide, g.max_kick, st);
                                    const float tight = gauss(g.sig_tight, g.max_kick, st);
                                    ka = wide * g.cos_g - tight * g.sin_g;
                                    ko = wide * g.sin_g + tight * g.cos_g;
                                    kz = 0.5f * wide;
                                }
                            }
                            const float dax = -sgn * std::fabs(sa) + ka, dot = so + ko, dz = 1.0f + kz;
                            if (!(dz > 1.0e-6f)) continue;
                            const float la = edge + dax / dz * back, lot = hit_other + dot / dz * back;
                            const float lx = axis == 0 ? la : lot, ly = axis == 0 ? lot : la;
                            float px = (lx / (2.0f * half_w) + 0.5f) * (float)fmt_w + (float)x0;
                            float py = (ly / (2.0f * half_h) + 0.5f) * (float)fmt_h + (float)y0;
                            if (!std::isfinite(px) || !std::isfinite(py)) continue;
                            px = std::fmax(-2.0e9f, std::fmin(2.0e9f, px));
                            py = std::fmax(-2.0e9f, std::fmin(2.0e9f, py));
                            const int ix = (int)std::floor(px - 0.5f), iy = (int)std::floor(py - 0.5f);
                            const float fx = (px - 0.5f) - (float)ix, fy = (py - 0.5f) - (float)iy;
                            splat(ix, iy, (1.0f - fx) * (1.0f - fy), cr, cg, cb);
                            splat(ix + 1, iy, fx * (1.0f - fy), cr, cg, cb);
                            splat(ix, iy + 1, (1.0f - fx) * fy, cr, cg, cb);
                            splat(ix + 1, iy + 1, fx * fy, cr, cg, cb);
                        }
                    }
                }
            }
        }
    }
    for (size_t i = 0; i < npx; ++i) { out_r[i] = (float)R[i]; out_g[i] = (float)G[i]; out_b[i] = (float)B[i]; }
    if (dir_r)
        for (size_t i = 0; i < npx; ++i) { dir_r[i] = (float)DR[i]; dir_g[i] = (float)DG[i]; dir_b[i] = (float)DB[i]; }
    if (reach_out) *reach_out = reach;
    if (counters) { counters[0] = traces; counters[1] = scrapes; }
    return 0;
}

}  // extern "C"
