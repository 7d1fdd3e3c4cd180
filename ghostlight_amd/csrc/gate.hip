// gate.hip — film-gate scatter layer.
//
// The gate opening is cut through a metal plate, so each of its four edges is
// a wall parallel to the optical axis, standoff_mm ahead of the sensor and
// thickness_mm deep.  A primary ray that would land just outside the opening
// strikes a wall at grazing incidence, reflects (Schlick reflectance to the
// wall normal), is scattered by the wall's machining lobe and lands back in
// frame near the mirror fold 2*edge - p.  Model, parameters and sampling follow
// gate.h:33-211 and gate_render.cu:73-280.
//
// MI355X shape: one lane per (source, pupil sample, wavelength) -- every such
// ray is independent (each scatter draw is a pure hash of stable identities),
// so the wavelength loop of the reference becomes parallelism and a single
// point source fills the chip.  Only the rays that scrape a wall (a thin band
// of the pupil) splat, so the sparse deposits go straight to global float
// atomics.
#include "gl_runtime.h"

#include <algorithm>
#include <cmath>
#include <cstring>

static_assert(sizeof(gl_gate_debug) == 120, "gl_gate_debug is 120 bytes (ctypes mirrors it)");

namespace gl {

namespace {

// Resolved runtime parameters (GpuGate, gate.h:36-52).
struct GateDev {
    float x_pos, x_neg, y_pos, y_neg;
    float t_mm, zs_mm;
    float sig_wide, sig_tight;
    float cos_g, sin_g;
    float r0, max_kick, inv_scatter, gain;
    int n_scatter, lobe;
};

struct GateArgs {
    const DevSurf* S;
    const float* poly;
    const float* arena;
    int n;
    const gl_source* src;
    int n_src;
    const gl_sample* grid;
    int n_grid;
    const float4* spec;   // lambda, cmf_r, cmf_g, cmf_b (already times d_lambda)
    int n_spec;
    const float* iat;     // n_spec x n
    float front_R, start_z, ray_weight;
    unsigned seed_base;
    float half_w, half_h;
    int fmt_w, fmt_h, x0, y0, width, height;
    float *out_r, *out_g, *out_b;
    unsigned long long* counters;   // [traces, scrapes] or null
    float *dir_r, *dir_g, *dir_b;   // direct image of the primary traces (telemetry) or null
    unsigned* reach;                // largest |fold - edge| of a scrape, float bits (telemetry) or null
    GateDev g;
};

// 32-bit xorshift uniform in [0, 1) (gate.h:66-72).
__host__ __device__ __forceinline__ float next_uniform(unsigned& s) {
    s ^= s << 13;
    s ^= s >> 17;
    s ^= s << 5;
    return (float)(s & 0x00FFFFFFu) * (1.0f / 16777216.0f);
}

// Clamped Cauchy (inverse CDF) and clamped Box-Muller Gaussian (gate.h:77-98).
__host__ __device__ __forceinline__ float draw_cauchy(float sigma, float max_kick, unsigned& s) {
    const float u = next_uniform(s);
    const float k = sigma * tanf(3.14159265358979323846f * (u - 0.5f));
    return fminf(fmaxf(k, -max_kick), max_kick);
}
__host__ __device__ __forceinline__ float draw_gauss(float sigma, float max_kick, unsigned& s) {
    const float u1 = fmaxf(next_uniform(s), 1.0e-7f);
    const float u2 = next_uniform(s);
    const float k = sigma * sqrtf(-2.0f * logf(u1)) * cosf(2.0f * 3.14159265358979323846f * u2);
    return fminf(fmaxf(k, -max_kick), max_kick);
}

// One lobe draw in the wall's own frame: along the wall normal, along the
// wall, down the axis (gate.h:116-131).
__host__ __device__ __forceinline__ void draw_lobe(const GateDev& g, unsigned& s, float& d_axis, float& d_other, float& d_z) {
    if (g.lobe == GL_GATE_CAUCHY_ISO) {
        d_axis = draw_cauchy(g.sig_wide, g.max_kick, s);
        d_other = draw_cauchy(g.sig_wide, g.max_kick, s);
        d_z = draw_cauchy(g.sig_wide, g.max_kick, s);
        return;
    }
    const float wide = draw_cauchy(g.sig_wide, g.max_kick, s);
    const float tight = draw_gauss(g.sig_tight, g.max_kick, s);
    d_axis = wide * g.cos_g - tight * g.sin_g;
    d_other = wide * g.sin_g + tight * g.cos_g;
    d_z = 0.5f * wide;
}

// Does a ray landing at p with slope s cross the wall at plane coordinate
// `edge` (sign sgn) between standoff and standoff + thickness?  (gate.h:146-159)
__device__ __forceinline__ bool wall_hit(const GateDev& g, float p, float s, float edge, float sgn, float& back,
                                         float& fold) {
    const float out = sgn * s;
    if (!(out > 1.0e-12f)) return false;
    const float over = sgn * (p - edge);
    if (!(over > 0.0f)) return false;
    if (over >= (g.zs_mm + g.t_mm) * out) return false;
    if (over <= g.zs_mm * out) return false;
    back = over / out;
    fold = 2.0f * edge - p;
    return true;
}

__device__ __forceinline__ void splat_px(const GateArgs& A, int x, int y, float w, float cr, float cg, float cb) {
    if (x < 0 || x >= A.width || y < 0 || y >= A.height) return;
    const size_t p = (size_t)y * A.width + x;
    if (fabsf(cr) > 1e-16f) atomicAdd(&A.out_r[p], cr * w);
    if (fabsf(cg) > 1e-16f) atomicAdd(&A.out_g[p], cg * w);
    if (fabsf(cb) > 1e-16f) atomicAdd(&A.out_b[p], cb * w);
}

__global__ __launch_bounds__(256) void gate_kernel(GateArgs A) {
    const int src_idx = blockIdx.y;
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long long)A.n_grid * A.n_spec) return;
    const int gi = (int)(idx / A.n_spec), si = (int)(idx % A.n_spec);
    const gl_source src = A.src[src_idx];
    const gl_sample gs = A.grid[gi];
    const GateDev& g = A.g;

    const float bx = tanf(src.angle_x), by = tanf(src.angle_y);
    const F3 dir0 = normalized(f3(bx, by, 1.0f));
    const float sdx = -20.0f * bx, sdy = -20.0f * by;
    const unsigned ray_seed = __float_as_uint(src.angle_x) * 0x9E3779B9u + __float_as_uint(src.angle_y) * 0x85EBCA6Bu +
                              (unsigned)gi * 0xC2B2AE35u + A.seed_base;
    const bool stochastic = g.sig_wide > 0.0f;
    const int n_k = stochastic ? g.n_scatter : 1;
    const float inv_k = stochastic ? g.inv_scatter : 1.0f;

    const float4 sp = A.spec[si];
    F3 d;
    const TraceRes res = trace_primary<true>(as_const(A.S), A.n, as_const(A.poly), A.arena,
                                             f3(gs.u * A.front_R + sdx, gs.v * A.front_R + sdy, A.start_z), dir0,
                                             sp.x, as_const(A.iat + (size_t)si * A.n), false, &d);
    if (A.counters) atomicAdd(&A.counters[0], 1ull);
    if (res.status != GL_TRACE_OK) return;
    if (!isfinite(res.px) || !isfinite(res.py)) return;
    if (!(fabsf(d.z) > 1.0e-12f)) return;
    const float w_base = res.w * A.ray_weight;
    // Telemetry: the direct image of the same traces, each landing added to
    // the pixel it falls in (gate_render.cu:158-172).
    if (A.dir_r) {
        const float dpx = (res.px / (2.0f * A.half_w) + 0.5f) * (float)A.fmt_w + (float)A.x0;
        const float dpy = (res.py / (2.0f * A.half_h) + 0.5f) * (float)A.fmt_h + (float)A.y0;
        const int dxi = (int)floorf(dpx), dyi = (int)floorf(dpy);
        if (dxi >= 0 && dxi < A.width && dyi >= 0 && dyi < A.height) {
            const size_t pix = (size_t)dyi * A.width + dxi;
            atomicAdd(&A.dir_r[pix], src.r * sp.y * w_base);
            atomicAdd(&A.dir_g[pix], src.g * sp.z * w_base);
            atomicAdd(&A.dir_b[pix], src.b * sp.w * w_base);
        }
    }
    const float sx = d.x / d.z, sy = d.y / d.z;

    for (int wall = 0; wall < 4; ++wall) {
        const int axis = wall >> 1;
        const float sgn = (wall & 1) ? -1.0f : 1.0f;
        float edge, p_axis, s_axis, p_other, s_other, o_lo, o_hi, n_axis;
        if (axis == 0) {
            edge = sgn > 0.0f ? g.x_pos : g.x_neg;
            p_axis = res.px; s_axis = sx; n_axis = d.x;
            p_other = res.py; s_other = sy;
            o_lo = g.y_neg; o_hi = g.y_pos;
        } else {
            edge = sgn > 0.0f ? g.y_pos : g.y_neg;
            p_axis = res.py; s_axis = sy; n_axis = d.y;
            p_other = res.px; s_other = sx;
            o_lo = g.x_neg; o_hi = g.x_pos;
        }
        float back, fold;
        if (!wall_hit(g, p_axis, s_axis, edge, sgn, back, fold)) continue;
        // Where the ray meets the wall along its length; beyond the opening on
        // that axis it misses the plate.
        const float hit_other = p_other - s_other * back;
        if (hit_other <= o_lo || hit_other >= o_hi) continue;
        if (A.counters) atomicAdd(&A.counters[1], 1ull);
        if (A.reach) atomicMax(A.reach, __float_as_uint(fabsf(fold - edge)));
        // Schlick to the wall normal: d is unit, so |n_axis| is the cosine.
        const float c = 1.0f - fminf(fmaxf(fabsf(n_axis), 0.0f), 1.0f);
        const float c2 = c * c;
        const float R = g.r0 + (1.0f - g.r0) * c2 * c2 * c;
        const float w_hit = w_base * R * g.gain * inv_k;
        if (!(w_hit > 0.0f)) continue;
        const float cr = src.r * sp.y * w_hit, cg = src.g * sp.z * w_hit, cb = src.b * sp.w * w_hit;
        for (int k = 0; k < n_k; ++k) {
            unsigned st = wang_hash(ray_seed ^ ((unsigned)si * 0x9E3779B9u) ^ ((unsigned)wall * 0x85EBCA6Bu) ^
                                    ((unsigned)k * 0x27D4EB2Fu));
            float k_axis = 0.0f, k_other = 0.0f, k_z = 0.0f;
            if (stochastic) draw_lobe(g, st, k_axis, k_other, k_z);
            const float d_ax = -sgn * fabsf(s_axis) + k_axis;
            const float d_ot = s_other + k_other;
            const float d_z = 1.0f + k_z;
            if (!(d_z > 1.0e-6f)) continue;
            const float land_axis = edge + d_ax / d_z * back;
            const float land_other = hit_other + d_ot / d_z * back;
            const float lx = axis == 0 ? land_axis : land_other;
            const float ly = axis == 0 ? land_other : land_axis;
            float px = (lx / (2.0f * A.half_w) + 0.5f) * (float)A.fmt_w + (float)A.x0;
            float py = (ly / (2.0f * A.half_h) + 0.5f) * (float)A.fmt_h + (float)A.y0;
            if (!isfinite(px) || !isfinite(py)) continue;
            px = fmaxf(-2.0e9f, fminf(2.0e9f, px));
            py = fmaxf(-2.0e9f, fminf(2.0e9f, py));
            const int ix = (int)floorf(px - 0.5f), iy = (int)floorf(py - 0.5f);
            const float fx = (px - 0.5f) - (float)ix, fy = (py - 0.5f) - (float)iy;
            splat_px(A, ix, iy, (1.0f - fx) * (1.0f - fy), cr, cg, cb);
            splat_px(A, ix + 1, iy, fx * (1.0f - fy), cr, cg, cb);
            splat_px(A, ix, iy + 1, (1.0f - fx) * fy, cr, cg, cb);
            splat_px(A, ix + 1, iy + 1, fx * fy, cr, cg, cb);
        }
    }
}

__global__ __launch_bounds__(256) void gate_zero_kernel(float* a, size_t n, unsigned long long* ctr,
                                                        unsigned* reach) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = 0.0f;
    if (ctr && i < 2) ctr[i] = 0ull;
    if (reach && i == 0) *reach = 0u;
}

// build_gpu_gate (gate.h:168-211): the opening from the effective sensor
// extent and the per-side offsets, every clamp applied.
GateDev resolve_gate(const gl_gate_params& gc, float half_w, float half_h) {
    auto clampf = [](float v, float lo, float hi) { return v < lo ? lo : (v > hi ? hi : v); };
    GateDev g{};
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
    g.n_scatter = std::min(64, std::max(1, gc.scatter_samples));
    g.inv_scatter = 1.0f / (float)g.n_scatter;
    g.lobe = gc.lobe;
    return g;
}

}  // namespace

// The pass behind gl_render_gate and gl_render_gate_debug (dbg may be null).
static int render_gate(gl_ctx* ctx, const gl_lens* lens, const gl_source* sources, int32_t n_sources, int32_t width,
                       int32_t height, int32_t fmt_w, int32_t fmt_h, int32_t fmt_x0, int32_t fmt_y0,
                       float sensor_half_w, float sensor_half_h, const gl_flare_params* rp, const gl_gate_params* gp,
                       float* out_r, float* out_g, float* out_b, uint64_t* counters_out, gl_gate_debug* dbg) {
    if (!lens || !rp || !gp || !out_r || !out_g || !out_b) return set_error(GL_ERR_INVALID, "null argument");
    if (width <= 0 || height <= 0 || fmt_w <= 0 || fmt_h <= 0)
        return set_error(GL_ERR_INVALID, "render_gate: degenerate output dimensions");
    if ((long long)width * height > (long long)INT32_MAX) return set_error(GL_ERR_INVALID, "render_gate: too many pixels");
    if (!(sensor_half_w > 0.0f) || !(sensor_half_h > 0.0f))
        return set_error(GL_ERR_INVALID, "render_gate: degenerate sensor half-size");
    if (n_sources < 0 || (n_sources && !sources)) return set_error(GL_ERR_INVALID, "bad source array");
    if (n_sources > 65535) return set_error(GL_ERR_INVALID, "render_gate: at most 65535 sources per call");
    if (gp->lobe != GL_GATE_CAUCHY_ISO && gp->lobe != GL_GATE_GROOVED)
        return set_error(GL_ERR_INVALID, "render_gate: unknown lobe %d", gp->lobe);
    const size_t npx = (size_t)width * height;
    std::fill(out_r, out_r + npx, 0.0f); std::fill(out_g, out_g + npx, 0.0f); std::fill(out_b, out_b + npx, 0.0f);
    if (counters_out) counters_out[0] = counters_out[1] = 0;
    if (dbg) {
        std::memset(&dbg->gate, 0, sizeof dbg->gate);
        dbg->traces = dbg->scrapes = 0;
        dbg->energy = 0.0;
        dbg->reach_mm = 0.0f;
        if (dbg->direct_r) {
            std::fill(dbg->direct_r, dbg->direct_r + npx, 0.0f);
            std::fill(dbg->direct_g, dbg->direct_g + npx, 0.0f);
            std::fill(dbg->direct_b, dbg->direct_b + npx, 0.0f);
        }
    }
    if (n_sources == 0) return GL_OK;
    int st;
    gl_ctx* c = resolve_ctx(ctx, &st);
    if (!c) return st;
    std::lock_guard<std::mutex> g(c->mu);
    gl_lens* L = const_cast<gl_lens*>(lens);
    const int n = (int)L->host.size();
    if (n <= 0) return set_error(GL_ERR_INVALID, "render_gate: lens has no surfaces");
    LensDev D;
    if ((st = lens_sync(L, rp->fp64_radius_min, c->stream, &D))) return st;

    // Pupil grid and the spectral table pre-multiplied by d_lambda
    // (gate_render.cu:320-360).
    std::vector<gl_sample> grid = build_samples(L->host.data(), n, rp->ray_grid, rp->pupil_jitter, rp->jitter_seed,
                                                rp->aperture_blades, rp->aperture_rotation);
    if (grid.empty()) return GL_OK;
    const int ns = std::max(3, gp->spectral_samples > 0 ? gp->spectral_samples : rp->spectral_samples);
    std::vector<gl_spec> spec;
    if ((st = spectral_table(ns, rp->output_cs, rp->custom_xyz_to_output, spec))) return st;
    if (ns != 3) {
        const float dl = (700.0f - 400.0f) / (float)ns;
        for (auto& s : spec) { s.cmf_r *= dl; s.cmf_g *= dl; s.cmf_b *= dl; }
    }
    std::vector<float> iat((size_t)ns * n), ibf((size_t)ns * n);
    ior_tables(L->host.data(), n, spec.data(), ns, iat.data(), ibf.data());
    std::vector<float4> spec4(ns);
    for (int i = 0; i < ns; ++i) spec4[i] = make_float4(spec[i].lambda, spec[i].cmf_r, spec[i].cmf_g, spec[i].cmf_b);

    DevBuf &bs = c->buf("gate_src"), &bg = c->buf("gate_grid"), &bsp = c->buf("gate_spec"), &bi = c->buf("gate_iat");
    DevBuf &bo = c->buf("gate_out"), &bc = c->buf("gate_ctr");
    if ((st = bs.ensure((size_t)n_sources * sizeof(gl_source))) || (st = bg.ensure(grid.size() * sizeof(gl_sample))) ||
        (st = bsp.ensure(spec4.size() * sizeof(float4))) || (st = bi.ensure(iat.size() * 4)) ||
        (st = bo.ensure((dbg && dbg->direct_r ? 6 : 3) * npx * 4)) || (st = bc.ensure(32)))
        return st;
    GL_HIP(hipMemcpyAsync(bs.p, sources, (size_t)n_sources * sizeof(gl_source), hipMemcpyHostToDevice, c->stream));
    GL_HIP(hipMemcpyAsync(bg.p, grid.data(), grid.size() * sizeof(gl_sample), hipMemcpyHostToDevice, c->stream));
    GL_HIP(hipMemcpyAsync(bsp.p, spec4.data(), spec4.size() * sizeof(float4), hipMemcpyHostToDevice, c->stream));
    GL_HIP(hipMemcpyAsync(bi.p, iat.data(), iat.size() * 4, hipMemcpyHostToDevice, c->stream));
    float* o = bo.as<float>();
    const bool direct = dbg && dbg->direct_r;
    const size_t planes = direct ? 6 : 3;
    unsigned long long* ctr = (counters_out || dbg) ? bc.as<unsigned long long>() : nullptr;
    unsigned* reach = dbg ? reinterpret_cast<unsigned*>(bc.as<char>() + 16) : nullptr;
    hipLaunchKernelGGL(gate_zero_kernel, dim3((unsigned)((planes * npx + 255) / 256)), dim3(256), 0, c->stream, o,
                       planes * npx, ctr, reach);

    GateArgs A;
    std::memset(&A, 0, sizeof A);
    A.S = D.surf; A.poly = D.poly; A.arena = D.arena; A.n = n;
    A.src = bs.as<gl_source>(); A.n_src = n_sources;
    A.grid = bg.as<gl_sample>(); A.n_grid = (int)grid.size();
    A.spec = bsp.as<float4>(); A.n_spec = ns; A.iat = bi.as<float>();
    A.front_R = L->host[0].semi_aperture; A.start_z = L->host[0].z - 20.0f;
    A.ray_weight = 1.0f / (float)grid.size();
    A.seed_base = (unsigned)rp->jitter_seed * 1000003u;
    A.half_w = sensor_half_w; A.half_h = sensor_half_h;
    A.fmt_w = fmt_w; A.fmt_h = fmt_h; A.x0 = fmt_x0; A.y0 = fmt_y0; A.width = width; A.height = height;
    A.out_r = o; A.out_g = o + npx; A.out_b = o + 2 * npx;
    A.counters = ctr;
    if (direct) { A.dir_r = o + 3 * npx; A.dir_g = o + 4 * npx; A.dir_b = o + 5 * npx; }
    A.reach = reach;
    A.g = resolve_gate(*gp, sensor_half_w, sensor_half_h);
    const long long lanes = (long long)A.n_grid * ns;
    hipLaunchKernelGGL(gate_kernel, dim3((unsigned)((lanes + 255) / 256), (unsigned)n_sources), dim3(256), 0, c->stream,
                       A);
    GL_HIP(hipGetLastError());
    GL_HIP(hipMemcpyAsync(out_r, o, npx * 4, hipMemcpyDeviceToHost, c->stream));
    GL_HIP(hipMemcpyAsync(out_g, o + npx, npx * 4, hipMemcpyDeviceToHost, c->stream));
    GL_HIP(hipMemcpyAsync(out_b, o + 2 * npx, npx * 4, hipMemcpyDeviceToHost, c->stream));
    unsigned long long ctr_h[2] = {0ull, 0ull};
    unsigned reach_h = 0u;
    if (ctr) GL_HIP(hipMemcpyAsync(ctr_h, ctr, 16, hipMemcpyDeviceToHost, c->stream));
    if (reach) GL_HIP(hipMemcpyAsync(&reach_h, reach, 4, hipMemcpyDeviceToHost, c->stream));
    if (direct) {
        GL_HIP(hipMemcpyAsync(dbg->direct_r, o + 3 * npx, npx * 4, hipMemcpyDeviceToHost, c->stream));
        GL_HIP(hipMemcpyAsync(dbg->direct_g, o + 4 * npx, npx * 4, hipMemcpyDeviceToHost, c->stream));
        GL_HIP(hipMemcpyAsync(dbg->direct_b, o + 5 * npx, npx * 4, hipMemcpyDeviceToHost, c->stream));
    }
    GL_HIP(hipStreamSynchronize(c->stream));
    if (counters_out) { counters_out[0] = ctr_h[0]; counters_out[1] = ctr_h[1]; }
    if (dbg) {
        static_assert(sizeof(GateDev) == sizeof(gl_gate_resolved), "gl_gate_resolved mirrors GateDev");
        std::memcpy(&dbg->gate, &A.g, sizeof A.g);
        dbg->traces = ctr_h[0];
        dbg->scrapes = ctr_h[1];
        std::memcpy(&dbg->reach_mm, &reach_h, sizeof reach_h);
        // The layer's energy as gate_render.cu:503-505 sums it.
        double e = 0.0;
        for (size_t i = 0; i < npx; ++i) e += out_r[i] + out_g[i] + out_b[i];
        dbg->energy = e;
    }
    return GL_OK;
}

}  // namespace gl

using namespace gl;

extern "C" {

int gl_render_gate(gl_ctx* ctx, const gl_lens* lens, const gl_source* sources, int32_t n_sources, int32_t width,
                   int32_t height, int32_t fmt_w, int32_t fmt_h, int32_t fmt_x0, int32_t fmt_y0, float sensor_half_w,
                   float sensor_half_h, const gl_flare_params* rp, const gl_gate_params* gp, float* out_r,
                   float* out_g, float* out_b, uint64_t* counters_out) {
    return render_gate(ctx, lens, sources, n_sources, width, height, fmt_w, fmt_h, fmt_x0, fmt_y0, sensor_half_w,
                       sensor_half_h, rp, gp, out_r, out_g, out_b, counters_out, nullptr);
}

int gl_render_gate_debug(gl_ctx* ctx, const gl_lens* lens, const gl_source* sources, int32_t n_sources,
                         int32_t width, int32_t height, int32_t fmt_w, int32_t fmt_h, int32_t fmt_x0, int32_t fmt_y0,
                         float sensor_half_w, float sensor_half_h, const gl_flare_params* rp,
                         const gl_gate_params* gp, float* out_r, float* out_g, float* out_b, gl_gate_debug* dbg) {
    if (!dbg) return set_error(GL_ERR_INVALID, "null debug record");
    if (dbg->direct_r && (!dbg->direct_g || !dbg->direct_b))
        return set_error(GL_ERR_INVALID, "direct planes: all three or none");
    return render_gate(ctx, lens, sources, n_sources, width, height, fmt_w, fmt_h, fmt_x0, fmt_y0, sensor_half_w,
                       sensor_half_h, rp, gp, out_r, out_g, out_b, nullptr, dbg);
}


int gl_gate_resolve(const gl_gate_params* gate, float half_w, float half_h, gl_gate_resolved* out) {
    if (!gate || !out) return set_error(GL_ERR_INVALID, "null argument");
    static_assert(sizeof(GateDev) == sizeof(gl_gate_resolved), "gl_gate_resolved mirrors GateDev");
    const GateDev g = resolve_gate(*gate, half_w, half_h);
    std::memcpy(out, &g, sizeof g);
    return GL_OK;
}

int gl_gate_sample(const gl_gate_params* gate, int32_t n, uint32_t seed, float* axis, float* other, float* z) {
    if (!gate || n < 0 || (n && (!axis || !other || !z))) return set_error(GL_ERR_INVALID, "null argument");
    const GateDev g = resolve_gate(*gate, 1.0f, 1.0f);
    for (int i = 0; i < n; ++i) {
        unsigned st = wang_hash(seed ^ (2654435761u * (uint32_t)i + 1u));
        if (st == 0u) st = 1u;
        draw_lobe(g, st, axis[i], other[i], z[i]);
    }
    return GL_OK;
}

}  // extern "C"
