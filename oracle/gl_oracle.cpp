// ORACLE — test infrastructure only.  Loaded by tests/, __graft_entry__.smoke()
// and bench.py's cpu_baseline leg as the CHECKER (ctypes, liboracle.so built by
// oracle/Makefile); never linked into, called by or shipped with the product.
//
// gl_oracle.cpp — CPU restatement of the reference's ghost path:
//   pupil sampler        ghostlight/src/aperture_sampler.h:42-223
//   IOR tables           ghostlight/src/ghost_render.cu:809-823
//   pair enumeration     ghostlight/src/ghost.cpp:17-30
//   pair pre-filter      ghostlight/src/ghost.cpp:34-138
//   ghost_kernel body    ghostlight/src/ghost_render.cu:249-490 (STATS and splat)
//   launcher semantics   ghostlight/src/ghost_render.cu:700-1353 (probe passes,
//                        refine, rects, adaptive budgets, spectral correction)
// Scalar type float = bit-exact oracle of the device path (statuses, masks,
// cull sets, landing pixels); double = FP64 image oracle.
//
// Parity pins (tests/): every function here is checked against the
// reference's own CPU code compiled into oracle/_ref (trace_ghost_ray,
// filter_ghost_pairs, build_aperture_samples) where it exists, against golden
// vectors in tests/golden/, and against closed forms (Fresnel normal
// incidence, energy conservation of the bilinear splat, linearity).
// The probe / concentration launcher logic has no CPU counterpart in the
// reference (it is CUDA host code) and is pinned by the cull-on == cull-off
// property the reference's validate_culling.py gates on.
#include "gl_oracle_trace.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

using namespace orc;

namespace {

// Planted faults for the tests that show the statistical gates of the
// concentrated estimator catch real mistakes (0 = none):
//   2  survivor rects shrunk by 10% of their extent on every side
//   3  adaptive budgets trace one 256-sample chunk fewer than they weight for
//   4  concentrated weights drop the rect-area factor
int g_plant_fault = 0;

// ---- pupil sampler (aperture_sampler.h) -----------------------------------
struct PMask {
    float aspect = 1.f;
    PolyPre poly;
};

PMask resolve_mask(const gl_surface* s, int n, int blades_override, float rotation_deg) {
    int nb = 0; float rot = 0.f, aspect = 1.f;
    Blade blade;
    for (int i = 0; i < n; ++i)
        if (s[i].is_stop) {
            aspect = s[i].aperture_aspect;
            if (s[i].aperture_shape == GL_APERTURE_POLYGON && s[i].aperture_blades >= 3) {
                nb = s[i].aperture_blades; rot = s[i].aperture_rotation_rad;
                blade = orc_blade_of(s[i]);
            }
            break;
        }
    // An override is a plain n-gon (aperture_sampler.h:112-120).
    if (blades_override > 0) { nb = blades_override; rot = rotation_deg * ((float)M_PI / 180.0f); blade = Blade{}; }
    PMask m; m.aspect = aspect;
    poly_precompute(nb, rot, 1.0f, m.poly);
    m.poly.blade = blade;
    return m;
}

inline bool mask_contains(const PMask& m, float u, float v) {
    const float ux = u / m.aspect;
    const float r2 = ux * ux + v * v;
    if (r2 > 1.0f) return false;
    if (m.poly.n >= 3 && m.poly.blade.deformed)
        return std::sqrt(r2) <= blade_radius<float>(m.poly.blade, std::atan2(v, ux));
    if (m.poly.n >= 3) {
        float mx = ux * m.poly.c[0] + v * m.poly.s[0];
        for (int k = 1; k < m.poly.n; ++k) { float d = ux * m.poly.c[k] + v * m.poly.s[k]; mx = d > mx ? d : mx; }
        if (mx > m.poly.apothem) return false;
    }
    return true;
}

inline uint32_t wang_hash(uint32_t s) {
    s = (s ^ 61u) ^ (s >> 16u); s *= 9u; s ^= s >> 4u; s *= 0x27d4eb2du; s ^= s >> 15u; return s;
}
inline float halton2(uint32_t n) {
    n = (n << 16u) | (n >> 16u);
    n = ((n & 0x00ff00ffu) << 8u) | ((n & 0xff00ff00u) >> 8u);
    n = ((n & 0x0f0f0f0fu) << 4u) | ((n & 0xf0f0f0f0u) >> 4u);
    n = ((n & 0x33333333u) << 2u) | ((n & 0xccccccccu) >> 2u);
    n = ((n & 0x55555555u) << 1u) | ((n & 0xaaaaaaaau) >> 1u);
    return (float)n * (1.0f / 4294967296.0f);
}
inline float halton3(uint32_t n) {
    float r = 0.0f, f = 1.0f / 3.0f;
    while (n > 0) { r += (float)(n % 3u) * f; n /= 3u; f /= 3.0f; }
    return r;
}

std::vector<gl_sample> build_samples(const gl_surface* s, int n, int N, int jitter, int seed, int bo, float rdeg) {
    std::vector<gl_sample> out;
    if (N <= 0) return out;
    const PMask m = resolve_mask(s, n, bo, rdeg);
    const uint32_t so = (uint32_t)seed * 1000003u;
    for (int k = 0; k < N * N; ++k) {
        const int gx = k % N, gy = k / N;
        float u, v;
        if (jitter == 2) { u = halton2((uint32_t)k) * 2.0f - 1.0f; v = halton3((uint32_t)k) * 2.0f - 1.0f; }
        else {
            float ju = (jitter == 1) ? (float)wang_hash((uint32_t)k + so) / 4294967296.0f : 0.5f;
            float jv = (jitter == 1) ? (float)wang_hash((uint32_t)k + (uint32_t)(N * N) + so) / 4294967296.0f : 0.5f;
            u = (((float)gx + ju) / (float)N) * 2.0f - 1.0f;
            v = (((float)gy + jv) / (float)N) * 2.0f - 1.0f;
        }
        if (mask_contains(m, u, v)) out.push_back({u, v});
    }
    return out;
}

// ---- per-sample ghost work (ghost_render.cu:249-490) ----------------------
struct Frame {
    float sw, sh;              // sensor half extents used by the pixel map
    float fmt_w, fmt_h, x0, y0;
    int width, height;
};

struct Acc {   // one thread's private accumulation (double, fixed order)
    std::vector<double> r, g, b;
};

struct Counters { uint64_t traces = 0, surv = 0, onsens = 0, segs = 0; };

// Concentration inputs (ConcParams, ghost_render.cu:175-189).
struct Conc {
    const float* rects = nullptr;  // 4 per ps
    PMask mask;
    int jitter = 2; uint32_t seed_offset = 0; int nr_lattice = 1; int n_r = 1;
    float area_norm = 0.f;
    const int* budgets = nullptr;  // 2 per ps
    float inv_aref = 0.f;
};

struct Job {
    Lens L;
    const gl_pair* pairs; int n_pairs;
    const gl_source* src; int n_src;
    const float* bxy;              // 2 per source: tan(angle) rounded once on the host
    const gl_spec* spec; int n_spec;
    const float* ior_at; const float* ior_before;
    float front_R, start_z;
    float gain, ray_weight;
    Hurb hurb;                     // edge diffraction (ghost_render.cu:358-392)
    uint32_t hurb_seed_base = 0;   // jitter_seed * 2654435761 + 1
    const gl_baffle* baffles = nullptr;
    int n_baffles = 0;
};

// The flare parameters' HURB switches and baffle stack into the job.
inline void job_diffraction(Job& J, const gl_flare_params* p) {
    J.hurb.on = p->hurb ? 1 : 0;
    J.hurb.dist = p->hurb_kick;
    J.hurb.min_sigma = p->hurb_min_sigma_rad;
    J.hurb.max_kick = p->hurb_max_kick_rad;
    J.hurb_seed_base = (uint32_t)p->jitter_seed * 2654435761u + 1u;
    J.baffles = p->baffles;
    J.n_baffles = std::max(0, std::min((int)p->n_baffles, GL_MAX_BAFFLES));
}

// Sample position + weight; returns false when the sample contributes nothing.
inline bool sample_pos(const Job& J, const Conc* conc, int ps, int gi, const gl_sample* grid, int n_grid,
                       float& u, float& v, float& w_ray) {
    if (gi >= n_grid) return false;
    w_ray = J.ray_weight;
    if (conc && conc->rects) {
        const float* rc = conc->rects + 4 * ps;
        if (rc[0] <= rc[2]) {
            int n_r_ps = conc->n_r, nr_lat = conc->nr_lattice;
            if (conc->budgets) { n_r_ps = conc->budgets[2 * ps]; nr_lat = conc->budgets[2 * ps + 1]; }
            if (gi >= n_r_ps) return false;
            if (g_plant_fault == 3 && conc->budgets && gi >= n_r_ps - 256) return false;
            float fu, fv;
            if (conc->jitter == 2) { fu = halton2((uint32_t)gi); fv = halton3((uint32_t)gi); }
            else {
                const int gx = gi % nr_lat, gy = gi / nr_lat, nn = nr_lat * nr_lat;
                const float ju = conc->jitter == 1 ? (float)wang_hash((uint32_t)gi + conc->seed_offset) / 4294967296.0f : 0.5f;
                const float jv = conc->jitter == 1 ? (float)wang_hash((uint32_t)gi + (uint32_t)nn + conc->seed_offset) / 4294967296.0f : 0.5f;
                fu = ((float)gx + ju) / (float)nr_lat;
                fv = ((float)gy + jv) / (float)nr_lat;
            }
            u = rc[0] + fu * (rc[2] - rc[0]);
            v = rc[1] + fv * (rc[3] - rc[1]);
            if (!mask_contains(conc->mask, u, v)) return false;
            w_ray = (rc[2] - rc[0]) * (rc[3] - rc[1]) *
                    (conc->budgets ? conc->inv_aref / (float)n_r_ps : conc->area_norm);
            if (g_plant_fault == 4) w_ray = 4.0f * (conc->budgets ? conc->inv_aref / (float)n_r_ps : conc->area_norm);
            return true;
        }
    }
    u = grid[gi].u; v = grid[gi].v;
    return true;
}

template <class S>
inline void splat(Acc& A, const Frame& F, S px, S py, S cr, S cg, S cb) {
    px = std::fmax(S(-2.0e9f), std::fmin(S(2.0e9f), px));
    py = std::fmax(S(-2.0e9f), std::fmin(S(2.0e9f), py));
    const int x0 = (int)std::floor(px - S(0.5f));
    const int y0 = (int)std::floor(py - S(0.5f));
    const S fx = (px - S(0.5f)) - S(x0), fy = (py - S(0.5f)) - S(y0);
    const S w[4] = {(S(1) - fx) * (S(1) - fy), fx * (S(1) - fy), (S(1) - fx) * fy, fx * fy};
    const int xs[4] = {x0, x0 + 1, x0, x0 + 1}, ys[4] = {y0, y0, y0 + 1, y0 + 1};
    for (int k = 0; k < 4; ++k) {
        if (xs[k] < 0 || xs[k] >= F.width || ys[k] < 0 || ys[k] >= F.height) continue;
        const size_t p = (size_t)ys[k] * F.width + xs[k];
        if (std::fabs(cr) > S(1e-14f)) A.r[p] += (double)(cr * w[k]);
        if (std::fabs(cg) > S(1e-14f)) A.g[p] += (double)(cg * w[k]);
        if (std::fabs(cb) > S(1e-14f)) A.b[p] += (double)(cb * w[k]);
    }
}

// One (ps, sample): the kernel body.  STATS mode tallies counters and the
// on-sensor flag; render mode splats.  Returns true when any lambda landed
// on the (margin) sensor (STATS bbox input).
template <class S, bool STATS>
bool do_sample(const Job& J, const Conc* conc, const Frame& F, int ps, int gi, const gl_sample* grid, int n_grid,
               const gl_spec* spec, int n_spec, const float* iat, const float* ibf,
               Acc* A, Counters& C, float* su, float* sv) {
    const int pair_idx = ps / J.n_src, src_idx = ps % J.n_src;
    const gl_pair& pr = J.pairs[pair_idx];
    if (pr.surf_a < 0 || pr.surf_b <= pr.surf_a || pr.surf_b >= J.L.n) return false;
    float u, v, w_ray;
    if (!sample_pos(J, conc, ps, gi, grid, n_grid, u, v, w_ray)) return false;
    if (su) { *su = u; *sv = v; }
    const gl_source& sr = J.src[src_idx];
    const float bx = J.bxy[2 * src_idx], by = J.bxy[2 * src_idx + 1];
    V3<float> bdf = normalized(V3<float>(bx, by, 1.0f));
    const float sdx = -20.0f * bx, sdy = -20.0f * by;
    Ray<S> base;
    base.o = V3<S>(S(u * J.front_R + sdx), S(v * J.front_R + sdy), S(J.start_z));
    if (sizeof(S) == 4) base.d = V3<S>(S(bdf.x), S(bdf.y), S(bdf.z));
    else base.d = normalized(V3<S>(S(bx), S(by), S(1)));
    // Front-of-lens baffles drop the ray before any wavelength (ghost_render.cu:351-356).
    const float ox = u * J.front_R + sdx, oy = v * J.front_R + sdy;
    if (J.n_baffles > 0 && baffles_block(J.baffles, J.n_baffles, ox, oy, bx, by)) return false;
    uint32_t hray = 0u;
    if (J.hurb.on) {
        uint32_t axb, ayb;
        std::memcpy(&axb, &sr.angle_x, 4);
        std::memcpy(&ayb, &sr.angle_y, 4);
        hray = axb * 0x9E3779B9u + ayb * 0x85EBCA6Bu + (uint32_t)gi * 0xC2B2AE35u +
               (uint32_t)pair_idx * 0x27D4EB2Fu + J.hurb_seed_base;
    }
    bool onsens_any = false;
    for (int si = 0; si < n_spec; ++si) {
        if (STATS) ++C.traces;
        const gl_spec& sp = spec[si];
        Ray<S> r = base;
        uint32_t hs = 0u;
        if (J.hurb.on) {
            hs = hurb_wang(hray + (uint32_t)si * 0x9E3779B9u);
            if (J.n_baffles > 0) {
                uint32_t bs = hurb_wang(hs ^ 0xBAFF1E5Du);
                V3<float> bn;
                const float bd = baffles_edge(J.baffles, J.n_baffles, ox, oy, bx, by, bn);
                hurb_apply<S>(J.hurb, r.d, V3<S>(S(bn.x), S(bn.y), S(bn.z)), S(bd), S(sp.lambda), bs);
            }
        }
        TraceOut<S> res = trace_ghost<S>(J.L, pr.surf_a, pr.surf_b, r, S(sp.lambda),
                                          iat + (size_t)si * J.L.n, ibf + (size_t)si * J.L.n, false,
                                          J.hurb.on ? &J.hurb : nullptr, &hs);
        C.segs += (uint64_t)res.segments;
        if (res.status != GL_TRACE_OK) continue;
        if (!std::isfinite(res.pos.x) || !std::isfinite(res.pos.y)) continue;
        if (STATS) {
            ++C.surv;
            S spx = (res.pos.x / (S(2) * S(F.sw)) + S(0.5f)) * S(F.fmt_w) + S(F.x0);
            S spy = (res.pos.y / (S(2) * S(F.sh)) + S(0.5f)) * S(F.fmt_h) + S(F.y0);
            if (std::isfinite(spx) && std::isfinite(spy) && spx >= S(0) && spx < S(F.width) &&
                spy >= S(0) && spy < S(F.height)) { ++C.onsens; onsens_any = true; }
            continue;
        }
        S base_w = res.weight * S(w_ray) * S(J.gain) * S(pr.area_boost);
        if (base_w < S(1e-14f)) continue;
        S px = (res.pos.x / (S(2) * S(F.sw)) + S(0.5f)) * S(F.fmt_w) + S(F.x0);
        S py = (res.pos.y / (S(2) * S(F.sh)) + S(0.5f)) * S(F.fmt_h) + S(F.y0);
        if (!std::isfinite(px) || !std::isfinite(py)) continue;
        S P = S(sr.r) * S(sp.cmf_r) + S(sr.g) * S(sp.cmf_g) + S(sr.b) * S(sp.cmf_b);
        S power = P * base_w;
        splat<S>(*A, F, px, py, power * S(sp.cmf_r), power * S(sp.cmf_g), power * S(sp.cmf_b));
    }
    return onsens_any;
}

struct ProbeOut {
    std::vector<uint64_t> pt, pss, po;     // per pair
    std::vector<float> umin, vmin, umax, vmax;
    std::vector<int> hits;                 // per ps
};

// One STATS pass over all ps (optionally masked per pair or per ps).
void probe_pass(const Job& J, const Frame& F, const gl_sample* grid, int n_grid,
                const gl_spec* spec, int n_spec, const float* iat, const float* ibf,
                const unsigned char* pair_mask, const unsigned char* ps_mask, bool want_bbox, ProbeOut& P,
                int threads) {
    const int n_ps = J.n_pairs * J.n_src;
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads)
    for (int ps = 0; ps < n_ps; ++ps) {
        const int pair_idx = ps / J.n_src;
        if (pair_mask && pair_mask[pair_idx] == 0) continue;
        if (ps_mask && ps_mask[ps] == 0) continue;
        Counters C;
        float umin = 1e30f, vmin = 1e30f, umax = -1e30f, vmax = -1e30f; int hits = 0;
        for (int gi = 0; gi < n_grid; ++gi) {
            float u = 0, v = 0;
            bool on = do_sample<float, true>(J, nullptr, F, ps, gi, grid, n_grid, spec, n_spec, iat, ibf,
                                             nullptr, C, &u, &v);
            if (on && want_bbox) {
                umin = std::min(umin, u); vmin = std::min(vmin, v);
                umax = std::max(umax, u); vmax = std::max(vmax, v); ++hits;
            }
        }
#pragma omp critical
        {
            P.pt[pair_idx] += C.traces; P.pss[pair_idx] += C.surv; P.po[pair_idx] += C.onsens;
            if (want_bbox && hits) {
                P.umin[ps] = std::min(P.umin[ps], umin); P.vmin[ps] = std::min(P.vmin[ps], vmin);
                P.umax[ps] = std::max(P.umax[ps], umax); P.vmax[ps] = std::max(P.vmax[ps], vmax);
                P.hits[ps] += hits;
            }
        }
    }
}

int threads_of(int t) {
#ifdef _OPENMP
    return t > 0 ? t : omp_get_max_threads();
#else
    (void)t; return 1;
#endif
}

}  // namespace

extern "C" {

void orc_plant_fault(int mode) { g_plant_fault = mode; }

// aperture_edge_distance (trace_core.h:206-290) of the oracle, float.
int orc_edge_distance(const gl_surface* s, const float* hit, float* d, float* nrm) {
    Lens L;
    lens_init(L, s, 1, nullptr, 0.0f);
    V3<float> n;
    *d = edge_distance<float>(V3<float>(hit[0], hit[1], hit[2]), s[0], L.poly[0], L.transformed[0], n);
    nrm[0] = n.x; nrm[1] = n.y; nrm[2] = n.z;
    return 0;
}

// HURB kicks of the oracle for the _hurb_sample_debug streams.
int orc_hurb_sample(float lam, const float* dd, int n, int dist, float max_kick, uint32_t seed, float* kicks,
                    float* sigma) {
    for (int i = 0; i < n; ++i) {
        const float sg = hurb_sigma<float>(lam, dd[i], dist);
        uint32_t st = hurb_wang(seed ^ (2654435761u * (uint32_t)i + 1u));
        if (st == 0u) st = 1u;
        kicks[i] = sg > 0.0f ? hurb_kick<float>(dist, sg, max_kick, st) : 0.0f;
        sigma[i] = sg;
    }
    return 0;
}

// The oracle's blade silhouette as gl_blade_profile, and its r(theta).
int orc_blade_profile(int blades, float rot, float curvature, float twist, float notch, float rake,
                      gl_blade_profile* out) {
    const Blade b = orc_blade_build(blades, rot, curvature, twist, notch, rake);
    *out = gl_blade_profile{};
    out->blades = b.blades; out->plain = (b.blades >= 3 && !b.deformed) || b.blades < 3 ? 1 : 0;
    out->facets = b.facets; out->rotation = b.rotation; out->sigma = b.sigma; out->half = b.half;
    out->r_w = b.r_w; out->phi_w = b.phi_w; out->u_w = b.u_w; out->u_rise = b.u_rise; out->qR = b.qR;
    out->qL = b.qL; out->p = b.p; out->beta = b.beta; out->eA = b.eA; out->eB = b.eB;
    out->area_frac = b.area_frac;
    return 0;
}
int orc_blade_radius(int blades, float rot, float curvature, float twist, float notch, float rake,
                     const float* theta, int n, float* out) {
    const Blade b = orc_blade_build(blades, rot, curvature, twist, notch, rake);
    for (int i = 0; i < n; ++i) out[i] = blade_radius<float>(b, theta[i]);
    return 0;
}

// Single ghost / primary ray trace.  dbl selects the scalar type.
int orc_trace_ghost(int dbl, const gl_surface* s, int n, const float* arena, int a, int b,
                    const float* ray7, const float* ior_at, const float* ior_before, int t_ge0,
                    float fp64_rmin, double* out_pos, double* out_w, int* out_status, int* out_segs) {
    Lens L; lens_init(L, s, n, arena, fp64_rmin);
    if (dbl) {
        Ray<double> r; r.o = V3<double>(ray7[0], ray7[1], ray7[2]); r.d = V3<double>(ray7[3], ray7[4], ray7[5]);
        TraceOut<double> o = trace_ghost<double>(L, a, b, r, ray7[6], ior_at, ior_before, t_ge0 != 0);
        out_pos[0] = o.pos.x; out_pos[1] = o.pos.y; out_pos[2] = o.pos.z; *out_w = o.weight;
        *out_status = o.status; if (out_segs) *out_segs = o.segments;
    } else {
        Ray<float> r; r.o = V3<float>(ray7[0], ray7[1], ray7[2]); r.d = V3<float>(ray7[3], ray7[4], ray7[5]);
        TraceOut<float> o = trace_ghost<float>(L, a, b, r, ray7[6], ior_at, ior_before, t_ge0 != 0);
        out_pos[0] = o.pos.x; out_pos[1] = o.pos.y; out_pos[2] = o.pos.z; *out_w = o.weight;
        *out_status = o.status; if (out_segs) *out_segs = o.segments;
    }
    return 0;
}

int orc_trace_primary(int dbl, const gl_surface* s, int n, const float* arena, const float* ray7,
                      const float* ior_at, int t_ge0, float fp64_rmin, double* out_pos, double* out_w,
                      int* out_status, int* out_segs) {
    Lens L; lens_init(L, s, n, arena, fp64_rmin);
    if (dbl) {
        Ray<double> r; r.o = V3<double>(ray7[0], ray7[1], ray7[2]); r.d = V3<double>(ray7[3], ray7[4], ray7[5]);
        TraceOut<double> o = trace_primary<double>(L, r, ray7[6], ior_at, t_ge0 != 0);
        out_pos[0] = o.pos.x; out_pos[1] = o.pos.y; out_pos[2] = o.pos.z; *out_w = o.weight; *out_status = o.status;
        if (out_segs) *out_segs = o.segments;
    } else {
        Ray<float> r; r.o = V3<float>(ray7[0], ray7[1], ray7[2]); r.d = V3<float>(ray7[3], ray7[4], ray7[5]);
        TraceOut<float> o = trace_primary<float>(L, r, ray7[6], ior_at, t_ge0 != 0);
        out_pos[0] = o.pos.x; out_pos[1] = o.pos.y; out_pos[2] = o.pos.z; *out_w = o.weight; *out_status = o.status;
        if (out_segs) *out_segs = o.segments;
    }
    return 0;
}

// FP32 primary survival of n launches (x, y) on the plane z0, all heading
// normalize(tan ax, tan ay, 1) (float tangents and Vec3f::normalized, as the
// reference's host code forms them) at one wavelength, with the CPU tracer's
// t >= 0 rule: 1 when the ray reaches the sensor with a finite landing.  The
// survivor centroid and the cat's-eye envelope of the starburst pupil
// (starburst_render.cu:393-582) are built from these answers.
int orc_primary_survival(const gl_surface* s, int n, const float* arena, const float* launch, int n_launch,
                         float ax, float ay, float z0, float lambda_nm, const float* ior_at,
                         unsigned char* out, int threads) {
    Lens L; lens_init(L, s, n, arena, 0.0f);
    const float bx = std::tan(ax), by = std::tan(ay);
    const float len = std::sqrt(bx * bx + by * by + 1.0f * 1.0f);
    const float inv = len > 1e-12f ? 1.0f / len : 0.0f;
    const float dx = bx * inv, dy = by * inv, dz = 1.0f * inv;
#pragma omp parallel for schedule(static) num_threads(threads > 0 ? threads : omp_get_max_threads())
    for (int i = 0; i < n_launch; ++i) {
        Ray<float> r;
        r.o = V3<float>(launch[2 * i], launch[2 * i + 1], z0);
        r.d = V3<float>(dx, dy, dz);
        TraceOut<float> o = trace_primary<float>(L, r, lambda_nm, ior_at, true);
        out[i] = (o.status == GL_TRACE_OK && std::isfinite(o.pos.x) && std::isfinite(o.pos.y)) ? 1 : 0;
    }
    return 0;
}

int orc_build_samples(const gl_surface* s, int n, int N, int jitter, int seed, int bo, float rdeg,
                      gl_sample* out, int cap, int* n_out) {
    std::vector<gl_sample> g = build_samples(s, n, N, jitter, seed, bo, rdeg);
    *n_out = (int)g.size();
    if ((int)g.size() > cap) return -1;
    std::copy(g.begin(), g.end(), out);
    return 0;
}

int orc_ior_tables(const gl_surface* s, int n, const gl_spec* spec, int ns, float* at, float* bf) {
    ior_tables(s, n, spec, ns, at, bf);
    return 0;
}

// filter_ghost_pairs (ghost.cpp:105-138) with the CPU tracer's semantics.
int orc_filter_ghost_pairs(const gl_surface* s, int n, const float* arena, float half_w, float half_h,
                           float min_intensity, int ghost_normalize, float max_area_boost,
                           gl_pair* out, int cap, int* n_out) {
    Lens L; lens_init(L, s, n, arena, 0.0f);
    const float probe_l[3] = {650.0f, 550.0f, 450.0f};
    std::vector<float> at(3 * (size_t)n), bf(3 * (size_t)n);
    gl_spec sp[3] = {{650.f, 0, 0, 0}, {550.f, 0, 0, 0}, {450.f, 0, 0, 0}};
    ior_tables(s, n, sp, 3, at.data(), bf.data());
    auto ior_before_d = [&](int idx) {
        for (int k = idx - 1; k >= 0; --k) if (s[k].is_active) return s[k].ior;
        return 1.0f;
    };
    int cnt = 0;
    for (int a = 0; a < n; ++a) {
        if (!s[a].is_active) continue;
        for (int b = a + 1; b < n; ++b) {
            if (!s[b].is_active) continue;
            if (std::fabs(ior_before_d(a) - s[a].ior) < 0.001f || std::fabs(ior_before_d(b) - s[b].ior) < 0.001f)
                continue;
            float total = 0.0f;
            for (int ch = 0; ch < 3; ++ch) {
                Ray<float> r; r.o = V3<float>(0, 0, s[0].z - 20.0f); r.d = V3<float>(0, 0, 1);
                TraceOut<float> o = trace_ghost<float>(L, a, b, r, probe_l[ch], &at[ch * n], &bf[ch * n], true);
                if (o.status == GL_TRACE_OK) total += o.weight;
            }
            float est = total / 3.0f;
            if (est < min_intensity) continue;
            float boost = 1.0f;
            if (ghost_normalize) {
                const int G = 8;
                const float fR = s[0].semi_aperture, z0 = s[0].z - 20.0f;
                float mnx = 1e30f, mxx = -1e30f, mny = 1e30f, mxy = -1e30f; int valid = 0;
                for (int gy = 0; gy < G; ++gy)
                    for (int gx = 0; gx < G; ++gx) {
                        float u = (((float)gx + 0.5f) / (float)G) * 2.0f - 1.0f;
                        float v = (((float)gy + 0.5f) / (float)G) * 2.0f - 1.0f;
                        if (u * u + v * v > 1.0f) continue;
                        Ray<float> r; r.o = V3<float>(u * fR, v * fR, z0); r.d = V3<float>(0, 0, 1);
                        TraceOut<float> o = trace_ghost<float>(L, a, b, r, 550.0f, &at[n], &bf[n], true);
                        if (o.status != GL_TRACE_OK) continue;
                        mnx = std::min(mnx, o.pos.x); mxx = std::max(mxx, o.pos.x);
                        mny = std::min(mny, o.pos.y); mxy = std::max(mxy, o.pos.y); ++valid;
                    }
                if (valid >= 2) {
                    float gw = std::max(mxx - mnx, 0.01f), gh = std::max(mxy - mny, 0.01f);
                    float ratio = (gw * gh) / ((2.0f * half_w) * (2.0f * half_h));
                    boost = std::min(std::max(ratio, 1.0f), max_area_boost);
                }
            }
            if (cnt < cap) out[cnt] = gl_pair{a, b, boost};
            ++cnt;
        }
    }
    *n_out = cnt;
    return cnt > cap ? -1 : 0;
}

// launch_ghost_render semantics (ghost_render.cu:630-1353) on the CPU.
// bxy: 2 floats per source = tan(angle) as the product computes it.
// C: 9-float spectral correction or NULL.  out planes: width*height each.
// counters_out (may be NULL): [traces, survived, on_sensor, segments] of the
// main render, pair_active_out (may be NULL): cull mask, rects_out (may be
// NULL): 4 floats per ps, budgets_out: 2 ints per ps.
int orc_render_ghosts(int dbl, const gl_surface* s, int n, const float* arena,
                      const gl_pair* pairs, int n_pairs, const gl_source* src, const float* bxy, int n_src,
                      const gl_spec* spec, int n_spec, const gl_flare_params* p,
                      float half_w, float half_h, int width, int height, int fmt_w, int fmt_h, int fx0, int fy0,
                      const float* C, float* out_r, float* out_g, float* out_b,
                      uint64_t* counters_out, unsigned char* pair_active_out, float* rects_out,
                      int* budgets_out, int threads) {
    threads = threads_of(threads);
    const size_t npx = (size_t)width * height;
    std::fill(out_r, out_r + npx, 0.f); std::fill(out_g, out_g + npx, 0.f); std::fill(out_b, out_b + npx, 0.f);
    if (n_pairs <= 0 || n_src <= 0 || n <= 0) return 0;
    Job J;
    lens_init(J.L, s, n, arena, p->fp64_radius_min);
    J.pairs = pairs; J.n_pairs = n_pairs; J.src = src; J.n_src = n_src; J.bxy = bxy;
    J.spec = spec; J.n_spec = n_spec;
    job_diffraction(J, p);
    std::vector<float> at((size_t)n_spec * n), bf((size_t)n_spec * n);
    ior_tables(s, n, spec, n_spec, at.data(), bf.data());
    J.ior_at = at.data(); J.ior_before = bf.data();
    J.front_R = s[0].semi_aperture; J.start_z = s[0].z - 20.0f;
    J.gain = p->flare_gain;
    std::vector<gl_sample> grid = build_samples(s, n, p->ray_grid, p->pupil_jitter, p->jitter_seed,
                                                p->aperture_blades, p->aperture_rotation);
    const int n_grid = (int)grid.size();
    if (n_grid == 0) return 0;
    J.ray_weight = 1.0f / (float)n_grid;
    const int ns = n_spec;
    const int n_ps = n_pairs * n_src;

    // ---- probe / cull / concentrate (ghost_render.cu:969-1238)
    const bool want_conc = p->concentrate_samples && n_pairs > 1 && ns != 3;
    bool cull_ready = false, rect_ready = false;
    std::vector<unsigned char> active(n_pairs, 1);
    std::vector<float> rects;
    if ((p->cull_dead_pairs || want_conc) && n_pairs > 1 && ns != 3) {
        Frame Fm{half_w * 1.25f, half_h * 1.25f, (float)fmt_w, (float)fmt_h, (float)fx0, (float)fy0, width, height};
        ProbeOut P;
        P.pt.assign(n_pairs, 0); P.pss.assign(n_pairs, 0); P.po.assign(n_pairs, 0);
        P.umin.assign(n_ps, 1e30f); P.vmin.assign(n_ps, 1e30f); P.umax.assign(n_ps, -1e30f); P.vmax.assign(n_ps, -1e30f);
        P.hits.assign(n_ps, 0);
        std::vector<unsigned char> cand(n_pairs, 0);
        bool pass0_ok = false;
        for (int pass = 0; pass < 3; ++pass) {
            std::vector<gl_sample> probe = build_samples(s, n, 40, 1, (int)(0x9E3779B9u * (unsigned)(pass + 1)),
                                                         p->aperture_blades, p->aperture_rotation);
            if (probe.empty()) { if (pass == 0) break; else continue; }
            const unsigned char* mask = (pass == 0 || want_conc) ? nullptr : cand.data();
            probe_pass(J, Fm, probe.data(), (int)probe.size(), spec, ns, at.data(), bf.data(), mask, nullptr,
                       want_conc, P, threads);
            if (pass == 0) {
                pass0_ok = true;
                if (!want_conc) {
                    int nc = 0;
                    for (int q = 0; q < n_pairs; ++q) if (P.pt[q] > 0 && P.po[q] == 0) { cand[q] = 1; ++nc; }
                    if (nc == 0) break;
                }
            }
        }
        if (want_conc && pass0_ok) {
            std::vector<unsigned char> refine(n_ps, 0); int n_ref = 0;
            for (int i = 0; i < n_ps; ++i) if (P.hits[i] < 16) { refine[i] = 1; ++n_ref; }
            const int rns = std::min(3, ns);
            const int li[3] = {0, ns / 2, ns - 1};
            if (n_ref > 0 && n_ref <= 4096) {
                std::vector<gl_sample> rp = build_samples(s, n, 128, 1, (int)0xB5297A4Du, p->aperture_blades,
                                                          p->aperture_rotation);
                if (!rp.empty()) {
                    std::vector<gl_spec> s3(rns);
                    std::vector<float> a3((size_t)rns * n), b3((size_t)rns * n);
                    for (int k = 0; k < rns; ++k) {
                        s3[k] = spec[li[k]];
                        std::copy_n(&at[(size_t)li[k] * n], n, &a3[(size_t)k * n]);
                        std::copy_n(&bf[(size_t)li[k] * n], n, &b3[(size_t)k * n]);
                    }
                    probe_pass(J, Fm, rp.data(), (int)rp.size(), s3.data(), rns, a3.data(), b3.data(), nullptr,
                               refine.data(), true, P, threads);
                }
            }
        }
        if (pass0_ok) {
            for (int q = 0; q < n_pairs; ++q) active[q] = P.po[q] > 0 ? 1 : 0;
            if (want_conc) {
                rects.assign((size_t)n_ps * 4, 0.f);
                const float dil = 1.5f * (2.0f / 40.0f);
                for (int i = 0; i < n_ps; ++i) {
                    float* r = &rects[4 * (size_t)i];
                    r[0] = 2.f; r[1] = 2.f; r[2] = -2.f; r[3] = -2.f;
                    if (P.hits[i] < 3) continue;
                    // Round-trip through the biased-float encoding (x + 2) the
                    // reference's atomics use (ghost_render.cu:194-200).
                    auto rt = [](float x) { float e = x + 2.0f; return e - 2.0f; };
                    float u0 = rt(P.umin[i]), v0 = rt(P.vmin[i]), u1 = rt(P.umax[i]), v1 = rt(P.vmax[i]);
                    if (!(u0 <= u1 && v0 <= v1)) continue;
                    r[0] = std::max(-1.0f, u0 - dil); r[1] = std::max(-1.0f, v0 - dil);
                    r[2] = std::min(1.0f, u1 + dil); r[3] = std::min(1.0f, v1 + dil);
                    if (g_plant_fault == 2) {
                        const float du = 0.1f * (r[2] - r[0]), dv = 0.1f * (r[3] - r[1]);
                        r[0] += du; r[2] -= du; r[1] += dv; r[3] -= dv;
                    }
                }
                rect_ready = true;
            }
        }
        cull_ready = p->cull_dead_pairs != 0;
    }

    // ---- concentration params + adaptive budgets (ghost_render.cu:1240-1304)
    Conc conc;
    std::vector<int> budgets;
    if (rect_ready) {
        conc.rects = rects.data();
        conc.mask = resolve_mask(s, n, p->aperture_blades, p->aperture_rotation);
        conc.jitter = p->pupil_jitter;
        conc.seed_offset = (uint32_t)p->jitter_seed * 1000003u;
        const int nr = std::max(1, (int)std::floor(std::sqrt((double)n_grid)));
        conc.nr_lattice = nr;
        conc.n_r = (p->pupil_jitter == 2) ? n_grid : nr * nr;
        const float N = (float)p->ray_grid;
        const float a_ref = 4.0f * (float)n_grid / (N * N);
        conc.area_norm = 1.0f / (a_ref * (float)conc.n_r);
        if (p->adaptive_sample_budgets) {
            budgets.assign((size_t)n_ps * 2, 0);
            const float boost = std::max(1.0f, p->adaptive_density_boost);
            int floor_s = std::max(256, p->adaptive_min_samples);
            floor_s = ((floor_s + 255) / 256) * 256;
            const int floor_nr = std::max(1, (int)std::lround(std::sqrt((double)floor_s)));
            for (int i = 0; i < n_ps; ++i) {
                budgets[2 * i] = conc.n_r; budgets[2 * i + 1] = conc.nr_lattice;
                const float* r = &rects[4 * (size_t)i];
                if (!(r[0] <= r[2])) continue;
                const float area = (r[2] - r[0]) * (r[3] - r[1]);
                const float raw = boost * (float)n_grid * area / a_ref;
                if (p->pupil_jitter == 2) {
                    int b = (int)std::ceil((double)raw);
                    b = ((b + 255) / 256) * 256;
                    b = std::max(floor_s, std::min(b, n_grid));
                    budgets[2 * i] = b; budgets[2 * i + 1] = conc.nr_lattice;
                } else {
                    int nrp = (int)std::lround(std::sqrt((double)std::max(0.0f, raw)));
                    nrp = std::max(floor_nr, std::min(nrp, conc.nr_lattice));
                    budgets[2 * i] = nrp * nrp; budgets[2 * i + 1] = nrp;
                }
            }
            conc.budgets = budgets.data();
            conc.inv_aref = 1.0f / a_ref;
        }
    }
    if (pair_active_out) for (int q = 0; q < n_pairs; ++q) pair_active_out[q] = cull_ready ? active[q] : 1;
    if (rects_out) {
        for (int i = 0; i < n_ps; ++i) {
            if (rect_ready) std::copy_n(&rects[4 * (size_t)i], 4, &rects_out[4 * (size_t)i]);
            else { rects_out[4 * i] = 2.f; rects_out[4 * i + 1] = 2.f; rects_out[4 * i + 2] = -2.f; rects_out[4 * i + 3] = -2.f; }
        }
    }
    if (budgets_out) {
        for (int i = 0; i < n_ps; ++i) {
            budgets_out[2 * i] = budgets.empty() ? conc.n_r : budgets[2 * i];
            budgets_out[2 * i + 1] = budgets.empty() ? conc.nr_lattice : budgets[2 * i + 1];
        }
    }

    // ---- main render (ghost_render.cu:1315-1346)
    Frame F{half_w, half_h, (float)fmt_w, (float)fmt_h, (float)fx0, (float)fy0, width, height};
    std::vector<double> R(npx, 0.0), G(npx, 0.0), B(npx, 0.0);
    uint64_t segs_total = 0;
#pragma omp parallel num_threads(threads) reduction(+ : segs_total)
    {
        Acc A;
        A.r.assign(npx, 0.0); A.g.assign(npx, 0.0); A.b.assign(npx, 0.0);
        Counters Cn;
#pragma omp for schedule(dynamic, 4)
        for (long long wi = 0; wi < (long long)n_ps * ((n_grid + 255) / 256); ++wi) {
            const int ps = (int)(wi / ((n_grid + 255) / 256));
            const int blk = (int)(wi % ((n_grid + 255) / 256));
            if (cull_ready && active[ps / n_src] == 0) continue;
            for (int gi = blk * 256; gi < std::min(n_grid, blk * 256 + 256); ++gi) {
                if (dbl) do_sample<double, false>(J, rect_ready ? &conc : nullptr, F, ps, gi, grid.data(), n_grid,
                                                  spec, ns, at.data(), bf.data(), &A, Cn, nullptr, nullptr);
                else do_sample<float, false>(J, rect_ready ? &conc : nullptr, F, ps, gi, grid.data(), n_grid,
                                             spec, ns, at.data(), bf.data(), &A, Cn, nullptr, nullptr);
            }
        }
        segs_total += Cn.segs;
        const std::vector<double>& rr = A.r;
        const std::vector<double>& gg = A.g;
        const std::vector<double>& bb = A.b;
#pragma omp critical
        for (size_t i = 0; i < npx; ++i) { R[i] += rr[i]; G[i] += gg[i]; B[i] += bb[i]; }
    }
    for (size_t i = 0; i < npx; ++i) {
        if (C) {
            const double r = R[i], g = G[i], b = B[i];
            R[i] = (double)C[0] * r + (double)C[1] * g + (double)C[2] * b;
            G[i] = (double)C[3] * r + (double)C[4] * g + (double)C[5] * b;
            B[i] = (double)C[6] * r + (double)C[7] * g + (double)C[8] * b;
        }
        out_r[i] = (float)R[i]; out_g[i] = (float)G[i]; out_b[i] = (float)B[i];
    }
    if (counters_out) { counters_out[0] = 0; counters_out[1] = 0; counters_out[2] = 0; counters_out[3] = segs_total; }
    return 0;
}

// Per-pair STATS counters over the full grid (the collect_stats launch,
// ghost_render.cu:1361-1438): traces, survived, on-sensor per pair and the
// segment total.  out: 3 * n_pairs.
int orc_ghost_stats(const gl_surface* s, int n, const float* arena, const gl_pair* pairs, int n_pairs,
                    const gl_source* src, const float* bxy, int n_src, const gl_spec* spec, int n_spec,
                    const gl_flare_params* p, float half_w, float half_h, int width, int height,
                    uint64_t* out, uint64_t* segs_out, int threads) {
    threads = threads_of(threads);
    Job J;
    lens_init(J.L, s, n, arena, p->fp64_radius_min);
    J.pairs = pairs; J.n_pairs = n_pairs; J.src = src; J.n_src = n_src; J.bxy = bxy;
    J.spec = spec; J.n_spec = n_spec;
    job_diffraction(J, p);
    std::vector<float> at((size_t)n_spec * n), bf((size_t)n_spec * n);
    ior_tables(s, n, spec, n_spec, at.data(), bf.data());
    J.ior_at = at.data(); J.ior_before = bf.data();
    J.front_R = s[0].semi_aperture; J.start_z = s[0].z - 20.0f;
    J.gain = p->flare_gain;
    std::vector<gl_sample> grid = build_samples(s, n, p->ray_grid, p->pupil_jitter, p->jitter_seed,
                                                p->aperture_blades, p->aperture_rotation);
    J.ray_weight = grid.empty() ? 0.f : 1.0f / (float)grid.size();
    Frame F{half_w, half_h, (float)width, (float)height, 0.f, 0.f, width, height};
    std::vector<uint64_t> acc(3 * (size_t)n_pairs, 0);
    uint64_t segs = 0;
    const int n_ps = n_pairs * n_src;
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads) reduction(+ : segs)
    for (int ps = 0; ps < n_ps; ++ps) {
        Counters Cn;
        for (int gi = 0; gi < (int)grid.size(); ++gi)
            do_sample<float, true>(J, nullptr, F, ps, gi, grid.data(), (int)grid.size(), spec, n_spec, at.data(),
                                   bf.data(), nullptr, Cn, nullptr, nullptr);
        segs += Cn.segs;
#pragma omp critical
        { acc[3 * (ps / n_src)] += Cn.traces; acc[3 * (ps / n_src) + 1] += Cn.surv; acc[3 * (ps / n_src) + 2] += Cn.onsens; }
    }
    for (int q = 0; q < n_pairs; ++q) { out[q] = acc[3 * q]; out[n_pairs + q] = acc[3 * q + 1]; out[2 * n_pairs + q] = acc[3 * q + 2]; }
    if (segs_out) *segs_out = segs;
    return 0;
}

}  // extern "C"
