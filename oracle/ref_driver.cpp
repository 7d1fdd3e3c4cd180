// ORACLE — test infrastructure only.  A C-ABI shim, written for this project,
// over the reference's own CPU functions.  oracle/Makefile compiles it together
// with the reference sources where they lie (/root/reference/ghostlight/src)
// into oracle/_ref/libglref.so; nothing here is shipped or called by the
// product.  Tests use it to pin the oracle restatement and the product's host
// preparation bit-for-bit against the reference.
#include "optical_system.h"
#include "trace.h"
#include "ghost.h"
#include "spectral.h"
#include "aperture_sampler.h"
#include "tmm.h"
#include "lens_calibration.h"
#include "source_map.h"
#include "aperture_sdf.h"

#include "../include/ghostlight_amd.h"

#include <algorithm>
#include <cstring>
#include <vector>

namespace {

OpticalSystem make_lens(const gl_surface* s, int n, const float* arena) {
    OpticalSystem L;
    L.surfaces.resize(n);
    L.coating_tables.resize(n);
    L.aperture_images.resize(n);
    L.surface_ids.resize(n);
    for (int i = 0; i < n; ++i) {
        Surface& d = L.surfaces[i];
        const gl_surface& g = s[i];
        d = Surface{};
        d.radius = g.radius; d.thickness = g.thickness; d.ior = g.ior; d.abbe_v = g.abbe_v;
        d.semi_aperture = g.semi_aperture; d.z = g.z;
        d.is_stop = g.is_stop != 0; d.is_active = g.is_active != 0;
        d.aperture_shape = g.aperture_shape; d.aperture_blades = g.aperture_blades;
        d.aperture_rotation_rad = g.aperture_rotation_rad; d.aperture_aspect = g.aperture_aspect;
        d.aperture_curvature = g.aperture_curvature; d.aperture_twist = g.aperture_twist;
        d.aperture_notch_rad = g.aperture_notch_rad; d.aperture_notch_angle_rad = g.aperture_notch_angle_rad;
        d.aperture_semi_diameter = g.aperture_semi_diameter;
        d.decenter_x = g.decenter_x; d.decenter_y = g.decenter_y;
        std::memcpy(d.rot, g.rot, sizeof d.rot);
        d.form = g.form; d.conic_k = g.conic_k;
        std::memcpy(d.asphere_terms, g.asphere_terms, sizeof d.asphere_terms);
        d.n_asphere_terms = g.n_asphere_terms; d.cyl_axis = g.cyl_axis;
        d.disp_model = g.disp_model;
        std::memcpy(d.sellmeier_B, g.sellmeier_B, sizeof d.sellmeier_B);
        std::memcpy(d.sellmeier_C, g.sellmeier_C, sizeof d.sellmeier_C);
        Coating& c = d.coating;
        c.model = (CoatingModel)g.coating_model;
        c.ar_layers = g.ar_layers;
        c.out_of_range_discard = g.out_of_range_discard != 0;
        c.angle_ref_ior = g.angle_ref_ior;
        c.tint_r = g.tint_r; c.tint_g = g.tint_g; c.tint_b = g.tint_b; c.tint_strength = g.tint_strength;
        c.gauss_sigma = g.gauss_sigma; c.gauss_background = g.gauss_background; c.gauss_peak = g.gauss_peak;
        c.gauss_decenter_x = g.gauss_decenter_x; c.gauss_decenter_y = g.gauss_decenter_y;
        CoatingTables& t = L.coating_tables[i];
        if (arena && g.table_count > 0)
            for (int k = 0; k < g.table_count; ++k)
                t.table.push_back({arena[g.table_offset + 2 * k], arena[g.table_offset + 2 * k + 1]});
        if (arena && g.sa_n_wavelengths > 0 && g.sa_n_angles > 0) {
            const float* wl = arena + g.sa_offset;
            t.sa_wavelengths.assign(wl, wl + g.sa_n_wavelengths);
            t.sa_angles.assign(wl + g.sa_n_wavelengths, wl + g.sa_n_wavelengths + g.sa_n_angles);
            const float* r = wl + g.sa_n_wavelengths + g.sa_n_angles;
            t.sa_r.assign(r, r + (size_t)g.sa_n_wavelengths * g.sa_n_angles);
        }
        d.refresh_aperture_profile();
    }
    L.sync_coating_pointers();
    return L;
}

// The reverse of make_lens: a reference OpticalSystem as the C ABI's surface
// array and coating arena.
int export_lens(const OpticalSystem& L, gl_surface* out, int cap, float* arena, int arena_cap, int* arena_len) {
    const int n = (int)L.surfaces.size();
    if (n > cap) return -1;
    std::vector<float> ar;
    for (int i = 0; i < n; ++i) {
        const Surface& d = L.surfaces[i];
        gl_surface& g = out[i];
        std::memset(&g, 0, sizeof g);
        g.radius = d.radius; g.thickness = d.thickness; g.ior = d.ior; g.abbe_v = d.abbe_v;
        g.semi_aperture = d.semi_aperture; g.z = d.z;
        g.is_stop = d.is_stop ? 1 : 0; g.is_active = d.is_active ? 1 : 0;
        g.aperture_shape = d.aperture_shape; g.aperture_blades = d.aperture_blades;
        g.aperture_rotation_rad = d.aperture_rotation_rad; g.aperture_aspect = d.aperture_aspect;
        g.aperture_curvature = d.aperture_curvature; g.aperture_twist = d.aperture_twist;
        g.aperture_notch_rad = d.aperture_notch_rad; g.aperture_notch_angle_rad = d.aperture_notch_angle_rad;
        g.aperture_semi_diameter = d.aperture_semi_diameter;
        g.decenter_x = d.decenter_x; g.decenter_y = d.decenter_y;
        std::memcpy(g.rot, d.rot, sizeof g.rot);
        g.form = d.form; g.conic_k = d.conic_k;
        std::memcpy(g.asphere_terms, d.asphere_terms, sizeof g.asphere_terms);
        g.n_asphere_terms = d.n_asphere_terms; g.cyl_axis = d.cyl_axis; g.disp_model = d.disp_model;
        std::memcpy(g.sellmeier_B, d.sellmeier_B, sizeof g.sellmeier_B);
        std::memcpy(g.sellmeier_C, d.sellmeier_C, sizeof g.sellmeier_C);
        const Coating& c = d.coating;
        g.coating_model = (int)c.model; g.ar_layers = c.ar_layers;
        g.out_of_range_discard = c.out_of_range_discard ? 1 : 0; g.angle_ref_ior = c.angle_ref_ior;
        g.tint_r = c.tint_r; g.tint_g = c.tint_g; g.tint_b = c.tint_b; g.tint_strength = c.tint_strength;
        g.gauss_sigma = c.gauss_sigma; g.gauss_background = c.gauss_background; g.gauss_peak = c.gauss_peak;
        g.gauss_decenter_x = c.gauss_decenter_x; g.gauss_decenter_y = c.gauss_decenter_y;
        if (i < (int)L.coating_tables.size()) {
            const CoatingTables& t = L.coating_tables[i];
            if (!t.table.empty()) {
                g.table_offset = (int)ar.size(); g.table_count = (int)t.table.size();
                for (const auto& e : t.table) { ar.push_back(e.key); ar.push_back(e.r); }
            }
            if (!t.sa_wavelengths.empty() && !t.sa_angles.empty()) {
                g.sa_offset = (int)ar.size();
                g.sa_n_wavelengths = (int)t.sa_wavelengths.size(); g.sa_n_angles = (int)t.sa_angles.size();
                ar.insert(ar.end(), t.sa_wavelengths.begin(), t.sa_wavelengths.end());
                ar.insert(ar.end(), t.sa_angles.begin(), t.sa_angles.end());
                ar.insert(ar.end(), t.sa_r.begin(), t.sa_r.end());
            }
        }
    }
    if ((int)ar.size() > arena_cap) return -2;
    std::copy(ar.begin(), ar.end(), arena);
    *arena_len = (int)ar.size();
    return n;
}

LensCalibration calib_of(const float* v) {
    LensCalibration c{};
    c.sensor_half_w = v[0]; c.sensor_half_h = v[1]; c.max_half_angle_h = v[2]; c.max_half_angle_v = v[3];
    c.image_circle_semi_w = v[4]; c.image_circle_semi_h = v[5]; c.focal_length_x = v[6]; c.focal_length_y = v[7];
    c.entrance_pupil_semi_x = v[8]; c.entrance_pupil_semi_y = v[9]; c.f_number_x = v[10]; c.f_number_y = v[11];
    c.pupil_area_frac = v[12];
    return c;
}

void put(const TraceResult& r, float* pos, float* w, int* st) {
    pos[0] = r.position.x; pos[1] = r.position.y; pos[2] = r.position.z;
    *w = r.weight; *st = (int)r.status;
}

}  // namespace

extern "C" {

// bake_aperture_sdf (aperture_sdf.cpp:110) of a W x H bitmap: W*H float4.
int ref_bake_aperture_sdf(const float* px, int W, int H, float semi, float aspect, float* out4) {
    ApertureImage img;
    img.width = W; img.height = H;
    img.pixels.assign(px, px + (size_t)W * H);
    ApertureSdfBake b;
    if (!bake_aperture_sdf(img, semi, aspect, b)) return -1;
    std::copy(b.texels.begin(), b.texels.end(), out4);
    return 0;
}

// make_aperture_profile (aperture_profile.h:285) as gl_blade_profile.
int ref_blade_profile(int blades, float rot, float curvature, float twist, float notch, float rake,
                      gl_blade_profile* out) {
    const ApertureProfile p = make_aperture_profile(blades, rot, ApertureShapeParams{curvature, twist, notch, rake});
    out->blades = p.blades; out->plain = p.plain; out->facets = p.facets; out->rotation = p.rotation;
    out->sigma = p.sigma; out->half = p.half; out->r_w = p.r_w; out->phi_w = p.phi_w;
    out->u_w = p.u_w; out->u_rise = p.u_rise; out->qR = p.qR; out->qL = p.qL;
    out->p = p.p; out->beta = p.beta; out->eA = p.eA; out->eB = p.eB; out->area_frac = p.area_frac;
    return 0;
}

// ApertureProfile::radius_at for n angles.
int ref_blade_radius(int blades, float rot, float curvature, float twist, float notch, float rake,
                     const float* theta, int n, float* out) {
    const ApertureProfile p = make_aperture_profile(blades, rot, ApertureShapeParams{curvature, twist, notch, rake});
    for (int i = 0; i < n; ++i) out[i] = p.radius_at(theta[i]);
    return 0;
}

// trace_ghost_ray (trace.cpp:171) for n rays (7 floats each).
int ref_trace_ghost_rays(const gl_surface* s, int n, const float* arena, int a, int b,
                         const float* rays, int n_rays, float* pos, float* w, int* st) {
    OpticalSystem L = make_lens(s, n, arena);
    for (int i = 0; i < n_rays; ++i) {
        Ray r; const float* q = rays + 7 * i;
        r.origin = Vec3f(q[0], q[1], q[2]); r.dir = Vec3f(q[3], q[4], q[5]); r.lambda = q[6];
        put(trace_ghost_ray(r, L, a, b), pos + 3 * i, w + i, st + i);
    }
    return 0;
}

// trace_primary_ray (trace.cpp:243).
int ref_trace_primary_rays(const gl_surface* s, int n, const float* arena, const float* rays, int n_rays,
                           float* pos, float* w, int* st) {
    OpticalSystem L = make_lens(s, n, arena);
    for (int i = 0; i < n_rays; ++i) {
        Ray r; const float* q = rays + 7 * i;
        r.origin = Vec3f(q[0], q[1], q[2]); r.dir = Vec3f(q[3], q[4], q[5]); r.lambda = q[6];
        put(trace_primary_ray(r, L), pos + 3 * i, w + i, st + i);
    }
    return 0;
}

// The RayPath overloads of trace_ghost_ray / trace_primary_ray
// (trace.cpp:264-395) for n rays; b < 0 traces the primary path.  Events go to
// events[i * max_events + k], k < n_events[i] (truncated at max_events).
int ref_trace_ray_paths(const gl_surface* s, int n, const float* arena, int a, int b, const float* rays, int n_rays,
                        int max_events, gl_trace_event* events, int* n_events, float* pos, float* w, int* st) {
    OpticalSystem L = make_lens(s, n, arena);
    for (int i = 0; i < n_rays; ++i) {
        Ray r; const float* q = rays + 7 * i;
        r.origin = Vec3f(q[0], q[1], q[2]); r.dir = Vec3f(q[3], q[4], q[5]); r.lambda = q[6];
        RayPath path;
        const TraceResult res = b >= 0 ? trace_ghost_ray(r, L, a, b, path) : trace_primary_ray(r, L, path);
        put(res, pos + 3 * i, w + i, st + i);
        const int m = (int)std::min<size_t>(path.events.size(), (size_t)max_events);
        n_events[i] = m;
        for (int k = 0; k < m; ++k) {
            const TraceEvent& e = path.events[k];
            gl_trace_event& o = events[(size_t)i * max_events + k];
            o.surface_index = e.surface_index;
            o.hit_point[0] = e.hit_point.x; o.hit_point[1] = e.hit_point.y; o.hit_point[2] = e.hit_point.z;
            o.surface_normal[0] = e.surface_normal.x; o.surface_normal[1] = e.surface_normal.y;
            o.surface_normal[2] = e.surface_normal.z;
            o.ior_before = e.ior_before; o.ior_after = e.ior_after; o.fresnel_weight = e.fresnel_weight;
            o.status = (int)e.status; o.reflected = e.reflected ? 1 : 0;
        }
    }
    return 0;
}

// filter_ghost_pairs (ghost.cpp:105).
int ref_filter_ghost_pairs(const gl_surface* s, int n, const float* arena, float hw, float hh,
                           float min_intensity, int normalize, float max_boost, gl_pair* out, int cap, int* n_out) {
    OpticalSystem L = make_lens(s, n, arena);
    FlareConfig cfg;
    cfg.min_ghost_intensity = min_intensity; cfg.ghost_normalize = normalize != 0; cfg.max_area_boost = max_boost;
    std::vector<GhostPair> pairs; std::vector<float> boosts;
    filter_ghost_pairs(L, hw, hh, cfg, pairs, boosts);
    *n_out = (int)pairs.size();
    for (size_t i = 0; i < pairs.size() && (int)i < cap; ++i) out[i] = gl_pair{pairs[i].surf_a, pairs[i].surf_b, boosts[i]};
    return 0;
}

// build_spectral_table (spectral.cpp:317) + resolve_output_matrix (:272).
int ref_build_spectral_table(int n, int output_cs, const float* custom9, gl_spec* out) {
    float custom[3][3] = {}; if (custom9) std::memcpy(custom, custom9, sizeof custom);
    float M[3][3];
    resolve_output_matrix((OutputColorSpace)output_cs, custom, M);
    std::vector<GPUSpectralSample> t = build_spectral_table(n, 400.0f, 700.0f, get_sensor_profile(SENSOR_CIE_1931), M);
    for (int i = 0; i < n; ++i) out[i] = gl_spec{t[i].lambda, t[i].cmf_r, t[i].cmf_g, t[i].cmf_b};
    return 0;
}

// compute_spectral_correction (spectral.cpp:419).
int ref_spectral_correction(const gl_spec* tab, int n, int output_cs, const float* custom9, float* C9) {
    float custom[3][3] = {}; if (custom9) std::memcpy(custom, custom9, sizeof custom);
    float M[3][3];
    resolve_output_matrix((OutputColorSpace)output_cs, custom, M);
    std::vector<GPUSpectralSample> t(n);
    for (int i = 0; i < n; ++i) t[i] = GPUSpectralSample{tab[i].lambda, tab[i].cmf_r, tab[i].cmf_g, tab[i].cmf_b};
    float C[3][3];
    compute_spectral_correction(t, 400.0f, 700.0f, get_sensor_profile(SENSOR_CIE_1931), M, C);
    std::memcpy(C9, C, sizeof C);
    return 0;
}

// resolve_input_matrix / resolve_output_matrix (spectral.cpp:203,272).
int ref_color_matrices(int in_cs, const float* cin9, int out_cs, const float* cout9, float* Min, float* Mout) {
    float ci[3][3] = {}, co[3][3] = {};
    if (cin9) std::memcpy(ci, cin9, sizeof ci);
    if (cout9) std::memcpy(co, cout9, sizeof co);
    float A[3][3], B[3][3];
    resolve_input_matrix((InputColorSpace)in_cs, ci, A);
    resolve_output_matrix((OutputColorSpace)out_cs, co, B);
    std::memcpy(Min, A, sizeof A); std::memcpy(Mout, B, sizeof B);
    return 0;
}

// build_aperture_samples (aperture_sampler.h:181).
int ref_build_aperture_samples(const gl_surface* s, int n, int ray_grid, int jitter, int seed, int blades,
                               float rot_deg, gl_sample* out, int cap, int* n_out) {
    OpticalSystem L = make_lens(s, n, nullptr);
    ApertureSamplerParams p;
    p.ray_grid = ray_grid; p.pupil_jitter = jitter; p.jitter_seed = seed;
    p.blades_override = blades; p.rotation_deg = rot_deg;
    std::vector<ApertureSample> g = build_aperture_samples(L, p);
    *n_out = (int)g.size();
    for (size_t i = 0; i < g.size() && (int)i < cap; ++i) out[i] = gl_sample{g[i].u, g[i].v};
    return 0;
}

// tmm::reflectance (tmm.h:89) for one (lambda, aoi).
float ref_tmm_reflectance(int n_layers, const float* n_re, const float* n_im, const float* thick,
                          float n_amb, float n_sub, float lambda_nm, float aoi_deg) {
    std::vector<tmm::Layer> layers(n_layers);
    for (int i = 0; i < n_layers; ++i) { layers[i].n = {n_re[i], n_im[i]}; layers[i].thickness_nm = thick[i]; }
    return tmm::reflectance(layers, {n_amb, 0.0f}, {n_sub, 0.0f}, lambda_nm, aoi_deg);
}

// calibrate_lens (lens_calibration.cpp:344): 13 floats in LensCalibration order.
int ref_calibrate_lens(const gl_surface* s, int n, const float* arena, float* out13) {
    OpticalSystem L = make_lens(s, n, arena);
    LensCalibration c = calibrate_lens(L);
    const float v[13] = {c.sensor_half_w, c.sensor_half_h, c.max_half_angle_h, c.max_half_angle_v,
                         c.image_circle_semi_w, c.image_circle_semi_h, c.focal_length_x, c.focal_length_y,
                         c.entrance_pupil_semi_x, c.entrance_pupil_semi_y, c.f_number_x, c.f_number_y,
                         c.pupil_area_frac};
    std::memcpy(out13, v, sizeof v);
    return 0;
}

// OpticalSystem::load (optical_system.cpp:1299) of a .lens file, exported.
// Returns the surface count, or a negative value on failure.
int ref_load_lens(const char* path, gl_surface* out, int cap, float* arena, int arena_cap, int* arena_len) {
    OpticalSystem L;
    if (!L.load(path)) return -3;
    return export_lens(L, out, cap, arena, arena_cap, arena_len);
}

// solve_source_angle (source_map.cpp:182) for a screen NDC position.
int ref_solve_source_angle(const gl_surface* s, int n, const float* arena, const float* calib13, float ndc_x,
                           float ndc_y, float eff_half_w, float eff_half_h, float* out2) {
    OpticalSystem L = make_lens(s, n, arena);
    const SourceMapSolve r = solve_source_angle(L, calib_of(calib13), ndc_x, ndc_y, eff_half_w, eff_half_h);
    out2[0] = r.angle_x; out2[1] = r.angle_y;
    return 0;
}

// solve_source_angle (source_map.cpp:182), the whole SourceMapSolve in
// gl_source_map field order: angle_x, angle_y, status, residual_mm,
// anchor_ax, anchor_ay, anchor_x, anchor_y, J[4], eff_half_w, eff_half_h,
// inv_tan_w, inv_tan_h, probe_evals (status and probe_evals as floats).
int ref_solve_source_map(const gl_surface* s, int n, const float* arena, const float* calib13, float ndc_x,
                         float ndc_y, float eff_half_w, float eff_half_h, float* out17) {
    OpticalSystem L = make_lens(s, n, arena);
    const SourceMapSolve r = solve_source_angle(L, calib_of(calib13), ndc_x, ndc_y, eff_half_w, eff_half_h);
    const float v[17] = {r.angle_x, r.angle_y, (float)(int)r.status, r.residual_mm, r.anchor_ax, r.anchor_ay,
                         r.anchor_x, r.anchor_y, r.J[0], r.J[1], r.J[2], r.J[3], r.eff_half_w, r.eff_half_h,
                         r.inv_tan_w, r.inv_tan_h, (float)r.probe_evals};
    std::memcpy(out17, v, sizeof v);
    return 0;
}

// source_map_screen (source_map.cpp:456) of a solve given as ref_solve_source_map's 17 floats.
int ref_source_map_screen(const float* solve17, float angle_x, float angle_y, float* out2) {
    SourceMapSolve b;
    b.angle_x = solve17[0]; b.angle_y = solve17[1];
    b.status = (SourceMapStatus)(int)solve17[2];
    b.residual_mm = solve17[3];
    b.anchor_ax = solve17[4]; b.anchor_ay = solve17[5]; b.anchor_x = solve17[6]; b.anchor_y = solve17[7];
    for (int i = 0; i < 4; ++i) b.J[i] = solve17[8 + i];
    b.eff_half_w = solve17[12]; b.eff_half_h = solve17[13];
    b.inv_tan_w = solve17[14]; b.inv_tan_h = solve17[15];
    source_map_screen(b, angle_x, angle_y, &out2[0], &out2[1]);
    return 0;
}

// source_map_landing (source_map.cpp:487): returns 1 with the landing, 0 when nothing survives.
int ref_source_map_landing(const gl_surface* s, int n, const float* arena, const float* calib13, float angle_x,
                           float angle_y, float* out2) {
    OpticalSystem L = make_lens(s, n, arena);
    return source_map_landing(L, calib_of(calib13), angle_x, angle_y, &out2[0], &out2[1]) ? 1 : 0;
}

}  // extern "C"
