/*
 * ghostlight_amd.h — C ABI of the MI355X flare hot path.
 *
 * Every entry point is plain C: POD structs, raw pointers and sizes, an int
 * status (GL_OK = 0) and a thread-local message from gl_last_error().  No
 * torch, no C++ types.  The Python package `ghostlight_amd` binds this with
 * ctypes; INTEGRATION.md shows the binding a maintainer of the reference
 * would add.
 *
 * Each declaration names the reference interface it replaces
 * (path:line relative to the reference repository root).
 *
 * Geometry convention (same as the reference): light travels in +z, the
 * sensor is the plane z = 0, surfaces sit at z <= 0, surfaces[0] is the front
 * element.  Lengths in mm, wavelengths in nm, angles in radians.
 */
#ifndef GHOSTLIGHT_AMD_H
#define GHOSTLIGHT_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GL_OK                 0
#define GL_ERR_INVALID       -1   /* bad argument (maps to Python ValueError)   */
#define GL_ERR_DEVICE        -2   /* HIP error / no device (RuntimeError)       */
#define GL_ERR_NO_MEMORY     -3   /* allocation or HBM-budget failure           */
#define GL_ERR_UNSUPPORTED   -4   /* feature outside this build (§ out of scope) */

#define GL_MAX_SURFACES       64
#define GL_MAX_ASPHERE_TERMS   8
#define GL_MAX_BLADES         32
#define GL_MAX_SPECTRAL      128

/* ---- enums (values equal the reference's, optical_system.h:54-97,
 *      render_config.h, spectral.h:16-38) ---------------------------------- */
enum { GL_FORM_SPHERE = 0, GL_FORM_ASPHERE = 1, GL_FORM_CYLINDRICAL = 2 };
enum { GL_CYL_AXIS_X = 0, GL_CYL_AXIS_Y = 1 };
enum { GL_APERTURE_CIRCLE = 0, GL_APERTURE_POLYGON = 1, GL_APERTURE_IMAGE = 2 };
enum { GL_DISP_AIR = 0, GL_DISP_ABBE = 1, GL_DISP_SELLMEIER = 2 };
enum { GL_COAT_SIMPLE = 0, GL_COAT_SPECTRAL = 1, GL_COAT_ANGULAR = 2,
       GL_COAT_SPECTRAL_ANGULAR = 3, GL_COAT_ATTENUATOR_GAUSS = 4,
       GL_COAT_ARTIST = 5 };
enum { GL_TRACE_OK = 0, GL_TRACE_VIGNETTED = 1, GL_TRACE_TIR = 2,
       GL_TRACE_MISSED_SURFACE = 3, GL_TRACE_INVALID_INPUT = 4 };
enum { GL_CS_ACESCG = 0, GL_CS_SRGB_LINEAR = 1, GL_CS_P3_D65 = 2,
       GL_CS_P3_D60 = 3, GL_CS_XYZ = 4, GL_CS_CUSTOM = 5 };
enum { GL_CS_INPUT_ACESCG = 0, GL_CS_INPUT_SRGB_LINEAR = 1, GL_CS_INPUT_CUSTOM = 2 };
enum { GL_AOV_NONE = 0, GL_AOV_PER_PAIR = 1 };
enum { GL_FILTER_ALL = 0, GL_FILTER_INCLUDE = 1, GL_FILTER_EXCLUDE = 2 };

/* One optical surface.  Field meaning follows `Surface` + `Coating`
 * (optical_system.h:109-249); table-backed coatings index a float arena
 * passed beside the surface array instead of holding raw pointers. */
typedef struct gl_surface {
    float   radius;           /* signed radius of curvature (0 = flat)            */
    float   thickness;        /* axial distance to the next surface               */
    float   ior;              /* n_d of the medium AFTER this surface             */
    float   abbe_v;           /* V_d of the medium after (0 = air)                */
    float   semi_aperture;    /* clear semi-diameter                              */
    float   z;                /* world z of the vertex                            */
    int32_t is_stop;
    int32_t is_active;
    int32_t aperture_shape;   /* GL_APERTURE_*                                    */
    int32_t aperture_blades;
    float   aperture_rotation_rad;
    float   aperture_aspect;
    float   decenter_x, decenter_y;
    float   rot[9];           /* row-major local->world rotation                  */
    int32_t form;             /* GL_FORM_*                                        */
    float   conic_k;
    float   asphere_terms[GL_MAX_ASPHERE_TERMS];
    int32_t n_asphere_terms;
    int32_t cyl_axis;
    int32_t disp_model;       /* GL_DISP_*                                        */
    float   sellmeier_B[3];
    float   sellmeier_C[3];
    /* coating */
    int32_t coating_model;    /* GL_COAT_*                                        */
    int32_t ar_layers;
    int32_t out_of_range_discard;
    float   angle_ref_ior;
    float   tint_r, tint_g, tint_b, tint_strength;
    float   gauss_sigma, gauss_background, gauss_peak;
    float   gauss_decenter_x, gauss_decenter_y;
    int32_t table_offset;     /* SPECTRAL/ANGULAR: arena[off + 2k] = key, [off+2k+1] = r */
    int32_t table_count;
    int32_t sa_offset;        /* SPECTRAL_ANGULAR: wl[nw] | ang[na] | r[nw*na]    */
    int32_t sa_n_wavelengths;
    int32_t sa_n_angles;
    /* blade shape of a POLYGON stop (optical_system.h:189-192): all zero is the
     * plain regular polygon */
    float   aperture_curvature;       /* [-1, +1] waist depth (-1 scalloped, +1 circle) */
    float   aperture_twist;           /* [-1, +1] waist position along the edge         */
    float   aperture_notch_rad;       /* [-pi/4, +pi/4] tip facet                       */
    float   aperture_notch_angle_rad; /* [0, pi/4] facet skew                           */
    /* IMAGE aperture (Surface::aperture_semi_diameter + ApertureImage,
     * optical_system.h:183, 262-269): row-major [0,1] luminance bitmap of
     * image_width x image_height floats at arena[image_offset]; width 0 = not
     * loaded (the surface passes every ray) */
    float   aperture_semi_diameter;   /* world radius at the bitmap's edge (mm)          */
    int32_t image_offset;
    int32_t image_width, image_height;
    /* the stop's bitmap only: 1 = arena[image_sdf_offset] holds its signed
     * distance field, [d_mm, nx, ny, 0] per texel (gl_bake_aperture_sdf), the
     * edge distance HURB kicks by */
    int32_t image_sdf;
    int32_t image_sdf_offset;
} gl_surface;

/* Derived blade silhouette of a bladed stop (ApertureProfile,
 * aperture_profile.h:71-239): boundary radius r(theta) in the frame normalised
 * to the blade-tip radius, each half blade edge a conic arc linear in 1/r,
 * optional straight tip facets.  Built on the host by gl_blade_profile_build;
 * the tracer, the pupil mask and the diffraction pupil all test against it. */
typedef struct gl_blade_profile {
    int32_t blades;       /* < 3: unused                                       */
    int32_t plain;        /* 1: no curvature / twist / notch (regular polygon) */
    int32_t facets;       /* 1: a tip facet is active                          */
    float   rotation;     /* radians                                           */
    float   sigma, half;  /* sector angle 2 pi / blades and its half           */
    float   r_w, phi_w;   /* waist radius and its angle within the sector      */
    float   u_w, u_rise;  /* 1 / r_w and u_w - 1                               */
    float   qR, qL;       /* 1 / sin^2 of half the waist-to-tip spans          */
    float   p, beta;      /* facet chord distance and normal direction         */
    float   eA, eB;       /* facet angular span either side of the tip         */
    float   area_frac;    /* open area / circumscribed disk (Simpson, 512)     */
} gl_blade_profile;

/* Pupil sample in normalised entrance-pupil coordinates
 * (ApertureSample, aperture_sampler.h:21). */
typedef struct gl_sample { float u, v; } gl_sample;

/* Spectral table row (GPUSpectralSample, spectral.h:45). */
typedef struct gl_spec { float lambda, cmf_r, cmf_g, cmf_b; } gl_spec;

/* Ghost pair with its area boost (GPUPair, ghost_render.cu:122). */
typedef struct gl_pair { int32_t surf_a, surf_b; float area_boost; } gl_pair;

/* Flare source by field angle (FlareSource, ghost.h:21). */
typedef struct gl_source { float angle_x, angle_y, r, g, b; } gl_source;

/* Front-of-lens occluder (GpuBaffle, baffle.h:33-41): a clipping plane dz mm
 * downstream of the spawn plane (20 mm ahead of the front vertex), so a
 * Baffle at z_front_mm ahead of the vertex has dz = 20 - z_front_mm.  A pupil
 * ray at spawn (px, py) heading (tan_ax, tan_ay) meets it at
 * (px + tan_ax dz, py + tan_ay dz) and is blocked outside the shape. */
#define GL_MAX_BAFFLES 8
enum { GL_BAFFLE_RECT = 0, GL_BAFFLE_CIRCLE = 1 };
typedef struct gl_baffle {
    int32_t shape;                  /* GL_BAFFLE_*                                 */
    float   dz;
    float   top, bottom, left, right;   /* RECT half-extents (mm)                  */
    float   cx, cy, radius, aspect;     /* CIRCLE / ellipse (y radius = r*aspect)  */
} gl_baffle;

/* HURB kick distributions (HurbKickDistribution, render_config.h:118). */
enum { GL_HURB_GAUSSIAN = 0, GL_HURB_LORENTZIAN = 1 };

/* The FlareConfig / RenderConfig fields the ghost path reads
 * (render_config.h:14-46, 273-337).  Defaults are set by gl_flare_params_init. */
typedef struct gl_flare_params {
    /* RenderConfig */
    int32_t ray_grid;
    int32_t spectral_samples;
    int32_t pupil_jitter;
    int32_t jitter_seed;
    int32_t sensor_model;
    int32_t input_cs;
    int32_t output_cs;
    float   custom_input_to_xyz[9];
    float   custom_xyz_to_output[9];
    float   sensor_half_w;          /* 0 = use the calibrated value               */
    float   sensor_half_h;
    int32_t verbose;
    int32_t collect_stats;
    int32_t spectral_correction;
    /* FlareConfig */
    int32_t aperture_blades;
    float   aperture_rotation;      /* degrees                                    */
    float   flare_gain;
    float   min_ghost_intensity;
    int32_t ghost_normalize;
    float   max_area_boost;
    int32_t cull_dead_pairs;
    int32_t concentrate_samples;
    int32_t adaptive_sample_budgets;
    float   adaptive_density_boost;
    int32_t adaptive_min_samples;
    int32_t aov_mode;
    int32_t aov_max_pairs;
    int32_t ghost_filter_mode;      /* GL_FILTER_*                                */
    int32_t ghost_filter_count;
    const int32_t* ghost_filter_pairs;  /* 2*count ints (surf_a, surf_b)          */
    /* Build-specific precision knob: spherical surfaces with |R| at or above
     * this use an FP64 ray-sphere solve (sag cancellation); <= 0 disables. */
    float   fp64_radius_min;
    /* Build-specific launch form of the main trace (no reference
     * counterpart): GL_BATCH_GRID = one workgroup per (pair-source, 256
     * samples); GL_BATCH_PERSISTENT = persistent workgroups over a
     * cost-ordered work list whose lambda split / worker count / order the
     * autotuner picks per workload shape; GL_BATCH_GRAPH = the same, with the
     * trace + splat + correction sequence captured once in a hipGraph and
     * replayed.  schedule: -1 = autotune, else a fixed candidate index. */
    int32_t batch_mode;
    int32_t schedule;
    /* Build-specific: 1 = every source renders into a frame of its own; the
     * output planes then hold n_sources x height x width floats each (the
     * reference needs one call per source for this). */
    int32_t frame_per_source;
    /* DiffractionConfig parts the ghost trace reads (render_config.h:156-188):
     * HURB edge diffraction (hurb.h) and the baffle stack, matte box first
     * (build_gpu_baffles, baffle.h:124-153). */
    int32_t hurb;                   /* 1: kick each ray at every edge it passes    */
    int32_t hurb_kick;              /* GL_HURB_*                                   */
    float   hurb_min_sigma_rad;
    float   hurb_max_kick_rad;
    int32_t n_baffles;
    gl_baffle baffles[GL_MAX_BAFFLES];
} gl_flare_params;

#define GL_BATCH_GRID        0
#define GL_BATCH_PERSISTENT  1
#define GL_BATCH_GRAPH       2

/* Per-render diagnostics (GhostRenderStats, flare_buffers.h:21-74).  Arrays
 * are caller-owned; *_cap gives their capacity, n_pairs / n_ps the fill. */
typedef struct gl_ghost_stats {
    float    ms_grid_build, ms_upload, ms_kernel, ms_download, ms_total;
    int32_t  n_pairs, n_sources, n_grid, n_spec;
    uint64_t traces_total, traces_survived, traces_on_sensor;
    uint64_t segments;              /* executed surface segments, main render    */
    int32_t  pair_cap;
    int32_t* pair_surf_a;
    int32_t* pair_surf_b;
    uint64_t* pair_traces;
    uint64_t* pair_survived;
    uint64_t* pair_on_sensor;
    int32_t  ps_cap;
    int32_t  n_ps;
    int32_t* ps_hits;
    float*   ps_rect;               /* 4 floats per (pair, source)               */
    int32_t* ps_budget;
} gl_ghost_stats;

/* Opaque handles. */
typedef struct gl_ctx  gl_ctx;   /* one HIP stream + its grow-only buffers        */
typedef struct gl_lens gl_lens;  /* device-resident prescription                  */

/* ---- runtime ------------------------------------------------------------ */

/* Replaces `_cuda_available` (bind_renderers.cpp:518-521): 1 when a gfx950
 * device is usable, else 0. */
int  gl_device_available(void);

/* Message for the last non-zero status on this thread (never NULL). */
const char* gl_last_error(void);

/* A render context: one HIP stream (default priority) and its buffers.  The
 * reference serialises every render behind process mutexes
 * (ghost_render.cu:504,656); contexts let independent renders overlap.
 * ctx = NULL in any call below means the process default context. */
int  gl_ctx_create(gl_ctx** out);
void gl_ctx_destroy(gl_ctx* ctx);
int  gl_ctx_synchronize(gl_ctx* ctx);

/* Device time (ms, HIP events on the context's stream) of the last main
 * ghost trace kernel run on ctx, and of the tiled splat resolve that follows
 * it.  Build-specific instrumentation for the roofline report; no
 * counterpart in the reference. */
float gl_ctx_last_kernel_ms(gl_ctx* ctx);
float gl_ctx_last_splat_ms(gl_ctx* ctx);

/* Deposit-record counts of the last ghost render on ctx (build-specific, no
 * reference counterpart): out[0] records the trace wrote, out[1] records
 * binned into tiles (a record touching k tiles counts k times), out[2]
 * resolve work chunks, out[3] tiles with several chunks (summed through the
 * fixed-point staging pass). */
int   gl_ctx_last_splat_counts(gl_ctx* ctx, uint64_t out[4]);

/* HBM accounting for admission control (build-specific; the reference has
 * no serving layer).  gl_device_memory: free / total bytes of the current
 * device.  gl_ctx_memory_bytes: device bytes a context's buffer cache holds.
 * gl_ctx_trim: release that cache (the context stays usable; buffers regrow
 * on demand).  Trimming waits for the context's stream first. */
int      gl_device_memory(uint64_t* free_bytes, uint64_t* total_bytes);

/* Schedule of the last ghost render on ctx (build-specific): out[0] lambda
 * values per work item, out[1] persistent workgroups, out[2] order (1 =
 * longest-first), out[3] candidate index, out[4] 1 while the autotuner is
 * still trying candidates for this workload shape, out[5] work items,
 * out[6] 1 if replayed from a captured hipGraph, out[7] candidates. */
int      gl_ctx_last_schedule(gl_ctx* ctx, int32_t out[8]);
uint64_t gl_ctx_memory_bytes(gl_ctx* ctx);
int      gl_ctx_trim(gl_ctx* ctx);

/* Device-resident lens (replaces the per-launch surface + coating-arena
 * upload, ghost_render.cu:919-954, coating_upload.cu:46-148). */
int  gl_lens_create(const gl_surface* surfaces, int32_t n_surfaces,
                    const float* arena, int32_t arena_len, gl_lens** out);
void gl_lens_destroy(gl_lens* lens);

/* ---- host preparation (bit-exact restatements of the reference's CPU code) */

/* gl_flare_params defaults = FlareConfig defaults (render_config.h:14-337). */
void gl_flare_params_init(gl_flare_params* p);

/* build_spectral_table (spectral.cpp:317) with resolve_output_matrix
 * (spectral.cpp:272); n == 3 gives the fixed RGB table
 * (ghost_render.cu:777-785).  Writes n rows. */
int  gl_build_spectral_table(int32_t n, int32_t output_cs,
                             const float custom_xyz_to_output[9], gl_spec* out);

/* compute_spectral_correction (spectral.cpp:419). */
int  gl_spectral_correction(const gl_spec* table, int32_t n, int32_t output_cs,
                            const float custom_xyz_to_output[9], float C_out[9]);

/* resolve_input_matrix / resolve_output_matrix (spectral.cpp:203, 272). */
int  gl_color_matrices(int32_t input_cs, const float custom_in[9],
                       int32_t output_cs, const float custom_out[9],
                       float M_in[9], float M_out[9]);

/* build_aperture_samples (aperture_sampler.h:181).  Writes at most cap
 * samples, *n_out = count. */
int  gl_build_aperture_samples(const gl_surface* surfaces, int32_t n_surfaces,
                               int32_t ray_grid, int32_t pupil_jitter,
                               int32_t jitter_seed, int32_t blades_override,
                               float rotation_deg, gl_sample* out, int32_t cap,
                               int32_t* n_out);

/* ior_at / ior_before tables [lambda][surface]
 * (ghost_render.cu:809-823; optical_system.h:230, 385-402). */
int  gl_build_ior_tables(const gl_surface* surfaces, int32_t n_surfaces,
                         const gl_spec* spec, int32_t n_spec,
                         float* ior_at, float* ior_before);

/* Bake a thin-film stack to a SPECTRAL_ANGULAR table, 31 lambda x 18 angles
 * (OpticalSystem::bake_coating_layers, optical_system.cpp:1483-1570, with
 * tmm::reflectance, tmm.h:89).  Layer i: complex index (n_re[i], n_im[i]) at
 * every wavelength, thickness_nm[i]; n_ambient/n_substrate per wavelength
 * (31 each).  out: wl[31] | ang[18] | r[31*18]. */
int  gl_bake_tmm_table(int32_t n_layers, const float* layer_n_re,
                       const float* layer_n_im, const float* thickness_nm,
                       const float* n_ambient31, const float* n_substrate31,
                       float* out);

/* make_aperture_profile (aperture_profile.h:285-363, with ap_area_fraction
 * :271-282): the blade silhouette of a stop with `blades` >= 3 blades at
 * `rotation_rad` and the four blade-shape controls (clamped to their
 * authored ranges).  blades < 3 gives the unused profile (area_frac 1). */
int  gl_blade_profile_build(int32_t blades, float rotation_rad, float curvature,
                            float twist, float notch_rad, float notch_angle_rad,
                            gl_blade_profile* out);

/* ApertureProfile::radius_at (aperture_profile.h:176-181): the normalised
 * boundary radius at absolute angle theta (1 when blades < 3). */
float gl_blade_radius_at(const gl_blade_profile* profile, float theta);
/* Analytic slope d(radius)/d(theta) of the same boundary, as the HURB edge
 * distance uses it (ApertureProfile::dr_dtheta, aperture_profile.h:235-255). */
float gl_blade_dr_dtheta(const gl_blade_profile* profile, float theta);

/* _hurb_sample_debug (bind_renderers.cpp:457-481): for each edge distance
 * d[i] the HURB kick scale sigma = lambda K / d and one kick drawn from an
 * independent stream seeded wang(seed ^ (2654435761 i + 1)) (hurb.h:49-96). */
int  gl_hurb_sample(float lambda_nm, const float* d_mm, int32_t n, int32_t dist,
                    float max_kick_rad, uint32_t seed, float* kicks, float* sigma);

/* _aperture_edge_distance_debug (bind_trace.cpp:93): perpendicular distance
 * from a world-space point on surface s to its clear-aperture edge and the
 * world-space edge normal (aperture_edge_distance, trace_core.h:206-290). */
int  gl_aperture_edge_distance(const gl_surface* s, const float hit[3], float* d_mm,
                               float normal[3]);

/* bake_aperture_sdf (aperture_sdf.cpp:110-189): signed distance (mm,
 * positive where the bitmap passes) from every texel centre to the 0.5 matte
 * edge and the unit edge normal, as W*H float4 [d_mm, nx, ny, 0], under the
 * trace's UV map (x spans 2 aspect semi, y spans 2 semi). */
int  gl_bake_aperture_sdf(const float* pixels, int32_t width, int32_t height,
                          float semi_diameter_mm, float aspect, float* out4);

/* ---- ghost path -------------------------------------------------------- */

/* enumerate_ghost_pairs (ghost.cpp:17): all active a < b.  Writes at most
 * cap pairs (area_boost = 1). */
int  gl_enumerate_ghost_pairs(const gl_lens* lens, gl_pair* out, int32_t cap,
                              int32_t* n_out);

/* filter_ghost_pairs (ghost.cpp:105): IOR-contrast gate, on-axis intensity
 * gate and area boost.  The 3 + 64 probe traces per pair run on the device. */
int  gl_filter_ghost_pairs(gl_ctx* ctx, const gl_lens* lens,
                           float sensor_half_w, float sensor_half_h,
                           const gl_flare_params* p,
                           gl_pair* out, int32_t cap, int32_t* n_out);

/* launch_ghost_render (ghost_render.h:75-94): probe / cull / concentrate /
 * adaptive budgets, then the ghost trace + splat and the spectral correction.
 * out_r/g/b: host planes (width*height floats), overwritten.  Pass NULL
 * host planes and non-NULL dev_out to keep the result on the device
 * (dev_out receives 3 device pointers owned by ctx, valid until the next
 * call on ctx).  stats may be NULL. */
int  gl_launch_ghost_render(gl_ctx* ctx, const gl_lens* lens,
                            const gl_pair* pairs, int32_t n_pairs,
                            const gl_source* sources, int32_t n_sources,
                            float sensor_half_w, float sensor_half_h,
                            int32_t width, int32_t height,
                            int32_t fmt_w, int32_t fmt_h,
                            int32_t fmt_x0_in_buf, int32_t fmt_y0_in_buf,
                            const gl_flare_params* p,
                            float* out_r, float* out_g, float* out_b,
                            float** dev_out, gl_ghost_stats* stats);

/* render_ghost_pipeline (ghost_render.cu:1448-1581): filter_ghost_pairs,
 * GhostFilter include/exclude, then launch_ghost_render.  In PER_PAIR AOV
 * mode aov_out receives n_layers x 3 planes (layer-major, r|g|b) and the
 * pair of each layer in aov_pairs (2 ints per layer); both caller-owned with
 * room for aov_cap layers, *aov_n = layers written. */
int  gl_render_ghost_pipeline(gl_ctx* ctx, const gl_lens* lens,
                              const gl_source* sources, int32_t n_sources,
                              float calib_half_w, float calib_half_h,
                              int32_t width, int32_t height,
                              int32_t fmt_w, int32_t fmt_h,
                              int32_t fmt_x0_in_buf, int32_t fmt_y0_in_buf,
                              const gl_flare_params* p,
                              float* out_r, float* out_g, float* out_b,
                              float* aov_out, int32_t* aov_pairs,
                              int32_t aov_cap, int32_t* aov_n,
                              gl_ghost_stats* stats);

/* Single-ray traces through the device tracer for tests and Designer-style
 * probes (trace_ghost_ray / trace_primary_ray, trace.h; GPU semantics of
 * trace_cuda.h:112,197).  rays: n x 7 floats (ox,oy,oz,dx,dy,dz,lambda).
 * out_pos: n x 3, out_w: n, out_status: n, out_segments: n (may be NULL).
 * t_ge0 = 1 selects the CPU tracer's t >= 0 sensor rule (trace.cpp:155). */
int  gl_trace_ghost_rays(gl_ctx* ctx, const gl_lens* lens, int32_t bounce_a,
                         int32_t bounce_b, const float* rays, int32_t n,
                         int32_t t_ge0, float fp64_radius_min,
                         float* out_pos, float* out_w, int32_t* out_status,
                         int32_t* out_segments);

/* gl_trace_ghost_rays with the device tracer's instantiation chosen:
 * level 0 is the generic surface step (what gl_trace_ghost_rays runs);
 * 1 and 2 are the predicated lean steps (1: flat and spherical surfaces,
 * untransformed, round or polygon apertures; 2: adds tilted or decentred
 * surfaces, cylinders and stretched apertures), with the refraction ratios
 * and coating rows taken from per-ray tables as a render takes its
 * per-lambda ones; -1 is the level a render of this lens runs.  A lean level
 * below the one the lens needs, or any lean level for a lens only the generic
 * step handles (aspheres, FP64 radii, image surfaces), is GL_ERR_INVALID.  Same results at every
 * level (the tests hold them bit-equal); no counterpart in the reference
 * beyond trace_ghost_ray / trace_primary_ray (trace.h). */
int  gl_trace_ghost_rays_level(gl_ctx* ctx, const gl_lens* lens, int32_t bounce_a,
                               int32_t bounce_b, const float* rays, int32_t n,
                               int32_t t_ge0, float fp64_radius_min, int32_t level,
                               float* out_pos, float* out_w, int32_t* out_status,
                               int32_t* out_segments);

/* One surface interaction of a diagnostic trace (TraceEvent,
 * trace_event.h:48-58): the hit point (the ray's origin when the surface was
 * missed), the surface normal (zero when missed), the index on each side,
 * the Fresnel factor of this surface (transmittance, or the reflectance at a
 * ghost bounce; 0 when the ray died here), the status and the bounce flag. */
typedef struct gl_trace_event {
    int32_t surface_index;
    float   hit_point[3];
    float   surface_normal[3];
    float   ior_before, ior_after;
    float   fresnel_weight;
    int32_t status;                     /* GL_TRACE_*                                 */
    int32_t reflected;                  /* 1 at a ghost bounce                        */
} gl_trace_event;

/* Self-check of the tracer's square root (no counterpart in the reference):
 * every float whose bits lie in [lo_bits, hi_bits] goes through the
 * tracer's corrected hardware square root and through IEEE sqrtf on the
 * device; *mismatches receives how many differ in any bit and
 * *first_mismatch the smallest such bit pattern (0xFFFFFFFF: none).  The
 * tracer uses that form only on arguments of 0, at least 2^-96, +inf or NaN,
 * so [0x0F800000, 0x7F800000] (2^-96 to +inf) and 0 must give no mismatch;
 * a NaN argument gives a NaN either way, its payload may differ. */
int  gl_selftest_sqrt_rn(gl_ctx* ctx, uint32_t lo_bits, uint32_t hi_bits, uint64_t* mismatches,
                         uint32_t* first_mismatch);
/* The same check for each of the tracer's fast correctly rounded forms
 * (no counterpart in the reference): op 0 = the square root above, op 1 =
 * the refined hardware reciprocal against 1.0f / x, op 2 = half of it
 * against 0.5f / x, op 3 = the square root of any x >= 0 (scaled below
 * 2^-96) against sqrtf, which must agree on all of [0x00000000, 0x7F800000].  The tracer uses the reciprocal forms only on normal
 * arguments whose quotient is normal, so [0x00800000, 0x7E800000] (op 1),
 * [0x00800000, 0x7DFFFFFF] (op 2) and their negative images must give no
 * mismatch. */
int  gl_selftest_rn(gl_ctx* ctx, int32_t op, uint32_t lo_bits, uint32_t hi_bits, uint64_t* mismatches,
                    uint32_t* first_mismatch);

/* The diagnostic overloads of trace_ghost_ray / trace_primary_ray
 * (trace.cpp:264-395; the reference's trace_ghost_ray_diagnostic and
 * trace_primary_ray_diagnostic, bind_trace.cpp:71-86) for n rays at once, one
 * lane each on the device: the landing as gl_trace_ghost_rays, plus every
 * surface the ray reached, in order, in events[i * max_events + k] for k <
 * n_events[i].  bounce_b < 0 traces the primary path.  max_events must be at
 * least 3 x the lens's surface count (the longest ghost path).  Bounce
 * indices out of order give INVALID_INPUT, weight 0 and no events, as the
 * reference's overload does. */
int  gl_trace_rays_diagnostic(gl_ctx* ctx, const gl_lens* lens, int32_t bounce_a, int32_t bounce_b,
                              const float* rays, int32_t n, int32_t t_ge0, float fp64_radius_min,
                              int32_t max_events, gl_trace_event* events, int32_t* n_events,
                              float* out_pos, float* out_w, int32_t* out_status);

/* ---- diffraction: FFT seam, starburst, veil --------------------------------- */

/* LensCalibration (lens_calibration.h:11-57), same field order. */
typedef struct gl_calibration {
    float sensor_half_w, sensor_half_h, max_half_angle_h, max_half_angle_v;
    float image_circle_semi_w, image_circle_semi_h;
    float focal_length_x, focal_length_y;
    float entrance_pupil_semi_x, entrance_pupil_semi_y;
    float f_number_x, f_number_y;
    float pupil_area_frac;
} gl_calibration;

enum { GL_STARBURST_SPRITE = 0, GL_STARBURST_MDFT = 1 };
enum { GL_MDFT_FP32 = 0, GL_MDFT_BF16X3 = 1 };

/* DiffractionConfig (render_config.h:131-189) + the RenderConfig / FlareConfig
 * fields the starburst and veil read. */
typedef struct gl_diffraction_params {
    int32_t starburst_engine;       /* GL_STARBURST_*                              */
    int32_t starburst_grid;
    int32_t starburst_grid_cap;
    float   pupil_fill;
    float   starburst_gain;
    float   scale_trim;
    int32_t spectral_samples;       /* 0 = inherit render_spectral_samples         */
    int32_t use_survivor_mask;
    int32_t survivor_grid;
    float   veil_gain;
    float   veil_spread;
    float   veil_falloff;
    int32_t render_spectral_samples;
    int32_t sensor_model;
    int32_t output_cs;
    float   custom_xyz_to_output[9];
    int32_t aperture_blades;
    float   aperture_rotation;      /* degrees                                     */
    int32_t mdft_precision;         /* GL_MDFT_*: build-specific (3xBF16 MFMA split) */
    int32_t n_baffles;              /* hard clip of the pupil (build_pupil_kernel) */
    gl_baffle baffles[GL_MAX_BAFFLES];
    int32_t use_surface_textures;   /* fold the first loaded IMAGE aperture's bitmap
                                     * into the pupil as graded transmission      */
} gl_diffraction_params;

/* StarburstDebug (diffraction.h:22-39).  Arrays are caller-owned with room
 * for grid_cap^2 (psf, pupil) and 3*grid_cap^2 (sprite_rgb); NULL skips. */
typedef struct gl_starburst_debug {
    int32_t grid;
    int32_t grid_cap;
    float*  psf;
    float*  sprite_rgb;
    float*  pupil;
    float   dx_mm_x, dx_mm_y, f_number_x, f_number_y;
    float   focal_length_x, focal_length_y;
    float   entrance_pupil_semi_x, entrance_pupil_semi_y;
    float   lambda_ref_nm, field_angle_x, field_angle_y;
    float   throughput, chief_offset_x, chief_offset_y;
} gl_starburst_debug;

/* VeilDebug (veil_render.h:31-39); reference is grid^2, sprite 3*grid^2. */
typedef struct gl_veil_debug {
    int32_t grid;
    float*  reference;
    float*  sprite_rgb;
    float   dx_mm_x, dx_mm_y, core_mm, falloff, core_texels;
} gl_veil_debug;

/* fft2d_c2c_forward_inplace (fft_backend.h:25): forward, unnormalised,
 * in-place 2-D C2C FFT of an N x N interleaved complex64 device array
 * (N a power of two, 2..16384), on ctx's stream. */
int  gl_fft2d_c2c_forward_inplace(gl_ctx* ctx, void* d_data, int32_t N);

/* Host-buffer convenience form of the same transform (tests, tools). */
int  gl_fft2d_c2c_forward_host(gl_ctx* ctx, const float* in, float* out, int32_t N);

/* Device time of the same transform (build-specific instrumentation, no
 * reference counterpart): `reps` back-to-back in-place transforms of an
 * N x N device array the context owns, bracketed by HIP events on ctx's
 * stream after one untimed transform; *ms_per_transform = elapsed / reps. */
int  gl_fft2d_time(gl_ctx* ctx, int32_t N, int32_t reps, float* ms_per_transform);

/* render_starburst (diffraction.h:54-72): effective pupil (stop x cat's-eye
 * envelope), LDS radix FFT, PSF, lambda bake, downsample and area resample
 * (SPRITE) or the matrix-DFT window (MDFT).  px/py: each source's buffer-pixel
 * position.  out planes: host, width*height, overwritten.  dbg may be NULL. */
int  gl_render_starburst(gl_ctx* ctx, const gl_lens* lens, const gl_calibration* calib,
                         const gl_source* sources, const float* px, const float* py,
                         int32_t n_sources, int32_t width, int32_t height,
                         int32_t fmt_w, int32_t fmt_h, float sensor_half_w,
                         float sensor_half_h, const gl_diffraction_params* p,
                         float* out_r, float* out_g, float* out_b,
                         gl_starburst_debug* dbg);

/* render_veil (veil_render.h:54-70): analytic Lorentzian glare-spread sprite
 * (cached per falloff) resampled per source.  dbg may be NULL. */
int  gl_render_veil(gl_ctx* ctx, const gl_source* sources, const float* px, const float* py,
                    int32_t n_sources, int32_t width, int32_t height, int32_t fmt_w,
                    int32_t fmt_h, float sensor_half_w, float sensor_half_h,
                    const gl_diffraction_params* p, float* out_r, float* out_g,
                    float* out_b, gl_veil_debug* dbg);

/* ---- calibration and source map ------------------------------------------- */

/* calibrate_lens (lens_calibration.cpp:344-443; Python binding
 * bind_calibration.cpp:40): paraxial focal lengths at a 1-degree field,
 * entrance-pupil support (marginal-ray bisection per azimuth), covered field
 * at 90% of the axial spawn-disc throughput and image circle at 5%, each with
 * its survivor-mean landing.  Every probe ray is traced on the device.
 * file_focal_length: the lens file's focal length, used only when the
 * covered field resolves to nothing (0 = none). */
int  gl_calibrate_lens(gl_ctx* ctx, const gl_lens* lens, float d_line_nm,
                       float file_focal_length, gl_calibration* out);

enum { GL_SMAP_SOLVED = 0, GL_SMAP_CONTINUED = 1, GL_SMAP_FALLBACK = 2 };

/* SourceMapSolve (source_map.h:20-52), same field order. */
typedef struct gl_source_map {
    float   angle_x, angle_y;           /* radians: the answer                        */
    int32_t status;                     /* GL_SMAP_*                                  */
    float   residual_mm;
    float   anchor_ax, anchor_ay;       /* angle whose landing was measured           */
    float   anchor_x, anchor_y;         /* its landing (mm)                           */
    float   J[4];                       /* d(landing)/d(angle), row-major             */
    float   eff_half_w, eff_half_h;
    float   inv_tan_w, inv_tan_h;       /* FALLBACK only                              */
    int32_t probe_evals;
} gl_source_map;

/* solve_source_angle (source_map.cpp:182-452): screen NDC (-1..1, 0 = axis)
 * of a frame with half-extents eff_half_w/h mm to the field angle whose
 * spawn-disc survivor mean lands there (Newton on device-traced probes),
 * with the arctangent continuation past the image circle and the
 * constant-focal-length fallback.  One-entry memo as the reference. */
int  gl_solve_source_angle(gl_ctx* ctx, const gl_lens* lens, const gl_calibration* calib,
                           float ndc_x, float ndc_y, float eff_half_w, float eff_half_h,
                           gl_source_map* out);

/* source_map_screen (source_map.cpp:456-484): the forward map of a solve at
 * any angle, as [0, 1] frame fractions.  Pure arithmetic, no device work. */
void gl_source_map_screen(const gl_source_map* base, float angle_x, float angle_y,
                          float* out_sx, float* out_sy);

/* source_map_landing (source_map.cpp:487-519): survivor-mean landing of the
 * spawn disc at one angle on the calibrated probe grid; *out_ok = 0 when no
 * ray survives. */
int  gl_source_map_landing(gl_ctx* ctx, const gl_lens* lens, const gl_calibration* calib,
                           float angle_x, float angle_y, float* out_x_mm, float* out_y_mm,
                           int32_t* out_ok);

/* The PSF pre-pass's deep-pupil retries (PSFRenderer.cpp:109-122, 203-247,
 * 383-410): for sources whose full-front-disc probes find nothing (a stop
 * much smaller than the front element), track the axial transmitting window
 * out to each source's field angle with the spawn-disc probe.
 * target_mode 0: the survivor-mean landing at the source angle;
 * target_mode 1: the field angle whose survivor mean lands on targets[2i..]
 * (Newton, 1 mrad finite differences, 5 um tolerance).
 * out6 per source: beam angle x, y, landing x, y (mm), residual, found (1/0). */
int  gl_psf_deep_pupil(gl_ctx* ctx, const gl_lens* lens, int32_t n, const float* angles,
                       const float* targets, int32_t target_mode, float* out6);

/* ---- geometric PSF grid --------------------------------------------------- */

/* PSFGpuSource (psf_render.h:17-24): a field-point source paired with its
 * chief landing (computed by the host pre-pass) and its tile origin. */
typedef struct gl_psf_source {
    float   angle_x, angle_y;           /* ray-spawn direction (radians)              */
    float   r, g, b;                    /* source weights                             */
    float   chief_x_mm, chief_y_mm;     /* chief landing on the sensor                */
    int32_t tile_x0, tile_y0;           /* tile origin in the composite (pixels)      */
} gl_psf_source;

/* PSFConfig's grid fields (render_config.h:364-383); the RenderConfig part
 * (ray grid, spectra, colour space, stop override) comes in gl_flare_params. */
typedef struct gl_psf_params {
    int32_t grid_nx, grid_ny, tile_w, tile_h;
    float   tile_extent_mm;
    int32_t monochromatic;
    float   splat_sigma_um;             /* <= 0: bilinear splat                       */
} gl_psf_params;

/* launch_psf_render (psf_render.h:30-42, psf_render.cu:225-465): primary
 * traces of every (source, pupil sample, wavelength) splatted relative to
 * the chief landing into the source's tile; composite of grid_nx*tile_w x
 * grid_ny*tile_h.  out_*: host, overwritten.  ms_out (may be NULL): kernel
 * time, HIP events. */
int  gl_render_psf_tiles(gl_ctx* ctx, const gl_lens* lens, const gl_psf_source* sources,
                         int32_t n_sources, const gl_flare_params* render,
                         const gl_psf_params* psf, float* out_r, float* out_g, float* out_b,
                         float* ms_out);

/* ---- film-gate scatter layer ---------------------------------------------- */

enum { GL_GATE_CAUCHY_ISO = 0, GL_GATE_GROOVED = 1 };

/* GateConfig (render_config.h:212-269), same fields and units. */
typedef struct gl_gate_params {
    float   thickness_mm, standoff_mm;
    float   offset_left_mm, offset_right_mm, offset_bottom_mm, offset_top_mm;
    int32_t lobe;                       /* GL_GATE_*                                  */
    float   roughness_rad, groove_aniso, groove_angle_deg, max_kick_rad, reflectance_r0;
    float   gain;
    int32_t scatter_samples;
    int32_t spectral_samples;           /* 0 = the render's spectral_samples          */
} gl_gate_params;

/* render_gate (gate_render.h:52-67, gate_render.cu:280-514): primary traces of
 * every (source, pupil sample, wavelength); a ray landing just outside the
 * frame opening that strikes the gate's cut wall (standoff_mm ahead of the
 * sensor, thickness_mm deep) reflects with Schlick reflectance, scatters by
 * the wall's lobe (scatter_samples draws) and is splatted back in frame.
 * sensor_half_w/h: the render's effective half extents (the opening).
 * render: ray grid, jitter, stop override, spectra, colour space, FP64
 * threshold.  out_*: host, width*height, overwritten.  counters_out (may be
 * NULL): [primary traces, wall scrapes]. */
int  gl_render_gate(gl_ctx* ctx, const gl_lens* lens, const gl_source* sources, int32_t n_sources,
                    int32_t width, int32_t height, int32_t fmt_w, int32_t fmt_h, int32_t fmt_x0,
                    int32_t fmt_y0, float sensor_half_w, float sensor_half_h,
                    const gl_flare_params* render, const gl_gate_params* gate,
                    float* out_r, float* out_g, float* out_b, uint64_t* counters_out);

/* The gate as the scatter pass resolves it (build_gpu_gate, gate.h:168-211):
 * the opening from the effective sensor half extents and the per-side
 * offsets, the clamped wall depth, roughness, groove frame, R0, kick cap and
 * scatter count.  Host only; backs the reference's _gate_params_debug
 * (bind_renderers.cpp:416-432). */
typedef struct gl_gate_resolved {
    float   x_pos, x_neg, y_pos, y_neg;
    float   t_mm, zs_mm;
    float   sig_wide, sig_tight;
    float   cos_g, sin_g;
    float   r0, max_kick, inv_scatter, gain;
    int32_t n_scatter, lobe;
} gl_gate_resolved;

int  gl_gate_resolve(const gl_gate_params* gate, float sensor_half_w, float sensor_half_h,
                     gl_gate_resolved* out);

/* The gate pass's telemetry (GateDebug, gate_render.h:27-39; backs the
 * reference's _render_gate_debug, bind_renderers.cpp:288-334): the resolved
 * gate, traces and scrapes, the layer's energy (sum of r + g + b over the
 * output, in double), the largest realised specular reach |fold - edge| of
 * any scrape, and the direct image of the same primary traces (each landing's
 * weight added to the pixel it falls in, width*height planes the caller
 * provides; skipped when direct_r is NULL). */
typedef struct gl_gate_debug {
    gl_gate_resolved gate;
    uint64_t traces, scrapes;
    double   energy;
    float    reach_mm;
    float*   direct_r;
    float*   direct_g;
    float*   direct_b;
} gl_gate_debug;

/* gl_render_gate with the telemetry filled into *dbg (which must not be NULL). */
int  gl_render_gate_debug(gl_ctx* ctx, const gl_lens* lens, const gl_source* sources, int32_t n_sources,
                          int32_t width, int32_t height, int32_t fmt_w, int32_t fmt_h, int32_t fmt_x0,
                          int32_t fmt_y0, float sensor_half_w, float sensor_half_h,
                          const gl_flare_params* render, const gl_gate_params* gate,
                          float* out_r, float* out_g, float* out_b, gl_gate_debug* dbg);

/* n draws of the wall's scatter lobe in its own frame (along the wall
 * normal, along the wall, down the axis), draw i from the stream seeded
 * wang(seed ^ (2654435761 i + 1)) with the device's xorshift, Cauchy and
 * Box-Muller draws (gate_sample_lobe, gate.h:116-131).  Host only; backs
 * the reference's _gate_sample_debug (bind_renderers.cpp:435-455). */
int  gl_gate_sample(const gl_gate_params* gate, int32_t n, uint32_t seed, float* axis, float* other, float* z);

/* Veil applied to a whole image by FFT convolution (the extended-source form
 * of render_veil, veil_render.h:54-70): out_c = veil_gain * (in_c (*) g) with
 * the unit-sum glare-spread kernel g(r) = (a^2/(r^2+a^2))^p of
 * veil_render.cu:94-106 (a = veil_spread * sensor_half_h mm, p =
 * veil_falloff), linear convolution over a zero-padded power-of-two grid.
 * device_ptrs = 1: in_* / out_* are device pointers on ctx's device; else
 * host buffers of width*height floats.  ms_out (may be NULL): device time of
 * the transform work, HIP events.  Build-specific: the reference splats the
 * kernel per point source instead. */
int  gl_veil_convolve(gl_ctx* ctx, const float* in_r, const float* in_g, const float* in_b,
                      int32_t width, int32_t height, float sensor_half_w, float sensor_half_h,
                      const gl_diffraction_params* p, int32_t device_ptrs,
                      float* out_r, float* out_g, float* out_b, float* ms_out);

/* gl_veil_convolve with device pointers, for callers that cannot allocate
 * (code under HIP graph capture, tight serving loops): no counterpart in the
 * reference.  It calls no hipMalloc / hipFree and no allocating library
 * routine; its work arrays live in one file-scope device array of 64 MiB
 * (per device, held until the process ends), so the padded grid is limited
 * to 2048^2: width and height <= 1024, larger images are rejected with
 * GL_ERR_INVALID.  ctx must have been created beforehand.  in_* / out_* are
 * device planes of width*height floats on ctx's device.  The work is queued
 * on ctx's stream and the call returns without waiting (gl_ctx_synchronize
 * before reading out_*).  Every call on one device shares that array and
 * keeps the last call's contents, so calls that overlap on one device (other
 * contexts, other threads) race. */
int  gl_veil_convolve_noalloc(gl_ctx* ctx, const float* in_r, const float* in_g, const float* in_b,
                              int32_t width, int32_t height, float sensor_half_w, float sensor_half_h,
                              const gl_diffraction_params* p, float* out_r, float* out_g, float* out_b);

/* Survival of primary rays launched from (x, y) on the spawn plane in one
 * direction (the CPU tracer's t >= 0 rule, trace.cpp:243): the kernel behind
 * the starburst's survivor centroid and cat's-eye envelope
 * (starburst_render.cu:393-582).  launch: n x 2 floats; out: n bytes (1 = OK). */
int  gl_primary_survival(gl_ctx* ctx, const gl_lens* lens, const float* launch, int32_t n,
                         float dir_x, float dir_y, float dir_z, float z_spawn,
                         float lambda_nm, unsigned char* out);

/* ---- matrix DFT on MFMA ----------------------------------------------------- */

/* Complex GEMM C = A . B on the matrix cores (row-major interleaved
 * complex64, host buffers), for tests of the MFMA engine.  precision:
 * GL_MDFT_FP32 = v_mfma_f32_32x32x2_f32, GL_MDFT_BF16X3 = 3-term BF16 split
 * (hi*hi + hi*lo + lo*hi) on v_mfma_f32_32x32x16_bf16, FP32 accumulate.
 * Build-specific; no counterpart in the reference. */
int  gl_cgemm_host(gl_ctx* ctx, int32_t M, int32_t N, int32_t K, const float* A,
                   const float* B, float* C, int32_t precision);

/* Matrix-DFT PSFs of one N x N complex pupil (host buffer) at n_lambda
 * wavelengths: E_l = A_y . P . A_x^T with A[k][n] = exp(-2 pi i (k - M/2)
 * (n - N/2) / (q_l N)), q_l = q_ref * lambda_l / lambda_ref (the lambda-scaled
 * sampling of the precision-starburst workload), PSF_l = |E_l|^2 (M x M).
 * Replaces the per-(source, lambda) contract_x / accumulate_y loops
 * (mdft_render.cu:137-175) with batched MFMA GEMMs.  out_psf: host,
 * n_lambda * M * M floats (may be NULL with dev_out); dev_out receives the
 * device array (owned by ctx).  ms_out (may be NULL): device time of the
 * GEMMs. */
int  gl_mft_psf(gl_ctx* ctx, const float* pupil, int32_t N, int32_t M, float q_ref,
                float lambda_ref, const float* lambdas, int32_t n_lambda, int32_t precision,
                float* out_psf, float** dev_out, float* ms_out);

#ifdef __cplusplus
}
#endif
#endif /* GHOSTLIGHT_AMD_H */
