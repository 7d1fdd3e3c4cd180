// sourcemap.hip — lens calibration (covered field, image circle, first-order
// optics) and the screen <-> field-angle source map, with every probe ray
// traced on the device.
//
// The searches follow lens_calibration.cpp:19-443, spawn_probe.cpp:29-245,
// source_map.cpp:22-519 and newton_aim.h:19-63.  Their control flow is
// sequential (every step depends on the probe before it), so it stays on the
// host; what each step measures is a batch of rays:
//
//   * a probe of the spawn disc -- one sampling window at one field angle -- is
//     ONE launch of disc_probe_kernel, which traces the window's n x n cell
//     centres AND every cell's ess x ess boundary sub-samples at once (a
//     boundary cell is known only after the centres, but tracing all
//     sub-samples costs nothing next to a second round trip);
//   * the marginal-ray bisections of every azimuth run side by side, one lane
//     per azimuth, in ONE launch of marginal_kernel.
//
// The survivor sums are formed on the host in the reference's order, so the
// area, the landing mean and every decision taken on them are the reference's
// to the bit wherever the per-ray landings are (the device tracer reproduces
// the CPU tracer's landings exactly).  The one difference: a polygonal stop's
// pupil mask is tested in half-plane form (mask_contains) rather than through
// atan2f/fmodf/cosf, which can move a sample lying on a blade edge to the
// other side.
#include "gl_runtime.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

namespace gl {

namespace {

constexpr float kSpawnOffset = 20.0f;   // spawn plane ahead of the front vertex (spawn_plane.h:36)

// One window of the spawn disc at one field angle.  Samples are numbered
// cell centres first (row-major), then every cell's sub-samples (cell
// row-major, sub-sample sj-major), ess*ess per cell.
struct ProbeDesc {
    float dx, dy, dz;        // launch direction, normalised on the host
    float sdx, sdy;          // spawn shift of the disc (spawn_plane.h:47-51)
    float u_org, v_org;      // lower corner of the sampled square (disc units)
    float step;              // cell size
    int32_t n, ess;          // cells per axis, sub-samples per cell axis
    int32_t out;             // index of the first sample in the output arrays
    int32_t pad;
};

__global__ __launch_bounds__(256) void disc_probe_kernel(const DevSurf* __restrict__ S, const float* __restrict__ poly,
                                                         const float* __restrict__ arena, int n_surf, DevMask mask,
                                                         const float* __restrict__ iat, float front_R, float z_spawn,
                                                         float lam, const ProbeDesc* __restrict__ P,
                                                         float2* __restrict__ land, unsigned char* __restrict__ ok) {
    const ProbeDesc p = P[blockIdx.y];
    const int nn = p.n * p.n, ee = p.ess * p.ess;
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nn * (1 + ee)) return;
    float u, v;
    if (idx < nn) {
        const int i = idx % p.n, j = idx / p.n;
        v = p.v_org + ((float)j + 0.5f) * p.step;
        u = p.u_org + ((float)i + 0.5f) * p.step;
    } else {
        const int r = idx - nn, k = r / ee, s = r % ee;
        const int i = k % p.n, j = k / p.n, si = s % p.ess, sj = s / p.ess;
        const float u0 = p.u_org + (float)i * p.step;
        const float v0 = p.v_org + (float)j * p.step;
        const float sub = p.step / (float)p.ess;
        v = v0 + ((float)sj + 0.5f) * sub;
        u = u0 + ((float)si + 0.5f) * sub;
    }
    float x = 0.0f, y = 0.0f;
    unsigned char good = 0;
    if (mask_contains(mask, u, v)) {
        const TraceRes r = trace_primary<false>(as_const(S), n_surf, as_const(poly), arena,
                                                f3(u * front_R + p.sdx, v * front_R + p.sdy, z_spawn),
                                                f3(p.dx, p.dy, p.dz), lam, as_const(iat), true, nullptr);
        if (r.status == GL_TRACE_OK && isfinite(r.px) && isfinite(r.py)) { good = 1; x = r.px; y = r.py; }
    }
    land[p.out + idx] = make_float2(x, y);
    ok[p.out + idx] = good;
}

// marginal_semi_along (lens_calibration.cpp:45-69), one lane per direction:
// the largest entrance-plane height along (ux, uy) at which an axial ray
// still reaches the sensor, by 40 bisection steps.
__global__ __launch_bounds__(64) void marginal_kernel(const DevSurf* __restrict__ S, const float* __restrict__ poly,
                                                      const float* __restrict__ arena, int n_surf,
                                                      const float* __restrict__ iat, float front_R, float z_spawn,
                                                      float lam, const float2* __restrict__ dirs, int k,
                                                      float* __restrict__ out) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= k) return;
    auto reaches = [&](float hx, float hy) {
        const TraceRes r = trace_primary<false>(as_const(S), n_surf, as_const(poly), arena, f3(hx, hy, z_spawn),
                                                f3(0.0f, 0.0f, 1.0f), lam, as_const(iat), true, nullptr);
        return r.status == GL_TRACE_OK;
    };
    const float ux = dirs[t].x, uy = dirs[t].y;
    float res;
    if (!reaches(0.0f, 0.0f)) {
        res = 0.0f;
    } else if (reaches(front_R * ux, front_R * uy)) {
        res = front_R;
    } else {
        float lo = 0.0f, hi = front_R;
        for (int it = 0; it < 40; ++it) {
            const float mid = 0.5f * (lo + hi);
            if (reaches(mid * ux, mid * uy)) lo = mid;
            else hi = mid;
        }
        res = lo;
    }
    out[t] = res;
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------

struct Window {
    float cu = 0.0f, cv = 0.0f;   // centre, normalised disc coordinates
    float r = 1.0f;               // half-extent of the sampled square
};

struct DiscArea {
    float area = 0.0f;            // coverage-weighted absolute disc area
    float mean_x = 0.0f, mean_y = 0.0f;
    bool at_edge = false;         // a survivor on the sampled rim
    Window fit;                   // survivor bounding box, padded a cell
};

// The device side of every probe: a lens variant, its pupil mask and the
// index slice at the probe wavelength.
struct Prober {
    gl_ctx* c = nullptr;
    const gl_lens* L = nullptr;
    LensDev D;
    int n = 0;
    DevMask mask{};
    float front_R = 0.0f, z_spawn = 0.0f, lam = 0.0f;
    const float* d_iat = nullptr;
    int launches = 0;
    std::vector<float2> land;
    std::vector<unsigned char> ok;

    int init(gl_ctx* ctx, const gl_lens* lens, float lambda_nm) {
        c = ctx; L = lens; lam = lambda_nm;
        n = (int)L->host.size();
        int st;
        if ((st = lens_sync(const_cast<gl_lens*>(L), 0.0f, c->stream, &D))) return st;
        mask = build_mask(L->host.data(), n, 0, 0.0f);
        front_R = L->host[0].semi_aperture;
        z_spawn = L->host[0].z - kSpawnOffset;
        std::vector<float> at(n), bf(n);
        gl_spec sp{lam, 0, 0, 0};
        ior_tables(L->host.data(), n, &sp, 1, at.data(), bf.data());
        DevBuf& bi = c->buf("probe_iat");
        if ((st = bi.ensure((size_t)n * 4))) return st;
        GL_HIP(hipMemcpyAsync(bi.p, at.data(), (size_t)n * 4, hipMemcpyHostToDevice, c->stream));
        GL_HIP(hipStreamSynchronize(c->stream));
        d_iat = bi.as<float>();
        return GL_OK;
    }

    static int samples_of(const ProbeDesc& d) { return d.n * d.n * (1 + d.ess * d.ess); }

    // Trace k probes in one launch; land/ok hold every sample afterwards.
    int run(std::vector<ProbeDesc>& descs) {
        int total = 0, widest = 0;
        for (auto& d : descs) { d.out = total; total += samples_of(d); widest = std::max(widest, samples_of(d)); }
        int st;
        DevBuf &bd = c->buf("probe_desc"), &bl = c->buf("probe_land"), &bo = c->buf("probe_ok");
        if ((st = bd.ensure(descs.size() * sizeof(ProbeDesc))) || (st = bl.ensure((size_t)total * sizeof(float2))) ||
            (st = bo.ensure((size_t)total)))
            return st;
        GL_HIP(hipMemcpyAsync(bd.p, descs.data(), descs.size() * sizeof(ProbeDesc), hipMemcpyHostToDevice, c->stream));
        hipLaunchKernelGGL(disc_probe_kernel, dim3((unsigned)((widest + 255) / 256), (unsigned)descs.size()), dim3(256),
                           0, c->stream, D.surf, D.poly, D.arena, n, mask, d_iat, front_R, z_spawn, lam,
                           bd.as<ProbeDesc>(), bl.as<float2>(), bo.as<unsigned char>());
        GL_HIP(hipGetLastError());
        land.resize(total);
        ok.resize(total);
        GL_HIP(hipMemcpyAsync(land.data(), bl.p, (size_t)total * sizeof(float2), hipMemcpyDeviceToHost, c->stream));
        GL_HIP(d2h_sync(c, ok.data(), bo.p, (size_t)total));
        ++launches;
        return GL_OK;
    }

    // marginal_semi_along for every direction at once (0 for a lens whose
    // front element has no aperture).
    int marginal(const std::vector<float2>& dirs, std::vector<float>& out) {
        out.assign(dirs.size(), 0.0f);
        if (dirs.empty() || front_R <= 0.0f) return GL_OK;
        int st;
        DevBuf &bd = c->buf("probe_dirs"), &bo = c->buf("probe_marg");
        if ((st = bd.ensure(dirs.size() * sizeof(float2))) || (st = bo.ensure(dirs.size() * 4))) return st;
        GL_HIP(hipMemcpyAsync(bd.p, dirs.data(), dirs.size() * sizeof(float2), hipMemcpyHostToDevice, c->stream));
        hipLaunchKernelGGL(marginal_kernel, dim3((unsigned)((dirs.size() + 63) / 64)), dim3(64), 0, c->stream, D.surf,
                           D.poly, D.arena, n, d_iat, front_R, z_spawn, lam, bd.as<float2>(), (int)dirs.size(),
                           bo.as<float>());
        GL_HIP(hipGetLastError());
        GL_HIP(d2h_sync(c, out.data(), bo.p, dirs.size() * 4));
        ++launches;
        return GL_OK;
    }
};

// Vec3f(tan(ah), tan(av), 1).normalized() with the reference's operation
// order (vec3.h:63-70): length = sqrt(x*x + y*y + z*z), then x * (1 / length).
void launch_dir(float ah, float av, float* d) {
    const float x = std::tan(ah), y = std::tan(av), z = 1.0f;
    const float l = std::sqrt(x * x + y * y + z * z);
    if (l > 1e-12f) {
        const float s = 1.0f / l;
        d[0] = x * s; d[1] = y * s; d[2] = z * s;
    } else {
        d[0] = d[1] = d[2] = 0.0f;
    }
}

ProbeDesc window_desc(float ah, float av, const Window& win, int grid, int ess) {
    ProbeDesc d{};
    d.n = grid;
    d.ess = ess;
    d.step = 2.0f * win.r / (float)grid;
    const float bx = std::tan(ah), by = std::tan(av);
    d.sdx = -kSpawnOffset * bx;
    d.sdy = -kSpawnOffset * by;
    launch_dir(ah, av, &d.dx);
    d.u_org = win.cu - win.r;
    d.v_org = win.cv - win.r;
    return d;
}

// A single ray launched from the entrance-plane origin at one field angle
// (trace_from_ok with zero height): a 1 x 1 window centred on (0, 0).
ProbeDesc origin_desc(float ah, float av) {
    ProbeDesc d{};
    d.n = 1; d.ess = 1;
    d.step = 1.0f;
    d.u_org = -0.5f; d.v_org = -0.5f;
    launch_dir(ah, av, &d.dx);
    return d;
}

// Pass 2 of probe_window (spawn_probe.cpp:108-183) over the traced samples.
DiscArea window_sums(const ProbeDesc& d, const float2* land, const unsigned char* ok) {
    const int n = d.n, ess = d.ess;
    const float step = d.step;
    const float u_org = d.u_org, v_org = d.v_org;
    DiscArea out;
    const float cell_area = step * step;
    const float sub_w = 1.0f / (float)(ess * ess);
    float lo_u = 1e30f, hi_u = -1e30f, lo_v = 1e30f, hi_v = -1e30f;
    auto note = [&](float u, float v, float w, float x, float y) {
        out.area += w * cell_area;
        out.mean_x += w * cell_area * x;
        out.mean_y += w * cell_area * y;
        lo_u = std::min(lo_u, u); hi_u = std::max(hi_u, u);
        lo_v = std::min(lo_v, v); hi_v = std::max(hi_v, v);
    };
    const int nn = n * n;
    for (int j = 0; j < n; ++j) {
        for (int i = 0; i < n; ++i) {
            const int k = j * n + i;
            const bool here = ok[k] != 0;
            const float uc = u_org + ((float)i + 0.5f) * step;
            const float vc = v_org + ((float)j + 0.5f) * step;
            const bool on_rim = (i == 0 || i == n - 1 || j == 0 || j == n - 1);
            if (here && on_rim) out.at_edge = true;
            bool edge = here && on_rim;
            const int di[4] = {1, -1, 0, 0};
            const int dj[4] = {0, 0, 1, -1};
            for (int q = 0; q < 4 && !edge; ++q) {
                const int ni = i + di[q], nj = j + dj[q];
                if (ni < 0 || ni >= n || nj < 0 || nj >= n) continue;
                if ((ok[nj * n + ni] != 0) != here) edge = true;
            }
            if (!edge) {
                if (here) note(uc, vc, 1.0f, land[k].x, land[k].y);
                continue;
            }
            const float u0 = u_org + (float)i * step;
            const float v0 = v_org + (float)j * step;
            const float sub = step / (float)ess;
            const int base = nn + k * ess * ess;
            for (int sj = 0; sj < ess; ++sj) {
                const float v = v0 + ((float)sj + 0.5f) * sub;
                for (int si = 0; si < ess; ++si) {
                    const float u = u0 + ((float)si + 0.5f) * sub;
                    const int q = base + sj * ess + si;
                    if (!ok[q]) continue;
                    note(u, v, sub_w, land[q].x, land[q].y);
                }
            }
        }
    }
    if (out.area > 0.0f) {
        out.mean_x /= out.area;
        out.mean_y /= out.area;
        out.fit.cu = 0.5f * (lo_u + hi_u);
        out.fit.cv = 0.5f * (lo_v + hi_v);
        out.fit.r = std::max(std::max(hi_u - out.fit.cu, hi_v - out.fit.cv), step) + step;
    }
    return out;
}

constexpr int kProbeGrid = 17;
constexpr int kProbeEdgeSS = 2;
constexpr float kDeepStopFill = 0.45f;
constexpr float kWindowDilation = 5.0f;

// initial_probe_window (spawn_probe.cpp:48-54).
Window initial_probe_window(float pupil_fill) {
    Window w{0.0f, 0.0f, 1.0f};
    if (pupil_fill > 0.0f && pupil_fill < kDeepStopFill) w.r = std::min(1.0f, pupil_fill * kWindowDilation);
    return w;
}

// probe_window (spawn_probe.cpp:57-183): one device launch, host sums.
int probe_window(Prober& P, float ah, float av, const Window& win, int grid, int ess, DiscArea* out) {
    std::vector<ProbeDesc> d{window_desc(ah, av, win, grid > 0 ? grid : kProbeGrid, ess > 0 ? ess : kProbeEdgeSS)};
    int st;
    if ((st = P.run(d))) return st;
    *out = window_sums(d[0], P.land.data(), P.ok.data());
    return GL_OK;
}

// probe_disc (spawn_probe.cpp:186-245): re-centre on the beam, widen when it
// outruns the window.
int probe_disc(Prober& P, float ah, float av, Window* track, int grid, int ess, DiscArea* out) {
    Window win = *track;
    if (win.r >= 1.0f) { win.cu = 0.0f; win.cv = 0.0f; win.r = 1.0f; }
    DiscArea best;
    int st;
    for (int attempt = 0; attempt < 6; ++attempt) {
        DiscArea d;
        if ((st = probe_window(P, ah, av, win, grid, ess, &d))) return st;
        if (d.area > 0.0f) best = d;
        if (d.area > 0.0f && !d.at_edge) {
            *track = d.fit;
            track->r = win.r;
            *out = d;
            return GL_OK;
        }
        if (win.r >= 1.0f) break;
        if (d.area > 0.0f && attempt < 3) {
            win.cu = d.fit.cu;
            win.cv = d.fit.cv;
            continue;
        }
        win.r = std::min(1.0f, win.r * 2.5f);
        if (win.r >= 1.0f) { win.cu = 0.0f; win.cv = 0.0f; }
    }
    track->cu = win.cu; track->cv = win.cv; track->r = win.r;
    *out = best;
    return GL_OK;
}

// ---- calibration (lens_calibration.cpp) -----------------------------------

const gl_surface* stop_surface(const gl_lens* L) {
    for (const auto& s : L->host)
        if (s.is_stop) return &s;
    return nullptr;
}

const gl_surface* bladed_stop(const gl_lens* L) {
    const gl_surface* s = stop_surface(L);
    return (s && s->aperture_shape == GL_APERTURE_POLYGON && s->aperture_blades >= 3) ? s : nullptr;
}

const float kLadderDeg[] = {0.0f, 4.0f, 8.0f, 14.0f, 22.0f, 32.0f, 45.0f, 60.0f};
constexpr int kNLadder = 8;
constexpr float kRefineTolDeg = 0.15f;
constexpr int kMaxRefine = 7;
constexpr float kCircleTolDeg = 1.0f;
constexpr int kMaxRefineCircle = 2;
constexpr float kMapThreshold = 0.90f;
constexpr float kImageCircleThreshold = 0.05f;

struct FieldSolve {
    float angle_rad = 0.0f;
    float half_mm = 0.0f;
    bool crossed = false;
};

// resolve_axis (lens_calibration.cpp:171-307): one walk out along the field
// angle brackets both thresholds, then Illinois regula falsi refines each.
int resolve_axis(Prober& P, float pupil_fill, bool horizontal, FieldSolve* map_out, FieldSolve* circle_out) {
    Window track = initial_probe_window(pupil_fill);
    const float DEG = (float)(M_PI / 180.0);
    *map_out = FieldSolve{};
    *circle_out = FieldSolve{};
    DiscArea axial;
    float axial_r = -1.0f;
    int st = GL_OK;
    auto at = [&](float angle, DiscArea* d) -> int {
        int e;
        if ((e = probe_disc(P, horizontal ? angle : 0.0f, horizontal ? 0.0f : angle, &track, 0, 0, d))) return e;
        if (track.r != axial_r) {
            Window ax{0.0f, 0.0f, track.r};
            if ((e = probe_window(P, 0.0f, 0.0f, ax, 0, 0, &axial))) return e;
            axial_r = track.r;
        }
        return GL_OK;
    };
    DiscArea d0;
    if ((st = at(0.0f, &d0))) return st;
    if (axial.area <= 0.0f) return GL_OK;

    int map_cross = -1, circle_cross = -1;
    float f_prev = 1.0f;
    float f_at_map[2] = {1.0f, 0.0f}, f_at_circle[2] = {1.0f, 0.0f};
    Window seed_map = track, seed_circle = track;
    for (int i = 1; i < kNLadder; ++i) {
        const Window before = track;
        DiscArea d;
        if ((st = at(kLadderDeg[i] * DEG, &d))) return st;
        const float f = d.area / axial.area;
        if (map_cross < 0 && f < kMapThreshold) {
            map_cross = i - 1;
            f_at_map[0] = f_prev; f_at_map[1] = f;
            seed_map = before;
        }
        if (circle_cross < 0 && f < kImageCircleThreshold) {
            circle_cross = i - 1;
            f_at_circle[0] = f_prev; f_at_circle[1] = f;
            seed_circle = before;
            break;
        }
        f_prev = f;
    }

    auto refine = [&](int cross, float thr, const float f_ends[2], const Window& seed, int max_steps, float tol_deg,
                      FieldSolve* r) -> int {
        *r = FieldSolve{};
        auto probe_at = [&](float angle, DiscArea* d) -> int {
            track.cu = seed.cu;
            track.cv = seed.cv;
            return at(angle, d);
        };
        int e;
        if (cross < 0) {
            DiscArea d;
            if ((e = probe_at(kLadderDeg[kNLadder - 1] * DEG, &d))) return e;
            r->angle_rad = kLadderDeg[kNLadder - 1] * DEG;
            r->half_mm = std::abs(horizontal ? d.mean_x : d.mean_y);
            r->crossed = false;
            return GL_OK;
        }
        float lo = kLadderDeg[cross] * DEG, hi = kLadderDeg[cross + 1] * DEG;
        float f_lo = f_ends[0] - thr, f_hi = f_ends[1] - thr;
        float t_lo = f_lo, t_hi = f_hi;
        const float tol = tol_deg * DEG;
        for (int it = 0; it < max_steps && (hi - lo) > tol; ++it) {
            const float denom = f_lo - f_hi;
            float mid = (std::abs(denom) < 1e-12f) ? 0.5f * (lo + hi) : lo + (hi - lo) * f_lo / denom;
            const float edge = 0.01f * (hi - lo);
            mid = std::min(std::max(mid, lo + edge), hi - edge);
            DiscArea d;
            if ((e = probe_at(mid, &d))) return e;
            const float f_mid = d.area / axial.area - thr;
            if (f_mid >= 0.0f) { lo = mid; f_lo = f_mid; t_lo = f_mid; f_hi *= 0.5f; }
            else { hi = mid; f_hi = f_mid; t_hi = f_mid; f_lo *= 0.5f; }
        }
        const float span = t_lo - t_hi;
        float root = (std::abs(span) < 1e-12f) ? 0.5f * (lo + hi) : lo + (hi - lo) * t_lo / span;
        root = std::min(std::max(root, lo), hi);
        DiscArea d;
        if ((e = probe_at(root, &d))) return e;
        if (d.area <= 0.0f) {
            root = lo;
            if ((e = probe_at(lo, &d))) return e;
        }
        r->angle_rad = root;
        r->half_mm = std::abs(horizontal ? d.mean_x : d.mean_y);
        r->crossed = true;
        return GL_OK;
    };
    if ((st = refine(map_cross, kMapThreshold, f_at_map, seed_map, kMaxRefine, kRefineTolDeg, map_out))) return st;
    if ((st = refine(circle_cross, kImageCircleThreshold, f_at_circle, seed_circle, kMaxRefineCircle, kCircleTolDeg,
                     circle_out)))
        return st;
    if (circle_out->angle_rad < map_out->angle_rad) *circle_out = *map_out;
    if (circle_out->half_mm < map_out->half_mm) circle_out->half_mm = map_out->half_mm;
    return GL_OK;
}

// axes_are_symmetric (lens_calibration.cpp:312-340).
bool axes_are_symmetric(const gl_lens* L) {
    static const float I[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    for (const auto& s : L->host) {
        if (s.form == GL_FORM_CYLINDRICAL) return false;
        if (s.aperture_aspect != 1.0f) return false;
        if (s.decenter_x != 0.0f || s.decenter_y != 0.0f) return false;
        for (int k = 0; k < 9; ++k)
            if (s.rot[k] != I[k]) return false;
        if (s.aperture_shape == GL_APERTURE_POLYGON) {
            if (s.aperture_blades % 4 != 0) return false;
            if (s.aperture_rotation_rad != 0.0f) return false;
        } else if (s.aperture_shape != GL_APERTURE_CIRCLE) {
            return false;
        }
    }
    return true;
}

// calibrate_lens (lens_calibration.cpp:344-443).
int calibrate(Prober& P, float d_line_nm, float file_focal_length, gl_calibration* out) {
    gl_calibration cal{};
    cal.pupil_area_frac = 1.0f;
    const gl_lens* L = P.L;
    if (L->host.empty()) { *out = cal; return GL_OK; }
    int st;
    // Paraxial focal lengths: one 1-degree ray per axis, both in one launch.
    {
        const float theta = (float)(1.0 * M_PI / 180.0);
        std::vector<ProbeDesc> d{origin_desc(theta, 0.0f), origin_desc(0.0f, theta)};
        if ((st = P.run(d))) return st;
        for (int a = 0; a < 2; ++a) {
            const int k = d[a].out;
            float efl = 0.0f;
            if (P.ok[k]) {
                const float img_h = a == 0 ? P.land[k].x : P.land[k].y;
                if (std::isfinite(img_h)) efl = std::abs(img_h) / std::tan(theta);
            }
            (a == 0 ? cal.focal_length_x : cal.focal_length_y) = efl;
        }
    }
    // Entrance pupil: per-axis support over a bladed stop's azimuths, or the
    // two principal meridians of a round one.
    if (const gl_surface* stop = bladed_stop(L)) {
        const int n = stop->aperture_blades * 8;
        const float rot = stop->aperture_rotation_rad;
        std::vector<float> th;
        for (int i = 0; i < n; ++i) th.push_back(rot + (float)(2.0 * M_PI) * i / (float)n);
        th.push_back(0.0f);
        th.push_back((float)(M_PI / 2.0));
        std::vector<float2> dirs(th.size());
        for (size_t i = 0; i < th.size(); ++i) dirs[i] = make_float2(std::cos(th[i]), std::sin(th[i]));
        std::vector<float> h;
        if ((st = P.marginal(dirs, h))) return st;
        float bx = 0.0f, by = 0.0f;
        for (size_t i = 0; i < th.size(); ++i) {
            bx = std::max(bx, h[i] * std::fabs(dirs[i].x));
            by = std::max(by, h[i] * std::fabs(dirs[i].y));
        }
        cal.entrance_pupil_semi_x = bx;
        cal.entrance_pupil_semi_y = by;
        gl_blade_profile prof;
        gl_blade_profile_build(stop->aperture_blades, stop->aperture_rotation_rad, stop->aperture_curvature,
                               stop->aperture_twist, stop->aperture_notch_rad, stop->aperture_notch_angle_rad, &prof);
        cal.pupil_area_frac = prof.area_frac;
    } else {
        std::vector<float> h;
        if ((st = P.marginal({make_float2(1.0f, 0.0f), make_float2(0.0f, 1.0f)}, h))) return st;
        cal.entrance_pupil_semi_x = h[0];
        cal.entrance_pupil_semi_y = h[1];
    }
    // Covered field and image circle.
    {
        const float front_R = P.front_R;
        const float fill_x = (front_R > 0.0f) ? cal.entrance_pupil_semi_x / front_R : 1.0f;
        const float fill_y = (front_R > 0.0f) ? cal.entrance_pupil_semi_y / front_R : 1.0f;
        FieldSolve map_h, map_v, circle_h, circle_v;
        if ((st = resolve_axis(P, std::min(fill_x, fill_y), true, &map_h, &circle_h))) return st;
        if (axes_are_symmetric(L)) {
            map_v = map_h;
            circle_v = circle_h;
        } else if ((st = resolve_axis(P, std::min(fill_x, fill_y), false, &map_v, &circle_v))) {
            return st;
        }
        cal.max_half_angle_h = map_h.angle_rad;
        cal.max_half_angle_v = map_v.angle_rad;
        cal.sensor_half_w = map_h.half_mm;
        cal.sensor_half_h = map_v.half_mm;
        cal.image_circle_semi_w = circle_h.half_mm;
        cal.image_circle_semi_h = circle_v.half_mm;
        if (!map_h.crossed || !map_v.crossed)
            std::fprintf(stderr,
                         "[ghostlight_amd] warning: covered-field probe never fell to %.0f%% of axial within %.0f deg; "
                         "the calibrated field is the search ceiling.\n",
                         (double)(kMapThreshold * 100.0f), (double)kLadderDeg[kNLadder - 1]);
        const float MIN_FIELD_RAD = (float)(0.05 * M_PI / 180.0);
        auto floor_axis = [&](float* angle, float* half) {
            if (*angle >= MIN_FIELD_RAD && *half > 0.0f) return;
            std::fprintf(stderr, "[ghostlight_amd] warning: covered field resolved to nothing; falling back to the "
                                 "lens file's focal length.\n");
            *angle = MIN_FIELD_RAD;
            *half = (file_focal_length > 0.0f) ? file_focal_length * std::tan(MIN_FIELD_RAD) : std::tan(MIN_FIELD_RAD);
        };
        floor_axis(&cal.max_half_angle_h, &cal.sensor_half_w);
        floor_axis(&cal.max_half_angle_v, &cal.sensor_half_h);
        if (cal.image_circle_semi_w <= 0.0f) cal.image_circle_semi_w = cal.sensor_half_w;
        if (cal.image_circle_semi_h <= 0.0f) cal.image_circle_semi_h = cal.sensor_half_h;
    }
    if (cal.entrance_pupil_semi_x > 0.0f) cal.f_number_x = cal.focal_length_x / (2.0f * cal.entrance_pupil_semi_x);
    if (cal.entrance_pupil_semi_y > 0.0f) cal.f_number_y = cal.focal_length_y / (2.0f * cal.entrance_pupil_semi_y);
    *out = cal;
    return GL_OK;
}

// ---- source map (source_map.cpp) ------------------------------------------

constexpr float kLambdaNm = 587.56f;
constexpr int kSolveGrid = 17;
constexpr int kMaxIter = 8;
constexpr float kFdStep = 5e-3f;
constexpr float kTolMm = 1e-3f;
constexpr float kAcceptFrac = 0.01f;
constexpr int kBisectSteps = 6;
constexpr float kBisectTolMm = 0.02f;
constexpr float kMaxExtendRad = 1.55f;
constexpr int kSolveEdgeSS = 2;
constexpr float kWarmStepRad = 0.10f;

// The landing function with a window that follows the pupil, and the last
// eight evaluations remembered with the window they were taken through
// (source_map.cpp:58-126).
struct SolveProbe {
    Prober* P;
    Window win;
    int evals = 0, calls = 0;
    int err = GL_OK;
    struct Entry { float ax, ay, cu, cv, r, x, y; Window out; bool ok; };
    static constexpr int kCache = 8;
    Entry cache[kCache]{};
    int n_cached = 0, next_slot = 0;

    bool at(float ax, float ay, float& x, float& y) {
        ++calls;
        if (err) return false;
        for (int i = 0; i < n_cached; ++i) {
            const Entry& e = cache[i];
            if (e.ax == ax && e.ay == ay && e.cu == win.cu && e.cv == win.cv && e.r == win.r) {
                if (!e.ok) return false;
                win = e.out; x = e.x; y = e.y;
                return true;
            }
        }
        ++evals;
        const Window in = win;
        Window w = in;
        DiscArea d;
        if ((err = probe_disc(*P, ax, ay, &w, kSolveGrid, kSolveEdgeSS, &d))) return false;
        const bool ok = (d.area > 0.0f) && std::isfinite(d.mean_x) && std::isfinite(d.mean_y);
        Entry& e = cache[next_slot];
        e.ax = ax; e.ay = ay;
        e.cu = in.cu; e.cv = in.cv; e.r = in.r;
        e.x = d.mean_x; e.y = d.mean_y; e.out = w; e.ok = ok;
        next_slot = (next_slot + 1) % kCache;
        if (n_cached < kCache) ++n_cached;
        if (!ok) return false;
        win = w;
        x = d.mean_x;
        y = d.mean_y;
        return true;
    }
};

// measure_jacobian (source_map.cpp:131-163).
bool measure_jacobian(SolveProbe& P, float ax, float ay, float* out_land, float* out_J) {
    float cx, cy;
    if (!P.at(ax, ay, cx, cy)) return false;
    auto slope = [&](float dax, float day, float* dx, float* dy) -> bool {
        float px, py;
        if (P.at(ax + dax, ay + day, px, py)) {
            *dx = (px - cx) / kFdStep;
            *dy = (py - cy) / kFdStep;
            return true;
        }
        if (P.at(ax - dax, ay - day, px, py)) {
            *dx = (cx - px) / kFdStep;
            *dy = (cy - py) / kFdStep;
            return true;
        }
        return false;
    };
    float dxdax, dydax, dxday, dyday;
    if (!slope(kFdStep, 0.0f, &dxdax, &dydax)) return false;
    if (!slope(0.0f, kFdStep, &dxday, &dyday)) return false;
    out_land[0] = cx;
    out_land[1] = cy;
    out_J[0] = dxdax; out_J[1] = dxday;
    out_J[2] = dydax; out_J[3] = dyday;
    return true;
}

// newton_aim (newton_aim.h:19-63): 2x2 finite-difference Newton with clamped
// steps, reverting to the best iterate.
template <typename LandingFn>
bool newton_aim(LandingFn&& landing, float target_x, float target_y, float& ax, float& ay, float& out_residual,
                int max_iter, float fd_step, float step_cap, float tol) {
    float best_ax = ax, best_ay = ay, best_res = 1e30f;
    auto record = [&](float x, float y) {
        const float rx = x - target_x, ry = y - target_y;
        const float res = std::sqrt(rx * rx + ry * ry);
        if (res < best_res) { best_res = res; best_ax = ax; best_ay = ay; }
        return res;
    };
    float cx, cy;
    if (!landing(ax, ay, cx, cy)) { out_residual = 1e30f; return false; }
    for (int it = 0; it < max_iter; ++it) {
        const float rx = cx - target_x, ry = cy - target_y;
        if (record(cx, cy) < tol) break;
        float xx, xy, yx, yy;
        if (!landing(ax + fd_step, ay, xx, xy)) break;
        if (!landing(ax, ay + fd_step, yx, yy)) break;
        const float Jxx = (xx - cx) / fd_step, Jyx = (xy - cy) / fd_step;
        const float Jxy = (yx - cx) / fd_step, Jyy = (yy - cy) / fd_step;
        const float det = Jxx * Jyy - Jxy * Jyx;
        if (std::fabs(det) < 1e-12f) break;
        float dax = (-rx * Jyy + ry * Jxy) / det;
        float day = (rx * Jyx - ry * Jxx) / det;
        dax = std::max(-step_cap, std::min(step_cap, dax));
        day = std::max(-step_cap, std::min(step_cap, day));
        ax += dax; ay += day;
        if (!landing(ax, ay, cx, cy)) break;
    }
    if (landing(ax, ay, cx, cy)) record(cx, cy);
    ax = best_ax; ay = best_ay;
    out_residual = best_res;
    return best_res < tol;
}

uint64_t fnv1a(const void* data, size_t n, uint64_t h = 1469598103934665603ULL) {
    const unsigned char* p = static_cast<const unsigned char*>(data);
    for (size_t i = 0; i < n; ++i) { h ^= p[i]; h *= 1099511628211ULL; }
    return h;
}

// One-entry memo (source_map.cpp:169-177), keyed by the prescription, the
// coating arena, the calibration and the request.
std::mutex g_memo_mu;
bool g_memo_valid = false;
uint64_t g_memo_key = 0;
gl_source_map g_memo_val;
bool g_warned_fallback = false;

// solve_source_angle (source_map.cpp:182-452).
int solve(Prober& Pr, const gl_calibration& calib, float ndc_x, float ndc_y, float eff_half_w, float eff_half_h,
          gl_source_map* result) {
    const gl_lens* L = Pr.L;
    uint64_t key = fnv1a(L->host.data(), L->host.size() * sizeof(gl_surface));
    key = fnv1a(L->arena.data(), L->arena.size() * sizeof(float), key);
    key = fnv1a(&calib, sizeof(gl_calibration), key);
    {
        const float req[4] = {ndc_x, ndc_y, eff_half_w, eff_half_h};
        key = fnv1a(req, sizeof(req), key);
    }
    {
        std::lock_guard<std::mutex> lk(g_memo_mu);
        if (g_memo_valid && g_memo_key == key) { *result = g_memo_val; return GL_OK; }
    }
    gl_source_map sm{};
    sm.eff_half_w = eff_half_w;
    sm.eff_half_h = eff_half_h;
    const float scale_w = (calib.sensor_half_w > 0.0f) ? eff_half_w / calib.sensor_half_w : 1.0f;
    const float scale_h = (calib.sensor_half_h > 0.0f) ? eff_half_h / calib.sensor_half_h : 1.0f;
    const float seed_ax = std::atan(ndc_x * scale_w * std::tan(calib.max_half_angle_h));
    const float seed_ay = std::atan(ndc_y * scale_h * std::tan(calib.max_half_angle_v));
    sm.inv_tan_w = 1.0f / (scale_w * std::tan(calib.max_half_angle_h));
    sm.inv_tan_h = 1.0f / (scale_h * std::tan(calib.max_half_angle_v));
    sm.angle_x = seed_ax;
    sm.angle_y = seed_ay;
    sm.status = GL_SMAP_FALLBACK;
    const int* evals_for_report = nullptr;

    auto memo = [&]() {
        std::lock_guard<std::mutex> lk(g_memo_mu);
        g_memo_valid = true; g_memo_key = key; g_memo_val = sm;
    };
    auto give_up = [&](const char* why) -> int {
        {
            std::lock_guard<std::mutex> lk(g_memo_mu);
            if (!g_warned_fallback) {
                g_warned_fallback = true;
                std::fprintf(stderr, "[ghostlight_amd] warning: source map could not be solved by tracing (%s); "
                                     "falling back to the assumed-focal-length form. Further occurrences are not "
                                     "reported.\n", why);
            }
        }
        sm.probe_evals = evals_for_report ? *evals_for_report : 0;
        memo();
        *result = sm;
        return GL_OK;
    };
    if (L->host.empty()) return give_up("lens has no surfaces");

    const float target_x = ndc_x * eff_half_w;
    const float target_y = ndc_y * eff_half_h;
    const float front_R = Pr.front_R;
    const float fill_x = (front_R > 0.0f) ? calib.entrance_pupil_semi_x / front_R : 1.0f;
    const float fill_y = (front_R > 0.0f) ? calib.entrance_pupil_semi_y / front_R : 1.0f;

    SolveProbe P{&Pr, initial_probe_window(std::min(fill_x, fill_y))};
    evals_for_report = &P.evals;
    auto landing = [&](float ax, float ay, float& x, float& y) -> bool { return P.at(ax, ay, x, y); };

    const Window win0 = initial_probe_window(std::min(fill_x, fill_y));
    auto reset_and_walk = [&](float ax, float ay) {
        P.win = win0;
        const float mag = std::sqrt(ax * ax + ay * ay);
        if (!(P.win.r < 1.0f) || !(mag > 0.0f)) return;
        const int steps = std::max(3, std::min(16, (int)std::ceil(mag / kWarmStepRad)));
        float tx, ty;
        for (int k = 1; k <= steps; ++k) {
            const float f = (float)k / (float)steps;
            P.at(ax * f, ay * f, tx, ty);
        }
    };
    const float seed_mag = std::sqrt(seed_ax * seed_ax + seed_ay * seed_ay);
    const float accept_mm = std::max(kTolMm, kAcceptFrac * std::min(std::fabs(eff_half_w), std::fabs(eff_half_h)));

    auto commit_solved = [&](float ax, float ay, float res) -> bool {
        float land[2], J[4];
        if (!measure_jacobian(P, ax, ay, land, J)) return false;
        sm.angle_x = ax;
        sm.angle_y = ay;
        sm.status = GL_SMAP_SOLVED;
        sm.residual_mm = res;
        sm.anchor_ax = ax; sm.anchor_ay = ay;
        sm.anchor_x = land[0]; sm.anchor_y = land[1];
        for (int i = 0; i < 4; ++i) sm.J[i] = J[i];
        sm.probe_evals = P.evals;
        return true;
    };
    float best_ax = seed_ax, best_ay = seed_ay, best_res = 1e30f;
    auto attempt = [&](float from_ax, float from_ay, float cap) -> bool {
        reset_and_walk(from_ax, from_ay);
        float ax = from_ax, ay = from_ay, res = 1e30f;
        newton_aim(landing, target_x, target_y, ax, ay, res, kMaxIter, kFdStep, cap, kTolMm);
        if (res < best_res) { best_res = res; best_ax = ax; best_ay = ay; }
        if (!(res < accept_mm)) return false;
        return commit_solved(ax, ay, res);
    };
    auto publish = [&]() -> int {
        if (P.err) return P.err;
        memo();
        *result = sm;
        return GL_OK;
    };

    if (attempt(seed_ax, seed_ay, std::max(0.02f, 0.5f * seed_mag))) return publish();
    if (P.err) return P.err;
    if (attempt(seed_ax, seed_ay, std::max(0.005f, 0.08f * seed_mag))) return publish();
    if (P.err) return P.err;

    // Continuation past the last landing the lens delivers.
    float ax_lo = 0.0f, ay_lo = 0.0f;
    bool have_anchor = false;
    float anchor_radius = 0.0f;
    Window anchor_win;
    const float tmag = std::sqrt(target_x * target_x + target_y * target_y);
    if (tmag > 0.0f) {
        const float ux = target_x / tmag, uy = target_y / tmag;
        const float rc = std::sqrt(ux * ux * calib.image_circle_semi_w * calib.image_circle_semi_w +
                                   uy * uy * calib.image_circle_semi_h * calib.image_circle_semi_h);
        float r_lo = 0.0f;
        float r_hi = 1.05f * rc;
        const float anchor_seed_ax = std::atan(0.98f * rc * ux * std::tan(calib.max_half_angle_h) /
                                               (calib.sensor_half_w > 0.0f ? calib.sensor_half_w : 1.0f));
        const float anchor_seed_ay = std::atan(0.98f * rc * uy * std::tan(calib.max_half_angle_v) /
                                               (calib.sensor_half_h > 0.0f ? calib.sensor_half_h : 1.0f));
        const float anchor_mag = std::sqrt(anchor_seed_ax * anchor_seed_ax + anchor_seed_ay * anchor_seed_ay);
        const float anchor_cap = std::max(0.02f, 0.5f * anchor_mag);
        reset_and_walk(anchor_seed_ax, anchor_seed_ay);
        auto try_target = [&](float tx, float ty, float& io_ax, float& io_ay) -> bool {
            float tax = io_ax, tay = io_ay, r = 1e30f;
            newton_aim(landing, tx, ty, tax, tay, r, kMaxIter, kFdStep, anchor_cap, kTolMm);
            if (!(r < accept_mm)) return false;
            io_ax = tax; io_ay = tay;
            return true;
        };
        {
            const float g = 0.98f * rc;
            float a = ax_lo, b = ay_lo;
            if (g > 0.0f && try_target(g * ux, g * uy, a, b)) {
                r_lo = g; ax_lo = a; ay_lo = b; have_anchor = true;
                anchor_radius = g; anchor_win = P.win;
            } else {
                r_hi = g;
            }
        }
        for (int i = 0; i < kBisectSteps && (r_hi - r_lo) > kBisectTolMm; ++i) {
            const float mid = 0.5f * (r_lo + r_hi);
            float a = ax_lo, b = ay_lo;
            if (try_target(mid * ux, mid * uy, a, b)) {
                r_lo = mid; ax_lo = a; ay_lo = b; have_anchor = true;
                anchor_radius = mid; anchor_win = P.win;
            } else {
                r_hi = mid;
            }
        }
    }
    if (P.err) return P.err;
    if (!have_anchor && !(seed_mag > 0.0f)) return give_up("nothing solves anywhere along the ray");
    if (have_anchor && attempt(ax_lo, ay_lo, std::max(0.02f, 0.5f * seed_mag))) return publish();
    if (P.err) return P.err;
    if (have_anchor && anchor_radius >= tmag && best_res < 1e29f) {
        reset_and_walk(best_ax, best_ay);
        if (commit_solved(best_ax, best_ay, best_res)) return publish();
    }
    if (P.err) return P.err;
    if (have_anchor) P.win = anchor_win;
    float land[2], J[4];
    if (!measure_jacobian(P, ax_lo, ay_lo, land, J)) {
        if (P.err) return P.err;
        return give_up("no convergent anchor to continue from");
    }
    const float det = J[0] * J[3] - J[1] * J[2];
    if (!(std::fabs(det) > 1e-12f)) return give_up("landing is singular in angle at the anchor");
    // theta(T) = theta_b + u_hat * atan(|u|), u = J^-1 (T - T_b).
    const float dx = target_x - land[0], dy = target_y - land[1];
    const float ux = (J[3] * dx - J[1] * dy) / det;
    const float uy = (-J[2] * dx + J[0] * dy) / det;
    const float um = std::sqrt(ux * ux + uy * uy);
    if (um > 1e-12f) {
        const float th = std::atan(um);
        sm.angle_x = ax_lo + (ux / um) * th;
        sm.angle_y = ay_lo + (uy / um) * th;
    } else {
        sm.angle_x = ax_lo;
        sm.angle_y = ay_lo;
    }
    sm.status = GL_SMAP_CONTINUED;
    sm.residual_mm = std::sqrt(dx * dx + dy * dy);
    sm.anchor_ax = ax_lo; sm.anchor_ay = ay_lo;
    sm.anchor_x = land[0]; sm.anchor_y = land[1];
    for (int i = 0; i < 4; ++i) sm.J[i] = J[i];
    sm.probe_evals = P.evals;
    memo();
    *result = sm;
    return GL_OK;
}

// ---- PSF pre-pass deep-pupil retries (PSFRenderer.cpp:109-122, 203-247, 383-410)

constexpr float kChiefLambdaNm = 587.56f;

// The axial transmitting window: one full-disc probe on axis; its survivor
// box seeds the tracked window of the deep-pupil retries.
int measure_axial(Prober& P, Window* win, bool* usable) {
    DiscArea d;
    int st;
    if ((st = probe_window(P, 0.0f, 0.0f, Window{0.0f, 0.0f, 1.0f}, 0, 0, &d))) return st;
    *usable = d.area > 0.0f;
    if (*usable) *win = d.fit;
    return GL_OK;
}

}  // namespace
}  // namespace gl

using namespace gl;

extern "C" {

int gl_calibrate_lens(gl_ctx* ctx, const gl_lens* lens, float d_line_nm, float file_focal_length, gl_calibration* out) {
    if (!lens || !out) return set_error(GL_ERR_INVALID, "null argument");
    if (!(d_line_nm > 0.0f)) return set_error(GL_ERR_INVALID, "probe wavelength must be positive");
    if (lens->host.empty()) {
        *out = gl_calibration{};
        out->pupil_area_frac = 1.0f;
        return GL_OK;
    }
    int st;
    gl_ctx* c = resolve_ctx(ctx, &st);
    if (!c) return st;
    std::lock_guard<std::mutex> g(c->mu);
    Prober P;
    if ((st = P.init(c, lens, d_line_nm))) return st;
    return calibrate(P, d_line_nm, file_focal_length, out);
}

int gl_solve_source_angle(gl_ctx* ctx, const gl_lens* lens, const gl_calibration* calib, float ndc_x, float ndc_y,
                          float eff_half_w, float eff_half_h, gl_source_map* out) {
    if (!lens || !calib || !out) return set_error(GL_ERR_INVALID, "null argument");
    if (!std::isfinite(ndc_x) || !std::isfinite(ndc_y)) return set_error(GL_ERR_INVALID, "source position must be finite");
    int st;
    gl_ctx* c = resolve_ctx(ctx, &st);
    if (!c) return st;
    std::lock_guard<std::mutex> g(c->mu);
    Prober P;
    if (!lens->host.empty() && (st = P.init(c, lens, kLambdaNm))) return st;
    P.L = lens;
    return solve(P, *calib, ndc_x, ndc_y, eff_half_w, eff_half_h, out);
}

int gl_psf_deep_pupil(gl_ctx* ctx, const gl_lens* lens, int32_t n, const float* angles, const float* targets,
                      int32_t target_mode, float* out6) {
    if (!lens || n < 0 || (n && (!angles || !out6)) || (target_mode && n && !targets))
        return set_error(GL_ERR_INVALID, "null argument");
    for (int i = 0; i < n; ++i) {
        float* o = out6 + 6 * (size_t)i;
        o[0] = angles[2 * i]; o[1] = angles[2 * i + 1];
        o[2] = o[3] = 0.0f; o[4] = 1e30f; o[5] = 0.0f;
    }
    if (n == 0 || lens->host.empty()) return GL_OK;
    int st;
    gl_ctx* c = resolve_ctx(ctx, &st);
    if (!c) return st;
    std::lock_guard<std::mutex> g(c->mu);
    Prober P;
    if ((st = P.init(c, lens, kChiefLambdaNm))) return st;
    Window axial_win;
    bool usable = false;
    if ((st = measure_axial(P, &axial_win, &usable))) return st;
    if (!usable) return GL_OK;
    for (int i = 0; i < n; ++i) {
        float* o = out6 + 6 * (size_t)i;
        const float sax = angles[2 * i], say = angles[2 * i + 1];
        const float mag = std::sqrt(sax * sax + say * say);
        Window track = axial_win;
        DiscArea d;
        if (!target_mode) {
            // Walk the window out to the angle, then the survivor mean there.
            if (track.r < 1.0f && mag > 0.0f) {
                const int steps = std::max(3, std::min(16, (int)std::ceil(mag / 0.10f)));
                for (int k = 1; k <= steps; ++k) {
                    const float f = (float)k / (float)steps;
                    if ((st = probe_disc(P, sax * f, say * f, &track, 0, 0, &d))) return st;
                }
            }
            if ((st = probe_disc(P, sax, say, &track, 0, 0, &d))) return st;
            if (d.area > 0.0f && std::isfinite(d.mean_x) && std::isfinite(d.mean_y)) {
                o[2] = d.mean_x; o[3] = d.mean_y; o[4] = 0.0f; o[5] = 1.0f;
            }
            continue;
        }
        // Aim the tracked window's survivor mean at the target.
        int err = GL_OK;
        auto landing = [&](float ax, float ay, float& x, float& y) -> bool {
            if (err) return false;
            Window w = track;
            DiscArea dd;
            if ((err = probe_disc(P, ax, ay, &w, 0, 0, &dd))) return false;
            if (!(dd.area > 0.0f) || !std::isfinite(dd.mean_x) || !std::isfinite(dd.mean_y)) return false;
            track = w;
            x = dd.mean_x; y = dd.mean_y;
            return true;
        };
        if (track.r < 1.0f && mag > 0.0f) {
            const int steps = std::max(3, std::min(16, (int)std::ceil(mag / 0.10f)));
            float tx, ty;
            for (int k = 1; k <= steps; ++k) {
                const float f = (float)k / (float)steps;
                landing(sax * f, say * f, tx, ty);
            }
        }
        float ax = sax, ay = say, tx, ty;
        if (landing(ax, ay, tx, ty)) {
            float res = 0.0f;
            const float cap = std::max(0.02f, 0.5f * mag);
            newton_aim(landing, targets[2 * i], targets[2 * i + 1], ax, ay, res, 8, 1e-3f, cap, 5e-3f);
            float x, y;
            if (landing(ax, ay, x, y)) {
                o[0] = ax; o[1] = ay; o[2] = x; o[3] = y; o[4] = res; o[5] = 1.0f;
            }
        }
        if (err) return err;
    }
    return GL_OK;
}

void gl_source_map_screen(const gl_source_map* base, float angle_x, float angle_y, float* out_sx, float* out_sy) {
    if (!base || !out_sx || !out_sy) return;
    if (base->status == GL_SMAP_FALLBACK) {
        *out_sx = 0.5f + 0.5f * std::tan(angle_x) * base->inv_tan_w;
        *out_sy = 0.5f + 0.5f * std::tan(angle_y) * base->inv_tan_h;
        return;
    }
    const float dax = angle_x - base->anchor_ax;
    const float day = angle_y - base->anchor_ay;
    const float d = std::sqrt(dax * dax + day * day);
    float ux = dax, uy = day;
    if (d > 1e-12f) {
        const float t = std::tan(std::min(d, kMaxExtendRad));
        ux = (dax / d) * t;
        uy = (day / d) * t;
    }
    const float x = base->anchor_x + base->J[0] * ux + base->J[1] * uy;
    const float y = base->anchor_y + base->J[2] * ux + base->J[3] * uy;
    *out_sx = (base->eff_half_w > 0.0f) ? 0.5f + 0.5f * x / base->eff_half_w : 0.5f;
    *out_sy = (base->eff_half_h > 0.0f) ? 0.5f + 0.5f * y / base->eff_half_h : 0.5f;
}

int gl_source_map_landing(gl_ctx* ctx, const gl_lens* lens, const gl_calibration* calib, float angle_x, float angle_y,
                          float* out_x_mm, float* out_y_mm, int32_t* out_ok) {
    if (!lens || !calib || !out_x_mm || !out_y_mm || !out_ok) return set_error(GL_ERR_INVALID, "null argument");
    *out_ok = 0;
    if (lens->host.empty()) return GL_OK;
    int st;
    gl_ctx* c = resolve_ctx(ctx, &st);
    if (!c) return st;
    std::lock_guard<std::mutex> g(c->mu);
    Prober P;
    if ((st = P.init(c, lens, kLambdaNm))) return st;
    const float front_R = P.front_R;
    const float fill_x = (front_R > 0.0f) ? calib->entrance_pupil_semi_x / front_R : 1.0f;
    const float fill_y = (front_R > 0.0f) ? calib->entrance_pupil_semi_y / front_R : 1.0f;
    Window win = initial_probe_window(std::min(fill_x, fill_y));
    const float mag = std::sqrt(angle_x * angle_x + angle_y * angle_y);
    DiscArea d;
    if (win.r < 1.0f && mag > 0.0f) {
        for (int k = 1; k <= 3; ++k) {
            const float f = (float)k / 4.0f;
            if ((st = probe_disc(P, angle_x * f, angle_y * f, &win, 0, 0, &d))) return st;
        }
    }
    if ((st = probe_disc(P, angle_x, angle_y, &win, 0, 0, &d))) return st;
    if (!(d.area > 0.0f) || !std::isfinite(d.mean_x) || !std::isfinite(d.mean_y)) return GL_OK;
    *out_x_mm = d.mean_x;
    *out_y_mm = d.mean_y;
    *out_ok = 1;
    return GL_OK;
}

}  // extern "C"
