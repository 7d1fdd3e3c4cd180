// sourcemap.hip — lens calibration (covered field, image circle, first-order
// optics) and the screen <-> field-angle source map, with every probe ray
// traced on the device.
//
// The reference defines the searches (lens_calibration.cpp:19-443,
// spawn_probe.cpp:29-245, source_map.cpp:22-519, newton_aim.h:19-63); the
// answers here are the same to the bit.  The searches are sequential (every
// step depends on the probe before it), so they run on the host as small
// objects -- a beam tracker, a field-axis walker with an Illinois bracket, a
// 2-D Newton on a 2x2 Cramer step, a source solver with explicit strategies --
// and what each step measures is a batch of rays:
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
#include <memory>
#include <mutex>
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
// Host side, part 1: what one probe measures
// ---------------------------------------------------------------------------

// A square sampling window over the unit spawn disc (disc coordinates).
struct DiscWindow {
    float cu = 0.0f, cv = 0.0f;   // centre
    float r = 1.0f;               // half-extent
    bool full() const { return r >= 1.0f; }
};

// The survivors of one window: coverage-weighted area (absolute disc units,
// so windows of different sizes compare), mean landing on the sensor,
// whether a survivor touches the window's rim, and the survivors' bounding
// box padded by one cell (the window a neighbouring field angle should use).
struct Survey {
    float area = 0.0f, mean_x = 0.0f, mean_y = 0.0f;
    bool rim = false;
    DiscWindow fit;
    bool lit() const { return area > 0.0f; }
    bool landed() const { return area > 0.0f && std::isfinite(mean_x) && std::isfinite(mean_y); }
};

// Running survivor integrals of one window.
struct Tally {
    float area = 0.0f, sx = 0.0f, sy = 0.0f;
    float u_lo = 1e30f, u_hi = -1e30f, v_lo = 1e30f, v_hi = -1e30f;
    void add(float u, float v, float wa, float x, float y) {
        area += wa;
        sx += wa * x;
        sy += wa * y;
        u_lo = std::min(u_lo, u); u_hi = std::max(u_hi, u);
        v_lo = std::min(v_lo, v); v_hi = std::max(v_hi, v);
    }
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

    // The largest axial entrance height that still reaches the sensor along
    // each direction, all directions in one launch (0 for a lens whose front
    // element has no aperture).
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

// Unit launch direction for field angles (ah, av): (tan ah, tan av, 1)
// scaled by 1 / |.| (vec3.h:63-70 forms it as x * (1 / length) with
// length = sqrt(x*x + y*y + z*z); the same float operations here).
void launch_dir(float ah, float av, float* d) {
    const float x = std::tan(ah), y = std::tan(av), z = 1.0f;
    const float len = std::sqrt(x * x + y * y + z * z);
    const float k = len > 1e-12f ? 1.0f / len : 0.0f;
    d[0] = len > 1e-12f ? x * k : 0.0f;
    d[1] = len > 1e-12f ? y * k : 0.0f;
    d[2] = len > 1e-12f ? z * k : 0.0f;
}

ProbeDesc window_desc(float ah, float av, const DiscWindow& win, int grid, int ess) {
    ProbeDesc d{};
    d.n = grid;
    d.ess = ess;
    d.step = 2.0f * win.r / (float)grid;
    // the disc follows the beam: shifted back along it to the spawn plane
    d.sdx = -kSpawnOffset * std::tan(ah);
    d.sdy = -kSpawnOffset * std::tan(av);
    launch_dir(ah, av, &d.dx);
    d.u_org = win.cu - win.r;
    d.v_org = win.cv - win.r;
    return d;
}

// One ray from the entrance-plane origin at a field angle: a 1 x 1 window.
ProbeDesc origin_desc(float ah, float av) {
    ProbeDesc d{};
    d.n = 1; d.ess = 1;
    d.step = 1.0f;
    d.u_org = -0.5f; d.v_org = -0.5f;
    launch_dir(ah, av, &d.dx);
    return d;
}

// Survivor integrals of one traced window (the sums of spawn_probe.cpp:108-183).
// Pass 1 marks the cells that straddle a boundary: a 4-neighbour inside the
// grid disagrees about survival, or a survivor lies on the window's rim (the
// outside counts as dark).  Pass 2 sums in raster order: whole cells at their
// centre, straddling cells by their ess x ess sub-samples in place.
Survey tally_window(const ProbeDesc& d, const float2* land, const unsigned char* ok) {
    const int n = d.n, ess = d.ess, nn = n * n;
    Survey s;
    std::vector<unsigned char> split((size_t)nn, 0);
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            const int k = j * n + i;
            const bool lit = ok[k] != 0;
            const bool border = i == 0 || j == 0 || i == n - 1 || j == n - 1;
            if (lit && border) { s.rim = true; split[k] = 1; continue; }
            const bool differs = (i + 1 < n && (ok[k + 1] != 0) != lit) || (i > 0 && (ok[k - 1] != 0) != lit) ||
                                 (j + 1 < n && (ok[k + n] != 0) != lit) || (j > 0 && (ok[k - n] != 0) != lit);
            split[k] = differs ? 1 : 0;
        }
    const float step = d.step;
    const float whole = 1.0f * (step * step);
    const float part = (1.0f / (float)(ess * ess)) * (step * step);
    const float sub = step / (float)ess;
    Tally t;
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            const int k = j * n + i;
            if (!split[k]) {
                if (ok[k])
                    t.add(d.u_org + ((float)i + 0.5f) * step, d.v_org + ((float)j + 0.5f) * step, whole, land[k].x,
                          land[k].y);
                continue;
            }
            const float cu0 = d.u_org + (float)i * step, cv0 = d.v_org + (float)j * step;
            const int first = nn + k * ess * ess;
            for (int sj = 0; sj < ess; ++sj)
                for (int si = 0; si < ess; ++si) {
                    const int q = first + sj * ess + si;
                    if (ok[q])
                        t.add(cu0 + ((float)si + 0.5f) * sub, cv0 + ((float)sj + 0.5f) * sub, part, land[q].x,
                              land[q].y);
                }
        }
    if (t.area > 0.0f) {
        s.area = t.area;
        s.mean_x = t.sx / t.area;
        s.mean_y = t.sy / t.area;
        s.fit.cu = 0.5f * (t.u_lo + t.u_hi);
        s.fit.cv = 0.5f * (t.v_lo + t.v_hi);
        s.fit.r = std::max(std::max(t.u_hi - s.fit.cu, t.v_hi - s.fit.cv), step) + step;
    }
    return s;
}

constexpr int kProbeGrid = 17;
constexpr int kProbeEdgeSS = 2;

// The first window of a search (spawn_probe.cpp:48-54): the whole disc, or
// five times the pupil fill around the axis when the stop admits less than
// 45% of the front element.
DiscWindow seed_window(float pupil_fill) {
    DiscWindow w;
    if (pupil_fill > 0.0f && pupil_fill < 0.45f) w.r = std::min(1.0f, pupil_fill * 5.0f);
    return w;
}

// One window at one field angle: one device launch, host sums.
int survey(Prober& P, float ah, float av, const DiscWindow& win, int grid, int ess, Survey& out) {
    std::vector<ProbeDesc> d{window_desc(ah, av, win, grid > 0 ? grid : kProbeGrid, ess > 0 ? ess : kProbeEdgeSS)};
    int st;
    if ((st = P.run(d))) return st;
    out = tally_window(d[0], P.land.data(), P.ok.data());
    return GL_OK;
}

// Follow the beam at one field angle (spawn_probe.cpp:186-245).  Up to six
// windows: a lit window whose survivors stay off its rim is the answer and
// re-centres the tracker (size kept); a lit one that touches the rim first
// slides onto the survivors (at most three times), otherwise the window
// grows 2.5x until it is the whole disc.  The last lit survey is returned
// when no window is clean.
int track_beam(Prober& P, float ah, float av, DiscWindow& track, int grid, int ess, Survey& out) {
    DiscWindow w = track.full() ? DiscWindow{} : track;
    Survey kept;
    for (int round = 0; round < 6; ++round) {
        Survey s;
        int st;
        if ((st = survey(P, ah, av, w, grid, ess, s))) return st;
        if (s.lit()) kept = s;
        if (s.lit() && !s.rim) {
            track = DiscWindow{s.fit.cu, s.fit.cv, w.r};
            out = s;
            return GL_OK;
        }
        if (w.full()) break;
        const bool slide = s.lit() && round < 3;
        if (slide) {
            w.cu = s.fit.cu;
            w.cv = s.fit.cv;
        } else {
            w.r = std::min(1.0f, w.r * 2.5f);
            if (w.full()) w = DiscWindow{0.0f, 0.0f, w.r};
        }
    }
    track = w;
    out = kept;
    return GL_OK;
}

// ---------------------------------------------------------------------------
// Host side, part 2: 2x2 linear algebra and the finite-difference Newton aim
// ---------------------------------------------------------------------------

// Row-major 2x2: [dx/dax dx/day; dy/dax dy/day] for a landing Jacobian.
struct Mat2 {
    float a00, a01, a10, a11;
    float det() const { return a00 * a11 - a01 * a10; }
    // Cramer's rule for M s = b.
    void solve(float b0, float b1, float dt, float& s0, float& s1) const {
        s0 = (b0 * a11 - a01 * b1) / dt;
        s1 = (a00 * b1 - a10 * b0) / dt;
    }
};

// Aim a landing function at a target (newton_aim.h:19-63 defines the
// search): forward-difference Jacobian with step h, the Newton step from
// Cramer's rule clamped to +-cap per axis, at most `iters` steps, stopping
// under `tol`.  The landing is evaluated once more at the last iterate, and
// the best iterate seen wins; returns whether its residual is under tol.
template <typename Land>
bool newton_2d(Land&& land, float tx, float ty, float& ax, float& ay, float& residual, int iters, float h,
               float cap, float tol) {
    float keep_ax = ax, keep_ay = ay, keep_res = 1e30f;
    auto score = [&](float x, float y) {
        const float ex = x - tx, ey = y - ty;
        const float r = std::sqrt(ex * ex + ey * ey);
        if (r < keep_res) { keep_res = r; keep_ax = ax; keep_ay = ay; }
        return r;
    };
    float x, y;
    if (!land(ax, ay, x, y)) { residual = 1e30f; return false; }
    for (int it = 0; it < iters; ++it) {
        const float ex = x - tx, ey = y - ty;
        if (score(x, y) < tol) break;
        float x1, y1, x2, y2;
        if (!land(ax + h, ay, x1, y1) || !land(ax, ay + h, x2, y2)) break;
        const Mat2 J{(x1 - x) / h, (x2 - x) / h, (y1 - y) / h, (y2 - y) / h};
        const float dt = J.det();
        if (std::fabs(dt) < 1e-12f) break;
        float sx, sy;
        J.solve(-ex, -ey, dt, sx, sy);
        ax += std::max(-cap, std::min(cap, sx));
        ay += std::max(-cap, std::min(cap, sy));
        if (!land(ax, ay, x, y)) break;
    }
    if (land(ax, ay, x, y)) score(x, y);
    ax = keep_ax;
    ay = keep_ay;
    residual = keep_res;
    return keep_res < tol;
}

// ---------------------------------------------------------------------------
// Host side, part 3: calibration (lens_calibration.cpp:19-443 defines it)
// ---------------------------------------------------------------------------

const gl_surface* stop_surface(const gl_lens* L) {
    for (const auto& s : L->host)
        if (s.is_stop) return &s;
    return nullptr;
}

const gl_surface* bladed_stop(const gl_lens* L) {
    const gl_surface* s = stop_surface(L);
    return (s && s->aperture_shape == GL_APERTURE_POLYGON && s->aperture_blades >= 3) ? s : nullptr;
}

// Field-angle ladder of the covered-field walk and the two thresholds it
// brackets, as fractions of the axial survivor area.
const float kFieldLadderDeg[] = {0.0f, 4.0f, 8.0f, 14.0f, 22.0f, 32.0f, 45.0f, 60.0f};
constexpr int kFieldRungs = 8;
constexpr float kDegToRad = (float)(M_PI / 180.0);

struct Threshold {
    float level;      // area fraction
    int max_steps;    // refinement evaluations
    float tol_deg;    // bracket width that ends refinement
};
constexpr Threshold kCoveredField{0.90f, 7, 0.15f};
constexpr Threshold kImageCircle{0.05f, 2, 1.0f};

struct FieldCrossing {
    float angle_rad = 0.0f;
    float half_mm = 0.0f;
    bool crossed = false;
};

// A ladder step whose far end fell below a threshold: the area fractions at
// its ends and where the beam's window sat at its near end.
struct Bracket {
    int rung = -1;
    float f_near = 1.0f, f_far = 0.0f;
    DiscWindow seed;
};

// Illinois-modified regula falsi on [lo, hi] with g(lo) >= 0 > g(hi): the
// end that keeps its place has its weight halved, while the measured values
// at the ends are kept apart for the final interpolation.
struct Illinois {
    float lo, hi;
    float w_lo, w_hi;     // weighted values that steer the next point
    float m_lo, m_hi;     // measured values at the current ends
    float next() const {
        const float denom = w_lo - w_hi;
        const float x = (std::abs(denom) < 1e-12f) ? 0.5f * (lo + hi) : lo + (hi - lo) * w_lo / denom;
        const float guard = 0.01f * (hi - lo);
        return std::min(std::max(x, lo + guard), hi - guard);
    }
    void take(float x, float g) {
        if (g >= 0.0f) { lo = x; w_lo = g; m_lo = g; w_hi *= 0.5f; }
        else { hi = x; w_hi = g; m_hi = g; w_lo *= 0.5f; }
    }
    float root() const {
        const float span = m_lo - m_hi;
        const float x = (std::abs(span) < 1e-12f) ? 0.5f * (lo + hi) : lo + (hi - lo) * m_lo / span;
        return std::min(std::max(x, lo), hi);
    }
};

// One axis of field angle: survivor area against the axial area measured
// through a window of the same size (re-measured whenever the tracker grows).
class FieldAxis {
public:
    FieldAxis(Prober& P, bool horizontal, float pupil_fill) : P_(P), horiz_(horizontal), track_(seed_window(pupil_fill)) {}

    int resolve(FieldCrossing& field, FieldCrossing& circle) {
        field = circle = FieldCrossing{};
        Survey s;
        int st;
        if ((st = at(0.0f, s))) return st;
        if (axial_.area <= 0.0f) return GL_OK;
        // An uncrossed bracket refines from the window the axial walk left.
        Bracket b_field, b_circle;
        b_field.seed = b_circle.seed = track_;
        float f_prev = 1.0f;
        for (int k = 1; k < kFieldRungs; ++k) {
            const DiscWindow before = track_;
            if ((st = at(kFieldLadderDeg[k] * kDegToRad, s))) return st;
            const float f = s.area / axial_.area;
            if (b_field.rung < 0 && f < kCoveredField.level) b_field = Bracket{k - 1, f_prev, f, before};
            if (b_circle.rung < 0 && f < kImageCircle.level) {
                b_circle = Bracket{k - 1, f_prev, f, before};
                break;                       // the image circle is the outer crossing
            }
            f_prev = f;
        }
        if ((st = refine(b_field, kCoveredField, field))) return st;
        if ((st = refine(b_circle, kImageCircle, circle))) return st;
        // The image circle never sits inside the covered field.
        if (circle.angle_rad < field.angle_rad) circle = field;
        if (circle.half_mm < field.half_mm) circle.half_mm = field.half_mm;
        return GL_OK;
    }

private:
    Prober& P_;
    bool horiz_;
    DiscWindow track_;
    Survey axial_;
    float axial_r_ = -1.0f;

    int at(float angle, Survey& s) {
        int st;
        if ((st = track_beam(P_, horiz_ ? angle : 0.0f, horiz_ ? 0.0f : angle, track_, 0, 0, s))) return st;
        if (track_.r != axial_r_) {
            if ((st = survey(P_, 0.0f, 0.0f, DiscWindow{0.0f, 0.0f, track_.r}, 0, 0, axial_))) return st;
            axial_r_ = track_.r;
        }
        return GL_OK;
    }
    // Every evaluation inside a bracket starts from where the beam was at its
    // near end (the window size is left as the walk grew it).
    int at_from(const DiscWindow& seed, float angle, Survey& s) {
        track_.cu = seed.cu;
        track_.cv = seed.cv;
        return at(angle, s);
    }
    float along(const Survey& s) const { return std::abs(horiz_ ? s.mean_x : s.mean_y); }

    int refine(const Bracket& b, const Threshold& thr, FieldCrossing& out) {
        Survey s;
        int st;
        if (b.rung < 0) {
            // never crossed on the ladder: report its ceiling, uncrossed
            const float top = kFieldLadderDeg[kFieldRungs - 1] * kDegToRad;
            if ((st = at_from(b.seed, top, s))) return st;
            out = FieldCrossing{top, along(s), false};
            return GL_OK;
        }
        const float g0 = b.f_near - thr.level, g1 = b.f_far - thr.level;
        Illinois il{kFieldLadderDeg[b.rung] * kDegToRad, kFieldLadderDeg[b.rung + 1] * kDegToRad, g0, g1, g0, g1};
        const float tol = thr.tol_deg * kDegToRad;
        for (int it = 0; it < thr.max_steps && (il.hi - il.lo) > tol; ++it) {
            const float x = il.next();
            if ((st = at_from(b.seed, x, s))) return st;
            il.take(x, s.area / axial_.area - thr.level);
        }
        float root = il.root();
        if ((st = at_from(b.seed, root, s))) return st;
        if (s.area <= 0.0f) {
            // an interpolated root past the last light: the near end has light
            root = il.lo;
            if ((st = at_from(b.seed, root, s))) return st;
        }
        out = FieldCrossing{root, along(s), true};
        return GL_OK;
    }
};

// axes_are_symmetric (lens_calibration.cpp:312-340): one axis serves both when
// every surface is rotationally symmetric up to a quarter turn.
bool quarter_turn_symmetric(const gl_lens* L) {
    static const float I[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    for (const auto& s : L->host) {
        const bool round_form = s.form != GL_FORM_CYLINDRICAL && s.aperture_aspect == 1.0f &&
                                s.decenter_x == 0.0f && s.decenter_y == 0.0f &&
                                std::equal(s.rot, s.rot + 9, I);
        if (!round_form) return false;
        const bool shape_ok = s.aperture_shape == GL_APERTURE_CIRCLE ||
                              (s.aperture_shape == GL_APERTURE_POLYGON && s.aperture_blades % 4 == 0 &&
                               s.aperture_rotation_rad == 0.0f);
        if (!shape_ok) return false;
    }
    return true;
}

// Paraxial focal lengths from one 1-degree ray per axis (both in one launch).
int first_order(Prober& P, gl_calibration& cal) {
    const float theta = (float)(1.0 * M_PI / 180.0);
    std::vector<ProbeDesc> d{origin_desc(theta, 0.0f), origin_desc(0.0f, theta)};
    int st;
    if ((st = P.run(d))) return st;
    float* efl[2] = {&cal.focal_length_x, &cal.focal_length_y};
    for (int a = 0; a < 2; ++a) {
        const int k = d[a].out;
        const float h = a == 0 ? P.land[k].x : P.land[k].y;
        *efl[a] = (P.ok[k] && std::isfinite(h)) ? std::abs(h) / std::tan(theta) : 0.0f;
    }
    return GL_OK;
}

// Entrance-pupil semi-axes: the support of the marginal heights over a
// bladed stop's azimuths (8 per blade plus the two axes), or the two
// principal meridians of a round stop.
int entrance_pupil(Prober& P, gl_calibration& cal) {
    std::vector<float2> dirs;
    const gl_surface* stop = bladed_stop(P.L);
    if (stop) {
        const int n = stop->aperture_blades * 8;
        std::vector<float> th;
        for (int i = 0; i < n; ++i) th.push_back(stop->aperture_rotation_rad + (float)(2.0 * M_PI) * i / (float)n);
        th.push_back(0.0f);
        th.push_back((float)(M_PI / 2.0));
        for (float t : th) dirs.push_back(make_float2(std::cos(t), std::sin(t)));
    } else {
        dirs = {make_float2(1.0f, 0.0f), make_float2(0.0f, 1.0f)};
    }
    std::vector<float> h;
    int st;
    if ((st = P.marginal(dirs, h))) return st;
    if (!stop) {
        cal.entrance_pupil_semi_x = h[0];
        cal.entrance_pupil_semi_y = h[1];
        return GL_OK;
    }
    float sx = 0.0f, sy = 0.0f;
    for (size_t i = 0; i < dirs.size(); ++i) {
        sx = std::max(sx, h[i] * std::fabs(dirs[i].x));
        sy = std::max(sy, h[i] * std::fabs(dirs[i].y));
    }
    cal.entrance_pupil_semi_x = sx;
    cal.entrance_pupil_semi_y = sy;
    gl_blade_profile prof;
    gl_blade_profile_build(stop->aperture_blades, stop->aperture_rotation_rad, stop->aperture_curvature,
                           stop->aperture_twist, stop->aperture_notch_rad, stop->aperture_notch_angle_rad, &prof);
    cal.pupil_area_frac = prof.area_frac;
    return GL_OK;
}

// A field axis that resolved to nothing falls back to a 0.05-degree field on
// the lens file's focal length (1 mm without one).
void floor_field(float* angle, float* half, float file_focal_length) {
    const float min_field = (float)(0.05 * M_PI / 180.0);
    if (*angle >= min_field && *half > 0.0f) return;
    std::fprintf(stderr, "[ghostlight_amd] warning: a covered-field axis resolved to nothing; using the lens file's "
                         "focal length instead.\n");
    *angle = min_field;
    *half = (file_focal_length > 0.0f) ? file_focal_length * std::tan(min_field) : std::tan(min_field);
}

int calibrate(Prober& P, float file_focal_length, gl_calibration* out) {
    gl_calibration cal{};
    cal.pupil_area_frac = 1.0f;
    if (P.L->host.empty()) { *out = cal; return GL_OK; }
    int st;
    if ((st = first_order(P, cal)) || (st = entrance_pupil(P, cal))) return st;
    const float fr = P.front_R;
    const float fill = std::min(fr > 0.0f ? cal.entrance_pupil_semi_x / fr : 1.0f,
                                fr > 0.0f ? cal.entrance_pupil_semi_y / fr : 1.0f);
    FieldCrossing fh, fv, ch, cv;
    {
        FieldAxis h(P, true, fill);
        if ((st = h.resolve(fh, ch))) return st;
    }
    if (quarter_turn_symmetric(P.L)) {
        fv = fh;
        cv = ch;
    } else {
        FieldAxis v(P, false, fill);
        if ((st = v.resolve(fv, cv))) return st;
    }
    cal.max_half_angle_h = fh.angle_rad;
    cal.max_half_angle_v = fv.angle_rad;
    cal.sensor_half_w = fh.half_mm;
    cal.sensor_half_h = fv.half_mm;
    cal.image_circle_semi_w = ch.half_mm;
    cal.image_circle_semi_h = cv.half_mm;
    if (!fh.crossed || !fv.crossed)
        std::fprintf(stderr, "[ghostlight_amd] warning: the survivor area never fell below %.0f%% of axial up to "
                             "%.0f deg; the covered field is capped there.\n",
                     (double)(kCoveredField.level * 100.0f), (double)kFieldLadderDeg[kFieldRungs - 1]);
    floor_field(&cal.max_half_angle_h, &cal.sensor_half_w, file_focal_length);
    floor_field(&cal.max_half_angle_v, &cal.sensor_half_h, file_focal_length);
    if (cal.image_circle_semi_w <= 0.0f) cal.image_circle_semi_w = cal.sensor_half_w;
    if (cal.image_circle_semi_h <= 0.0f) cal.image_circle_semi_h = cal.sensor_half_h;
    if (cal.entrance_pupil_semi_x > 0.0f) cal.f_number_x = cal.focal_length_x / (2.0f * cal.entrance_pupil_semi_x);
    if (cal.entrance_pupil_semi_y > 0.0f) cal.f_number_y = cal.focal_length_y / (2.0f * cal.entrance_pupil_semi_y);
    *out = cal;
    return GL_OK;
}

// ---------------------------------------------------------------------------
// Host side, part 4: the traced source map (source_map.cpp:22-519 defines it)
// ---------------------------------------------------------------------------

constexpr float kMapLambdaNm = 587.56f;   // the calibration's d line
constexpr int kMapGrid = 17;              // the calibration's probe grid ...
constexpr int kMapEdgeSS = 2;             // ... and boundary sub-sampling
constexpr int kAimIters = 8;
constexpr float kAimStep = 5e-3f;         // finite-difference step, rad
constexpr float kAimTolMm = 1e-3f;
constexpr float kAcceptOfFrame = 0.01f;   // acceptance, fraction of the smaller half-frame
constexpr int kAnchorHalvings = 6;
constexpr float kAnchorTolMm = 0.02f;
constexpr float kExtendMaxRad = 1.55f;    // tan() clamp of the continuation
constexpr float kWalkStepRad = 0.10f;     // warm-up walk step

// The request a solve answers, as the bytes that decide it: prescription,
// coating arena, calibration, position and frame.
std::vector<uint8_t> solve_key(const gl_lens* L, const gl_calibration& calib, float nx, float ny, float hw, float hh) {
    std::vector<uint8_t> k;
    auto put = [&](const void* p, size_t n) {
        const uint8_t* b = static_cast<const uint8_t*>(p);
        k.insert(k.end(), b, b + n);
    };
    put(L->host.data(), L->host.size() * sizeof(gl_surface));
    put(L->arena.data(), L->arena.size() * sizeof(float));
    put(&calib, sizeof calib);
    const float req[4] = {nx, ny, hw, hh};
    put(req, sizeof req);
    return k;
}

// Last answer, shared by the chunks of an extended-source render.
struct LastSolve {
    std::mutex mu;
    std::vector<uint8_t> key;
    gl_source_map value{};
    bool valid = false;
    bool warned = false;
};
LastSolve& last_solve() {
    static LastSolve s;
    return s;
}

// The survivor-mean landing as a function of field angle, through a window
// that follows the beam from call to call.  The last eight evaluations are
// remembered with the window they went through and the window they left, so
// a repeated evaluation replays both (the same answers, no trace).
class BeamLanding {
public:
    BeamLanding(Prober& P, const DiscWindow& w) : P_(P), win(w) {}
    DiscWindow win;
    int traced = 0;
    int err = GL_OK;

    bool operator()(float ax, float ay, float& x, float& y) {
        if (err) return false;
        for (int i = 0; i < used_; ++i) {
            const Memo& m = memo_[i];
            if (m.ax == ax && m.ay == ay && m.in.cu == win.cu && m.in.cv == win.cv && m.in.r == win.r) {
                if (!m.ok) return false;
                win = m.out; x = m.x; y = m.y;
                return true;
            }
        }
        ++traced;
        Memo m{ax, ay, win, win, 0.0f, 0.0f, false};
        Survey s;
        if ((err = track_beam(P_, ax, ay, m.out, kMapGrid, kMapEdgeSS, s))) return false;
        m.ok = s.landed();
        m.x = s.mean_x;
        m.y = s.mean_y;
        memo_[slot_] = m;
        slot_ = (slot_ + 1) % kMemo;
        used_ = std::min(used_ + 1, kMemo);
        if (!m.ok) return false;
        win = m.out; x = m.x; y = m.y;
        return true;
    }

private:
    struct Memo { float ax, ay; DiscWindow in, out; float x, y; bool ok; };
    static constexpr int kMemo = 8;
    Prober& P_;
    Memo memo_[kMemo]{};
    int used_ = 0, slot_ = 0;
};

class SourceSolver {
public:
    SourceSolver(Prober& P, const gl_calibration& cal, float nx, float ny, float hw, float hh)
        : P_(P), cal_(cal), nx_(nx), ny_(ny), hw_(hw), hh_(hh) {}

    int run(gl_source_map* out) {
        const gl_lens* L = P_.L;
        const std::vector<uint8_t> key = solve_key(L, cal_, nx_, ny_, hw_, hh_);
        LastSolve& ls = last_solve();
        {
            std::lock_guard<std::mutex> lk(ls.mu);
            if (ls.valid && ls.key == key) { *out = ls.value; return GL_OK; }
        }
        start_fallback();
        if (L->host.empty()) return fallback(key, "the lens has no surfaces", out);

        tx_ = nx_ * hw_;
        ty_ = ny_ * hh_;
        const float fr = P_.front_R;
        win0_ = seed_window(std::min(fr > 0.0f ? cal_.entrance_pupil_semi_x / fr : 1.0f,
                                     fr > 0.0f ? cal_.entrance_pupil_semi_y / fr : 1.0f));
        beam_.reset(new BeamLanding(P_, win0_));
        accept_ = std::max(kAimTolMm, kAcceptOfFrame * std::min(std::fabs(hw_), std::fabs(hh_)));

        // Newton from the closed-form seed: first with room for a large
        // correction, then with short steps for a quantisation-noisy Jacobian.
        for (const float cap : {std::max(0.02f, 0.5f * seed_mag_), std::max(0.005f, 0.08f * seed_mag_)}) {
            if (aim_from(seed_ax_, seed_ay_, cap)) return publish(key, out);
            if (beam_->err) return beam_->err;
        }
        // Beyond the last landing the lens delivers: the furthest target along
        // the ray that still solves anchors a continuation.
        const Anchor A = find_anchor();
        if (beam_->err) return beam_->err;
        if (!A.found && !(seed_mag_ > 0.0f)) return fallback(key, "no field angle lands anywhere along the ray", out);
        if (A.found) {
            if (aim_from(A.ax, A.ay, std::max(0.02f, 0.5f * seed_mag_))) return publish(key, out);
            if (beam_->err) return beam_->err;
            if (A.radius >= std::sqrt(tx_ * tx_ + ty_ * ty_) && best_res_ < 1e29f) {
                walk_out(best_ax_, best_ay_);
                if (record_solved(best_ax_, best_ay_, best_res_)) return publish(key, out);
            }
            if (beam_->err) return beam_->err;
            beam_->win = A.win;
        }
        return continue_from(A, key, out);
    }

private:
    struct Anchor {
        bool found = false;
        float ax = 0.0f, ay = 0.0f, radius = 0.0f;
        DiscWindow win;
    };

    Prober& P_;
    const gl_calibration& cal_;
    float nx_, ny_, hw_, hh_;
    float tx_ = 0.0f, ty_ = 0.0f, accept_ = 0.0f;
    float seed_ax_ = 0.0f, seed_ay_ = 0.0f, seed_mag_ = 0.0f;
    float best_ax_ = 0.0f, best_ay_ = 0.0f, best_res_ = 1e30f;
    DiscWindow win0_;
    std::unique_ptr<BeamLanding> beam_;
    gl_source_map sm_{};

    // The constant-focal-length answer, which is also the solve's seed.
    void start_fallback() {
        sm_ = gl_source_map{};
        sm_.eff_half_w = hw_;
        sm_.eff_half_h = hh_;
        const float kw = (cal_.sensor_half_w > 0.0f) ? hw_ / cal_.sensor_half_w : 1.0f;
        const float kh = (cal_.sensor_half_h > 0.0f) ? hh_ / cal_.sensor_half_h : 1.0f;
        const float tw = std::tan(cal_.max_half_angle_h), th = std::tan(cal_.max_half_angle_v);
        seed_ax_ = std::atan(nx_ * kw * tw);
        seed_ay_ = std::atan(ny_ * kh * th);
        seed_mag_ = std::sqrt(seed_ax_ * seed_ax_ + seed_ay_ * seed_ay_);
        best_ax_ = seed_ax_;
        best_ay_ = seed_ay_;
        sm_.inv_tan_w = 1.0f / (kw * tw);
        sm_.inv_tan_h = 1.0f / (kh * th);
        sm_.angle_x = seed_ax_;
        sm_.angle_y = seed_ay_;
        sm_.status = GL_SMAP_FALLBACK;
    }

    void remember(const std::vector<uint8_t>& key) {
        LastSolve& ls = last_solve();
        std::lock_guard<std::mutex> lk(ls.mu);
        ls.key = key;
        ls.value = sm_;
        ls.valid = true;
    }

    int fallback(const std::vector<uint8_t>& key, const char* why, gl_source_map* out) {
        {
            LastSolve& ls = last_solve();
            std::lock_guard<std::mutex> lk(ls.mu);
            if (!ls.warned) {
                ls.warned = true;
                std::fprintf(stderr, "[ghostlight_amd] warning: the traced source map has no solution here (%s); "
                                     "using the constant-focal-length map (reported once).\n", why);
            }
        }
        sm_.probe_evals = beam_ ? beam_->traced : 0;
        remember(key);
        *out = sm_;
        return GL_OK;
    }

    int publish(const std::vector<uint8_t>& key, gl_source_map* out) {
        if (beam_->err) return beam_->err;
        remember(key);
        *out = sm_;
        return GL_OK;
    }

    // Restart from the seed window and carry it out to (ax, ay) in steps of
    // at most 0.1 rad (3 to 16 of them), so a deep-stop window finds the beam.
    void walk_out(float ax, float ay) {
        BeamLanding& b = *beam_;
        b.win = win0_;
        const float mag = std::sqrt(ax * ax + ay * ay);
        if (!(b.win.r < 1.0f) || !(mag > 0.0f)) return;
        const int steps = std::max(3, std::min(16, (int)std::ceil(mag / kWalkStepRad)));
        for (int k = 1; k <= steps; ++k) {
            const float f = (float)k / (float)steps;
            float x, y;
            b(ax * f, ay * f, x, y);
        }
    }

    // Landing and its angle derivatives at (ax, ay): one-sided differences,
    // forward where the forward probe lands, else backward (the continuation
    // anchor is the last angle that lands).
    bool slopes(float ax, float ay, float land[2], Mat2& J) {
        BeamLanding& b = *beam_;
        float cx, cy;
        if (!b(ax, ay, cx, cy)) return false;
        auto column = [&](float dax, float day, float& dx, float& dy) -> bool {
            float px, py;
            if (b(ax + dax, ay + day, px, py)) { dx = (px - cx) / kAimStep; dy = (py - cy) / kAimStep; return true; }
            if (b(ax - dax, ay - day, px, py)) { dx = (cx - px) / kAimStep; dy = (cy - py) / kAimStep; return true; }
            return false;
        };
        if (!column(kAimStep, 0.0f, J.a00, J.a10)) return false;
        if (!column(0.0f, kAimStep, J.a01, J.a11)) return false;
        land[0] = cx;
        land[1] = cy;
        return true;
    }

    void set_map(float ax, float ay, const float land[2], const Mat2& J) {
        sm_.anchor_ax = ax; sm_.anchor_ay = ay;
        sm_.anchor_x = land[0]; sm_.anchor_y = land[1];
        sm_.J[0] = J.a00; sm_.J[1] = J.a01; sm_.J[2] = J.a10; sm_.J[3] = J.a11;
        sm_.probe_evals = beam_->traced;
    }

    bool record_solved(float ax, float ay, float res) {
        float land[2];
        Mat2 J{};
        if (!slopes(ax, ay, land, J)) return false;
        sm_.angle_x = ax;
        sm_.angle_y = ay;
        sm_.status = GL_SMAP_SOLVED;
        sm_.residual_mm = res;
        set_map(ax, ay, land, J);
        return true;
    }

    bool aim(float tx, float ty, float& ax, float& ay, float& res, float cap) {
        return newton_2d(*beam_, tx, ty, ax, ay, res, kAimIters, kAimStep, cap, kAimTolMm);
    }

    // One solve attempt at the real target from (ax0, ay0); the best iterate
    // of all attempts is kept for the in-range fallback.
    bool aim_from(float ax0, float ay0, float cap) {
        walk_out(ax0, ay0);
        float ax = ax0, ay = ay0, res = 1e30f;
        aim(tx_, ty_, ax, ay, res, cap);
        if (res < best_res_) { best_res_ = res; best_ax_ = ax; best_ay_ = ay; }
        return res < accept_ && record_solved(ax, ay, res);
    }

    // Halve the radius along the direction of the target, inside 1.05x the
    // image circle, for the furthest target that still solves.  The seed, the
    // Newton cap and the window depend only on the lens and the direction, so
    // requests along one ray share their anchor.
    Anchor find_anchor() {
        Anchor A;
        const float tmag = std::sqrt(tx_ * tx_ + ty_ * ty_);
        if (!(tmag > 0.0f)) return A;
        const float ux = tx_ / tmag, uy = ty_ / tmag;
        const float icw = cal_.image_circle_semi_w, ich = cal_.image_circle_semi_h;
        const float rc = std::sqrt(ux * ux * icw * icw + uy * uy * ich * ich);
        const float aax = std::atan(0.98f * rc * ux * std::tan(cal_.max_half_angle_h) /
                                    (cal_.sensor_half_w > 0.0f ? cal_.sensor_half_w : 1.0f));
        const float aay = std::atan(0.98f * rc * uy * std::tan(cal_.max_half_angle_v) /
                                    (cal_.sensor_half_h > 0.0f ? cal_.sensor_half_h : 1.0f));
        const float cap = std::max(0.02f, 0.5f * std::sqrt(aax * aax + aay * aay));
        walk_out(aax, aay);
        float inside = 0.0f, outside = 1.05f * rc;
        auto reaches = [&](float radius) -> bool {
            float ax = A.ax, ay = A.ay, res = 1e30f;
            aim(radius * ux, radius * uy, ax, ay, res, cap);
            if (!(res < accept_)) return false;
            A = Anchor{true, ax, ay, radius, beam_->win};
            return true;
        };
        const float g = 0.98f * rc;
        if (g > 0.0f && reaches(g)) inside = g;
        else outside = g;
        for (int i = 0; i < kAnchorHalvings && (outside - inside) > kAnchorTolMm; ++i) {
            const float mid = 0.5f * (inside + outside);
            if (reaches(mid)) inside = mid;
            else outside = mid;
        }
        return A;
    }

    // theta(T) = theta_a + u / |u| atan(|u|), u = J^-1 (T - T_a): value and
    // slope meet the solved map at the anchor, monotone, bounded by 90 deg.
    int continue_from(const Anchor& A, const std::vector<uint8_t>& key, gl_source_map* out) {
        float land[2];
        Mat2 J{};
        if (!slopes(A.ax, A.ay, land, J)) {
            if (beam_->err) return beam_->err;
            return fallback(key, "no anchor to continue from", out);
        }
        const float dt = J.det();
        if (!(std::fabs(dt) > 1e-12f)) return fallback(key, "the landing is singular in angle at the anchor", out);
        const float dx = tx_ - land[0], dy = ty_ - land[1];
        float ux, uy;
        J.solve(dx, dy, dt, ux, uy);
        const float um = std::sqrt(ux * ux + uy * uy);
        sm_.angle_x = A.ax;
        sm_.angle_y = A.ay;
        if (um > 1e-12f) {
            const float turn = std::atan(um);
            sm_.angle_x = A.ax + (ux / um) * turn;
            sm_.angle_y = A.ay + (uy / um) * turn;
        }
        sm_.status = GL_SMAP_CONTINUED;
        sm_.residual_mm = std::sqrt(dx * dx + dy * dy);
        set_map(A.ax, A.ay, land, J);
        remember(key);
        *out = sm_;
        return GL_OK;
    }
};

// ---- PSF pre-pass deep-pupil retries (PSFRenderer.cpp:109-122, 203-247, 383-410)

constexpr float kChiefLambdaNm = 587.56f;

// The axial transmitting window: one full-disc survey on axis; its survivor
// box seeds the tracked window of the deep-pupil retries.
int measure_axial(Prober& P, DiscWindow* win, bool* usable) {
    Survey s;
    int st;
    if ((st = survey(P, 0.0f, 0.0f, DiscWindow{}, 0, 0, s))) return st;
    *usable = s.lit();
    if (*usable) *win = s.fit;
    return GL_OK;
}

// Steps of the warm-up walk out to a field angle of magnitude `mag`.
int walk_steps(float mag) { return std::max(3, std::min(16, (int)std::ceil(mag / kWalkStepRad))); }

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
    return calibrate(P, file_focal_length, out);
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
    if (!lens->host.empty() && (st = P.init(c, lens, kMapLambdaNm))) return st;
    P.L = lens;
    SourceSolver S(P, *calib, ndc_x, ndc_y, eff_half_w, eff_half_h);
    return S.run(out);
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
    DiscWindow axial_win;
    bool usable = false;
    if ((st = measure_axial(P, &axial_win, &usable))) return st;
    if (!usable) return GL_OK;
    for (int i = 0; i < n; ++i) {
        float* o = out6 + 6 * (size_t)i;
        const float sax = angles[2 * i], say = angles[2 * i + 1];
        const float mag = std::sqrt(sax * sax + say * say);
        DiscWindow track = axial_win;
        // The survivor mean through the tracked window (the window moves on
        // only when the probe lands).
        int err = GL_OK;
        auto landing = [&](float ax, float ay, float& x, float& y) -> bool {
            if (err) return false;
            DiscWindow w = track;
            Survey s;
            if ((err = track_beam(P, ax, ay, w, 0, 0, s))) return false;
            if (!s.landed()) return false;
            track = w;
            x = s.mean_x; y = s.mean_y;
            return true;
        };
        if (!target_mode) {
            // Walk out to the angle, then read the survivor mean there.
            Survey s;
            if (track.r < 1.0f && mag > 0.0f) {
                const int steps = walk_steps(mag);
                for (int k = 1; k <= steps; ++k) {
                    const float f = (float)k / (float)steps;
                    if ((st = track_beam(P, sax * f, say * f, track, 0, 0, s))) return st;
                }
            }
            if ((st = track_beam(P, sax, say, track, 0, 0, s))) return st;
            if (s.landed()) { o[2] = s.mean_x; o[3] = s.mean_y; o[4] = 0.0f; o[5] = 1.0f; }
            continue;
        }
        // Aim the tracked window's survivor mean at the target.
        if (track.r < 1.0f && mag > 0.0f) {
            const int steps = walk_steps(mag);
            float tx, ty;
            for (int k = 1; k <= steps; ++k) {
                const float f = (float)k / (float)steps;
                landing(sax * f, say * f, tx, ty);
            }
        }
        float ax = sax, ay = say, tx, ty;
        if (landing(ax, ay, tx, ty)) {
            float res = 0.0f;
            newton_2d(landing, targets[2 * i], targets[2 * i + 1], ax, ay, res, 8, 1e-3f, std::max(0.02f, 0.5f * mag),
                      5e-3f);
            float x, y;
            if (landing(ax, ay, x, y)) { o[0] = ax; o[1] = ay; o[2] = x; o[3] = y; o[4] = res; o[5] = 1.0f; }
        }
        if (err) return err;
    }
    return GL_OK;
}

void gl_source_map_screen(const gl_source_map* base, float angle_x, float angle_y, float* out_sx, float* out_sy) {
    if (!base || !out_sx || !out_sy) return;
    if (base->status == GL_SMAP_FALLBACK) {
        // constant focal length
        *out_sx = 0.5f + 0.5f * std::tan(angle_x) * base->inv_tan_w;
        *out_sy = 0.5f + 0.5f * std::tan(angle_y) * base->inv_tan_h;
        return;
    }
    // Linear in the tangent-stretched offset from the anchor, the stretch
    // saturating at 1.55 rad.
    float ox = angle_x - base->anchor_ax, oy = angle_y - base->anchor_ay;
    const float len = std::sqrt(ox * ox + oy * oy);
    if (len > 1e-12f) {
        const float stretch = std::tan(std::min(len, kExtendMaxRad));
        ox = (ox / len) * stretch;
        oy = (oy / len) * stretch;
    }
    const float x = base->anchor_x + base->J[0] * ox + base->J[1] * oy;
    const float y = base->anchor_y + base->J[2] * ox + base->J[3] * oy;
    *out_sx = base->eff_half_w > 0.0f ? 0.5f + 0.5f * x / base->eff_half_w : 0.5f;
    *out_sy = base->eff_half_h > 0.0f ? 0.5f + 0.5f * y / base->eff_half_h : 0.5f;
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
    if ((st = P.init(c, lens, kMapLambdaNm))) return st;
    const float fr = P.front_R;
    DiscWindow win = seed_window(std::min(fr > 0.0f ? calib->entrance_pupil_semi_x / fr : 1.0f,
                                          fr > 0.0f ? calib->entrance_pupil_semi_y / fr : 1.0f));
    // The calibration's grid (not the solve's), walked out in quarters.
    const float mag = std::sqrt(angle_x * angle_x + angle_y * angle_y);
    Survey s;
    if (win.r < 1.0f && mag > 0.0f)
        for (int k = 1; k <= 3; ++k) {
            const float f = (float)k / 4.0f;
            if ((st = track_beam(P, angle_x * f, angle_y * f, win, 0, 0, s))) return st;
        }
    if ((st = track_beam(P, angle_x, angle_y, win, 0, 0, s))) return st;
    if (!s.landed()) return GL_OK;
    *out_x_mm = s.mean_x;
    *out_y_mm = s.mean_y;
    *out_ok = 1;
    return GL_OK;
}

}  // extern "C"
