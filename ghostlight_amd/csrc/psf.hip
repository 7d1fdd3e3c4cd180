// psf.hip — geometric PSF grid: primary traces of every (source, pupil
// sample, wavelength), splatted relative to the source's chief landing into
// the source's own tile of the composite (psf_kernel / launch_psf_render,
// ghostlight/src/psf_render.cu:34-465).
//
// Every splat of a source lands inside that source's tile, so a workgroup
// (one source, 256 pupil samples) accumulates into an LDS copy of the tile
// with LDS atomics and writes it out once: one global atomic per touched
// tile pixel per workgroup instead of four (bilinear) or (2r+1)^2 (Gaussian)
// per ray and wavelength.  Tiles too large for LDS take the direct path.
#include "gl_runtime.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

namespace gl {

static constexpr int kPsfBlock = 1024;
static constexpr size_t kPsfMaxLdsTile = 128 * 1024;

struct PsfArgs {
    const DevSurf* __restrict__ S;
    const float* __restrict__ poly;
    const float* __restrict__ arena;
    int n;
    const gl_psf_source* __restrict__ src;
    int n_src;
    const gl_sample* __restrict__ grid;
    int n_grid;
    float front_R, start_z, mm_per_px;
    int tile_w, tile_h, comp_w, comp_h;
    float* __restrict__ out_r;
    float* __restrict__ out_g;
    float* __restrict__ out_b;
    float ray_weight;
    int mono;
    const float4* __restrict__ spec;     // lambda, cmf_r, cmf_g, cmf_b (x d-lambda)
    int n_spec;
    const float* __restrict__ iat;       // n_spec x n
    float sigma_px;
    int radius_px;
};

// Adds v to tile pixel (x, y) (composite coordinates) of plane c.
template <bool LDS>
__device__ __forceinline__ void psf_add(const PsfArgs& A, float* tile, int x0t, int y0t, int c, int x, int y, float v) {
    if (LDS) {
        atomicAdd(&tile[(size_t)c * A.tile_w * A.tile_h + (size_t)(y - y0t) * A.tile_w + (x - x0t)], v);
    } else {
        float* o = c == 0 ? A.out_r : (c == 1 ? A.out_g : A.out_b);
        atomicAdd(o + (size_t)y * A.comp_w + x, v);
    }
}

// LF: the lens's lean level (lean_level).
template <bool LDS, int LF>
__global__ __launch_bounds__(kPsfBlock) void psf_kernel(PsfArgs A) {
    extern __shared__ float tile[];
    const int si_src = blockIdx.x;
    const gl_psf_source s = A.src[si_src];
    const int planes = A.mono ? 1 : 3;
    const int TT = A.tile_w * A.tile_h;
    if (LDS) {
        for (int i = threadIdx.x; i < planes * TT; i += blockDim.x) tile[i] = 0.0f;
        __syncthreads();
    }
    const int gi = (int)blockIdx.y * kPsfBlock + (int)threadIdx.x;
    const int tx_min = s.tile_x0, ty_min = s.tile_y0;
    const int tx_max = s.tile_x0 + A.tile_w, ty_max = s.tile_y0 + A.tile_h;
    if (gi < A.n_grid) {
        const gl_sample g = A.grid[gi];
        const float bx = tanf(s.angle_x), by = tanf(s.angle_y);
        const F3 dir = normalized(f3(bx, by, 1.0f));
        const F3 o = f3(g.u * A.front_R + (-20.0f * bx), g.v * A.front_R + (-20.0f * by), A.start_z);
        const float src_avg = (s.r + s.g + s.b) * (1.0f / 3.0f);
        for (int si = 0; si < A.n_spec; ++si) {
            const float4 sp = A.spec[si];
            TraceRes res = trace_primary<true, LF>(as_const(A.S), A.n, as_const(A.poly), A.arena, o, dir, sp.x,
                                               as_const(A.iat) + (size_t)si * A.n, false, nullptr);
            if (res.status != GL_TRACE_OK) continue;
            if (!isfinite(res.px) || !isfinite(res.py)) continue;
            const float tx = (res.px - s.chief_x_mm) / A.mm_per_px + (float)A.tile_w * 0.5f;
            const float ty = (res.py - s.chief_y_mm) / A.mm_per_px + (float)A.tile_h * 0.5f;
            const float margin = (float)(A.radius_px > 0 ? A.radius_px + 1 : 1);
            if (tx < -margin || tx > (float)A.tile_w + margin || ty < -margin || ty > (float)A.tile_h + margin) continue;
            const float px = tx + (float)s.tile_x0, py = ty + (float)s.tile_y0;
            const float w = res.w * A.ray_weight;
            const float lum = src_avg * sp.z * w;
            const float cr = s.r * sp.y * w, cg = s.g * sp.z * w, cb = s.b * sp.w * w;
            auto put = [&](int xi, int yi, float wt) {
                if (xi < tx_min || xi >= tx_max || yi < ty_min || yi >= ty_max || xi < 0 || xi >= A.comp_w || yi < 0 ||
                    yi >= A.comp_h)
                    return;
                if (A.mono) {
                    if (fabsf(lum) > 1e-14f) psf_add<LDS>(A, tile, tx_min, ty_min, 0, xi, yi, lum * wt);
                } else {
                    if (fabsf(cr) > 1e-14f) psf_add<LDS>(A, tile, tx_min, ty_min, 0, xi, yi, cr * wt);
                    if (fabsf(cg) > 1e-14f) psf_add<LDS>(A, tile, tx_min, ty_min, 1, xi, yi, cg * wt);
                    if (fabsf(cb) > 1e-14f) psf_add<LDS>(A, tile, tx_min, ty_min, 2, xi, yi, cb * wt);
                }
            };
            if (A.sigma_px <= 0.0f) {
                const int x0 = (int)floorf(px - 0.5f), y0 = (int)floorf(py - 0.5f);
                const float fx = (px - 0.5f) - (float)x0, fy = (py - 0.5f) - (float)y0;
                put(x0, y0, (1.0f - fx) * (1.0f - fy));
                put(x0 + 1, y0, fx * (1.0f - fy));
                put(x0, y0 + 1, (1.0f - fx) * fy);
                put(x0 + 1, y0 + 1, fx * fy);
            } else {
                // Normalised Gaussian over (2r+1)^2 cells around the hit cell.
                const float inv_2sig2 = 1.0f / (2.0f * A.sigma_px * A.sigma_px);
                const int cx = (int)floorf(px), cy = (int)floorf(py);
                const float ox = px - ((float)cx + 0.5f), oy = py - ((float)cy + 0.5f);
                const int R = A.radius_px;
                // The normalising sum factors into row and column sums
                // (exp(-(a + b)) = exp(-a) exp(-b)): 2(2R+1) exponentials
                // instead of (2R+1)^2; it differs from the cell-by-cell sum by
                // rounding only (a few 1e-8 relative).
                float sx = 0.0f, sy = 0.0f;
                for (int q = -R; q <= R; ++q) {
                    const float dxr = (float)q - ox, dyr = (float)q - oy;
                    sx += expf(-(dxr * dxr) * inv_2sig2);
                    sy += expf(-(dyr * dyr) * inv_2sig2);
                }
                const float wsum = sx * sy;
                const float inv_wsum = (wsum > 1e-20f) ? (1.0f / wsum) : 0.0f;
                // Cell weights from the hardware base-2 exponential:
                // exp(t) = 2^(t log2 e), within a few 1e-7 relative for the
                // |t| <= 2 (2.5 + 1)^2 / 2 this footprint reaches.  Each lane
                // walks its row from its own column (lane mod (2R + 1)), so
                // the wave's LDS atomics spread over the footprint instead
                // of landing on the same few cells at once.
                const int W2 = 2 * R + 1;
                const int dx0 = -R + (int)(threadIdx.x & 63) % W2;
                for (int dy = -R; dy <= R; ++dy) {
                    const int yi = cy + dy;
                    if (yi < ty_min || yi >= ty_max || yi < 0 || yi >= A.comp_h) continue;
                    const float dyr = (float)dy - oy;
                    int dx = dx0;
                    for (int j = 0; j < W2; ++j) {
                        const float dxr = (float)dx - ox;
                        const float t = -(dxr * dxr + dyr * dyr) * inv_2sig2;
                        const float wt = __builtin_amdgcn_exp2f(t * 1.44269504088896341f) * inv_wsum;
                        if (wt >= 1e-6f) put(cx + dx, yi, wt);
                        dx = dx == R ? -R : dx + 1;
                    }
                }
            }
        }
    }
    if (LDS) {
        __syncthreads();
        for (int i = threadIdx.x; i < planes * TT; i += blockDim.x) {
            const float v = tile[i];
            if (v == 0.0f) continue;
            const int c = i / TT, p = i % TT;
            const int x = tx_min + p % A.tile_w, y = ty_min + p / A.tile_w;
            if (x < 0 || x >= A.comp_w || y < 0 || y >= A.comp_h) continue;
            float* o = c == 0 ? A.out_r : (c == 1 ? A.out_g : A.out_b);
            atomicAdd(o + (size_t)y * A.comp_w + x, v);
        }
    }
}

__global__ void psf_zero_kernel(float* __restrict__ a, float* __restrict__ b, float* __restrict__ c, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) { a[i] = 0.f; b[i] = 0.f; c[i] = 0.f; }
}

}  // namespace gl

using namespace gl;

extern "C" {

int gl_render_psf_tiles(gl_ctx* ctx, const gl_lens* lens, const gl_psf_source* sources, int32_t n_sources,
                        const gl_flare_params* rp, const gl_psf_params* pp, float* out_r, float* out_g, float* out_b,
                        float* ms_out) {
    if (!lens || !rp || !pp || !out_r || !out_g || !out_b) return set_error(GL_ERR_INVALID, "null argument");
    if (pp->grid_nx <= 0 || pp->grid_ny <= 0 || pp->tile_w <= 0 || pp->tile_h <= 0)
        return set_error(GL_ERR_INVALID, "render_psf: grid and tile dimensions must be positive");
    if (!(pp->tile_extent_mm > 0.0f) || !std::isfinite(pp->tile_extent_mm))
        return set_error(GL_ERR_INVALID, "render_psf: tile_extent_mm must be finite and positive");
    if (n_sources < 0 || (n_sources && !sources)) return set_error(GL_ERR_INVALID, "bad source array");
    if ((long long)n_sources > (long long)pp->grid_nx * pp->grid_ny)
        return set_error(GL_ERR_INVALID, "render_psf: %d sources exceed the %d tiles", n_sources, pp->grid_nx * pp->grid_ny);
    const long long cw = (long long)pp->grid_nx * pp->tile_w, ch = (long long)pp->grid_ny * pp->tile_h;
    if (cw * ch > (long long)INT32_MAX) return set_error(GL_ERR_INVALID, "render_psf: composite has too many pixels");
    const size_t npx = (size_t)(cw * ch);
    std::fill(out_r, out_r + npx, 0.0f); std::fill(out_g, out_g + npx, 0.0f); std::fill(out_b, out_b + npx, 0.0f);
    if (n_sources == 0) return GL_OK;
    int st;
    gl_ctx* c = resolve_ctx(ctx, &st);
    if (!c) return st;
    std::lock_guard<std::mutex> g(c->mu);
    gl_lens* L = const_cast<gl_lens*>(lens);
    const int n = (int)L->host.size();
    LensDev D;
    if ((st = lens_sync(L, rp->fp64_radius_min, c->stream, &D))) return st;

    // Pupil grid and spectral table (psf_render.cu:262-330).
    std::vector<gl_sample> grid = build_samples(L->host.data(), n, rp->ray_grid, rp->pupil_jitter, rp->jitter_seed,
                                                rp->aperture_blades, rp->aperture_rotation);
    if (grid.empty()) return GL_OK;
    std::vector<gl_spec> spec;
    if (pp->monochromatic) {
        spec = {{587.56f, 0.0f, 1.0f, 0.0f}};
    } else {
        const int ns = std::max(3, rp->spectral_samples);
        if ((st = spectral_table(ns, rp->output_cs, rp->custom_xyz_to_output, spec))) return st;
        if (ns != 3) {
            const float dl = (700.0f - 400.0f) / (float)ns;
            for (auto& s : spec) { s.cmf_r *= dl; s.cmf_g *= dl; s.cmf_b *= dl; }
        }
    }
    const int ns = (int)spec.size();
    std::vector<float> iat((size_t)ns * n), ibf((size_t)ns * n);
    ior_tables(L->host.data(), n, spec.data(), ns, iat.data(), ibf.data());
    std::vector<float4> spec4(ns);
    for (int i = 0; i < ns; ++i) spec4[i] = make_float4(spec[i].lambda, spec[i].cmf_r, spec[i].cmf_g, spec[i].cmf_b);

    DevBuf &bs = c->buf("psf_src"), &bg = c->buf("psf_grid"), &bsp = c->buf("psf_spec"), &bi = c->buf("psf_iat");
    DevBuf &bo = c->buf("psf_out");
    if ((st = bs.ensure((size_t)n_sources * sizeof(gl_psf_source))) || (st = bg.ensure(grid.size() * sizeof(gl_sample))) ||
        (st = bsp.ensure(spec4.size() * sizeof(float4))) || (st = bi.ensure(iat.size() * 4)) ||
        (st = bo.ensure(3 * npx * 4)))
        return st;
    GL_HIP(hipMemcpyAsync(bs.p, sources, (size_t)n_sources * sizeof(gl_psf_source), hipMemcpyHostToDevice, c->stream));
    GL_HIP(hipMemcpyAsync(bg.p, grid.data(), grid.size() * sizeof(gl_sample), hipMemcpyHostToDevice, c->stream));
    GL_HIP(hipMemcpyAsync(bsp.p, spec4.data(), spec4.size() * sizeof(float4), hipMemcpyHostToDevice, c->stream));
    GL_HIP(hipMemcpyAsync(bi.p, iat.data(), iat.size() * 4, hipMemcpyHostToDevice, c->stream));
    float* o = bo.as<float>();
    hipLaunchKernelGGL(psf_zero_kernel, dim3((unsigned)((npx + 255) / 256)), dim3(256), 0, c->stream, o, o + npx,
                       o + 2 * npx, npx);

    PsfArgs A;
    std::memset(&A, 0, sizeof A);
    A.S = D.surf; A.poly = D.poly; A.arena = D.arena; A.n = n;
    A.src = bs.as<gl_psf_source>(); A.n_src = n_sources;
    A.grid = bg.as<gl_sample>(); A.n_grid = (int)grid.size();
    A.front_R = L->host[0].semi_aperture; A.start_z = L->host[0].z - 20.0f;
    A.mm_per_px = pp->tile_extent_mm / (float)pp->tile_w;
    A.tile_w = pp->tile_w; A.tile_h = pp->tile_h; A.comp_w = (int)cw; A.comp_h = (int)ch;
    A.out_r = o; A.out_g = o + npx; A.out_b = o + 2 * npx;
    A.ray_weight = 1.0f / (float)grid.size();
    A.mono = pp->monochromatic ? 1 : 0;
    A.spec = bsp.as<float4>(); A.n_spec = ns;
    A.iat = bi.as<float>();
    // Splat footprint: sigma in sensor micrometres -> pixels (psf_render.cu:426-438).
    A.sigma_px = A.mm_per_px > 0.0f ? pp->splat_sigma_um / (A.mm_per_px * 1000.0f) : 0.0f;
    A.radius_px = 0;
    if (A.sigma_px > 0.0f) A.radius_px = std::min(16, std::max(1, (int)std::ceil(2.5f * A.sigma_px)));

    const size_t tile_bytes = (size_t)(A.mono ? 1 : 3) * pp->tile_w * pp->tile_h * sizeof(float);
    const dim3 grid_dim((unsigned)n_sources, (unsigned)((grid.size() + kPsfBlock - 1) / kPsfBlock));
    GL_HIP(hipEventRecord(c->ev0, c->stream));
    int lds_max = 0;
    if (hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, 0) != hipSuccess) lds_max = 65536;
    const bool lds = tile_bytes <= std::min<size_t>(kPsfMaxLdsTile, (size_t)lds_max);
    if (lds && D.lean == 1) hipLaunchKernelGGL((psf_kernel<true, 1>), grid_dim, dim3(kPsfBlock), tile_bytes, c->stream, A);
    else if (lds && D.lean == 2)
        hipLaunchKernelGGL((psf_kernel<true, 2>), grid_dim, dim3(kPsfBlock), tile_bytes, c->stream, A);
    else if (lds) hipLaunchKernelGGL((psf_kernel<true, 0>), grid_dim, dim3(kPsfBlock), tile_bytes, c->stream, A);
    else if (D.lean == 1) hipLaunchKernelGGL((psf_kernel<false, 1>), grid_dim, dim3(kPsfBlock), 0, c->stream, A);
    else if (D.lean == 2) hipLaunchKernelGGL((psf_kernel<false, 2>), grid_dim, dim3(kPsfBlock), 0, c->stream, A);
    else hipLaunchKernelGGL((psf_kernel<false, 0>), grid_dim, dim3(kPsfBlock), 0, c->stream, A);
    GL_HIP(hipGetLastError());
    GL_HIP(hipEventRecord(c->ev1, c->stream));
    GL_HIP(hipMemcpyAsync(out_r, o, npx * 4, hipMemcpyDeviceToHost, c->stream));
    GL_HIP(hipMemcpyAsync(out_g, o + npx, npx * 4, hipMemcpyDeviceToHost, c->stream));
    GL_HIP(hipMemcpyAsync(out_b, o + 2 * npx, npx * 4, hipMemcpyDeviceToHost, c->stream));
    GL_HIP(hipStreamSynchronize(c->stream));
    if (ms_out) GL_HIP(hipEventElapsedTime(ms_out, c->ev0, c->ev1));
    return GL_OK;
}

}  // extern "C"
