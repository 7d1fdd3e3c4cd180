// diffraction.hip — the aperture starburst (sprite engine), the veiling-glare
// pass and the FFT seam on gfx950.
//
//   FFT   (fft_backend.h:25)       2-D C2C in place: register-resident Stockham
//                                  row transforms of radix 16 staged through LDS
//                                  (gl_fft.h), columns by an in-place tile-pair
//                                  transpose, so the transform needs no scratch
//                                  array.  The last transpose of the PSF path
//                                  fuses |F|^2 and the fftshift, so the spectrum
//                                  is never re-read.
//   reductions                     two-level, FP64 partials, result left on the
//                                  device and consumed by the next kernel (no
//                                  host round trip, unlike starburst_render.cu:979-996).
//   resample                       gather form of splat_resample_kernel
//                                  (starburst_render.cu:316-384): one lane per
//                                  output pixel loops over the sources, so the
//                                  deposit is deterministic and atomic-free.
#include "gl_runtime.h"
#include "gl_fft.h"

#include <algorithm>
#include <cmath>
#include <cstdio>

namespace gl {

static constexpr float kLambdaRef = 550.0f;
static constexpr int kSpriteOversample = 3;
static constexpr int kSurvivorEdgeSS = 10;
static constexpr int kSurvivorCenterSS = 8;
static constexpr int kVeilGrid = 1024;
static constexpr int kVeilCoreDiv = 32;

// ---------------------------------------------------------------------------
// FFT
// ---------------------------------------------------------------------------

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// Radix-2 Stockham stage in global memory for rows longer than one
// workgroup's LDS: half-size `ns` (1, 2, 4, ..., N/2), out of place, all rows.
__global__ __launch_bounds__(256) void fft_stage_global_kernel(const float2* __restrict__ x, float2* __restrict__ y,
                                                               int N, int ns, int tshift,
                                                               const float2* __restrict__ tw) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int half = N >> 1;
    if (t >= (size_t)N * half) return;
    const size_t row = t / half;
    const int j = (int)(t % half);
    const float2* xr = x + row * N;
    float2* yr = y + row * N;
    const int k = j & (ns - 1);
    const float2 w = tw[k << tshift];
    const float2 a = xr[j];
    const float2 b = cmul(xr[j + half], w);
    const int base = ((j / ns) * 2 * ns) + k;
    yr[base] = make_float2(a.x + b.x, a.y + b.y);
    yr[base + ns] = make_float2(a.x - b.x, a.y - b.y);
}

// Final transpose of the PSF path fused with |F|^2 and the fftshift
// (psf_from_fft_kernel, starburst_render.cu:202-211):
//   psf[j][i] = |F[(j + N/2) % N][(i + N/2) % N]|^2, where F = T^T.
__global__ __launch_bounds__(256) void transpose_power_shift_kernel(const float2* __restrict__ T, float* __restrict__ psf,
                                                                    int N) {
    __shared__ float tile[32][33];
    const int bx = blockIdx.x * 32, by = blockIdx.y * 32;   // tile of T: rows by.., cols bx..
    for (int r = threadIdx.y; r < 32; r += 8) {
        const int y = by + r, x = bx + threadIdx.x;
        if (x < N && y < N) { const float2 z = T[(size_t)y * N + x]; tile[r][threadIdx.x] = z.x * z.x + z.y * z.y; }
    }
    __syncthreads();
    const int h = N >> 1;
    for (int r = threadIdx.y; r < 32; r += 8) {
        const int fr = bx + r, fc = by + threadIdx.x;       // F[fr][fc] = T[fc][fr]
        if (fr < N && fc < N) {
            const int j = (fr + h) & (N - 1), i = (fc + h) & (N - 1);
            psf[(size_t)j * N + i] = tile[threadIdx.x][r];
        }
    }
}

static int fft_twiddles(gl_ctx* c, int N, const float2** out) {
    char name[32];
    std::snprintf(name, sizeof name, "fft_tw_%d", N);
    DevBuf& b = c->buf(name);
    if (!b.p) {
        std::vector<float2> tw(std::max(1, N / 2));
        for (int k = 0; k < N / 2; ++k) {
            const double a = -2.0 * M_PI * (double)k / (double)N;
            tw[k] = make_float2((float)std::cos(a), (float)std::sin(a));
        }
        int st = b.ensure(tw.size() * sizeof(float2));
        if (st) return st;
        GL_HIP(hipMemcpyAsync(b.p, tw.data(), tw.size() * sizeof(float2), hipMemcpyHostToDevice, c->stream));
    }
    *out = b.as<float2>();
    return GL_OK;
}

static int ilog2(int N) { int l = 0; while ((1 << l) < N) ++l; return l; }

static int max_lds_bytes(gl_ctx* c) {
    if (c->max_lds <= 0) {
        int v = 0;
        if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, c->device) != hipSuccess || v <= 0)
            v = 65536;
        c->max_lds = v;
    }
    return c->max_lds;
}

// True when the LDS row transform of length N fits one workgroup.
bool fft_rows_fit(gl_ctx* c, int N) {
    return (long long)fft::rows_lds_bytes(ilog2(N)) <= (long long)max_lds_bytes(c);
}

static int fft_rows(gl_ctx* c, float2* d, int N) {
    if (!fft_rows_fit(c, N)) {
        // Row longer than a workgroup's LDS: ping-pong Stockham stages in HBM.
        const float2* tw;
        int st = fft_twiddles(c, N, &tw);
        if (st) return st;
        DevBuf& pp = c->buf("fft_pingpong");
        if ((st = pp.ensure((size_t)N * N * sizeof(float2)))) return st;
        float2 *x = d, *y = pp.as<float2>();
        const int logN = ilog2(N);
        const size_t work = (size_t)N * (N / 2);
        for (int s = 0; s < logN; ++s) {
            const int ns = 1 << s;
            hipLaunchKernelGGL(fft_stage_global_kernel, dim3((unsigned)((work + 255) / 256)), dim3(256), 0, c->stream,
                               x, y, N, ns, logN - 1 - s, tw);
            std::swap(x, y);
        }
        if (x != d) GL_HIP(hipMemcpyAsync(d, x, (size_t)N * N * sizeof(float2), hipMemcpyDeviceToDevice, c->stream));
        GL_HIP(hipGetLastError());
        return GL_OK;
    }
    GL_HIP(fft::rows(c->stream, ilog2(N), fft::RowsInPlace{d, N}, N));
    return GL_OK;
}

// 2-D forward FFT in place: rows, transpose, rows, transpose.  When psf !=
// nullptr the result is written as the shifted power spectrum instead (data
// is clobbered).
static int fft2d(gl_ctx* c, float2* data, int N, float* psf) {
    if (N < 2 || (N & (N - 1)) || N > 16384) return set_error(GL_ERR_INVALID, "FFT size %d is not a power of two in [2, 16384]", N);
    int st;
    if ((st = fft_rows(c, data, N))) return st;
    GL_HIP(fft::transpose_square_inplace(c->stream, data, N));
    if ((st = fft_rows(c, data, N))) return st;
    if (psf) {
        const dim3 tb(32, 8), tg((unsigned)((N + 31) / 32), (unsigned)((N + 31) / 32));
        hipLaunchKernelGGL(transpose_power_shift_kernel, tg, tb, 0, c->stream, data, psf, N);
        GL_HIP(hipGetLastError());
    } else {
        GL_HIP(fft::transpose_square_inplace(c->stream, data, N));
    }
    return GL_OK;
}

int fft2d_dev(gl_ctx* c, float2* data, int N, float* psf) { return fft2d(c, data, N, psf); }

// ---------------------------------------------------------------------------
// Reductions: FP64 partials, result stays on the device.
// ---------------------------------------------------------------------------

template <class T>
__global__ __launch_bounds__(256) void reduce_partial_kernel(const T* __restrict__ in, size_t n, int mode,
                                                             double* __restrict__ partial) {
    // mode 0: sum(x); mode 1: sum(|z|^2) for float2 inputs viewed as floats pairs.
    __shared__ double sm[256];
    double acc = 0.0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        if (mode == 0) acc += (double)in[i];
        else { const float a = in[2 * i], b = in[2 * i + 1]; acc += (double)(a * a + b * b); }
    }
    sm[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sm[threadIdx.x] += sm[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = sm[0];
}

__global__ __launch_bounds__(256) void reduce_final_kernel(const double* __restrict__ partial, int n,
                                                           double* __restrict__ out) {
    __shared__ double sm[256];
    double acc = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) acc += partial[i];
    sm[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sm[threadIdx.x] += sm[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = sm[0];
}

static constexpr int kReduceBlocks = 1024;

static int reduce_to(gl_ctx* c, const float* d_in, size_t n, int mode, double* d_out) {
    DevBuf& bp = c->buf("reduce_partial");
    int st = bp.ensure(kReduceBlocks * sizeof(double));
    if (st) return st;
    const int blocks = (int)std::min<size_t>(kReduceBlocks, (n + 255) / 256);
    hipLaunchKernelGGL(reduce_partial_kernel<float>, dim3((unsigned)std::max(1, blocks)), dim3(256), 0, c->stream, d_in,
                       n, mode, bp.as<double>());
    hipLaunchKernelGGL(reduce_final_kernel, dim3(1), dim3(256), 0, c->stream, bp.as<double>(), std::max(1, blocks), d_out);
    GL_HIP(hipGetLastError());
    return GL_OK;
}

// x *= 1 / sum   (sum read from the device; no-op for a non-positive sum)
__global__ void scale_by_inv_sum_kernel(float* __restrict__ d, size_t n, const double* __restrict__ sum) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double s = *sum;
    if (s > 0.0) d[i] *= (float)(1.0 / s);
}

// ---------------------------------------------------------------------------
// Starburst kernels
// ---------------------------------------------------------------------------

__device__ __forceinline__ float sample_env(const float* __restrict__ env, int ME, float u, float v) {
    float fx = (u * 0.5f + 0.5f) * (float)ME - 0.5f;
    float fy = (v * 0.5f + 0.5f) * (float)ME - 0.5f;
    fx = fminf(fmaxf(fx, 0.0f), (float)(ME - 1));
    fy = fminf(fmaxf(fy, 0.0f), (float)(ME - 1));
    const int x0 = (int)floorf(fx), y0 = (int)floorf(fy);
    const int x1 = min(x0 + 1, ME - 1), y1 = min(y0 + 1, ME - 1);
    const float tx = fx - (float)x0, ty = fy - (float)y0;
    const float a = env[(size_t)y0 * ME + x0] * (1.0f - tx) + env[(size_t)y0 * ME + x1] * tx;
    const float b = env[(size_t)y1 * ME + x0] * (1.0f - tx) + env[(size_t)y1 * ME + x1] * tx;
    return a * (1.0f - ty) + b * ty;
}

// What modifies the stop silhouette (PupilExtras, starburst_render.cu:104-133):
// the cat's-eye envelope and the front-of-lens baffle stack, the latter met by
// the pupil ray at (u ep_x, v ep_y) heading (tan_ax, tan_ay).  All zero: the
// plain stop.
struct PupilExtra {
    const float* env;
    int ME;
    float ep_x, ep_y, tan_ax, tan_ay;
    DevBaffles bf;
    // front-glass IMAGE aperture as graded transmission: the pupil ray crosses
    // that surface tex_d0 downstream of the spawn plane
    const float* tex;
    int img_w, img_h;
    float tex_aspect, tex_semi, tex_d0;
};

// Effective pupil A(u,v): 4x4-supersampled stop silhouette x cat's-eye
// envelope x baffle clip (build_pupil_kernel, starburst_render.cu:155-191).
__global__ __launch_bounds__(256) void pupil_kernel(float2* __restrict__ out, int N, float fill, DevMask mask,
                                                    PupilExtra ex) {
    const float* __restrict__ env = ex.env;
    const int ME = ex.ME;
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= N * N) return;
    const int i = idx % N, j = idx / N;
    float re = 0.0f;
#pragma unroll
    for (int sy = 0; sy < 4; ++sy)
#pragma unroll
        for (int sx = 0; sx < 4; ++sx) {
            const float u = (((float)i + ((float)sx + 0.5f) * 0.25f) / (float)N * 2.0f - 1.0f) / fill;
            const float v = (((float)j + ((float)sy + 0.5f) * 0.25f) / (float)N * 2.0f - 1.0f) / fill;
            if (!mask_contains(mask, u, v)) continue;
            float a = 1.0f;
            if (env != nullptr) a *= sample_env(env, ME, u, v);
            if (ex.bf.n > 0 && baffles_block(ex.bf, u * ex.ep_x, v * ex.ep_y, ex.tan_ax, ex.tan_ay)) continue;
            if (ex.tex != nullptr) {
                const float hx = (u * ex.ep_x + ex.tan_ax * ex.tex_d0) / ex.tex_aspect;
                const float hy = v * ex.ep_y + ex.tan_ay * ex.tex_d0;
                const float tu = 0.5f + hx / (2.0f * ex.tex_semi);
                const float tv = 0.5f + hy / (2.0f * ex.tex_semi);
                a *= image_sample(ex, ex.tex, tu, tv);
            }
            re += a;
        }
    out[idx] = make_float2(re * (1.0f / 16.0f), 0.0f);
}

__device__ __forceinline__ float bilinear_psf(const float* __restrict__ psf, float fx, float fy, int N) {
    if (fx < 0.0f || fy < 0.0f || fx >= (float)(N - 1) || fy >= (float)(N - 1)) return 0.0f;
    const int x0 = (int)floorf(fx), y0 = (int)floorf(fy);
    const float tx = fx - (float)x0, ty = fy - (float)y0;
    const float* r0 = psf + (size_t)y0 * N;
    const float* r1 = r0 + N;
    const float a = r0[x0] * (1.0f - tx) + r0[x0 + 1] * tx;
    const float b = r1[x0] * (1.0f - tx) + r1[x0 + 1] * tx;
    return a * (1.0f - ty) + b * ty;
}

// lambda bake by the scaling theorem (bake_sprite_kernel, starburst_render.cu:255-278).
__global__ __launch_bounds__(256) void bake_sprite_kernel(const float* __restrict__ psf, int N,
                                                          const float4* __restrict__ spec, int ns, float lambda_ref,
                                                          float* __restrict__ sprite) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= N * N) return;
    const int i = idx % N, j = idx / N;
    const float dxr = (float)(i - N / 2), dyr = (float)(j - N / 2), half = (float)(N / 2);
    float cr = 0.0f, cg = 0.0f, cb = 0.0f;
    for (int si = 0; si < ns; ++si) {
        const float4 s = spec[si];
        const float k = lambda_ref / s.x;
        const float val = bilinear_psf(psf, half + dxr * k, half + dyr * k, N) * k * k;
        cr += s.y * val; cg += s.z * val; cb += s.w * val;
    }
    sprite[3 * (size_t)idx] = cr;
    sprite[3 * (size_t)idx + 1] = cg;
    sprite[3 * (size_t)idx + 2] = cb;
}

// Box-sum downsample of a C-channel sprite (downsample_kernel, starburst_render.cu:283-299).
__global__ __launch_bounds__(256) void downsample_kernel(const float* __restrict__ fine, int N, int ds, int C,
                                                         float* __restrict__ coarse, int M) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= M * M) return;
    const int a = idx % M, b = idx / M;
    float s[3] = {0.f, 0.f, 0.f};
    for (int jj = 0; jj < ds; ++jj)
        for (int ii = 0; ii < ds; ++ii) {
            const float* p = fine + (size_t)C * ((size_t)(b * ds + jj) * N + (a * ds + ii));
            for (int c = 0; c < C; ++c) s[c] += p[c];
        }
    for (int c = 0; c < C; ++c) coarse[(size_t)C * idx + c] = s[c];
}

using StarSrc = StarSource;

// Area-weighted resample of a sprite onto the output, gather form: one lane
// per output pixel of the union footprint, summing every source whose
// footprint covers it (splat_resample_kernel, starburst_render.cu:316-384;
// veil_render.cu:170-212).  C = 3 (RGB sprite) or 1 (mono sprite).
__global__ __launch_bounds__(256) void resample_gather_kernel(const float* __restrict__ sprite, int M, int C,
                                                              float sx_ds, float sy_ds, float px_mm_x, float px_mm_y,
                                                              int rx, int ry, const StarSrc* __restrict__ src,
                                                              int n_src, float gain, int bx0, int by0, int bw, int bh,
                                                              float* __restrict__ out_r, float* __restrict__ out_g,
                                                              float* __restrict__ out_b, int width, int height) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= bw * bh) return;
    const int ix = bx0 + t % bw, iy = by0 + t / bw;
    if (ix < 0 || ix >= width || iy < 0 || iy >= height) return;
    float vr = 0.f, vg = 0.f, vb = 0.f;
    for (int q = 0; q < n_src; ++q) {
        const StarSrc s = src[q];
        const int cx = (int)floorf(s.px), cy = (int)floorf(s.py);
        if (ix < cx - rx || ix > cx + rx || iy < cy - ry || iy > cy + ry) continue;
        const float ax0 = (((float)ix - s.px) * px_mm_x) / sx_ds + 0.5f * (float)M - 0.5f;
        const float ax1 = (((float)ix + 1.0f - s.px) * px_mm_x) / sx_ds + 0.5f * (float)M - 0.5f;
        const float by0_ = (((float)iy - s.py) * px_mm_y) / sy_ds + 0.5f * (float)M - 0.5f;
        const float by1_ = (((float)iy + 1.0f - s.py) * px_mm_y) / sy_ds + 0.5f * (float)M - 0.5f;
        const float xlo = ax0 + 0.5f, xhi = ax1 + 0.5f, ylo = by0_ + 0.5f, yhi = by1_ + 0.5f;
        int i0 = (int)floorf(xlo), i1 = (int)floorf(xhi - 1e-6f);
        int j0 = (int)floorf(ylo), j1 = (int)floorf(yhi - 1e-6f);
        if (i0 < 0) i0 = 0;
        if (i1 > M - 1) i1 = M - 1;
        if (j0 < 0) j0 = 0;
        if (j1 > M - 1) j1 = M - 1;
        if (i1 < i0 || j1 < j0) continue;
        if (i1 - i0 > 64) i1 = i0 + 64;
        if (j1 - j0 > 64) j1 = j0 + 64;
        float ar = 0.f, ag = 0.f, ab = 0.f;
        for (int j = j0; j <= j1; ++j) {
            const float yov = fminf(yhi, (float)(j + 1)) - fmaxf(ylo, (float)j);
            if (yov <= 0.0f) continue;
            for (int i = i0; i <= i1; ++i) {
                const float xov = fminf(xhi, (float)(i + 1)) - fmaxf(xlo, (float)i);
                if (xov <= 0.0f) continue;
                const float w = xov * yov;
                const float* p = sprite + (size_t)C * ((size_t)j * M + i);
                if (C == 3) { ar += p[0] * w; ag += p[1] * w; ab += p[2] * w; }
                else { const float m = p[0] * w; ar += m; ag += m; ab += m; }
            }
        }
        vr += ar * s.r * gain; vg += ag * s.g * gain; vb += ab * s.b * gain;
    }
    const size_t pix = (size_t)iy * width + ix;
    out_r[pix] += vr;
    out_g[pix] += vg;
    out_b[pix] += vb;
}

// Primary-ray survival for the survivor centroid / envelope.
__global__ __launch_bounds__(256) void survive_kernel(const DevSurf* __restrict__ S, const float* __restrict__ poly,
                                                      const float* __restrict__ arena, int n,
                                                      const float2* __restrict__ launch, int n_launch, float dx,
                                                      float dy, float dz, float z0, float lam,
                                                      const float* __restrict__ iat, unsigned char* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_launch) return;
    const float2 l = launch[i];
    TraceRes r = trace_primary<false>(as_const(S), n, as_const(poly), arena, f3(l.x, l.y, z0), f3(dx, dy, dz), lam,
                               as_const(iat), true, nullptr);
    out[i] = (r.status == GL_TRACE_OK && isfinite(r.px) && isfinite(r.py)) ? 1 : 0;
}

// Veil glare-spread reference g(r) = (a^2 / (r^2 + a^2))^p (veil_render.cu:94-106).
__global__ __launch_bounds__(256) void veil_reference_kernel(float* __restrict__ ref, int N, float a_texels, float p) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= N * N) return;
    const float dx = (float)(idx % N - N / 2), dy = (float)(idx / N - N / 2);
    const float a2 = a_texels * a_texels;
    ref[idx] = powf(a2 / (dx * dx + dy * dy + a2), p);
}

__global__ void zero3_kernel(float* __restrict__ a, float* __restrict__ b, float* __restrict__ c, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) { a[i] = 0.f; b[i] = 0.f; c[i] = 0.f; }
}

// ---------------------------------------------------------------------------
// Host helpers
// ---------------------------------------------------------------------------

static int po2_floor(double x, int hi) {
    int p = 1;
    while (p * 2 <= x && p * 2 <= hi) p *= 2;
    return p < 1 ? 1 : p;
}

// Survival of launches (x, y) on the spawn plane, CPU-tracer semantics.
static int survive(gl_ctx* c, gl_lens* L, const std::vector<float2>& launch, float ax, float ay, float lam,
                   std::vector<unsigned char>& out) {
    out.assign(launch.size(), 0);
    if (launch.empty()) return GL_OK;
    int st;
    LensDev D;
    if ((st = lens_sync(L, 0.0f, c->stream, &D))) return st;
    const int n = (int)L->host.size();
    const float bx = std::tan(ax), by = std::tan(ay);   // float tangent, as the reference's host code
    float l = std::sqrt(bx * bx + by * by + 1.0f * 1.0f);
    const float inv = (l > 1e-12f) ? 1.0f / l : 0.0f;
    const float dx = bx * inv, dy = by * inv, dz = 1.0f * inv;
    gl_spec sp{lam, 0, 0, 0};
    std::vector<float> at(n), bf(n);
    ior_tables(L->host.data(), n, &sp, 1, at.data(), bf.data());
    DevBuf &bl = c->buf("sv_launch"), &bo = c->buf("sv_out"), &bi = c->buf("sv_iat");
    if ((st = bl.ensure(launch.size() * sizeof(float2))) || (st = bo.ensure(launch.size())) ||
        (st = bi.ensure(n * sizeof(float))))
        return st;
    GL_HIP(hipMemcpyAsync(bl.p, launch.data(), launch.size() * sizeof(float2), hipMemcpyHostToDevice, c->stream));
    GL_HIP(hipMemcpyAsync(bi.p, at.data(), n * sizeof(float), hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(survive_kernel, dim3((unsigned)((launch.size() + 255) / 256)), dim3(256), 0, c->stream, D.surf,
                       D.poly, D.arena, n, bl.as<float2>(), (int)launch.size(), dx, dy, dz,
                       L->host[0].z - 20.0f, lam, bi.as<float>(), bo.as<unsigned char>());
    GL_HIP(hipGetLastError());
    GL_HIP(hipMemcpyAsync(out.data(), bo.p, launch.size(), hipMemcpyDeviceToHost, c->stream));
    GL_HIP(hipStreamSynchronize(c->stream));
    return GL_OK;
}

// Survivor-crescent centroid (solve_survivor_center, starburst_render.cu:503-582).
static int survivor_center(gl_ctx* c, gl_lens* L, float ax, float ay, float lam, float& cx, float& cy, bool& any) {
    cx = cy = 0.0f; any = false;
    const float front_R = L->host[0].semi_aperture;
    if (front_R <= 0.0f) return GL_OK;
    const float bx = std::tan(ax), by = std::tan(ay);
    const float span_x = 1.1f * front_R + 20.0f * std::fabs(bx);
    const float span_y = 1.1f * front_R + 20.0f * std::fabs(by);
    const int S = 81;
    std::vector<float2> pts((size_t)S * S);
    for (int j = 0; j < S; ++j) {
        const float ly = span_y * (2.0f * (float)j / (float)(S - 1) - 1.0f);
        for (int i = 0; i < S; ++i) pts[(size_t)j * S + i] = make_float2(span_x * (2.0f * (float)i / (float)(S - 1) - 1.0f), ly);
    }
    std::vector<unsigned char> sv;
    int st = survive(c, L, pts, ax, ay, lam, sv);
    if (st) return st;
    const float hx = span_x / (float)(S - 1), hy = span_y / (float)(S - 1);
    const int K = kSurvivorCenterSS;
    const float inv_k2 = 1.0f / (float)(K * K);
    // Pass 1 marks the cells whose 8-neighbourhood changes state and lists
    // their K x K supersamples; one launch traces them all.
    std::vector<unsigned char> bnd((size_t)S * S, 0);
    std::vector<float2> ss;
    for (int j = 0; j < S; ++j)
        for (int i = 0; i < S; ++i) {
            const unsigned char s = sv[(size_t)j * S + i];
            bool b = false;
            for (int dj = -1; dj <= 1 && !b; ++dj)
                for (int di = -1; di <= 1; ++di) {
                    if (!di && !dj) continue;
                    const int ni = i + di, nj = j + dj;
                    if (ni < 0 || ni >= S || nj < 0 || nj >= S) continue;
                    if (sv[(size_t)nj * S + ni] != s) { b = true; break; }
                }
            if (!b) continue;
            bnd[(size_t)j * S + i] = 1;
            const float2 p = pts[(size_t)j * S + i];
            for (int sy = 0; sy < K; ++sy) {
                const float lly = p.y + hy * ((2.0f * ((float)sy + 0.5f) / (float)K) - 1.0f);
                for (int sx = 0; sx < K; ++sx)
                    ss.push_back(make_float2(p.x + hx * ((2.0f * ((float)sx + 0.5f) / (float)K) - 1.0f), lly));
            }
        }
    std::vector<unsigned char> sv2;
    if ((st = survive(c, L, ss, ax, ay, lam, sv2))) return st;
    // Pass 2 sums in raster order, each boundary cell's supersamples in place
    // of its centre, so the FP64 sums round as the reference's do.
    double sxs = 0.0, sys = 0.0, ws = 0.0;
    size_t k = 0;
    for (int j = 0; j < S; ++j)
        for (int i = 0; i < S; ++i) {
            const size_t idx = (size_t)j * S + i;
            if (!bnd[idx]) {
                if (sv[idx]) { sxs += pts[idx].x; sys += pts[idx].y; ws += 1.0; }
                continue;
            }
            for (int q = 0; q < K * K; ++q, ++k)
                if (sv2[k]) { sxs += (double)ss[k].x * inv_k2; sys += (double)ss[k].y * inv_k2; ws += inv_k2; }
        }
    if (ws > 0.0) { cx = (float)(sxs / ws); cy = (float)(sys / ws); any = true; }
    return GL_OK;
}

// Cat's-eye envelope (build_survivor_envelope, starburst_render.cu:393-487).
static int survivor_envelope(gl_ctx* c, gl_lens* L, const DevMask& stop, float ep_x, float ep_y, float cx, float cy,
                             float ax, float ay, float lam, int ME, std::vector<float>& env) {
    env.assign((size_t)ME * ME, -1.0f);
    std::vector<float2> pts;
    std::vector<int> where;
    for (int j = 0; j < ME; ++j)
        for (int i = 0; i < ME; ++i) {
            const float u = ((float)i + 0.5f) / (float)ME * 2.0f - 1.0f;
            const float v = ((float)j + 0.5f) / (float)ME * 2.0f - 1.0f;
            if (!mask_contains_host(stop, u, v)) continue;
            pts.push_back(make_float2(cx + u * ep_x, cy + v * ep_y));
            where.push_back(j * ME + i);
        }
    std::vector<unsigned char> sv;
    int st = survive(c, L, pts, ax, ay, lam, sv);
    if (st) return st;
    for (size_t k = 0; k < where.size(); ++k) env[where[k]] = sv[k] ? 1.0f : 0.0f;
    const std::vector<float> base = env;
    const int K = kSurvivorEdgeSS;
    const float du = 2.0f / (float)ME;
    std::vector<float2> ss;
    std::vector<int> cell, owner;
    std::vector<int> inside_cnt;
    for (int j = 0; j < ME; ++j)
        for (int i = 0; i < ME; ++i) {
            const size_t idx = (size_t)j * ME + i;
            const float e = base[idx];
            if (e < 0.0f) continue;
            bool bnd = false;
            for (int dj = -1; dj <= 1 && !bnd; ++dj)
                for (int di = -1; di <= 1; ++di) {
                    if (!di && !dj) continue;
                    const int ni = i + di, nj = j + dj;
                    if (ni < 0 || ni >= ME || nj < 0 || nj >= ME) continue;
                    const float ne = base[(size_t)nj * ME + ni];
                    if (ne >= 0.0f && ((ne > 0.5f) != (e > 0.5f))) { bnd = true; break; }
                }
            if (!bnd) continue;
            const float u0 = ((float)i + 0.5f) / (float)ME * 2.0f - 1.0f;
            const float v0 = ((float)j + 0.5f) / (float)ME * 2.0f - 1.0f;
            int inside = 0;
            for (int sy = 0; sy < K; ++sy)
                for (int sx = 0; sx < K; ++sx) {
                    const float uu = u0 + du * (((float)sx + 0.5f) / (float)K - 0.5f);
                    const float vv = v0 + du * (((float)sy + 0.5f) / (float)K - 0.5f);
                    if (!mask_contains_host(stop, uu, vv)) continue;
                    ++inside;
                    ss.push_back(make_float2(cx + uu * ep_x, cy + vv * ep_y));
                    owner.push_back((int)cell.size());
                }
            cell.push_back((int)idx);
            inside_cnt.push_back(inside);
        }
    std::vector<unsigned char> sv2;
    if ((st = survive(c, L, ss, ax, ay, lam, sv2))) return st;
    std::vector<int> surv(cell.size(), 0);
    for (size_t k = 0; k < ss.size(); ++k) surv[owner[k]] += sv2[k];
    for (size_t q = 0; q < cell.size(); ++q)
        if (inside_cnt[q] > 0) env[cell[q]] = (float)surv[q] / (float)inside_cnt[q];
    const int R = 2;
    std::vector<float> filled = env;
    for (int j = 0; j < ME; ++j)
        for (int i = 0; i < ME; ++i) {
            if (env[(size_t)j * ME + i] >= 0.0f) continue;
            float best = 0.0f;
            for (int dj = -R; dj <= R; ++dj)
                for (int di = -R; di <= R; ++di) {
                    const int ni = i + di, nj = j + dj;
                    if (ni < 0 || ni >= ME || nj < 0 || nj >= ME) continue;
                    const float e = env[(size_t)nj * ME + ni];
                    if (e >= 0.0f && e > best) best = e;
                }
            filled[(size_t)j * ME + i] = best;
        }
    env.swap(filled);
    return GL_OK;
}

// Starburst spectral table: output-space CMF x d_lambda, white-balanced per
// channel (starburst_render.cu:709-730).
static int star_spectrum(const gl_diffraction_params* p, std::vector<float4>& out) {
    int ns = p->spectral_samples > 0 ? p->spectral_samples : p->render_spectral_samples;
    if (ns < 3) ns = 3;
    // The starburst always integrates the CMF table, also for three samples.
    std::vector<gl_spec> spec;
    int st = spectral_table_cmf(ns, p->output_cs, p->custom_xyz_to_output, spec);
    if (st) return st;
    const float dl = (700.0f - 400.0f) / (float)ns;
    for (auto& s : spec) { s.cmf_r *= dl; s.cmf_g *= dl; s.cmf_b *= dl; }
    double sr = 0, sg = 0, sb = 0;
    for (const auto& s : spec) { sr += s.cmf_r; sg += s.cmf_g; sb += s.cmf_b; }
    for (auto& s : spec) {
        if (sr > 0.0) s.cmf_r /= (float)sr;
        if (sg > 0.0) s.cmf_g /= (float)sg;
        if (sb > 0.0) s.cmf_b /= (float)sb;
    }
    out.resize(spec.size());
    for (size_t i = 0; i < spec.size(); ++i) out[i] = make_float4(spec[i].lambda, spec[i].cmf_r, spec[i].cmf_g, spec[i].cmf_b);
    return GL_OK;
}

static int resample(gl_ctx* c, const float* d_sprite, int M, int C, float sx_ds, float sy_ds, float px_mm_x,
                    float px_mm_y, const std::vector<StarSrc>& src, float gain, float* o_r, float* o_g, float* o_b,
                    int width, int height) {
    int rx = (int)std::ceil(0.5 * M * sx_ds / px_mm_x) + 1;
    int ry = (int)std::ceil(0.5 * M * sy_ds / px_mm_y) + 1;
    if (rx > width) rx = width;
    if (ry > height) ry = height;
    int x0 = width, y0 = height, x1 = -1, y1 = -1;
    for (const auto& s : src) {
        const int cxp = (int)std::floor(s.px), cyp = (int)std::floor(s.py);
        x0 = std::min(x0, cxp - rx); x1 = std::max(x1, cxp + rx);
        y0 = std::min(y0, cyp - ry); y1 = std::max(y1, cyp + ry);
    }
    x0 = std::max(x0, 0); y0 = std::max(y0, 0); x1 = std::min(x1, width - 1); y1 = std::min(y1, height - 1);
    if (x1 < x0 || y1 < y0) return GL_OK;
    const int bw = x1 - x0 + 1, bh = y1 - y0 + 1;
    void* d_src;
    DevBuf& b = c->buf("star_src");
    int st = b.ensure(src.size() * sizeof(StarSrc));
    if (st) return st;
    d_src = b.p;
    GL_HIP(hipMemcpyAsync(d_src, src.data(), src.size() * sizeof(StarSrc), hipMemcpyHostToDevice, c->stream));
    const long long total = (long long)bw * bh;
    hipLaunchKernelGGL(resample_gather_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, c->stream, d_sprite,
                       M, C, sx_ds, sy_ds, px_mm_x, px_mm_y, rx, ry, static_cast<const StarSrc*>(d_src),
                       (int)src.size(), gain, x0, y0, bw, bh, o_r, o_g, o_b, width, height);
    GL_HIP(hipGetLastError());
    return GL_OK;
}

static int out_planes(gl_ctx* c, const char* tag, size_t npx, float** r, float** g, float** b) {
    char n1[48], n2[48], n3[48];
    std::snprintf(n1, sizeof n1, "%s_r", tag);
    std::snprintf(n2, sizeof n2, "%s_g", tag);
    std::snprintf(n3, sizeof n3, "%s_b", tag);
    DevBuf &br = c->buf(n1), &bg = c->buf(n2), &bb = c->buf(n3);
    int st;
    if ((st = br.ensure(npx * 4)) || (st = bg.ensure(npx * 4)) || (st = bb.ensure(npx * 4))) return st;
    *r = br.as<float>(); *g = bg.as<float>(); *b = bb.as<float>();
    hipLaunchKernelGGL(zero3_kernel, dim3((unsigned)((npx + 255) / 256)), dim3(256), 0, c->stream, *r, *g, *b, npx);
    GL_HIP(hipGetLastError());
    return GL_OK;
}

static int download3(gl_ctx* c, const float* r, const float* g, const float* b, size_t npx, float* o_r, float* o_g,
                     float* o_b) {
    GL_HIP(hipMemcpyAsync(o_r, r, npx * 4, hipMemcpyDeviceToHost, c->stream));
    GL_HIP(hipMemcpyAsync(o_g, g, npx * 4, hipMemcpyDeviceToHost, c->stream));
    GL_HIP(hipMemcpyAsync(o_b, b, npx * 4, hipMemcpyDeviceToHost, c->stream));
    GL_HIP(hipStreamSynchronize(c->stream));
    return GL_OK;
}


}  // namespace gl

using namespace gl;

extern "C" {

int gl_fft2d_c2c_forward_inplace(gl_ctx* ctx, void* d, int32_t N) {
    int st;
    gl_ctx* c = resolve_ctx(ctx, &st);
    if (!c) return st;
    if (!d) return set_error(GL_ERR_INVALID, "null data");
    // The long-row form keeps its ping-pong plane and twiddles in the
    // context's buffer cache: one call at a time per context.
    std::lock_guard<std::mutex> g(c->mu);
    return fft2d(c, static_cast<float2*>(d), N, nullptr);
}

int gl_fft2d_c2c_forward_host(gl_ctx* ctx, const float* in, float* out, int32_t N) {
    int st;
    gl_ctx* c = resolve_ctx(ctx, &st);
    if (!c) return st;
    std::lock_guard<std::mutex> g(c->mu);
    if (N < 2 || (N & (N - 1)) || N > 16384) return set_error(GL_ERR_INVALID, "FFT size %d is not a power of two in [2, 16384]", N);
    const size_t bytes = (size_t)N * N * sizeof(float2);
    DevBuf& d = c->buf("fft_host_data");
    if ((st = d.ensure(bytes))) return st;
    GL_HIP(hipMemcpyAsync(d.p, in, bytes, hipMemcpyHostToDevice, c->stream));
    if ((st = fft2d(c, d.as<float2>(), N, nullptr))) return st;
    GL_HIP(hipMemcpyAsync(out, d.p, bytes, hipMemcpyDeviceToHost, c->stream));
    GL_HIP(hipStreamSynchronize(c->stream));
    return GL_OK;
}

int gl_fft2d_time(gl_ctx* ctx, int32_t N, int32_t reps, float* ms_per_transform) {
    if (!ms_per_transform || reps < 1) return set_error(GL_ERR_INVALID, "null out or reps < 1");
    int st;
    gl_ctx* c = resolve_ctx(ctx, &st);
    if (!c) return st;
    std::lock_guard<std::mutex> g(c->mu);
    if (N < 2 || (N & (N - 1)) || N > 16384) return set_error(GL_ERR_INVALID, "FFT size %d is not a power of two in [2, 16384]", N);
    const size_t bytes = (size_t)N * N * sizeof(float2);
    DevBuf& d = c->buf("fft_time_data");
    if ((st = d.ensure(bytes))) return st;
    GL_HIP(hipMemsetAsync(d.p, 0, bytes, c->stream));
    if ((st = fft2d(c, d.as<float2>(), N, nullptr))) return st;
    GL_HIP(hipEventRecord(c->ev0, c->stream));
    for (int i = 0; i < reps; ++i)
        if ((st = fft2d(c, d.as<float2>(), N, nullptr))) return st;
    GL_HIP(hipEventRecord(c->ev1, c->stream));
    GL_HIP(hipEventSynchronize(c->ev1));
    float ms = 0.0f;
    GL_HIP(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    *ms_per_transform = ms / (float)reps;
    return GL_OK;
}

int gl_primary_survival(gl_ctx* ctx, const gl_lens* lens, const float* launch, int32_t n, float dx, float dy, float dz,
                        float z_spawn, float lam, unsigned char* out) {
    if (!lens || !launch || !out || n < 0) return set_error(GL_ERR_INVALID, "null argument");
    int st;
    gl_ctx* c = resolve_ctx(ctx, &st);
    if (!c) return st;
    std::lock_guard<std::mutex> g(c->mu);
    gl_lens* L = const_cast<gl_lens*>(lens);
    if (n == 0) return GL_OK;
    LensDev D;
    if ((st = lens_sync(L, 0.0f, c->stream, &D))) return st;
    const int ns = (int)L->host.size();
    gl_spec sp{lam, 0, 0, 0};
    std::vector<float> at(ns), bf(ns);
    ior_tables(L->host.data(), ns, &sp, 1, at.data(), bf.data());
    DevBuf &bl = c->buf("sv_launch"), &bo = c->buf("sv_out"), &bi = c->buf("sv_iat");
    if ((st = bl.ensure((size_t)n * sizeof(float2))) || (st = bo.ensure(n)) || (st = bi.ensure(ns * sizeof(float))))
        return st;
    GL_HIP(hipMemcpyAsync(bl.p, launch, (size_t)n * sizeof(float2), hipMemcpyHostToDevice, c->stream));
    GL_HIP(hipMemcpyAsync(bi.p, at.data(), ns * sizeof(float), hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(survive_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, D.surf, D.poly,
                       D.arena, ns, bl.as<float2>(), n, dx, dy, dz, z_spawn, lam, bi.as<float>(),
                       bo.as<unsigned char>());
    GL_HIP(hipGetLastError());
    GL_HIP(hipMemcpyAsync(out, bo.p, n, hipMemcpyDeviceToHost, c->stream));
    GL_HIP(hipStreamSynchronize(c->stream));
    return GL_OK;
}

int gl_render_starburst(gl_ctx* ctx, const gl_lens* lens, const gl_calibration* calib, const gl_source* sources,
                        const float* src_px, const float* src_py, int32_t n_src, int32_t width, int32_t height,
                        int32_t fmt_w, int32_t fmt_h, float sensor_half_w, float sensor_half_h,
                        const gl_diffraction_params* p, float* out_r, float* out_g, float* out_b,
                        gl_starburst_debug* dbg) {
    if (!lens || !calib || !p || !out_r || !out_g || !out_b) return set_error(GL_ERR_INVALID, "null argument");
    if (width <= 0 || height <= 0 || fmt_w <= 0 || fmt_h <= 0) return set_error(GL_ERR_INVALID, "degenerate output dimensions");
    const size_t npx = (size_t)width * height;
    std::fill(out_r, out_r + npx, 0.0f); std::fill(out_g, out_g + npx, 0.0f); std::fill(out_b, out_b + npx, 0.0f);
    if (n_src <= 0) return GL_OK;
    int st;
    gl_ctx* c = resolve_ctx(ctx, &st);
    if (!c) return st;
    std::lock_guard<std::mutex> g(c->mu);
    gl_lens* L = const_cast<gl_lens*>(lens);

    // ---- first-order optics (starburst_render.cu:619-635)
    const float front_R = L->host[0].semi_aperture;
    float efl_x = calib->focal_length_x, efl_y = calib->focal_length_y;
    float ep_x = calib->entrance_pupil_semi_x, ep_y = calib->entrance_pupil_semi_y;
    if (efl_x <= 0.0f && calib->max_half_angle_h > 0.0f) efl_x = sensor_half_w / std::tan(calib->max_half_angle_h);
    if (efl_y <= 0.0f && calib->max_half_angle_v > 0.0f) efl_y = sensor_half_h / std::tan(calib->max_half_angle_v);
    if (ep_x <= 0.0f) ep_x = front_R;
    if (ep_y <= 0.0f) ep_y = front_R;
    if (efl_x <= 0.0f || efl_y <= 0.0f || ep_x <= 0.0f || ep_y <= 0.0f)
        return set_error(GL_ERR_INVALID, "degenerate first-order optics (no focal length / pupil)");
    const float fnum_x = efl_x / (2.0f * ep_x), fnum_y = efl_y / (2.0f * ep_y);

    int N = p->starburst_grid;
    if (N < 256 || (N & (N - 1)) != 0) return set_error(GL_ERR_INVALID, "starburst_grid must be a power of two >= 256");
    const float fill = (p->pupil_fill > 0.02f && p->pupil_fill < 0.9f) ? p->pupil_fill : 0.30f;
    const float lambda_ref_mm = kLambdaRef * 1e-6f;
    const float dx_fine_x = lambda_ref_mm * fnum_x * fill, dx_fine_y = lambda_ref_mm * fnum_y * fill;

    // ---- grid growth / engine routing (starburst_render.cu:648-702)
    int engine = p->starburst_engine;
    {
        const float trim0 = p->scale_trim > 0.0f ? p->scale_trim : 1.0f;
        const float dx0_x = trim0 * dx_fine_x, dx0_y = trim0 * dx_fine_y;
        const float frame_w = 2.0f * sensor_half_w, frame_h = 2.0f * sensor_half_h;
        const float frac = std::fmax(frame_w > 0.0f ? (float)N * dx0_x / frame_w : 0.0f,
                                     frame_h > 0.0f ? (float)N * dx0_y / frame_h : 0.0f);
        int grid_cap = p->starburst_grid_cap >= 256 ? p->starburst_grid_cap : 8192;
        if (grid_cap > 16384) grid_cap = 16384;
        { int q = 256; while ((q << 1) <= grid_cap) q <<= 1; grid_cap = q; }
        if (grid_cap < N) grid_cap = N;
        const float circle_w = calib->image_circle_semi_w > 0.0f ? calib->image_circle_semi_w : calib->sensor_half_w;
        const float circle_h = calib->image_circle_semi_h > 0.0f ? calib->image_circle_semi_h : calib->sensor_half_h;
        const bool oversized = (circle_w > 0.0f && sensor_half_w > 1.5f * circle_w) ||
                               (circle_h > 0.0f && sensor_half_h > 1.5f * circle_h);
        const bool prominent = frac >= 0.20f;
        if (prominent) engine = GL_STARBURST_SPRITE;
        if (frac < 1.0f && (prominent || oversized) && dx0_x > 0.0f && dx0_y > 0.0f && grid_cap > N) {
            float reach_x = 0.0f, reach_y = 0.0f;
            for (int i = 0; i < n_src; ++i) {
                reach_x = std::fmax(reach_x, std::fmax(src_px[i], (float)width - src_px[i]));
                reach_y = std::fmax(reach_y, std::fmax(src_py[i], (float)height - src_py[i]));
            }
            const float pbx = frame_w / (float)width, pby = frame_h / (float)height;
            const double covN = std::fmax(2.0 * reach_x * pbx / dx0_x, 2.0 * reach_y * pby / dx0_y);
            int want = covN > (double)N ? (int)std::ceil(covN) : N;
            if (want > grid_cap) want = grid_cap;
            int Ne = 256; while (Ne < want) Ne <<= 1;
            if (Ne > N) { N = Ne; engine = GL_STARBURST_SPRITE; }
        }
    }

    std::vector<float4> spec;
    if ((st = star_spectrum(p, spec))) return st;

    // Pupil silhouette, aspect forced to 1 (the pitch is per axis).
    DevMask mask = build_mask(L->host.data(), (int)L->host.size(), p->aperture_blades, p->aperture_rotation);
    mask.aspect = 1.0f;

    double sax = 0.0, say = 0.0;
    for (int i = 0; i < n_src; ++i) { sax += sources[i].angle_x; say += sources[i].angle_y; }
    const float field_ax = (float)(sax / (double)n_src), field_ay = (float)(say / (double)n_src);
    float chief_cx = 0.0f, chief_cy = 0.0f;
    std::vector<float> env;
    int env_N = 0;
    const bool off_axis = std::fabs(field_ax) > 1e-5f || std::fabs(field_ay) > 1e-5f;
    const bool want_env = p->use_survivor_mask && off_axis;
    auto fill_dbg = [&](float T) {
        if (!dbg) return;
        dbg->grid = N;
        dbg->f_number_x = fnum_x; dbg->f_number_y = fnum_y;
        dbg->focal_length_x = efl_x; dbg->focal_length_y = efl_y;
        dbg->entrance_pupil_semi_x = ep_x; dbg->entrance_pupil_semi_y = ep_y;
        dbg->lambda_ref_nm = kLambdaRef;
        dbg->field_angle_x = field_ax; dbg->field_angle_y = field_ay;
        dbg->throughput = T;
        dbg->chief_offset_x = chief_cx; dbg->chief_offset_y = chief_cy;
    };
    if (want_env) {
        bool any = false;
        if ((st = survivor_center(c, L, field_ax, field_ay, kLambdaRef, chief_cx, chief_cy, any))) return st;
        if (!any) { fill_dbg(0.0f); return GL_OK; }
        int ME = p->survivor_grid;
        ME = ME < 16 ? 16 : (ME > 1024 ? 1024 : ME);
        DevMask real_stop = build_mask(L->host.data(), (int)L->host.size(), 0, 0.0f);
        real_stop.aspect = 1.0f;
        if ((st = survivor_envelope(c, L, real_stop, ep_x, ep_y, chief_cx, chief_cy, field_ax, field_ay, kLambdaRef,
                                    ME, env)))
            return st;
        env_N = ME;
    }

    const size_t NN = (size_t)N * N;
    DevBuf &bp = c->buf("star_pupil"), &bpsf = c->buf("star_psf");
    DevBuf &bsp = c->buf("star_sprite"), &bds = c->buf("star_sprite_ds"), &bspec = c->buf("star_spec");
    DevBuf &benv = c->buf("star_env"), &bsum = c->buf("star_sum");
    if ((st = bp.ensure(NN * 8)) || (st = bpsf.ensure(NN * 4)) ||
        (st = bsp.ensure(NN * 12)) || (st = bds.ensure(NN * 12)) || (st = bspec.ensure(spec.size() * 16)) ||
        (st = bsum.ensure(4 * sizeof(double))))
        return st;
    GL_HIP(hipMemcpyAsync(bspec.p, spec.data(), spec.size() * 16, hipMemcpyHostToDevice, c->stream));
    const float* d_env = nullptr;
    if (env_N) {
        if ((st = benv.ensure(env.size() * 4))) return st;
        GL_HIP(hipMemcpyAsync(benv.p, env.data(), env.size() * 4, hipMemcpyHostToDevice, c->stream));
        d_env = benv.as<float>();
    }
    const unsigned blocks = (unsigned)((NN + 255) / 256);
    double T = 1.0;
    double* dsum = bsum.as<double>();
    PupilExtra plain{}, eff{};
    eff.env = d_env; eff.ME = env_N;
    eff.ep_x = ep_x; eff.ep_y = ep_y;
    eff.tan_ax = std::tan(field_ax); eff.tan_ay = std::tan(field_ay);
    eff.bf.n = std::max(0, std::min((int)p->n_baffles, GL_MAX_BAFFLES));
    for (int k = 0; k < eff.bf.n; ++k) eff.bf.b[k] = p->baffles[k];
    if (p->use_surface_textures) {
        // the first surface carrying a loaded bitmap (starburst_render.cu:807-846)
        const std::vector<gl_surface>& hs = L->host;
        for (size_t k = 0; k < hs.size(); ++k) {
            const gl_surface& g = hs[k];
            if (g.aperture_shape != GL_APERTURE_IMAGE || g.image_width <= 0 || g.image_height <= 0) continue;
            if ((long long)g.image_offset + (long long)g.image_width * g.image_height > (long long)L->arena.size())
                break;
            if (g.aperture_semi_diameter > 0.0f) {
                LensDev D;
                if ((st = lens_sync(L, 0.0f, c->stream, &D))) return st;
                eff.tex = D.arena + g.image_offset;
                eff.img_w = g.image_width; eff.img_h = g.image_height;
                eff.tex_aspect = g.aperture_aspect > 0.0f ? g.aperture_aspect : 1.0f;
                eff.tex_semi = g.aperture_semi_diameter;
                eff.tex_d0 = g.z - (hs[0].z - 20.0f);
            }
            break;
        }
    }
    if (d_env || eff.bf.n > 0 || eff.tex) {
        hipLaunchKernelGGL(pupil_kernel, dim3(blocks), dim3(256), 0, c->stream, bp.as<float2>(), N, fill, mask, plain);
        if ((st = reduce_to(c, bp.as<float>(), NN, 1, dsum))) return st;
        hipLaunchKernelGGL(pupil_kernel, dim3(blocks), dim3(256), 0, c->stream, bp.as<float2>(), N, fill, mask, eff);
        if ((st = reduce_to(c, bp.as<float>(), NN, 1, dsum + 1))) return st;
        double s2[2];
        GL_HIP(hipMemcpyAsync(s2, dsum, 2 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        GL_HIP(hipStreamSynchronize(c->stream));
        T = s2[0] > 0.0 ? s2[1] / s2[0] : 1.0;
        T = T < 0.0 ? 0.0 : (T > 1.0 ? 1.0 : T);
    } else {
        hipLaunchKernelGGL(pupil_kernel, dim3(blocks), dim3(256), 0, c->stream, bp.as<float2>(), N, fill, mask, plain);
    }
    GL_HIP(hipGetLastError());
    if (dbg && dbg->pupil && dbg->grid_cap >= N) {
        std::vector<float2> hp(NN);
        GL_HIP(hipMemcpyAsync(hp.data(), bp.p, NN * 8, hipMemcpyDeviceToHost, c->stream));
        GL_HIP(hipStreamSynchronize(c->stream));
        for (size_t k = 0; k < NN; ++k) dbg->pupil[k] = hp[k].x;
    }
    if (want_env && T < 1e-3) { fill_dbg((float)T); return GL_OK; }

    const float px_mm_x = 2.0f * sensor_half_w / (float)fmt_w, px_mm_y = 2.0f * sensor_half_h / (float)fmt_h;
    const float trim = p->scale_trim > 0.0f ? p->scale_trim : 1.0f;
    const float aperture_scale = (2.8f * 2.8f) * calib->pupil_area_frac / std::fmax(fnum_x * fnum_y, 1e-6f);
    const float gain = p->starburst_gain * (float)T * aperture_scale;
    float *o_r, *o_g, *o_b;
    if ((st = out_planes(c, "star_out", npx, &o_r, &o_g, &o_b))) return st;

    if (engine == GL_STARBURST_MDFT) {
        std::vector<StarSource> ms(n_src);
        for (int i = 0; i < n_src; ++i) ms[i] = {src_px[i], src_py[i], sources[i].r, sources[i].g, sources[i].b};
        if ((st = mdft_render(c, bp.as<float2>(), N, fill, spec, kLambdaRef, trim * dx_fine_x, trim * dx_fine_y, px_mm_x,
                              px_mm_y, ms, gain, o_r, o_g, o_b, width, height, p->mdft_precision)))
            return st;
        fill_dbg((float)T);
        if (dbg) { dbg->dx_mm_x = trim * dx_fine_x; dbg->dx_mm_y = trim * dx_fine_y; }
        return download3(c, o_r, o_g, o_b, npx, out_r, out_g, out_b);
    }

    // ---- FFT -> shifted power spectrum -> unit sum -> lambda bake
    if ((st = fft2d(c, bp.as<float2>(), N, bpsf.as<float>()))) return st;
    if ((st = reduce_to(c, bpsf.as<float>(), NN, 0, dsum + 2))) return st;
    hipLaunchKernelGGL(scale_by_inv_sum_kernel, dim3(blocks), dim3(256), 0, c->stream, bpsf.as<float>(), NN, dsum + 2);
    hipLaunchKernelGGL(bake_sprite_kernel, dim3(blocks), dim3(256), 0, c->stream, bpsf.as<float>(), N,
                       bspec.as<float4>(), (int)spec.size(), kLambdaRef, bsp.as<float>());
    GL_HIP(hipGetLastError());

    const float sx = trim * dx_fine_x, sy = trim * dx_fine_y;
    const double ds_target = std::min((double)px_mm_x / (sx * kSpriteOversample), (double)px_mm_y / (sy * kSpriteOversample));
    const int ds = po2_floor(ds_target, N / 8);
    const int M = N / ds;
    const float* splat_src = bsp.as<float>();
    if (ds > 1) {
        hipLaunchKernelGGL(downsample_kernel, dim3((unsigned)(((size_t)M * M + 255) / 256)), dim3(256), 0, c->stream,
                           bsp.as<float>(), N, ds, 3, bds.as<float>(), M);
        GL_HIP(hipGetLastError());
        splat_src = bds.as<float>();
    }
    std::vector<StarSrc> ss(n_src);
    for (int i = 0; i < n_src; ++i) ss[i] = {src_px[i], src_py[i], sources[i].r, sources[i].g, sources[i].b};
    double sum_host = 0.0;
    GL_HIP(hipMemcpyAsync(&sum_host, dsum + 2, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    if ((st = resample(c, splat_src, M, 3, sx * ds, sy * ds, px_mm_x, px_mm_y, ss, gain, o_r, o_g, o_b, width, height)))
        return st;
    if ((st = download3(c, o_r, o_g, o_b, npx, out_r, out_g, out_b))) return st;
    if (!(sum_host > 0.0)) {   // degenerate pupil: clean no-op (starburst_render.cu:988-995)
        std::fill(out_r, out_r + npx, 0.0f); std::fill(out_g, out_g + npx, 0.0f); std::fill(out_b, out_b + npx, 0.0f);
    }
    fill_dbg((float)T);
    if (dbg) {
        dbg->dx_mm_x = sx; dbg->dx_mm_y = sy;
        if (dbg->grid_cap >= N) {
            if (dbg->psf) GL_HIP(d2h_sync(c, dbg->psf, bpsf.p, NN * 4));
            if (dbg->sprite_rgb) GL_HIP(d2h_sync(c, dbg->sprite_rgb, bsp.p, NN * 12));
        }
    }
    return GL_OK;
}

int gl_render_veil(gl_ctx* ctx, const gl_source* sources, const float* src_px, const float* src_py, int32_t n_src,
                   int32_t width, int32_t height, int32_t fmt_w, int32_t fmt_h, float sensor_half_w,
                   float sensor_half_h, const gl_diffraction_params* p, float* out_r, float* out_g, float* out_b,
                   gl_veil_debug* dbg) {
    if (!p || !out_r || !out_g || !out_b) return set_error(GL_ERR_INVALID, "null argument");
    if (width <= 0 || height <= 0 || fmt_w <= 0 || fmt_h <= 0) return set_error(GL_ERR_INVALID, "degenerate output dimensions");
    const size_t npx = (size_t)width * height;
    std::fill(out_r, out_r + npx, 0.0f); std::fill(out_g, out_g + npx, 0.0f); std::fill(out_b, out_b + npx, 0.0f);
    if (n_src <= 0) return GL_OK;
    const float half_ref = sensor_half_h > 0.0f ? sensor_half_h : sensor_half_w;
    if (half_ref <= 0.0f) return set_error(GL_ERR_INVALID, "degenerate sensor half-size");
    float spread = p->veil_spread;
    if (!(spread > 0.0f)) spread = 0.12f;
    spread = spread < 1e-3f ? 1e-3f : (spread > 4.0f ? 4.0f : spread);
    const float core_mm = spread * half_ref;
    float fp = p->veil_falloff;
    if (!(fp >= 1.0f)) fp = 1.0f;
    if (fp > 3.0f) fp = 3.0f;
    const int N = kVeilGrid;
    const float a_texels = (float)N / (float)kVeilCoreDiv;
    const float dx_fine = core_mm / a_texels;
    const size_t NN = (size_t)N * N;
    int st;
    gl_ctx* c = resolve_ctx(ctx, &st);
    if (!c) return st;
    std::lock_guard<std::mutex> g(c->mu);
    DevBuf &bref = c->buf("veil_ref"), &bds = c->buf("veil_ds"), &bsum = c->buf("veil_sum");
    if ((st = bref.ensure(NN * 4)) || (st = bds.ensure(NN * 4)) || (st = bsum.ensure(sizeof(double)))) return st;
    // The reference is cached per falloff (veil_render.cu:331-347).
    float& baked = c->veil_baked_p;
    if (baked != fp) {
        hipLaunchKernelGGL(veil_reference_kernel, dim3((unsigned)((NN + 255) / 256)), dim3(256), 0, c->stream,
                           bref.as<float>(), N, a_texels, fp);
        if ((st = reduce_to(c, bref.as<float>(), NN, 0, bsum.as<double>()))) return st;
        hipLaunchKernelGGL(scale_by_inv_sum_kernel, dim3((unsigned)((NN + 255) / 256)), dim3(256), 0, c->stream,
                           bref.as<float>(), NN, bsum.as<double>());
        GL_HIP(hipGetLastError());
        baked = fp;
    }
    const float px_mm_x = 2.0f * sensor_half_w / (float)fmt_w, px_mm_y = 2.0f * sensor_half_h / (float)fmt_h;
    const double ds_target = std::min((double)px_mm_x / (dx_fine * kSpriteOversample),
                                      (double)px_mm_y / (dx_fine * kSpriteOversample));
    const int ds = po2_floor(ds_target, N / 8);
    const int M = N / ds;
    const float* sp = bref.as<float>();
    if (ds > 1) {
        hipLaunchKernelGGL(downsample_kernel, dim3((unsigned)(((size_t)M * M + 255) / 256)), dim3(256), 0, c->stream,
                           bref.as<float>(), N, ds, 1, bds.as<float>(), M);
        GL_HIP(hipGetLastError());
        sp = bds.as<float>();
    }
    float *o_r, *o_g, *o_b;
    if ((st = out_planes(c, "veil_out", npx, &o_r, &o_g, &o_b))) return st;
    std::vector<StarSrc> ss(n_src);
    for (int i = 0; i < n_src; ++i) ss[i] = {src_px[i], src_py[i], sources[i].r, sources[i].g, sources[i].b};
    if ((st = resample(c, sp, M, 1, dx_fine * ds, dx_fine * ds, px_mm_x, px_mm_y, ss, p->veil_gain, o_r, o_g, o_b,
                       width, height)))
        return st;
    if ((st = download3(c, o_r, o_g, o_b, npx, out_r, out_g, out_b))) return st;
    if (dbg) {
        dbg->grid = N;
        dbg->dx_mm_x = dx_fine; dbg->dx_mm_y = dx_fine;
        dbg->core_mm = core_mm; dbg->falloff = fp; dbg->core_texels = a_texels;
        if (dbg->reference || dbg->sprite_rgb) {
            std::vector<float> h(NN);
            GL_HIP(d2h_sync(c, h.data(), bref.p, NN * 4));
            if (dbg->reference) std::copy(h.begin(), h.end(), dbg->reference);
            if (dbg->sprite_rgb)
                for (size_t k = 0; k < NN; ++k) { dbg->sprite_rgb[3 * k] = h[k]; dbg->sprite_rgb[3 * k + 1] = h[k]; dbg->sprite_rgb[3 * k + 2] = h[k]; }
        }
    }
    return GL_OK;
}

}  // extern "C"
