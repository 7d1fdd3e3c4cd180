// mdft.hip — the matrix-DFT starburst on the gfx950 matrix cores.
//
// Complex GEMM engine.  A complex product is four real products; each runs on
// the MFMA units either as
//   FP32   v_mfma_f32_32x32x2_f32 (full-precision operands), or
//   3xBF16 v_mfma_f32_32x32x16_bf16 with every FP32 operand split as
//          x = hi + lo (hi = bf16(x), lo = bf16(x - hi)) and the product formed
//          as hi*hi + hi*lo + lo*hi (the lo*lo term, ~2^-16 relative, is
//          dropped), accumulated in FP32.
// Tiles: 64x64 per workgroup (4 waves, one 32x32 MFMA block each), K step 16,
// operands staged through LDS already split into bf16 hi/lo planes so each
// element is converted once per workgroup.  batch = gridDim.z.
//
// DFT operands are generated on the device from FP64-reduced phases (the
// integer lag product is reduced mod 1 turn in FP64 before the float
// sin/cos), so large-argument twiddles keep full accuracy.
//
// mdft_render follows mdft_starburst (mdft_render.cu:199-340): the
// autocorrelation C of the pupil from two FFTs, then per source and per
// wavelength  tmp = C . Px,  R = Re(Py . tmp),  window += cmf * k^2 * R; the
// wavelengths of one source run as one batched GEMM pair.
#include "gl_runtime.h"

#include <algorithm>
#include <cmath>
#include <cstdio>

namespace gl {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef short s16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

enum { EPI_COMPLEX = 0, EPI_REAL = 1, EPI_POWER = 2 };

__device__ __forceinline__ bf16x8 neg8(bf16x8 a) {
    s16x8 b = __builtin_bit_cast(s16x8, a);
    b ^= (short)0x8000;
    return __builtin_bit_cast(bf16x8, b);
}

// C = alpha * A . B.  A: M x K, B: K x N, row-major interleaved complex64 with
// leading dimensions in complex elements; per-batch strides sA/sB/sC (0 =
// shared operand).  Output per EPI: complex, real part (float), or |.|^2 (float).
template <int PREC, int EPI>
__global__ __launch_bounds__(256) void cgemm_kernel(const float2* __restrict__ A, int lda, long long sA,
                                                    const float2* __restrict__ B, int ldb, long long sB,
                                                    void* __restrict__ Cv, int ldc, long long sC, int M, int N, int K,
                                                    float alpha) {
    constexpr bool NEED_IM = (EPI != EPI_REAL);
    const int bz = blockIdx.z;
    A += (size_t)bz * sA;
    B += (size_t)bz * sB;
    const int m0 = blockIdx.y * 64, n0 = blockIdx.x * 64;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    f32x16 acc_r, acc_i;
    for (int v = 0; v < 16; ++v) { acc_r[v] = 0.0f; acc_i[v] = 0.0f; }

    // LDS: A planes [row][k], B planes transposed [col][k].
    //   PREC 1: bf16 planes re_hi, re_lo, im_hi, im_lo;  PREC 0: float re, im.
    __shared__ __attribute__((aligned(16))) char smem[PREC == 1 ? 2 * 4 * 64 * 16 * 2 : 2 * 2 * 64 * 17 * 4];
    __bf16(*sAh)[64][16] = reinterpret_cast<__bf16(*)[64][16]>(smem);
    __bf16(*sBh)[64][16] = reinterpret_cast<__bf16(*)[64][16]>(smem + 4 * 64 * 16 * 2);
    float(*sAf)[64][17] = reinterpret_cast<float(*)[64][17]>(smem);
    float(*sBf)[64][17] = reinterpret_cast<float(*)[64][17]>(smem + 2 * 64 * 17 * 4);

    const int a_row = tid >> 2, a_k = (tid & 3) * 4;     // A tile: 64 rows x 16 k
    const int b_k = tid >> 4, b_n = (tid & 15) * 4;      // B tile: 16 k x 64 cols
    for (int k0 = 0; k0 < K; k0 += 16) {
        {
            const int gr = m0 + a_row;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int gk = k0 + a_k + j;
                const float2 x = (gr < M && gk < K) ? A[(size_t)gr * lda + gk] : make_float2(0.f, 0.f);
                if (PREC == 1) {
                    const __bf16 rh = (__bf16)x.x, ih = (__bf16)x.y;
                    sAh[0][a_row][a_k + j] = rh;
                    sAh[1][a_row][a_k + j] = (__bf16)(x.x - (float)rh);
                    sAh[2][a_row][a_k + j] = ih;
                    sAh[3][a_row][a_k + j] = (__bf16)(x.y - (float)ih);
                } else {
                    sAf[0][a_row][a_k + j] = x.x;
                    sAf[1][a_row][a_k + j] = x.y;
                }
            }
            const int gk = k0 + b_k;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int gn = n0 + b_n + j;
                const float2 x = (gk < K && gn < N) ? B[(size_t)gk * ldb + gn] : make_float2(0.f, 0.f);
                if (PREC == 1) {
                    const __bf16 rh = (__bf16)x.x, ih = (__bf16)x.y;
                    sBh[0][b_n + j][b_k] = rh;
                    sBh[1][b_n + j][b_k] = (__bf16)(x.x - (float)rh);
                    sBh[2][b_n + j][b_k] = ih;
                    sBh[3][b_n + j][b_k] = (__bf16)(x.y - (float)ih);
                } else {
                    sBf[0][b_n + j][b_k] = x.x;
                    sBf[1][b_n + j][b_k] = x.y;
                }
            }
        }
        __syncthreads();
        const int r = lane & 31, kh = lane >> 5;
        if (PREC == 1) {
            // 16 k-slots per MFMA: lane half kh carries k = 8*kh .. 8*kh+7 for
            // both operands (the same k in the same slot).
            const bf16x8 arh = *reinterpret_cast<const bf16x8*>(&sAh[0][wm + r][8 * kh]);
            const bf16x8 arl = *reinterpret_cast<const bf16x8*>(&sAh[1][wm + r][8 * kh]);
            const bf16x8 aih = *reinterpret_cast<const bf16x8*>(&sAh[2][wm + r][8 * kh]);
            const bf16x8 ail = *reinterpret_cast<const bf16x8*>(&sAh[3][wm + r][8 * kh]);
            const bf16x8 brh = *reinterpret_cast<const bf16x8*>(&sBh[0][wn + r][8 * kh]);
            const bf16x8 brl = *reinterpret_cast<const bf16x8*>(&sBh[1][wn + r][8 * kh]);
            const bf16x8 bih = *reinterpret_cast<const bf16x8*>(&sBh[2][wn + r][8 * kh]);
            const bf16x8 bil = *reinterpret_cast<const bf16x8*>(&sBh[3][wn + r][8 * kh]);
            const bf16x8 nbih = neg8(bih), nbil = neg8(bil);
            // Re: Ar*Br - Ai*Bi
            acc_r = __builtin_amdgcn_mfma_f32_32x32x16_bf16(arh, brh, acc_r, 0, 0, 0);
            acc_r = __builtin_amdgcn_mfma_f32_32x32x16_bf16(arh, brl, acc_r, 0, 0, 0);
            acc_r = __builtin_amdgcn_mfma_f32_32x32x16_bf16(arl, brh, acc_r, 0, 0, 0);
            acc_r = __builtin_amdgcn_mfma_f32_32x32x16_bf16(aih, nbih, acc_r, 0, 0, 0);
            acc_r = __builtin_amdgcn_mfma_f32_32x32x16_bf16(aih, nbil, acc_r, 0, 0, 0);
            acc_r = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ail, nbih, acc_r, 0, 0, 0);
            if (NEED_IM) {
                // Im: Ar*Bi + Ai*Br
                acc_i = __builtin_amdgcn_mfma_f32_32x32x16_bf16(arh, bih, acc_i, 0, 0, 0);
                acc_i = __builtin_amdgcn_mfma_f32_32x32x16_bf16(arh, bil, acc_i, 0, 0, 0);
                acc_i = __builtin_amdgcn_mfma_f32_32x32x16_bf16(arl, bih, acc_i, 0, 0, 0);
                acc_i = __builtin_amdgcn_mfma_f32_32x32x16_bf16(aih, brh, acc_i, 0, 0, 0);
                acc_i = __builtin_amdgcn_mfma_f32_32x32x16_bf16(aih, brl, acc_i, 0, 0, 0);
                acc_i = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ail, brh, acc_i, 0, 0, 0);
            }
        } else {
#pragma unroll
            for (int kk = 0; kk < 16; kk += 2) {
                const float ar = sAf[0][wm + r][kk + kh], ai = sAf[1][wm + r][kk + kh];
                const float br = sBf[0][wn + r][kk + kh], bi = sBf[1][wn + r][kk + kh];
                acc_r = __builtin_amdgcn_mfma_f32_32x32x2f32(ar, br, acc_r, 0, 0, 0);
                acc_r = __builtin_amdgcn_mfma_f32_32x32x2f32(ai, -bi, acc_r, 0, 0, 0);
                if (NEED_IM) {
                    acc_i = __builtin_amdgcn_mfma_f32_32x32x2f32(ar, bi, acc_i, 0, 0, 0);
                    acc_i = __builtin_amdgcn_mfma_f32_32x32x2f32(ai, br, acc_i, 0, 0, 0);
                }
            }
        }
        __syncthreads();
    }
    // 32x32 accumulator layout: lane l holds column l % 32 and rows
    // 8*(v/4) + 4*(l/32) + v%4 for v = 0..15.
    const int col = n0 + wn + (lane & 31);
    if (col >= N) return;
#pragma unroll
    for (int v = 0; v < 16; ++v) {
        const int row = m0 + wm + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
        if (row >= M) continue;
        const size_t o = (size_t)bz * sC + (size_t)row * ldc + col;
        if (EPI == EPI_COMPLEX) {
            static_cast<float2*>(Cv)[o] = make_float2(alpha * acc_r[v], alpha * acc_i[v]);
        } else if (EPI == EPI_REAL) {
            static_cast<float*>(Cv)[o] = alpha * acc_r[v];
        } else {
            const float re = alpha * acc_r[v], im = alpha * acc_i[v];
            static_cast<float*>(Cv)[o] = re * re + im * im;
        }
    }
}

template <int PREC, int EPI>
static int cgemm_launch(hipStream_t s, const float2* A, int lda, long long sA, const float2* B, int ldb, long long sB,
                        void* C, int ldc, long long sC, int M, int N, int K, int batch, float alpha) {
    dim3 grid((unsigned)((N + 63) / 64), (unsigned)((M + 63) / 64), (unsigned)batch);
    hipLaunchKernelGGL((cgemm_kernel<PREC, EPI>), grid, dim3(256), 0, s, A, lda, sA, B, ldb, sB, C, ldc, sC, M, N, K,
                       alpha);
    GL_HIP(hipGetLastError());
    return GL_OK;
}

static int cgemm(hipStream_t s, int prec, int epi, const float2* A, int lda, long long sA, const float2* B, int ldb,
                 long long sB, void* C, int ldc, long long sC, int M, int N, int K, int batch, float alpha) {
    if (prec == GL_MDFT_BF16X3) {
        if (epi == EPI_COMPLEX) return cgemm_launch<1, EPI_COMPLEX>(s, A, lda, sA, B, ldb, sB, C, ldc, sC, M, N, K, batch, alpha);
        if (epi == EPI_REAL) return cgemm_launch<1, EPI_REAL>(s, A, lda, sA, B, ldb, sB, C, ldc, sC, M, N, K, batch, alpha);
        return cgemm_launch<1, EPI_POWER>(s, A, lda, sA, B, ldb, sB, C, ldc, sC, M, N, K, batch, alpha);
    }
    if (epi == EPI_COMPLEX) return cgemm_launch<0, EPI_COMPLEX>(s, A, lda, sA, B, ldb, sB, C, ldc, sC, M, N, K, batch, alpha);
    if (epi == EPI_REAL) return cgemm_launch<0, EPI_REAL>(s, A, lda, sA, B, ldb, sB, C, ldc, sC, M, N, K, batch, alpha);
    return cgemm_launch<0, EPI_POWER>(s, A, lda, sA, B, ldb, sB, C, ldc, sC, M, N, K, batch, alpha);
}

// exp(-2 pi i t) for a phase t in turns given as an exact-ish double, reduced
// mod 1 in FP64 before the float sin/cos.
__device__ __forceinline__ float2 cis_turns(double t) {
    t -= rint(t);
    float s, c;
    sincosf((float)(-2.0 * M_PI * t), &s, &c);
    return make_float2(c, s);
}

// MFT operand: mat[r][c] = exp(-2 pi i (a(r) b(c)) / (q N)), a = r - rows/2 or
// c - cols/2 by `swap` (so both A_y (k rows) and A_x^T (n rows) come from one
// kernel).  batch z: q = q_tab[z].
__global__ __launch_bounds__(256) void mft_matrix_kernel(float2* __restrict__ out, int rows, int cols, long long stride,
                                                         double half_r, double half_c, int Npupil,
                                                         const double* __restrict__ q_tab) {
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (size_t)rows * cols) return;
    const int r = (int)(idx / cols), c = (int)(idx % cols);
    const double inv = 1.0 / (q_tab[blockIdx.z] * (double)Npupil);
    out[(size_t)blockIdx.z * stride + idx] = cis_turns(((double)r - half_r) * ((double)c - half_c) * inv);
}

// MDFT axis factors (build_px_kernel / build_py_kernel, mdft_render.cu:109-135):
//   F[m][k] = sinc(m px_mm / S_l) exp(-2 pi i m ((k0 + k) + 0.5 - src) px_mm / S_l)
// written as Px[b][kx] (rows = lag) or, transposed, Py[ky][a].
__global__ __launch_bounds__(256) void mdft_factor_kernel(float2* __restrict__ out, int Mc, int Kn, int mc,
                                                          long long stride, const double* __restrict__ S_l, double px_mm,
                                                          int k0, double src, int lag_rows) {
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (size_t)Mc * Kn) return;
    int lag, k;
    if (lag_rows) { lag = (int)(idx / Kn); k = (int)(idx % Kn); }
    else { k = (int)(idx / Mc); lag = (int)(idx % Mc); }
    const double S = S_l[blockIdx.z];
    const double m = (double)(lag - mc);
    const double x = m * px_mm / S;
    double w = 1.0;
    if (fabs(x) >= 1e-7) w = sin(M_PI * x) / (M_PI * x);
    const double rel = ((double)(k0 + k) + 0.5 - src) * px_mm;
    const float2 e = cis_turns(m * rel / S);
    out[(size_t)blockIdx.z * stride + idx] = make_float2((float)w * e.x, (float)w * e.y);
}

__global__ void magsq_complex_kernel(float2* __restrict__ d, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float2 z = d[i];
    d[i] = make_float2(z.x * z.x + z.y * z.y, 0.0f);
}

// C[a][b] = conj(G[(a - mc) mod N][(b - mc) mod N]) (extract_C_kernel, mdft_render.cu:88-98).
__global__ void extract_c_kernel(float2* __restrict__ C, const float2* __restrict__ G, int N, int mc) {
    const int Mc = 2 * mc + 1;
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= Mc * Mc) return;
    const int b = idx % Mc, a = idx / Mc;
    const int gx = ((b - mc) % N + N) % N, gy = ((a - mc) % N + N) % N;
    const float2 g = G[(size_t)gy * N + gx];
    C[idx] = make_float2(g.x, -g.y);
}

__global__ void scale_complex_kernel(float2* __restrict__ C, int n, float k) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float2 z = C[i];
    C[i] = make_float2(z.x * k, z.y * k);
}

// window: out += gain * src * sum_l w_c(l) R_l  (the per-lambda accumulate of
// accumulate_y_kernel plus splat_window_kernel, mdft_render.cu:156-194).
__global__ __launch_bounds__(256) void mdft_combine_kernel(const float* __restrict__ R, int n_l, int Kx, int Ky,
                                                           const float4* __restrict__ w, int kx0, int ky0, int width,
                                                           int height, float gain_r, float gain_g, float gain_b,
                                                           float* __restrict__ o_r, float* __restrict__ o_g,
                                                           float* __restrict__ o_b) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= Kx * Ky) return;
    float sr = 0.f, sg = 0.f, sb = 0.f;
    for (int l = 0; l < n_l; ++l) {
        const float v = R[(size_t)l * Kx * Ky + idx];
        const float4 ww = w[l];
        sr += ww.y * v; sg += ww.z * v; sb += ww.w * v;
    }
    const int ix = kx0 + idx % Kx, iy = ky0 + idx / Kx;
    if (ix < 0 || ix >= width || iy < 0 || iy >= height) return;
    const size_t p = (size_t)iy * width + ix;
    o_r[p] += sr * gain_r;
    o_g[p] += sg * gain_g;
    o_b[p] += sb * gain_b;
}

int fft2d_dev(gl_ctx* c, float2* data, int N, float* psf);

int mdft_render(gl_ctx* c, const float2* d_pupil, int N, float fill_in, const std::vector<float4>& spec,
                float lambda_ref, float dx_x, float dx_y, float px_mm_x, float px_mm_y,
                const std::vector<StarSource>& src, float gain, float* o_r, float* o_g, float* o_b, int width,
                int height, int precision) {
    if (src.empty() || N <= 0) return GL_OK;
    const float Sx = (float)N * dx_x, Sy = (float)N * dx_y;
    const float fill = (fill_in > 0.02f && fill_in < 0.9f) ? fill_in : 0.30f;
    int mc = (int)std::ceil(fill * (double)N) + 4;
    if (mc > N / 2 - 1) mc = N / 2 - 1;
    const int Mc = 2 * mc + 1;
    const size_t NN = (size_t)N * N;
    int st;
    DevBuf &bw = c->buf("mdft_work"), &bc = c->buf("mdft_C");
    if ((st = bw.ensure(NN * 8)) || (st = bc.ensure((size_t)Mc * Mc * 8))) return st;
    GL_HIP(hipMemcpyAsync(bw.p, d_pupil, NN * 8, hipMemcpyDeviceToDevice, c->stream));
    if ((st = fft2d_dev(c, bw.as<float2>(), N, nullptr))) return st;
    hipLaunchKernelGGL(magsq_complex_kernel, dim3((unsigned)((NN + 255) / 256)), dim3(256), 0, c->stream, bw.as<float2>(), NN);
    if ((st = fft2d_dev(c, bw.as<float2>(), N, nullptr))) return st;
    hipLaunchKernelGGL(extract_c_kernel, dim3((unsigned)((Mc * Mc + 255) / 256)), dim3(256), 0, c->stream,
                       bc.as<float2>(), bw.as<float2>(), N, mc);
    float2 c00;
    GL_HIP(hipMemcpyAsync(&c00, bc.as<float2>() + (size_t)mc * Mc + mc, sizeof c00, hipMemcpyDeviceToHost, c->stream));
    GL_HIP(hipStreamSynchronize(c->stream));
    if (!(c00.x > 0.0f)) return GL_OK;   // empty pupil: no starburst
    // Normalise to the unit-total pattern: C /= (C00 N^2) (mdft_render.cu:254-262).
    const float cs = 1.0f / (c00.x * (float)N * (float)N);
    hipLaunchKernelGGL(scale_complex_kernel, dim3((unsigned)((Mc * Mc + 255) / 256)), dim3(256), 0, c->stream,
                       bc.as<float2>(), Mc * Mc, cs);
    GL_HIP(hipGetLastError());
    int half_x = (int)std::floor(0.5 * (double)Sx / (double)px_mm_x);
    int half_y = (int)std::floor(0.5 * (double)Sy / (double)px_mm_y);
    if (half_x < 1) half_x = 1;
    if (half_y < 1) half_y = 1;
    const float knorm_gain = (px_mm_x * px_mm_y) / (dx_x * dx_y) * gain;
    const int nl = (int)spec.size();
    std::vector<double> Sxl(nl), Syl(nl);
    std::vector<float4> w(nl);
    for (int l = 0; l < nl; ++l) {
        const float lam_scale = spec[l].x / lambda_ref;
        const float k2 = (lambda_ref / spec[l].x) * (lambda_ref / spec[l].x);
        Sxl[l] = (double)(Sx * lam_scale);
        Syl[l] = (double)(Sy * lam_scale);
        w[l] = make_float4(spec[l].x, spec[l].y * k2, spec[l].z * k2, spec[l].w * k2);
    }
    DevBuf &bsx = c->buf("mdft_Sx"), &bsy = c->buf("mdft_Sy"), &bww = c->buf("mdft_w");
    if ((st = bsx.ensure(nl * 8)) || (st = bsy.ensure(nl * 8)) || (st = bww.ensure(nl * 16))) return st;
    GL_HIP(hipMemcpyAsync(bsx.p, Sxl.data(), nl * 8, hipMemcpyHostToDevice, c->stream));
    GL_HIP(hipMemcpyAsync(bsy.p, Syl.data(), nl * 8, hipMemcpyHostToDevice, c->stream));
    GL_HIP(hipMemcpyAsync(bww.p, w.data(), nl * 16, hipMemcpyHostToDevice, c->stream));
    for (const StarSource& s : src) {
        const int kx0 = std::max(0, (int)std::floor(s.px) - half_x);
        const int kx1 = std::min(width, (int)std::floor(s.px) + half_x + 1);
        const int ky0 = std::max(0, (int)std::floor(s.py) - half_y);
        const int ky1 = std::min(height, (int)std::floor(s.py) + half_y + 1);
        const int Kx = kx1 - kx0, Ky = ky1 - ky0;
        if (Kx <= 0 || Ky <= 0) continue;
        DevBuf &bpx = c->buf("mdft_Px"), &bpy = c->buf("mdft_Py"), &btm = c->buf("mdft_tmp"), &bR = c->buf("mdft_R");
        if ((st = bpx.ensure((size_t)nl * Mc * Kx * 8)) || (st = bpy.ensure((size_t)nl * Ky * Mc * 8)) ||
            (st = btm.ensure((size_t)nl * Mc * Kx * 8)) || (st = bR.ensure((size_t)nl * Kx * Ky * 4)))
            return st;
        hipLaunchKernelGGL(mdft_factor_kernel, dim3((unsigned)(((size_t)Mc * Kx + 255) / 256), 1, (unsigned)nl), dim3(256),
                           0, c->stream, bpx.as<float2>(), Mc, Kx, mc, (long long)Mc * Kx, bsx.as<double>(),
                           (double)px_mm_x, kx0, (double)s.px, 1);
        hipLaunchKernelGGL(mdft_factor_kernel, dim3((unsigned)(((size_t)Mc * Ky + 255) / 256), 1, (unsigned)nl), dim3(256),
                           0, c->stream, bpy.as<float2>(), Mc, Ky, mc, (long long)Mc * Ky, bsy.as<double>(),
                           (double)px_mm_y, ky0, (double)s.py, 0);
        GL_HIP(hipGetLastError());
        // tmp_l = C . Px_l   (Mc x Mc) . (Mc x Kx)
        if ((st = cgemm(c->stream, precision, EPI_COMPLEX, bc.as<float2>(), Mc, 0, bpx.as<float2>(), Kx,
                        (long long)Mc * Kx, btm.p, Kx, (long long)Mc * Kx, Mc, Kx, Mc, nl, 1.0f)))
            return st;
        // R_l = Re(Py_l . tmp_l)   (Ky x Mc) . (Mc x Kx)
        if ((st = cgemm(c->stream, precision, EPI_REAL, bpy.as<float2>(), Mc, (long long)Ky * Mc, btm.as<float2>(), Kx,
                        (long long)Mc * Kx, bR.p, Kx, (long long)Kx * Ky, Ky, Kx, Mc, nl, 1.0f)))
            return st;
        hipLaunchKernelGGL(mdft_combine_kernel, dim3((unsigned)((Kx * Ky + 255) / 256)), dim3(256), 0, c->stream,
                           bR.as<float>(), nl, Kx, Ky, bww.as<float4>(), kx0, ky0, width, height, knorm_gain * s.r,
                           knorm_gain * s.g, knorm_gain * s.b, o_r, o_g, o_b);
        GL_HIP(hipGetLastError());
    }
    return GL_OK;
}

}  // namespace gl

using namespace gl;

// The pre-split 3xBF16 kernel applies to tile-aligned shapes on a device
// whose workgroups may hold its 80 KB of LDS.
static bool nt3_ok(int M, int N, int K) {
    if (!cgemm_nt3_fits(M, N, K)) return false;
    static int lds_max = -1;
    if (lds_max < 0) {
        int dev = 0, v = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
            v = 0;
        lds_max = v;
    }
    return (size_t)lds_max >= cgemm_nt3_lds_bytes();
}

extern "C" {

int gl_cgemm_host(gl_ctx* ctx, int32_t M, int32_t N, int32_t K, const float* A, const float* B, float* C, int32_t prec) {
    if (M <= 0 || N <= 0 || K <= 0 || !A || !B || !C) return set_error(GL_ERR_INVALID, "bad GEMM arguments");
    if (prec != GL_MDFT_FP32 && prec != GL_MDFT_BF16X3) return set_error(GL_ERR_INVALID, "precision %d", prec);
    int st;
    gl_ctx* c = resolve_ctx(ctx, &st);
    if (!c) return st;
    std::lock_guard<std::mutex> g(c->mu);
    DevBuf &ba = c->buf("gemm_A"), &bb = c->buf("gemm_B"), &bcc = c->buf("gemm_C");
    if ((st = ba.ensure((size_t)M * K * 8)) || (st = bb.ensure((size_t)K * N * 8)) || (st = bcc.ensure((size_t)M * N * 8)))
        return st;
    GL_HIP(hipMemcpyAsync(ba.p, A, (size_t)M * K * 8, hipMemcpyHostToDevice, c->stream));
    GL_HIP(hipMemcpyAsync(bb.p, B, (size_t)K * N * 8, hipMemcpyHostToDevice, c->stream));
    if (prec == GL_MDFT_BF16X3 && nt3_ok(M, N, K)) {
        // Pre-split operands: A as is, B transposed to N x K.
        DevBuf &bas = c->buf("gemm_As"), &bbs = c->buf("gemm_Bs");
        if ((st = bas.ensure((size_t)M * K * 8)) || (st = bbs.ensure((size_t)K * N * 8))) return st;
        if ((st = split_planes(c->stream, ba.as<float2>(), K, 0, M, K, 0, 1, bas.as<__bf16>())) ||
            (st = split_planes(c->stream, bb.as<float2>(), N, 0, K, N, 1, 1, bbs.as<__bf16>())) ||
            (st = cgemm_nt3(c->stream, 0, bas.as<__bf16>(), 0, bbs.as<__bf16>(), 0, bcc.p, N, 0, M, N, K, 1, 1.0f)))
            return st;
    } else if ((st = cgemm(c->stream, prec, EPI_COMPLEX, ba.as<float2>(), K, 0, bb.as<float2>(), N, 0, bcc.p, N, 0, M,
                           N, K, 1, 1.0f))) {
        return st;
    }
    GL_HIP(hipMemcpyAsync(C, bcc.p, (size_t)M * N * 8, hipMemcpyDeviceToHost, c->stream));
    GL_HIP(hipStreamSynchronize(c->stream));
    return GL_OK;
}

int gl_mft_psf(gl_ctx* ctx, const float* pupil, int32_t N, int32_t M, float q_ref, float lambda_ref,
               const float* lambdas, int32_t nl, int32_t prec, float* out_psf, float** dev_out, float* ms_out) {
    if (!pupil || N <= 0 || M <= 0 || nl <= 0 || !lambdas || !(q_ref > 0.0f) || !(lambda_ref > 0.0f))
        return set_error(GL_ERR_INVALID, "bad MFT arguments");
    if (prec != GL_MDFT_FP32 && prec != GL_MDFT_BF16X3) return set_error(GL_ERR_INVALID, "precision %d", prec);
    int st;
    gl_ctx* c = resolve_ctx(ctx, &st);
    if (!c) return st;
    std::lock_guard<std::mutex> g(c->mu);
    const size_t NN = (size_t)N * N, MN = (size_t)M * N, MM = (size_t)M * M;
    DevBuf &bp = c->buf("mft_pupil"), &bax = c->buf("mft_AxT"), &bay = c->buf("mft_Ay"), &btt = c->buf("mft_T");
    DevBuf &bo = c->buf("mft_psf"), &bq = c->buf("mft_q");
    if ((st = bp.ensure(NN * 8)) || (st = bax.ensure((size_t)nl * MN * 8)) || (st = bay.ensure((size_t)nl * MN * 8)) ||
        (st = btt.ensure((size_t)nl * MN * 8)) || (st = bo.ensure((size_t)nl * MM * 4)) || (st = bq.ensure(nl * 8)))
        return st;
    std::vector<double> q(nl);
    for (int l = 0; l < nl; ++l) q[l] = (double)q_ref * (double)lambdas[l] / (double)lambda_ref;
    GL_HIP(hipMemcpyAsync(bp.p, pupil, NN * 8, hipMemcpyHostToDevice, c->stream));
    GL_HIP(hipMemcpyAsync(bq.p, q.data(), nl * 8, hipMemcpyHostToDevice, c->stream));
    const unsigned gb = (unsigned)((MN + 255) / 256);
    // A_x^T: N rows (pupil index n) x M cols (field index k);  A_y: M x N.
    hipLaunchKernelGGL(mft_matrix_kernel, dim3(gb, 1, (unsigned)nl), dim3(256), 0, c->stream, bax.as<float2>(), N, M,
                       (long long)MN, (double)N / 2.0, (double)M / 2.0, N, bq.as<double>());
    hipLaunchKernelGGL(mft_matrix_kernel, dim3(gb, 1, (unsigned)nl), dim3(256), 0, c->stream, bay.as<float2>(), M, N,
                       (long long)MN, (double)M / 2.0, (double)N / 2.0, N, bq.as<double>());
    GL_HIP(hipGetLastError());
    if (prec == GL_MDFT_BF16X3 && nt3_ok(M, N, N) && nt3_ok(M, M, N)) {
        // Split once: P (shared), A_y per lambda.  T_l^T = A_y P^T (M x N)
        // comes out split, ready to be the B^T operand of PSF_l = |A_y T_l|^2.
        DevBuf &bps = c->buf("mft_Ps"), &bys = c->buf("mft_Ays"), &bts = c->buf("mft_Ts");
        if ((st = bps.ensure(NN * 8)) || (st = bys.ensure((size_t)nl * MN * 8)) || (st = bts.ensure((size_t)nl * MN * 8)))
            return st;
        GL_HIP(hipEventRecord(c->ev0, c->stream));
        if ((st = split_planes(c->stream, bp.as<float2>(), N, 0, N, N, 0, 1, bps.as<__bf16>())) ||
            (st = split_planes(c->stream, bay.as<float2>(), N, (long long)MN, M, N, 0, nl, bys.as<__bf16>())) ||
            (st = cgemm_nt3(c->stream, 2, bys.as<__bf16>(), 4 * (long long)MN, bps.as<__bf16>(), 0, bts.p, N,
                            4 * (long long)MN, M, N, N, nl, 1.0f)) ||
            (st = cgemm_nt3(c->stream, 1, bys.as<__bf16>(), 4 * (long long)MN, bts.as<__bf16>(), 4 * (long long)MN,
                            bo.p, M, (long long)MM, M, M, N, nl, 1.0f)))
            return st;
        GL_HIP(hipEventRecord(c->ev1, c->stream));
    } else {
        GL_HIP(hipEventRecord(c->ev0, c->stream));
        // T_l = P . A_x^T (N x M);  PSF_l = |A_y . T_l|^2 (M x M)
        if ((st = cgemm(c->stream, prec, EPI_COMPLEX, bp.as<float2>(), N, 0, bax.as<float2>(), M, (long long)MN, btt.p,
                        M, (long long)MN, N, M, N, nl, 1.0f)))
            return st;
        if ((st = cgemm(c->stream, prec, EPI_POWER, bay.as<float2>(), N, (long long)MN, btt.as<float2>(), M,
                        (long long)MN, bo.p, M, (long long)MM, M, M, N, nl, 1.0f)))
            return st;
        GL_HIP(hipEventRecord(c->ev1, c->stream));
    }
    if (out_psf) GL_HIP(hipMemcpyAsync(out_psf, bo.p, (size_t)nl * MM * 4, hipMemcpyDeviceToHost, c->stream));
    GL_HIP(hipStreamSynchronize(c->stream));
    if (ms_out) GL_HIP(hipEventElapsedTime(ms_out, c->ev0, c->ev1));
    if (dev_out) *dev_out = bo.as<float>();
    return GL_OK;
}

}  // extern "C"
