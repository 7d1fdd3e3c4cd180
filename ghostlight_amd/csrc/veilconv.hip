// veilconv.hip — veil glare spread applied to a whole image by FFT
// convolution (the extended-source form of the veil pass).
//
// out_c = gain * (in_c (*) g), with g the glare-spread function of
// veil_render.cu:94-106, g(r) = (a^2 / (r^2 + a^2))^p, sampled at integer
// pixel offsets (r in mm from the per-axis pixel pitch; a = veil_spread *
// sensor_half_h mm) and normalised to unit sum over its support.  The
// convolution is linear, not circular: the image is zero-padded to a P x P
// power of two with P >= width + half-support and height + half-support, and
// the kernel support is the whole P x P grid centred on the origin.
//
// g is real and even, so its transform G is real: two colour planes ride in
// one complex FFT (R + iG), are multiplied by G, and come back through the
// same forward FFT by the conjugation identity ifft(X) = conj(fft(conj(X))) / P^2.
// Per call: 2 forward + 2 inverse P x P transforms (LDS radix rows +
// LDS-tiled transposes) for three planes; G is cached per (P, a, p, pitch).
#include "gl_runtime.h"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace gl {

int fft2d_dev(gl_ctx* c, float2* data, float2* tmp, int N, float* psf);

__global__ __launch_bounds__(256) void gsf_kernel(float2* __restrict__ k, int P, float px_mm_x, float px_mm_y,
                                                  float a_mm, float p) {
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (size_t)P * P) return;
    const int x = (int)(idx % P), y = (int)(idx / P);
    const int dx = x < P / 2 ? x : x - P, dy = y < P / 2 ? y : y - P;
    const float rx = (float)dx * px_mm_x, ry = (float)dy * px_mm_y;
    const float a2 = a_mm * a_mm;
    k[idx] = make_float2(powf(a2 / (rx * rx + ry * ry + a2), p), 0.0f);
}

// Sum of the real parts (FP64 partials).
__global__ __launch_bounds__(256) void sum_re_kernel(const float2* __restrict__ k, size_t n, double* __restrict__ out) {
    __shared__ double sm[256];
    double acc = 0.0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        acc += (double)k[i].x;
    sm[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sm[threadIdx.x] += sm[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(out, sm[0]);
}

// G = Re(FFT(g)) / sum(g), stored as float (the imaginary part of an even
// kernel's transform is rounding noise).
__global__ __launch_bounds__(256) void gsf_finish_kernel(const float2* __restrict__ K, float* __restrict__ G, size_t n,
                                                         const double* __restrict__ sum) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double s = *sum;
    G[i] = s > 0.0 ? (float)((double)K[i].x / s) : 0.0f;
}

__global__ void zero_d_kernel(double* d) { *d = 0.0; }

// Pack two planes (b may be null) into the zero-padded complex buffer.
__global__ __launch_bounds__(256) void pack2_kernel(float2* __restrict__ z, int P, const float* __restrict__ a,
                                                    const float* __restrict__ b, int W, int H) {
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (size_t)P * P) return;
    const int x = (int)(idx % P), y = (int)(idx / P);
    float re = 0.f, im = 0.f;
    if (x < W && y < H) {
        const size_t o = (size_t)y * W + x;
        re = a[o];
        if (b) im = b[o];
    }
    z[idx] = make_float2(re, im);
}

// X <- conj(X * G): the multiply and the first conjugation of the inverse.
__global__ __launch_bounds__(256) void mul_conj_kernel(float2* __restrict__ z, const float* __restrict__ G, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float g = G[i];
    const float2 v = z[i];
    z[i] = make_float2(v.x * g, -(v.y * g));
}

// out_a = s * Re(conj(Z)) = s * Re(Z), out_b = s * Im(conj(Z)) = -s * Im(Z).
__global__ __launch_bounds__(256) void unpack2_kernel(const float2* __restrict__ z, int P, float s,
                                                      float* __restrict__ a, float* __restrict__ b, int W, int H) {
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (size_t)W * H) return;
    const int x = (int)(idx % W), y = (int)(idx / W);
    const float2 v = z[(size_t)y * P + x];
    a[idx] = s * v.x;
    if (b) b[idx] = -(s * v.y);
}

static int veil_conv_dev(gl_ctx* c, const float* in_r, const float* in_g, const float* in_b, int W, int H,
                         float px_mm_x, float px_mm_y, float a_mm, float fp, float gain, float* o_r, float* o_g,
                         float* o_b, int* P_out) {
    // Half-support large enough that the kernel tail is resolved across the
    // whole frame: P >= 2 * max(W, H).
    int P = 2;
    while (P < 2 * std::max(W, H)) P <<= 1;
    if (P > 16384) return set_error(GL_ERR_INVALID, "image %dx%d too large for the FFT veil (P = %d)", W, H, P);
    *P_out = P;
    const size_t PP = (size_t)P * P;
    int st;
    DevBuf &bk = c->buf("vconv_G"), &bz = c->buf("vconv_z"), &bt = c->buf("fft_tmp"), &bs = c->buf("vconv_sum");
    if ((st = bk.ensure(PP * 4)) || (st = bz.ensure(PP * 8)) || (st = bt.ensure(PP * 8)) || (st = bs.ensure(8))) return st;
    const unsigned blocks = (unsigned)((PP + 255) / 256);
    // Kernel transform, cached per geometry.
    std::vector<uint8_t> key(sizeof(int) + 4 * sizeof(float));
    std::memcpy(key.data(), &P, sizeof P);
    const float kv[4] = {px_mm_x, px_mm_y, a_mm, fp};
    std::memcpy(key.data() + sizeof P, kv, sizeof kv);
    if (c->vconv_key != key) {
        hipLaunchKernelGGL(gsf_kernel, dim3(blocks), dim3(256), 0, c->stream, bz.as<float2>(), P, px_mm_x, px_mm_y, a_mm,
                           fp);
        hipLaunchKernelGGL(zero_d_kernel, dim3(1), dim3(1), 0, c->stream, bs.as<double>());
        hipLaunchKernelGGL(sum_re_kernel, dim3(1024), dim3(256), 0, c->stream, bz.as<float2>(), PP, bs.as<double>());
        if ((st = fft2d_dev(c, bz.as<float2>(), bt.as<float2>(), P, nullptr))) return st;
        hipLaunchKernelGGL(gsf_finish_kernel, dim3(blocks), dim3(256), 0, c->stream, bz.as<float2>(), bk.as<float>(), PP,
                           bs.as<double>());
        GL_HIP(hipGetLastError());
        c->vconv_key = key;
    }
    const float s = gain / (float)PP;
    const float* ins[2][2] = {{in_r, in_g}, {in_b, nullptr}};
    float* outs[2][2] = {{o_r, o_g}, {o_b, nullptr}};
    for (int k = 0; k < 2; ++k) {
        hipLaunchKernelGGL(pack2_kernel, dim3(blocks), dim3(256), 0, c->stream, bz.as<float2>(), P, ins[k][0], ins[k][1],
                           W, H);
        if ((st = fft2d_dev(c, bz.as<float2>(), bt.as<float2>(), P, nullptr))) return st;
        hipLaunchKernelGGL(mul_conj_kernel, dim3(blocks), dim3(256), 0, c->stream, bz.as<float2>(), bk.as<float>(), PP);
        if ((st = fft2d_dev(c, bz.as<float2>(), bt.as<float2>(), P, nullptr))) return st;
        hipLaunchKernelGGL(unpack2_kernel, dim3((unsigned)(((size_t)W * H + 255) / 256)), dim3(256), 0, c->stream,
                           bz.as<float2>(), P, s, outs[k][0], outs[k][1], W, H);
        GL_HIP(hipGetLastError());
    }
    return GL_OK;
}

}  // namespace gl

using namespace gl;

extern "C" {

int gl_veil_convolve(gl_ctx* ctx, const float* in_r, const float* in_g, const float* in_b, int32_t width,
                     int32_t height, float sensor_half_w, float sensor_half_h, const gl_diffraction_params* p,
                     int32_t device_ptrs, float* out_r, float* out_g, float* out_b, float* ms_out) {
    if (!in_r || !in_g || !in_b || !out_r || !out_g || !out_b || !p) return set_error(GL_ERR_INVALID, "null argument");
    if (width <= 0 || height <= 0) return set_error(GL_ERR_INVALID, "degenerate image");
    if (!(sensor_half_w > 0.0f) || !(sensor_half_h > 0.0f)) return set_error(GL_ERR_INVALID, "sensor half extents must be positive");
    int st;
    gl_ctx* c = resolve_ctx(ctx, &st);
    if (!c) return st;
    std::lock_guard<std::mutex> g(c->mu);
    float spread = p->veil_spread;
    if (!(spread > 0.0f)) spread = 0.12f;
    spread = spread < 1e-3f ? 1e-3f : (spread > 4.0f ? 4.0f : spread);
    const float a_mm = spread * sensor_half_h;
    float fp = p->veil_falloff;
    if (!(fp >= 1.0f)) fp = 1.0f;
    if (fp > 3.0f) fp = 3.0f;
    const float px_mm_x = 2.0f * sensor_half_w / (float)width, px_mm_y = 2.0f * sensor_half_h / (float)height;
    const size_t npx = (size_t)width * height;
    const float *ir = in_r, *ig = in_g, *ib = in_b;
    float *o_r = out_r, *o_g = out_g, *o_b = out_b;
    if (!device_ptrs) {
        DevBuf &bi = c->buf("vconv_in"), &bo = c->buf("vconv_out");
        if ((st = bi.ensure(3 * npx * 4)) || (st = bo.ensure(3 * npx * 4))) return st;
        float* di = bi.as<float>();
        GL_HIP(hipMemcpyAsync(di, in_r, npx * 4, hipMemcpyHostToDevice, c->stream));
        GL_HIP(hipMemcpyAsync(di + npx, in_g, npx * 4, hipMemcpyHostToDevice, c->stream));
        GL_HIP(hipMemcpyAsync(di + 2 * npx, in_b, npx * 4, hipMemcpyHostToDevice, c->stream));
        ir = di; ig = di + npx; ib = di + 2 * npx;
        o_r = bo.as<float>(); o_g = o_r + npx; o_b = o_r + 2 * npx;
    }
    int P = 0;
    GL_HIP(hipEventRecord(c->ev0, c->stream));
    if ((st = veil_conv_dev(c, ir, ig, ib, width, height, px_mm_x, px_mm_y, a_mm, fp, p->veil_gain, o_r, o_g, o_b, &P)))
        return st;
    GL_HIP(hipEventRecord(c->ev1, c->stream));
    if (!device_ptrs) {
        GL_HIP(hipMemcpyAsync(out_r, o_r, npx * 4, hipMemcpyDeviceToHost, c->stream));
        GL_HIP(hipMemcpyAsync(out_g, o_g, npx * 4, hipMemcpyDeviceToHost, c->stream));
        GL_HIP(hipMemcpyAsync(out_b, o_b, npx * 4, hipMemcpyDeviceToHost, c->stream));
    }
    GL_HIP(hipStreamSynchronize(c->stream));
    if (ms_out) GL_HIP(hipEventElapsedTime(ms_out, c->ev0, c->ev1));
    return GL_OK;
}

}  // extern "C"
