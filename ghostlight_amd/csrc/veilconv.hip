// veilconv.hip — veil glare spread applied to a whole image by FFT
// convolution (the extended-source form of the veil pass).
//
// out_c = gain * (in_c (*) g), with g the glare-spread function of
// veil_render.cu:94-106, g(r) = (a^2 / (r^2 + a^2))^p, sampled at integer
// pixel offsets (r in mm from the per-axis pixel pitch; a = veil_spread *
// sensor_half_h mm) and normalised to unit sum over its support.  The
// convolution is linear, not circular: the image is zero-padded to a P x P
// power of two with P >= 2 max(width, height), and the kernel support is the
// whole P x P grid centred on the origin.
//
// g is real and even, so its transform G is real: two colour planes ride in
// one complex transform (R + iG), are multiplied by G, and come back through
// the same forward transform by the conjugation identity
// ifft(X) = conj(fft(conj(X))) / P^2.  One plane pair costs six passes:
//
//   1. row FFT of the H image rows, the zero-padded pack fused in  -> A (H x P)
//   2. transpose A -> B (P x P); rows >= H of the padded image are zero
//   3. row FFT of B (the column transform), epilogue conj(X * G^T)
//   4. row FFT of B (first half of the inverse, along y)
//   5. transpose the first H columns of B -> A
//   6. row FFT of A's H rows (along x), epilogue writes the W output
//      pixels of each row, scaled, into the two output planes
//
// Every pass streams its array once, and the two row passes over the image
// rows touch H / P of the grid.  G^T (the transform left transposed, as
// pass 3 wants it) is cached per (P, a, p, pitch) on the context.
#include "gl_runtime.h"
#include "gl_fft.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>

namespace gl {

int fft2d_dev(gl_ctx* c, float2* data, int N, float* psf);
bool fft_rows_fit(gl_ctx* c, int N);

namespace {

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

// Sum of the real parts (FP64 partials, one atomic per workgroup).
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

// ---- fused row policies of the six-pass pipeline ----------------------------

// Pass 1: reads image row y (x < W; b may be null), zero beyond W.
struct IOPackRows {
    const float* a;
    const float* b;
    int w;
    float2* out;   // pitch p
    int p;
    __device__ float2 load(int row, int i) const {
        if (i >= w) return make_float2(0.0f, 0.0f);
        const size_t o = (size_t)row * w + i;
        return make_float2(a[o], b ? b[o] : 0.0f);
    }
    __device__ void store(int row, int i, float2 v) const { out[(size_t)row * p + i] = v; }
};

// Pass 3: in place, epilogue X <- conj(X * Gt) (the multiply and the first
// conjugation of the inverse).
struct IOMulConj {
    float2* d;
    const float* g;
    int p;
    __device__ float2 load(int row, int i) const { return d[(size_t)row * p + i]; }
    __device__ void store(int row, int i, float2 v) const {
        const float s = g[(size_t)row * p + i];
        d[(size_t)row * p + i] = make_float2(v.x * s, -(v.y * s));
    }
};

// Pass 6: out_a = s Re(conj(Z)) = s Re(Z), out_b = s Im(conj(Z)) = -s Im(Z)
// for the W pixels of each image row.
struct IOUnpackRows {
    const float2* d;
    int p;
    float s;
    float* a;
    float* b;
    int w;
    __device__ float2 load(int row, int i) const { return d[(size_t)row * p + i]; }
    __device__ void store(int row, int i, float2 v) const {
        if (i >= w) return;
        const size_t o = (size_t)row * w + i;
        a[o] = s * v.x;
        if (b) b[o] = -(s * v.y);
    }
};

// Fallback element-wise forms for grids whose rows exceed one workgroup's LDS.
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

// Natural-layout multiply: z[ky][kx] <- conj(z * Gt[kx][ky]).
__global__ __launch_bounds__(256) void mul_conj_t_kernel(float2* __restrict__ z, const float* __restrict__ Gt, int P) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)P * P) return;
    const int x = (int)(i % P), y = (int)(i / P);
    const float g = Gt[(size_t)x * P + y];
    const float2 v = z[i];
    z[i] = make_float2(v.x * g, -(v.y * g));
}

__global__ __launch_bounds__(256) void unpack2_kernel(const float2* __restrict__ z, int P, float s,
                                                      float* __restrict__ a, float* __restrict__ b, int W, int H) {
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (size_t)W * H) return;
    const int x = (int)(idx % W), y = (int)(idx / W);
    const float2 v = z[(size_t)y * P + x];
    a[idx] = s * v.x;
    if (b) b[idx] = -(s * v.y);
}

}  // namespace

// Work arrays of one convolution: A >= H x P complex, B = P x P complex,
// Gt = P x P float, sum = one double.
struct VeilWork {
    float2* A;
    float2* B;
    float* Gt;
    double* sum;
};

static int ilog2i(int n) { int l = 0; while ((1 << l) < n) ++l; return l; }

static int veil_pad(int W, int H) {
    int P = 2;
    while (P < 2 * std::max(W, H)) P <<= 1;
    return P;
}

// G^T into w.Gt: g on the grid (B), unit-sum scale, rows / transpose / rows,
// i.e. the 2-D transform left transposed.
static int veil_kernel_transform(gl_ctx* c, hipStream_t s, int P, float px_mm_x, float px_mm_y, float a_mm, float fp,
                                 const VeilWork& w) {
    const size_t PP = (size_t)P * P;
    const unsigned blocks = (unsigned)((PP + 255) / 256);
    hipLaunchKernelGGL(gsf_kernel, dim3(blocks), dim3(256), 0, s, w.B, P, px_mm_x, px_mm_y, a_mm, fp);
    hipLaunchKernelGGL(zero_d_kernel, dim3(1), dim3(1), 0, s, w.sum);
    hipLaunchKernelGGL(sum_re_kernel, dim3(1024), dim3(256), 0, s, w.B, PP, w.sum);
    GL_HIP(hipGetLastError());
    if (c && !fft_rows_fit(c, P)) {
        int st = fft2d_dev(c, w.B, P, nullptr);
        if (st) return st;
        GL_HIP(fft::transpose_square_inplace(s, w.B, P));
    } else {
        const int L = ilog2i(P);
        GL_HIP(fft::rows(s, L, fft::RowsInPlace{w.B, P}, P));
        GL_HIP(fft::transpose_square_inplace(s, w.B, P));
        GL_HIP(fft::rows(s, L, fft::RowsInPlace{w.B, P}, P));
    }
    hipLaunchKernelGGL(gsf_finish_kernel, dim3(blocks), dim3(256), 0, s, w.B, w.Gt, PP, w.sum);
    GL_HIP(hipGetLastError());
    return GL_OK;
}

// The six passes for one plane pair (b / ob may be null).
static int veil_pair(hipStream_t s, int P, const float* a, const float* b, int W, int H, float scale, float* oa,
                     float* ob, const VeilWork& w) {
    const int L = ilog2i(P);
    GL_HIP(fft::rows(s, L, IOPackRows{a, b, W, w.A, P}, H));
    GL_HIP(fft::transpose_limited(s, w.A, H, w.B, P, P));
    GL_HIP(fft::rows(s, L, IOMulConj{w.B, w.Gt, P}, P));
    GL_HIP(fft::rows(s, L, fft::RowsInPlace{w.B, P}, P));
    GL_HIP(fft::transpose_limited(s, w.B, P, w.A, H, P));
    GL_HIP(fft::rows(s, L, IOUnpackRows{w.A, P, scale, oa, ob, W}, H));
    return GL_OK;
}

// Grids too long for the LDS row transform: generic 2-D transforms on B.
static int veil_pair_generic(gl_ctx* c, int P, const float* a, const float* b, int W, int H, float scale, float* oa,
                             float* ob, const VeilWork& w) {
    const size_t PP = (size_t)P * P;
    const unsigned blocks = (unsigned)((PP + 255) / 256);
    hipLaunchKernelGGL(pack2_kernel, dim3(blocks), dim3(256), 0, c->stream, w.B, P, a, b, W, H);
    int st;
    if ((st = fft2d_dev(c, w.B, P, nullptr))) return st;
    hipLaunchKernelGGL(mul_conj_t_kernel, dim3(blocks), dim3(256), 0, c->stream, w.B, w.Gt, P);
    if ((st = fft2d_dev(c, w.B, P, nullptr))) return st;
    hipLaunchKernelGGL(unpack2_kernel, dim3((unsigned)(((size_t)W * H + 255) / 256)), dim3(256), 0, c->stream, w.B, P,
                       scale, oa, ob, W, H);
    GL_HIP(hipGetLastError());
    return GL_OK;
}

struct VeilGeom {
    float a_mm, fp, px_mm_x, px_mm_y;
};

static VeilGeom veil_geom(const gl_diffraction_params* p, int W, int H, float shw, float shh) {
    float spread = p->veil_spread;
    if (!(spread > 0.0f)) spread = 0.12f;
    spread = spread < 1e-3f ? 1e-3f : (spread > 4.0f ? 4.0f : spread);
    float fp = p->veil_falloff;
    if (!(fp >= 1.0f)) fp = 1.0f;
    if (fp > 3.0f) fp = 3.0f;
    return {spread * shh, fp, 2.0f * shw / (float)W, 2.0f * shh / (float)H};
}

static int veil_conv_dev(gl_ctx* c, const float* in_r, const float* in_g, const float* in_b, int W, int H,
                         const VeilGeom& g, float gain, float* o_r, float* o_g, float* o_b) {
    const int P = veil_pad(W, H);
    if (P > 16384) return set_error(GL_ERR_INVALID, "image %dx%d too large for the FFT veil (P = %d)", W, H, P);
    const size_t PP = (size_t)P * P;
    int st;
    DevBuf &bk = c->buf("vconv_G"), &ba = c->buf("vconv_a"), &bb = c->buf("vconv_b"), &bs = c->buf("vconv_sum");
    if ((st = bk.ensure(PP * 4)) || (st = ba.ensure((size_t)H * P * 8)) || (st = bb.ensure(PP * 8)) ||
        (st = bs.ensure(8)))
        return st;
    const VeilWork w{ba.as<float2>(), bb.as<float2>(), bk.as<float>(), bs.as<double>()};
    std::vector<uint8_t> key(sizeof(int) + 4 * sizeof(float));
    std::memcpy(key.data(), &P, sizeof P);
    const float kv[4] = {g.px_mm_x, g.px_mm_y, g.a_mm, g.fp};
    std::memcpy(key.data() + sizeof P, kv, sizeof kv);
    if (c->vconv_key != key) {
        if ((st = veil_kernel_transform(c, c->stream, P, g.px_mm_x, g.px_mm_y, g.a_mm, g.fp, w))) return st;
        c->vconv_key = key;
    }
    const float s = gain / (float)PP;
    const bool fused = fft_rows_fit(c, P);
    const float* ins[2][2] = {{in_r, in_g}, {in_b, nullptr}};
    float* outs[2][2] = {{o_r, o_g}, {o_b, nullptr}};
    for (int k = 0; k < 2; ++k) {
        st = fused ? veil_pair(c->stream, P, ins[k][0], ins[k][1], W, H, s, outs[k][0], outs[k][1], w)
                   : veil_pair_generic(c, P, ins[k][0], ins[k][1], W, H, s, outs[k][0], outs[k][1], w);
        if (st) return st;
    }
    return GL_OK;
}

// ---- the entry point that allocates nothing ---------------------------------
//
// Work arrays in one file-scope device array, sized for grids up to
// kNoAllocMaxP (images up to kNoAllocMaxP / 2 on their longer side):
//   B  kNoAllocMaxP^2 complex            32 MiB
//   A  kNoAllocMaxP^2 / 2 complex        16 MiB   (H <= P / 2 rows)
//   Gt kNoAllocMaxP^2 float              16 MiB
//   sum                                  256 B
// Each device the library runs on holds its own copy until the process ends.
// Its device address is looked up once per device (hipGetSymbolAddress, with
// that device current) and only that address is handed to kernels.
static constexpr int kNoAllocMaxP = 2048;
static constexpr size_t kNoAllocB = (size_t)kNoAllocMaxP * kNoAllocMaxP;   // float2 units
static constexpr size_t kNoAllocA = kNoAllocB / 2;
static constexpr size_t kNoAllocG = kNoAllocB / 2;                          // P^2 floats
static constexpr size_t kNoAllocF2 = kNoAllocB + kNoAllocA + kNoAllocG + 32;
static __device__ __align__(256) float2 scratch[kNoAllocF2];

static constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static float2* g_scratch_ptr[kMaxDevices] = {};

static int scratch_ptr(int device, float2** out) {
    if (device < 0 || device >= kMaxDevices) return set_error(GL_ERR_INVALID, "device %d out of range", device);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        int prev = 0;
        GL_HIP(hipGetDevice(&prev));
        if (prev != device) GL_HIP(hipSetDevice(device));
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (prev != device) GL_HIP(hipSetDevice(prev));
        if (e != hipSuccess || !p)
            return set_error(GL_ERR_DEVICE, "hipGetSymbolAddress(scratch) -- %s", hipGetErrorString(e));
        g_scratch_ptr[device] = static_cast<float2*>(p);
    }
    *out = g_scratch_ptr[device];
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
    const VeilGeom geom = veil_geom(p, width, height, sensor_half_w, sensor_half_h);
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
    GL_HIP(hipEventRecord(c->ev0, c->stream));
    if ((st = veil_conv_dev(c, ir, ig, ib, width, height, geom, p->veil_gain, o_r, o_g, o_b))) return st;
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

int gl_veil_convolve_noalloc(gl_ctx* ctx, const float* in_r, const float* in_g, const float* in_b, int32_t width,
                             int32_t height, float sensor_half_w, float sensor_half_h,
                             const gl_diffraction_params* p, float* out_r, float* out_g, float* out_b) {
    if (!ctx) return set_error(GL_ERR_INVALID, "gl_veil_convolve_noalloc needs a context created beforehand");
    if (!in_r || !in_g || !in_b || !out_r || !out_g || !out_b || !p) return set_error(GL_ERR_INVALID, "null argument");
    if (width <= 0 || height <= 0) return set_error(GL_ERR_INVALID, "degenerate image");
    if (!(sensor_half_w > 0.0f) || !(sensor_half_h > 0.0f)) return set_error(GL_ERR_INVALID, "sensor half extents must be positive");
    const int P = veil_pad(width, height);
    if (P > kNoAllocMaxP)
        return set_error(GL_ERR_INVALID, "image %dx%d needs a %d^2 grid; gl_veil_convolve_noalloc holds at most %d^2 "
                         "(images up to %d on the longer side)", width, height, P, kNoAllocMaxP, kNoAllocMaxP, kNoAllocMaxP / 2);
    std::lock_guard<std::mutex> g(ctx->mu);
    float2* base = nullptr;
    int st = scratch_ptr(ctx->device, &base);
    if (st) return st;
    // Carve the arrays at this call's pitch P (A holds height <= P / 2 rows).
    const size_t PP = (size_t)P * P;
    VeilWork w;
    w.B = base;
    w.A = base + PP;
    w.Gt = reinterpret_cast<float*>(base + PP + PP / 2);
    w.sum = reinterpret_cast<double*>(base + PP + PP / 2 + PP / 2);
    const VeilGeom geom = veil_geom(p, width, height, sensor_half_w, sensor_half_h);
    hipStream_t s = ctx->stream;
    // The kernel transform is rebuilt on every call: the array keeps whatever
    // the last call on this device left in it.
    if ((st = veil_kernel_transform(nullptr, s, P, geom.px_mm_x, geom.px_mm_y, geom.a_mm, geom.fp, w))) return st;
    const float scale = p->veil_gain / (float)PP;
    if ((st = veil_pair(s, P, in_r, in_g, width, height, scale, out_r, out_g, w))) return st;
    return veil_pair(s, P, in_b, nullptr, width, height, scale, out_b, nullptr, w);
}

}  // extern "C"
