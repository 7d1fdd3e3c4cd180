// cgemm3.hip — complex GEMM in 3xBF16 on pre-split operands, for large
// aligned shapes (the MFT of the precision starburst: 2048 x 1024 x 1024 and
// 2048 x 2048 x 1024 per wavelength).
//
// Operands are split ONCE into four bf16 planes per matrix -- re_hi, re_lo,
// im_hi, im_lo with hi = bf16(x), lo = bf16(x - hi) -- in "k-contiguous"
// layout: A as M x K rows, B as its transpose N x K (the NT form), so one
// 16-byte load is exactly one lane's 8-element MFMA fragment.  The product is
// hi*hi + hi*lo + lo*hi per real product (the lo*lo term, ~2^-16 relative, is
// dropped) on v_mfma_f32_32x32x16_bf16 with FP32 accumulation: the same
// arithmetic as the split-on-load kernel in mdft.hip, without re-splitting
// every element once per output tile.
//
// Tile: 128 x 128 per workgroup, 8 waves each owning a 64 x 32 block (2 x 1
// MFMA blocks of 32 x 32, re and im accumulators: 64 accumulator registers),
// K step 32.  Two LDS stages of 80 KB (rows padded by 16 bytes against bank
// conflicts): while a step's 96 MFMAs per wave run out of one stage, the next
// step's operands go from registers into the other and the step after that is
// fetched from HBM.  One workgroup per CU, two waves per SIMD, so one wave's
// LDS reads, stash and barrier overlap with the other's MFMAs.
// Epilogues: complex float2, |z|^2 float, or the result split again into bf16
// planes (so a GEMM can feed the next one as its B operand).
#include "gl_runtime.h"

namespace gl {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef short s16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x16_t __attribute__((ext_vector_type(16)));
typedef unsigned u32x4_t __attribute__((ext_vector_type(4)));

namespace {

constexpr int kT = 128;         // tile edge
constexpr int kBK = 32;         // K step
constexpr int kRow = kBK + 8;   // padded LDS row (bf16 elements)

__device__ __forceinline__ bf16x8_t neg_bf16x8(bf16x8_t a) {
    s16x8_t b = __builtin_bit_cast(s16x8_t, a);
    b ^= (short)0x8000;
    return __builtin_bit_cast(bf16x8_t, b);
}

// Split a batch of complex matrices X (R x C, leading dimension ldx, batch
// stride sX complex elements) into 4 bf16 planes, optionally transposed:
// out[b][p][r][c] (or [p][c][r]), plane size R*C, batch stride 4*R*C.
__global__ __launch_bounds__(256) void split_planes_kernel(const float2* __restrict__ X, int ldx, long long sX, int R,
                                                           int C, int transpose, __bf16* __restrict__ out) {
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t RC = (size_t)R * C;
    if (idx >= RC) return;
    const int r = (int)(idx / C), c = (int)(idx % C);
    const float2 x = X[(size_t)blockIdx.z * sX + (size_t)r * ldx + c];
    const __bf16 rh = (__bf16)x.x, ih = (__bf16)x.y;
    const size_t o = transpose ? (size_t)c * R + r : idx;
    __bf16* base = out + (size_t)blockIdx.z * 4 * RC;
    base[o] = rh;
    base[RC + o] = (__bf16)(x.x - (float)rh);
    base[2 * RC + o] = ih;
    base[3 * RC + o] = (__bf16)(x.y - (float)ih);
}

// C = alpha * A . B with A (M x K) and Bt = B^T (N x K) as 4-plane bf16
// splits (plane strides M*K and N*K, batch strides sA / sB elements, 0 =
// shared).  EPI 0: float2 C, 1: float |z|^2, 2: 4-plane bf16 split of C
// (row-major M x N, batch stride sC elements of the plane set).
template <int EPI>
__global__ __launch_bounds__(512) void cgemm_nt3_kernel(const __bf16* __restrict__ A, long long sA,
                                                           const __bf16* __restrict__ Bt, long long sB,
                                                           void* __restrict__ Cv, int ldc, long long sC, int M, int N,
                                                           int K, float alpha) {
    extern __shared__ __attribute__((aligned(16))) __bf16 lds[];
    constexpr int kStage = 2 * 4 * kT * kRow;  // A [4][128][kRow] then B, per stage
    const int bz = blockIdx.z;
    const __bf16* Ab = A + (size_t)bz * sA;
    const __bf16* Bb = Bt + (size_t)bz * sB;
    const size_t pa = (size_t)M * K, pb = (size_t)N * K;
    const int m0 = blockIdx.y * kT, n0 = blockIdx.x * kT;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int wm = (wave >> 2) * 64, wn = (wave & 3) * 32;

    f32x16_t acc[2][1][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 1; ++j)
#pragma unroll
            for (int v = 0; v < 16; ++v) { acc[i][j][0][v] = 0.0f; acc[i][j][1][v] = 0.0f; }

    // Loader: per operand 4 planes x 128 rows x 32 k = 2048 vectors of 8;
    // thread t moves vectors t + 512 j (j < 4): plane (v >> 9), row
    // ((v >> 2) & 127), k offset 8 (v & 3).
    u32x4_t ra[4], rb[4];
    auto fetch = [&](int k0) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int v = tid + 512 * j;
            const int p = v >> 9, row = (v >> 2) & 127, kq = (v & 3) * 8;
            ra[j] = *reinterpret_cast<const u32x4_t*>(Ab + p * pa + (size_t)(m0 + row) * K + k0 + kq);
            rb[j] = *reinterpret_cast<const u32x4_t*>(Bb + p * pb + (size_t)(n0 + row) * K + k0 + kq);
        }
    };
    auto stash = [&](int stage) {
        __bf16* sa = lds + stage * kStage;
        __bf16* sb = sa + 4 * kT * kRow;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int v = tid + 512 * j;
            const int p = v >> 9, row = (v >> 2) & 127, kq = (v & 3) * 8;
            *reinterpret_cast<u32x4_t*>(sa + (p * kT + row) * kRow + kq) = ra[j];
            *reinterpret_cast<u32x4_t*>(sb + (p * kT + row) * kRow + kq) = rb[j];
        }
    };

    // Two LDS stages: step k computes out of one while step k+1's operands
    // (already in registers) are written to the other and step k+2's are
    // fetched; one barrier per step.
    fetch(0);
    stash(0);
    __syncthreads();
    if (kBK < K) fetch(kBK);
    const int r = lane & 31, kh = lane >> 5;
    int stage = 0;
    for (int k0 = 0; k0 < K; k0 += kBK, stage ^= 1) {
        const __bf16* sa = lds + stage * kStage;
        const __bf16* sb = sa + 4 * kT * kRow;
#pragma unroll
        for (int ks = 0; ks < kBK; ks += 16) {
            bf16x8_t a[2][4], b[1][4];
#pragma unroll
            for (int p = 0; p < 4; ++p) {
#pragma unroll
                for (int i = 0; i < 2; ++i)
                    a[i][p] = *reinterpret_cast<const bf16x8_t*>(sa + (p * kT + wm + 32 * i + r) * kRow + ks + 8 * kh);
                b[0][p] = *reinterpret_cast<const bf16x8_t*>(sb + (p * kT + wn + r) * kRow + ks + 8 * kh);
            }
#pragma unroll
            for (int j = 0; j < 1; ++j) {
                const bf16x8_t nbih = neg_bf16x8(b[j][2]), nbil = neg_bf16x8(b[j][3]);
#pragma unroll
                for (int i = 0; i < 2; ++i) {
                    f32x16_t re = acc[i][j][0], im = acc[i][j][1];
                    re = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[i][0], b[j][0], re, 0, 0, 0);
                    im = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[i][0], b[j][2], im, 0, 0, 0);
                    re = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[i][0], b[j][1], re, 0, 0, 0);
                    im = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[i][0], b[j][3], im, 0, 0, 0);
                    re = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[i][1], b[j][0], re, 0, 0, 0);
                    im = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[i][1], b[j][2], im, 0, 0, 0);
                    re = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[i][2], nbih, re, 0, 0, 0);
                    im = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[i][2], b[j][0], im, 0, 0, 0);
                    re = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[i][2], nbil, re, 0, 0, 0);
                    im = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[i][2], b[j][1], im, 0, 0, 0);
                    re = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[i][3], nbih, re, 0, 0, 0);
                    im = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[i][3], b[j][0], im, 0, 0, 0);
                    acc[i][j][0] = re;
                    acc[i][j][1] = im;
                }
            }
        }
        if (k0 + kBK < K) {
            stash(stage ^ 1);
            if (k0 + 2 * kBK < K) fetch(k0 + 2 * kBK);
        }
        __syncthreads();
    }

    // 32 x 32 accumulator layout: lane l holds column l % 32 and rows
    // 8 (v / 4) + 4 (l / 32) + v % 4, v = 0..15.
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 1; ++j) {
            const int col = n0 + wn + 32 * j + r;
#pragma unroll
            for (int v = 0; v < 16; ++v) {
                const int row = m0 + wm + 32 * i + 8 * (v >> 2) + 4 * kh + (v & 3);
                const float re = alpha * acc[i][j][0][v], im = alpha * acc[i][j][1][v];
                if (EPI == 0) {
                    static_cast<float2*>(Cv)[(size_t)bz * sC + (size_t)row * ldc + col] = make_float2(re, im);
                } else if (EPI == 1) {
                    static_cast<float*>(Cv)[(size_t)bz * sC + (size_t)row * ldc + col] = re * re + im * im;
                } else {
                    __bf16* o = static_cast<__bf16*>(Cv) + (size_t)bz * sC;
                    const size_t pc = (size_t)M * ldc, e = (size_t)row * ldc + col;
                    const __bf16 rh = (__bf16)re, ih = (__bf16)im;
                    o[e] = rh;
                    o[pc + e] = (__bf16)(re - (float)rh);
                    o[2 * pc + e] = ih;
                    o[3 * pc + e] = (__bf16)(im - (float)ih);
                }
            }
        }
}

}  // namespace

size_t cgemm_nt3_lds_bytes() { return (size_t)2 * 2 * 4 * kT * kRow * sizeof(__bf16); }

bool cgemm_nt3_fits(int M, int N, int K) { return M > 0 && N > 0 && K > 0 && M % kT == 0 && N % kT == 0 && K % kBK == 0; }

int split_planes(hipStream_t s, const float2* X, int ldx, long long sX, int R, int C, int transpose, int batch,
                 __bf16* out) {
    const size_t RC = (size_t)R * C;
    hipLaunchKernelGGL(split_planes_kernel, dim3((unsigned)((RC + 255) / 256), 1, (unsigned)batch), dim3(256), 0, s, X,
                       ldx, sX, R, C, transpose, out);
    GL_HIP(hipGetLastError());
    return GL_OK;
}

int cgemm_nt3(hipStream_t s, int epi, const __bf16* A, long long sA, const __bf16* Bt, long long sB, void* C, int ldc,
              long long sC, int M, int N, int K, int batch, float alpha) {
    if (!cgemm_nt3_fits(M, N, K)) return set_error(GL_ERR_INVALID, "cgemm_nt3: %d x %d x %d not tile aligned", M, N, K);
    const dim3 grid((unsigned)(N / kT), (unsigned)(M / kT), (unsigned)batch);
    const size_t lds = cgemm_nt3_lds_bytes();
    if (epi == 0)
        hipLaunchKernelGGL(cgemm_nt3_kernel<0>, grid, dim3(512), lds, s, A, sA, Bt, sB, C, ldc, sC, M, N, K, alpha);
    else if (epi == 1)
        hipLaunchKernelGGL(cgemm_nt3_kernel<1>, grid, dim3(512), lds, s, A, sA, Bt, sB, C, ldc, sC, M, N, K, alpha);
    else
        hipLaunchKernelGGL(cgemm_nt3_kernel<2>, grid, dim3(512), lds, s, A, sA, Bt, sB, C, ldc, sC, M, N, K, alpha);
    GL_HIP(hipGetLastError());
    return GL_OK;
}

}  // namespace gl
