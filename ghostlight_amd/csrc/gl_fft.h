// gl_fft.h — FFT building blocks for gfx950: register-resident Stockham row
// transforms staged through LDS, and square tile transposes.
//
// Row transform of length N = 2^LOGN (one row per N/16 lanes):
//   * stages of radix 16 (the last one radix 2^(LOGN mod 4) when LOGN is not a
//     multiple of 4), each an in-register DFT of 16 values per lane;
//   * Stockham autosort (self-sorting): stage s reads x[j + r N/R], twiddles
//     by exp(-2 pi i r k / (Ns R)) with k = j mod Ns, and writes
//     y[(j / Ns) Ns R + k + q Ns]; the first stage reads the source row and the
//     last one writes the destination row, both with unit stride across lanes,
//     so only the LOGN/4 - 1 exchanges in between touch LDS;
//   * LDS slots are XOR-swizzled within each group of 16 (i ^ ((i / 16) mod 16)),
//     which keeps the stride-16 writes of the first exchange free of bank
//     conflicts without growing the row (N complex per row);
//   * twiddles from sincospi of the exact fraction 2k/(Ns R) (Ns R is a power of
//     two), then products of depth <= 4 for the powers 2..15.
// The row source and destination are a policy object (`IO`): plain in-place
// rows, or fused pack / multiply / unpack steps of the veil convolution.
#pragma once

#include <hip/hip_runtime.h>

namespace gl {
namespace fft {

__device__ __forceinline__ float2 c_add(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 c_sub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 c_mul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// x * W^m with W = exp(-2 pi i / 16), m in [0, 8).
template <int M>
__device__ __forceinline__ float2 w16(float2 x) {
    constexpr float C1 = 0.923879532511286756f;   // cos(pi / 8)
    constexpr float S1 = 0.382683432365089772f;   // sin(pi / 8)
    constexpr float H = 0.707106781186547524f;    // sqrt(1 / 2)
    if constexpr (M == 0) {
        return x;
    } else if constexpr (M == 4) {
        return make_float2(x.y, -x.x);
    } else if constexpr (M == 2) {
        return make_float2(H * (x.x + x.y), H * (x.y - x.x));
    } else if constexpr (M == 6) {
        return make_float2(H * (x.y - x.x), -(H * (x.x + x.y)));
    } else {
        constexpr float c = M == 1 ? C1 : M == 3 ? S1 : M == 5 ? -S1 : -C1;
        constexpr float s = M == 1 ? -S1 : M == 3 ? -C1 : M == 5 ? -C1 : -S1;
        return make_float2(x.x * c - x.y * s, x.x * s + x.y * c);
    }
}

// o[k] *= exp(-2 pi i k / R) for k < R/2 (R <= 16).
template <int R, int K = 0>
__device__ __forceinline__ void dit_twiddle(float2* o) {
    if constexpr (K < R / 2) {
        o[K] = w16<K * (16 / R)>(o[K]);
        dit_twiddle<R, K + 1>(o);
    }
}

// In-register forward DFT of R values (radix-2 decimation in time, unrolled).
template <int R>
struct Dft {
    static __device__ __forceinline__ void run(float2* v) {
        float2 e[R / 2], o[R / 2];
#pragma unroll
        for (int i = 0; i < R / 2; ++i) {
            e[i] = v[2 * i];
            o[i] = v[2 * i + 1];
        }
        Dft<R / 2>::run(e);
        Dft<R / 2>::run(o);
        dit_twiddle<R>(o);
#pragma unroll
        for (int k = 0; k < R / 2; ++k) {
            v[k] = c_add(e[k], o[k]);
            v[k + R / 2] = c_sub(e[k], o[k]);
        }
    }
};
template <>
struct Dft<1> {
    static __device__ __forceinline__ void run(float2*) {}
};

// v[r] *= exp(-2 pi i r k / M), M a power of two.
template <int R>
__device__ __forceinline__ void stage_twiddle(float2* v, int k, int M) {
    if constexpr (R > 1) {
        float s, c;
        sincospif(-2.0f * (float)k / (float)M, &s, &c);
        float2 p[R];
        p[1] = make_float2(c, s);
#pragma unroll
        for (int r = 2; r < R; ++r) {
            const int hb = (r & 8) ? 8 : (r & 4) ? 4 : (r & 2) ? 2 : 1;   // highest power of two in r
            p[r] = (r == hb) ? c_mul(p[hb / 2], p[hb / 2]) : c_mul(p[hb], p[r - hb]);
        }
#pragma unroll
        for (int r = 1; r < R; ++r) v[r] = c_mul(v[r], p[r]);
    }
}

template <int LOGN>
struct RowPlan {
    static constexpr int N = 1 << LOGN;
    static constexpr int NST = LOGN < 4 ? 1 : (LOGN + 3) / 4;                  // stages
    static constexpr int RL = LOGN < 4 ? N : (1 << (LOGN - 4 * (NST - 1)));    // last radix
    static constexpr int T = LOGN < 4 ? 1 : N / 16;                            // lanes per row
    static constexpr int ROWS = T >= 256 ? 1 : 256 / T;                        // rows per workgroup
    static constexpr int THREADS = T * ROWS;
    static constexpr int LDS = N;                                              // complex per row
    static constexpr size_t lds_bytes() { return NST > 1 ? (size_t)ROWS * LDS * sizeof(float2) : 0; }
};

__device__ __forceinline__ int lds_slot(int i) { return i ^ ((i >> 4) & 15); }

template <int LOGN, int S, class IO>
__device__ __forceinline__ void row_stages(const IO& io, int row, bool live, int t, float2* lds) {
    using P = RowPlan<LOGN>;
    constexpr int R = (S == P::NST - 1) ? P::RL : 16;
    constexpr int NB = P::N / R;        // butterflies per row
    constexpr int PER = NB / P::T;      // per lane; PER * R = 16 (or N when N < 16)
    constexpr int NS = 1 << (4 * S);    // size of the sub-transforms done so far
    constexpr bool FIRST = S == 0, LAST = S == P::NST - 1;
    float2 v[PER][R];
#pragma unroll
    for (int b = 0; b < PER; ++b) {
        const int j = t + b * P::T;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int i = j + r * NB;
            if constexpr (FIRST) v[b][r] = live ? io.load(row, i) : make_float2(0.0f, 0.0f);
            else v[b][r] = lds[lds_slot(i)];
        }
    }
    if constexpr (!FIRST && !LAST) __syncthreads();   // every lane has read before any writes
#pragma unroll
    for (int b = 0; b < PER; ++b) {
        const int j = t + b * P::T;
        const int k = j & (NS - 1);
        if constexpr (!FIRST) stage_twiddle<R>(v[b], k, NS * R);
        Dft<R>::run(v[b]);
        const int base = (j - k) * R + k;
#pragma unroll
        for (int q = 0; q < R; ++q) {
            const int o = base + q * NS;
            if constexpr (LAST) {
                if (live) io.store(row, o, v[b][q]);
            } else {
                lds[lds_slot(o)] = v[b][q];
            }
        }
    }
    if constexpr (!LAST) {
        __syncthreads();
        row_stages<LOGN, S + 1>(io, row, live, t, lds);
    }
}

// One workgroup transforms RowPlan::ROWS rows; rows >= n_rows only take part
// in the barriers.
template <int LOGN, class IO>
__global__ __launch_bounds__(RowPlan<LOGN>::THREADS) void rows_kernel(IO io, int n_rows) {
    using P = RowPlan<LOGN>;
    extern __shared__ float2 fft_lds[];
    const int lr = (int)threadIdx.x / P::T;
    const int t = (int)threadIdx.x - lr * P::T;
    const int row = (int)blockIdx.x * P::ROWS + lr;
    row_stages<LOGN, 0>(io, row, row < n_rows, t, fft_lds + (size_t)lr * P::LDS);
}

// ---- row source / destination policies ------------------------------------

// In place on a row-major N x N array (row pitch n).
struct RowsInPlace {
    float2* d;
    int n;
    __device__ float2 load(int row, int i) const { return d[(size_t)row * n + i]; }
    __device__ void store(int row, int i, float2 v) const { d[(size_t)row * n + i] = v; }
};

template <int LOGN, class IO>
inline hipError_t launch_rows(hipStream_t s, const IO& io, int n_rows) {
    using P = RowPlan<LOGN>;
    const unsigned grid = (unsigned)((n_rows + P::ROWS - 1) / P::ROWS);
    if (grid == 0) return hipSuccess;
    hipLaunchKernelGGL((rows_kernel<LOGN, IO>), dim3(grid), dim3(P::THREADS), P::lds_bytes(), s, io, n_rows);
    return hipGetLastError();
}

// LDS bytes the row transform of 2^logn needs per workgroup.
inline size_t rows_lds_bytes(int logn) {
    switch (logn) {
#define GL_FFT_LDS(L) case L: return RowPlan<L>::lds_bytes();
        GL_FFT_LDS(1) GL_FFT_LDS(2) GL_FFT_LDS(3) GL_FFT_LDS(4) GL_FFT_LDS(5) GL_FFT_LDS(6) GL_FFT_LDS(7)
        GL_FFT_LDS(8) GL_FFT_LDS(9) GL_FFT_LDS(10) GL_FFT_LDS(11) GL_FFT_LDS(12) GL_FFT_LDS(13) GL_FFT_LDS(14)
#undef GL_FFT_LDS
    }
    return ~(size_t)0;
}

// Dispatch on log2(N) in [1, 14].
template <class IO>
inline hipError_t rows(hipStream_t s, int logn, const IO& io, int n_rows) {
    switch (logn) {
#define GL_FFT_CASE(L) case L: return launch_rows<L>(s, io, n_rows);
        GL_FFT_CASE(1) GL_FFT_CASE(2) GL_FFT_CASE(3) GL_FFT_CASE(4) GL_FFT_CASE(5) GL_FFT_CASE(6) GL_FFT_CASE(7)
        GL_FFT_CASE(8) GL_FFT_CASE(9) GL_FFT_CASE(10) GL_FFT_CASE(11) GL_FFT_CASE(12) GL_FFT_CASE(13) GL_FFT_CASE(14)
#undef GL_FFT_CASE
    }
    return hipErrorInvalidValue;
}

// ---- transposes ------------------------------------------------------------

// In-place transpose of a square n x n array by tile pairs: workgroup p
// swaps tile (I, J) with tile (J, I), I <= J, p = J (J + 1) / 2 + I.
static __global__ __launch_bounds__(256) void transpose_square_inplace_kernel(float2* __restrict__ d, int n) {
    __shared__ float2 ta[32][33];
    __shared__ float2 tb[32][33];
    const int p = (int)blockIdx.x;
    int J = (int)((sqrtf(8.0f * (float)p + 1.0f) - 1.0f) * 0.5f);
    while ((J + 1) * (J + 2) / 2 <= p) ++J;
    while (J * (J + 1) / 2 > p) --J;
    const int I = p - J * (J + 1) / 2;
    const int tx = (int)threadIdx.x & 31, ty = (int)threadIdx.x >> 5;
    const int r0 = I * 32, c0 = J * 32;   // tile A at rows r0.., cols c0..; tile B is its mirror
    for (int r = ty; r < 32; r += 8) {
        const int y = r0 + r, x = c0 + tx;
        if (y < n && x < n) ta[r][tx] = d[(size_t)y * n + x];
        const int y2 = c0 + r, x2 = r0 + tx;
        if (I != J && y2 < n && x2 < n) tb[r][tx] = d[(size_t)y2 * n + x2];
    }
    __syncthreads();
    for (int r = ty; r < 32; r += 8) {
        // new A[r][c] = old B[c][r]; new B[r][c] = old A[c][r]
        const int y = r0 + r, x = c0 + tx;
        if (y < n && x < n) d[(size_t)y * n + x] = I != J ? tb[tx][r] : ta[tx][r];
        const int y2 = c0 + r, x2 = r0 + tx;
        if (I != J && y2 < n && x2 < n) d[(size_t)y2 * n + x2] = ta[tx][r];
    }
}

inline hipError_t transpose_square_inplace(hipStream_t s, float2* d, int n) {
    const int nt = (n + 31) / 32;
    const unsigned pairs = (unsigned)((size_t)nt * (nt + 1) / 2);
    hipLaunchKernelGGL(transpose_square_inplace_kernel, dim3(pairs), dim3(256), 0, s, d, n);
    return hipGetLastError();
}

// out[r][c] = in[c][r] for r < out_rows, c < n; rows of `in` at or beyond
// in_rows read as zero (they are never touched).  Both arrays have pitch n.
static __global__ __launch_bounds__(256) void transpose_limited_kernel(const float2* __restrict__ in, int in_rows,
                                                                float2* __restrict__ out, int out_rows, int n) {
    __shared__ float2 tile[32][33];
    const int ob_c = (int)blockIdx.x * 32;   // out column block = in row block
    const int ob_r = (int)blockIdx.y * 32;   // out row block = in column block
    const int tx = (int)threadIdx.x & 31, ty = (int)threadIdx.x >> 5;
    for (int r = ty; r < 32; r += 8) {
        const int y = ob_c + r, x = ob_r + tx;
        tile[r][tx] = (y < in_rows && x < n) ? in[(size_t)y * n + x] : make_float2(0.0f, 0.0f);
    }
    __syncthreads();
    for (int r = ty; r < 32; r += 8) {
        const int y = ob_r + r, x = ob_c + tx;
        if (y < out_rows && x < n) out[(size_t)y * n + x] = tile[tx][r];
    }
}

inline hipError_t transpose_limited(hipStream_t s, const float2* in, int in_rows, float2* out, int out_rows, int n) {
    const dim3 grid((unsigned)((n + 31) / 32), (unsigned)((out_rows + 31) / 32));
    if (grid.y == 0) return hipSuccess;
    hipLaunchKernelGGL(transpose_limited_kernel, grid, dim3(256), 0, s, in, in_rows, out, out_rows, n);
    return hipGetLastError();
}

}  // namespace fft
}  // namespace gl
