// ghost.hip — the ghost path on gfx950: pair pre-filter probes, the
// probe / cull / concentration passes, the ghost trace + splat, the spectral
// colour correction and the host launcher that sequences them on a context
// stream (launch_ghost_render / render_ghost_pipeline semantics,
// ghostlight/src/ghost_render.cu:630-1581).
//
// Work mapping.  blockIdx.x = (pair, source) slot, blockIdx.y = 256-sample
// chunk of the pupil grid; one lane per pupil sample loops over the spectral
// samples (the reference's decomposition, ghost_render.cu:249-254,372), so a
// wave shares its pair and walks the prescription with uniform indices.
#include "gl_runtime.h"
#include "gl_splat.h"

#include <algorithm>
#include <cmath>
#include <cstdio>

namespace gl {

static constexpr int kBlock = 256;
static constexpr int kProbeGrid = 40;
static constexpr float kCullMargin = 1.25f;
static constexpr int kProbePasses = 3;
static constexpr float kConcDilateCells = 1.5f;
static constexpr int kConcMinHits = 3;
static constexpr int kRefineGrid = 128;
static constexpr int kRefineHitsMax = 16;
static constexpr int kRefineMaxEntries = 4096;
static constexpr int kStage = 256;            // deposit records staged per wave in LDS
static constexpr long long kProbeBlocks = 16384;   // workgroups a probe launch aims at

// Persistent-launch schedule candidates the autotuner times: lambda range
// per work item (ns / lam_div), workgroups per CU, longest-first order.  The
// lambda range is also the number of wavelengths one pass of 64 lanes holds
// (deposit_lanes), which decides how coherent a pass is against how many
// distinct per-lambda table rows it gathers: the 24-surface 31-lambda render
// is fastest at 4 wavelengths a pass, the 30-surface 16-lambda sweep at 16.
struct Sched { int lam_div, blocks_per_cu, longest_first; };
static constexpr Sched kSched[] = {{1, 4, 1}, {1, 8, 1}, {2, 8, 1}, {4, 8, 1}, {8, 8, 1}, {16, 8, 1}};
static constexpr int kNumSched = (int)(sizeof(kSched) / sizeof(kSched[0]));

// hseed: the source-angle part of the per-ray HURB seed, float bits of
// angle_x * 0x9E3779B9 + angle_y * 0x85EBCA6B (ghost_render.cu:362-368).
struct SrcDev { float bx, by, dx, dy, dz, r, g, b; uint32_t hseed; };

struct ConcDev {
    const float4* rects;      // nullptr = off
    DevMask mask;
    int jitter;
    uint32_t seed_offset;
    int nr_lattice, n_r;
    float area_norm;
    const int2* budgets;      // nullptr = uniform budget
    float inv_aref;
};

struct GhostArgs {
    const DevSurf* __restrict__ S;
    const float* __restrict__ poly;
    const float* __restrict__ arena;
    int n;
    const gl_pair* __restrict__ pairs;
    int n_src;
    const SrcDev* __restrict__ src;
    const gl_sample* __restrict__ grid;
    int n_grid;
    float front_R, start_z;
    float sw, sh;
    int width, height;
    float fmt_w, fmt_h, x0, y0;
    float* __restrict__ out_r;
    float* __restrict__ out_g;
    float* __restrict__ out_b;
    float gain, ray_weight;
    const float4* __restrict__ spec;   // lambda, cmf_r, cmf_g, cmf_b
    int n_spec;
    const float* __restrict__ iat;
    const float* __restrict__ ibf;
    const unsigned char* __restrict__ pair_active;
    const unsigned char* __restrict__ ps_refine;
    const unsigned* __restrict__ refine_n;      // refine pass: slots the mask holds (run only for 1..kRefineMaxEntries)
    ConcDev conc;
    int* __restrict__ bbox;                     // STATS: mins | maxs | hits
    unsigned long long* __restrict__ cnt;       // STATS: traces | survived | on-sensor (per pair)
    unsigned long long* __restrict__ segs;      // optional segment counter
    const int2* __restrict__ lrow;              // SPLAT: per-(lambda, surface) coating lookups
    const float2* __restrict__ rat;             // per-(lambda, surface) refraction ratios (trace_ghost)
    const LamRec* __restrict__ lrec;            // iat, ibf, ratios and lrow interleaved (trace_ghost_rec)
    int lam_group;                              // STATS: lanes per sample (one lambda each), 0 = loop
    int probe_iters;                            // STATS: sample chunks per workgroup (>= 1)
    int frame_mode;                             // SPLAT: records carry a frame: 1 = source, 2 = pair
    float4* __restrict__ rec;                   // SPLAT: deposit records
    unsigned long long rec_cap;
    unsigned long long* __restrict__ rec_count;
    HurbDev hurb;                               // edge diffraction (HURB launches only)
    uint32_t hurb_seed_base;                    // jitter_seed * 2654435761 + 1
    DevBaffles baffles;                         // front-of-lens clip (n = 0: none)
};

__device__ __forceinline__ int enc_pupil(float x) { return __float_as_int(x + 2.0f); }

__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}
__device__ __forceinline__ float wave_min(float v) {
    for (int off = 32; off > 0; off >>= 1) v = fminf(v, __shfl_xor(v, off, 64));
    return v;
}
__device__ __forceinline__ float wave_max(float v) {
    for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, 64));
    return v;
}

// Concentration lattice index -> lattice index walked in 8 x 8 tiles, so the
// 64 lanes of a wave trace a compact pupil patch instead of a 64-sample strip:
// a vignetting edge then leaves fewer waves partly dead.  A bijection on
// [0, lim): full 8-row bands go tile by tile (the last tile of a band is
// nr % 8 wide), the rows after them keep row-major order.  The sample set,
// and so the image, is unchanged; only the lane assignment moves.  `lim` is
// the number of lattice indices traced (a budget lattice larger than the pupil
// grid is cut at the grid size): only the 8-row bands wholly below it are
// reordered, so the traced set stays [0, lim).
__device__ __forceinline__ int tile_order(int idx, int nr, int lim) {
    const int band_len = 8 * nr, full_bands = lim / band_len;
    if (idx >= full_bands * band_len) return idx;
    const int band = idx / band_len, r = idx - band * band_len, full = nr >> 3;
    int gx, gy;
    if (r < full * 64) {
        gx = ((r >> 6) << 3) + (r & 7);
        gy = (r >> 3) & 7;
    } else {
        const int rem = nr - 8 * full, q = r - full * 64;
        gy = q / rem;
        gx = 8 * full + (q - gy * rem);
    }
    return (band * 8 + gy) * nr + gx;
}

// Sample position and estimator weight (ghost_render.cu:280-335).
// gi comes back as the index that names the sample (the lattice index once a
// concentrated lattice is walked in tiles); HURB seeds with it.
__device__ __forceinline__ bool sample_at(const GhostArgs& A, int ps, int& gi, float& u, float& v, float& w_ray) {
    w_ray = A.ray_weight;
    if (A.conc.rects != nullptr) {
        const float4 rc = A.conc.rects[ps];
        if (rc.x <= rc.z) {
            int n_r_ps = A.conc.n_r, nr_lat = A.conc.nr_lattice;
            if (A.conc.budgets) { const int2 b = A.conc.budgets[ps]; n_r_ps = b.x; nr_lat = b.y; }
            if (gi >= n_r_ps) return false;
            float fu, fv;
            if (A.conc.jitter == 2) { fu = halton2((uint32_t)gi); fv = halton3((uint32_t)gi); }
            else {
                gi = tile_order(gi, nr_lat, min(n_r_ps, A.n_grid));
                const int gx = gi % nr_lat, gy = gi / nr_lat, nn = nr_lat * nr_lat;
                const float ju = A.conc.jitter == 1 ? (float)wang_hash((uint32_t)gi + A.conc.seed_offset) / 4294967296.0f : 0.5f;
                const float jv = A.conc.jitter == 1 ? (float)wang_hash((uint32_t)gi + (uint32_t)nn + A.conc.seed_offset) / 4294967296.0f : 0.5f;
                fu = ((float)gx + ju) / (float)nr_lat;
                fv = ((float)gy + jv) / (float)nr_lat;
            }
            u = rc.x + fu * (rc.z - rc.x);
            v = rc.y + fv * (rc.w - rc.y);
            if (!mask_contains(A.conc.mask, u, v)) return false;
            w_ray = (rc.z - rc.x) * (rc.w - rc.y) * (A.conc.budgets ? A.conc.inv_aref / (float)n_r_ps : A.conc.area_norm);
            return true;
        }
    }
    const gl_sample g = A.grid[gi];
    u = g.u; v = g.v;
    return true;
}

// Flush a wave's staged deposit records to HBM: one 64-bit reservation, then
// the 64 lanes copy the run out with consecutive (coalesced) 16-byte stores.
// Records past the capacity are counted but not written; the host regrows
// the buffer and renders again.
__device__ __forceinline__ void flush_stage(const GhostArgs& A, float4* stage, int n, int lane) {
    unsigned long long base = 0;
    if (lane == 0) base = atomicAdd(A.rec_count, (unsigned long long)n);
    base = __shfl(base, 0, 64);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    for (int i = lane; i < n; i += 64) {
        const unsigned long long d = base + (unsigned long long)i;
        if (d < A.rec_cap) A.rec[d] = stage[i];
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Trace one lane's pupil sample `gi` of (pair, source) slot `ps` over the
// wavelengths [si0, si1) and stage a deposit record for every landing
// (ghost_render.cu:372-462).  The wave must be converged: the lambda loop is
// uniform and the staging uses wave ballots.
//
// Front-of-lens baffles drop the ray before any wavelength (achromatic).
// H = HURB: each wavelength gets its own kick stream, seeded from the source
// angle, sample, pair and jitter seed; a baffle edge kicks the entering ray on
// a stream of its own (ghost_render.cu:351-392).
__device__ __forceinline__ uint32_t hurb_ray_seed(const GhostArgs& A, const SrcDev& sr, int gi, int pair_idx) {
    return sr.hseed + (uint32_t)gi * 0xC2B2AE35u + (uint32_t)pair_idx * 0x27D4EB2Fu + A.hurb_seed_base;
}
template <bool H>
__device__ __forceinline__ F3 hurb_entry(const GhostArgs& A, const SrcDev& sr, F3 o0, F3 d0, float lam, uint32_t ray,
                                         int si, uint32_t& hs) {
    hs = hurb_hash(ray + (uint32_t)si * 0x9E3779B9u);
    if (A.baffles.n > 0) {
        uint32_t bs = hurb_hash(hs ^ 0xBAFF1E5Du);
        F3 bn;
        const float bd = baffles_edge(A.baffles, o0.x, o0.y, sr.bx, sr.by, bn);
        hurb_apply(A.hurb, d0, bn, bd, lam, bs);
    }
    return d0;
}

// p with its value hidden from the optimiser: what is read through the
// result is loaded where it is used (scalar loads from the kernel argument
// segment) instead of being hoisted and held in registers across the trace.
template <class T> __device__ __forceinline__ const T GLC* fresh(const T GLC* p) {
    asm volatile("" : "+s"(p));
    return p;
}

// The launch's GhostArgs where the kernel receives it: every kernel that
// deposits takes GhostArgs as its first argument, at offset 0 of the kernel
// argument segment.  (Taking the parameter's address instead would make the
// compiler copy the whole struct to scratch.)
__device__ __forceinline__ const GhostArgs GLC* kernel_args() {
    return (const GhostArgs GLC*)__builtin_amdgcn_kernarg_segment_ptr();
}

// One (pupil sample, wavelength) trace of a deposit pass: true with the
// deposit record in rec when the landing deposits (ghost_render.cu:372-462).
// Only the wavelength is needed during the trace; the colour weights and the
// launch's scalars are read after it (through fresh(Ap)), so they hold no
// registers across the surface loop.
// REC: the wavelength's tables come from the interleaved records (the
// lambda-across-lanes form, no HURB).
template <bool H, int LF, int CM, bool REC = false>
__device__ __forceinline__ bool trace_record(const GhostArgs& A, const GhostArgs GLC* Ap, const gl_pair& pr,
                                             int pair_idx, int src_idx, const SrcDev& sr, F3 o0, F3 d0, float w_ray,
                                             uint32_t hray, int si, float4& rec) {
    const float GLC* spf = as_const(reinterpret_cast<const float*>(A.spec)) + 4 * si;
    const float lam = spf[0];
    TraceRes res;
    if constexpr (REC) {
        static_assert(!H, "the record form has no HURB");
        res = trace_ghost_rec<LF, CM>(as_const(A.S), A.n, as_const(A.poly), A.arena, pr.surf_a, pr.surf_b, o0, d0,
                                      lam, as_const(A.lrec) + (size_t)si * A.n, false);
    } else {
        uint32_t hs = 0u;
        const F3 dl = H ? hurb_entry<H>(A, sr, o0, d0, lam, hray, si, hs) : d0;
        res = trace_ghost<true, H, LF, CM, true>(as_const(A.S), A.n, as_const(A.poly), A.arena, pr.surf_a, pr.surf_b,
                                                 o0, dl, lam, as_const(A.iat) + (size_t)si * A.n,
                                                 as_const(A.ibf) + (size_t)si * A.n, false,
                                                 A.lrow ? as_const(A.lrow) + (size_t)si * A.n : nullptr, &A.hurb, &hs,
                                                 as_const(A.rat) + (size_t)si * A.n);
    }
    if (res.status != GL_TRACE_OK) return false;
    if (!isfinite(res.px) || !isfinite(res.py)) return false;
    const GhostArgs GLC& B = *fresh(Ap);
    const SrcDev GLC& srb = as_const(B.src)[src_idx];
    float base = res.w * w_ray * B.gain * as_const(B.pairs)[pair_idx].area_boost;
    if (base < 1e-14f) return false;
    float px = (res.px / (2.0f * B.sw) + 0.5f) * B.fmt_w + B.x0;
    float py = (res.py / (2.0f * B.sh) + 0.5f) * B.fmt_h + B.y0;
    if (!isfinite(px) || !isfinite(py)) return false;
    px = fmaxf(-2.0e9f, fminf(2.0e9f, px));
    py = fmaxf(-2.0e9f, fminf(2.0e9f, py));
    const int x0 = (int)floorf(px - 0.5f), y0 = (int)floorf(py - 0.5f);
    if (x0 + 1 < 0 || x0 >= B.width || y0 + 1 < 0 || y0 >= B.height) return false;   // footprint off the image
    const float GLC* spc = as_const(reinterpret_cast<const float*>(B.spec)) + 4 * si;
    const float cr = spc[1], cg = spc[2], cb = spc[3];
    const float P = srb.r * cr + srb.g * cg + srb.b * cb;
    const int frame = B.frame_mode == 2 ? pair_idx : (B.frame_mode ? src_idx : 0);
    rec = make_float4(px, py, P * base, __int_as_float(si | (frame << 8)));
    return true;
}

// Stage the records of the wave's depositing lanes (a ballot and a prefix
// count give each its slot), flushing first when they would overflow.
__device__ __forceinline__ void stage_records(const GhostArgs& A, float4* stage, int& staged, int lane, bool dep,
                                              const float4& rec) {
    const unsigned long long bal = __ballot(dep);
    const int cnt = (int)__popcll(bal);
    if (cnt) {
        if (staged + cnt > kStage) { flush_stage(A, stage, staged, lane); staged = 0; }
        if (dep) {
            const int slot = staged + (int)__builtin_amdgcn_mbcnt_hi((unsigned)(bal >> 32),
                                                                    __builtin_amdgcn_mbcnt_lo((unsigned)bal, 0u));
            stage[slot] = rec;
        }
        staged += cnt;
    }
}

template <bool H, int LF = 0, int CM = CC_ALL>
__device__ __forceinline__ void deposit_range(const GhostArgs& A, int ps, int gi, int si0, int si1, float4* stage,
                                              int& staged, int lane) {
    const GhostArgs GLC* Ap = kernel_args();
    const int pair_idx = ps / A.n_src, src_idx = ps % A.n_src;
    const gl_pair pr = A.pairs[pair_idx];
    float u = 0.f, v = 0.f, w_ray = 0.f;
    bool live = gi < A.n_grid;
    if (live) live = sample_at(A, ps, gi, u, v, w_ray);
    const SrcDev sr = A.src[src_idx];
    const F3 o0 = f3(u * A.front_R + (-20.0f * sr.bx), v * A.front_R + (-20.0f * sr.by), A.start_z);
    const F3 d0 = f3(sr.dx, sr.dy, sr.dz);
    if (A.baffles.n > 0 && live && baffles_block(A.baffles, o0.x, o0.y, sr.bx, sr.by)) live = false;
    const uint32_t hray = H ? hurb_ray_seed(A, sr, gi, pair_idx) : 0u;
    for (int si = si0; si < si1; ++si) {
        float4 rec = make_float4(0.f, 0.f, 0.f, 0.f);
        const bool dep = live && trace_record<H, LF, CM>(A, Ap, pr, pair_idx, src_idx, sr, o0, d0, w_ray, hray, si, rec);
        stage_records(A, stage, staged, lane, dep, rec);
    }
}

// Whether the lambda-across-lanes trace reads the interleaved records: where
// the coating weight reads coating rows (a third per-lambda table), one
// record address serves all three; with only Airy and Fresnel weights the two
// separate tables measured faster.
template <int CM> static constexpr bool kRecForm = CM != (CC_AIRY | CC_FRESNEL);

// The same traces with the wavelengths across lanes: the wave's 64 pupil
// samples (one per lane, as deposit_range takes them) times the L = si1 - si0
// wavelengths are 64 L (sample, wavelength) traces, run in L passes of 64;
// trace q = 64 p + lane of pass p is sample lane q / L at wavelength
// si0 + q % L, its origin and weight fetched from that sample's lane.  The
// lanes of one pass hold a few neighbouring samples at consecutive
// wavelengths, which mostly meet the same fate, so a pass lasts as long as
// its longest few samples rather than the longest of 64.
template <int LF, int CM>
__device__ __forceinline__ void deposit_lanes(const GhostArgs& A, int ps, int gi, int si0, int si1, float4* stage,
                                              int& staged, int lane) {
    const GhostArgs GLC* Ap = kernel_args();
    const int pair_idx = ps / A.n_src, src_idx = ps % A.n_src;
    const gl_pair pr = A.pairs[pair_idx];
    float u = 0.f, v = 0.f, w_ray = 0.f;
    bool live = gi < A.n_grid;
    if (live) live = sample_at(A, ps, gi, u, v, w_ray);
    const SrcDev sr = A.src[src_idx];
    const float ox = u * A.front_R + (-20.0f * sr.bx), oy = v * A.front_R + (-20.0f * sr.by);
    if (A.baffles.n > 0 && live && baffles_block(A.baffles, ox, oy, sr.bx, sr.by)) live = false;
    const F3 d0 = f3(sr.dx, sr.dy, sr.dz);
    const int L = si1 - si0;
    if (L <= 0) return;
    // (sl, r) = (q / L, q % L), advanced by 64 = dq L + mq per pass.
    const int dq = 64 / L, mq = 64 - dq * L;
    int sl = lane / L, r = lane - (lane / L) * L;
    for (int p = 0; p < L; ++p) {
        const float qx = __shfl(ox, sl, 64), qy = __shfl(oy, sl, 64), qw = __shfl(w_ray, sl, 64);
        const bool ql = __shfl((int)live, sl, 64) != 0;
        float4 rec = make_float4(0.f, 0.f, 0.f, 0.f);
        const bool dep = ql && trace_record<false, LF, CM, kRecForm<CM>>(A, Ap, pr, pair_idx, src_idx, sr,
                                                                         f3(qx, qy, A.start_z), d0, qw, 0u, si0 + r,
                                                                         rec);
        stage_records(A, stage, staged, lane, dep, rec);
        sl += dq; r += mq;
        if (r >= L) { r -= L; ++sl; }
    }
}

// Persistent main trace: a fixed set of workgroups pulls (slot, sample
// chunk, lambda range) items from a device work list, longest first when the
// schedule says so; deposit staging carries over between items.  clk[0] /
// clk[1] receive the device wall clock at the first start / last finish.
struct PersistQ {
    const int4* __restrict__ work;      // {ps, chunk, si0, si1}
    unsigned* __restrict__ ctl;         // [0] queue head, [1] item count
    unsigned long long* __restrict__ clk;
};

// Eight waves per SIMD: the trace is VALU-bound with long dependent chains,
// and the extra waves hide more latency than the few spilled registers cost.
// LF: the lens's lean level (lean_level); CM: the coating
// classes its surfaces use (coating_classes).
template <int LF, int CM>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(8, 8)))
void ghost_persist_kernel(GhostArgs A, PersistQ Q) {
    __shared__ float4 stage_all[kBlock / 64][kStage];
    __shared__ unsigned s_item;
    const int lane = threadIdx.x & 63;
    float4* stage = stage_all[threadIdx.x >> 6];
    int staged = 0;
    if (threadIdx.x == 0) atomicMin(Q.clk, (unsigned long long)wall_clock64());
    const unsigned n_items = Q.ctl[1];
    for (;;) {
        if (threadIdx.x == 0) s_item = atomicAdd(Q.ctl, 1u);
        __syncthreads();
        const unsigned item = s_item;
        __syncthreads();
        if (item >= n_items) break;
        const int4 w = Q.work[item];
        deposit_lanes<LF, CM>(A, w.x, w.y * kBlock + (int)threadIdx.x, w.z, w.w, stage, staged, lane);
    }
    if (staged) flush_stage(A, stage, staged, lane);
    __syncthreads();
    if (threadIdx.x == 0) atomicMax(Q.clk + 1, (unsigned long long)wall_clock64());
}

__global__ void persist_reset_kernel(PersistQ Q) {
    Q.ctl[0] = 0u;
    Q.clk[0] = ~0ull;
    Q.clk[1] = 0ull;
}

// STATS = probe / diagnostic pass (counters, survivor bbox); otherwise the
// render (SPLAT) and/or the segment count (COUNT).
// LF: the lens's lean level (lean_level), as for the
// persistent kernel.
template <bool STATS, bool SPLAT, bool COUNT, bool H, int LF = 0>
__global__ __launch_bounds__(kBlock) void ghost_kernel(GhostArgs A) {
    const int ps = blockIdx.x;
    const int pair_idx = ps / A.n_src, src_idx = ps % A.n_src;
    const int lane = threadIdx.x & 63;
    // STATS with lam_group = G > 0: G lanes per pupil sample, one wavelength
    // each (the probe passes are short; spreading lambda fills the device).
    const int G = STATS ? A.lam_group : 0;
    const int chunk0 = (int)blockIdx.y * (STATS ? A.probe_iters : 1);   // first sample chunk of this workgroup
    int gi, si0 = 0, si1 = A.n_spec;
    if (STATS && G > 0) {
        const int Gs = G > 0 ? G : 1;                // keeps the non-STATS instantiations free of a zero divisor
        gi = chunk0 * (kBlock / Gs) + (int)(threadIdx.x >> 6) * (64 / Gs) + lane / Gs;
        si0 = lane % Gs;
        si1 = si0 < A.n_spec ? si0 + 1 : si0;
    } else {
        gi = chunk0 * kBlock + (int)threadIdx.x;
    }
    // Whole-block exits are uniform; per-lane validity stays a flag so the
    // wave reductions below see every lane.
    if (A.pair_active != nullptr && A.pair_active[pair_idx] == 0) return;
    if (A.refine_n != nullptr) {
        const unsigned nr = *A.refine_n;
        if (nr == 0u || nr > (unsigned)kRefineMaxEntries) return;
    }
    if (A.ps_refine != nullptr && A.ps_refine[ps] == 0) return;
    const gl_pair pr = A.pairs[pair_idx];
    if (pr.surf_a < 0 || pr.surf_b <= pr.surf_a || pr.surf_b >= A.n) return;

    if (SPLAT) {
        __shared__ float4 stage_all[kBlock / 64][kStage];
        float4* stage = stage_all[threadIdx.x >> 6];
        int staged = 0;                              // wave-uniform
        deposit_range<H, LF>(A, ps, gi, 0, A.n_spec, stage, staged, lane);
        if (staged) flush_stage(A, stage, staged, lane);
        return;
    }
    unsigned long long n_tr = 0, n_sv = 0, n_on = 0, n_seg = 0;
    // STATS: a workgroup takes probe_iters consecutive sample chunks (chunk
    // blockIdx.y * iters + it) and keeps its counts, its on-sensor box and
    // its hits in registers across them, so the per-pair counters and the
    // per-slot box take one set of atomics per wave instead of one per chunk.
    const int iters = STATS ? A.probe_iters : 1;
    const int chunk_lanes = (STATS && G > 0) ? kBlock / G : kBlock;
    float bumn = 3.0f, bvmn = 3.0f, bumx = -3.0f, bvmx = -3.0f;
    int hits_acc = 0;
    unsigned long long any_on = 0;
    for (int it = 0; it < iters; ++it) {
    if (it > 0) gi += chunk_lanes;
    const unsigned long long on0 = n_on;
    float u = 0.f, v = 0.f, w_ray = 0.f;
    int gs = gi;
    bool live = gs < A.n_grid;
    if (live) live = sample_at(A, ps, gs, u, v, w_ray);
    const SrcDev sr = A.src[src_idx];
    const F3 o0 = f3(u * A.front_R + (-20.0f * sr.bx), v * A.front_R + (-20.0f * sr.by), A.start_z);
    const F3 d0 = f3(sr.dx, sr.dy, sr.dz);
    if (A.baffles.n > 0 && live && baffles_block(A.baffles, o0.x, o0.y, sr.bx, sr.by)) live = false;
    const uint32_t hray = H ? hurb_ray_seed(A, sr, gs, pair_idx) : 0u;
    if (live)
        for (int si = si0; si < si1; ++si) {
            const float GLC* spf = as_const(reinterpret_cast<const float*>(A.spec)) + 4 * si;
            if (STATS) ++n_tr;
            uint32_t hs = 0u;
            const F3 dl = H ? hurb_entry<H>(A, sr, o0, d0, spf[0], hray, si, hs) : d0;
            TraceRes res = trace_ghost<false, H, LF, CC_ALL, true>(
                as_const(A.S), A.n, as_const(A.poly), A.arena, pr.surf_a, pr.surf_b, o0, dl, spf[0],
                as_const(A.iat) + (size_t)si * A.n, as_const(A.ibf) + (size_t)si * A.n, false, nullptr, &A.hurb, &hs,
                as_const(A.rat) + (size_t)si * A.n);
            n_seg += (unsigned long long)res.segs;
            if (!STATS) continue;
            if (res.status != GL_TRACE_OK) continue;
            if (!isfinite(res.px) || !isfinite(res.py)) continue;
            ++n_sv;
            float spx = (res.px / (2.0f * A.sw) + 0.5f) * A.fmt_w + A.x0;
            float spy = (res.py / (2.0f * A.sh) + 0.5f) * A.fmt_h + A.y0;
            if (isfinite(spx) && isfinite(spy) && spx >= 0.0f && spx < (float)A.width && spy >= 0.0f &&
                spy < (float)A.height) ++n_on;
        }
    if (STATS && A.bbox != nullptr) {
        // This chunk's box and hits (samples with any wavelength on the sensor).
        const bool on = n_on > on0;
        const unsigned long long bal = __ballot(on);
        int hits = (int)__popcll(bal);
        if (G > 0) {
            hits = 0;
            const unsigned long long gm = G == 64 ? ~0ull : ((1ull << G) - 1ull);
            for (int g0 = 0; g0 < 64; g0 += G) hits += ((bal >> g0) & gm) != 0ull ? 1 : 0;
        }
        hits_acc += hits;
        any_on |= bal;
        if (on) { bumn = fminf(bumn, u); bvmn = fminf(bvmn, v); bumx = fmaxf(bumx, u); bvmx = fmaxf(bvmx, v); }
    }
    }
    if (STATS) {
        const unsigned long long t = wave_sum_u64(n_tr), s = wave_sum_u64(n_sv), o = wave_sum_u64(n_on);
        const unsigned long long sg = wave_sum_u64(n_seg);
        const bool lane0 = (threadIdx.x & 63) == 0;
        if (lane0) {
            if (t) atomicAdd(A.cnt + pair_idx, t);
            if (s) atomicAdd(A.cnt + gridDim.x / A.n_src + pair_idx, s);
            if (o) atomicAdd(A.cnt + 2 * (gridDim.x / A.n_src) + pair_idx, o);
            if (A.segs && sg) atomicAdd(A.segs, sg);
        }
        if (A.bbox != nullptr) {
            const float umn = wave_min(bumn), vmn = wave_min(bvmn);
            const float umx = wave_max(bumx), vmx = wave_max(bvmx);
            if (lane0 && any_on) {
                const int n_ps = (int)gridDim.x;
                atomicMin(A.bbox + 2 * ps, enc_pupil(umn));
                atomicMin(A.bbox + 2 * ps + 1, enc_pupil(vmn));
                atomicMax(A.bbox + 2 * n_ps + 2 * ps, enc_pupil(umx));
                atomicMax(A.bbox + 2 * n_ps + 2 * ps + 1, enc_pupil(vmx));
                atomicAdd(A.bbox + 4 * n_ps + ps, hits_acc);
            }
        }
    } else if (COUNT) {
        const unsigned long long sg = wave_sum_u64(n_seg);
        if ((threadIdx.x & 63) == 0 && sg) atomicAdd(A.segs, sg);
    }
}

// The counters a render reads back, gathered into one block for one copy:
// records made and binned, resolve chunks | staged tiles << 32, and the
// device wall clock at the persistent trace's first start and last finish.
__global__ void render_summary_kernel(SplatBufs B, int n_tiles, const unsigned long long* __restrict__ clk,
                                      unsigned long long* __restrict__ out) {
    out[0] = *B.rec_count;
    out[1] = *B.bin_total;
    out[2] = (unsigned long long)B.chunk_start[n_tiles] | ((unsigned long long)*B.mc_count << 32);
    out[3] = clk ? clk[0] : 0ull;
    out[4] = clk ? clk[1] : 0ull;
}

// Refine mask of the (pair, source) slots: hits below `hmax`, and how many
// (one workgroup of 1024 lanes).
__global__ __launch_bounds__(1024) void refine_mask_kernel(const int* __restrict__ hits, int n, int hmax,
                                                           unsigned char* __restrict__ mask, unsigned* __restrict__ count) {
    __shared__ unsigned s_n;
    if (threadIdx.x == 0) s_n = 0u;
    __syncthreads();
    unsigned k = 0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const bool r = hits[i] < hmax;
        mask[i] = r ? 1 : 0;
        k += r ? 1u : 0u;
    }
    atomicAdd(&s_n, k);
    __syncthreads();
    if (threadIdx.x == 0) *count = s_n;
}

__global__ void fill3_kernel(float* __restrict__ a, float* __restrict__ b, float* __restrict__ c, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) { a[i] = 0.f; b[i] = 0.f; c[i] = 0.f; }
}

// Output-space 3x3 colour correction (ghost_render.cu:610-622); contraction
// is off for this file so each product and sum rounds separately.
__global__ void spectral_correct_kernel(float* __restrict__ r, float* __restrict__ g, float* __restrict__ b, size_t n,
                                        float c00, float c01, float c02, float c10, float c11, float c12,
                                        float c20, float c21, float c22) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float rr = r[i], gg = g[i], bb = b[i];
    r[i] = (c00 * rr + c01 * gg) + c02 * bb;
    g[i] = (c10 * rr + c11 * gg) + c12 * bb;
    b[i] = (c20 * rr + c21 * gg) + c22 * bb;
}

// Pre-filter probes (ghost.cpp:34-103) for every pair: lanes 0-2 the on-axis
// intensity rays at 650/550/450 nm, lanes 3-66 the 8x8 green spread grid.
// CPU-tracer semantics (sensor t >= 0).  out: float4 {x, y, w, status}.
__global__ void filter_probe_kernel(const DevSurf* __restrict__ S, const float* __restrict__ poly,
                                    const float* __restrict__ arena, int n, const int2* __restrict__ pairs,
                                    const float* __restrict__ iat3, const float* __restrict__ ibf3, float front_R,
                                    float z0, float4* __restrict__ out) {
    const int p = blockIdx.x, t = threadIdx.x;
    if (t >= 67) return;
    const int2 pr = pairs[p];
    F3 o, d = f3(0.f, 0.f, 1.f);
    float lam;
    int slot;
    if (t < 3) {
        o = f3(0.f, 0.f, z0);
        lam = t == 0 ? 650.0f : (t == 1 ? 550.0f : 450.0f);
        slot = t;
    } else {
        const int k = t - 3, gx = k % 8, gy = k / 8;
        const float u = (((float)gx + 0.5f) / 8.0f) * 2.0f - 1.0f;
        const float v = (((float)gy + 0.5f) / 8.0f) * 2.0f - 1.0f;
        if (u * u + v * v > 1.0f) { out[(size_t)p * 67 + t] = make_float4(0.f, 0.f, 0.f, -1.0f); return; }
        o = f3(u * front_R, v * front_R, z0);
        lam = 550.0f;
        slot = 1;
    }
    TraceRes r = trace_ghost(as_const(S), n, as_const(poly), arena, pr.x, pr.y, o, d, lam,
                             as_const(iat3) + (size_t)slot * n, as_const(ibf3) + (size_t)slot * n, true);
    out[(size_t)p * 67 + t] = make_float4(r.px, r.py, r.w, (float)r.status);
}

// Single-ray traces for tests and probes.  <0, CC_ALL, false> is the generic
// surface step; the lean levels (1, 2) with a coating class and RT read each
// ray's wavelength tables as the persistent render kernel reads the
// per-lambda ones -- interleaved records (lrec) where kRecForm<CM>, else the
// separate ratio (rat) and coating-row (lrow) tables -- so they run the trace
// a render of this lens runs.
template <int LF, int CM, bool RT>
__global__ void ray_kernel(const DevSurf* __restrict__ S, const float* __restrict__ poly, const float* __restrict__ arena,
                           int n, int a, int b, const float* __restrict__ rays, int n_rays, const float* __restrict__ iat,
                           const float* __restrict__ ibf, const LamRec* __restrict__ lrec,
                           const float2* __restrict__ rat, const int2* __restrict__ lrow, int t_ge0,
                           float* __restrict__ pos, float* __restrict__ w, int* __restrict__ st,
                           int* __restrict__ segs) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_rays) return;
    const float* q = rays + 7 * (size_t)i;
    const size_t row = (size_t)i * n;
    // Per-ray IOR slices: lane-varying, so plain global pointers.
    TraceRes r;
    if (b >= 0 && RT && kRecForm<CM>)
        r = trace_ghost_rec<LF, CM>(as_const(S), n, as_const(poly), arena, a, b, f3(q[0], q[1], q[2]),
                                    f3(q[3], q[4], q[5]), q[6], as_const(lrec + row), t_ge0 != 0);
    else if (b >= 0 && RT)
        r = trace_ghost<true, false, LF, CM, true>(as_const(S), n, as_const(poly), arena, a, b, f3(q[0], q[1], q[2]),
                                                   f3(q[3], q[4], q[5]), q[6], as_const(iat + row),
                                                   as_const(ibf + row), t_ge0 != 0, as_const(lrow + row), nullptr,
                                                   nullptr, as_const(rat + row));
    else
        r = (b >= 0) ? trace_ghost<true, false, LF, CM, false>(as_const(S), n, as_const(poly), arena, a, b,
                                                               f3(q[0], q[1], q[2]), f3(q[3], q[4], q[5]), q[6],
                                                               as_const(iat + row), as_const(ibf + row), t_ge0 != 0)
                          : trace_primary<true, LF>(as_const(S), n, as_const(poly), arena, f3(q[0], q[1], q[2]),
                                                    f3(q[3], q[4], q[5]), q[6], as_const(iat + row), t_ge0 != 0, nullptr);
    pos[3 * i] = r.px; pos[3 * i + 1] = r.py; pos[3 * i + 2] = r.pz;
    w[i] = r.w; st[i] = r.status;
    if (segs) segs[i] = r.segs;
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------

static int upload(gl_ctx* c, const char* name, const void* src, size_t bytes, void** dptr) {
    DevBuf& b = c->buf(name);
    int st = b.ensure(bytes);
    if (st) return st;
    if (bytes) GL_HIP(hipMemcpyAsync(b.p, src, bytes, hipMemcpyHostToDevice, c->stream));
    *dptr = b.p;
    return GL_OK;
}

// Pinned staging (gl_ctx::pin): pin_begin sizes it at the start of a render
// (no transfer of the context is in flight then), pin_take hands out 64-byte
// aligned slices, upload_pinned stages a small upload through one (falling
// back to a pageable copy if the slice does not fit).
static int pin_begin(gl_ctx* c, size_t bytes) {
    c->pin_used = 0;
    if (bytes <= c->pin_cap) return GL_OK;
    std::lock_guard<std::mutex> gm(device_mem_mutex());
    if (c->pin) (void)hipHostFree(c->pin);
    c->pin = nullptr;
    c->pin_cap = 0;
    const size_t cap = std::max<size_t>(bytes, (size_t)64 << 10);
    void* h = nullptr;
    if (hipHostMalloc(&h, cap, hipHostMallocDefault) != hipSuccess || !h)
        return set_error(GL_ERR_NO_MEMORY, "hipHostMalloc of %zu bytes of staging failed", cap);
    c->pin = static_cast<char*>(h);
    c->pin_cap = cap;
    return GL_OK;
}
static void* pin_take(gl_ctx* c, size_t bytes) {
    const size_t off = (c->pin_used + 63) & ~(size_t)63;
    if (!c->pin || off + bytes > c->pin_cap) return nullptr;
    c->pin_used = off + bytes;
    return c->pin + off;
}
static int upload_pinned(gl_ctx* c, const char* name, const void* src, size_t bytes, void** dptr) {
    void* h = bytes ? pin_take(c, bytes) : nullptr;
    if (!h) return upload(c, name, src, bytes, dptr);
    DevBuf& b = c->buf(name);
    int st = b.ensure(bytes);
    if (st) return st;
    std::memcpy(h, src, bytes);
    GL_HIP(hipMemcpyAsync(b.p, h, bytes, hipMemcpyHostToDevice, c->stream));
    *dptr = b.p;
    return GL_OK;
}

template <class T> static void key_append(std::vector<uint8_t>& k, const T& v) {
    const uint8_t* p = reinterpret_cast<const uint8_t*>(&v);
    k.insert(k.end(), p, p + sizeof(T));
}

static SrcDev make_src(const gl_source& s) {
    SrcDev d;
    d.bx = (float)std::tan((double)s.angle_x);
    d.by = (float)std::tan((double)s.angle_y);
    // normalize(bx, by, 1) with the tracer's float operation order (vec3.h:67-71).
    float l = std::sqrt(d.bx * d.bx + d.by * d.by + 1.0f * 1.0f);
    float inv = (l > 1e-12f) ? 1.0f / l : 0.0f;
    d.dx = d.bx * inv; d.dy = d.by * inv; d.dz = 1.0f * inv;
    d.r = s.r; d.g = s.g; d.b = s.b;
    uint32_t ax, ay;
    std::memcpy(&ax, &s.angle_x, 4);
    std::memcpy(&ay, &s.angle_y, 4);
    d.hseed = ax * 0x9E3779B9u + ay * 0x85EBCA6Bu;
    return d;
}

// Lanes per sample for the lambda-parallel probe: the smallest power of two
// holding every wavelength, or 0 (per-lane loop) above 64.
static int lam_group(int ns) {
    if (ns > 64) return 0;
    int g = 1;
    while (g < ns) g <<= 1;
    return g;
}

// lean: the lens's lean level (no HURB launch takes a lean form).
template <bool STATS, bool SPLAT, bool COUNT>
static int launch(const GhostArgs& A_in, int n_ps, int n_samples, hipStream_t s, int lean) {
    if (n_ps <= 0 || n_samples <= 0) return GL_OK;
    GhostArgs A = A_in;
    A.lam_group = STATS ? lam_group(A.n_spec) : 0;
    const int per_block = A.lam_group > 0 ? kBlock / A.lam_group : kBlock;
    const long long chunks = (n_samples + per_block - 1) / per_block;
    // Probe passes: enough chunks per workgroup that the launch holds about
    // kProbeBlocks workgroups (the counters and boxes then take one set of
    // atomics per wave for all of them); other passes one chunk each.
    int iters = 1;
    if (STATS) {
        const long long want = ((long long)n_ps * chunks + kProbeBlocks - 1) / kProbeBlocks;
        iters = (int)std::max(1ll, std::min(chunks, want));
    }
    A.probe_iters = iters;
    dim3 grid((unsigned)n_ps, (unsigned)((chunks + iters - 1) / iters), 1);
    if (A.hurb.on) hipLaunchKernelGGL((ghost_kernel<STATS, SPLAT, COUNT, true>), grid, dim3(kBlock), 0, s, A);
    else if (lean == 1) hipLaunchKernelGGL((ghost_kernel<STATS, SPLAT, COUNT, false, 1>), grid, dim3(kBlock), 0, s, A);
    else if (lean == 2) hipLaunchKernelGGL((ghost_kernel<STATS, SPLAT, COUNT, false, 2>), grid, dim3(kBlock), 0, s, A);
    else hipLaunchKernelGGL((ghost_kernel<STATS, SPLAT, COUNT, false>), grid, dim3(kBlock), 0, s, A);
    GL_HIP(hipGetLastError());
    return GL_OK;
}

// The interleaved per-(lambda, surface) records (LamRec) from the separate
// tables: lrows is lambda_rows' output, the iat.size() entries followed by the
// coating rows.  The rows follow the records here, and each entry's
// self-relative row offset is re-based to its record's lrow field (word
// 8q + 4 of the record array; rows start on a whole record, so they stay
// 8-byte aligned).
static std::vector<LamRec> lam_records(const std::vector<float>& iat, const std::vector<float>& ibf,
                                       const std::vector<float2>& rats, const std::vector<int32_t>& lrows) {
    const size_t m = iat.size();
    const size_t row_words = lrows.size() - 2 * m;
    std::vector<LamRec> out(m + (row_words + 7) / 8);
    for (size_t q = 0; q < m; ++q) {
        LamRec& r = out[q];
        r.iat = iat[q]; r.ibf = ibf[q]; r.rfw = rats[q].x; r.rbw = rats[q].y;
        int2 e = make_int2(lrows[2 * q], lrows[2 * q + 1]);
        if (e.x >= 0) {      // a row: old word 2q + e.x, past the 2m entry words
            const size_t row = 8 * m + (2 * q + (size_t)e.x - 2 * m);
            e.x = (int)(row - (8 * q + 4));
        }
        r.lrow = e;
        r.pad = make_int2(0, 0);
    }
    std::memcpy(reinterpret_cast<int32_t*>(out.data()) + 8 * m, lrows.data() + 2 * m, row_words * sizeof(int32_t));
    return out;
}

struct Launch {
    gl_ctx* c;
    gl_lens* L;
    const gl_flare_params* p;
    std::vector<gl_pair> pairs;
    std::vector<SrcDev> src;
    int n_src, width, height, fmt_w, fmt_h, fx0, fy0;
    float sw, sh;
    // PER_PAIR AOV: one output frame per pair, and the pairs are rendered as
    // the reference's single-pair launches are (no probe, cull or
    // concentration, ghost_render.cu:1540-1553 with n_pairs = 1).
    bool frame_by_pair = false;
};

// Host caches: pupil grid (GridKey) and spectral table (SpecKey).
static int prepare_tables(gl_ctx* c, const gl_lens* L, const gl_flare_params* p) {
    std::vector<uint8_t> gk;
    key_append(gk, p->ray_grid); key_append(gk, p->pupil_jitter); key_append(gk, p->jitter_seed);
    key_append(gk, p->aperture_blades); key_append(gk, p->aperture_rotation);
    for (const auto& s : L->host)
        if (s.is_stop) {
            key_append(gk, s.aperture_shape); key_append(gk, s.aperture_blades);
            key_append(gk, s.aperture_aspect); key_append(gk, s.aperture_rotation_rad);
            break;
        }
    if (gk != c->grid_key) {
        c->grid = build_samples(L->host.data(), (int)L->host.size(), p->ray_grid, p->pupil_jitter, p->jitter_seed,
                                p->aperture_blades, p->aperture_rotation);
        c->grid_key = gk;
        c->grid_uploaded = false;
    }
    const int ns = std::max(3, p->spectral_samples);
    std::vector<uint8_t> sk;
    key_append(sk, ns); key_append(sk, p->sensor_model); key_append(sk, p->output_cs);
    key_append(sk, p->spectral_correction); key_append(sk, p->custom_xyz_to_output);
    if (sk != c->spec_key) {
        int st = spectral_table(ns, p->output_cs, p->custom_xyz_to_output, c->spec);
        if (st) return st;
        c->spec_do_correct = (ns != 3) && p->spectral_correction;
        if (c->spec_do_correct) spectral_correction(c->spec, p->output_cs, p->custom_xyz_to_output, c->spec_C);
        c->spec_key = sk;
        c->spec_uploaded = false;
    }
    return GL_OK;
}

static int do_launch(Launch& Lc, float* out_r, float* out_g, float* out_b, float** dev_out, gl_ghost_stats* stats) {
    gl_ctx* c = Lc.c;
    gl_lens* L = Lc.L;
    const gl_flare_params* p = Lc.p;
    const double t_fn0 = now_ms();
    const int n = (int)L->host.size();
    const int n_pairs = (int)Lc.pairs.size(), n_src = Lc.n_src;
    int st;
    const double t_prep0 = now_ms();
    if ((st = prepare_tables(c, L, p))) return st;
    const int n_grid = (int)c->grid.size();
    const int ns = (int)c->spec.size();
    const double t_prep1 = now_ms();

    // Frame mode: one output frame per source (or per pair), frame-major planes.
    const int frames = Lc.frame_by_pair ? std::max(1, n_pairs) : (p->frame_per_source ? std::max(1, n_src) : 1);
    const size_t npx = (size_t)Lc.width * Lc.height * (size_t)frames;
    DevBuf &bo_r = c->buf("out_r"), &bo_g = c->buf("out_g"), &bo_b = c->buf("out_b");
    if ((st = bo_r.ensure(npx * 4)) || (st = bo_g.ensure(npx * 4)) || (st = bo_b.ensure(npx * 4))) return st;
    auto zero_out = [&]() -> int {
        hipLaunchKernelGGL(fill3_kernel, dim3((unsigned)((npx + 255) / 256)), dim3(256), 0, c->stream, bo_r.as<float>(),
                           bo_g.as<float>(), bo_b.as<float>(), npx);
        GL_HIP(hipGetLastError());
        return GL_OK;
    };
    if ((st = zero_out())) return st;
    if (dev_out) { dev_out[0] = bo_r.as<float>(); dev_out[1] = bo_g.as<float>(); dev_out[2] = bo_b.as<float>(); }
    if (stats) {
        stats->n_pairs = n_pairs; stats->n_sources = n_src; stats->n_grid = n_grid; stats->n_spec = ns;
        stats->ms_grid_build = (float)(t_prep1 - t_prep0);
    }
    auto finish_empty = [&]() -> int {
        if (out_r) {
            GL_HIP(hipMemcpyAsync(out_r, bo_r.p, npx * 4, hipMemcpyDeviceToHost, c->stream));
            GL_HIP(hipMemcpyAsync(out_g, bo_g.p, npx * 4, hipMemcpyDeviceToHost, c->stream));
            GL_HIP(hipMemcpyAsync(out_b, bo_b.p, npx * 4, hipMemcpyDeviceToHost, c->stream));
        }
        GL_HIP(hipStreamSynchronize(c->stream));
        return GL_OK;
    };
    if (n_pairs == 0 || n_src == 0 || n_grid == 0) return finish_empty();

    const double t_up0 = now_ms();
    const int n_ps = n_pairs * n_src;
    {
        const size_t slack = 64 * 12;
        const size_t pin_bytes = (size_t)n_pairs * sizeof(gl_pair) + (size_t)n_src * sizeof(SrcDev) +
                                 3 * (size_t)n_pairs + (size_t)n_ps * (sizeof(float4) + sizeof(int2) + 5 * 4) +
                                 3 * (size_t)n_pairs * 8 + 2 * 64 + slack;
        if ((st = pin_begin(c, pin_bytes))) return st;
    }
    LensDev D;
    if ((st = lens_sync(L, p->fp64_radius_min, c->stream, &D))) return st;
    void *d_pairs, *d_src, *d_grid = c->buf("grid").p, *d_spec = c->buf("spec").p;
    if ((st = upload_pinned(c, "pairs", Lc.pairs.data(), n_pairs * sizeof(gl_pair), &d_pairs))) return st;
    if ((st = upload_pinned(c, "src", Lc.src.data(), n_src * sizeof(SrcDev), &d_src))) return st;
    // Tables of this (lens, spectral table): per-lambda indices after / before
    // each surface, coating lookup rows, refraction ratios (forward ibf / iat,
    // backward iat / ibf: the same IEEE quotients the tracer would form) and
    // their 3-lambda slices for the refine probe.  Built and uploaded only
    // when the lens or the spectral table changes.
    const int rns = std::min(3, ns);
    const int li[3] = {0, ns / 2, ns - 1};
    DevBuf &b_iat = c->buf("iat"), &b_ibf = c->buf("ibf"), &b_lrow = c->buf("lrow"), &b_rat = c->buf("rat");
    DevBuf& b_lrec = c->buf("lrec");
    DevBuf &b_s3 = c->buf("spec3"), &b_a3 = c->buf("iat3"), &b_b3 = c->buf("ibf3"), &b_r3 = c->buf("rat3");
    {
        std::vector<uint8_t> tk;
        key_append(tk, L->uid);
        tk.insert(tk.end(), c->spec_key.begin(), c->spec_key.end());
        if (tk != c->ltab_key || !b_iat.p || !b_ibf.p || !b_lrow.p || !b_rat.p || !b_lrec.p || !b_s3.p || !b_a3.p ||
            !b_b3.p || !b_r3.p) {
            c->ltab_key.clear();
            std::vector<float> iat((size_t)ns * n), ibf((size_t)ns * n);
            ior_tables(L->host.data(), n, c->spec.data(), ns, iat.data(), ibf.data());
            std::vector<int32_t> lrows;
            lambda_rows(L->host.data(), n, L->arena.data(), c->spec.data(), ns, lrows);
            std::vector<float2> rats((size_t)ns * n);
            for (size_t q = 0; q < rats.size(); ++q) rats[q] = make_float2(ibf[q] / iat[q], iat[q] / ibf[q]);
            const std::vector<LamRec> lrec = lam_records(iat, ibf, rats, lrows);
            std::vector<float4> s3(rns);
            std::vector<float> a3((size_t)rns * n), b3((size_t)rns * n);
            std::vector<float2> r3((size_t)rns * n);
            for (int k = 0; k < rns; ++k) {
                const gl_spec& q = c->spec[li[k]];
                s3[k] = make_float4(q.lambda, q.cmf_r, q.cmf_g, q.cmf_b);
                std::copy_n(&iat[(size_t)li[k] * n], n, &a3[(size_t)k * n]);
                std::copy_n(&ibf[(size_t)li[k] * n], n, &b3[(size_t)k * n]);
                std::copy_n(&rats[(size_t)li[k] * n], n, &r3[(size_t)k * n]);
            }
            void* d;
            if ((st = upload(c, "iat", iat.data(), iat.size() * 4, &d)) ||
                (st = upload(c, "ibf", ibf.data(), ibf.size() * 4, &d)) ||
                (st = upload(c, "lrow", lrows.data(), lrows.size() * sizeof(int32_t), &d)) ||
                (st = upload(c, "rat", rats.data(), rats.size() * sizeof(float2), &d)) ||
                (st = upload(c, "lrec", lrec.data(), lrec.size() * sizeof(LamRec), &d)) ||
                (st = upload(c, "spec3", s3.data(), rns * sizeof(float4), &d)) ||
                (st = upload(c, "iat3", a3.data(), a3.size() * 4, &d)) ||
                (st = upload(c, "ibf3", b3.data(), b3.size() * 4, &d)) ||
                (st = upload(c, "rat3", r3.data(), r3.size() * sizeof(float2), &d)))
                return st;
            c->ltab_key = tk;
        }
    }
    void *d_iat = b_iat.p, *d_ibf = b_ibf.p, *d_lrow = b_lrow.p, *d_rat = b_rat.p;
    if (!c->grid_uploaded || !d_grid || c->buf("grid").cap < n_grid * sizeof(gl_sample)) {
        if ((st = upload(c, "grid", c->grid.data(), n_grid * sizeof(gl_sample), &d_grid))) return st;
        c->grid_uploaded = true;
    }
    if (!c->spec_uploaded || !d_spec || c->buf("spec").cap < ns * sizeof(gl_spec)) {
        if ((st = upload(c, "spec", c->spec.data(), ns * sizeof(gl_spec), &d_spec))) return st;
        c->spec_uploaded = true;
    }
    const double t_up1 = now_ms();

    GhostArgs A;
    std::memset(&A, 0, sizeof A);           // padding too: the bytes key the captured graph
    A.S = D.surf; A.poly = D.poly; A.arena = D.arena; A.n = n;
    A.pairs = static_cast<const gl_pair*>(d_pairs);
    A.n_src = n_src; A.src = static_cast<const SrcDev*>(d_src);
    A.grid = static_cast<const gl_sample*>(d_grid); A.n_grid = n_grid;
    A.front_R = L->host[0].semi_aperture; A.start_z = L->host[0].z - 20.0f;
    A.sw = Lc.sw; A.sh = Lc.sh; A.width = Lc.width; A.height = Lc.height;
    A.fmt_w = (float)Lc.fmt_w; A.fmt_h = (float)Lc.fmt_h; A.x0 = (float)Lc.fx0; A.y0 = (float)Lc.fy0;
    A.out_r = bo_r.as<float>(); A.out_g = bo_g.as<float>(); A.out_b = bo_b.as<float>();
    A.gain = p->flare_gain; A.ray_weight = 1.0f / (float)n_grid;
    A.spec = static_cast<const float4*>(d_spec); A.n_spec = ns;
    A.iat = static_cast<const float*>(d_iat); A.ibf = static_cast<const float*>(d_ibf);
    A.lrow = static_cast<const int2*>(d_lrow);
    A.rat = static_cast<const float2*>(d_rat);
    A.lrec = static_cast<const LamRec*>(b_lrec.p);
    A.conc.rects = nullptr;
    A.hurb.on = p->hurb ? 1 : 0;
    A.hurb.dist = p->hurb_kick;
    A.hurb.min_sigma = p->hurb_min_sigma_rad;
    A.hurb.max_kick = p->hurb_max_kick_rad;
    A.hurb_seed_base = (uint32_t)p->jitter_seed * 2654435761u + 1u;
    A.baffles.n = std::max(0, std::min((int)p->n_baffles, GL_MAX_BAFFLES));
    for (int k = 0; k < A.baffles.n; ++k) A.baffles.b[k] = p->baffles[k];

    // ---- probe / cull / concentrate (ghost_render.cu:969-1238)
    const bool probes_allowed = n_pairs > 1 && ns != 3 && !Lc.frame_by_pair;
    const bool want_conc = p->concentrate_samples && probes_allowed;
    bool cull_ready = false, rect_ready = false;
    std::vector<float4> rects;
    std::vector<unsigned char> active(n_pairs, 1);
    if ((p->cull_dead_pairs || want_conc) && probes_allowed) {
        DevBuf &bc = c->buf("probe_cnt"), &bb = c->buf("probe_bbox");
        DevBuf &bm = c->buf("pair_mask"), &brf = c->buf("refine_mask"), &bnr = c->buf("refine_n");
        if ((st = bc.ensure(3 * n_pairs * 8)) || (st = bb.ensure((size_t)n_ps * 5 * 4)) || (st = bm.ensure(n_pairs)) ||
            (st = brf.ensure(n_ps)) || (st = bnr.ensure(4)))
            return st;
        // Probe sample sets of this lens: three jittered 40^2 passes (stored
        // one after another) and the 128^2 refine grid, built on the host
        // once per (lens, blade override, rotation) and kept on the device.
        DevBuf &bpa = c->buf("probe_all"), &brg = c->buf("refine_grid");
        {
            std::vector<uint8_t> pk;
            key_append(pk, L->uid); key_append(pk, p->aperture_blades); key_append(pk, p->aperture_rotation);
            if (pk != c->probe_key || !bpa.p || !brg.p || (int)c->probe_n.size() != kProbePasses) {
                c->probe_key.clear();
                std::vector<gl_sample> all;
                c->probe_n.assign(kProbePasses, 0);
                for (int pass = 0; pass < kProbePasses; ++pass) {
                    const std::vector<gl_sample> probe = build_samples(L->host.data(), n, kProbeGrid, 1,
                                                                       (int)(0x9E3779B9u * (unsigned)(pass + 1)),
                                                                       p->aperture_blades, p->aperture_rotation);
                    c->probe_n[pass] = (int)probe.size();
                    all.insert(all.end(), probe.begin(), probe.end());
                }
                const std::vector<gl_sample> rp = build_samples(L->host.data(), n, kRefineGrid, 1, (int)0xB5297A4Du,
                                                                p->aperture_blades, p->aperture_rotation);
                c->refine_n = (int)rp.size();
                void* d;
                if ((st = upload(c, "probe_all", all.data(), std::max<size_t>(1, all.size()) * sizeof(gl_sample), &d)) ||
                    (st = upload(c, "refine_grid", rp.data(), std::max<size_t>(1, rp.size()) * sizeof(gl_sample), &d)))
                    return st;
                c->probe_key = pk;
            }
        }
        GL_HIP(hipMemsetAsync(bc.p, 0, 3 * n_pairs * 8, c->stream));
        if (want_conc) {
            GL_HIP(hipMemsetAsync(bb.p, 0x7F, (size_t)n_ps * 2 * 4, c->stream));
            GL_HIP(hipMemsetAsync(bb.as<int>() + 2 * n_ps, 0, (size_t)n_ps * 3 * 4, c->stream));
        }
        GhostArgs P = A;
        P.sw = Lc.sw * kCullMargin; P.sh = Lc.sh * kCullMargin;
        P.cnt = bc.as<unsigned long long>();
        P.bbox = want_conc ? bb.as<int>() : nullptr;
        P.segs = nullptr;
        P.ps_refine = nullptr;
        std::vector<unsigned long long> cnt(3 * n_pairs);
        bool pass0_ok = false;
        const gl_sample* probe_all = bpa.as<gl_sample>();
        if (c->probe_n[0] > 0 && want_conc && !A.hurb.on) {
            // With concentration every pass probes every pair, and the pass
            // statistics are integer sums, minima and maxima: the three
            // passes run as one launch over their concatenated samples with
            // the same counters.  (HURB seeds its kicks with the sample index,
            // so HURB probes keep one launch per pass.)
            int total = 0;
            for (int pass = 0; pass < kProbePasses; ++pass) total += c->probe_n[pass];
            P.grid = probe_all; P.n_grid = total;
            P.pair_active = nullptr;
            if ((st = launch<true, false, false>(P, n_ps, P.n_grid, c->stream, D.lean))) return st;
            pass0_ok = true;
        } else {
            std::vector<unsigned char> cand(n_pairs, 0);
            size_t off = 0;
            for (int pass = 0; pass < kProbePasses; ++pass) {
                const int np = c->probe_n[pass];
                if (np == 0) { if (pass == 0) break; else continue; }
                P.grid = probe_all + off; P.n_grid = np;
                off += (size_t)np;
                P.pair_active = (pass == 0 || want_conc) ? nullptr : bm.as<unsigned char>();
                if ((st = launch<true, false, false>(P, n_ps, P.n_grid, c->stream, D.lean))) return st;
                if (pass == 0) {
                    pass0_ok = true;
                    if (!want_conc) {
                        GL_HIP(d2h_sync(c, cnt.data(), bc.p, 3 * n_pairs * 8));
                        int nc = 0;
                        for (int q = 0; q < n_pairs; ++q)
                            if (cnt[q] > 0 && cnt[2 * n_pairs + q] == 0) { cand[q] = 1; ++nc; }
                        if (nc == 0) break;
                        void* d;
                        if ((st = upload_pinned(c, "pair_mask", cand.data(), n_pairs, &d))) return st;
                    }
                }
            }
        }
        if (want_conc && pass0_ok && c->refine_n > 0) {
            // Refine the (pair, source) slots with few hits at three
            // wavelengths on the 128^2 grid; the slot mask and its count are
            // formed on the device, and the refine pass runs only when 1 to
            // kRefineMaxEntries slots qualify (checked by its workgroups).
            hipLaunchKernelGGL(refine_mask_kernel, dim3(1), dim3(1024), 0, c->stream, bb.as<int>() + 4 * n_ps, n_ps,
                               kRefineHitsMax, brf.as<unsigned char>(), bnr.as<unsigned>());
            GL_HIP(hipGetLastError());
            GhostArgs R = P;
            R.grid = brg.as<gl_sample>(); R.n_grid = c->refine_n;
            R.spec = b_s3.as<float4>(); R.n_spec = rns;
            R.iat = b_a3.as<float>(); R.ibf = b_b3.as<float>();
            R.rat = b_r3.as<float2>();
            R.pair_active = nullptr; R.ps_refine = brf.as<unsigned char>();
            R.refine_n = bnr.as<unsigned>();
            if ((st = launch<true, false, false>(R, n_ps, R.n_grid, c->stream, D.lean))) return st;
        }
        if (pass0_ok) {
            std::vector<int> bbh(want_conc ? (size_t)n_ps * 5 : 0);
            void* h_cnt = pin_take(c, 3 * (size_t)n_pairs * 8);
            void* h_bbh = want_conc ? pin_take(c, (size_t)n_ps * 5 * 4) : nullptr;
            GL_HIP(hipMemcpyAsync(h_cnt ? h_cnt : cnt.data(), bc.p, 3 * n_pairs * 8, hipMemcpyDeviceToHost, c->stream));
            if (want_conc)
                GL_HIP(hipMemcpyAsync(h_bbh ? h_bbh : bbh.data(), bb.p, (size_t)n_ps * 5 * 4, hipMemcpyDeviceToHost,
                                      c->stream));
            GL_HIP(hipStreamSynchronize(c->stream));
            if (h_cnt) std::memcpy(cnt.data(), h_cnt, 3 * (size_t)n_pairs * 8);
            if (h_bbh) std::memcpy(bbh.data(), h_bbh, (size_t)n_ps * 5 * 4);
            for (int q = 0; q < n_pairs; ++q) active[q] = cnt[2 * n_pairs + q] > 0 ? 1 : 0;
            if (want_conc) {
                rects.assign(n_ps, make_float4(2.f, 2.f, -2.f, -2.f));
                const float dil = kConcDilateCells * (2.0f / (float)kProbeGrid);
                auto dec = [](int i) { float f; std::memcpy(&f, &i, 4); return f - 2.0f; };
                for (int i = 0; i < n_ps; ++i) {
                    if (bbh[(size_t)4 * n_ps + i] < kConcMinHits) continue;
                    float u0 = dec(bbh[2 * (size_t)i]), v0 = dec(bbh[2 * (size_t)i + 1]);
                    float u1 = dec(bbh[2 * (size_t)n_ps + 2 * i]), v1 = dec(bbh[2 * (size_t)n_ps + 2 * i + 1]);
                    if (!(u0 <= u1 && v0 <= v1)) continue;
                    u0 = std::max(-1.0f, u0 - dil); v0 = std::max(-1.0f, v0 - dil);
                    u1 = std::min(1.0f, u1 + dil); v1 = std::min(1.0f, v1 + dil);
                    rects[i] = make_float4(u0, v0, u1, v1);
                }
                if (stats && stats->ps_cap >= n_ps) {
                    stats->n_ps = n_ps;
                    for (int i = 0; i < n_ps; ++i) {
                        if (stats->ps_hits) stats->ps_hits[i] = bbh[(size_t)4 * n_ps + i];
                        if (stats->ps_rect) {
                            stats->ps_rect[4 * i] = rects[i].x; stats->ps_rect[4 * i + 1] = rects[i].y;
                            stats->ps_rect[4 * i + 2] = rects[i].z; stats->ps_rect[4 * i + 3] = rects[i].w;
                        }
                    }
                }
                rect_ready = true;
            }
        }
        void* d;
        if ((st = upload_pinned(c, "pair_mask", active.data(), n_pairs, &d))) return st;
        cull_ready = p->cull_dead_pairs != 0;
    }

    // ---- concentration + adaptive budgets (ghost_render.cu:1240-1304)
    std::vector<int2> budgets;
    if (rect_ready) {
        void* d_rects;
        if ((st = upload_pinned(c, "rects", rects.data(), n_ps * sizeof(float4), &d_rects))) return st;
        A.conc.rects = static_cast<const float4*>(d_rects);
        A.conc.mask = build_mask(L->host.data(), n, p->aperture_blades, p->aperture_rotation);
        A.conc.jitter = p->pupil_jitter;
        A.conc.seed_offset = (uint32_t)p->jitter_seed * 1000003u;
        const int nr = std::max(1, (int)std::floor(std::sqrt((double)n_grid)));
        A.conc.nr_lattice = nr;
        A.conc.n_r = (p->pupil_jitter == 2) ? n_grid : nr * nr;
        const float N = (float)p->ray_grid;
        const float a_ref = 4.0f * (float)n_grid / (N * N);
        A.conc.area_norm = 1.0f / (a_ref * (float)A.conc.n_r);
        A.conc.budgets = nullptr;
        if (p->adaptive_sample_budgets) {
            budgets.assign(n_ps, make_int2(A.conc.n_r, A.conc.nr_lattice));
            const float boost = std::max(1.0f, p->adaptive_density_boost);
            int floor_s = std::max(256, p->adaptive_min_samples);
            floor_s = ((floor_s + 255) / 256) * 256;
            const int floor_nr = std::max(1, (int)std::lround(std::sqrt((double)floor_s)));
            for (int i = 0; i < n_ps; ++i) {
                const float4 r = rects[i];
                if (!(r.x <= r.z)) continue;
                const float area = (r.z - r.x) * (r.w - r.y);
                const float raw = boost * (float)n_grid * area / a_ref;
                if (p->pupil_jitter == 2) {
                    int b = (int)std::ceil((double)raw);
                    b = ((b + 255) / 256) * 256;
                    b = std::max(floor_s, std::min(b, n_grid));
                    budgets[i] = make_int2(b, A.conc.nr_lattice);
                } else {
                    int nrp = (int)std::lround(std::sqrt((double)std::max(0.0f, raw)));
                    nrp = std::max(floor_nr, std::min(nrp, A.conc.nr_lattice));
                    budgets[i] = make_int2(nrp * nrp, nrp);
                }
            }
            void* d_b;
            if ((st = upload_pinned(c, "budgets", budgets.data(), n_ps * sizeof(int2), &d_b))) return st;
            A.conc.budgets = static_cast<const int2*>(d_b);
            A.conc.inv_aref = 1.0f / a_ref;
            if (stats && stats->ps_cap >= n_ps && stats->ps_budget)
                for (int i = 0; i < n_ps; ++i) stats->ps_budget[i] = budgets[i].x;
        }
    }
    A.pair_active = cull_ready ? c->buf("pair_mask").as<unsigned char>() : nullptr;

    if (p->verbose) {
        std::printf("ghostlight_amd: %d pairs x %d sources x %d samples -> grid (%d, %d)\n", n_pairs, n_src, n_grid,
                    n_ps, (n_grid + kBlock - 1) / kBlock);
        std::fflush(stdout);
    }
    // Deposit-record capacity: the previous render's count with headroom,
    // bounded by the number of (sample, lambda) traces this launch can make.
    unsigned long long max_rec = 0;
    for (int i = 0; i < n_ps; ++i) {
        unsigned long long ns_i = (unsigned long long)n_grid;
        if (rect_ready && rects[i].x <= rects[i].z) ns_i = budgets.empty() ? (unsigned long long)A.conc.n_r
                                                                           : (unsigned long long)budgets[i].x;
        max_rec += ns_i;
    }
    max_rec *= (unsigned long long)ns;
    const SplatGeom sg = splat_geom(Lc.width, Lc.height, frames);
    A.frame_mode = frames > 1 ? (Lc.frame_by_pair ? 2 : 1) : 0;
    unsigned long long want = std::max<unsigned long long>(c->splat_hint + c->splat_hint / 4, 1ull << 22);
    want = std::max<unsigned long long>(1, std::min(want, max_rec));
    // ---- launch form: grid, or persistent over a cost-ordered work list
    // (schedule from the autotuner), optionally replayed from a hipGraph.
    // HURB traces run one grid launch: the persistent kernel has no kick path.
    const bool persistent = p->batch_mode != GL_BATCH_GRID && !p->hurb;
    const bool use_graph = p->batch_mode == GL_BATCH_GRAPH;
    int cand = -1;
    bool trial = false;
    gl_ctx::TuneState* ts = nullptr;
    PersistQ Q;
    std::memset(&Q, 0, sizeof Q);
    int workers = 0, lam_grp = ns;
    if (persistent) {
        if (c->n_cu <= 0) {
            GL_HIP(hipDeviceGetAttribute(&c->n_cu, hipDeviceAttributeMultiprocessorCount, 0));
            GL_HIP(hipDeviceGetAttribute(&c->wall_khz, hipDeviceAttributeWallClockRate, 0));
            if (c->n_cu <= 0) c->n_cu = 1;
            if (c->wall_khz <= 0) c->wall_khz = 100000;
        }
        std::vector<uint8_t> tk;
        key_append(tk, n_ps); key_append(tk, ns); key_append(tk, n_grid); key_append(tk, Lc.width);
        key_append(tk, Lc.height); key_append(tk, n); key_append(tk, rect_ready);
        ts = &c->tune[tk];
        if (ts->ms.empty()) ts->ms.assign(kNumSched, -1.0f);
        if (p->schedule >= 0 && p->schedule < kNumSched) cand = p->schedule;
        else {
            // Trial order: a lens whose coatings read coating rows (the
            // interleaved-record form) starts from few wavelengths a pass, one
            // without them from all of them on half the workers (which leaves
            // room for a concurrent render on another stream) -- the schedule
            // a one-off shape (a render never repeated) keeps.
            static constexpr int kOrderRec[kNumSched] = {4, 3, 5, 2, 1, 0};
            const bool rec_order = D.coat_classes != (CC_AIRY | CC_FRESNEL);
            for (int q = 0; q < kNumSched && cand < 0; ++q) {
                const int k = rec_order ? kOrderRec[q] : q;
                if (ts->ms[k] < 0.0f) { cand = k; trial = true; }
            }
            if (cand < 0) cand = (int)(std::min_element(ts->ms.begin(), ts->ms.end()) - ts->ms.begin());
        }
        const Sched& sc = kSched[cand];
        lam_grp = (ns + sc.lam_div - 1) / sc.lam_div;
        workers = c->n_cu * sc.blocks_per_cu;
        // Work items {slot, chunk, lambda range}; cost ~ surfaces per trace x
        // rays.  Rebuilt only when its inputs change.
        std::vector<uint8_t> wk;
        key_append(wk, cand); key_append(wk, ns); key_append(wk, n_grid); key_append(wk, n_src); key_append(wk, n);
        key_append(wk, cull_ready); key_append(wk, rect_ready); key_append(wk, A.conc.n_r);
        wk.insert(wk.end(), active.begin(), active.end());
        for (const auto& pq : Lc.pairs) { key_append(wk, pq.surf_a); key_append(wk, pq.surf_b); }
        if (rect_ready) for (const auto& r : rects) key_append(wk, r);
        for (const auto& b : budgets) key_append(wk, b);
        DevBuf &bw = c->buf("persist_work"), &bq = c->buf("persist_ctl");
        if ((st = bq.ensure(64))) return st;
        if (wk != c->work_key || !bw.p) {
            std::vector<int4> work;
            std::vector<float> cost;
            for (int q = 0; q < n_pairs; ++q) {
                if (cull_ready && !active[q]) continue;
                const gl_pair& pq = Lc.pairs[q];
                if (pq.surf_a < 0 || pq.surf_b <= pq.surf_a || pq.surf_b >= n) continue;
                for (int si = 0; si < n_src; ++si) {
                    const int ps = q * n_src + si;
                    int ns_ps = n_grid;
                    if (rect_ready && rects[ps].x <= rects[ps].z) ns_ps = budgets.empty() ? A.conc.n_r : budgets[ps].x;
                    const int chunks = (ns_ps + kBlock - 1) / kBlock;
                    for (int ch = 0; ch < chunks; ++ch)
                        for (int l0 = 0; l0 < ns; l0 += lam_grp) {
                            const int l1 = std::min(ns, l0 + lam_grp);
                            work.push_back(make_int4(ps, ch, l0, l1));
                            const int rays = std::min(kBlock, ns_ps - ch * kBlock);
                            cost.push_back((float)(n + 2 * (pq.surf_b - pq.surf_a)) * (float)rays * (float)(l1 - l0));
                        }
                }
            }
            if (sc.longest_first) {
                std::vector<int> idx(work.size());
                for (size_t i = 0; i < idx.size(); ++i) idx[i] = (int)i;
                std::stable_sort(idx.begin(), idx.end(), [&](int x, int y) { return cost[x] > cost[y]; });
                std::vector<int4> w2(work.size());
                for (size_t i = 0; i < idx.size(); ++i) w2[i] = work[idx[i]];
                work.swap(w2);
            }
            if ((st = bw.ensure(std::max<size_t>(1, work.size()) * sizeof(int4)))) return st;
            if (!work.empty())
                GL_HIP(hipMemcpyAsync(bw.p, work.data(), work.size() * sizeof(int4), hipMemcpyHostToDevice, c->stream));
            const unsigned ctl[2] = {0u, (unsigned)work.size()};
            GL_HIP(hipMemcpyAsync(bq.p, ctl, sizeof ctl, hipMemcpyHostToDevice, c->stream));
            c->work_host.swap(work);
            c->work_key.swap(wk);
        }
        Q.work = bw.as<int4>();
        Q.ctl = bq.as<unsigned>();
        Q.clk = reinterpret_cast<unsigned long long*>(bq.as<char>() + 16);
        c->last_sched[5] = (int)c->work_host.size();
    }

    const double t_k0 = now_ms();
    bool graphed = false;
    for (int attempt = 0; attempt < 2; ++attempt) {
        SplatBufs sb;
        std::memset(&sb, 0, sizeof sb);
        if ((st = splat_prepare(c, sg, want, sb))) return st;
        DevBuf& b_sum = c->buf("render_summary");
        if ((st = b_sum.ensure(64))) return st;
        A.rec = sb.rec; A.rec_cap = sb.rec_cap; A.rec_count = sb.rec_count;
        // The whole device sequence of one render after the probes: no host
        // synchronisation and no memset nodes, so it can be captured.
        auto enqueue = [&](bool events) -> int {
            int e;
            if ((e = zero_out())) return e;
            if ((e = splat_clear(c->stream, sg, sb))) return e;
            if (persistent) hipLaunchKernelGGL(persist_reset_kernel, dim3(1), dim3(1), 0, c->stream, Q);
            if (events) GL_HIP(hipEventRecord(c->ev0, c->stream));
            if (persistent) {
                const dim3 gw((unsigned)workers), bw(kBlock);
                if (D.lean == 1 && D.coat_classes == (CC_SA | CC_FRESNEL))
                    hipLaunchKernelGGL((ghost_persist_kernel<1, CC_SA | CC_FRESNEL>), gw, bw, 0, c->stream, A, Q);
                else if (D.lean == 1 && D.coat_classes == (CC_AIRY | CC_FRESNEL))
                    hipLaunchKernelGGL((ghost_persist_kernel<1, CC_AIRY | CC_FRESNEL>), gw, bw, 0, c->stream, A, Q);
                else if (D.lean == 1)
                    hipLaunchKernelGGL((ghost_persist_kernel<1, CC_ALL>), gw, bw, 0, c->stream, A, Q);
                else if (D.lean == 2 && D.coat_classes == (CC_SA | CC_FRESNEL))
                    hipLaunchKernelGGL((ghost_persist_kernel<2, CC_SA | CC_FRESNEL>), gw, bw, 0, c->stream, A, Q);
                else if (D.lean == 2 && D.coat_classes == (CC_AIRY | CC_FRESNEL))
                    hipLaunchKernelGGL((ghost_persist_kernel<2, CC_AIRY | CC_FRESNEL>), gw, bw, 0, c->stream, A, Q);
                else if (D.lean == 2)
                    hipLaunchKernelGGL((ghost_persist_kernel<2, CC_ALL>), gw, bw, 0, c->stream, A, Q);
                else
                    hipLaunchKernelGGL((ghost_persist_kernel<0, CC_ALL>), gw, bw, 0, c->stream, A, Q);
                GL_HIP(hipGetLastError());
            } else if ((e = launch<false, true, false>(A, n_ps, n_grid, c->stream, D.lean))) {
                return e;
            }
            if (events) GL_HIP(hipEventRecord(c->ev1, c->stream));
            if ((e = splat_resolve(c->stream, sg, sb, A.spec, A.out_r, A.out_g, A.out_b))) return e;
            if (events) GL_HIP(hipEventRecord(c->ev2, c->stream));
            if (c->spec_do_correct) {
                const float* C = c->spec_C;
                hipLaunchKernelGGL(spectral_correct_kernel, dim3((unsigned)((npx + 255) / 256)), dim3(256), 0, c->stream,
                                   A.out_r, A.out_g, A.out_b, npx, C[0], C[1], C[2], C[3], C[4], C[5], C[6], C[7], C[8]);
                GL_HIP(hipGetLastError());
            }
            hipLaunchKernelGGL(render_summary_kernel, dim3(1), dim3(1), 0, c->stream, sb, sg.n_tiles,
                               persistent ? Q.clk : nullptr, b_sum.as<unsigned long long>());
            GL_HIP(hipGetLastError());
            return GL_OK;
        };
        if (use_graph) {
            std::vector<uint8_t> gk;
            key_append(gk, A); key_append(gk, sb); key_append(gk, Q); key_append(gk, workers); key_append(gk, sg);
            key_append(gk, c->spec_do_correct); key_append(gk, c->spec_C); key_append(gk, npx);
            key_append(gk, b_sum.p);
            if (!c->graph || gk != c->graph_key) {
                c->graph_key.clear();
                hipGraph_t gph = nullptr;
                int e = 0;
                hipError_t ec;
                {
                    std::lock_guard<std::mutex> gm(device_mem_mutex());
                    if (c->graph) { (void)hipGraphExecDestroy(c->graph); c->graph = nullptr; }
                    GL_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
                    e = enqueue(false);
                    ec = hipStreamEndCapture(c->stream, &gph);
                }
                if (e) { if (gph) (void)hipGraphDestroy(gph); return e; }
                if (ec != hipSuccess) return set_error(GL_ERR_DEVICE, "hipStreamEndCapture -- %s", hipGetErrorString(ec));
                const hipError_t ei = hipGraphInstantiate(&c->graph, gph, nullptr, nullptr, 0);
                (void)hipGraphDestroy(gph);
                if (ei != hipSuccess) { c->graph = nullptr; return set_error(GL_ERR_DEVICE, "hipGraphInstantiate -- %s", hipGetErrorString(ei)); }
                c->graph_key = gk;
            }
            GL_HIP(hipEventRecord(c->ev0, c->stream));
            GL_HIP(hipGraphLaunch(c->graph, c->stream));
            GL_HIP(hipEventRecord(c->ev2, c->stream));
            graphed = true;
        } else if ((st = enqueue(true))) {
            return st;
        }
        unsigned long long sum_local[5] = {0, 0, 0, 0, 0};
        unsigned long long* h_sum = static_cast<unsigned long long*>(pin_take(c, sizeof sum_local));
        GL_HIP(hipMemcpyAsync(h_sum ? h_sum : sum_local, b_sum.p, sizeof sum_local, hipMemcpyDeviceToHost, c->stream));
        if (out_r) {
            GL_HIP(hipMemcpyAsync(out_r, A.out_r, npx * 4, hipMemcpyDeviceToHost, c->stream));
            GL_HIP(hipMemcpyAsync(out_g, A.out_g, npx * 4, hipMemcpyDeviceToHost, c->stream));
            GL_HIP(hipMemcpyAsync(out_b, A.out_b, npx * 4, hipMemcpyDeviceToHost, c->stream));
        }
        GL_HIP(hipStreamSynchronize(c->stream));
        if (h_sum) std::memcpy(sum_local, h_sum, sizeof sum_local);
        const unsigned long long counts[4] = {sum_local[0], sum_local[1], sum_local[3], sum_local[4]};
        const unsigned chunks[2] = {(unsigned)(sum_local[2] & 0xFFFFFFFFull), (unsigned)(sum_local[2] >> 32)};
        c->splat_hint = counts[0];
        c->splat_counts[0] = counts[0]; c->splat_counts[1] = counts[1];
        c->splat_counts[2] = chunks[0]; c->splat_counts[3] = chunks[1];
        // Trace-kernel time: HIP events around the launch when it is issued
        // eagerly; inside a replayed graph, the device wall clock its first
        // and last workgroups read.
        float kms = 0.f, sms = 0.f;
        if (graphed) {
            kms = counts[3] > counts[2] ? (float)((double)(counts[3] - counts[2]) / (double)c->wall_khz) : 0.0f;
        } else {
            GL_HIP(hipEventElapsedTime(&kms, c->ev0, c->ev1));
        }
        if (graphed) {
            float tot = 0.f;
            GL_HIP(hipEventElapsedTime(&tot, c->ev0, c->ev2));
            sms = std::max(0.0f, tot - kms);
        } else {
            GL_HIP(hipEventElapsedTime(&sms, c->ev1, c->ev2));
        }
        c->last_kernel_ms = kms;
        c->last_splat_ms = sms;
        const bool overflow = counts[0] > sb.rec_cap || counts[1] > sb.bin_cap;
        if (!overflow) {
            if (ts && cand >= 0) {
                if (ts->ms[cand] < 0.0f || kms < ts->ms[cand]) ts->ms[cand] = kms;
            }
            break;
        }
        if (attempt == 1) return set_error(GL_ERR_DEVICE, "splat record buffer overflow after regrow");
        // The binned copy holds each record once per touched tile (at most 4).
        want = std::max<unsigned long long>(counts[0], (counts[1] * 2 + 2) / 3) + (1ull << 20);
    }
    const double t_d1 = now_ms();
    if (persistent) {
        const Sched& sc = kSched[cand];
        c->last_sched[0] = lam_grp; c->last_sched[1] = workers; c->last_sched[2] = sc.longest_first;
        c->last_sched[3] = cand; c->last_sched[4] = trial ? 1 : 0; c->last_sched[6] = graphed ? 1 : 0;
        c->last_sched[7] = kNumSched;
    } else {
        for (int i = 0; i < 8; ++i) c->last_sched[i] = 0;
    }

    if (stats) {
        stats->ms_upload = (float)(t_up1 - t_up0);
        stats->ms_kernel = c->last_kernel_ms;
        stats->ms_download = (float)(t_d1 - t_k0);
        // Segment count of the main render (same inputs, no splat), then the
        // pre-cull per-pair picture over the full grid (ghost_render.cu:1361-1438).
        DevBuf& bsg = c->buf("segs");
        DevBuf& bcn = c->buf("stat_cnt");
        if ((st = bsg.ensure(8)) || (st = bcn.ensure(3 * n_pairs * 8))) return st;
        GL_HIP(hipMemsetAsync(bsg.p, 0, 8, c->stream));
        GL_HIP(hipMemsetAsync(bcn.p, 0, 3 * n_pairs * 8, c->stream));
        GhostArgs Cg = A;
        Cg.segs = bsg.as<unsigned long long>();
        if ((st = launch<false, false, true>(Cg, n_ps, n_grid, c->stream, D.lean))) return st;
        GhostArgs Sg = A;
        Sg.pair_active = nullptr; Sg.conc.rects = nullptr; Sg.bbox = nullptr; Sg.segs = nullptr;
        Sg.cnt = bcn.as<unsigned long long>();
        if ((st = launch<true, false, false>(Sg, n_ps, n_grid, c->stream, D.lean))) return st;
        std::vector<unsigned long long> h(3 * n_pairs);
        unsigned long long segs = 0;
        GL_HIP(hipMemcpyAsync(h.data(), bcn.p, 3 * n_pairs * 8, hipMemcpyDeviceToHost, c->stream));
        GL_HIP(hipMemcpyAsync(&segs, bsg.p, 8, hipMemcpyDeviceToHost, c->stream));
        GL_HIP(hipStreamSynchronize(c->stream));
        stats->segments = segs;
        stats->traces_total = stats->traces_survived = stats->traces_on_sensor = 0;
        for (int q = 0; q < n_pairs; ++q) {
            stats->traces_total += h[q]; stats->traces_survived += h[n_pairs + q];
            stats->traces_on_sensor += h[2 * n_pairs + q];
            if (q < stats->pair_cap) {
                if (stats->pair_surf_a) stats->pair_surf_a[q] = Lc.pairs[q].surf_a;
                if (stats->pair_surf_b) stats->pair_surf_b[q] = Lc.pairs[q].surf_b;
                if (stats->pair_traces) stats->pair_traces[q] = h[q];
                if (stats->pair_survived) stats->pair_survived[q] = h[n_pairs + q];
                if (stats->pair_on_sensor) stats->pair_on_sensor[q] = h[2 * n_pairs + q];
            }
        }
        stats->ms_total = (float)(now_ms() - t_fn0);
    }
    return GL_OK;
}

// Combined image of a frame-per-pair render: out_c[k] = sum over frames f, in
// frame order, of in_c[f * npx + k] (a fixed order, so the sum is the same on
// every run).
__global__ __launch_bounds__(256) void sum_frames_kernel(const float* __restrict__ r, const float* __restrict__ g,
                                                         const float* __restrict__ b, int frames, size_t npx,
                                                         float* __restrict__ out) {
    const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= npx) return;
    float sr = 0.0f, sg = 0.0f, sb = 0.0f;
    for (int f = 0; f < frames; ++f) {
        const size_t o = (size_t)f * npx + k;
        sr += r[o]; sg += g[o]; sb += b[o];
    }
    out[k] = sr; out[npx + k] = sg; out[2 * npx + k] = sb;
}

static int validate(const gl_lens* L, const gl_flare_params* p, int width, int height) {
    if (!L || !p) return set_error(GL_ERR_INVALID, "null lens or params");
    if (width <= 0 || height <= 0) return set_error(GL_ERR_INVALID, "width and height must be positive");
    if ((size_t)width * height > (size_t)INT32_MAX) return set_error(GL_ERR_INVALID, "image has too many pixels");
    if (p->ray_grid <= 0 || p->ray_grid > 4096) return set_error(GL_ERR_INVALID, "ray_grid %d outside [1, 4096]", p->ray_grid);
    if (std::max(3, p->spectral_samples) > GL_MAX_SPECTRAL)
        return set_error(GL_ERR_INVALID, "spectral_samples %d above %d", p->spectral_samples, GL_MAX_SPECTRAL);
    return GL_OK;
}

}  // namespace gl

using namespace gl;

extern "C" {

int gl_launch_ghost_render(gl_ctx* ctx, const gl_lens* lens, const gl_pair* pairs, int32_t n_pairs,
                           const gl_source* sources, int32_t n_sources, float sw, float sh, int32_t width,
                           int32_t height, int32_t fmt_w, int32_t fmt_h, int32_t fx0, int32_t fy0,
                           const gl_flare_params* p, float* out_r, float* out_g, float* out_b, float** dev_out,
                           gl_ghost_stats* stats) {
    int st = validate(lens, p, width, height);
    if (st) return st;
    if (n_pairs < 0 || n_sources < 0 || (n_pairs && !pairs) || (n_sources && !sources))
        return set_error(GL_ERR_INVALID, "bad pair/source arrays");
    if (!(sw > 0.0f) || !(sh > 0.0f)) return set_error(GL_ERR_INVALID, "sensor half extents must be positive");
    gl_ctx* c = resolve_ctx(ctx, &st);
    if (!c) return st;
    std::lock_guard<std::mutex> g(c->mu);
    gl_lens* L = const_cast<gl_lens*>(lens);
    Launch Lc{c, L, p, std::vector<gl_pair>(pairs, pairs + n_pairs), {}, n_sources, width, height,
              fmt_w, fmt_h, fx0, fy0, sw, sh};
    Lc.src.reserve(n_sources);
    for (int i = 0; i < n_sources; ++i) Lc.src.push_back(make_src(sources[i]));
    return do_launch(Lc, out_r, out_g, out_b, dev_out, stats);
}

int gl_filter_ghost_pairs(gl_ctx* ctx, const gl_lens* lens, float hw, float hh, const gl_flare_params* p,
                          gl_pair* out, int32_t cap, int32_t* n_out) {
    if (!lens || !p || !n_out) return set_error(GL_ERR_INVALID, "null argument");
    const std::vector<gl_surface>& s = lens->host;
    const int n = (int)s.size();
    // IOR-contrast gate on n_d (ghost.cpp:112-125), then device probes.
    auto ior_before = [&](int idx) {
        for (int k = idx - 1; k >= 0; --k) if (s[k].is_active) return s[k].ior;
        return 1.0f;
    };
    std::vector<int2> cand;
    for (int a = 0; a < n; ++a) {
        if (!s[a].is_active) continue;
        for (int b = a + 1; b < n; ++b) {
            if (!s[b].is_active) continue;
            if (std::fabs(ior_before(a) - s[a].ior) < 0.001f || std::fabs(ior_before(b) - s[b].ior) < 0.001f) continue;
            cand.push_back(make_int2(a, b));
        }
    }
    *n_out = 0;
    if (cand.empty()) return GL_OK;
    int st;
    gl_ctx* c = resolve_ctx(ctx, &st);
    if (!c) return st;
    std::lock_guard<std::mutex> g(c->mu);
    gl_lens* L = const_cast<gl_lens*>(lens);
    LensDev D;
    if ((st = lens_sync(L, 0.0f, c->stream, &D))) return st;
    const gl_spec s3[3] = {{650.f, 0, 0, 0}, {550.f, 0, 0, 0}, {450.f, 0, 0, 0}};
    std::vector<float> a3(3 * (size_t)n), b3(3 * (size_t)n);
    ior_tables(s.data(), n, s3, 3, a3.data(), b3.data());
    void *d_c, *d_a, *d_b;
    if ((st = upload(c, "f_pairs", cand.data(), cand.size() * sizeof(int2), &d_c)) ||
        (st = upload(c, "f_iat", a3.data(), a3.size() * 4, &d_a)) ||
        (st = upload(c, "f_ibf", b3.data(), b3.size() * 4, &d_b)))
        return st;
    DevBuf& bo = c->buf("f_out");
    if ((st = bo.ensure(cand.size() * 67 * sizeof(float4)))) return st;
    hipLaunchKernelGGL(filter_probe_kernel, dim3((unsigned)cand.size()), dim3(128), 0, c->stream, D.surf, D.poly,
                       D.arena, n, static_cast<const int2*>(d_c), static_cast<const float*>(d_a),
                       static_cast<const float*>(d_b), s[0].semi_aperture, s[0].z - 20.0f, bo.as<float4>());
    GL_HIP(hipGetLastError());
    std::vector<float4> res(cand.size() * 67);
    GL_HIP(hipMemcpyAsync(res.data(), bo.p, res.size() * sizeof(float4), hipMemcpyDeviceToHost, c->stream));
    GL_HIP(hipStreamSynchronize(c->stream));
    int cnt = 0;
    for (size_t i = 0; i < cand.size(); ++i) {
        const float4* r = &res[i * 67];
        float total = 0.0f;
        for (int ch = 0; ch < 3; ++ch) if (r[ch].w == 0.0f) total += r[ch].z;
        const float est = total / 3.0f;
        if (est < p->min_ghost_intensity) continue;
        float boost = 1.0f;
        if (p->ghost_normalize) {
            float mnx = 1e30f, mxx = -1e30f, mny = 1e30f, mxy = -1e30f;
            int valid = 0;
            for (int k = 0; k < 64; ++k) {
                const float4 q = r[3 + k];
                if (q.w != 0.0f) continue;
                mnx = std::min(mnx, q.x); mxx = std::max(mxx, q.x);
                mny = std::min(mny, q.y); mxy = std::max(mxy, q.y);
                ++valid;
            }
            if (valid >= 2) {
                const float gw = std::max(mxx - mnx, 0.01f), gh = std::max(mxy - mny, 0.01f);
                const float ratio = (gw * gh) / ((2.0f * hw) * (2.0f * hh));
                boost = std::min(std::max(ratio, 1.0f), p->max_area_boost);
            }
        }
        if (cnt < cap && out) out[cnt] = gl_pair{cand[i].x, cand[i].y, boost};
        ++cnt;
    }
    *n_out = cnt;
    return cnt > cap ? set_error(GL_ERR_INVALID, "pair capacity %d < %d", cap, cnt) : GL_OK;
}

int gl_render_ghost_pipeline(gl_ctx* ctx, const gl_lens* lens, const gl_source* sources, int32_t n_sources,
                             float calib_hw, float calib_hh, int32_t width, int32_t height, int32_t fmt_w,
                             int32_t fmt_h, int32_t fx0, int32_t fy0, const gl_flare_params* p, float* out_r,
                             float* out_g, float* out_b, float* aov_out, int32_t* aov_pairs, int32_t aov_cap,
                             int32_t* aov_n, gl_ghost_stats* stats) {
    int st = validate(lens, p, width, height);
    if (st) return st;
    if (p && p->frame_per_source)
        return set_error(GL_ERR_INVALID, "frame_per_source renders go through gl_launch_ghost_render");
    if (aov_n) *aov_n = 0;
    const size_t npx = (size_t)width * height;
    if (out_r) { std::fill(out_r, out_r + npx, 0.0f); std::fill(out_g, out_g + npx, 0.0f); std::fill(out_b, out_b + npx, 0.0f); }
    if (n_sources <= 0) return GL_OK;
    const float hw = (p->sensor_half_w > 0.0f) ? p->sensor_half_w : calib_hw;
    const float hh = (p->sensor_half_h > 0.0f) ? p->sensor_half_h : calib_hh;
    std::vector<gl_pair> pairs(GL_MAX_SURFACES * GL_MAX_SURFACES / 2 + 1);
    int32_t np = 0;
    if ((st = gl_filter_ghost_pairs(ctx, lens, hw, hh, p, pairs.data(), (int32_t)pairs.size(), &np))) return st;
    pairs.resize(np);
    if (p->ghost_filter_mode != GL_FILTER_ALL && p->ghost_filter_count > 0 && p->ghost_filter_pairs) {
        const bool include = p->ghost_filter_mode == GL_FILTER_INCLUDE;
        std::vector<gl_pair> kept;
        for (const auto& q : pairs) {
            bool hit = false;
            for (int k = 0; k < p->ghost_filter_count; ++k)
                if (p->ghost_filter_pairs[2 * k] == q.surf_a && p->ghost_filter_pairs[2 * k + 1] == q.surf_b) { hit = true; break; }
            if (include ? hit : !hit) kept.push_back(q);
        }
        pairs.swap(kept);
    }
    if (pairs.empty()) return GL_OK;
    if (p->aov_mode == GL_AOV_PER_PAIR && aov_out) {
        // All layers from one launch: each deposit record carries its pair as
        // the frame, the resolve writes frame-major layers, and the combined
        // image is their sum (ghost_render.cu:1526-1567 renders the pairs one
        // launch each and sums on the host).  Pairs past aov_max_pairs are
        // left out of both, as there.
        const int n_layers = p->aov_max_pairs < 0 ? (int)pairs.size() : std::min((int)pairs.size(), p->aov_max_pairs);
        if (n_layers > aov_cap) return set_error(GL_ERR_INVALID, "AOV capacity %d < %d layers", aov_cap, n_layers);
        if (n_layers <= 0) return GL_OK;
        pairs.resize(n_layers);
        gl_ctx* c = resolve_ctx(ctx, &st);
        if (!c) return st;
        std::lock_guard<std::mutex> g(c->mu);
        Launch Lc{c, const_cast<gl_lens*>(lens), p, pairs, {}, n_sources, width, height, fmt_w, fmt_h, fx0, fy0, hw, hh};
        Lc.frame_by_pair = true;
        Lc.src.reserve(n_sources);
        for (int i = 0; i < n_sources; ++i) Lc.src.push_back(make_src(sources[i]));
        float* dev[3] = {nullptr, nullptr, nullptr};
        if ((st = do_launch(Lc, nullptr, nullptr, nullptr, dev, nullptr))) return st;
        if (out_r) {
            DevBuf& bs = c->buf("aov_sum");
            if ((st = bs.ensure(3 * npx * sizeof(float)))) return st;
            hipLaunchKernelGGL(sum_frames_kernel, dim3((unsigned)((npx + 255) / 256)), dim3(256), 0, c->stream, dev[0],
                               dev[1], dev[2], n_layers, npx, bs.as<float>());
            GL_HIP(hipGetLastError());
            GL_HIP(hipMemcpyAsync(out_r, bs.as<float>(), npx * 4, hipMemcpyDeviceToHost, c->stream));
            GL_HIP(hipMemcpyAsync(out_g, bs.as<float>() + npx, npx * 4, hipMemcpyDeviceToHost, c->stream));
            GL_HIP(hipMemcpyAsync(out_b, bs.as<float>() + 2 * npx, npx * 4, hipMemcpyDeviceToHost, c->stream));
        }
        // layer i, channel ch -> aov_out[(i * 3 + ch) * npx ...]
        for (int ch = 0; ch < 3; ++ch)
            GL_HIP(hipMemcpy2DAsync(aov_out + (size_t)ch * npx, 3 * npx * sizeof(float), dev[ch], npx * sizeof(float),
                                    npx * sizeof(float), (size_t)n_layers, hipMemcpyDeviceToHost, c->stream));
        GL_HIP(hipStreamSynchronize(c->stream));
        if (aov_pairs)
            for (int i = 0; i < n_layers; ++i) { aov_pairs[2 * i] = pairs[i].surf_a; aov_pairs[2 * i + 1] = pairs[i].surf_b; }
        if (aov_n) *aov_n = n_layers;
        return GL_OK;
    }
    return gl_launch_ghost_render(ctx, lens, pairs.data(), (int32_t)pairs.size(), sources, n_sources, hw, hh, width,
                                  height, fmt_w, fmt_h, fx0, fy0, p, out_r, out_g, out_b, nullptr, stats);
}

int gl_trace_ghost_rays_level(gl_ctx* ctx, const gl_lens* lens, int32_t a, int32_t b, const float* rays,
                              int32_t n_rays, int32_t t_ge0, float fp64_rmin, int32_t level, float* out_pos,
                              float* out_w, int32_t* out_status, int32_t* out_segments) {
    if (!lens || !rays || n_rays < 0 || !out_pos || !out_w || !out_status) return set_error(GL_ERR_INVALID, "null argument");
    if (level < -1 || level > 2) return set_error(GL_ERR_INVALID, "level must be -1, 0, 1 or 2");
    if (n_rays == 0) return GL_OK;
    const int n = (int)lens->host.size();
    if (b >= 0 && (a < 0 || b <= a || b >= n)) {
        for (int i = 0; i < n_rays; ++i) { out_status[i] = GL_TRACE_INVALID_INPUT; out_w[i] = 0.0f; }
        return GL_OK;
    }
    int st;
    gl_ctx* c = resolve_ctx(ctx, &st);
    if (!c) return st;
    std::lock_guard<std::mutex> g(c->mu);
    gl_lens* L = const_cast<gl_lens*>(lens);
    LensDev D;
    if ((st = lens_sync(L, fp64_rmin, c->stream, &D))) return st;
    const int lf = level < 0 ? D.lean : level;
    // A lean level handles the surfaces of its own level and below only;
    // D.lean == 0 is a lens only the generic step handles.
    if (lf != 0 && (D.lean == 0 || lf < D.lean))
        return set_error(GL_ERR_INVALID, "lean level %d does not handle this lens (its level is %d)", lf, D.lean);
    // Per-ray tables (each ray carries its own wavelength): indices after and
    // before each surface, and for the lean levels the refraction ratios and
    // coating rows the render tables hold per lambda.
    std::vector<gl_spec> sp((size_t)n_rays);
    for (int i = 0; i < n_rays; ++i) sp[i] = gl_spec{rays[7 * (size_t)i + 6], 0, 0, 0};
    std::vector<float> iat((size_t)n_rays * n), ibf((size_t)n_rays * n);
    ior_tables(L->host.data(), n, sp.data(), n_rays, iat.data(), ibf.data());
    void *d_r, *d_a, *d_b, *d_rec = nullptr, *d_rat = nullptr, *d_lrow = nullptr;
    if ((st = upload(c, "r_rays", rays, (size_t)n_rays * 7 * 4, &d_r)) ||
        (st = upload(c, "r_iat", iat.data(), iat.size() * 4, &d_a)) ||
        (st = upload(c, "r_ibf", ibf.data(), ibf.size() * 4, &d_b)))
        return st;
    if (lf != 0) {
        std::vector<float2> rats(iat.size());
        for (size_t q = 0; q < rats.size(); ++q) rats[q] = make_float2(ibf[q] / iat[q], iat[q] / ibf[q]);
        std::vector<int32_t> lrows;
        lambda_rows(L->host.data(), n, L->arena.data(), sp.data(), n_rays, lrows);
        const std::vector<LamRec> lrec = lam_records(iat, ibf, rats, lrows);
        if ((st = upload(c, "r_lrec", lrec.data(), lrec.size() * sizeof(LamRec), &d_rec)) ||
            (st = upload(c, "r_rat", rats.data(), rats.size() * sizeof(float2), &d_rat)) ||
            (st = upload(c, "r_lrow", lrows.data(), lrows.size() * sizeof(int32_t), &d_lrow)))
            return st;
    }
    DevBuf &bp = c->buf("r_pos"), &bw = c->buf("r_w"), &bs = c->buf("r_st"), &bg = c->buf("r_seg");
    if ((st = bp.ensure((size_t)n_rays * 12)) || (st = bw.ensure((size_t)n_rays * 4)) ||
        (st = bs.ensure((size_t)n_rays * 4)) || (st = bg.ensure((size_t)n_rays * 4)))
        return st;
    const dim3 gr((unsigned)((n_rays + 255) / 256)), bl(256);
    const float *fr = static_cast<const float*>(d_r), *fa = static_cast<const float*>(d_a),
                *fb = static_cast<const float*>(d_b);
    const LamRec* lq = static_cast<const LamRec*>(d_rec);
    const float2* rt = static_cast<const float2*>(d_rat);
    const int2* lr = static_cast<const int2*>(d_lrow);
    const int cc = D.coat_classes;
#define GL_RAY_LAUNCH(LF_, CM_, RT_)                                                                                 \
    hipLaunchKernelGGL((ray_kernel<LF_, CM_, RT_>), gr, bl, 0, c->stream, D.surf, D.poly, D.arena, n, a, b, fr,      \
                       n_rays, fa, fb, lq, rt, lr, t_ge0, bp.as<float>(), bw.as<float>(), bs.as<int>(), bg.as<int>())
    if (lf == 1 && cc == (CC_SA | CC_FRESNEL)) GL_RAY_LAUNCH(1, CC_SA | CC_FRESNEL, true);
    else if (lf == 1 && cc == (CC_AIRY | CC_FRESNEL)) GL_RAY_LAUNCH(1, CC_AIRY | CC_FRESNEL, true);
    else if (lf == 1) GL_RAY_LAUNCH(1, CC_ALL, true);
    else if (lf == 2 && cc == (CC_SA | CC_FRESNEL)) GL_RAY_LAUNCH(2, CC_SA | CC_FRESNEL, true);
    else if (lf == 2 && cc == (CC_AIRY | CC_FRESNEL)) GL_RAY_LAUNCH(2, CC_AIRY | CC_FRESNEL, true);
    else if (lf == 2) GL_RAY_LAUNCH(2, CC_ALL, true);
    else GL_RAY_LAUNCH(0, CC_ALL, false);
#undef GL_RAY_LAUNCH
    GL_HIP(hipGetLastError());
    GL_HIP(hipMemcpyAsync(out_pos, bp.p, (size_t)n_rays * 12, hipMemcpyDeviceToHost, c->stream));
    GL_HIP(hipMemcpyAsync(out_w, bw.p, (size_t)n_rays * 4, hipMemcpyDeviceToHost, c->stream));
    GL_HIP(hipMemcpyAsync(out_status, bs.p, (size_t)n_rays * 4, hipMemcpyDeviceToHost, c->stream));
    if (out_segments) GL_HIP(hipMemcpyAsync(out_segments, bg.p, (size_t)n_rays * 4, hipMemcpyDeviceToHost, c->stream));
    GL_HIP(hipStreamSynchronize(c->stream));
    return GL_OK;
}

int gl_trace_ghost_rays(gl_ctx* ctx, const gl_lens* lens, int32_t a, int32_t b, const float* rays, int32_t n_rays,
                        int32_t t_ge0, float fp64_rmin, float* out_pos, float* out_w, int32_t* out_status,
                        int32_t* out_segments) {
    return gl_trace_ghost_rays_level(ctx, lens, a, b, rays, n_rays, t_ge0, fp64_rmin, 0, out_pos, out_w, out_status,
                                     out_segments);
}

}  // extern "C"
