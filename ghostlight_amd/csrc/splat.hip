// splat.hip — the ghost splat without global float atomics in the hot loop,
// and with an image that does not depend on the order records arrive in.
//
// The trace kernel does not add into the image.  Each surviving (ray, lambda)
// becomes one 16-byte deposit record {px, py, power, lambda index}, staged per
// wave in LDS and flushed to HBM in dense runs (one 64-bit atomic reservation
// per ~256 records).  The image is then built tile by tile:
//
//   bin_count    per-workgroup LDS histogram of records over screen tiles
//                (a bilinear footprint touches 1, 2 or 4 tiles), merged into
//                per-tile counts and also kept per workgroup for the scatter;
//                also the largest |power x colour weight| a tile receives
//   bin_scan     one workgroup: exclusive scan -> tile starts / cursors, the
//                work-chunk list (a tile with many records is split into
//                chunks) and a staging slot per tile with several chunks
//   bin_scatter  copies each record into its tiles' ranges (each workgroup
//                takes the same record slice as in bin_count and reserves its
//                ranges from that slice's stored histogram)
//   resolve      persistent workgroups pull chunks from a queue and sum the
//                chunk's bilinear contributions into an LDS tile; a tile with
//                one chunk is written out by that workgroup (read-add-write),
//                a tile with several adds its chunk sums into its staging slot
//   finish       one workgroup per staged tile writes it out and zeroes the slot
//
// Every output pixel therefore has one writer, and the sums are kept in fixed
// point: each contribution is rounded to a multiple of 2^s, with s chosen per
// tile from its record count and largest contribution so that no int64 sum
// can overflow (count x largest < 2^62).  Integer addition is associative, so
// the LDS and staging atomics give the same bits in any order: a render is
// bit-identical from run to run and across launch schedules, however the
// trace's workgroups interleave their record flushes.  The quantum is 2^-38
// or finer relative to the tile's largest contribution (2^-(62 - log2 count)),
// far below float32 resolution for any pixel the tile's bright records reach.
//
// Frame mode: a record's lambda word also carries its frame (one frame per
// source, rec.w bits 8..), tiles are numbered frame-major and the resolve
// writes into that frame's planes (out + frame * width * height).  With more
// tiles than an LDS histogram holds, the binning passes count straight into
// the global tile counters.
//
// Each contribution is the same product the reference's splat adds
// (ghost_render.cu:430-462): colour = power * cmf_c(lambda), weight = the
// bilinear cell overlap, skipped when |colour| <= 1e-14.
#include "gl_runtime.h"
#include "gl_splat.h"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace gl {

static constexpr int kBinBlocks = 1024;
static constexpr int kBinThreads = 256;
static constexpr int kChunk = 8192;          // records per resolve work item
static constexpr int kResolveBlocks = 1536;     // six 24 KiB LDS tiles per CU
static constexpr int kMaxTiles = 8192;       // LDS: 8 B a tile in bin_count, 12 B in bin_scatter
static constexpr int kFinishBlocks = 512;

// Bilinear footprint of a record: pixels x0..x0+1, y0..y0+1 clipped to the
// image (the splat's cell arithmetic, ghost_render.cu:430-441).
struct Foot {
    int x0, y0;
    float fx, fy;
};

__device__ __forceinline__ Foot foot_of(float px, float py) {
    Foot f;
    f.x0 = (int)floorf(px - 0.5f);
    f.y0 = (int)floorf(py - 0.5f);
    f.fx = (px - 0.5f) - (float)f.x0;
    f.fy = (py - 0.5f) - (float)f.y0;
    return f;
}

// Tiles touched by a record (clipped footprint), as a [tx0, tx1] x [ty0, ty1]
// box within frame `base / tiles_per_frame` (base = first tile id of the frame).
__device__ __forceinline__ bool tiles_of(const float4& r, const SplatGeom& g, int& tx0, int& tx1, int& ty0, int& ty1,
                                         int& base) {
    const Foot f = foot_of(r.x, r.y);
    const int xa = max(f.x0, 0), xb = min(f.x0 + 1, g.width - 1);
    const int ya = max(f.y0, 0), yb = min(f.y0 + 1, g.height - 1);
    if (xa > xb || ya > yb) return false;
    const int frame = __float_as_int(r.w) >> 8;
    if (frame >= g.frames) return false;
    base = frame * g.tiles_per_frame;
    tx0 = xa >> g.tile_shift; tx1 = xb >> g.tile_shift;
    ty0 = ya >> g.tile_shift; ty1 = yb >> g.tile_shift;
    return true;
}

// The up-to-four tiles of a record's footprint box as a fixed sequence of
// slots j = 0..3: (tx0, ty0), (tx1, ty0), (tx0, ty1), (tx1, ty1), a slot
// valid only where its corner is a new tile.
__device__ __forceinline__ bool tile_slot(int j, int tx0, int tx1, int ty0, int ty1, int base, int tiles_x, int& t) {
    const int tx = (j & 1) ? tx1 : tx0, ty = (j & 2) ? ty1 : ty0;
    const bool ok = ((j & 1) == 0 || tx1 > tx0) && ((j & 2) == 0 || ty1 > ty0);
    t = base + ty * tiles_x + tx;
    return ok;
}

// Wave-aggregated global tile updates: the active lanes that hold the same
// tile t form one group, served by one atomic from its lowest lane.  The
// groups are taken one at a time (the lowest pending lane's tile first); the
// loop runs once per distinct tile among the wave's lanes.  COUNT adds the
// group's size to cnt[t] and raises mag[t] to the group's largest m; CURSOR
// reserves the group's slots in cur[t] and returns each lane's own slot.
template <bool CURSOR>
__device__ __forceinline__ unsigned long long wave_tile_update(bool valid, int t, unsigned m, unsigned* cnt,
                                                               unsigned* mag, unsigned long long* cur) {
    const int lane = (int)(threadIdx.x & 63);
    unsigned long long pending = __ballot(valid);
    unsigned long long slot = 0;
    while (pending) {
        const int leader = __builtin_ctzll(pending);
        const int tl = __shfl(t, leader, 64);
        const bool mine = valid && t == tl;
        const unsigned long long grp = __ballot(mine);
        const unsigned n = (unsigned)__popcll(grp);
        if (CURSOR) {
            unsigned long long b = 0;
            if (lane == leader) b = atomicAdd(&cur[tl], (unsigned long long)n);
            b = __shfl(b, leader, 64);
            const unsigned rank = __builtin_amdgcn_mbcnt_hi((unsigned)(grp >> 32),
                                                            __builtin_amdgcn_mbcnt_lo((unsigned)grp, 0u));
            if (mine) slot = b + rank;
        } else {
            unsigned mm = mine ? m : 0u;
            for (int o = 32; o > 0; o >>= 1) mm = max(mm, (unsigned)__shfl_xor((int)mm, o, 64));
            if (lane == leader) {
                atomicAdd(&cnt[tl], n);
                atomicMax(&mag[tl], mm);
            }
        }
        pending &= ~grp;
    }
    return slot;
}

// Records a thread loads before it processes them in the binning passes:
// several loads in flight per lane.  A thread still takes its records in the
// order r0 + tid, r0 + tid + blockDim, ..., so the passes see the same
// per-thread sequence as a plain strided loop.
static constexpr int kBinUnroll = 4;

template <class F>
__device__ __forceinline__ void for_records(const float4* __restrict__ rec, unsigned long long r0,
                                            unsigned long long r1, F&& f) {
    const unsigned long long step = blockDim.x;
    for (unsigned long long i0 = r0 + threadIdx.x; i0 < r1; i0 += kBinUnroll * step) {
        float4 rr[kBinUnroll];
#pragma unroll
        for (int u = 0; u < kBinUnroll; ++u) {
            const unsigned long long i = i0 + (unsigned long long)u * step;
            if (i < r1) rr[u] = rec[i];
        }
#pragma unroll
        for (int u = 0; u < kBinUnroll; ++u)
            if (i0 + (unsigned long long)u * step < r1) f(rr[u]);
    }
}

__device__ __forceinline__ unsigned long long n_records(const SplatBufs& B) {
    const unsigned long long n = *B.rec_count;
    return n < B.rec_cap ? n : B.rec_cap;
}

__global__ __launch_bounds__(256) void splat_clear_kernel(SplatBufs B, int n_tiles) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n_tiles) { B.tile_count[i] = 0u; B.tile_mag[i] = 0u; }
    if (i == 0) { *B.rec_count = 0ull; *B.queue = 0u; *B.bin_total = 0ull; *B.mc_count = 0u; }
}

// Upper bound on every contribution a record makes: |power| x the largest
// |colour weight| of its lambda, rounded like the products the resolve forms
// (rounding is monotonic, and the bilinear weights are at most 1).  Returned
// as float bits, which order like the non-negative floats they encode.
__device__ __forceinline__ unsigned record_mag(const float4& r, const float4* __restrict__ spec) {
    const float4 s = spec[__float_as_int(r.w) & 0xFF];
    const float m = fmaxf(fabsf(s.y), fmaxf(fabsf(s.z), fabsf(s.w)));
    return __float_as_uint(fabsf(r.z) * m);
}

// LDS: per-workgroup histograms in LDS (g.lds_hist), else straight into the
// global tile counters.  The two forms are separate instantiations so that
// the LDS one issues LDS atomics (a pointer that may be either would make
// every atomic a flat one).
template <bool LDS>
__global__ __launch_bounds__(kBinThreads) void bin_count_kernel(SplatBufs B, SplatGeom g,
                                                                const float4* __restrict__ spec) {
    extern __shared__ unsigned hist[];            // counts, then magnitudes (LDS form)
    unsigned* lmag = hist + g.n_tiles;
    if (LDS) {
        for (int t = threadIdx.x; t < 2 * g.n_tiles; t += blockDim.x) hist[t] = 0u;
        __syncthreads();
    }
    unsigned* h = LDS ? hist : B.tile_count;
    unsigned* hm = LDS ? lmag : B.tile_mag;
    const unsigned long long n = n_records(B);
    const unsigned long long per = (n + gridDim.x - 1) / gridDim.x;
    const unsigned long long r0 = per * blockIdx.x, r1 = min(n, r0 + per);
    if (LDS) {
        for_records(B.rec, r0, r1, [&](const float4& r) {
            int tx0, tx1, ty0, ty1, base;
            if (!tiles_of(r, g, tx0, tx1, ty0, ty1, base)) return;
            const unsigned m = record_mag(r, spec);
            for (int ty = ty0; ty <= ty1; ++ty)
                for (int tx = tx0; tx <= tx1; ++tx) {
                    const int t = base + ty * g.tiles_x + tx;
                    atomicAdd(&h[t], 1u);
                    atomicMax(&hm[t], m);
                }
        });
    } else {
        // Global counters, one atomic per distinct tile of the wave.  The
        // loop bound is uniform across the workgroup, so every lane of a wave
        // is active at each aggregation step (lanes past the range hold no
        // tile).
        const unsigned long long span = (r1 > r0) ? r1 - r0 : 0ull;
        const unsigned long long steps = (span + blockDim.x - 1) / blockDim.x;
        for (unsigned long long k = 0; k < steps; ++k) {
            const unsigned long long i = r0 + k * blockDim.x + threadIdx.x;
            int tx0 = 0, tx1 = 0, ty0 = 0, ty1 = 0, base = 0;
            bool in = false;
            unsigned m = 0u;
            if (i < r1) {
                const float4 r = B.rec[i];
                in = tiles_of(r, g, tx0, tx1, ty0, ty1, base);
                if (in) m = record_mag(r, spec);
            }
            for (int j = 0; j < 4; ++j) {
                int t = 0;
                const bool ok = in && tile_slot(j, tx0, tx1, ty0, ty1, base, g.tiles_x, t);
                wave_tile_update<false>(ok, t, m, h, hm, nullptr);
            }
        }
    }
    if (LDS) {
        __syncthreads();
        unsigned* mine = B.blk_hist + (size_t)blockIdx.x * g.n_tiles;
        for (int t = threadIdx.x; t < g.n_tiles; t += blockDim.x) {
            const unsigned c = hist[t];
            mine[t] = c;
            if (c) {
                atomicAdd(&B.tile_count[t], c);
                atomicMax(&B.tile_mag[t], lmag[t]);
            }
        }
    }
}

// Exclusive scan of tile counts (one workgroup of 1024 lanes, a contiguous
// span of tiles each): tile_start, tile_cursor, chunk_start, the staging slots
// of tiles with several chunks, and the totals.
__global__ __launch_bounds__(1024) void bin_scan_kernel(SplatBufs B, SplatGeom g) {
    __shared__ unsigned long long s_rec[1024];
    __shared__ unsigned s_chk[1024];
    __shared__ unsigned s_mc[1024];
    const int per = (g.n_tiles + 1023) / 1024;
    const int t0 = threadIdx.x * per, t1 = min(g.n_tiles, t0 + per);
    unsigned long long rs = 0;
    unsigned cs = 0, ms = 0;
    for (int t = t0; t < t1; ++t) {
        const unsigned c = B.tile_count[t];
        rs += c;
        cs += (c + kChunk - 1) / kChunk;
        ms += c > (unsigned)kChunk ? 1u : 0u;
    }
    s_rec[threadIdx.x] = rs;
    s_chk[threadIdx.x] = cs;
    s_mc[threadIdx.x] = ms;
    __syncthreads();
    // Hillis-Steele inclusive scan over the 1024 lane totals.
    for (int off = 1; off < 1024; off <<= 1) {
        const bool in = threadIdx.x >= (unsigned)off;
        const unsigned long long a = in ? s_rec[threadIdx.x - off] : 0ull;
        const unsigned b = in ? s_chk[threadIdx.x - off] : 0u;
        const unsigned m = in ? s_mc[threadIdx.x - off] : 0u;
        __syncthreads();
        s_rec[threadIdx.x] += a;
        s_chk[threadIdx.x] += b;
        s_mc[threadIdx.x] += m;
        __syncthreads();
    }
    unsigned long long run = threadIdx.x ? s_rec[threadIdx.x - 1] : 0ull;
    unsigned crun = threadIdx.x ? s_chk[threadIdx.x - 1] : 0u;
    unsigned mrun = threadIdx.x ? s_mc[threadIdx.x - 1] : 0u;
    for (int t = t0; t < t1; ++t) {
        const unsigned c = B.tile_count[t];
        B.tile_start[t] = run;
        B.tile_cursor[t] = run;
        B.chunk_start[t] = crun;
        run += c;
        crun += (c + kChunk - 1) / kChunk;
        if (c > (unsigned)kChunk) {
            B.mc_slot[t] = mrun;
            if (mrun < B.stage_tiles) B.mc_tile[mrun] = (unsigned)t;
            ++mrun;
        }
    }
    if (threadIdx.x == 1023) {
        *B.bin_total = s_rec[1023];
        B.chunk_start[g.n_tiles] = s_chk[1023];
        *B.mc_count = s_mc[1023];
    }
}

__global__ __launch_bounds__(kBinThreads) void bin_scatter_kernel(SplatBufs B, SplatGeom g) {
    extern __shared__ unsigned smem[];
    const unsigned long long n = n_records(B);
    const unsigned long long per = (n + gridDim.x - 1) / gridDim.x;
    const unsigned long long r0 = per * blockIdx.x, r1 = min(n, r0 + per);
    if (!g.lds_hist) {
        // Global form: cursor reservations aggregated per distinct tile of
        // the wave (the loop bound is uniform across the workgroup, so whole
        // waves run each step together).
        const unsigned long long span = (r1 > r0) ? r1 - r0 : 0ull;
        const unsigned long long steps = (span + blockDim.x - 1) / blockDim.x;
        for (unsigned long long k = 0; k < steps; ++k) {
            const unsigned long long i = r0 + k * blockDim.x + threadIdx.x;
            float4 r = make_float4(0.f, 0.f, 0.f, 0.f);
            int tx0 = 0, tx1 = 0, ty0 = 0, ty1 = 0, fb = 0;
            bool in = false;
            if (i < r1) {
                r = B.rec[i];
                in = tiles_of(r, g, tx0, tx1, ty0, ty1, fb);
            }
            for (int j = 0; j < 4; ++j) {
                int t = 0;
                const bool ok = in && tile_slot(j, tx0, tx1, ty0, ty1, fb, g.tiles_x, t);
                const unsigned long long slot = wave_tile_update<true>(ok, t, 0u, nullptr, nullptr, B.tile_cursor);
                if (ok && slot < B.bin_cap) B.binned[slot] = r;
            }
        }
        return;
    }
    unsigned* hist = smem;                       // per-tile cursor within this block's ranges
    unsigned long long* base = reinterpret_cast<unsigned long long*>(smem + ((g.n_tiles + 1) & ~1));
    // This block's slice counts, stored by bin_count for the same slice.
    const unsigned* mine = B.blk_hist + (size_t)blockIdx.x * g.n_tiles;
    for (int t = threadIdx.x; t < g.n_tiles; t += blockDim.x) {
        const unsigned c = mine[t];
        base[t] = c ? atomicAdd(&B.tile_cursor[t], (unsigned long long)c) : 0ull;
        hist[t] = 0u;
    }
    __syncthreads();
    for (unsigned long long i = r0 + threadIdx.x; i < r1; i += blockDim.x) {
        const float4 r = B.rec[i];
        int tx0, tx1, ty0, ty1, fb;
        if (!tiles_of(r, g, tx0, tx1, ty0, ty1, fb)) continue;
        for (int ty = ty0; ty <= ty1; ++ty)
            for (int tx = tx0; tx <= tx1; ++tx) {
                const int t = fb + ty * g.tiles_x + tx;
                const unsigned long long slot = base[t] + atomicAdd(&hist[t], 1u);
                if (slot < B.bin_cap) B.binned[slot] = r;
            }
    }
}

// Fixed-point exponent of a tile: every contribution is below 2^e (e from the
// tile's largest one) and a pixel receives at most `count` of them, so with
// quantum 2^s, s = e + bitlength(count) - 62, every int64 sum stays below 2^62.
__device__ __forceinline__ int tile_quantum(unsigned count, unsigned mag_bits) {
    int e;
    (void)frexpf(__uint_as_float(mag_bits), &e);
    return e + (32 - __clz(count)) - 62;
}

// A contribution rounded to the nearest multiple of 2^s, in units of 2^s.
// ldexpf is exact here (the scaled value is at most 2^62), so the rounding
// happens once, in the conversion.
__device__ __forceinline__ long long to_fixed(float x, int s) { return __float2ll_rn(ldexpf(x, -s)); }

__device__ __forceinline__ float from_fixed(long long v, int s) { return (float)ldexp((double)v, s); }

// One work item = up to kChunk records of one tile.
__global__ __launch_bounds__(256) void resolve_kernel(SplatBufs B, SplatGeom g, const float4* __restrict__ spec,
                                                      float* __restrict__ out_r, float* __restrict__ out_g,
                                                      float* __restrict__ out_b) {
    extern __shared__ unsigned long long acc[];   // 3 fixed-point planes of TS x TS
    __shared__ unsigned s_item;
    const int TS = 1 << g.tile_shift, TT = TS * TS;
    const unsigned n_items = B.chunk_start[g.n_tiles];
    const unsigned long long bin_n = min(*B.bin_total, B.bin_cap);
    for (;;) {
        if (threadIdx.x == 0) s_item = atomicAdd(B.queue, 1u);
        __syncthreads();
        const unsigned item = s_item;
        __syncthreads();
        if (item >= n_items) break;
        // Tile of this item: last t with chunk_start[t] <= item.
        int lo = 0, hi = g.n_tiles - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (B.chunk_start[mid] <= item) lo = mid; else hi = mid - 1;
        }
        const int t = lo;
        const unsigned k = item - B.chunk_start[t];
        const unsigned long long rb = B.tile_start[t] + (unsigned long long)k * kChunk;
        const unsigned long long re = min(min(B.tile_start[t] + B.tile_count[t], rb + kChunk), bin_n);
        const int nchunks = (int)(B.chunk_start[t + 1] - B.chunk_start[t]);
        const int sq = tile_quantum(B.tile_count[t], B.tile_mag[t]);
        for (int i = threadIdx.x; i < 3 * TT; i += blockDim.x) acc[i] = 0ull;
        __syncthreads();
        const int frame = t / g.tiles_per_frame, tf = t % g.tiles_per_frame;
        const int ox = (tf % g.tiles_x) << g.tile_shift, oy = (tf / g.tiles_x) << g.tile_shift;
        const size_t fo = (size_t)frame * g.width * g.height;
        for (unsigned long long i = rb + threadIdx.x; i < re; i += blockDim.x) {
            const float4 r = B.binned[i];
            const Foot f = foot_of(r.x, r.y);
            const float4 s = spec[__float_as_int(r.w) & 0xFF];
            const float cr = r.z * s.y, cg = r.z * s.z, cb = r.z * s.w;
            const float w00 = (1.0f - f.fx) * (1.0f - f.fy), w10 = f.fx * (1.0f - f.fy);
            const float w01 = (1.0f - f.fx) * f.fy, w11 = f.fx * f.fy;
            const int xs[4] = {f.x0, f.x0 + 1, f.x0, f.x0 + 1}, ys[4] = {f.y0, f.y0, f.y0 + 1, f.y0 + 1};
            const float ws[4] = {w00, w10, w01, w11};
            const bool ur = fabsf(cr) > 1e-14f, ug = fabsf(cg) > 1e-14f, ub = fabsf(cb) > 1e-14f;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int lx = xs[q] - ox, ly = ys[q] - oy;
                if (xs[q] < 0 || xs[q] >= g.width || ys[q] < 0 || ys[q] >= g.height) continue;
                if (lx < 0 || lx >= TS || ly < 0 || ly >= TS) continue;
                const int p = ly * TS + lx;
                const long long vr = ur ? to_fixed(cr * ws[q], sq) : 0ll;
                const long long vg = ug ? to_fixed(cg * ws[q], sq) : 0ll;
                const long long vb = ub ? to_fixed(cb * ws[q], sq) : 0ll;
                if (vr) atomicAdd(&acc[p], (unsigned long long)vr);
                if (vg) atomicAdd(&acc[TT + p], (unsigned long long)vg);
                if (vb) atomicAdd(&acc[2 * TT + p], (unsigned long long)vb);
            }
        }
        __syncthreads();
        if (nchunks == 1) {
            for (int p = threadIdx.x; p < TT; p += blockDim.x) {
                const int x = ox + (p & (TS - 1)), y = oy + (p >> g.tile_shift);
                if (x >= g.width || y >= g.height) continue;
                const size_t o = fo + (size_t)y * g.width + x;
                const long long vr = (long long)acc[p], vg = (long long)acc[TT + p], vb = (long long)acc[2 * TT + p];
                if (vr) out_r[o] += from_fixed(vr, sq);
                if (vg) out_g[o] += from_fixed(vg, sq);
                if (vb) out_b[o] += from_fixed(vb, sq);
            }
        } else {
            const unsigned slot = B.mc_slot[t];
            if (slot < B.stage_tiles) {
                unsigned long long* st = B.stage + (size_t)slot * 3 * TT;
                for (int i = threadIdx.x; i < 3 * TT; i += blockDim.x)
                    if (acc[i]) atomicAdd(st + i, acc[i]);
            }
        }
        __syncthreads();
    }
}

// Tiles with several chunks: one workgroup each converts the staged sums,
// adds them to the image and leaves the slot zero for the next render.
__global__ __launch_bounds__(256) void resolve_finish_kernel(SplatBufs B, SplatGeom g, float* __restrict__ out_r,
                                                             float* __restrict__ out_g, float* __restrict__ out_b) {
    const int TS = 1 << g.tile_shift, TT = TS * TS;
    const unsigned n = min(*B.mc_count, B.stage_tiles);
    for (unsigned i = blockIdx.x; i < n; i += gridDim.x) {
        const int t = (int)B.mc_tile[i];
        const int sq = tile_quantum(B.tile_count[t], B.tile_mag[t]);
        const int frame = t / g.tiles_per_frame, tf = t % g.tiles_per_frame;
        const int ox = (tf % g.tiles_x) << g.tile_shift, oy = (tf / g.tiles_x) << g.tile_shift;
        const size_t fo = (size_t)frame * g.width * g.height;
        unsigned long long* st = B.stage + (size_t)i * 3 * TT;
        for (int p = threadIdx.x; p < TT; p += blockDim.x) {
            const long long vr = (long long)st[p], vg = (long long)st[TT + p], vb = (long long)st[2 * TT + p];
            if (!(vr | vg | vb)) continue;
            st[p] = 0ull; st[TT + p] = 0ull; st[2 * TT + p] = 0ull;
            const int x = ox + (p & (TS - 1)), y = oy + (p >> g.tile_shift);
            if (x >= g.width || y >= g.height) continue;
            const size_t o = fo + (size_t)y * g.width + x;
            if (vr) out_r[o] += from_fixed(vr, sq);
            if (vg) out_g[o] += from_fixed(vg, sq);
            if (vb) out_b[o] += from_fixed(vb, sq);
        }
    }
}

SplatGeom splat_geom(int width, int height, int frames) {
    SplatGeom g;
    std::memset(&g, 0, sizeof g);
    g.width = width;
    g.height = height;
    g.frames = frames < 1 ? 1 : frames;
    g.tile_shift = 5;
    for (;;) {
        const int TS = 1 << g.tile_shift;
        g.tiles_x = (width + TS - 1) / TS;
        g.tiles_y = (height + TS - 1) / TS;
        const long long nt = (long long)g.tiles_x * g.tiles_y;
        if (nt <= kMaxTiles || g.tile_shift >= 6) { g.tiles_per_frame = (int)nt; break; }
        ++g.tile_shift;
    }
    const long long all = (long long)g.tiles_per_frame * g.frames;
    g.n_tiles = all > (long long)INT32_MAX / 2 ? INT32_MAX / 2 : (int)all;
    g.lds_hist = g.n_tiles <= kMaxTiles ? 1 : 0;
    return g;
}

int splat_prepare(gl_ctx* c, const SplatGeom& g, unsigned long long rec_cap, SplatBufs& B) {
    if (g.tiles_per_frame > kMaxTiles)
        return set_error(GL_ERR_INVALID, "image too large for the tiled splat (%d tiles a frame)", g.tiles_per_frame);
    DevBuf &brec = c->buf("splat_rec"), &bbin = c->buf("splat_bin"), &bt = c->buf("splat_tiles");
    DevBuf& bst = c->buf("splat_stage");
    int st;
    const unsigned long long bin_cap = rec_cap + rec_cap / 2;
    if ((st = brec.ensure((size_t)rec_cap * sizeof(float4))) || (st = bbin.ensure((size_t)bin_cap * sizeof(float4))))
        return st;
    // tile arrays: start/cursor (u64), scalars, count (u32), chunk_start
    // (u32, n+1), magnitude, staging slot, staged tile (u32 each), then the
    // binning workgroups' slice histograms (LDS form)
    const size_t nt = (size_t)g.n_tiles;
    const size_t bytes = nt * 8 * 2 + 64 + nt * 4 + (nt + 1) * 4 + nt * 4 * 3;
    const size_t blk_bytes = g.lds_hist ? (size_t)kBinBlocks * nt * 4 : 0;
    if ((st = bt.ensure(bytes + 256 + blk_bytes))) return st;
    char* p = bt.as<char>();
    B.rec = brec.as<float4>();
    B.rec_cap = brec.cap / sizeof(float4);
    B.binned = bbin.as<float4>();
    B.bin_cap = bbin.cap / sizeof(float4);
    B.tile_start = reinterpret_cast<unsigned long long*>(p);
    B.tile_cursor = B.tile_start + nt;
    B.rec_count = B.tile_cursor + nt;
    B.bin_total = B.rec_count + 1;
    B.tile_count = reinterpret_cast<unsigned*>(B.bin_total + 1);
    B.chunk_start = B.tile_count + nt;
    B.queue = B.chunk_start + nt + 1;
    B.mc_count = B.queue + 1;
    B.tile_mag = B.mc_count + 1;
    B.mc_slot = B.tile_mag + nt;
    B.mc_tile = B.mc_slot + nt;
    B.blk_hist = g.lds_hist ? reinterpret_cast<unsigned*>(p + ((bytes + 255) & ~(size_t)255)) : nullptr;
    // Staging for tiles with several chunks: each holds more than kChunk of
    // the binned records, so there are at most bin_cap / kChunk of them (a
    // render whose binning overflows is redone with larger buffers; its
    // excess tiles are skipped, never written out of bounds).
    const int TS = 1 << g.tile_shift;
    const unsigned long long mc_max = std::min<unsigned long long>(nt, B.bin_cap / kChunk + 1);
    const size_t stage_bytes = (size_t)mc_max * 3 * TS * TS * sizeof(unsigned long long);
    const bool fresh = !(bst.p && stage_bytes <= bst.cap);
    if ((st = bst.ensure(stage_bytes))) return st;
    // The finish pass zeroes every slot it drains, so only a new allocation
    // needs clearing (here, outside any captured sequence).
    if (fresh) GL_HIP(hipMemsetAsync(bst.p, 0, bst.cap, c->stream));
    B.stage = bst.as<unsigned long long>();
    B.stage_tiles = (unsigned)(bst.cap / ((size_t)3 * TS * TS * sizeof(unsigned long long)));
    return GL_OK;
}

int splat_clear(hipStream_t s, const SplatGeom& g, const SplatBufs& B) {
    hipLaunchKernelGGL(splat_clear_kernel, dim3((unsigned)((g.n_tiles + 255) / 256)), dim3(256), 0, s, B, g.n_tiles);
    GL_HIP(hipGetLastError());
    return GL_OK;
}

int splat_resolve(hipStream_t s, const SplatGeom& g, const SplatBufs& B, const float4* spec, float* out_r, float* out_g,
                  float* out_b) {
    const size_t hist_bytes = g.lds_hist ? (size_t)g.n_tiles * 8 : 0;
    const size_t scat_bytes = g.lds_hist ? (size_t)((g.n_tiles + 1) & ~1) * 4 + (size_t)g.n_tiles * 8 : 0;
    if (g.lds_hist) hipLaunchKernelGGL(bin_count_kernel<true>, dim3(kBinBlocks), dim3(kBinThreads), hist_bytes, s, B, g, spec);
    else hipLaunchKernelGGL(bin_count_kernel<false>, dim3(kBinBlocks), dim3(kBinThreads), 0, s, B, g, spec);
    hipLaunchKernelGGL(bin_scan_kernel, dim3(1), dim3(1024), 0, s, B, g);
    hipLaunchKernelGGL(bin_scatter_kernel, dim3(kBinBlocks), dim3(kBinThreads), scat_bytes, s, B, g);
    const int TS = 1 << g.tile_shift;
    hipLaunchKernelGGL(resolve_kernel, dim3(kResolveBlocks), dim3(256), (size_t)3 * TS * TS * sizeof(unsigned long long),
                       s, B, g, spec, out_r, out_g, out_b);
    hipLaunchKernelGGL(resolve_finish_kernel, dim3(kFinishBlocks), dim3(256), 0, s, B, g, out_r, out_g, out_b);
    GL_HIP(hipGetLastError());
    return GL_OK;
}

}  // namespace gl
