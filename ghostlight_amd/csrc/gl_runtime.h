// gl_runtime.h — host-side runtime shared by the HIP translation units:
// error state, contexts (stream + grow-only device buffers), device lenses.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/ghostlight_amd.h"
#include "gl_device.h"

namespace gl {

int set_error(int code, const char* fmt, ...);
#define GL_HIP(call)                                                                    \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess) return ::gl::set_error(GL_ERR_DEVICE, "%s -- %s", #call,  \
                                                     hipGetErrorString(e_));            \
    } while (0)

// Grow-only device allocation owned by a context.
struct DevBuf {
    void* p = nullptr;
    size_t cap = 0;
    int ensure(size_t bytes);
    template <class T> T* as() const { return static_cast<T*>(p); }
    void release();
};

struct HostCacheKey;

struct gl_ctx_impl;
}  // namespace gl

// One render context: a default-priority HIP stream, its buffers, caches.
struct gl_ctx {
    hipStream_t stream = nullptr;
    int device = 0;
    std::mutex mu;                       // one render at a time per context
    std::map<std::string, gl::DevBuf> bufs;
    hipEvent_t ev0 = nullptr, ev1 = nullptr, ev2 = nullptr;
    float last_kernel_ms = 0.0f;         // main trace kernel, event-timed
    float last_splat_ms = 0.0f;          // tiled splat resolve after it, event-timed
    unsigned long long splat_hint = 0;   // deposit records the last render produced
    unsigned long long splat_counts[4] = {0, 0, 0, 0};  // records, binned records, chunks, staged tiles
    // Persistent-launch autotuner: per workload shape, the measured kernel
    // time of each schedule candidate (-1 = not tried yet).
    struct TuneState { std::vector<float> ms; };
    std::map<std::vector<uint8_t>, TuneState> tune;
    std::vector<int4> work_host;         // work list currently on the device
    std::vector<uint8_t> work_key;       // inputs it was built from
    hipGraphExec_t graph = nullptr;      // captured trace + splat sequence
    std::vector<uint8_t> graph_key;
    int last_sched[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    int n_cu = 0, wall_khz = 0;
    std::vector<uint8_t> vconv_key;      // geometry of the cached veil-kernel transform
    // Host caches (pupil grid, spectral table) keyed like the reference's
    // GridKey / SpecKey (ghost_render.cu:515-556).
    std::vector<uint8_t> grid_key, spec_key;
    std::vector<gl_sample> grid;
    std::vector<gl_spec> spec;
    float spec_C[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    bool spec_do_correct = false;
    bool grid_uploaded = false, spec_uploaded = false;
    float veil_baked_p = -1.0f;          // falloff the cached veil reference holds
    // Ghost-render tables of one (lens, spectral table) -- per-lambda IOR,
    // coating lookup rows, refraction ratios and their 3-lambda refine
    // slices -- and the probe sample sets of one (lens, blade override,
    // rotation), on the device while their inputs stay the same.
    std::vector<uint8_t> ltab_key, probe_key;
    std::vector<int> probe_n;            // samples in each probe pass (passes concatenated)
    int refine_n = 0;                    // samples in the refine grid
    // Pinned host staging of a ghost render's small transfers (pairs,
    // sources, masks, rects, budgets up; probe and render counters down):
    // slices are handed out in order within one render, which ends with a
    // stream synchronisation, so the next render starts on an idle buffer.
    char* pin = nullptr;
    size_t pin_cap = 0, pin_used = 0;
    int max_lds = 0;                     // per-workgroup LDS bytes (device attribute)
    gl::DevBuf& buf(const char* name) { return bufs[name]; }
};

namespace gl {
// Device-to-host copy ordered on the context's stream.  The library never
// touches the legacy (null) stream: a synchronous hipMemcpy there would have to
// wait on every blocking stream, including one another thread is capturing.
inline hipError_t d2h_sync(gl_ctx* c, void* dst, const void* src, size_t bytes) {
    const hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream);
    return e != hipSuccess ? e : hipStreamSynchronize(c->stream);
}

// Held around every hipMalloc/hipFree of the library and around each graph
// capture, so that no allocation or release (hipFree synchronises the device)
// falls inside another thread's capture sequence.
std::mutex& device_mem_mutex();

// Device view of a lens as the kernels read it.
struct LensDev {
    const DevSurf* surf = nullptr;
    const float* poly = nullptr;             // per polygon surface: c[32] | s[32]
    const float* arena = nullptr;
    int lean = 0;                            // the lens's lean level (lean_level)
    int coat_classes = CC_ALL;               // coating-weight specialisation (coating_classes)
};
}  // namespace gl

// Device-resident prescription.  The host copy never changes after creation;
// device variants (one per FP64-radius threshold in use) are created on first
// use and never modified, so concurrent renders on one lens need the mutex
// only while a variant is looked up or created.
struct gl_lens {
    uint64_t uid = 0;                        // unique per created lens: keys the contexts' per-lens caches
    std::vector<gl_surface> host;
    std::vector<float> arena;
    int device = -1;
    float* d_arena = nullptr;
    struct Variant {
        float fp64_rmin;
        gl::DevSurf* d_surf;
        float* d_poly;
        int lean;
        int coat_classes;
    };
    std::vector<Variant> variants;
    std::mutex mu;
};

namespace gl {
gl_ctx* resolve_ctx(gl_ctx* ctx, int* status);
int lens_sync(gl_lens* lens, float fp64_rmin, hipStream_t stream, LensDev* out);
void build_devsurfs(const gl_surface* s, int n, float fp64_rmin, const float* arena, int arena_len,
                    std::vector<DevSurf>& out, std::vector<float>& poly);
// The tracer instantiation that covers a lens: 1 when every active surface
// is flat or spherical, untransformed, solved in FP32, with a round or
// regular-polygon aperture of aspect 1; 2 when cylinders, decentred / tilted
// surfaces or stretched apertures are present as well; else 0 (the generic
// tracer: FP64 radii, aspheres, deformed blades, bitmaps or attenuators).
// Levels 1 and 2 take the predicated surface step.
int lean_level(const std::vector<DevSurf>& ds);
// The coating-weight specialisation a lens can use: CC_SA | CC_FRESNEL when
// every active surface is spectral-angular or a bare interface,
// CC_AIRY | CC_FRESNEL when every active surface is SIMPLE, else CC_ALL.
int coating_classes(const std::vector<DevSurf>& ds);
// Per-lambda coating lookups of the lambda-only part of surface_weight: one
// int2 entry per (lambda, surface) -- {-2, R bits} for a lambda-only value,
// {-1, 0} for a discarded wavelength, {offset, 0} for a spectral-angular
// table whose lambda-blended {R, dR} row sits `offset` words after the entry
// -- followed by those rows, in one buffer of 4-byte words.
void lambda_rows(const gl_surface* s, int n, const float* arena, const gl_spec* spec, int ns,
                 std::vector<int32_t>& buf);
DevMask build_mask(const gl_surface* s, int n, int blades_override, float rotation_deg);

// host preparation (hostprep.cpp)
std::vector<gl_sample> build_samples(const gl_surface* s, int n, int N, int jitter, int seed, int bo, float rdeg);
void ior_tables(const gl_surface* s, int n, const gl_spec* spec, int ns, float* at, float* bf);
float ior_at(const gl_surface& s, float lambda_nm);
int spectral_table(int n, int output_cs, const float* custom9, std::vector<gl_spec>& out);
int spectral_correction(const std::vector<gl_spec>& t, int output_cs, const float* custom9, float C[9]);
int spectral_table_cmf(int n, int output_cs, const float* custom9, std::vector<gl_spec>& out);
bool mask_contains_host(const DevMask& m, float u, float v);

// A source as the diffraction passes see it: buffer-pixel position + colour.
struct StarSource { float px, py, r, g, b; };

// Matrix-DFT starburst window (mdft.hip).
int mdft_render(gl_ctx* c, const float2* d_pupil, int N, float fill, const std::vector<float4>& spec, float lambda_ref,
                float dx_x, float dx_y, float px_mm_x, float px_mm_y, const std::vector<StarSource>& src, float gain,
                float* o_r, float* o_g, float* o_b, int width, int height, int precision);
void poly_normals(int n_blades, float rot_rad, float scale, float* c, float* s, float* apothem);

// 3xBF16 complex GEMM on pre-split operands (cgemm3.hip).  split_planes
// writes the 4 bf16 planes (re_hi, re_lo, im_hi, im_lo) of a batch of R x C
// complex matrices, transposed on request; cgemm_nt3 computes C = alpha A B
// from A (M x K) and B^T (N x K) planes, epi 0: float2, 1: |.|^2, 2: planes.
bool cgemm_nt3_fits(int M, int N, int K);
size_t cgemm_nt3_lds_bytes();
int split_planes(hipStream_t s, const float2* X, int ldx, long long sX, int R, int C, int transpose, int batch,
                 __bf16* out);
int cgemm_nt3(hipStream_t s, int epi, const __bf16* A, long long sA, const __bf16* Bt, long long sB, void* C, int ldc,
              long long sC, int M, int N, int K, int batch, float alpha);
double now_ms();
}  // namespace gl
