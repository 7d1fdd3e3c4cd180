// trace_diag.hip — diagnostic single-ray traces: the landing plus one record
// per surface the ray reaches (the RayPath overloads of trace_ghost_ray /
// trace_primary_ray, trace.cpp:264-395).
//
// One lane per ray.  The surface interaction is the tracer's own (intersect,
// image-aperture clip, coating weight, attenuator, reflect / refract with the
// same arguments and float operations as step() in gl_device.h), with a
// record written at each outcome the reference's recorder sees:
//   missed / clipped by the aperture  -> VIGNETTED at the ray's origin, zero
//                                        normal, Fresnel factor 0
//   clipped by an image aperture      -> VIGNETTED at the hit, with its normal
//   total internal reflection         -> TIR at the hit, Fresnel factor 0
//   reflected at a ghost bounce       -> OK, reflectance, reflected = 1
//   transmitted                       -> OK, transmittance
// Muted surfaces are passed through with no record, as in the reference.
#include "gl_runtime.h"

#include <algorithm>
#include <cstring>

static_assert(sizeof(gl_trace_event) == 48, "gl_trace_event is 48 bytes (ctypes mirrors it)");

namespace gl {
namespace {

struct EventSink {
    gl_trace_event* ev;
    int cap;
    int n;
    __device__ void put(int s, F3 hit, F3 nrm, float n1, float n2, float fw, bool reflected, int status) {
        if (n >= cap) return;
        gl_trace_event e;
        e.surface_index = s;
        e.hit_point[0] = hit.x; e.hit_point[1] = hit.y; e.hit_point[2] = hit.z;
        e.surface_normal[0] = nrm.x; e.surface_normal[1] = nrm.y; e.surface_normal[2] = nrm.z;
        e.ior_before = n1;
        e.ior_after = n2;
        e.fresnel_weight = fw;
        e.status = status;
        e.reflected = reflected ? 1 : 0;
        ev[n++] = e;
    }
};

__device__ bool diag_step(const DevSurf GLC& s, int k, const float GLC* poly, const float* __restrict__ arena, F3& o,
                          F3& d, float lam, float n2, bool reflecting, float reflect_ior, float& cur, TraceRes& res,
                          EventSink& sink) {
    float t; F3 nrm;
    ++res.segs;
    if (!intersect<false>(o, d, s, poly, t, nrm)) {
        sink.put(k, o, f3(0.0f, 0.0f, 0.0f), cur, n2, 0.0f, false, GL_TRACE_VIGNETTED);
        return false;
    }
    const F3 hit = add(o, mul(d, t));
    o = hit;
    if ((s.flags & SF_IMAGE) && !image_passes(s, arena, hit)) {
        sink.put(k, hit, nrm, cur, n2, 0.0f, false, GL_TRACE_VIGNETTED);
        return false;
    }
    const float n1 = cur;
    if (reflecting) {
        const float fw = surface_weight<CC_ALL>(fabsf(dot3(nrm, d)), n1, n2, n1 / n2, s, arena, lam, true, nullptr);
        res.w *= fw;
        res.w *= attenuator(hit, s);
        sink.put(k, hit, nrm, n1, n2, fw, true, GL_TRACE_OK);
        d = reflect(d, nrm);
        cur = reflect_ior;
    } else {
        F3 nd;
        if (!refract_ratio(d, nrm, n1 / n2, nd)) {
            sink.put(k, hit, nrm, n1, n2, 0.0f, false, GL_TRACE_TIR);
            res.status = GL_TRACE_TIR;
            return false;
        }
        const float fw = surface_weight<CC_ALL>(fabsf(dot3(nrm, d)), n1, n2, n1 / n2, s, arena, lam, false, nullptr);
        res.w *= fw;
        res.w *= attenuator(hit, s);
        sink.put(k, hit, nrm, n1, n2, fw, false, GL_TRACE_OK);
        d = nd;
        cur = n2;
    }
    return true;
}

// Three legs as the tracer runs them: forward 0..b reflecting at b, backward
// b-1..a against the index before each surface reflecting at a, forward
// a+1..n-1; b < 0 is the primary path (forward 0..n-1).
__global__ __launch_bounds__(64) void diag_kernel(const DevSurf* __restrict__ S, const float* __restrict__ poly,
                                                  const float* __restrict__ arena, int n, int a, int b,
                                                  const float* __restrict__ rays, int n_rays,
                                                  const float* __restrict__ iat_all, const float* __restrict__ ibf_all,
                                                  int t_ge0, int max_events, gl_trace_event* __restrict__ events,
                                                  int* __restrict__ n_events, float* __restrict__ pos,
                                                  float* __restrict__ w, int* __restrict__ st) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_rays) return;
    const float* q = rays + 7 * (size_t)i;
    const float* iat = iat_all + (size_t)i * n;
    const float* ibf = ibf_all + (size_t)i * n;
    const DevSurf GLC* SC = as_const(S);
    const float GLC* PC = as_const(poly);
    EventSink sink{events + (size_t)i * max_events, max_events, 0};
    TraceRes res; res.status = GL_TRACE_VIGNETTED; res.w = 1.0f; res.segs = 0; res.px = res.py = res.pz = 0.0f;
    F3 o = f3(q[0], q[1], q[2]), d = f3(q[3], q[4], q[5]);
    const float lam = q[6];
    float cur = 1.0f;
    bool alive = true;
    if (b < 0) {
        for (int k = 0; k < n && alive; ++k) {
            if (!(SC[k].flags & SF_ACTIVE)) continue;
            alive = diag_step(SC[k], k, PC, arena, o, d, lam, iat[k], false, 0.0f, cur, res, sink);
        }
    } else {
        for (int k = 0; k <= b && alive; ++k) {
            if (!(SC[k].flags & SF_ACTIVE)) continue;
            alive = diag_step(SC[k], k, PC, arena, o, d, lam, iat[k], k == b, cur, cur, res, sink);
        }
        for (int k = b - 1; k >= a && alive; --k) {
            if (!(SC[k].flags & SF_ACTIVE)) continue;
            const float rior = k == a ? iat[a] : cur;
            alive = diag_step(SC[k], k, PC, arena, o, d, lam, ibf[k], k == a, rior, cur, res, sink);
        }
        for (int k = a + 1; k < n && alive; ++k) {
            if (!(SC[k].flags & SF_ACTIVE)) continue;
            alive = diag_step(SC[k], k, PC, arena, o, d, lam, iat[k], false, cur, cur, res, sink);
        }
    }
    if (alive) to_sensor(o, d, t_ge0 != 0, res);
    pos[3 * i] = res.px; pos[3 * i + 1] = res.py; pos[3 * i + 2] = res.pz;
    w[i] = res.w;
    st[i] = res.status;
    n_events[i] = sink.n;
}

// Every float bit pattern in [lo, hi] through one of the tracer's fast
// correctly rounded forms and through the IEEE operation it stands for (this
// unit is built with IEEE division and square root); counts the patterns
// whose bits differ and keeps the smallest one.
//   OP 0: sqrt_rn_big(x) against sqrtf(x)
//   OP 1: rcp_rn(x) against 1.0f / x
//   OP 2: 0.5f * rcp_rn(x) against 0.5f / x (the sphere's 1 / 2a)
//   OP 3: sqrt_rn_pos(x) against sqrtf(x)
template <int OP>
__global__ __launch_bounds__(256) void rn_check_kernel(uint32_t lo, uint32_t hi, unsigned long long* bad,
                                                      unsigned* first) {
    unsigned long long nbad = 0;
    unsigned fb = 0xFFFFFFFFu;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t u = (uint64_t)lo + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; u <= (uint64_t)hi; u += stride) {
        const float x = __uint_as_float((uint32_t)u);
        float a, b;
        if (OP == 0) { a = sqrt_rn_big(x); b = sqrtf(x); }
        else if (OP == 1) { a = rcp_rn(x); b = 1.0f / x; }
        else if (OP == 2) { a = 0.5f * rcp_rn(x); b = 0.5f / x; }
        else { a = sqrt_rn_pos(x); b = sqrtf(x); }
        if (__float_as_uint(a) != __float_as_uint(b)) {
            ++nbad;
            fb = min(fb, (uint32_t)u);
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        nbad += __shfl_xor(nbad, off, 64);
        fb = min(fb, (unsigned)__shfl_xor((int)fb, off, 64));
    }
    if ((threadIdx.x & 63) == 0 && nbad) {
        atomicAdd(bad, nbad);
        atomicMin(first, fb);
    }
}

}  // namespace
}  // namespace gl

using namespace gl;

extern "C" {

int gl_trace_rays_diagnostic(gl_ctx* ctx, const gl_lens* lens, int32_t a, int32_t b, const float* rays,
                             int32_t n_rays, int32_t t_ge0, float fp64_rmin, int32_t max_events,
                             gl_trace_event* events, int32_t* n_events, float* out_pos, float* out_w,
                             int32_t* out_status) {
    if (!lens || !rays || n_rays < 0 || !events || !n_events || !out_pos || !out_w || !out_status)
        return set_error(GL_ERR_INVALID, "null argument");
    const int n = (int)lens->host.size();
    if (max_events < 3 * n) return set_error(GL_ERR_INVALID, "max_events %d < 3 x %d surfaces", max_events, n);
    if (n_rays == 0) return GL_OK;
    if (b >= 0 && (a < 0 || b <= a || b >= n)) {
        for (int i = 0; i < n_rays; ++i) {
            out_status[i] = GL_TRACE_INVALID_INPUT;
            out_w[i] = 0.0f;
            out_pos[3 * i] = out_pos[3 * i + 1] = out_pos[3 * i + 2] = 0.0f;
            n_events[i] = 0;
        }
        return GL_OK;
    }
    int st;
    gl_ctx* c = resolve_ctx(ctx, &st);
    if (!c) return st;
    std::lock_guard<std::mutex> g(c->mu);
    gl_lens* L = const_cast<gl_lens*>(lens);
    LensDev D;
    if ((st = lens_sync(L, fp64_rmin, c->stream, &D))) return st;
    // Per-ray index slices (each ray carries its own wavelength).
    std::vector<float> iat((size_t)n_rays * n), ibf((size_t)n_rays * n);
    for (int i = 0; i < n_rays; ++i) {
        gl_spec sp{rays[7 * (size_t)i + 6], 0, 0, 0};
        ior_tables(L->host.data(), n, &sp, 1, &iat[(size_t)i * n], &ibf[(size_t)i * n]);
    }
    const size_t ev_bytes = (size_t)n_rays * max_events * sizeof(gl_trace_event);
    DevBuf &br = c->buf("diag_rays"), &ba = c->buf("diag_iat"), &bb = c->buf("diag_ibf"), &be = c->buf("diag_ev");
    DevBuf &bn = c->buf("diag_n"), &bp = c->buf("diag_pos"), &bw = c->buf("diag_w"), &bs = c->buf("diag_st");
    if ((st = br.ensure((size_t)n_rays * 28)) || (st = ba.ensure(iat.size() * 4)) || (st = bb.ensure(ibf.size() * 4)) ||
        (st = be.ensure(ev_bytes)) || (st = bn.ensure((size_t)n_rays * 4)) || (st = bp.ensure((size_t)n_rays * 12)) ||
        (st = bw.ensure((size_t)n_rays * 4)) || (st = bs.ensure((size_t)n_rays * 4)))
        return st;
    GL_HIP(hipMemcpyAsync(br.p, rays, (size_t)n_rays * 28, hipMemcpyHostToDevice, c->stream));
    GL_HIP(hipMemcpyAsync(ba.p, iat.data(), iat.size() * 4, hipMemcpyHostToDevice, c->stream));
    GL_HIP(hipMemcpyAsync(bb.p, ibf.data(), ibf.size() * 4, hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(diag_kernel, dim3((unsigned)((n_rays + 63) / 64)), dim3(64), 0, c->stream, D.surf, D.poly,
                       D.arena, n, a, b, br.as<float>(), n_rays, ba.as<float>(), bb.as<float>(), t_ge0, max_events,
                       be.as<gl_trace_event>(), bn.as<int>(), bp.as<float>(), bw.as<float>(), bs.as<int>());
    GL_HIP(hipGetLastError());
    GL_HIP(hipMemcpyAsync(events, be.p, ev_bytes, hipMemcpyDeviceToHost, c->stream));
    GL_HIP(hipMemcpyAsync(n_events, bn.p, (size_t)n_rays * 4, hipMemcpyDeviceToHost, c->stream));
    GL_HIP(hipMemcpyAsync(out_pos, bp.p, (size_t)n_rays * 12, hipMemcpyDeviceToHost, c->stream));
    GL_HIP(hipMemcpyAsync(out_w, bw.p, (size_t)n_rays * 4, hipMemcpyDeviceToHost, c->stream));
    GL_HIP(hipMemcpyAsync(out_status, bs.p, (size_t)n_rays * 4, hipMemcpyDeviceToHost, c->stream));
    GL_HIP(hipStreamSynchronize(c->stream));
    return GL_OK;
}

int gl_selftest_rn(gl_ctx* ctx, int32_t op, uint32_t lo_bits, uint32_t hi_bits, uint64_t* mismatches,
                   uint32_t* first_mismatch) {
    if (!mismatches || !first_mismatch) return set_error(GL_ERR_INVALID, "null argument");
    if (lo_bits > hi_bits) return set_error(GL_ERR_INVALID, "empty bit range");
    if (op < 0 || op > 3)
        return set_error(GL_ERR_INVALID, "op must be 0 (sqrt), 1 (reciprocal), 2 (0.5 / x) or 3 (sqrt of x >= 0)");
    int st;
    gl_ctx* c = resolve_ctx(ctx, &st);
    if (!c) return st;
    std::lock_guard<std::mutex> g(c->mu);
    DevBuf& b = c->buf("rn_check");
    if ((st = b.ensure(16))) return st;
    const uint32_t init[4] = {0u, 0u, 0xFFFFFFFFu, 0u};
    GL_HIP(hipMemcpyAsync(b.p, init, 16, hipMemcpyHostToDevice, c->stream));
    unsigned long long* d_bad = b.as<unsigned long long>();
    unsigned* d_first = reinterpret_cast<unsigned*>(b.as<char>() + 8);
    const dim3 grid(8192), block(256);
    if (op == 0) hipLaunchKernelGGL(rn_check_kernel<0>, grid, block, 0, c->stream, lo_bits, hi_bits, d_bad, d_first);
    else if (op == 1) hipLaunchKernelGGL(rn_check_kernel<1>, grid, block, 0, c->stream, lo_bits, hi_bits, d_bad, d_first);
    else if (op == 2) hipLaunchKernelGGL(rn_check_kernel<2>, grid, block, 0, c->stream, lo_bits, hi_bits, d_bad, d_first);
    else hipLaunchKernelGGL(rn_check_kernel<3>, grid, block, 0, c->stream, lo_bits, hi_bits, d_bad, d_first);
    GL_HIP(hipGetLastError());
    uint32_t out[4];
    GL_HIP(hipMemcpyAsync(out, b.p, 16, hipMemcpyDeviceToHost, c->stream));
    GL_HIP(hipStreamSynchronize(c->stream));
    *mismatches = (uint64_t)out[0] | ((uint64_t)out[1] << 32);
    *first_mismatch = out[2];
    return GL_OK;
}

int gl_selftest_sqrt_rn(gl_ctx* ctx, uint32_t lo_bits, uint32_t hi_bits, uint64_t* mismatches,
                        uint32_t* first_mismatch) {
    return gl_selftest_rn(ctx, 0, lo_bits, hi_bits, mismatches, first_mismatch);
}

}  // extern "C"
