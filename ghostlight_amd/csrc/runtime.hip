// runtime.hip — error state, device discovery, render contexts and
// device-resident lenses.
#include "gl_runtime.h"

#include <algorithm>
#include <atomic>

#include <cstdio>

namespace gl {

static thread_local std::string t_err = "no error";

int set_error(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    t_err = buf;
    return code;
}

std::mutex& device_mem_mutex() {
    static std::mutex m;
    return m;
}

int DevBuf::ensure(size_t bytes) {
    if (bytes <= cap && p) return GL_OK;
    std::lock_guard<std::mutex> g(device_mem_mutex());
    if (p) (void)hipFree(p);
    p = nullptr; cap = 0;
    size_t want = bytes < 256 ? 256 : bytes;
    hipError_t e = hipMalloc(&p, want);
    if (e != hipSuccess) {
        p = nullptr;
        return set_error(e == hipErrorOutOfMemory ? GL_ERR_NO_MEMORY : GL_ERR_DEVICE,
                         "hipMalloc(%zu) -- %s", want, hipGetErrorString(e));
    }
    cap = want;
    return GL_OK;
}

void DevBuf::release() {
    if (!p) return;
    std::lock_guard<std::mutex> g(device_mem_mutex());
    (void)hipFree(p);
    p = nullptr; cap = 0;
}

static int device_ok(int* dev) {
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n <= 0)
        return set_error(GL_ERR_DEVICE, "no HIP device available (%s)",
                         e == hipSuccess ? "device count 0" : hipGetErrorString(e));
    GL_HIP(hipGetDevice(dev));
    return GL_OK;
}

static int ctx_init(gl_ctx* c) {
    int st = device_ok(&c->device);
    if (st) return st;
    GL_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    GL_HIP(hipEventCreate(&c->ev0));
    GL_HIP(hipEventCreate(&c->ev1));
    GL_HIP(hipEventCreate(&c->ev2));
    return GL_OK;
}

gl_ctx* resolve_ctx(gl_ctx* ctx, int* status) {
    *status = GL_OK;
    if (ctx) return ctx;
    // Process default context: created once, never destroyed (the runtime may
    // already be torn down at exit).
    static std::mutex m;
    static gl_ctx* def = nullptr;
    std::lock_guard<std::mutex> g(m);
    if (!def) {
        gl_ctx* c = new gl_ctx();
        int st = ctx_init(c);
        if (st) { delete c; *status = st; return nullptr; }
        def = c;
    }
    return def;
}

int lens_sync(gl_lens* L, float fp64_rmin, hipStream_t stream, LensDev* out) {
    int dev = 0;
    GL_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> g(L->mu);
    if (L->device != dev) {
        if (L->device >= 0) return set_error(GL_ERR_INVALID, "lens was created for device %d, used on device %d",
                                             L->device, dev);
        {
            std::lock_guard<std::mutex> ga(device_mem_mutex());
            GL_HIP(hipMalloc(&L->d_arena, std::max<size_t>(1, L->arena.size()) * sizeof(float)));
        }
        GL_HIP(hipMemcpyAsync(L->d_arena, L->arena.data(), L->arena.size() * sizeof(float), hipMemcpyHostToDevice,
                              stream));
        L->device = dev;
    }
    const gl_lens::Variant* v = nullptr;
    for (const auto& x : L->variants)
        if (x.fp64_rmin == fp64_rmin) { v = &x; break; }
    if (!v) {
        std::vector<DevSurf> ds;
        std::vector<float> poly;
        build_devsurfs(L->host.data(), (int)L->host.size(), fp64_rmin, L->arena.data(), (int)L->arena.size(), ds, poly);
        gl_lens::Variant nv{fp64_rmin, nullptr, nullptr, lean_level(ds), coating_classes(ds)};
        {
            std::lock_guard<std::mutex> ga(device_mem_mutex());
            GL_HIP(hipMalloc(&nv.d_surf, std::max<size_t>(1, ds.size()) * sizeof(DevSurf)));
            GL_HIP(hipMalloc(&nv.d_poly, std::max<size_t>(1, poly.size()) * sizeof(float)));
        }
        GL_HIP(hipMemcpyAsync(nv.d_surf, ds.data(), ds.size() * sizeof(DevSurf), hipMemcpyHostToDevice, stream));
        GL_HIP(hipMemcpyAsync(nv.d_poly, poly.data(), poly.size() * sizeof(float), hipMemcpyHostToDevice, stream));
        GL_HIP(hipStreamSynchronize(stream));
        L->variants.push_back(nv);
        v = &L->variants.back();
    }
    out->surf = v->d_surf;
    out->poly = v->d_poly;
    out->arena = L->d_arena;
    out->lean = v->lean;
    out->coat_classes = v->coat_classes;
    return GL_OK;
}

int coating_classes(const std::vector<DevSurf>& ds) {
    int cls = 0;
    bool all_simple = true;
    for (const DevSurf& d : ds) {
        if (!(d.flags & SF_ACTIVE)) continue;
        const int m = d.coat_model;
        if (m != GL_COAT_SIMPLE) all_simple = false;
        if (m == GL_COAT_SIMPLE) cls |= d.ar_layers <= 0 ? CC_FRESNEL : CC_AIRY;
        else if (m == GL_COAT_SPECTRAL_ANGULAR) cls |= (d.sa_nw > 0 && d.sa_na > 0) ? CC_SA : CC_FRESNEL;
        else if (m == GL_COAT_SPECTRAL) cls |= d.tab_n <= 0 ? CC_FRESNEL : CC_ROW;
        else if (m == GL_COAT_ANGULAR) cls |= d.tab_n <= 0 ? CC_FRESNEL : CC_TABLE;
        else if (m == GL_COAT_ARTIST || m == GL_COAT_ATTENUATOR_GAUSS) cls |= CC_ROW;
        else cls |= CC_FRESNEL;
    }
    if ((cls & ~(CC_SA | CC_FRESNEL)) == 0) return CC_SA | CC_FRESNEL;
    if (all_simple) return CC_AIRY | CC_FRESNEL;
    return CC_ALL;
}

int lean_level(const std::vector<DevSurf>& ds) {
    const int32_t heavy = SF_FP64 | SF_DEFORMED | SF_IMAGE;
    int level = 1;
    for (const DevSurf& d : ds) {
        if (!(d.flags & SF_ACTIVE)) continue;
        if (d.flags & heavy) return 0;
        if (d.coat_model == GL_COAT_ATTENUATOR_GAUSS) return 0;
        const bool flat = (d.flags & SF_FLAT) != 0;
        if (!flat && d.form != GL_FORM_SPHERE && d.form != GL_FORM_CYLINDRICAL) return 0;
        if ((d.flags & (SF_TRANSFORMED | SF_ASPECT)) || (!flat && d.form == GL_FORM_CYLINDRICAL)) level = 2;
    }
    return level;
}

}  // namespace gl

extern "C" {

const char* gl_last_error(void) { return gl::t_err.c_str(); }

int gl_device_available(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return 0;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, 0) != hipSuccess) return 0;
    return std::strncmp(prop.gcnArchName, "gfx950", 6) == 0 ? 1 : 0;
}

int gl_ctx_create(gl_ctx** out) {
    if (!out) return gl::set_error(GL_ERR_INVALID, "null out");
    gl_ctx* c = new gl_ctx();
    int st = gl::ctx_init(c);
    if (st) { delete c; *out = nullptr; return st; }
    *out = c;
    return GL_OK;
}

void gl_ctx_destroy(gl_ctx* c) {
    if (!c) return;
    {
        std::lock_guard<std::mutex> g(c->mu);
        if (c->stream) (void)hipStreamSynchronize(c->stream);
        for (auto& kv : c->bufs) kv.second.release();
        if (c->ev0) (void)hipEventDestroy(c->ev0);
        if (c->ev1) (void)hipEventDestroy(c->ev1);
        if (c->ev2) (void)hipEventDestroy(c->ev2);
        {
            std::lock_guard<std::mutex> gm(gl::device_mem_mutex());
            if (c->graph) (void)hipGraphExecDestroy(c->graph);
            if (c->pin) (void)hipHostFree(c->pin);
            if (c->stream) (void)hipStreamDestroy(c->stream);
        }
    }
    delete c;
}

int gl_ctx_synchronize(gl_ctx* ctx) {
    int st;
    gl_ctx* c = gl::resolve_ctx(ctx, &st);
    if (!c) return st;
    GL_HIP(hipStreamSynchronize(c->stream));
    return GL_OK;
}

int gl_device_memory(uint64_t* free_bytes, uint64_t* total_bytes) {
    size_t f = 0, t = 0;
    GL_HIP(hipMemGetInfo(&f, &t));
    if (free_bytes) *free_bytes = f;
    if (total_bytes) *total_bytes = t;
    return GL_OK;
}

uint64_t gl_ctx_memory_bytes(gl_ctx* ctx) {
    int st;
    gl_ctx* c = gl::resolve_ctx(ctx, &st);
    if (!c) return 0;
    std::lock_guard<std::mutex> g(c->mu);
    uint64_t n = 0;
    for (const auto& kv : c->bufs) n += kv.second.cap;
    return n;
}

int gl_ctx_trim(gl_ctx* ctx) {
    int st;
    gl_ctx* c = gl::resolve_ctx(ctx, &st);
    if (!c) return st;
    std::lock_guard<std::mutex> g(c->mu);
    GL_HIP(hipStreamSynchronize(c->stream));
    for (auto& kv : c->bufs) kv.second.release();
    c->bufs.clear();
    {
        std::lock_guard<std::mutex> gm(gl::device_mem_mutex());
        if (c->graph) (void)hipGraphExecDestroy(c->graph);
        c->graph = nullptr;
        if (c->pin) (void)hipHostFree(c->pin);
        c->pin = nullptr;
        c->pin_cap = c->pin_used = 0;
    }
    c->graph_key.clear();
    c->work_host.clear();
    c->work_key.clear();
    c->vconv_key.clear();
    c->ltab_key.clear();
    c->probe_key.clear();
    c->grid_uploaded = false;
    c->spec_uploaded = false;
    c->veil_baked_p = -1.0f;
    return GL_OK;
}

int gl_ctx_last_schedule(gl_ctx* ctx, int32_t out[8]) {
    if (!out) return gl::set_error(GL_ERR_INVALID, "null out");
    int st;
    gl_ctx* c = gl::resolve_ctx(ctx, &st);
    if (!c) return st;
    for (int i = 0; i < 8; ++i) out[i] = c->last_sched[i];
    return GL_OK;
}

float gl_ctx_last_splat_ms(gl_ctx* ctx) {
    int st;
    gl_ctx* c = gl::resolve_ctx(ctx, &st);
    return c ? c->last_splat_ms : -1.0f;
}

int gl_ctx_last_splat_counts(gl_ctx* ctx, uint64_t out[4]) {
    if (!out) return gl::set_error(GL_ERR_INVALID, "null out");
    int st;
    gl_ctx* c = gl::resolve_ctx(ctx, &st);
    if (!c) return st;
    for (int i = 0; i < 4; ++i) out[i] = c->splat_counts[i];
    return GL_OK;
}

float gl_ctx_last_kernel_ms(gl_ctx* ctx) {
    int st;
    gl_ctx* c = gl::resolve_ctx(ctx, &st);
    return c ? c->last_kernel_ms : -1.0f;
}

int gl_lens_create(const gl_surface* s, int32_t n, const float* arena, int32_t arena_len, gl_lens** out) {
    if (!out) return gl::set_error(GL_ERR_INVALID, "null out");
    *out = nullptr;
    if (!s || n <= 0 || n > GL_MAX_SURFACES)
        return gl::set_error(GL_ERR_INVALID, "surface count %d outside [1, %d]", n, GL_MAX_SURFACES);
    for (int i = 0; i < n; ++i) {
        const gl_surface& g = s[i];
        if (g.table_count > 0 && (!arena || g.table_offset < 0 || g.table_offset + 2 * g.table_count > arena_len))
            return gl::set_error(GL_ERR_INVALID, "surface %d: coating table outside the arena", i);
        if (g.sa_n_wavelengths > 0 &&
            (!arena || g.sa_offset < 0 ||
             g.sa_offset + g.sa_n_wavelengths + g.sa_n_angles + g.sa_n_wavelengths * g.sa_n_angles > arena_len))
            return gl::set_error(GL_ERR_INVALID, "surface %d: spectral-angular table outside the arena", i);
        if (g.aperture_shape == GL_APERTURE_IMAGE && g.image_width > 0 &&
            (!arena || g.image_offset < 0 || g.image_height <= 0 ||
             (long long)g.image_offset + (long long)g.image_width * g.image_height > (long long)arena_len))
            return gl::set_error(GL_ERR_INVALID, "surface %d: aperture bitmap outside the arena", i);
    }
    static std::atomic<uint64_t> next_uid{1};
    gl_lens* L = new gl_lens();
    L->uid = next_uid.fetch_add(1);
    L->host.assign(s, s + n);
    if (arena && arena_len > 0) L->arena.assign(arena, arena + arena_len);
    else L->arena.assign(1, 0.0f);
    *out = L;
    return GL_OK;
}

void gl_lens_destroy(gl_lens* L) {
    if (!L) return;
    {
        std::lock_guard<std::mutex> g(gl::device_mem_mutex());
        for (auto& v : L->variants) { (void)hipFree(v.d_surf); (void)hipFree(v.d_poly); }
        if (L->d_arena) (void)hipFree(L->d_arena);
    }
    delete L;
}

int gl_enumerate_ghost_pairs(const gl_lens* L, gl_pair* out, int32_t cap, int32_t* n_out) {
    if (!L || !n_out) return gl::set_error(GL_ERR_INVALID, "null argument");
    const int n = (int)L->host.size();
    int cnt = 0;
    for (int a = 0; a < n; ++a) {
        if (!L->host[a].is_active) continue;
        for (int b = a + 1; b < n; ++b) {
            if (!L->host[b].is_active) continue;
            if (cnt < cap && out) out[cnt] = gl_pair{a, b, 1.0f};
            ++cnt;
        }
    }
    *n_out = cnt;
    return cnt > cap ? gl::set_error(GL_ERR_INVALID, "pair capacity %d < %d", cap, cnt) : GL_OK;
}

}  // extern "C"
