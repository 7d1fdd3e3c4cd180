// ORACLE — test infrastructure only (bench.py's roofline leg and tests).
// Algorithmic FP32 operation count of the ghost tracer: the oracle's tracer
// template instantiated with a counting scalar.  Every +, -, *, /, sqrt and
// transcendental (sin, cos, asin, atan2, fmod, exp) counts one operation; comparisons, fabs,
// fmin/fmax and conversions count none.  The count is what the reference's
// per-segment arithmetic costs (trace_core.h intersect / aperture / refract /
// coating), independent of how a compiler lowers it.
#include <cmath>
#include <cstdint>

namespace orc {
struct CF;
}
namespace std {
orc::CF sqrt(orc::CF);
orc::CF fabs(orc::CF);
orc::CF fmin(orc::CF, orc::CF);
orc::CF fmax(orc::CF, orc::CF);
orc::CF cos(orc::CF);
orc::CF asin(orc::CF);
orc::CF exp(orc::CF);
orc::CF sin(orc::CF);
orc::CF atan2(orc::CF, orc::CF);
orc::CF fmod(orc::CF, orc::CF);
orc::CF tan(orc::CF);
orc::CF log(orc::CF);
}  // namespace std

namespace orc {
thread_local uint64_t g_ops = 0;
struct CF {
    float v;
    CF() : v(0.f) {}
    CF(float x) : v(x) {}
    CF(double x) : v((float)x) {}
    CF(int x) : v((float)x) {}
    explicit operator double() const { return v; }
    explicit operator float() const { return v; }
    explicit operator int() const { return (int)v; }
    CF operator+(CF o) const { ++g_ops; return CF(v + o.v); }
    CF operator-(CF o) const { ++g_ops; return CF(v - o.v); }
    CF operator*(CF o) const { ++g_ops; return CF(v * o.v); }
    CF operator/(CF o) const { ++g_ops; return CF(v / o.v); }
    CF operator-() const { return CF(-v); }
    CF& operator+=(CF o) { ++g_ops; v += o.v; return *this; }
    CF& operator-=(CF o) { ++g_ops; v -= o.v; return *this; }
    CF& operator*=(CF o) { ++g_ops; v *= o.v; return *this; }
    bool operator<(CF o) const { return v < o.v; }
    bool operator>(CF o) const { return v > o.v; }
    bool operator<=(CF o) const { return v <= o.v; }
    bool operator>=(CF o) const { return v >= o.v; }
    bool operator==(CF o) const { return v == o.v; }
};
}  // namespace orc

namespace std {
orc::CF sqrt(orc::CF a) { ++orc::g_ops; return orc::CF(std::sqrt(a.v)); }
orc::CF fabs(orc::CF a) { return orc::CF(std::fabs(a.v)); }
orc::CF fmin(orc::CF a, orc::CF b) { return orc::CF(std::fmin(a.v, b.v)); }
orc::CF fmax(orc::CF a, orc::CF b) { return orc::CF(std::fmax(a.v, b.v)); }
orc::CF cos(orc::CF a) { ++orc::g_ops; return orc::CF(std::cos(a.v)); }
orc::CF asin(orc::CF a) { ++orc::g_ops; return orc::CF(std::asin(a.v)); }
orc::CF exp(orc::CF a) { ++orc::g_ops; return orc::CF(std::exp(a.v)); }
orc::CF sin(orc::CF a) { ++orc::g_ops; return orc::CF(std::sin(a.v)); }
orc::CF atan2(orc::CF a, orc::CF b) { ++orc::g_ops; return orc::CF(std::atan2(a.v, b.v)); }
orc::CF fmod(orc::CF a, orc::CF b) { ++orc::g_ops; return orc::CF(std::fmod(a.v, b.v)); }
orc::CF tan(orc::CF a) { ++orc::g_ops; return orc::CF(std::tan(a.v)); }
orc::CF log(orc::CF a) { ++orc::g_ops; return orc::CF(std::log(a.v)); }
}  // namespace std

#include "gl_oracle_trace.h"

extern "C" {

// Trace n rays (7 floats each; ior slices per ray: n x n_surf) through pair
// (a, b) with counting arithmetic.  out[0] = total ops, out[1] = total
// segments, out[2] = surviving traces.
int orc_count_ops(const gl_surface* s, int n, const float* arena, int a, int b, const float* rays, int n_rays,
                  const float* ior_at, const float* ior_before, uint64_t* out) {
    orc::Lens L;
    orc::lens_init(L, s, n, arena, 0.0f);
    uint64_t segs = 0, ok = 0;
    orc::g_ops = 0;
    for (int i = 0; i < n_rays; ++i) {
        const float* q = rays + 7 * (size_t)i;
        orc::Ray<orc::CF> r;
        r.o = orc::V3<orc::CF>(q[0], q[1], q[2]);
        r.d = orc::V3<orc::CF>(q[3], q[4], q[5]);
        orc::TraceOut<orc::CF> o = orc::trace_ghost<orc::CF>(L, a, b, r, orc::CF(q[6]), ior_at + (size_t)i * n,
                                                             ior_before + (size_t)i * n, false);
        segs += (uint64_t)o.segments;
        ok += o.status == GL_TRACE_OK;
    }
    out[0] = orc::g_ops; out[1] = segs; out[2] = ok;
    return 0;
}

}  // extern "C"
