// Device-side 2-D float math with the operation order of PlayRho's value types,
// so that every rounding matches the reference float build.  Compiled with
// -ffp-contract=off: no multiply-add is ever fused, as on x86-64 SSE.
//
// Reference definitions followed:
//   Dot / GetMagnitudeSquared start their sums from zero  (Math.hpp:274-334)
//   Cross = a.x*b.y - a.y*b.x                            (Math.hpp:365-386)
//   Vector / scalar multiplies by the reciprocal          (Vector.hpp:511-521)
//   Rotate / InverseRotate / Transform / MulT            (d2/Math.hpp:107-229)
//   UnitVec::Get(x, y) with the hypot fallback           (d2/UnitVec.hpp:175-236)
//   UnitVec::Get(angle) = {cos, sin}                     (UnitVec.cpp:27-30)
//   std::min / std::max / std::clamp comparison forms
//   AlmostEqual via nextafter, 4 ULP                      (Math.hpp:186-194)
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "sincosf.hpp"

namespace prxd {

struct V2 {
    float x, y;
};

__device__ __forceinline__ V2 mk(float x, float y) { return V2{x, y}; }
__device__ __forceinline__ V2 operator+(V2 a, V2 b) { return V2{a.x + b.x, a.y + b.y}; }
__device__ __forceinline__ V2 operator-(V2 a, V2 b) { return V2{a.x - b.x, a.y - b.y}; }
__device__ __forceinline__ V2 operator-(V2 a) { return V2{-a.x, -a.y}; }
__device__ __forceinline__ V2 mul(V2 a, float s) { return V2{a.x * s, a.y * s}; }
__device__ __forceinline__ V2 mul(float s, V2 a) { return V2{s * a.x, s * a.y}; }
__device__ __forceinline__ V2 divs(V2 a, float s) {
    const float inv = 1.0f / s;
    return V2{a.x * inv, a.y * inv};
}
__device__ __forceinline__ bool eq(V2 a, V2 b) { return a.x == b.x && a.y == b.y; }

__device__ __forceinline__ float dot(V2 a, V2 b) {
    float r = 0.0f;
    r += a.x * b.x;
    r += a.y * b.y;
    return r;
}
__device__ __forceinline__ float cross(V2 a, V2 b) {
    const float m = a.x * b.y;
    const float s = a.y * b.x;
    return m - s;
}
__device__ __forceinline__ float magsq(V2 v) {
    float r = 0.0f;
    r += v.x * v.x;
    r += v.y * v.y;
    return r;
}
__device__ __forceinline__ V2 revperp(V2 v) { return V2{-v.y, v.x}; }
__device__ __forceinline__ V2 fwdperp(V2 v) { return V2{v.y, -v.x}; }

__device__ __forceinline__ V2 rotate(V2 v, V2 q) {
    const float nx = (q.x * v.x) - (q.y * v.y);
    const float ny = (q.y * v.x) + (q.x * v.y);
    return V2{nx, ny};
}
__device__ __forceinline__ V2 invrotate(V2 v, V2 q) {
    const float nx = (q.x * v.x) + (q.y * v.y);
    const float ny = (q.x * v.y) - (q.y * v.x);
    return V2{nx, ny};
}
__device__ __forceinline__ V2 rotate_unit(V2 a, V2 amount) {
    return V2{a.x * amount.x - a.y * amount.y, a.y * amount.x + a.x * amount.y};
}

struct Xf {
    V2 p, q;
};
__device__ __forceinline__ V2 transform(V2 v, const Xf& xf) { return rotate(v, xf.q) + xf.p; }
__device__ __forceinline__ V2 invtransform(V2 v, const Xf& xf) { return invrotate(v - xf.p, xf.q); }
__device__ __forceinline__ Xf mulT(const Xf& A, const Xf& B) {
    const V2 dp = B.p - A.p;
    return Xf{invrotate(dp, A.q), rotate_unit(B.q, V2{A.q.x, -A.q.y})};
}

__device__ __forceinline__ V2 unit_from_angle(float a) {
    float sn, cs;
    prx_math::SinCosF<true>(a, &sn, &cs);
    return V2{cs, sn};
}

// GetTransformation(pos, localCenter)
__device__ __forceinline__ Xf get_xf(float px, float py, float pa, V2 lc) {
    const V2 q = unit_from_angle(pa);
    return Xf{V2{px, py} - rotate(lc, q), q};
}

__device__ __forceinline__ float hypot_f(float x, float y) {
    // Single rounding of the double-precision result, as the C library's
    // float hypot computes it.
    const double dx = double(x), dy = double(y);
    return float(__builtin_sqrt(dx * dx + dy * dy));
}

__device__ __forceinline__ V2 unit_vector(V2 v, V2 fallback) {
    const float m2 = v.x * v.x + v.y * v.y;
    if (__builtin_isnormal(m2)) {
        const float mag = __builtin_sqrtf(m2);
        const float inv = 1.0f / mag;
        return V2{v.x * inv, v.y * inv};
    }
    const float mag = hypot_f(v.x, v.y);
    if (!__builtin_isnormal(mag)) return fallback;
    return V2{v.x / mag, v.y / mag};
}

// UnitVec::Get(x, y) with the zero fallback, also giving the magnitude (zero
// when no direction is found) (UnitVec.hpp:175-236).
__device__ __forceinline__ V2 unit_polar(V2 v, float& mag) {
    const float m2 = v.x * v.x + v.y * v.y;
    if (__builtin_isnormal(m2)) {
        mag = __builtin_sqrtf(m2);
        const float inv = 1.0f / mag;
        return V2{v.x * inv, v.y * inv};
    }
    const float h = hypot_f(v.x, v.y);
    if (!__builtin_isnormal(h)) {
        mag = 0.0f;
        return V2{0.0f, 0.0f};
    }
    mag = h;
    return V2{v.x / h, v.y / h};
}

__device__ __forceinline__ float fmax_std(float a, float b) { return (a < b) ? b : a; }
__device__ __forceinline__ float fmin_std(float a, float b) { return (b < a) ? b : a; }
__device__ __forceinline__ float clamp_std(float v, float lo, float hi) {
    return (v < lo) ? lo : ((hi < v) ? hi : v);
}

__device__ __forceinline__ float nextafter_f(float a, float b) {
    if (a != a || b != b) return a + b;
    if (a == b) return b;
    if (a == 0.0f) {
        const std::uint32_t tiny = 1u;
        const float t = __uint_as_float(tiny);
        return (b > 0.0f) ? t : -t;
    }
    std::uint32_t u = __float_as_uint(a);
    if ((a < b) == (a > 0.0f)) ++u;
    else --u;
    return __uint_as_float(u);
}
__device__ __forceinline__ bool almost_equal(float a, float b) {
    int ulp = 4;
    for (; (a != b) && (ulp > 0); --ulp) a = nextafter_f(a, b);
    return a == b;
}
__device__ __forceinline__ bool almost_zero(float v) { return __builtin_fabsf(v) < 1.17549435e-38f; }

}  // namespace prxd
