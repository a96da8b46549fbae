// Single-precision sine and cosine that reproduce the results of the C
// library's sinf/cosf on the x86-64 hosts PlayRho runs on (glibc 2.28+), so
// that UnitVec::Get(angle) (UnitVec.cpp:27-30) gives bit-identical rotations on
// the GPU.  glibc's float sine/cosine are not correctly rounded; their result
// is the single rounding of a double-precision evaluation:
//
//   * |x| < 2^-12     : sin -> x,  cos -> 1
//   * |x| < pi/4      : odd/even minimax polynomials in x^2 evaluated in double
//   * |x| < 120       : quadrant n = round(x * 2/pi) from a 2^24-scaled product,
//                       r = x - n * pi/2, polynomial selected by n
//   * |x| >= 120      : Payne-Hanek style reduction with a 4/pi bit table
//
// This is the published method of ARM's optimized-routines sincosf (Szabolcs
// Nagy, 2018), which glibc adopted.  The x86-64 glibc selects a build of it
// compiled with FMA contraction on CPUs that support FMA; PRX_SINCOS_FMA picks
// that evaluation (fused multiply-adds in the reduction and the polynomials).
// Checks: tests/test_oracle_golden.py compares this restatement on the host
// with the host C library over a dense sample of angles; the device build is
// compared with the host library in tests/test_gpu_parity.py
// (test_device_sincos_matches_host_library).
//
// The polynomial coefficients, the reduction constants and the 4/pi bit table
// below are those of optimized-routines math/sincosf.h and
// math/sincosf_data.c, used under its licence:
//
//   Copyright (c) 2018, Arm Limited.
//   SPDX-License-Identifier: MIT OR Apache-2.0 WITH LLVM-exception
//
//   MIT License: Permission is hereby granted, free of charge, to any person
//   obtaining a copy of this software and associated documentation files (the
//   "Software"), to deal in the Software without restriction, including
//   without limitation the rights to use, copy, modify, merge, publish,
//   distribute, sublicense, and/or sell copies of the Software, and to permit
//   persons to whom the Software is furnished to do so, subject to the
//   following conditions: The above copyright notice and this permission
//   notice shall be included in all copies or substantial portions of the
//   Software.  THE SOFTWARE IS PROVIDED "AS IS", WITHOUT WARRANTY OF ANY KIND,
//   EXPRESS OR IMPLIED, INCLUDING BUT NOT LIMITED TO THE WARRANTIES OF
//   MERCHANTABILITY, FITNESS FOR A PARTICULAR PURPOSE AND NONINFRINGEMENT. IN
//   NO EVENT SHALL THE AUTHORS OR COPYRIGHT HOLDERS BE LIABLE FOR ANY CLAIM,
//   DAMAGES OR OTHER LIABILITY, WHETHER IN AN ACTION OF CONTRACT, TORT OR
//   OTHERWISE, ARISING FROM, OUT OF OR IN CONNECTION WITH THE SOFTWARE OR THE
//   USE OR OTHER DEALINGS IN THE SOFTWARE.
#pragma once

#include <cstdint>
#include <cstring>

#ifdef __HIPCC__
#define PRX_HD __host__ __device__ __forceinline__
#else
#define PRX_HD inline
#endif

namespace prx_math {

// The published method's constants: 2/pi * 2^24 and pi/2 for the reduction,
// the cosine (c0..c4) and sine (s1..s3) polynomial coefficients.
struct SinCosTab {
    double hpi_inv;  // 2/pi * 2^24
    double hpi;      // pi/2
    double c0, c1, c2, c3, c4;
    double s1, s2, s3;
};

// By value from a constant expression, so the coefficients fold into the
// code as immediates.
PRX_HD constexpr SinCosTab Tab() {
    return SinCosTab{0x1.45F306DC9C883p+23,
                     0x1.921FB54442D18p0,
                     0x1p0,
                     -0x1.ffffffd0c621cp-2,
                     0x1.55553e1068f19p-5,
                     -0x1.6c087e89a359dp-10,
                     0x1.99343027bf8c3p-16,
                     -0x1.555545995a603p-3,
                     0x1.1107605230bc4p-7,
                     -0x1.994eb3774cf24p-13};
}

PRX_HD std::uint32_t InvPio4(int i) {
#ifdef __HIP_DEVICE_COMPILE__
    __constant__ static const std::uint32_t a[24] = {
#else
    static const std::uint32_t a[24] = {
#endif
        0xa2,       0xa2f9,     0xa2f983,   0xa2f9836e, 0xf9836e4e, 0x836e4e44,
        0x6e4e4415, 0x4e441529, 0x441529fc, 0x1529fc27, 0x29fc2757, 0xfc2757d1,
        0x2757d1f5, 0x57d1f534, 0xd1f534dd, 0xf534ddc0, 0x34ddc0db, 0xddc0db62,
        0xc0db6295, 0xdb629599, 0x6295993c, 0x95993c43, 0x993c4390, 0x3c439041};
    return a[i];
}

PRX_HD std::uint32_t AsU32(float f) {
#ifdef __HIP_DEVICE_COMPILE__
    return __float_as_uint(f);
#else
    std::uint32_t u;
    std::memcpy(&u, &f, 4);
    return u;
#endif
}

PRX_HD std::uint32_t AbsTop12(float x) { return (AsU32(x) >> 20) & 0x7ff; }

template <bool Fma>
PRX_HD double MulAdd(double a, double b, double c) {
    if constexpr (Fma) {
#ifdef __HIP_DEVICE_COMPILE__
        return __builtin_fma(a, b, c);
#else
        return __builtin_fma(a, b, c);
#endif
    } else {
        const double p = a * b;
        return c + p;
    }
}

// Polynomial for sin (n even) or cos (n odd) on the reduced argument.  The
// second table of the published method holds the cosine coefficients
// negated (and the same sine coefficients), so evaluating with it equals
// negating the cosine polynomial of the first: every step is a product or a
// sum of sign-symmetric roundings, and the cosine polynomial on the reduced
// range stays near cos(r) >= 0.7, never 0.  `negCos` selects that table.
// Selecting by a flag instead of a table pointer keeps every coefficient a
// compile-time constant: no per-lane table load inside the solver chains.
template <bool Fma>
PRX_HD float Poly(double x, double x2, const SinCosTab& p, int n, bool negCos) {
    if ((n & 1) == 0) {
        const double x3 = x * x2;
        const double s1 = MulAdd<Fma>(x2, p.s3, p.s2);
        const double x7 = x3 * x2;
        const double s = MulAdd<Fma>(x3, p.s1, x);
        return float(MulAdd<Fma>(x7, s1, s));
    }
    const double x4 = x2 * x2;
    const double c2 = MulAdd<Fma>(x2, p.c4, p.c3);
    const double c1 = MulAdd<Fma>(x2, p.c1, p.c0);
    const double x6 = x4 * x2;
    const double c = MulAdd<Fma>(x4, p.c2, c1);
    const float r = float(MulAdd<Fma>(x6, c2, c));
    return negCos ? -r : r;
}

// The published sign table {1, -1, -1, 1}[i & 3].
PRX_HD double QuadrantSign(int i) { return ((i + 1) & 2) ? -1.0 : 1.0; }

template <bool Fma>
PRX_HD double ReduceFast(double x, const SinCosTab& p, int* np) {
    const double r = x * p.hpi_inv;
    const int n = (int(std::int32_t(r)) + 0x800000) >> 24;
    *np = n;
    return MulAdd<Fma>(-double(n), p.hpi, x);
}

PRX_HD double ReduceLarge(std::uint32_t xi, int* np) {
    const int base = int((xi >> 26) & 15);
    const int shift = int((xi >> 23) & 7);
    xi = (xi & 0xffffff) | 0x800000;
    xi <<= shift;
    std::uint64_t res0 = std::uint64_t(std::uint32_t(xi * InvPio4(base)));
    const std::uint64_t res1 = std::uint64_t(xi) * InvPio4(base + 4);
    const std::uint64_t res2 = std::uint64_t(xi) * InvPio4(base + 8);
    res0 = (res2 >> 32) | (res0 << 32);
    res0 += res1;
    const std::uint64_t n = (res0 + (1ULL << 61)) >> 62;
    res0 -= n << 62;
    const double x = double(std::int64_t(res0));
    *np = int(n);
    return x * 0x1.921FB54442D18p-62;
}

template <bool Fma>
PRX_HD float SinF(float y) {
    double x = y;
    constexpr SinCosTab p = Tab();
    int n;
    const std::uint32_t t = AbsTop12(y);
    if (t < AbsTop12(0x1.921FB6p-1f)) {
        const double s = x * x;
        if (t < AbsTop12(0x1p-12f)) return y;
        return Poly<Fma>(x, s, p, 0, false);
    }
    if (t < AbsTop12(120.0f)) {
        x = ReduceFast<Fma>(x, p, &n);
        const double s = QuadrantSign(n);
        return Poly<Fma>(x * s, x * x, p, n, (n & 2) != 0);
    }
    const std::uint32_t xi = AsU32(y);
    const int sign = int(xi >> 31);
    x = ReduceLarge(xi, &n);
    const double s = QuadrantSign(n + sign);
    return Poly<Fma>(x * s, x * x, p, n, ((n + sign) & 2) != 0);
}

template <bool Fma>
PRX_HD float CosF(float y) {
    double x = y;
    constexpr SinCosTab p = Tab();
    int n;
    const std::uint32_t t = AbsTop12(y);
    if (t < AbsTop12(0x1.921FB6p-1f)) {
        const double x2 = x * x;
        if (t < AbsTop12(0x1p-12f)) return 1.0f;
        return Poly<Fma>(x, x2, p, 1, false);
    }
    if (t < AbsTop12(120.0f)) {
        x = ReduceFast<Fma>(x, p, &n);
        const double s = QuadrantSign(n);
        return Poly<Fma>(x * s, x * x, p, n ^ 1, (n & 2) != 0);
    }
    const std::uint32_t xi = AsU32(y);
    const int sign = int(xi >> 31);
    x = ReduceLarge(xi, &n);
    const double s = QuadrantSign(n + sign);
    return Poly<Fma>(x * s, x * x, p, n ^ 1, ((n + sign) & 2) != 0);
}

// Both values with one argument reduction; each equals SinF / CosF above
// (same reduction, same table, polynomials of opposite parity).  Below pi/4
// the fast reduction gives quadrant 0 and leaves the argument unchanged
// (fma(-0, pi/2, x) = x), so the small-argument polynomials are the fast
// path's and one path serves every |y| < 120 (the tiny-argument results are
// selected at the end): lanes of a wave with angles on either side of pi/4
// run the same instructions.
template <bool Fma>
PRX_HD void SinCosF(float y, float* sinOut, float* cosOut) {
    double x = y;
    constexpr SinCosTab p = Tab();
    int n;
    const std::uint32_t t = AbsTop12(y);
    double s;
    bool neg;
    if (t < AbsTop12(120.0f)) {
        x = ReduceFast<Fma>(x, p, &n);
        s = QuadrantSign(n);
        neg = (n & 2) != 0;
    } else {
        const std::uint32_t xi = AsU32(y);
        const int sign = int(xi >> 31);
        x = ReduceLarge(xi, &n);
        s = QuadrantSign(n + sign);
        neg = ((n + sign) & 2) != 0;
    }
    const double xs = x * s, x2 = x * x;
    const bool tiny = t < AbsTop12(0x1p-12f);
    const float sv = Poly<Fma>(xs, x2, p, n, neg);
    const float cv = Poly<Fma>(xs, x2, p, n ^ 1, neg);
    *sinOut = tiny ? y : sv;
    *cosOut = tiny ? 1.0f : cv;
}

}  // namespace prx_math
