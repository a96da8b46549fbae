// ORACLE — test infrastructure only.  Dual numbers for forward-mode
// differentiation of the restatement.  In Dual<float> (Tf) `v` is the float
// value and every operation computes it exactly as the float instantiation
// does (same operation, same order, float rounding), so a World<Tf> takes the
// float world's discrete path bit for bit; `d` carries the derivative of `v`
// along that path in double precision.  The result is the derivative of the
// float step as executed, branch by branch, at any horizon: the quantity a
// reverse (adjoint) pass of the same step computes.  Dual<double> (Td) is the
// same over the double instantiation; its derivatives are those of the FP64
// step, which central differences of the FP64 step approach wherever no event
// lies inside the difference, so Td checks the differentiation rules apart
// from where float and double paths part.
//
// Derivatives of the library calls: sqrt, hypot, sin, cos, atan2 from their
// closed forms evaluated in double at the float argument; abs by the sign of
// the value; trunc has derivative 0 and nextafter passes the derivative of its
// first argument through (a one-ulp move); comparisons, min, max and clamps
// look at the values only, so the derivative follows the branch taken.
#pragma once

#include <cmath>
#include <limits>
#include <type_traits>

namespace oracle {

template <class V, class G = double>
struct Dual {
    V v = V(0);
    G d = G(0);

    constexpr Dual() = default;
    template <class T, std::enable_if_t<std::is_arithmetic_v<T>, int> = 0>
    constexpr Dual(T x) : v(V(x)), d(G(0)) {}
    constexpr Dual(V x, G dx, int) : v(x), d(dx) {}
    template <class T, std::enable_if_t<std::is_arithmetic_v<T>, int> = 0>
    explicit constexpr operator T() const {
        return T(v);
    }

    Dual& operator+=(Dual b) { return *this = Dual(v + b.v, d + b.d, 0); }
    Dual& operator-=(Dual b) { return *this = Dual(v - b.v, d - b.d, 0); }
    Dual& operator*=(Dual b) { return *this = Dual(v * b.v, d * G(b.v) + G(v) * b.d, 0); }
    Dual& operator/=(Dual b) {
        const G bv = G(b.v);
        return *this = Dual(v / b.v, (d * bv - G(v) * b.d) / (bv * bv), 0);
    }

    // Friends, so that an arithmetic operand on either side converts.
    friend Dual operator+(Dual a, Dual b) { return a += b; }
    friend Dual operator-(Dual a, Dual b) { return a -= b; }
    friend Dual operator*(Dual a, Dual b) { return a *= b; }
    friend Dual operator/(Dual a, Dual b) { return a /= b; }
    friend Dual operator-(Dual a) { return Dual(-a.v, -a.d, 0); }
    friend Dual operator+(Dual a) { return a; }
    friend bool operator<(Dual a, Dual b) { return a.v < b.v; }
    friend bool operator>(Dual a, Dual b) { return a.v > b.v; }
    friend bool operator<=(Dual a, Dual b) { return a.v <= b.v; }
    friend bool operator>=(Dual a, Dual b) { return a.v >= b.v; }
    friend bool operator==(Dual a, Dual b) { return a.v == b.v; }
    friend bool operator!=(Dual a, Dual b) { return a.v != b.v; }
};

using Tf = Dual<float>;
using Td = Dual<double>;
using Tff = Dual<float, float>;  // derivatives in float too: their sensitivity to rounding

}  // namespace oracle

namespace std {

template <class V, class G>
struct numeric_limits<oracle::Dual<V, G>> : numeric_limits<V> {
    using D = oracle::Dual<V, G>;
    static constexpr D min() noexcept { return D(numeric_limits<V>::min()); }
    static constexpr D max() noexcept { return D(numeric_limits<V>::max()); }
    static constexpr D lowest() noexcept { return D(numeric_limits<V>::lowest()); }
    static constexpr D epsilon() noexcept { return D(numeric_limits<V>::epsilon()); }
    static constexpr D infinity() noexcept { return D(numeric_limits<V>::infinity()); }
    static constexpr D quiet_NaN() noexcept { return D(numeric_limits<V>::quiet_NaN()); }
    static constexpr D denorm_min() noexcept { return D(numeric_limits<V>::denorm_min()); }
};

template <class V, class G> inline oracle::Dual<V, G> abs(oracle::Dual<V, G> a) {
    return oracle::Dual<V, G>(std::abs(a.v), std::signbit(a.v) ? -a.d : a.d, 0);
}
template <class V, class G> inline oracle::Dual<V, G> sqrt(oracle::Dual<V, G> a) {
    const V r = std::sqrt(a.v);
    return oracle::Dual<V, G>(r, a.v > V(0) ? a.d / (G(2) * std::sqrt(G(a.v))) : G(0), 0);
}
template <class V, class G> inline oracle::Dual<V, G> hypot(oracle::Dual<V, G> x, oracle::Dual<V, G> y) {
    const V h = std::hypot(x.v, y.v);
    const G hd = std::hypot(G(x.v), G(y.v));
    return oracle::Dual<V, G>(h, hd > G(0) ? (G(x.v) * x.d + G(y.v) * y.d) / hd : G(0), 0);
}
template <class V, class G> inline oracle::Dual<V, G> sin(oracle::Dual<V, G> a) {
    return oracle::Dual<V, G>(std::sin(a.v), std::cos(G(a.v)) * a.d, 0);
}
template <class V, class G> inline oracle::Dual<V, G> cos(oracle::Dual<V, G> a) {
    return oracle::Dual<V, G>(std::cos(a.v), -std::sin(G(a.v)) * a.d, 0);
}
template <class V, class G> inline oracle::Dual<V, G> atan2(oracle::Dual<V, G> y, oracle::Dual<V, G> x) {
    const G xv = G(x.v), yv = G(y.v), r2 = xv * xv + yv * yv;
    return oracle::Dual<V, G>(std::atan2(y.v, x.v), r2 > G(0) ? (xv * y.d - yv * x.d) / r2 : G(0), 0);
}
template <class V, class G> inline oracle::Dual<V, G> trunc(oracle::Dual<V, G> a) { return oracle::Dual<V, G>(std::trunc(a.v), G(0), 0); }
template <class V, class G> inline oracle::Dual<V, G> nextafter(oracle::Dual<V, G> a, oracle::Dual<V, G> b) {
    return oracle::Dual<V, G>(std::nextafter(a.v, b.v), a.d, 0);
}
template <class V, class G> inline bool isnormal(oracle::Dual<V, G> a) { return std::isnormal(a.v); }
template <class V, class G> inline bool isfinite(oracle::Dual<V, G> a) { return std::isfinite(a.v); }
template <class V, class G> inline bool signbit(oracle::Dual<V, G> a) { return std::signbit(a.v); }

}  // namespace std
