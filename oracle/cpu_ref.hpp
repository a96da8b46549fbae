// ORACLE — test infrastructure only.  A CPU restatement of PlayRho's per-step
// hot path, `Step(AabbTreeWorld&, const StepConf&)` with doToi == false, used as
// the checker for the HIP implementation in playrho_amd/.  Nothing in the
// product links, loads or calls this code; tests/, __graft_entry__.smoke() and
// bench.py's cpu_baseline leg use it as the checker / CPU baseline only.
//
// Templated on the real type: the float instantiation is pinned bit-for-bit to
// the reference float build (oracle/_ref, tests/golden); the double
// instantiation gives the FP64 comparison figure.
//
// Every function cites the reference file:line it restates (paths relative to
// the PlayRho checkout).  Operation order inside each arithmetic expression
// follows the reference so that float results are bit-identical; compile with
// -ffp-contract=off and without -ffast-math.
//
// Pinning (what checks each part):
//  - whole step, float: per-step digests of full state vs the reference build
//    for the hello / pyramid / drop / pile / billiards goldens (tests/golden);
//  - ContactFeature / manifold: the same digests include every manifold;
//  - double instantiation: closed-form free fall (v = a t, p += v h), and
//    agreement with the float instantiation to float precision on short runs.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <tuple>
#include <vector>

namespace oracle {

template <class R>
struct V2 {
    R x = 0, y = 0;
};
template <class R> inline V2<R> operator+(V2<R> a, V2<R> b) { return {a.x + b.x, a.y + b.y}; }
template <class R> inline V2<R> operator-(V2<R> a, V2<R> b) { return {a.x - b.x, a.y - b.y}; }
template <class R> inline V2<R> operator-(V2<R> a) { return {-a.x, -a.y}; }
// Vector * scalar, element-wise (Vector.hpp:496-507).
template <class R> inline V2<R> operator*(V2<R> a, R s) { return {a.x * s, a.y * s}; }
// Scalar * vector (Vector.hpp:479-490).
template <class R> inline V2<R> operator*(R s, V2<R> a) { return {s * a.x, s * a.y}; }
// Vector / scalar multiplies by the reciprocal (Vector.hpp:511-521).
template <class R> inline V2<R> operator/(V2<R> a, R s) {
    const R inv = R(1) / s;
    return {a.x * inv, a.y * inv};
}
template <class R> inline bool operator==(V2<R> a, V2<R> b) { return a.x == b.x && a.y == b.y; }

// Dot starts its sum from zero (Math.hpp:320-334).
template <class R> inline R Dot(V2<R> a, V2<R> b) {
    R r = 0;
    r += a.x * b.x;
    r += a.y * b.y;
    return r;
}
// 2-D cross (Math.hpp:365-386).
template <class R> inline R Cross(V2<R> a, V2<R> b) {
    const R m = a.x * b.y;
    const R s = a.y * b.x;
    return m - s;
}
// Math.hpp:274-283.
template <class R> inline R MagSq(V2<R> v) {
    R r = 0;
    r += v.x * v.x;
    r += v.y * v.y;
    return r;
}
template <class R> inline V2<R> RevPerp(V2<R> v) { return {-v.y, v.x}; }  // UnitVec.hpp:364-368
template <class R> inline V2<R> FwdPerp(V2<R> v) { return {v.y, -v.x}; }  // UnitVec.hpp:374-378

// Rotate / InverseRotate (d2/Math.hpp:107-134); q = (cos, sin).
template <class R> inline V2<R> Rotate(V2<R> v, V2<R> q) {
    const R nx = (q.x * v.x) - (q.y * v.y);
    const R ny = (q.y * v.x) + (q.x * v.y);
    return {nx, ny};
}
template <class R> inline V2<R> InvRotate(V2<R> v, V2<R> q) {
    const R nx = (q.x * v.x) + (q.y * v.y);
    const R ny = (q.x * v.y) - (q.y * v.x);
    return {nx, ny};
}
// UnitVec::Rotate (UnitVec.hpp:353-358).
template <class R> inline V2<R> RotateUnit(V2<R> a, V2<R> amount) {
    return {a.x * amount.x - a.y * amount.y, a.y * amount.x + a.x * amount.y};
}

template <class R>
struct Xf {
    V2<R> p;
    V2<R> q;  // cos, sin
};
// d2/Math.hpp:194-212
template <class R> inline V2<R> Transform(V2<R> v, const Xf<R>& xf) { return Rotate(v, xf.q) + xf.p; }
template <class R> inline V2<R> InvTransform(V2<R> v, const Xf<R>& xf) { return InvRotate(v - xf.p, xf.q); }
// MulT (d2/Math.hpp:224-229).
template <class R> inline Xf<R> MulT(const Xf<R>& A, const Xf<R>& B) {
    const V2<R> dp = B.p - A.p;
    return Xf<R>{InvRotate(dp, A.q), RotateUnit(B.q, V2<R>{A.q.x, -A.q.y})};
}

// UnitVec::Get(angle) = {cos, sin} (UnitVec.cpp:27-30).
template <class R> inline V2<R> UnitFromAngle(R a) { return {std::cos(a), std::sin(a)}; }

template <class R>
struct Pos {
    V2<R> lin;
    R ang = 0;
};
template <class R> inline Pos<R> operator+(Pos<R> a, Pos<R> b) { return {a.lin + b.lin, a.ang + b.ang}; }
template <class R> inline Pos<R> operator-(Pos<R> a) { return {-a.lin, -a.ang}; }
template <class R> inline bool operator!=(Pos<R> a, Pos<R> b) { return !(a.lin == b.lin) || a.ang != b.ang; }

template <class R>
struct Vel {
    V2<R> lin;
    R ang = 0;
};

// GetTransformation(pos, localCenter) (d2/Math.hpp:232-245).
template <class R> inline Xf<R> GetXf(const Pos<R>& pos, V2<R> lc) {
    const V2<R> q = UnitFromAngle(pos.ang);
    return Xf<R>{pos.lin - Rotate(lc, q), q};
}

// UnitVec::Get(x, y, fallback) (UnitVec.hpp:175-236).
template <class R> inline V2<R> UnitVector(V2<R> v, V2<R> fallback) {
    const R m2 = v.x * v.x + v.y * v.y;
    if (std::isnormal(m2)) {
        const R mag = std::sqrt(m2);
        const R inv = R(1) / mag;
        return {v.x * inv, v.y * inv};
    }
    const R mag = std::hypot(v.x, v.y);
    if (!std::isnormal(mag)) return fallback;
    return {v.x / mag, v.y / mag};
}

// UnitVec::Get(x, y) with the default (zero) fallback, returning the
// magnitude as well (UnitVec.hpp:175-236): zero magnitude when no direction.
template <class R> inline V2<R> UnitPolar(V2<R> v, R& mag) {
    const R m2 = v.x * v.x + v.y * v.y;
    if (std::isnormal(m2)) {
        mag = std::sqrt(m2);
        const R inv = R(1) / mag;
        return {v.x * inv, v.y * inv};
    }
    const R h = std::hypot(v.x, v.y);
    if (!std::isnormal(h)) {
        mag = R(0);
        return {R(0), R(0)};
    }
    mag = h;
    return {v.x / h, v.y / h};
}

// std::max / std::min / std::clamp comparison forms.
template <class R> inline R Max(R a, R b) { return (a < b) ? b : a; }
template <class R> inline R Min(R a, R b) { return (b < a) ? b : a; }
template <class R> inline R Clamp(R v, R lo, R hi) { return (v < lo) ? lo : ((hi < v) ? hi : v); }

// AlmostEqual with 4 ULP via nextafter (Math.hpp:186-194).
template <class R> inline bool AlmostEqual(R a, R b, int ulp = 4) {
    for (; (a != b) && (ulp > 0); --ulp) a = std::nextafter(a, b);
    return a == b;
}
template <class R> inline bool AlmostZero(R v) { return std::abs(v) < std::numeric_limits<R>::min(); }

// ---------------------------------------------------------------------------
// ContactFeature (ContactFeature.hpp:30-85), packed as
// typeA | typeB << 1 | indexA << 8 | indexB << 16.
using Cf = std::uint32_t;
constexpr std::uint32_t kVertex = 0, kFace = 1, kInvalidVertex = 255;
inline Cf MakeCf(std::uint32_t tA, std::uint32_t iA, std::uint32_t tB, std::uint32_t iB) {
    return tA | (tB << 1) | ((iA & 0xffu) << 8) | ((iB & 0xffu) << 16);
}
inline Cf FlipCf(Cf c) {
    return MakeCf((c >> 1) & 1u, (c >> 16) & 0xffu, c & 1u, (c >> 8) & 0xffu);
}
inline std::uint32_t CfIndexB(Cf c) { return (c >> 16) & 0xffu; }
constexpr Cf kDefaultCf = (kInvalidVertex << 8) | (kInvalidVertex << 16);

enum ManifoldType : std::uint32_t { kUnset = 0, kCircles = 1, kFaceA = 2, kFaceB = 3 };

template <class R>
struct MPoint {
    V2<R> lp;
    Cf cf = kDefaultCf;
    R ni = 0, ti = 0;
};

template <class R>
struct Manifold {
    std::uint32_t type = kUnset;
    std::uint32_t n = 0;
    V2<R> ln;  // local normal
    V2<R> lpt; // local point
    MPoint<R> pts[2];
    void Add(const MPoint<R>& p) { pts[n++] = p; }
};
// Manifold factories (d2/Manifold.hpp:168-243).
template <class R> inline Manifold<R> ForCircles(V2<R> vA, std::uint32_t iA, V2<R> vB, std::uint32_t iB) {
    Manifold<R> m;
    m.type = kCircles;
    m.lpt = vA;
    m.Add(MPoint<R>{vB, MakeCf(kVertex, iA, kVertex, iB)});
    return m;
}
template <class R> inline Manifold<R> ForFace(bool faceB, V2<R> ln, V2<R> lp) {
    Manifold<R> m;
    m.type = faceB ? kFaceB : kFaceA;
    m.ln = ln;
    m.lpt = lp;
    return m;
}

// DistanceProxy: vertices, forward normals, vertex radius.
template <class R>
struct Proxy {
    const V2<R>* v = nullptr;
    const V2<R>* nrm = nullptr;
    std::uint32_t n = 0;
    R radius = 0;
};

inline std::uint32_t ModNext(std::uint32_t i, std::uint32_t n) { return (i + 1 < n) ? i + 1 : 0; }
inline std::uint32_t ModPrev(std::uint32_t i, std::uint32_t n) { return i ? i - 1 : n - 1; }

template <class R>
struct ClipV {
    V2<R> v;
    Cf cf;
};
template <class R>
struct ClipList {
    ClipV<R> e[2];
    std::uint32_t n = 0;
    void push(const ClipV<R>& c) { e[n++] = c; }
};

// Manifold.cpp:61-102
template <class R>
ClipList<R> ClipSegmentToLine(const ClipList<R>& vIn, V2<R> normal, R offset, std::uint32_t indexA) {
    ClipList<R> out;
    if (vIn.n == 2) {
        const R d0 = Dot(normal, vIn.e[0].v) - offset;
        const R d1 = Dot(normal, vIn.e[1].v) - offset;
        if (d0 <= R(0) || AlmostZero(d0)) out.push(vIn.e[0]);
        if (d1 <= R(0) || AlmostZero(d1)) out.push(vIn.e[1]);
        if (out.n < 2 && std::signbit(d0) != std::signbit(d1)) {
            const R interp = d0 / (d0 - d1);
            const V2<R> vertex = vIn.e[0].v + (vIn.e[1].v - vIn.e[0].v) * interp;
            out.push(ClipV<R>{vertex, MakeCf(kVertex, indexA, kFace, CfIndexB(vIn.e[0].cf))});
        }
    }
    return out;
}

// Manifold.cpp:122-141
template <class R>
inline void MostAntiParallelEdge(V2<R> n0, const Xf<R>& xf0, const Proxy<R>& s1, const Xf<R>& xf1,
                                 std::uint32_t first, std::uint32_t second,
                                 std::uint32_t& e0, std::uint32_t& e1) {
    if (second == kInvalidVertex) {
        const V2<R> normal = InvRotate(Rotate(n0, xf0.q), xf1.q);
        const std::uint32_t prev = ModPrev(first, s1.n);
        const R prevDot = Dot(normal, s1.nrm[prev]);
        const R currDot = Dot(normal, s1.nrm[first]);
        if (prevDot < currDot) {
            e0 = prev;
            e1 = first;
        } else {
            e0 = first;
            e1 = ModNext(first, s1.n);
        }
        return;
    }
    if ((second > first) && ((first + 1) == second)) {
        e0 = first;
        e1 = second;
    } else {
        e0 = second;
        e1 = first;
    }
}

// Polygon-polygon manifold (Manifold.cpp:168-291).
template <class R>
Manifold<R> PolyManifold(bool flipped, const Proxy<R>& s0, const Xf<R>& xf0, std::uint32_t idx0,
                         const Proxy<R>& s1, const Xf<R>& xf1, std::uint32_t i1a, std::uint32_t i1b,
                         R maxCirclesRatio) {
    const R r0 = s0.radius;
    const R totalRadius = r0 + s1.radius;
    const std::uint32_t idx0Next = ModNext(idx0, s0.n);
    const V2<R> rv0 = s0.v[idx0];
    const V2<R> rv1 = s0.v[idx0Next];
    const V2<R> av0 = Transform(rv0, xf0);
    const V2<R> av1 = Transform(rv1, xf0);
    const R lenEdge0 = MagSq(rv1 - rv0);
    const V2<R> rn0 = s0.nrm[idx0];
    const V2<R> re0dir = RevPerp(rn0);
    std::uint32_t e1first, e1second;
    MostAntiParallelEdge(rn0, xf0, s1, xf1, i1a, i1b, e1first, e1second);
    const V2<R> s1rv0 = s1.v[e1first];
    const V2<R> s1av0 = Transform(s1rv0, xf1);
    const V2<R> s1rv1 = s1.v[e1second];
    const V2<R> s1av1 = Transform(s1rv1, xf1);
    {
        const V2<R> ae0dir = Rotate(re0dir, xf0.q);
        // GetClipPoints (Manifold.cpp:143-156)
        ClipList<R> ie;
        ie.push(ClipV<R>{s1av0, MakeCf(kFace, idx0, kVertex, e1first)});
        ie.push(ClipV<R>{s1av1, MakeCf(kFace, idx0, kVertex, e1second)});
        const R dp0 = -Dot(ae0dir, av0);
        const R dp1 = +Dot(ae0dir, av1);
        const ClipList<R> pts = ClipSegmentToLine(ie, -ae0dir, dp0, idx0);
        const ClipList<R> clip = ClipSegmentToLine(pts, ae0dir, dp1, idx0Next);
        if (clip.n == 2) {
            const V2<R> absNormal = FwdPerp(ae0dir);
            const V2<R> relMid = (rv0 + rv1) / R(2);
            const R absOffset = Dot(absNormal, av0);
            const V2<R> normal = FwdPerp(re0dir);
            Manifold<R> m = ForFace(flipped, normal, relMid);
            for (std::uint32_t k = 0; k < clip.n; ++k) {
                const auto& cp = clip.e[k];
                if ((Dot(absNormal, cp.v) - absOffset) <= totalRadius) {
                    m.Add(MPoint<R>{InvTransform(cp.v, xf1), flipped ? FlipCf(cp.cf) : cp.cf});
                }
            }
            if (m.n > 0) return m;
        }
    }
    const R trSq = totalRadius * totalRadius;
    const R cr = maxCirclesRatio * r0;
    const bool mustFace = lenEdge0 > cr * cr;
    auto vertexCase = [&](V2<R> v0rel, std::uint32_t i0, V2<R> v1rel, std::uint32_t i1) {
        if (mustFace) {
            Manifold<R> m = ForFace(flipped, FwdPerp(re0dir), v0rel);
            m.Add(MPoint<R>{v1rel, flipped ? MakeCf(kVertex, i1, kFace, i0) : MakeCf(kFace, i0, kVertex, i1)});
            return m;
        }
        return !flipped ? ForCircles(v0rel, i0, v1rel, i1) : ForCircles(v1rel, i1, v0rel, i0);
    };
    if (MagSq(av0 - s1av0) <= trSq) return vertexCase(rv0, idx0, s1rv0, e1first);
    if (MagSq(av1 - s1av1) <= trSq) return vertexCase(rv1, idx0Next, s1rv1, e1second);
    if (MagSq(av0 - s1av1) <= trSq) return vertexCase(rv0, idx0, s1rv1, e1second);
    if (MagSq(av1 - s1av0) <= trSq) return vertexCase(rv1, idx0Next, s1rv0, e1first);
    return Manifold<R>{};
}

// Polygon-circle manifold (Manifold.cpp:293-372).
template <class R>
Manifold<R> PolyCircleManifold(bool flipped, R totalRadius, const Proxy<R>& s, const Xf<R>& sxf,
                               V2<R> point, const Xf<R>& xf) {
    const std::uint32_t n = s.n;
    const V2<R> cLocal = InvTransform(Transform(point, xf), sxf);
    std::uint32_t indexOfMax = 0;
    R maxSep = -std::numeric_limits<R>::max();  // -MaxFloat (Settings.hpp:64)
    for (std::uint32_t i = 0; i < n; ++i) {
        const R sep = Dot(s.nrm[i], cLocal - s.v[i]);
        if (sep > totalRadius) return Manifold<R>{};
        if (maxSep < sep) {
            maxSep = sep;
            indexOfMax = i;
        }
    }
    const std::uint32_t indexOfMax2 = ModNext(indexOfMax, n);
    const V2<R> v1 = s.v[indexOfMax];
    const V2<R> v2 = s.v[indexOfMax2];
    auto face = [&](V2<R> fc) {
        Manifold<R> m = ForFace(flipped, s.nrm[indexOfMax], fc);
        m.Add(MPoint<R>{point, flipped ? MakeCf(kVertex, 0, kFace, indexOfMax) : MakeCf(kFace, indexOfMax, kVertex, 0)});
        return m;
    };
    if (maxSep < R(0)) return face((v1 + v2) / R(2));
    const V2<R> cv1 = cLocal - v1;
    if (Dot(cv1, v2 - v1) <= R(0)) {
        if (MagSq(cv1) > totalRadius * totalRadius) return Manifold<R>{};
        return !flipped ? ForCircles(v1, indexOfMax, point, 0u) : ForCircles(point, 0u, v1, indexOfMax);
    }
    const V2<R> cv2 = cLocal - v2;
    if (Dot(cv2, v1 - v2) <= R(0)) {
        if (MagSq(cv2) > totalRadius * totalRadius) return Manifold<R>{};
        return !flipped ? ForCircles(v2, indexOfMax2, point, 0u) : ForCircles(point, 0u, v2, indexOfMax2);
    }
    const V2<R> fc = (v1 + v2) / R(2);
    if (Dot(cLocal - fc, s.nrm[indexOfMax]) > totalRadius) return Manifold<R>{};
    return face(fc);
}

// GetMinSeparationInfo (ShapeSeparation.cpp:39-64)
template <class R>
inline void MinSepInfo(V2<R> origin, V2<R> dir, const V2<R>* verts, std::uint32_t n,
                       R& minSep, std::uint32_t& first, std::uint32_t& second) {
    minSep = std::numeric_limits<R>::infinity();
    first = kInvalidVertex;
    second = kInvalidVertex;
    for (std::uint32_t i = 0; i < n; ++i) {
        const R s = Dot(dir, verts[i] - origin);
        if (minSep > s) {
            minSep = s;
            first = i;
            second = kInvalidVertex;
        } else if (minSep == s) {
            second = i;
        }
    }
}

template <class R>
struct SepInfo {
    R distance;
    std::uint32_t first, second0, second1;
};

// GetMaxSeparation4x4 (ShapeSeparation.cpp:66-104)
template <class R>
SepInfo<R> MaxSep4x4(const Proxy<R>& p1, const Xf<R>& xf1, const Proxy<R>& p2, const Xf<R>& xf2) {
    SepInfo<R> out{-std::numeric_limits<R>::infinity(), kInvalidVertex, kInvalidVertex, kInvalidVertex};
    const Xf<R> xf = MulT(xf1, xf2);
    V2<R> verts[4];
    for (int k = 0; k < 4; ++k) verts[k] = Transform(p2.v[k], xf);
    for (std::uint32_t i = 0; i < 4; ++i) {
        R d;
        std::uint32_t f, s;
        MinSepInfo(p1.v[i], p1.nrm[i], verts, 4u, d, f, s);
        if (out.distance < d) {
            out.distance = d;
            out.second0 = f;
            out.second1 = s;
            out.first = i;
        }
    }
    return out;
}

// GetMaxSeparation (ShapeSeparation.cpp:106-128)
template <class R>
SepInfo<R> MaxSep(const Proxy<R>& p1, const Xf<R>& xf1, const Proxy<R>& p2, const Xf<R>& xf2) {
    SepInfo<R> out{-std::numeric_limits<R>::infinity(), kInvalidVertex, kInvalidVertex, kInvalidVertex};
    const Xf<R> xf = MulT(xf2, xf1);
    for (std::uint32_t i = 0; i < p1.n; ++i) {
        const V2<R> origin = Transform(p1.v[i], xf);
        const V2<R> normal = Rotate(p1.nrm[i], xf.q);
        R d;
        std::uint32_t f, s;
        MinSepInfo(origin, normal, p2.v, p2.n, d, f, s);
        if (out.distance < d) {
            out.distance = d;
            out.second0 = f;
            out.second1 = s;
            out.first = i;
        }
    }
    return out;
}

// CollideShapes (Manifold.cpp:393-438)
template <class R>
Manifold<R> CollideShapes(const Proxy<R>& A, const Xf<R>& xfA, const Proxy<R>& B, const Xf<R>& xfB,
                          R linearSlop, R maxCirclesRatio) {
    const R totalRadius = A.radius + B.radius;
    if (A.n == 1 && B.n == 1) {
        // Manifold.cpp:374-386
        const V2<R> pA = Transform(A.v[0], xfA);
        const V2<R> pB = Transform(B.v[0], xfB);
        const R lenSq = MagSq(pB - pA);
        const R totSq = totalRadius * totalRadius;
        if (lenSq > totSq) return Manifold<R>{};
        return ForCircles(A.v[0], 0u, B.v[0], 0u);
    }
    if (A.n == 1) return PolyCircleManifold(true, totalRadius, B, xfB, A.v[0], xfA);
    if (B.n == 1) return PolyCircleManifold(false, totalRadius, A, xfA, B.v[0], xfB);
    const bool do4x4 = (A.n == 4) && (B.n == 4);
    const SepInfo<R> sa = do4x4 ? MaxSep4x4(A, xfA, B, xfB) : MaxSep(A, xfA, B, xfB);
    if (sa.distance > totalRadius) return Manifold<R>{};
    const SepInfo<R> sb = do4x4 ? MaxSep4x4(B, xfB, A, xfA) : MaxSep(B, xfB, A, xfA);
    if (sb.distance > totalRadius) return Manifold<R>{};
    const R kTol = linearSlop / R(10);
    return (sb.distance > (sa.distance + kTol))
               ? PolyManifold(true, B, xfB, sb.first, A, xfA, sb.second0, sb.second1, maxCirclesRatio)
               : PolyManifold(false, A, xfA, sa.first, B, xfB, sa.second0, sa.second1, maxCirclesRatio);
}

// ---------------------------------------------------------------------------
// World manifold (WorldManifold.cpp:37-130)
template <class R>
struct WorldManifold {
    V2<R> normal;
    std::uint32_t n = 0;
    V2<R> p[2];
    R ni[2] = {0, 0}, ti[2] = {0, 0};
    R sep[2] = {0, 0};
};

template <class R>
WorldManifold<R> GetWorldManifold(const Manifold<R>& m, const Xf<R>& xfA, R rA, const Xf<R>& xfB, R rB) {
    WorldManifold<R> w;
    if (m.type == kCircles) {
        const V2<R> pA = Transform(m.lpt, xfA);
        const V2<R> pB = Transform(m.pts[0].lp, xfB);
        const V2<R> normal = UnitVector(pB - pA, V2<R>{R(1), R(0)});
        const V2<R> cA = pA + (rA * normal);
        const V2<R> cB = pB - (rB * normal);
        w.normal = normal;
        w.n = 1;
        w.p[0] = (cA + cB) / R(2);
        w.ni[0] = m.pts[0].ni;
        w.ti[0] = m.pts[0].ti;
        w.sep[0] = Dot(cB - cA, normal);
        return w;
    }
    if (m.type == kFaceA) {
        const V2<R> normal = Rotate(m.ln, xfA.q);
        const V2<R> plane = Transform(m.lpt, xfA);
        w.normal = normal;
        w.n = m.n;
        for (std::uint32_t i = 0; i < m.n; ++i) {
            const V2<R> clip = Transform(m.pts[i].lp, xfB);
            const V2<R> cA = clip + (rA - Dot(clip - plane, normal)) * normal;
            const V2<R> cB = clip - (rB * normal);
            w.p[i] = (cA + cB) / R(2);
            w.ni[i] = m.pts[i].ni;
            w.ti[i] = m.pts[i].ti;
            w.sep[i] = Dot(cB - cA, normal);
        }
        return w;
    }
    if (m.type == kFaceB) {
        const V2<R> normal = Rotate(m.ln, xfB.q);
        const V2<R> plane = Transform(m.lpt, xfB);
        w.normal = -normal;
        w.n = m.n;
        for (std::uint32_t i = 0; i < m.n; ++i) {
            const V2<R> clip = Transform(m.pts[i].lp, xfA);
            const V2<R> cB = clip + (rB - Dot(clip - plane, normal)) * normal;
            const V2<R> cA = clip - (rA * normal);
            w.p[i] = (cA + cB) / R(2);
            w.ni[i] = m.pts[i].ni;
            w.ti[i] = m.pts[i].ti;
            w.sep[i] = Dot(cA - cB, normal);
        }
        return w;
    }
    return w;
}

// Position solver manifold (PositionSolverManifold.cpp:42-121)
template <class R>
struct PSM {
    V2<R> normal, point;
    R sep;
};
template <class R>
PSM<R> GetPSM(const Manifold<R>& m, std::uint32_t idx, const Xf<R>& xfA, const Xf<R>& xfB) {
    if (m.type == kCircles) {
        const V2<R> pA = Transform(m.lpt, xfA);
        const V2<R> pB = Transform(m.pts[idx].lp, xfB);
        const V2<R> delta = pB - pA;
        const V2<R> normal = UnitVector(delta, V2<R>{R(0), R(0)});
        const V2<R> mid = (pA + pB) / R(2);
        return PSM<R>{normal, mid, Dot(delta, normal)};
    }
    if (m.type == kFaceA) {
        const V2<R> plane = Transform(m.lpt, xfA);
        const V2<R> normal = Rotate(m.ln, xfA.q);
        const V2<R> clip = Transform(m.pts[idx].lp, xfB);
        return PSM<R>{normal, clip, Dot(clip - plane, normal)};
    }
    if (m.type == kFaceB) {
        const V2<R> plane = Transform(m.lpt, xfB);
        const V2<R> normal = Rotate(m.ln, xfB.q);
        const V2<R> clip = Transform(m.pts[idx].lp, xfA);
        return PSM<R>{-normal, clip, Dot(clip - plane, normal)};
    }
    return PSM<R>{V2<R>{}, V2<R>{}, std::numeric_limits<R>::quiet_NaN()};
}

// ---------------------------------------------------------------------------
// Step configuration: the StepConf fields the regular step uses
// (StepConf.hpp:49-392).
template <class R>
struct StepConf {
    R dt, dtRatio, minStillTimeToSleep, linearSlop, angularSlop, regResolutionRate,
        regMinSeparation, regMinMomentum, velocityThreshold, maxTranslation, maxRotation,
        maxLinearCorrection, maxAngularCorrection, linearSleepTolerance, angularSleepTolerance,
        displaceMultiplier, aabbExtension, maxCirclesRatio;
    std::uint32_t velIters, posIters;
    bool warmStart, toi, blockSolve;
    // TOI phase (StepConf.hpp:84-158, 216-374).
    R toiResolutionRate = R(0.75f), toiMinSeparation = R(-0.005f * 1.5f), toiMinMomentum = 0,
      targetDepth = R(0.005f * 3.0f), tolerance = R(0.005f / 4.0f);
    std::uint32_t toiVelIters = 8, toiPosIters = 20, maxToiRootIters = 30, maxToiIters = 20, maxDistanceIters = 20,
                  maxSubSteps = 8;
};

// Body flags as in the dump format.
enum BodyFlag : std::uint32_t {
    kSpeedable = 1, kAccelerable = 2, kImpenetrable = 4, kAwake = 8, kAutoSleep = 16,
    kEnabled = 32, kFixedRotation = 64
};
// Contact flags as in the dump format.
enum ContactFlag : std::uint32_t {
    kTouching = 1, kCEnabled = 2, kSensor = 4, kCImpenetrable = 8, kNeedsUpdating = 16,
    kNeedsFiltering = 32
};

template <class R>
struct Shape {
    std::vector<V2<R>> v, nrm;
    R radius, density, friction, restitution;
    std::uint32_t sensor, catmask;
    std::int32_t group;
    Proxy<R> proxy() const { return Proxy<R>{v.data(), nrm.data(), std::uint32_t(v.size()), radius}; }
};

template <class R>
struct Body {
    std::uint32_t type = 0, flags = 0;
    Pos<R> pos0, pos1;
    V2<R> lc;
    Vel<R> vel;
    V2<R> lacc;
    R aacc = 0, ldamp = 0, adamp = 0, invMass = 0, invRotI = 0, underActive = 0;
    R alpha0 = 0;  // Sweep::alpha0 (Sweep.hpp:67)
    Xf<R> xf;  // cached transform of pos1 (Body.hpp:697-715)
    bool destroyed = false;  // Destroy(World&, BodyID) took it out of the world
    std::vector<std::uint32_t> shapes;
    std::vector<std::uint32_t> proxies;
    bool awake() const { return (flags & kAwake) != 0; }
    bool speedable() const { return (flags & kSpeedable) != 0; }
    bool accelerable() const { return (flags & kAccelerable) != 0; }
    // Body::SetAwake (Body.cpp:163-171)
    void SetAwake() {
        if (flags & kSpeedable) {
            flags |= kAwake;
            underActive = 0;
        }
    }
    void SetAwakeFlag() { flags |= kAwake; }
    // Unawaken -> UnsetAwake (Body.hpp:1203-1210, Body.cpp:173-182)
    bool Unawaken() {
        if ((flags & kAwake) && (flags & kAutoSleep)) {
            flags &= ~kAwake;
            underActive = 0;
            vel = Vel<R>{};
            return true;
        }
        return false;
    }
};

template <class R>
struct Leaf {
    bool used = false;
    std::uint32_t body = 0, shape = 0, child = 0;
    R lo[2], hi[2];
};

template <class R>
struct Contact {
    std::uint32_t kmin = 0, kmax = 0;
    std::uint32_t bodyA = 0, bodyB = 0, shapeA = 0, shapeB = 0;
    std::uint32_t flags = 0;
    R friction = 0, restitution = 0, tangentSpeed = 0;
    // TOI state (Contact.hpp:334-338, 451-480): value, e_toiFlag, sub-step count.
    R toi = 0;
    bool toiValid = false;
    std::uint32_t toiCount = 0;
};

// PreStepStats and RegStepStats (StepStats.hpp:38-46, 69-82); the two real
// fields are kept in double (exact for either instantiation).
struct StepStats {
    std::uint32_t preProxiesCreated = 0, preProxiesMoved = 0, preDestroyed = 0, preAdded = 0, preUpdated = 0,
                  preSkipped = 0;
    double minSeparation = std::numeric_limits<double>::infinity(), maxIncImpulse = 0;
    std::uint32_t islandsFound = 0, islandsSolved = 0, bodiesSlept = 0, maxIslandBodies = 0, regAdded = 0,
                  regUpdated = 0, regSkipped = 0, regProxiesMoved = 0, sumPosIters = 0, sumVelIters = 0;
    // ToiStepStats (StepStats.hpp:117-151).
    double toiMinSeparation = std::numeric_limits<double>::infinity(), toiMaxIncImpulse = 0;
    std::uint32_t toiIslandsFound = 0, toiIslandsSolved = 0, toiContactsFound = 0, toiAtMaxSubSteps = 0,
                  toiUpdatedToi = 0, toiUpdatedTouching = 0, toiSkippedTouching = 0, toiAdded = 0,
                  toiProxiesMoved = 0, toiSumPosIters = 0, toiSumVelIters = 0, toiMaxDistIters = 0,
                  toiMaxToiIters = 0, toiMaxRootIters = 0;
};

// IslandStats of one island solve (StepStats.hpp:149-163).
struct IslandStats {
    double minSeparation = std::numeric_limits<double>::infinity(), maxIncImpulse = 0;
    bool solved = false;
    std::uint32_t positionIters = 0, velocityIters = 0, bodiesSlept = 0;
};
// The contact counts ProcessContactsForTOI adds (AabbTreeWorld.cpp:2059-2142).
struct ToiIslandStats : IslandStats {
    std::uint32_t contactsUpdated = 0, contactsSkipped = 0;
};

// Solver scratch records.
template <class R>
struct BodyConstraint {
    R invMass = 0, invRotI = 0;
    V2<R> lc;
    Pos<R> pos;
    Vel<R> vel;
};

template <class R>
struct VcPoint {
    V2<R> relA, relB;
    R ni = 0, ti = 0, normalMass = 0, tangentMass = 0, velocityBias = 0;
};

template <class R>
struct VelocityConstraint {
    std::uint32_t bodyA = 0, bodyB = 0, n = 0;
    V2<R> normal;
    R friction = 0, restitution = 0, tangentSpeed = 0;
    VcPoint<R> p[2];
    R K[3] = {0, 0, 0};           // k00, k11, k01
    R normalMass[3] = {0, 0, 0};  // inverse of K (m00, m11, m01)
    V2<R> tangent() const { return FwdPerp(normal); }
};

template <class R>
struct PositionConstraint {
    Manifold<R> m;
    std::uint32_t bodyA, bodyB;
    R totalRadius;
};

// Revolute joint (RevoluteJointConf.hpp:27-123): its configuration, the
// state carried from step to step (impulse, angular motor impulse, limit
// state) and the per-step solver data (rA, rB, mass, angularMass).
enum : std::uint32_t { kLimitInactive = 0, kLimitLower = 1, kLimitUpper = 2, kLimitEqual = 3 };

template <class R>
struct V3 {
    R x = 0, y = 0, z = 0;
};
template <class R> inline R Dot3(V3<R> a, V3<R> b) {  // Math.hpp:320-333
    R r = 0;
    r += a.x * b.x;
    r += a.y * b.y;
    r += a.z * b.z;
    return r;
}
template <class R> inline V3<R> Cross3(V3<R> a, V3<R> b) {  // Math.hpp:397-409
    return V3<R>{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
// Mat33 as its three vectors GetX, GetY, GetZ (symmetric here).
template <class R>
struct M33 {
    V3<R> ex, ey, ez;
};
template <class R> inline V3<R> Solve33(const M33<R>& m, V3<R> b) {  // Math.hpp:446-456
    const R dp = Dot3(m.ex, Cross3(m.ey, m.ez));
    const R det = (dp != R(0)) ? (R(1) / dp) : dp;
    return V3<R>{det * Dot3(b, Cross3(m.ey, m.ez)), det * Dot3(m.ex, Cross3(b, m.ez)),
                 det * Dot3(m.ex, Cross3(m.ey, b))};
}
template <class R> inline V2<R> Solve22(const M33<R>& m, V2<R> b) {  // Math.hpp:461-474
    const R xx = m.ex.x, xy = m.ey.x, yx = m.ex.y, yy = m.ey.y;
    const R cp = xx * yy - xy * yx;
    const R det = (cp != R(0)) ? (R(1) / cp) : cp;
    return V2<R>{det * (yy * b.x - xy * b.y), det * (xx * b.y - yx * b.x)};
}

template <class R>
struct Joint {
    std::uint32_t bodyA = 0, bodyB = 0;
    bool destroyed = false;  // Destroy(World&, JointID) took it out of the world
    bool enableLimit = false, enableMotor = false, collideConnected = false;
    V2<R> laA, laB;
    R referenceAngle = 0, lower = 0, upper = 0, motorSpeed = 0, maxMotorTorque = 0;
    V3<R> impulse;
    R motorImpulse = 0;
    std::uint32_t limitState = kLimitInactive;
    V2<R> rA, rB;
    M33<R> mass;
    R angularMass = 0;
    // kind: 0 revolute, 1 distance, 2 weld.  Distance keeps its impulse in
    // impulse.x; length is the distance joint's rest length.
    std::uint32_t kind = 0;
    R length = 0, frequency = 0, dampingRatio = 0;
    V2<R> u;
    R dmass = 0, invGamma = 0, bias = 0, gamma = 0;
    // kind 3 rope: length is the maximum length, curLength the length found
    // by InitVelocity.  kind 4 friction: linear impulse in impulse.x, y and
    // angular impulse in impulse.z.
    R curLength = 0, maxForce = 0, maxTorque = 0;
    R lm00 = 0, lm01 = 0, lm10 = 0, lm11 = 0;
    // kind 5 motor: linear offset in laA, angular offset in referenceAngle,
    // impulses as friction.
    R correctionFactor = 0, angularError = 0;
    V2<R> linearError;
    // kind 6 wheel: impulse.x impulse, impulse.y spring impulse, impulse.z
    // angular (motor) impulse; dmass the point-to-line mass.
    V2<R> localXAxis, ax, ay;
    R sAx = 0, sAy = 0, sBx = 0, sBy = 0, springMass = 0;
    // kind 7 pulley: ground anchors, lengths (length, lengthB), ratio; uA in
    // u, uB in ay; impulse in impulse.x, mass in dmass.
    V2<R> groundA, groundB;
    R lengthB = 0, ratio = 1;
    // kind 8 prismatic: local x axis of A in localXAxis; translation limits in
    // lower, upper; the maximum motor force in maxMotorTorque.  Solver data:
    // axis in ax, perp in ay, s1, s2, a1, a2, K in mass, motorMass.
    R s1 = 0, s2 = 0, a1 = 0, a2 = 0, motorMass = 0;
    // kind 9 target: no body A (bodyA = kNoBody); the target in groundA, the
    // anchor on B in laB, maxForce, frequency, dampingRatio.  Solver data: rB,
    // gamma, the effective mass in lm00..lm11, C in linearError; the impulse
    // in impulse.x, y.
    // kind 10 gear: bodies A, B (in bodyA, bodyB) and C, D; per side (0: the
    // first joint's A-C, 1: the second's B-D) whether it is prismatic, its
    // local anchors A, B and local x axis (prismatic) or reference angle
    // (revolute); the ratio in ratio and the constant.  Solver data: the
    // Jacobians JvAC, JvBD, JwA..JwD, the mass in dmass, the impulse in
    // impulse.x.
    std::uint32_t bodyC = 0, bodyD = 0;
    bool gearPris[2] = {false, false};
    V2<R> gearLaA[2], gearLaB[2], gearAxis[2];
    R gearRef[2] = {0, 0};
    R constant = 0;
    V2<R> JvAC, JvBD;
    R JwA = 0, JwB = 0, JwC = 0, JwD = 0;
};

// InvalidBodyID's index (BodyID.hpp:50): a target joint's body A.
constexpr std::uint32_t kNoBody = 0xffffu;

template <class R>
class World {
public:
    std::vector<Joint<R>> joints;
    // m_bodyJoints: per body (other body, joint) in creation order.
    std::vector<std::vector<std::pair<std::uint32_t, std::uint32_t>>> bodyJoints;
    std::vector<char> islandedJoints;
    std::vector<Shape<R>> shapes;
    std::vector<Body<R>> bodies;
    std::vector<std::uint32_t> bodyOrder;  // m_bodies
    std::vector<Leaf<R>> tree;             // by proxy (leaf) id
    std::uint32_t leafCount = 0;
    std::vector<std::pair<std::uint32_t, std::uint32_t>> fixturesForProxies;  // (body, shape)
    std::vector<std::uint32_t> bodiesForSync;  // m_bodiesForSync (AabbTreeWorld.hpp)
    std::vector<std::uint32_t> proxiesForContacts;
    // Contact pool (ObjectPool.hpp:62-116) with manifolds at the same index.
    std::vector<Contact<R>> contactBuf;
    std::vector<Manifold<R>> manifoldBuf;
    std::vector<std::uint32_t> contactFree;
    // m_contacts and per-body contact lists: (kmin, kmax, id).
    struct KeyedId {
        std::uint32_t kmin, kmax, id;
    };
    std::vector<KeyedId> contacts;
    std::vector<std::vector<KeyedId>> bodyContacts;
    std::vector<char> islandedBodies, islandedContacts;

    // Leaf id allocation of DynamicTree (DynamicTree.cpp:456-537): the node
    // free list is a stack.  A leaf and its new parent take two nodes
    // (CreateLeaf :517-537) and DestroyLeaf (:539-569) frees the parent and
    // then the leaf, so a new leaf takes the node of the leaf destroyed last;
    // with no freed node, the first leaf takes node 0 and every later leaf the
    // next unused node, followed by its parent.
    std::vector<std::uint32_t> leafFree;
    std::uint32_t AllocLeafId() {
        if (!leafFree.empty()) {
            const std::uint32_t id = leafFree.back();
            leafFree.pop_back();
            return id;
        }
        const std::uint32_t id = (leafCount == 0) ? 0u : 2u * leafCount - 1u;
        ++leafCount;
        if (tree.size() <= id) tree.resize(id + 1);
        return id;
    }
    void FreeLeaf(std::uint32_t id) {
        tree[id] = Leaf<R>{};
        leafFree.push_back(id);
    }

    // ComputeAABB(proxy, xf0, xf1) (AABB.cpp:51-62) + GetFattenedAABB +
    // GetDisplacedAABB (detail/AABB.hpp:305-319, Interval.hpp:186-220).
    void FatAabb(const Proxy<R>& px, const Xf<R>& xf0, const Xf<R>& xf1, const StepConf<R>& conf,
                 R lo[2], R hi[2], bool fatten) const {
        R mn[2] = {std::numeric_limits<R>::infinity(), std::numeric_limits<R>::infinity()};
        R mx[2] = {-std::numeric_limits<R>::infinity(), -std::numeric_limits<R>::infinity()};
        auto include = [&](V2<R> p) {
            mn[0] = Min(p.x, mn[0]);
            mx[0] = Max(p.x, mx[0]);
            mn[1] = Min(p.y, mn[1]);
            mx[1] = Max(p.y, mx[1]);
        };
        for (std::uint32_t i = 0; i < px.n; ++i) {
            include(Transform(px.v[i], xf0));
            include(Transform(px.v[i], xf1));
        }
        for (int d = 0; d < 2; ++d) {
            mn[d] = mn[d] - px.radius;
            mx[d] = mx[d] + px.radius;
        }
        if (fatten) {
            for (int d = 0; d < 2; ++d) {
                mn[d] = mn[d] - conf.aabbExtension;
                mx[d] = mx[d] + conf.aabbExtension;
            }
            const V2<R> disp = conf.displaceMultiplier * (xf1.p - xf0.p);
            const R dd[2] = {disp.x, disp.y};
            for (int d = 0; d < 2; ++d) {
                if (dd[d] < R(0)) mn[d] += dd[d];
                else mx[d] += dd[d];
            }
        }
        for (int d = 0; d < 2; ++d) {
            lo[d] = mn[d];
            hi[d] = mx[d];
        }
    }

    static bool Overlap(const R alo[2], const R ahi[2], const R blo[2], const R bhi[2]) {
        for (int d = 0; d < 2; ++d) {
            const R maxOfMins = Max(alo[d], blo[d]);
            const R minOfMaxs = Min(ahi[d], bhi[d]);
            if (!(minOfMaxs >= maxOfMins)) return false;
        }
        return true;
    }

    Xf<R> Xf0(const Body<R>& b) const { return GetXf(b.pos0, b.lc); }

    // CreateProxies (AabbTreeWorld.cpp:557-579)
    void CreateProxies(std::uint32_t bodyId, std::uint32_t shapeId, const StepConf<R>& conf) {
        Body<R>& b = bodies[bodyId];
        const Xf<R> xf0 = Xf0(b);
        const Xf<R> xf1 = b.xf;
        const std::uint32_t id = AllocLeafId();
        Leaf<R>& lf = tree[id];
        lf.used = true;
        lf.body = bodyId;
        lf.shape = shapeId;
        lf.child = 0;
        FatAabb(shapes[shapeId].proxy(), xf0, xf1, conf, lf.lo, lf.hi, true);
        b.proxies.push_back(id);
        proxiesForContacts.push_back(id);
    }

    // Synchronize (AabbTreeWorld.cpp:2564-2585)
    std::uint32_t Synchronize(std::uint32_t bodyId, const StepConf<R>& conf) {
        Body<R>& b = bodies[bodyId];
        const Xf<R> xf0 = Xf0(b);
        const Xf<R> xf1 = b.xf;
        std::uint32_t moved = 0;
        for (std::uint32_t id : b.proxies) {
            Leaf<R>& lf = tree[id];
            R lo[2], hi[2];
            FatAabb(shapes[lf.shape].proxy(), xf0, xf1, conf, lo, hi, false);
            const bool contains = lf.lo[0] <= lo[0] && lf.hi[0] >= hi[0] && lf.lo[1] <= lo[1] && lf.hi[1] >= hi[1];
            if (!contains) {
                FatAabb(shapes[lf.shape].proxy(), xf0, xf1, conf, lf.lo, lf.hi, true);
                proxiesForContacts.push_back(id);
                ++moved;
            }
        }
        return moved;
    }

    // Contact listener calls of the current step in call order (the
    // reference calls the begin / end / pre-solve / post-solve listeners at
    // AabbTreeWorld.cpp:2249-2252, 2703-2719 and 1671-1674), serialised as
    // u32 kind (0 begin, 1 end, 2 pre-solve, 3 post-solve), u32 contact id,
    // then the pre-solve manifold (u32 type | count << 8, two contact
    // features, 12 floats) or the post-solve impulses (u32 count, u32 solved
    // position iteration, normal [2], tangent [2]).
    std::vector<std::uint32_t> events;
    std::uint32_t nevents = 0;
    void EventHead(std::uint32_t kind, std::uint32_t cid) {
        events.push_back(kind);
        events.push_back(cid);
        ++nevents;
    }
    static std::uint32_t FloatBits(R v) {
        const float f = float(v);
        std::uint32_t u;
        std::memcpy(&u, &f, 4);
        return u;
    }
    void EventManifold(const Manifold<R>& m) {
        events.push_back(m.type | (m.n << 8));
        events.push_back(m.n > 0 ? m.pts[0].cf : kDefaultCf);
        events.push_back(m.n > 1 ? m.pts[1].cf : kDefaultCf);
        R f[12] = {};
        if (m.type == kFaceA || m.type == kFaceB) {
            f[0] = m.ln.x;
            f[1] = m.ln.y;
        }
        if (m.type != kUnset) {
            f[2] = m.lpt.x;
            f[3] = m.lpt.y;
        }
        for (std::uint32_t p = 0; p < m.n; ++p) {
            f[4 + 4 * p + 0] = m.pts[p].lp.x;
            f[4 + 4 * p + 1] = m.pts[p].lp.y;
            f[4 + 4 * p + 2] = m.pts[p].ni;
            f[4 + 4 * p + 3] = m.pts[p].ti;
        }
        for (R v : f) events.push_back(FloatBits(v));
    }

    // InternalDestroy (AabbTreeWorld.cpp:2245-2279), called by DestroyContacts
    // and, with `from` the body being destroyed, by Destroy(World&, BodyID).
    void InternalDestroy(std::uint32_t cid, std::uint32_t from = 0xffffffffu) {
        Contact<R>& c = contactBuf[cid];
        if (c.flags & kTouching) EventHead(1, cid);  // :2249-2252
        auto eraseFrom = [&](std::uint32_t body) {
            auto& lst = bodyContacts[body];
            for (auto it = lst.begin(); it != lst.end(); ++it) {
                if (it->id == cid) {
                    lst.erase(it);
                    break;
                }
            }
        };
        if (c.bodyA != from) eraseFrom(c.bodyA);
        if (c.bodyB != from) eraseFrom(c.bodyB);
        if (manifoldBuf[cid].n > 0 && !(c.flags & kSensor)) {
            bodies[c.bodyA].SetAwake();
            bodies[c.bodyB].SetAwake();
        }
        contactFree.push_back(cid);
        contactBuf[cid] = Contact<R>{};
        manifoldBuf[cid] = Manifold<R>{};
    }

    // DestroyContacts (AabbTreeWorld.cpp:2290-2331); filtering is never flagged
    // by the operations this restatement supports.
    std::uint32_t DestroyContacts() {
        std::uint32_t destroyed = 0;
        std::vector<KeyedId> kept;
        kept.reserve(contacts.size());
        for (const auto& kc : contacts) {
            const Leaf<R>& a = tree[kc.kmin];
            const Leaf<R>& b = tree[kc.kmax];
            if (!Overlap(a.lo, a.hi, b.lo, b.hi)) {
                InternalDestroy(kc.id);
                ++destroyed;
            } else {
                kept.push_back(kc);
            }
        }
        contacts.swap(kept);
        return destroyed;
    }

    // Update (AabbTreeWorld.cpp:2587-2721) for a non-sensor contact.
    void Update(std::uint32_t cid, const StepConf<R>& conf) {
        Contact<R>& c = contactBuf[cid];
        Manifold<R>& manifold = manifoldBuf[cid];
        const Manifold<R> old = manifold;
        const bool oldTouching = (c.flags & kTouching) != 0;
        bool newTouching = false;
        const Body<R>& bA = bodies[c.bodyA];
        const Body<R>& bB = bodies[c.bodyB];
        if (c.flags & kSensor) {
            // Sensors are not produced by the supported scene API.
            manifold = Manifold<R>{};
        } else {
            Manifold<R> nm = CollideShapes(shapes[c.shapeA].proxy(), bA.xf, shapes[c.shapeB].proxy(), bB.xf,
                                           conf.linearSlop, conf.maxCirclesRatio);
            const std::uint32_t oldN = old.n, newN = nm.n;
            newTouching = newN > 0;
            bool found[2] = {false, newN < 2};
            for (std::uint32_t i = 0; i < newN; ++i) {
                for (std::uint32_t j = 0; j < oldN; ++j) {
                    if (nm.pts[i].cf == old.pts[j].cf) {
                        found[i] = true;
                        nm.pts[i].ni = old.pts[j].ni;
                        nm.pts[i].ti = old.pts[j].ti;
                        break;
                    }
                }
            }
            for (std::uint32_t i = 0; i < newN; ++i) {
                if (!found[i]) {
                    R least = std::numeric_limits<R>::infinity();
                    for (std::uint32_t j = 0; j < oldN; ++j) {
                        const R d = MagSq(old.pts[j].lp - nm.pts[i].lp);
                        if (least > d) {
                            least = d;
                            nm.pts[i].ni = old.pts[j].ni;
                            nm.pts[i].ti = old.pts[j].ti;
                        }
                    }
                }
            }
            manifold = nm;
        }
        c.flags &= ~kNeedsUpdating;
        if (!oldTouching && newTouching) {
            c.flags |= kTouching;
            EventHead(0, cid);
        } else if (oldTouching && !newTouching) {
            c.flags &= ~kTouching;
            EventHead(1, cid);
        }
        if (!(c.flags & kSensor) && newTouching) {
            EventHead(2, cid);
            EventManifold(old);
        }
    }

    // UpdateContacts (AabbTreeWorld.cpp:2333-2421)
    std::uint32_t UpdateContacts(const StepConf<R>& conf, std::uint32_t* skipped = nullptr) {
        std::uint32_t updated = 0, skip = 0;
        for (const auto& kc : contacts) {
            if (contactBuf[kc.id].flags & kNeedsUpdating) {
                Update(kc.id, conf);
                ++updated;
            } else {
                ++skip;
            }
        }
        if (skipped) *skipped = skip;
        return updated;
    }

    // FindContacts (AabbTreeWorld.cpp:709-751) + AddContacts (:2424-2547).
    std::uint32_t FindAndAddContacts(const StepConf<R>& conf) {
        struct PK {
            std::uint32_t kmin, kmax;
        };
        std::vector<PK> keys;
        for (std::uint32_t pid : proxiesForContacts) {
            const Leaf<R>& l0 = tree[pid];
            for (std::uint32_t nid = 0; nid < tree.size(); ++nid) {
                const Leaf<R>& l1 = tree[nid];
                if (!l1.used) continue;
                if (!Overlap(l0.lo, l0.hi, l1.lo, l1.hi)) continue;
                if (nid != pid && l0.body != l1.body) {
                    keys.push_back(PK{std::min(pid, nid), std::max(pid, nid)});
                }
            }
        }
        proxiesForContacts.clear();
        std::sort(keys.begin(), keys.end(), [](const PK& a, const PK& b) {
            return (a.kmin < b.kmin) || (a.kmin == b.kmin && a.kmax < b.kmax);
        });
        keys.erase(std::unique(keys.begin(), keys.end(),
                               [](const PK& a, const PK& b) { return a.kmin == b.kmin && a.kmax == b.kmax; }),
                   keys.end());
        std::uint32_t added = 0;
        for (const PK& k : keys) {
            const Leaf<R>& la = tree[k.kmin];
            const Leaf<R>& lb = tree[k.kmax];
            Body<R>& bA = bodies[la.body];
            Body<R>& bB = bodies[lb.body];
            const Shape<R>& sA = shapes[la.shape];
            const Shape<R>& sB = shapes[lb.shape];
            if (!(bA.accelerable() || bB.accelerable())) continue;
            // ShouldCollide(joints) (AabbTreeWorld.cpp:479-490).
            {
                bool jointed = false;
                for (const auto& bj : bodyJoints[la.body]) {
                    if (bj.first == lb.body && !joints[bj.second].collideConnected) jointed = true;
                }
                if (jointed) continue;
            }
            // ShouldCollide(shapes) (Filter.hpp:81-88).
            {
                const std::int32_t gA = sA.group, gB = sB.group;
                bool collide;
                if (gA == gB && gA != 0) {
                    collide = gA > 0;
                } else {
                    const std::uint32_t catA = sA.catmask & 0xffffu, maskA = sA.catmask >> 16;
                    const std::uint32_t catB = sB.catmask & 0xffffu, maskB = sB.catmask >> 16;
                    collide = ((maskA & catB) != 0) && ((maskB & catA) != 0);
                }
                if (!collide) continue;
            }
            auto& lA = bodyContacts[la.body];
            auto& lB = bodyContacts[lb.body];
            const auto& search = (lA.size() < lB.size()) ? lA : lB;
            bool exists = false;
            for (const auto& e : search) {
                if (e.kmin == k.kmin && e.kmax == k.kmax) {
                    exists = true;
                    break;
                }
            }
            if (exists) continue;
            std::uint32_t cid;
            if (!contactFree.empty()) {
                cid = contactFree.back();
                contactFree.pop_back();
            } else {
                cid = std::uint32_t(contactBuf.size());
                contactBuf.emplace_back();
                manifoldBuf.emplace_back();
            }
            Contact<R>& c = contactBuf[cid];
            c = Contact<R>{};
            manifoldBuf[cid] = Manifold<R>{};
            c.kmin = k.kmin;
            c.kmax = k.kmax;
            c.bodyA = la.body;
            c.bodyB = lb.body;
            c.shapeA = la.shape;
            c.shapeB = lb.shape;
            c.flags = kCEnabled | kNeedsUpdating;
            if ((bA.flags & kImpenetrable) || (bB.flags & kImpenetrable)) c.flags |= kCImpenetrable;
            if (sA.sensor || sB.sensor) c.flags |= kSensor;
            c.friction = std::sqrt(sA.friction * sB.friction);                       // Contact.hpp:57-60
            c.restitution = (sA.restitution > sB.restitution) ? sA.restitution : sB.restitution;  // :66-69
            contacts.push_back(KeyedId{k.kmin, k.kmax, cid});
            lA.push_back(KeyedId{k.kmin, k.kmax, cid});
            lB.push_back(KeyedId{k.kmin, k.kmax, cid});
            if (!(c.flags & kSensor)) {
                if (bA.speedable()) bA.SetAwakeFlag();
                if (bB.speedable()) bB.SetAwakeFlag();
            }
            Update(cid, conf);
            ++added;
        }
        return added;
    }

    // GetVelocity(body, h) (Body.cpp:290-309) and Cap (Velocity.cpp:27-45).
    Vel<R> IntegratedVelocity(const Body<R>& b, const StepConf<R>& conf) const {
        Vel<R> v = b.vel;
        const R h = conf.dt;
        if (b.accelerable()) {
            v.lin = v.lin + h * b.lacc;
            v.ang = v.ang + h * b.aacc;
            const R ldiv = R(1) + h * b.ldamp;
            const R linv = R(1) / ldiv;
            v.lin = V2<R>{v.lin.x * linv, v.lin.y * linv};
            v.ang = v.ang / (R(1) + h * b.adamp);
        }
        const V2<R> translation = h * v.lin;
        const R lsq = MagSq(translation);
        if (lsq > conf.maxTranslation * conf.maxTranslation) {
            const R ratio = conf.maxTranslation / std::sqrt(lsq);
            v.lin = V2<R>{v.lin.x * ratio, v.lin.y * ratio};
        }
        const R absRot = std::abs(h * v.ang);
        if (absRot > conf.maxRotation) {
            const R ratio = conf.maxRotation / absRot;
            v.ang = v.ang * ratio;
        }
        return v;
    }

    // GetContactRelVelocity (d2/Math.cpp:73-103)
    static V2<R> RelVel(const Vel<R>& vA, V2<R> relA, const Vel<R>& vB, V2<R> relB) {
        const V2<R> velBrot = RevPerp(relB) * (vB.ang / R(1));
        const V2<R> velArot = RevPerp(relA) * (vA.ang / R(1));
        return (vB.lin + velBrot) - (vA.lin + velArot);
    }

    // VelocityConstraint ctor + GetPoint + ComputeK (VelocityConstraint.cpp:38-185)
    VelocityConstraint<R> MakeVC(const Contact<R>& c, const WorldManifold<R>& wm,
                                 const std::vector<BodyConstraint<R>>& bcs, const StepConf<R>& conf) const {
        VelocityConstraint<R> vc;
        vc.bodyA = c.bodyA;
        vc.bodyB = c.bodyB;
        vc.normal = wm.normal;
        vc.friction = c.friction;
        vc.restitution = c.restitution;
        vc.tangentSpeed = c.tangentSpeed;
        const BodyConstraint<R>& A = bcs[c.bodyA];
        const BodyConstraint<R>& B = bcs[c.bodyB];
        const R dtRatio = conf.warmStart ? conf.dtRatio : R(0);
        const R invMass = A.invMass + B.invMass;
        for (std::uint32_t j = 0; j < wm.n; ++j) {
            VcPoint<R>& p = vc.p[vc.n];
            const V2<R> relA = wm.p[j] - A.pos.lin;
            const V2<R> relB = wm.p[j] - B.pos.lin;
            p.ni = wm.ni[j] * dtRatio;
            p.ti = wm.ti[j] * dtRatio;
            {
                const V2<R> dv = RelVel(A.vel, relA, B.vel, relB);
                const R vn = Dot(dv, vc.normal);
                p.velocityBias = (vn < -conf.velocityThreshold) ? -vc.restitution * vn : R(0);
            }
            p.relA = relA;
            p.relB = relB;
            {
                const R cA = Cross(relA, vc.normal);
                const R cB = Cross(relB, vc.normal);
                const R rmA = A.invRotI * (cA * cA) / R(1);
                const R rmB = B.invRotI * (cB * cB) / R(1);
                const R value = invMass + rmA + rmB;
                p.normalMass = (value != R(0)) ? R(1) / value : R(0);
            }
            {
                const V2<R> t = vc.tangent();
                const R cA = Cross(relA, t);
                const R cB = Cross(relB, t);
                const R rmA = A.invRotI * (cA * cA) / R(1);
                const R rmB = B.invRotI * (cB * cB) / R(1);
                const R value = invMass + rmA + rmB;
                p.tangentMass = (value != R(0)) ? R(1) / value : R(0);
            }
            ++vc.n;
        }
        if (conf.blockSolve && vc.n == 2) {
            const R rnA0 = Cross(vc.p[0].relA, vc.normal);
            const R rnB0 = Cross(vc.p[0].relB, vc.normal);
            const R rnA1 = Cross(vc.p[1].relA, vc.normal);
            const R rnB1 = Cross(vc.p[1].relB, vc.normal);
            const R irmA0 = (A.invRotI * (rnA0 * rnA0)) / R(1);
            const R irmA1 = (A.invRotI * (rnA1 * rnA1)) / R(1);
            const R irmB0 = (B.invRotI * (rnB0 * rnB0)) / R(1);
            const R irmB1 = (B.invRotI * (rnB1 * rnB1)) / R(1);
            const R irmA = (A.invRotI * rnA0 * rnA1) / R(1);
            const R irmB = (B.invRotI * rnB0 * rnB1) / R(1);
            const R k00 = invMass + irmA0 + irmB0;
            const R k11 = invMass + irmA1 + irmB1;
            const R k01 = invMass + irmA + irmB;
            const R k00sq = k00 * k00;
            const R kdiff = k00 * k11 - k01 * k01;
            if (k00sq < R(1000) * kdiff) {
                vc.K[0] = k00;
                vc.K[1] = k11;
                vc.K[2] = k01;
                // Invert (Math.hpp:428-442) of [[k00, k01], [k01, k11]].
                const R cp = k00 * k11 - k01 * k01;
                if (!AlmostZero(cp)) {
                    const R inv = R(1) / cp;
                    vc.normalMass[0] = k11 * inv;
                    vc.normalMass[1] = k00 * inv;
                    vc.normalMass[2] = -k01 * inv;
                }
            } else {
                vc.n = 1;  // RemovePoint
            }
        }
        return vc;
    }

    // SolveTangentConstraint / SeqSolveNormalConstraint (ContactSolver.cpp:337-454)
    static R SolveDirectional(VelocityConstraint<R>& vc, std::vector<BodyConstraint<R>>& bcs, bool tangent) {
        R maxInc = 0;
        BodyConstraint<R>& A = bcs[vc.bodyA];
        BodyConstraint<R>& B = bcs[vc.bodyB];
        const V2<R> dir = tangent ? vc.tangent() : vc.normal;
        Vel<R> vA = A.vel, vB = B.vel;
        auto proc = [&](std::uint32_t i) {
            VcPoint<R>& p = vc.p[i];
            const V2<R> closing = RelVel(vA, p.relA, vB, p.relB);
            R oldImpulse, newImpulse;
            if (tangent) {
                const R dirVel = vc.tangentSpeed - Dot(closing, dir);
                const R lambda = p.tangentMass * dirVel;
                const R maxImpulse = vc.friction * p.ni;
                oldImpulse = p.ti;
                newImpulse = Clamp(oldImpulse + lambda, -maxImpulse, maxImpulse);
            } else {
                const R dirVel = Dot(closing, dir);
                const R lambda = p.normalMass * (p.velocityBias - dirVel);
                oldImpulse = p.ni;
                newImpulse = Max(oldImpulse + lambda, R(0));
            }
            const R inc = newImpulse - oldImpulse;
            const V2<R> P = inc * dir;
            const R LA = Cross(p.relA, P) / R(1);
            const R LB = Cross(p.relB, P) / R(1);
            vA.lin = vA.lin - A.invMass * P;
            vA.ang = vA.ang - A.invRotI * LA;
            vB.lin = vB.lin + B.invMass * P;
            vB.ang = vB.ang + B.invRotI * LB;
            maxInc = Max(maxInc, std::abs(inc));
            if (tangent) p.ti = oldImpulse + inc;
            else p.ni = oldImpulse + inc;
        };
        if (vc.n == 2) proc(1);
        proc(0);
        A.vel = vA;
        B.vel = vB;
        return maxInc;
    }

    // BlockSolveNormalConstraint and its cases (ContactSolver.cpp:65-335)
    static R BlockSolve(VelocityConstraint<R>& vc, std::vector<BodyConstraint<R>>& bcs) {
        BodyConstraint<R>& A = bcs[vc.bodyA];
        BodyConstraint<R>& B = bcs[vc.bodyB];
        const V2<R> normal = vc.normal;
        const R k00 = vc.K[0], k11 = vc.K[1], k01 = vc.K[2];
        V2<R> bprime;
        {
            const V2<R> dv0 = RelVel(A.vel, vc.p[0].relA, B.vel, vc.p[0].relB);
            const V2<R> dv1 = RelVel(A.vel, vc.p[1].relA, B.vel, vc.p[1].relB);
            const R vn0 = Dot(dv0, normal);
            const R vn1 = Dot(dv1, normal);
            const V2<R> b{vn0 - vc.p[0].velocityBias, vn1 - vc.p[1].velocityBias};
            // Transform(impulses, K) = K * impulses with K rows {k00,k01},{k01,k11}
            // (Vector.hpp:458-470); InvMass3 unpacks to a 2x2 (Math.hpp).
            R t0 = 0;
            t0 += k00 * vc.p[0].ni;
            t0 += k01 * vc.p[1].ni;
            R t1 = 0;
            t1 += k01 * vc.p[0].ni;
            t1 += k11 * vc.p[1].ni;
            bprime = V2<R>{b.x - t0, b.y - t1};
        }
        auto update = [&](R n0, R n1) {
            // BlockSolveUpdate + GetVelocityDelta (ContactSolver.cpp:65-111)
            const R d0 = n0 - vc.p[0].ni;
            const R d1 = n1 - vc.p[1].ni;
            const V2<R> P0 = d0 * normal;
            const V2<R> P1 = d1 * normal;
            const V2<R> P = P0 + P1;
            const R LA = (Cross(vc.p[0].relA, P0) + Cross(vc.p[1].relA, P1)) / R(1);
            const R LB = (Cross(vc.p[0].relB, P0) + Cross(vc.p[1].relB, P1)) / R(1);
            const V2<R> dvAlin = -(A.invMass * P);
            const R dvAang = -(A.invRotI * LA);
            const V2<R> dvBlin = B.invMass * P;
            const R dvBang = B.invRotI * LB;
            A.vel.lin = A.vel.lin + dvAlin;
            A.vel.ang = A.vel.ang + dvAang;
            B.vel.lin = B.vel.lin + dvBlin;
            B.vel.ang = B.vel.ang + dvBang;
            vc.p[0].ni = n0;
            vc.p[1].ni = n1;
            return Max(std::abs(n0), std::abs(n1));
        };
        {
            // Case 1: -Transform(b', normalMass) (ContactSolver.cpp:113-141)
            const R m00 = vc.normalMass[0], m11 = vc.normalMass[1], m01 = vc.normalMass[2];
            R x0 = 0;
            x0 += m00 * bprime.x;
            x0 += m01 * bprime.y;
            R x1 = 0;
            x1 += m01 * bprime.x;
            x1 += m11 * bprime.y;
            const R n0 = -x0, n1 = -x1;
            if (n0 >= R(0) && n1 >= R(0)) return update(n0, n1);
        }
        {
            // Case 2 (ContactSolver.cpp:143-172)
            const R n0 = -vc.p[0].normalMass * bprime.x;
            const R vn2 = k01 * n0 + bprime.y;
            if (n0 >= R(0) && vn2 >= R(0)) return update(n0, R(0));
        }
        {
            // Case 3 (ContactSolver.cpp:174-203)
            const R n1 = -vc.p[1].normalMass * bprime.y;
            const R vn1 = k01 * n1 + bprime.x;
            if (n1 >= R(0) && vn1 >= R(0)) return update(R(0), n1);
        }
        {
            // Case 4 (ContactSolver.cpp:205-216)
            if (bprime.x >= R(0) && bprime.y >= R(0)) return update(R(0), R(0));
        }
        return R(0);
    }

    // GaussSeidel::SolveVelocityConstraint (ContactSolver.cpp:456-492)
    static R SolveVC(VelocityConstraint<R>& vc, std::vector<BodyConstraint<R>>& bcs) {
        R maxInc = 0;
        maxInc = Max(maxInc, SolveDirectional(vc, bcs, true));
        R n;
        if (vc.n == 1 || (vc.K[0] == R(0) && vc.K[1] == R(0) && vc.K[2] == R(0))) {
            n = SolveDirectional(vc, bcs, false);
        } else {
            n = BlockSolve(vc, bcs);
        }
        maxInc = Max(maxInc, n);
        return maxInc;
    }

    // GaussSeidel::SolvePositionConstraint (ContactSolver.cpp:494-643)
    static R SolvePC(const PositionConstraint<R>& pc, std::vector<BodyConstraint<R>>& bcs, const StepConf<R>& conf) {
        BodyConstraint<R>& A = bcs[pc.bodyA];
        BodyConstraint<R>& B = bcs[pc.bodyB];
        const R invMassA = A.invMass, invRotIA = A.invRotI;
        const R invMassB = B.invMass, invRotIB = B.invRotI;
        const V2<R> lcA = A.lc, lcB = B.lc;
        const R invMassTotal = invMassA + invMassB;
        const R totalRadius = pc.totalRadius;
        struct Sol {
            Pos<R> a, b;
            R sep;
        };
        auto solve = [&](const PSM<R>& psm, V2<R> pA, V2<R> pB) {
            const R separation = psm.sep - totalRadius;
            const V2<R> rA = psm.point - pA;
            const V2<R> rB = psm.point - pB;
            const R crA = Cross(rA, psm.normal) / R(1);
            const R crB = Cross(rB, psm.normal) / R(1);
            const R K = invMassTotal + invRotIA * (crA * crA) + invRotIB * (crB * crB);
            const R C = -Clamp(conf.regResolutionRate * (separation + conf.linearSlop), -conf.maxLinearCorrection, R(0));
            const V2<R> P = (psm.normal * C) / K;
            const R LA = Cross(rA, P) / R(1);
            const R LB = Cross(rB, P) / R(1);
            return Sol{-Pos<R>{invMassA * P, invRotIA * LA}, Pos<R>{invMassB * P, invRotIB * LB}, separation};
        };
        Pos<R> posA = A.pos, posB = B.pos;
        R minSep = std::numeric_limits<R>::infinity();
        if (pc.m.n == 1) {
            const PSM<R> psm0 = GetPSM(pc.m, 0, GetXf(posA, lcA), GetXf(posB, lcB));
            const Sol s = solve(psm0, posA.lin, posB.lin);
            posA = posA + s.a;
            posB = posB + s.b;
            minSep = R(0) + s.sep;
        } else if (pc.m.n == 2) {
            const Xf<R> xfA = GetXf(posA, lcA);
            const Xf<R> xfB = GetXf(posB, lcB);
            const PSM<R> psm0 = GetPSM(pc.m, 0, xfA, xfB);
            const PSM<R> psm1 = GetPSM(pc.m, 1, xfA, xfB);
            if (AlmostEqual(psm0.sep, psm1.sep)) {
                const Sol s0 = solve(psm0, posA.lin, posB.lin);
                const Sol s1 = solve(psm1, posA.lin, posB.lin);
                posA = posA + s0.a + s1.a;
                posB = posB + s0.b + s1.b;
                minSep = s0.sep;
            } else if (psm0.sep < psm1.sep) {
                const Sol s0 = solve(psm0, posA.lin, posB.lin);
                posA = posA + s0.a;
                posB = posB + s0.b;
                const PSM<R> p1 = GetPSM(pc.m, 1, GetXf(posA, lcA), GetXf(posB, lcB));
                const Sol s1 = solve(p1, posA.lin, posB.lin);
                posA = posA + s1.a;
                posB = posB + s1.b;
                minSep = s0.sep;
            } else if (psm1.sep < psm0.sep) {
                const Sol s1 = solve(psm1, posA.lin, posB.lin);
                posA = posA + s1.a;
                posB = posB + s1.b;
                const PSM<R> p0 = GetPSM(pc.m, 0, GetXf(posA, lcA), GetXf(posB, lcB));
                const Sol s0 = solve(p0, posA.lin, posB.lin);
                posA = posA + s0.a;
                posB = posB + s0.b;
                minSep = s1.sep;
            }
        }
        A.pos = posA;
        B.pos = posB;
        return minSep;
    }

    struct Island {
        std::vector<std::uint32_t> bodies, contacts, joints;
    };

    // GetMat33 (RevoluteJointConf.cpp:36-62).
    static M33<R> JointMass(R mA, V2<R> rA, R iA, R mB, V2<R> rB, R iB) {
        const R totInvI = iA + iB;
        const R exx = mA + ((rA.y * rA.y) * iA / R(1)) + mB + ((rB.y * rB.y) * iB / R(1));
        const R eyx = (-rA.y * rA.x * iA / R(1)) + (-rB.y * rB.x * iB / R(1));
        const R ezx = (-rA.y * iA * R(1) / R(1)) + (-rB.y * iB * R(1) / R(1));
        const R eyy = mA + ((rA.x * rA.x) * iA / R(1)) + mB + ((rB.x * rB.x) * iB / R(1));
        const R ezy = (rA.x * iA * R(1) / R(1)) + (rB.x * iB * R(1) / R(1));
        return M33<R>{V3<R>{exx, eyx, ezx}, V3<R>{eyx, eyy, ezy}, V3<R>{ezx, ezy, totInvI}};
    }

    // GetInverse22(Mat33) (Math.hpp:478-490).
    static M33<R> Inverse22(const M33<R>& k) {
        const R a = k.ex.x, b = k.ey.x, c = k.ex.y, d = k.ey.y;
        R det = (a * d) - (b * c);
        if (det != R(0)) det = R(1) / det;
        return M33<R>{V3<R>{det * d, -det * c, R(0)}, V3<R>{-det * b, det * a, R(0)}, V3<R>{R(0), R(0), R(0)}};
    }
    // GetSymInverse33(Mat33) (Math.hpp:493-510).
    static M33<R> SymInverse33(const M33<R>& k) {
        const R invDet = Dot3(k.ex, Cross3(k.ey, k.ez));
        const R det = (invDet != R(0)) ? (R(1) / invDet) : invDet;
        const R a11 = k.ex.x, a12 = k.ey.x, a13 = k.ez.x, a22 = k.ey.y, a23 = k.ez.y, a33 = k.ez.z;
        const R exy = det * (a13 * a23 - a12 * a33);
        const R eyz = det * (a13 * a12 - a11 * a23);
        const R exz = det * (a12 * a23 - a13 * a22);
        return M33<R>{V3<R>{det * (a22 * a33 - a23 * a23), exy, exz}, V3<R>{exy, det * (a11 * a33 - a13 * a13), eyz},
                      V3<R>{exz, eyz, det * (a11 * a22 - a12 * a12)}};
    }
    // Matrix times vector, row by row from zero (Vector.hpp:455-470).
    static V3<R> MulRows(const M33<R>& m, V3<R> v) {
        auto row = [&](V3<R> r) {
            R e = 0;
            e += r.x * v.x;
            e += r.y * v.y;
            e += r.z * v.z;
            return e;
        };
        return V3<R>{row(m.ex), row(m.ey), row(m.ez)};
    }
    static void ApplyImpulse(BodyConstraint<R>& A, BodyConstraint<R>& B, V2<R> P, R LA, R LB, Vel<R>& velA,
                             Vel<R>& velB) {
        velA.lin = velA.lin - A.invMass * P;
        velA.ang = velA.ang - A.invRotI * LA;
        velB.lin = velB.lin + B.invMass * P;
        velB.ang = velB.ang + B.invRotI * LB;
    }

    // InitVelocity(DistanceJointConf&) (DistanceJointConf.cpp:81-162).
    static void DistanceInitVelocity(Joint<R>& j, std::vector<BodyConstraint<R>>& bcs, const StepConf<R>& conf) {
        BodyConstraint<R>& A = bcs[j.bodyA];
        BodyConstraint<R>& B = bcs[j.bodyB];
        const R mA = A.invMass, iA = A.invRotI, mB = B.invMass, iB = B.invRotI;
        Vel<R> velA = A.vel, velB = B.vel;
        j.rA = Rotate(j.laA - A.lc, UnitFromAngle(A.pos.ang));
        j.rB = Rotate(j.laB - B.lc, UnitFromAngle(B.pos.ang));
        R length;
        j.u = UnitPolar((B.pos.lin + j.rB) - (A.pos.lin + j.rA), length);
        const R crAu = Cross(j.rA, j.u) / R(1);
        const R crBu = Cross(j.rB, j.u) / R(1);
        R invMass = mA + iA * (crAu * crAu) + mB + iB * (crBu * crBu);
        j.dmass = (invMass != R(0)) ? R(1) / invMass : R(0);
        if (j.frequency > R(0)) {
            const R C = length - j.length;
            const R omega = R(2) * kPi * j.frequency;
            const R d = R(2) * j.dmass * j.dampingRatio * omega;
            const R k = j.dmass * (omega * omega);
            const R h = conf.dt;
            const R gamma = h * (d + h * k);
            j.invGamma = (gamma != R(0)) ? R(1) / gamma : R(0);
            j.bias = C * h * k * j.invGamma;
            invMass += j.invGamma;
            j.dmass = (invMass != R(0)) ? R(1) / invMass : R(0);
        } else {
            j.invGamma = 0;
            j.bias = 0;
        }
        if (conf.warmStart) {
            j.impulse.x *= conf.dtRatio;
            const V2<R> P = j.impulse.x * j.u;
            ApplyImpulse(A, B, P, Cross(j.rA, P) / R(1), Cross(j.rB, P) / R(1), velA, velB);
        } else {
            j.impulse.x = 0;
        }
        A.vel = velA;
        B.vel = velB;
    }

    // SolveVelocity(DistanceJointConf&) (DistanceJointConf.cpp:164-194).
    static bool DistanceSolveVelocity(Joint<R>& j, std::vector<BodyConstraint<R>>& bcs) {
        BodyConstraint<R>& A = bcs[j.bodyA];
        BodyConstraint<R>& B = bcs[j.bodyB];
        Vel<R> velA = A.vel, velB = B.vel;
        const V2<R> vpA = velA.lin + RevPerp(j.rA) * (velA.ang / R(1));
        const V2<R> vpB = velB.lin + RevPerp(j.rB) * (velB.ang / R(1));
        const R Cdot = Dot(j.u, vpB - vpA);
        const R impulse = -j.dmass * (Cdot + j.bias + j.invGamma * j.impulse.x);
        j.impulse.x += impulse;
        const V2<R> P = impulse * j.u;
        ApplyImpulse(A, B, P, Cross(j.rA, P) / R(1), Cross(j.rB, P) / R(1), velA, velB);
        A.vel = velA;
        B.vel = velB;
        return impulse == R(0);
    }

    // SolvePosition(const DistanceJointConf&) (DistanceJointConf.cpp:196-245).
    static bool DistanceSolvePosition(const Joint<R>& j, std::vector<BodyConstraint<R>>& bcs, const StepConf<R>& conf) {
        if (j.frequency > R(0)) return true;
        BodyConstraint<R>& A = bcs[j.bodyA];
        BodyConstraint<R>& B = bcs[j.bodyB];
        Pos<R> posA = A.pos, posB = B.pos;
        const V2<R> rA = Rotate(j.laA - A.lc, UnitFromAngle(posA.ang));
        const V2<R> rB = Rotate(j.laB - B.lc, UnitFromAngle(posB.ang));
        R length;
        const V2<R> u = UnitPolar((posB.lin + rB) - (posA.lin + rA), length);
        const R C = Clamp(length - j.length, -conf.maxLinearCorrection, conf.maxLinearCorrection);
        const R impulse = -j.dmass * C;
        const V2<R> P = impulse * u;
        posA.lin = posA.lin - A.invMass * P;
        posA.ang = posA.ang - A.invRotI * (Cross(rA, P) / R(1));
        posB.lin = posB.lin + B.invMass * P;
        posB.ang = posB.ang + B.invRotI * (Cross(rB, P) / R(1));
        A.pos = posA;
        B.pos = posB;
        return std::abs(C) < conf.linearSlop;
    }

    // InitVelocity(WeldJointConf&) (WeldJointConf.cpp:115-213).
    static void WeldInitVelocity(Joint<R>& j, std::vector<BodyConstraint<R>>& bcs, const StepConf<R>& conf) {
        BodyConstraint<R>& A = bcs[j.bodyA];
        BodyConstraint<R>& B = bcs[j.bodyB];
        const R mA = A.invMass, iA = A.invRotI, mB = B.invMass, iB = B.invRotI;
        const R aA = A.pos.ang, aB = B.pos.ang;
        Vel<R> velA = A.vel, velB = B.vel;
        j.rA = Rotate(j.laA - A.lc, UnitFromAngle(aA));
        j.rB = Rotate(j.laB - B.lc, UnitFromAngle(aB));
        const M33<R> K = JointMass(mA, j.rA, iA, mB, j.rB, iB);
        if (j.frequency > R(0)) {
            j.mass = Inverse22(K);
            R invRotI = iA + iB;
            const R rotI = (invRotI > R(0)) ? R(1) / invRotI : R(0);
            const R C = aB - aA - j.referenceAngle;
            const R omega = R(2) * kPi * j.frequency;
            const R d = R(2) * rotI * j.dampingRatio * omega;
            const R k = rotI * omega * omega;
            const R h = conf.dt;
            const R invGamma = h * (d + h * k);
            j.gamma = (invGamma != R(0)) ? R(1) / invGamma : R(0);
            j.bias = C * h * k * j.gamma;
            invRotI += j.gamma;
            j.mass.ez.z = (invRotI != R(0)) ? R(1) / invRotI : R(0);
        } else if (K.ez.z == R(0)) {
            j.mass = Inverse22(K);
            j.gamma = 0;
            j.bias = 0;
        } else {
            j.mass = SymInverse33(K);
            j.gamma = 0;
            j.bias = 0;
        }
        if (conf.warmStart) {
            j.impulse.x *= conf.dtRatio;
            j.impulse.y *= conf.dtRatio;
            j.impulse.z *= conf.dtRatio;
            const V2<R> P{j.impulse.x, j.impulse.y};
            const R L = j.impulse.z;
            ApplyImpulse(A, B, P, L + Cross(j.rA, P) / R(1), L + Cross(j.rB, P) / R(1), velA, velB);
        } else {
            j.impulse = V3<R>{};
        }
        A.vel = velA;
        B.vel = velB;
    }

    // SolveVelocity(WeldJointConf&) (WeldJointConf.cpp:215-303).
    static bool WeldSolveVelocity(Joint<R>& j, std::vector<BodyConstraint<R>>& bcs) {
        BodyConstraint<R>& A = bcs[j.bodyA];
        BodyConstraint<R>& B = bcs[j.bodyB];
        const Vel<R> oldVelA = A.vel, oldVelB = B.vel;
        Vel<R> velA = oldVelA, velB = oldVelB;
        const R iA = A.invRotI, iB = B.invRotI;
        if (j.frequency > R(0)) {
            const R Cdot2 = velB.ang - velA.ang;
            const R g = j.gamma * j.impulse.z;
            const R impulse2 = -j.mass.ez.z * (Cdot2 + j.bias + g);
            j.impulse.z += impulse2;
            velA.ang = velA.ang - iA * impulse2;
            velB.ang = velB.ang + iB * impulse2;
            const V2<R> vb = velB.lin + RevPerp(j.rB) * (velB.ang / R(1));
            const V2<R> va = velA.lin + RevPerp(j.rA) * (velA.ang / R(1));
            const V2<R> Cdot1 = vb - va;
            // Transform(Vec2, Mat33) (Math.hpp:522-527).
            const V2<R> t{j.mass.ex.x * Cdot1.x + j.mass.ey.x * Cdot1.y, j.mass.ex.y * Cdot1.x + j.mass.ey.y * Cdot1.y};
            const V2<R> impulse1 = -t;
            j.impulse.x += impulse1.x;
            j.impulse.y += impulse1.y;
            ApplyImpulse(A, B, impulse1, Cross(j.rA, impulse1) / R(1), Cross(j.rB, impulse1) / R(1), velA, velB);
        } else {
            const V2<R> vb = velB.lin + RevPerp(j.rB) * (velB.ang / R(1));
            const V2<R> va = velA.lin + RevPerp(j.rA) * (velA.ang / R(1));
            const V2<R> Cdot1 = vb - va;
            const R Cdot2 = (velB.ang - velA.ang) / R(1);
            const V3<R> t = MulRows(j.mass, V3<R>{Cdot1.x, Cdot1.y, Cdot2});
            const V3<R> impulse{-t.x, -t.y, -t.z};
            j.impulse.x += impulse.x;
            j.impulse.y += impulse.y;
            j.impulse.z += impulse.z;
            const V2<R> P{impulse.x, impulse.y};
            const R L = impulse.z;
            ApplyImpulse(A, B, P, L + Cross(j.rA, P) / R(1), L + Cross(j.rB, P) / R(1), velA, velB);
        }
        const bool changed = !(velA.lin == oldVelA.lin) || velA.ang != oldVelA.ang || !(velB.lin == oldVelB.lin) ||
                             velB.ang != oldVelB.ang;
        if (changed) {
            A.vel = velA;
            B.vel = velB;
            return false;
        }
        return true;
    }

    // SolvePosition(const WeldJointConf&) (WeldJointConf.cpp:305-375).
    static bool WeldSolvePosition(const Joint<R>& j, std::vector<BodyConstraint<R>>& bcs, const StepConf<R>& conf) {
        BodyConstraint<R>& A = bcs[j.bodyA];
        BodyConstraint<R>& B = bcs[j.bodyB];
        Pos<R> posA = A.pos, posB = B.pos;
        const R mA = A.invMass, iA = A.invRotI, mB = B.invMass, iB = B.invRotI;
        const V2<R> rA = Rotate(j.laA - A.lc, UnitFromAngle(posA.ang));
        const V2<R> rB = Rotate(j.laB - B.lc, UnitFromAngle(posB.ang));
        const M33<R> K = JointMass(mA, rA, iA, mB, rB, iB);
        const V2<R> C1 = (posB.lin + rB) - (posA.lin + rA);
        const R positionError = std::sqrt(MagSq(C1));
        R angularError = 0;
        V2<R> P;
        R L = 0;
        if (j.frequency > R(0)) {
            P = -Solve22(K, C1);
        } else {
            const R C2 = posB.ang - posA.ang - j.referenceAngle;
            angularError = std::abs(C2);
            V3<R> impulse;
            if (K.ez.z > R(0)) {
                const V3<R> s = Solve33(K, V3<R>{C1.x, C1.y, C2});
                impulse = V3<R>{-s.x, -s.y, -s.z};
            } else {
                const V2<R> s = Solve22(K, C1);
                impulse = V3<R>{-s.x, -s.y, R(0)};
            }
            P = V2<R>{impulse.x, impulse.y};
            L = impulse.z;
        }
        const R LA = (j.frequency > R(0)) ? Cross(rA, P) / R(1) : L + Cross(rA, P) / R(1);
        const R LB = (j.frequency > R(0)) ? Cross(rB, P) / R(1) : L + Cross(rB, P) / R(1);
        posA.lin = posA.lin - mA * P;
        posA.ang = posA.ang - iA * LA;
        posB.lin = posB.lin + mB * P;
        posB.ang = posB.ang + iB * LB;
        A.pos = posA;
        B.pos = posB;
        return (positionError <= conf.linearSlop) && (angularError <= conf.angularSlop);
    }

    // InitVelocity(RopeJointConf&) (RopeJointConf.cpp:59-128).
    static void RopeInitVelocity(Joint<R>& j, std::vector<BodyConstraint<R>>& bcs, const StepConf<R>& conf) {
        BodyConstraint<R>& A = bcs[j.bodyA];
        BodyConstraint<R>& B = bcs[j.bodyB];
        const R mA = A.invMass, iA = A.invRotI, mB = B.invMass, iB = B.invRotI;
        Vel<R> velA = A.vel, velB = B.vel;
        j.rA = Rotate(j.laA - A.lc, UnitFromAngle(A.pos.ang));
        j.rB = Rotate(j.laB - B.lc, UnitFromAngle(B.pos.ang));
        R length;
        const V2<R> uv = UnitPolar((B.pos.lin + j.rB) - (A.pos.lin + j.rA), length);
        j.curLength = length;
        const R C = length - j.length;
        j.limitState = (C > R(0)) ? kLimitUpper : kLimitInactive;
        if (length > conf.linearSlop) {
            j.u = uv;
        } else {
            j.u = V2<R>{R(0), R(0)};
            j.dmass = 0;
            j.impulse.x = 0;
            return;
        }
        const R crA = Cross(j.rA, j.u) / R(1);
        const R crB = Cross(j.rB, j.u) / R(1);
        const R invMass = mA + mB + iA * (crA * crA) + iB * (crB * crB);
        j.dmass = (invMass != R(0)) ? R(1) / invMass : R(0);
        if (conf.warmStart) {
            j.impulse.x *= conf.dtRatio;
            const V2<R> P = j.impulse.x * j.u;
            ApplyImpulse(A, B, P, Cross(j.rA, P) / R(1), Cross(j.rB, P) / R(1), velA, velB);
        } else {
            j.impulse.x = 0;
        }
        A.vel = velA;
        B.vel = velB;
    }

    // SolveVelocity(RopeJointConf&) (RopeJointConf.cpp:130-167).
    static bool RopeSolveVelocity(Joint<R>& j, std::vector<BodyConstraint<R>>& bcs, const StepConf<R>& conf) {
        BodyConstraint<R>& A = bcs[j.bodyA];
        BodyConstraint<R>& B = bcs[j.bodyB];
        Vel<R> velA = A.vel, velB = B.vel;
        const V2<R> vpA = velA.lin + RevPerp(j.rA) * (velA.ang / R(1));
        const V2<R> vpB = velB.lin + RevPerp(j.rB) * (velB.ang / R(1));
        const R C = j.curLength - j.length;
        const R invH = (conf.dt != R(0)) ? R(1) / conf.dt : R(0);
        const R Cdot = Dot(j.u, vpB - vpA) + ((C < R(0)) ? invH * C : R(0));
        R local = -j.dmass * Cdot;
        const R old = j.impulse.x;
        j.impulse.x = Min(R(0), j.impulse.x + local);
        local = j.impulse.x - old;
        const V2<R> P = local * j.u;
        ApplyImpulse(A, B, P, Cross(j.rA, P) / R(1), Cross(j.rB, P) / R(1), velA, velB);
        A.vel = velA;
        B.vel = velB;
        return local == R(0);
    }

    // SolvePosition(const RopeJointConf&) (RopeJointConf.cpp:169-212).
    static bool RopeSolvePosition(const Joint<R>& j, std::vector<BodyConstraint<R>>& bcs, const StepConf<R>& conf) {
        BodyConstraint<R>& A = bcs[j.bodyA];
        BodyConstraint<R>& B = bcs[j.bodyB];
        Pos<R> posA = A.pos, posB = B.pos;
        const V2<R> rA = Rotate(j.laA - A.lc, UnitFromAngle(posA.ang));
        const V2<R> rB = Rotate(j.laB - B.lc, UnitFromAngle(posB.ang));
        R length;
        const V2<R> uv = UnitPolar((posB.lin + rB) - (posA.lin + rA), length);
        const R C = Clamp(length - j.length, R(0), conf.maxLinearCorrection);
        const R local = -j.dmass * C;
        const V2<R> P = local * uv;
        posA.lin = posA.lin - A.invMass * P;
        posA.ang = posA.ang - A.invRotI * (Cross(rA, P) / R(1));
        posB.lin = posB.lin + B.invMass * P;
        posB.ang = posB.ang + B.invRotI * (Cross(rB, P) / R(1));
        A.pos = posA;
        B.pos = posB;
        return (length - j.length) < conf.linearSlop;
    }

    // InitVelocity(FrictionJointConf&) (FrictionJointConf.cpp:76-139).
    static void FrictionInitVelocity(Joint<R>& j, std::vector<BodyConstraint<R>>& bcs, const StepConf<R>& conf) {
        BodyConstraint<R>& A = bcs[j.bodyA];
        BodyConstraint<R>& B = bcs[j.bodyB];
        const R mA = A.invMass, iA = A.invRotI, mB = B.invMass, iB = B.invRotI;
        Vel<R> velA = A.vel, velB = B.vel;
        j.rA = Rotate(j.laA - A.lc, UnitFromAngle(A.pos.ang));
        j.rB = Rotate(j.laB - B.lc, UnitFromAngle(B.pos.ang));
        const V2<R> rA = j.rA, rB = j.rB;
        const R exx = mA + iA * (rA.y * rA.y) / R(1) + mB + iB * (rB.y * rB.y) / R(1);
        const R exy = -iA * rA.x * rA.y / R(1) + -iB * rB.x * rB.y / R(1);
        const R eyy = mA + iA * (rA.x * rA.x) / R(1) + mB + iB * (rB.x * rB.x) / R(1);
        // Invert(Matrix22) (Math.hpp:429-442) of columns (exx, exy), (exy, eyy).
        const R cp = Cross(V2<R>{exx, exy}, V2<R>{exy, eyy});
        if (!AlmostZero(cp)) {
            const R inverse = R(1) / cp;
            j.lm00 = eyy * inverse;
            j.lm01 = -exy * inverse;
            j.lm10 = -exy * inverse;
            j.lm11 = exx * inverse;
        } else {
            j.lm00 = j.lm01 = j.lm10 = j.lm11 = 0;
        }
        const R invRotI = iA + iB;
        j.angularMass = (invRotI > R(0)) ? R(1) / invRotI : R(0);
        if (conf.warmStart) {
            j.impulse.x *= conf.dtRatio;
            j.impulse.y *= conf.dtRatio;
            j.impulse.z *= conf.dtRatio;
            const V2<R> P{j.impulse.x, j.impulse.y};
            ApplyImpulse(A, B, P, Cross(rA, P) / R(1) + j.impulse.z, Cross(rB, P) / R(1) + j.impulse.z, velA, velB);
        } else {
            j.impulse = V3<R>{};
        }
        A.vel = velA;
        B.vel = velB;
    }

    // SolveVelocity(FrictionJointConf&) (FrictionJointConf.cpp:141-231).
    static bool FrictionSolveVelocity(Joint<R>& j, std::vector<BodyConstraint<R>>& bcs, const StepConf<R>& conf) {
        BodyConstraint<R>& A = bcs[j.bodyA];
        BodyConstraint<R>& B = bcs[j.bodyB];
        Vel<R> velA = A.vel, velB = B.vel;
        const R iA = A.invRotI, iB = B.invRotI;
        const R h = conf.dt;
        bool solved = true;
        {
            const R angularImpulse = -j.angularMass * (velB.ang - velA.ang);
            const R old = j.impulse.z;
            const R maxAngular = h * j.maxTorque;
            j.impulse.z = Clamp(j.impulse.z + angularImpulse, -maxAngular, maxAngular);
            const R inc = j.impulse.z - old;
            if (inc != R(0)) solved = false;
            velA.ang = velA.ang - iA * inc;
            velB.ang = velB.ang + iB * inc;
        }
        {
            const V2<R> vb = velB.lin + RevPerp(j.rB) * (velB.ang / R(1));
            const V2<R> va = velA.lin + RevPerp(j.rA) * (velA.ang / R(1));
            const V2<R> dv = vb - va;
            // Transform(v, Mat22): the matrix's vectors dotted with v from zero
            // (Vector.hpp:455-470).
            R tx = 0, ty = 0;
            tx += j.lm00 * dv.x;
            tx += j.lm01 * dv.y;
            ty += j.lm10 * dv.x;
            ty += j.lm11 * dv.y;
            const V2<R> old{j.impulse.x, j.impulse.y};
            V2<R> li{old.x + -tx, old.y + -ty};
            const R maxImpulse = h * j.maxForce;
            if (MagSq(li) > maxImpulse * maxImpulse) {
                const V2<R> dir = UnitVector(li, V2<R>{R(0), R(0)});
                li = V2<R>{dir.x * maxImpulse, dir.y * maxImpulse};
            }
            j.impulse.x = li.x;
            j.impulse.y = li.y;
            const V2<R> inc = li - old;
            if (inc.x != R(0) || inc.y != R(0)) solved = false;
            ApplyImpulse(A, B, inc, Cross(j.rA, inc) / R(1), Cross(j.rB, inc) / R(1), velA, velB);
        }
        A.vel = velA;
        B.vel = velB;
        return solved;
    }

    // InitVelocity(MotorJointConf&) (MotorJointConf.cpp:79-158).
    static void MotorInitVelocity(Joint<R>& j, std::vector<BodyConstraint<R>>& bcs, const StepConf<R>& conf) {
        BodyConstraint<R>& A = bcs[j.bodyA];
        BodyConstraint<R>& B = bcs[j.bodyB];
        const R mA = A.invMass, iA = A.invRotI, mB = B.invMass, iB = B.invRotI;
        Vel<R> velA = A.vel, velB = B.vel;
        j.rA = Rotate(j.laA - A.lc, UnitFromAngle(A.pos.ang));
        j.rB = Rotate(-B.lc, UnitFromAngle(B.pos.ang));
        const V2<R> rA = j.rA, rB = j.rB;
        const R exx = mA + mB + iA * (rA.y * rA.y) / R(1) + iB * (rB.y * rB.y) / R(1);
        const R exy = -iA * rA.x * rA.y / R(1) + -iB * rB.x * rB.y / R(1);
        const R eyy = mA + mB + iA * (rA.x * rA.x) / R(1) + iB * (rB.x * rB.x) / R(1);
        const R cp = Cross(V2<R>{exx, exy}, V2<R>{exy, eyy});
        if (!AlmostZero(cp)) {
            const R inverse = R(1) / cp;
            j.lm00 = eyy * inverse;
            j.lm01 = -exy * inverse;
            j.lm10 = -exy * inverse;
            j.lm11 = exx * inverse;
        } else {
            j.lm00 = j.lm01 = j.lm10 = j.lm11 = 0;
        }
        const R invRotI = iA + iB;
        j.angularMass = (invRotI > R(0)) ? R(1) / invRotI : R(0);
        j.linearError = (B.pos.lin + rB) - (A.pos.lin + rA);
        j.angularError = (B.pos.ang - A.pos.ang) - j.referenceAngle;
        if (conf.warmStart) {
            j.impulse.x *= conf.dtRatio;
            j.impulse.y *= conf.dtRatio;
            j.impulse.z *= conf.dtRatio;
            const V2<R> P{j.impulse.x, j.impulse.y};
            ApplyImpulse(A, B, P, Cross(rA, P) / R(1) + j.impulse.z, Cross(rB, P) / R(1) + j.impulse.z, velA, velB);
        } else {
            j.impulse = V3<R>{};
        }
        A.vel = velA;
        B.vel = velB;
    }

    // SolveVelocity(MotorJointConf&) (MotorJointConf.cpp:160-229).
    static bool MotorSolveVelocity(Joint<R>& j, std::vector<BodyConstraint<R>>& bcs, const StepConf<R>& conf) {
        BodyConstraint<R>& A = bcs[j.bodyA];
        BodyConstraint<R>& B = bcs[j.bodyB];
        Vel<R> velA = A.vel, velB = B.vel;
        const R mA = A.invMass, iA = A.invRotI, mB = B.invMass, iB = B.invRotI;
        const R h = conf.dt;
        const R invH = (h != R(0)) ? R(1) / h : R(0);
        bool solved = true;
        {
            const R Cdot = (velB.ang - velA.ang) + invH * j.correctionFactor * j.angularError;
            const R angularImpulse = -j.angularMass * Cdot;
            const R old = j.impulse.z;
            const R maxAngular = h * j.maxTorque;
            const R now = Clamp(old + angularImpulse, -maxAngular, maxAngular);
            j.impulse.z = now;
            const R inc = now - old;
            if (inc != R(0)) {
                // Velocity{LinearVelocity2{}, ...}: the linear parts take a zero.
                velA.lin = velA.lin - V2<R>{R(0), R(0)};
                velA.ang = velA.ang - iA * inc;
                velB.lin = velB.lin + V2<R>{R(0), R(0)};
                velB.ang = velB.ang + iB * inc;
                solved = false;
            }
        }
        {
            const V2<R> vb = velB.lin + RevPerp(j.rB) * (velB.ang / R(1));
            const V2<R> va = velA.lin + RevPerp(j.rA) * (velA.ang / R(1));
            const R k = invH * j.correctionFactor;
            const V2<R> Cdot = (vb - va) + V2<R>{k * j.linearError.x, k * j.linearError.y};
            R tx = 0, ty = 0;
            tx += j.lm00 * Cdot.x;
            tx += j.lm01 * Cdot.y;
            ty += j.lm10 * Cdot.x;
            ty += j.lm11 * Cdot.y;
            const V2<R> old{j.impulse.x, j.impulse.y};
            V2<R> li{old.x + -tx, old.y + -ty};
            const R maxImpulse = h * j.maxForce;
            if (MagSq(li) > maxImpulse * maxImpulse) {
                const V2<R> dir = UnitVector(li, V2<R>{R(0), R(0)});
                li = V2<R>{dir.x * maxImpulse, dir.y * maxImpulse};
            }
            j.impulse.x = li.x;
            j.impulse.y = li.y;
            const V2<R> inc = li - old;
            const R angA = Cross(j.rA, inc) / R(1);
            const R angB = Cross(j.rB, inc) / R(1);
            if (inc.x != R(0) || inc.y != R(0) || angA != R(0) || angB != R(0)) {
                velA.lin = velA.lin - mA * inc;
                velA.ang = velA.ang - iA * angA;
                velB.lin = velB.lin + mB * inc;
                velB.ang = velB.ang + iB * angB;
                solved = false;
            }
        }
        if (solved) return true;
        A.vel = velA;
        B.vel = velB;
        return false;
    }

    // InitVelocity(WheelJointConf&) (WheelJointConf.cpp:105-230).
    static void WheelInitVelocity(Joint<R>& j, std::vector<BodyConstraint<R>>& bcs, const StepConf<R>& conf) {
        BodyConstraint<R>& A = bcs[j.bodyA];
        BodyConstraint<R>& B = bcs[j.bodyB];
        const R mA = A.invMass, iA = A.invRotI, mB = B.invMass, iB = B.invRotI;
        Vel<R> velA = A.vel, velB = B.vel;
        const V2<R> qA = UnitFromAngle(A.pos.ang);
        const V2<R> rA = Rotate(j.laA - A.lc, qA);
        const V2<R> rB = Rotate(j.laB - B.lc, UnitFromAngle(B.pos.ang));
        const V2<R> dd = (B.pos.lin + rB) - (A.pos.lin + rA);
        {
            j.ay = Rotate(RevPerp(j.localXAxis), qA);
            j.sAy = Cross(dd + rA, j.ay);
            j.sBy = Cross(rB, j.ay);
            const R invMass = mA + mB + iA * (j.sAy * j.sAy) / R(1) + iB * (j.sBy * j.sBy) / R(1);
            j.dmass = (invMass > R(0)) ? R(1) / invMass : R(0);
        }
        j.springMass = 0;
        j.bias = 0;
        j.gamma = 0;
        if (j.frequency > R(0)) {
            j.ax = Rotate(j.localXAxis, qA);
            j.sAx = Cross(dd + rA, j.ax);
            j.sBx = Cross(rB, j.ax);
            const R invMass = mA + mB + iA * (j.sAx * j.sAx) / R(1) + iB * (j.sBx * j.sBx) / R(1);
            if (invMass > R(0)) {
                j.springMass = R(1) / invMass;
                const R C = Dot(dd, j.ax);
                const R omega = R(2) * kPi * j.frequency;
                const R d = R(2) * j.springMass * j.dampingRatio * omega;
                const R k = j.springMass * omega * omega;
                const R h = conf.dt;
                const R invGamma = h * (d + h * k);
                j.gamma = (invGamma > R(0)) ? R(1) / invGamma : R(0);
                j.bias = C * h * k * j.gamma;
                const R total = invMass + j.gamma;
                j.springMass = (total > R(0)) ? R(1) / total : R(0);
            }
        } else {
            j.impulse.y = 0;
            j.ax = V2<R>{R(0), R(0)};
            j.sAx = 0;
            j.sBx = 0;
        }
        if (j.enableMotor) {
            const R invRotI = iA + iB;
            j.angularMass = (invRotI > R(0)) ? R(1) / invRotI : R(0);
        } else {
            j.angularMass = 0;
            j.impulse.z = 0;
        }
        if (conf.warmStart) {
            j.impulse.x *= conf.dtRatio;
            j.impulse.y *= conf.dtRatio;
            j.impulse.z *= conf.dtRatio;
            const V2<R> P = j.impulse.x * j.ay + j.impulse.y * j.ax;
            const R LA = (j.impulse.x * j.sAy + j.impulse.y * j.sAx) / R(1) + j.impulse.z;
            const R LB = (j.impulse.x * j.sBy + j.impulse.y * j.sBx) / R(1) + j.impulse.z;
            ApplyImpulse(A, B, P, LA, LB, velA, velB);
        } else {
            j.impulse = V3<R>{};
        }
        A.vel = velA;
        B.vel = velB;
    }

    // SolveVelocity(WheelJointConf&) (WheelJointConf.cpp:232-307).
    static bool WheelSolveVelocity(Joint<R>& j, std::vector<BodyConstraint<R>>& bcs, const StepConf<R>& conf) {
        BodyConstraint<R>& A = bcs[j.bodyA];
        BodyConstraint<R>& B = bcs[j.bodyB];
        const Vel<R> oldVelA = A.vel, oldVelB = B.vel;
        Vel<R> velA = oldVelA, velB = oldVelB;
        const R iA = A.invRotI, iB = B.invRotI;
        {
            const R dot = Dot(j.ax, velB.lin - velA.lin);
            const R Cdot = dot + j.sBx * velB.ang / R(1) - j.sAx * velA.ang / R(1);
            const R impulse = -j.springMass * (Cdot + j.bias + j.gamma * j.impulse.y);
            j.impulse.y += impulse;
            const V2<R> P = impulse * j.ax;
            ApplyImpulse(A, B, P, impulse * j.sAx / R(1), impulse * j.sBx / R(1), velA, velB);
        }
        {
            const R Cdot = velB.ang - velA.ang - j.motorSpeed;
            R impulse = -j.angularMass * Cdot;
            const R old = j.impulse.z;
            const R maxImpulse = conf.dt * j.maxMotorTorque;
            j.impulse.z = Clamp(j.impulse.z + impulse, -maxImpulse, maxImpulse);
            impulse = j.impulse.z - old;
            velA.ang = velA.ang - iA * impulse;
            velB.ang = velB.ang + iB * impulse;
        }
        {
            const R dot = Dot(j.ay, velB.lin - velA.lin);
            const R Cdot = dot + j.sBy * velB.ang / R(1) - j.sAy * velA.ang / R(1);
            const R impulse = -j.dmass * Cdot;
            j.impulse.x += impulse;
            const V2<R> P = impulse * j.ay;
            ApplyImpulse(A, B, P, impulse * j.sAy / R(1), impulse * j.sBy / R(1), velA, velB);
        }
        const bool changed = !(velA.lin == oldVelA.lin) || velA.ang != oldVelA.ang || !(velB.lin == oldVelB.lin) ||
                             velB.ang != oldVelB.ang;
        if (changed) {
            A.vel = velA;
            B.vel = velB;
            return false;
        }
        return true;
    }

    // SolvePosition(const WheelJointConf&) (WheelJointConf.cpp:309-359).  The
    // effective mass uses the sAy, sBy found by InitVelocity, as there.
    static bool WheelSolvePosition(const Joint<R>& j, std::vector<BodyConstraint<R>>& bcs, const StepConf<R>& conf) {
        BodyConstraint<R>& A = bcs[j.bodyA];
        BodyConstraint<R>& B = bcs[j.bodyB];
        Pos<R> posA = A.pos, posB = B.pos;
        const R mA = A.invMass, iA = A.invRotI, mB = B.invMass, iB = B.invRotI;
        const V2<R> qA = UnitFromAngle(posA.ang);
        const V2<R> rA = Rotate(j.laA - A.lc, qA);
        const V2<R> rB = Rotate(j.laB - B.lc, UnitFromAngle(posB.ang));
        const V2<R> d = (posB.lin - posA.lin) + (rB - rA);
        const V2<R> ay = Rotate(RevPerp(j.localXAxis), qA);
        const R sAy = Cross(d + rA, ay);
        const R sBy = Cross(rB, ay);
        const R C = Dot(d, ay);
        const R k = mA + mB + iA * (j.sAy * j.sAy) / R(1) + iB * (j.sBy * j.sBy) / R(1);
        const R impulse = (k != R(0)) ? -(C / k) : R(0);
        const V2<R> P = impulse * ay;
        const R LA = impulse * sAy / R(1);
        const R LB = impulse * sBy / R(1);
        posA.lin = posA.lin - mA * P;
        posA.ang = posA.ang - iA * LA;
        posB.lin = posB.lin + mB * P;
        posB.ang = posB.ang + iB * LB;
        A.pos = posA;
        B.pos = posB;
        return std::abs(C) <= conf.linearSlop;
    }

    // InitVelocity(PulleyJointConf&) (PulleyJointConf.cpp:92-153).
    static void PulleyInitVelocity(Joint<R>& j, std::vector<BodyConstraint<R>>& bcs, const StepConf<R>& conf) {
        BodyConstraint<R>& A = bcs[j.bodyA];
        BodyConstraint<R>& B = bcs[j.bodyB];
        const R mA = A.invMass, iA = A.invRotI, mB = B.invMass, iB = B.invRotI;
        Vel<R> velA = A.vel, velB = B.vel;
        j.rA = Rotate(j.laA - A.lc, UnitFromAngle(A.pos.ang));
        j.rB = Rotate(j.laB - B.lc, UnitFromAngle(B.pos.ang));
        j.u = UnitVector((A.pos.lin + j.rA) - j.groundA, V2<R>{R(0), R(0)});
        j.ay = UnitVector((B.pos.lin + j.rB) - j.groundB, V2<R>{R(0), R(0)});
        const R ruA = Cross(j.rA, j.u);
        const R ruB = Cross(j.rB, j.ay);
        const R tA = mA + (iA * (ruA * ruA)) / R(1);
        const R tB = mB + (iB * (ruB * ruB)) / R(1);
        const R total = tA + j.ratio * j.ratio * tB;
        j.dmass = (total > R(0)) ? R(1) / total : R(0);
        if (conf.warmStart) {
            j.impulse.x *= conf.dtRatio;
            const V2<R> PA = (-j.impulse.x) * j.u;
            const V2<R> PB = (-j.ratio * j.impulse.x) * j.ay;
            velA.lin = velA.lin + mA * PA;
            velA.ang = velA.ang + iA * Cross(j.rA, PA) / R(1);
            velB.lin = velB.lin + mB * PB;
            velB.ang = velB.ang + iB * Cross(j.rB, PB) / R(1);
        } else {
            j.impulse.x = 0;
        }
        A.vel = velA;
        B.vel = velB;
    }

    // SolveVelocity(PulleyJointConf&) (PulleyJointConf.cpp:155-190).
    static bool PulleySolveVelocity(Joint<R>& j, std::vector<BodyConstraint<R>>& bcs) {
        BodyConstraint<R>& A = bcs[j.bodyA];
        BodyConstraint<R>& B = bcs[j.bodyB];
        Vel<R> velA = A.vel, velB = B.vel;
        const V2<R> vpA = velA.lin + RevPerp(j.rA) * (velA.ang / R(1));
        const V2<R> vpB = velB.lin + RevPerp(j.rB) * (velB.ang / R(1));
        const R Cdot = -Dot(j.u, vpA) - j.ratio * Dot(j.ay, vpB);
        const R impulse = -j.dmass * Cdot;
        j.impulse.x += impulse;
        const V2<R> PA = -impulse * j.u;
        const V2<R> PB = -j.ratio * impulse * j.ay;
        velA.lin = velA.lin + A.invMass * PA;
        velA.ang = velA.ang + A.invRotI * Cross(j.rA, PA) / R(1);
        velB.lin = velB.lin + B.invMass * PB;
        velB.ang = velB.ang + B.invRotI * Cross(j.rB, PB) / R(1);
        A.vel = velA;
        B.vel = velB;
        return impulse == R(0);
    }

    // SolvePosition(const PulleyJointConf&) (PulleyJointConf.cpp:192-253).
    static bool PulleySolvePosition(const Joint<R>& j, std::vector<BodyConstraint<R>>& bcs, const StepConf<R>& conf) {
        BodyConstraint<R>& A = bcs[j.bodyA];
        BodyConstraint<R>& B = bcs[j.bodyB];
        const R mA = A.invMass, iA = A.invRotI, mB = B.invMass, iB = B.invRotI;
        Pos<R> posA = A.pos, posB = B.pos;
        const V2<R> rA = Rotate(j.laA - A.lc, UnitFromAngle(posA.ang));
        const V2<R> rB = Rotate(j.laB - B.lc, UnitFromAngle(posB.ang));
        R lengthA, lengthB;
        const V2<R> uA = UnitPolar((posA.lin + rA) - j.groundA, lengthA);
        const V2<R> uB = UnitPolar((posB.lin + rB) - j.groundB, lengthB);
        const R ruA = Cross(rA, uA);
        const R ruB = Cross(rB, uB);
        const R tA = mA + iA * (ruA * ruA) / R(1);
        const R tB = mB + iB * (ruB * ruB) / R(1);
        const R total = tA + (j.ratio * j.ratio) * tB;
        const R mass = (total > R(0)) ? R(1) / total : R(0);
        const R src = j.length + j.ratio * j.lengthB;
        const R dst = lengthA + j.ratio * lengthB;
        const R C = src - dst;
        const R linearError = std::abs(C);
        const R impulse = -mass * C;
        const V2<R> PA = -impulse * uA;
        const V2<R> PB = -j.ratio * impulse * uB;
        posA.lin = posA.lin + mA * PA;
        posA.ang = posA.ang + iA * Cross(rA, PA) / R(1);
        posB.lin = posB.lin + mB * PB;
        posB.ang = posB.ang + iB * Cross(rB, PB) / R(1);
        A.pos = posA;
        B.pos = posB;
        return linearError < conf.linearSlop;
    }

    // InitVelocity(PrismaticJointConf&) (PrismaticJointConf.cpp:155-282).
    static void PrismaticInitVelocity(Joint<R>& j, std::vector<BodyConstraint<R>>& bcs, const StepConf<R>& conf) {
        BodyConstraint<R>& A = bcs[j.bodyA];
        BodyConstraint<R>& B = bcs[j.bodyB];
        const R mA = A.invMass, iA = A.invRotI, mB = B.invMass, iB = B.invRotI;
        Vel<R> velA = A.vel, velB = B.vel;
        const V2<R> qA = UnitFromAngle(A.pos.ang);
        const V2<R> qB = UnitFromAngle(B.pos.ang);
        const V2<R> rA = Rotate(j.laA - A.lc, qA);
        const V2<R> rB = Rotate(j.laB - B.lc, qB);
        const V2<R> d = ((B.pos.lin - A.pos.lin) + rB) - rA;
        j.ax = Rotate(j.localXAxis, qA);
        j.a1 = Cross(d + rA, j.ax);
        j.a2 = Cross(rB, j.ax);
        const R invRotMassA = iA * (j.a1 * j.a1) / R(1);
        const R invRotMassB = iB * (j.a2 * j.a2) / R(1);
        const R totalInvMass = mA + mB + invRotMassA + invRotMassB;
        j.motorMass = (totalInvMass > R(0)) ? R(1) / totalInvMass : R(0);
        {
            j.ay = Rotate(RevPerp(j.localXAxis), qA);
            j.s1 = Cross(d + rA, j.ay);
            j.s2 = Cross(rB, j.ay);
            const R invRotMassA2 = iA * (j.s1 * j.s1) / R(1);
            const R invRotMassB2 = iB * (j.s2 * j.s2) / R(1);
            const R k11 = mA + mB + invRotMassA2 + invRotMassB2;
            const R k12 = (iA * j.s1 + iB * j.s2);
            const R k13 = (iA * j.s1 * j.a1 + iB * j.s2 * j.a2) / R(1);
            const R totalInvRotI = iA + iB;
            const R k22 = (totalInvRotI == R(0)) ? R(1) : totalInvRotI;
            const R k23 = (iA * j.a1 + iB * j.a2);
            const R k33 = totalInvMass;
            j.mass = M33<R>{V3<R>{k11, k12, k13}, V3<R>{k12, k22, k23}, V3<R>{k13, k23, k33}};
        }
        if (j.enableLimit) {
            const R translation = Dot(j.ax, d);
            if (std::abs(j.upper - j.lower) < (conf.linearSlop * R(2))) {
                j.limitState = kLimitEqual;
            } else if (translation <= j.lower) {
                if (j.limitState != kLimitLower) {
                    j.limitState = kLimitLower;
                    j.impulse.z = 0;
                }
            } else if (translation >= j.upper) {
                if (j.limitState != kLimitUpper) {
                    j.limitState = kLimitUpper;
                    j.impulse.z = 0;
                }
            } else {
                j.limitState = kLimitInactive;
                j.impulse.z = 0;
            }
        } else {
            j.limitState = kLimitInactive;
            j.impulse.z = 0;
        }
        if (!j.enableMotor) j.motorImpulse = 0;
        if (conf.warmStart) {
            j.impulse.x *= conf.dtRatio;
            j.impulse.y *= conf.dtRatio;
            j.impulse.z *= conf.dtRatio;
            j.motorImpulse *= conf.dtRatio;
            const V2<R> Px = j.impulse.x * j.ay;
            const R Pxs1 = j.impulse.x * j.s1;
            const R Pxs2 = j.impulse.x * j.s2;
            const R PzLength = j.motorImpulse + j.impulse.z;
            const V2<R> Pz = PzLength * j.ax;
            const V2<R> P = Px + Pz;
            const R L = j.impulse.y;
            const R LA = L + (Pxs1 + PzLength * j.a1);
            const R LB = L + (Pxs2 + PzLength * j.a2);
            ApplyImpulse(A, B, P, LA, LB, velA, velB);
        } else {
            j.impulse = V3<R>{};
            j.motorImpulse = 0;
        }
        A.vel = velA;
        B.vel = velB;
    }

    // SolveVelocity(PrismaticJointConf&) (PrismaticJointConf.cpp:284-398).
    static bool PrismaticSolveVelocity(Joint<R>& j, std::vector<BodyConstraint<R>>& bcs, const StepConf<R>& conf) {
        BodyConstraint<R>& A = bcs[j.bodyA];
        BodyConstraint<R>& B = bcs[j.bodyB];
        const Vel<R> oldVelA = A.vel, oldVelB = B.vel;
        Vel<R> velA = oldVelA, velB = oldVelB;
        if (j.enableMotor && j.limitState != kLimitEqual) {
            const R vDot = Dot(j.ax, velB.lin - velA.lin);
            const R Cdot = vDot + (j.a2 * velB.ang - j.a1 * velA.ang);
            R impulse = j.motorMass * (j.motorSpeed - Cdot);
            const R oldImpulse = j.motorImpulse;
            const R maxImpulse = conf.dt * j.maxMotorTorque;
            j.motorImpulse = Clamp(j.motorImpulse + impulse, -maxImpulse, maxImpulse);
            impulse = j.motorImpulse - oldImpulse;
            const V2<R> P = impulse * j.ax;
            ApplyImpulse(A, B, P, impulse * j.a1, impulse * j.a2, velA, velB);
        }
        const V2<R> velDelta = velB.lin - velA.lin;
        const R sRotSpeed = (j.s2 * velB.ang - j.s1 * velA.ang);
        const V2<R> Cdot1{Dot(j.ay, velDelta) + sRotSpeed, velB.ang - velA.ang};
        if (j.enableLimit && j.limitState != kLimitInactive) {
            const R deltaDot = Dot(j.ax, velDelta);
            const R aRotSpeed = (j.a2 * velB.ang - j.a1 * velA.ang);
            const R Cdot2 = deltaDot + aRotSpeed;
            const V3<R> f1 = j.impulse;
            const V3<R> s = Solve33(j.mass, V3<R>{-Cdot1.x, -Cdot1.y, -Cdot2});
            j.impulse = V3<R>{j.impulse.x + s.x, j.impulse.y + s.y, j.impulse.z + s.z};
            if (j.limitState == kLimitLower) j.impulse.z = Max(j.impulse.z, R(0));
            else if (j.limitState == kLimitUpper) j.impulse.z = Min(j.impulse.z, R(0));
            const R dz = j.impulse.z - f1.z;
            const V2<R> b{-Cdot1.x - dz * j.mass.ez.x, -Cdot1.y - dz * j.mass.ez.y};
            const V2<R> f2r = Solve22(j.mass, b) + V2<R>{f1.x, f1.y};
            j.impulse.x = f2r.x;
            j.impulse.y = f2r.y;
            const V3<R> df{j.impulse.x - f1.x, j.impulse.y - f1.y, j.impulse.z - f1.z};
            const V2<R> P = df.x * j.ay + df.z * j.ax;
            const R LA = (df.x * j.s1 + df.y + df.z * j.a1);
            const R LB = (df.x * j.s2 + df.y + df.z * j.a2);
            ApplyImpulse(A, B, P, LA, LB, velA, velB);
        } else {
            const V2<R> df = Solve22(j.mass, V2<R>{-Cdot1.x, -Cdot1.y});
            j.impulse.x += df.x;
            j.impulse.y += df.y;
            const V2<R> P = df.x * j.ay;
            const R LA = (df.x * j.s1 + df.y);
            const R LB = (df.x * j.s2 + df.y);
            ApplyImpulse(A, B, P, LA, LB, velA, velB);
        }
        const bool changed = !(velA.lin == oldVelA.lin) || velA.ang != oldVelA.ang || !(velB.lin == oldVelB.lin) ||
                             velB.ang != oldVelB.ang;
        if (changed) {
            A.vel = velA;
            B.vel = velB;
            return false;
        }
        return true;
    }

    // SolvePosition(const PrismaticJointConf&) (PrismaticJointConf.cpp:409-531).
    static bool PrismaticSolvePosition(const Joint<R>& j, std::vector<BodyConstraint<R>>& bcs,
                                       const StepConf<R>& conf) {
        BodyConstraint<R>& A = bcs[j.bodyA];
        BodyConstraint<R>& B = bcs[j.bodyB];
        Pos<R> posA = A.pos, posB = B.pos;
        const R mA = A.invMass, iA = A.invRotI, mB = B.invMass, iB = B.invRotI;
        const V2<R> qA = UnitFromAngle(posA.ang);
        const V2<R> qB = UnitFromAngle(posB.ang);
        const V2<R> rA = Rotate(j.laA - A.lc, qA);
        const V2<R> rB = Rotate(j.laB - B.lc, qB);
        const V2<R> d = (posB.lin + rB) - (posA.lin + rA);
        const V2<R> axis = Rotate(j.localXAxis, qA);
        const R a1 = Cross(d + rA, axis);
        const R a2 = Cross(rB, axis);
        const V2<R> perp = Rotate(RevPerp(j.localXAxis), qA);
        const R s1 = Cross(d + rA, perp);
        const R s2 = Cross(rB, perp);
        const V2<R> C1{Dot(perp, d), posB.ang - posA.ang - j.referenceAngle};
        R linearError = std::abs(C1.x);
        const R angularError = std::abs(C1.y);
        bool active = false;
        R C2 = 0;
        if (j.enableLimit) {
            const R translation = Dot(axis, d);
            if (std::abs(j.upper - j.lower) < (R(2) * conf.linearSlop)) {
                C2 = Clamp(translation, -conf.maxLinearCorrection, conf.maxLinearCorrection);
                linearError = Max(linearError, std::abs(translation));
                active = true;
            } else if (translation <= j.lower) {
                C2 = Clamp(translation - j.lower + conf.linearSlop, -conf.maxLinearCorrection, R(0));
                linearError = Max(linearError, j.lower - translation);
                active = true;
            } else if (translation >= j.upper) {
                C2 = Clamp(translation - j.upper - conf.linearSlop, R(0), conf.maxLinearCorrection);
                linearError = Max(linearError, translation - j.upper);
                active = true;
            }
        }
        V3<R> impulse;
        if (active) {
            const R k11 = mA + iA * (s1 * s1) / R(1) + mB + iB * (s2 * s2) / R(1);
            const R k12 = iA * s1 + iB * s2;
            const R k13 = iA * s1 * a1 / R(1) + iB * s2 * a2 / R(1);
            R k22 = iA + iB;
            if (k22 == R(0)) k22 = R(1);
            const R k23 = iA * a1 + iB * a2;
            const R k33 = mA + iA * (a1 * a1) / R(1) + mB + iB * (a2 * a2) / R(1);
            const M33<R> K{V3<R>{k11, k12, k13}, V3<R>{k12, k22, k23}, V3<R>{k13, k23, k33}};
            impulse = Solve33(K, V3<R>{-C1.x, -C1.y, -C2});
        } else {
            const R k11 = mA + iA * s1 * s1 / R(1) + mB + iB * s2 * s2 / R(1);
            const R k12 = iA * s1 + iB * s2;
            R k22 = iA + iB;
            if (k22 == R(0)) k22 = R(1);
            // Solve(Mat22, Vec2) (Math.hpp:413-426) with the columns (k11,
            // k12) and (k12, k22).
            const R cp = k11 * k22 - k12 * k12;
            const V2<R> b{-C1.x, -C1.y};
            V2<R> sol{R(0), R(0)};
            if (!AlmostZero(cp)) {
                const R inverse = R(1) / cp;
                sol = V2<R>{(k22 * b.x - k12 * b.y) * inverse, (k11 * b.y - k12 * b.x) * inverse};
            }
            impulse = V3<R>{sol.x, sol.y, R(0)};
        }
        const V2<R> P = impulse.x * perp + impulse.z * axis;
        const R LA = (impulse.x * s1 + impulse.y + impulse.z * a1);
        const R LB = (impulse.x * s2 + impulse.y + impulse.z * a2);
        posA.lin = posA.lin - mA * P;
        posA.ang = posA.ang - iA * LA;
        posB.lin = posB.lin + mB * P;
        posB.ang = posB.ang + iB * LB;
        A.pos = posA;
        B.pos = posB;
        return (linearError <= conf.linearSlop) && (angularError <= conf.angularSlop);
    }

    // InitVelocity(TargetJointConf&) (TargetJointConf.cpp:85-141), with
    // GetEffectiveMassMatrix (TargetJointConf.cpp:60-83).
    static void TargetInitVelocity(Joint<R>& j, std::vector<BodyConstraint<R>>& bcs, const StepConf<R>& conf) {
        BodyConstraint<R>& B = bcs[j.bodyB];
        Vel<R> velB = B.vel;
        const V2<R> qB = UnitFromAngle(B.pos.ang);
        const R mass = (B.invMass != R(0)) ? (R(1) / B.invMass) : R(0);
        const R omega = R(2) * kPi * j.frequency;
        const R d = R(2) * mass * j.dampingRatio * omega;
        const R k = mass * (omega * omega);
        const R h = conf.dt;
        const R tmp = d + h * k;
        const R invGamma = h * tmp;
        j.gamma = (invGamma != R(0)) ? R(1) / invGamma : R(0);
        const R beta = h * k * j.gamma;
        j.rB = Rotate(j.laB - B.lc, qB);
        {
            const R invMass = B.invMass, invRotI = B.invRotI;
            const R exx = invMass + (invRotI * (j.rB.y * j.rB.y) / R(1)) + j.gamma;
            const R exy = -invRotI * j.rB.x * j.rB.y / R(1);
            const R eyy = invMass + (invRotI * (j.rB.x * j.rB.x) / R(1)) + j.gamma;
            // Invert(Matrix22) (Math.hpp:429-442) of columns (exx, exy), (exy, eyy).
            const R cp = Cross(V2<R>{exx, exy}, V2<R>{exy, eyy});
            if (!AlmostZero(cp)) {
                const R inverse = R(1) / cp;
                j.lm00 = eyy * inverse;
                j.lm01 = -exy * inverse;
                j.lm10 = -exy * inverse;
                j.lm11 = exx * inverse;
            } else {
                j.lm00 = j.lm01 = j.lm10 = j.lm11 = 0;
            }
        }
        j.linearError = ((B.pos.lin + j.rB) - j.groundA) * beta;
        velB.ang *= R(0.98f);
        if (conf.warmStart) {
            j.impulse.x *= conf.dtRatio;
            j.impulse.y *= conf.dtRatio;
            const V2<R> P{j.impulse.x, j.impulse.y};
            const R crossBP = Cross(j.rB, P) / R(1);
            velB.lin = velB.lin + B.invMass * P;
            velB.ang = velB.ang + B.invRotI * crossBP;
        } else {
            j.impulse = V3<R>{};
        }
        B.vel = velB;
    }

    // SolveVelocity(TargetJointConf&) (TargetJointConf.cpp:143-178).
    static bool TargetSolveVelocity(Joint<R>& j, std::vector<BodyConstraint<R>>& bcs, const StepConf<R>& conf) {
        BodyConstraint<R>& B = bcs[j.bodyB];
        Vel<R> velB = B.vel;
        const V2<R> Cdot = velB.lin + RevPerp(j.rB) * (velB.ang / R(1));
        const V2<R> ev = Cdot + (j.linearError + j.gamma * V2<R>{j.impulse.x, j.impulse.y});
        const V2<R> old{j.impulse.x, j.impulse.y};
        // Transform(-ev, mass): the matrix's vectors dotted with -ev from zero
        // (Vector.hpp:455-470).
        const V2<R> nev = -ev;
        R tx = 0, ty = 0;
        tx += j.lm00 * nev.x;
        tx += j.lm01 * nev.y;
        ty += j.lm10 * nev.x;
        ty += j.lm11 * nev.y;
        V2<R> imp{old.x + tx, old.y + ty};
        const R maxImpulse = conf.dt * j.maxForce;
        if (MagSq(imp) > maxImpulse * maxImpulse) {
            const V2<R> dir = UnitVector(imp, V2<R>{R(0), R(0)});
            imp = V2<R>{dir.x * maxImpulse, dir.y * maxImpulse};
        }
        j.impulse.x = imp.x;
        j.impulse.y = imp.y;
        const V2<R> inc = imp - old;
        const R angImpulseB = Cross(j.rB, inc) / R(1);
        velB.lin = velB.lin + B.invMass * inc;
        velB.ang = velB.ang + B.invRotI * angImpulseB;
        B.vel = velB;
        return inc.x == R(0) && inc.y == R(0);
    }

    // InitVelocity(GearJointConf&) (GearJointConf.cpp:171-255): per side the
    // Jacobian of its coordinate (a prismatic side the axis and the arms'
    // crosses, a revolute one the unit angular terms, side B-D scaled by the
    // ratio) and the summed inverse mass, then the warm start on all four.
    static void GearInitVelocity(Joint<R>& j, std::vector<BodyConstraint<R>>& bcs, const StepConf<R>& conf) {
        BodyConstraint<R>& A = bcs[j.bodyA];
        BodyConstraint<R>& B = bcs[j.bodyB];
        BodyConstraint<R>& C = bcs[j.bodyC];
        BodyConstraint<R>& D = bcs[j.bodyD];
        Vel<R> velA = A.vel, velB = B.vel, velC = C.vel, velD = D.vel;
        R invMass = 0;
        if (j.gearPris[0]) {
            const V2<R> qA = UnitFromAngle(A.pos.ang);
            const V2<R> qC = UnitFromAngle(C.pos.ang);
            const V2<R> u = Rotate(j.gearAxis[0], qC);
            const V2<R> rC = Rotate(j.gearLaA[0] - C.lc, qC);
            const V2<R> rA = Rotate(j.gearLaB[0] - A.lc, qA);
            j.JvAC = V2<R>{R(1) * u.x, R(1) * u.y};
            j.JwC = Cross(rC, u);
            j.JwA = Cross(rA, u);
            const R invRotMassC = C.invRotI * (j.JwC * j.JwC) / R(1);
            const R invRotMassA = A.invRotI * (j.JwA * j.JwA) / R(1);
            invMass += C.invMass + A.invMass + invRotMassC + invRotMassA;
        } else {
            j.JvAC = V2<R>{R(0), R(0)};
            j.JwA = R(1);
            j.JwC = R(1);
            invMass += A.invRotI + C.invRotI;
        }
        if (j.gearPris[1]) {
            const V2<R> qB = UnitFromAngle(B.pos.ang);
            const V2<R> qD = UnitFromAngle(D.pos.ang);
            const V2<R> u = Rotate(j.gearAxis[1], qD);
            const V2<R> rD = Rotate(j.gearLaA[1] - D.lc, qD);
            const V2<R> rB = Rotate(j.gearLaB[1] - B.lc, qB);
            j.JvBD = V2<R>{j.ratio * u.x, j.ratio * u.y};
            j.JwD = j.ratio * Cross(rD, u);
            j.JwB = j.ratio * Cross(rB, u);
            const R invRotMassD = D.invRotI * (j.JwD * j.JwD) / R(1);
            const R invRotMassB = B.invRotI * (j.JwB * j.JwB) / R(1);
            invMass += (j.ratio * j.ratio) * (D.invMass + B.invMass) + invRotMassD + invRotMassB;
        } else {
            j.JvBD = V2<R>{R(0), R(0)};
            j.JwB = j.ratio * R(1);
            j.JwD = j.ratio * R(1);
            invMass += (j.ratio * j.ratio) * (B.invRotI + D.invRotI);
        }
        j.dmass = (invMass > R(0)) ? R(1) / invMass : R(0);
        if (conf.warmStart) {
            const R imp = j.impulse.x;
            velA.lin = velA.lin + (A.invMass * imp) * j.JvAC;
            velA.ang = velA.ang + A.invRotI * imp * j.JwA / R(1);
            velB.lin = velB.lin + (B.invMass * imp) * j.JvBD;
            velB.ang = velB.ang + B.invRotI * imp * j.JwB / R(1);
            velC.lin = velC.lin - (C.invMass * imp) * j.JvAC;
            velC.ang = velC.ang - C.invRotI * imp * j.JwC / R(1);
            velD.lin = velD.lin - (D.invMass * imp) * j.JvBD;
            velD.ang = velD.ang - D.invRotI * imp * j.JwD / R(1);
        } else {
            j.impulse.x = R(0);
        }
        A.vel = velA;
        B.vel = velB;
        C.vel = velC;
        D.vel = velD;
    }

    // SolveVelocity(GearJointConf&) (GearJointConf.cpp:257-296).
    static bool GearSolveVelocity(Joint<R>& j, std::vector<BodyConstraint<R>>& bcs) {
        BodyConstraint<R>& A = bcs[j.bodyA];
        BodyConstraint<R>& B = bcs[j.bodyB];
        BodyConstraint<R>& C = bcs[j.bodyC];
        BodyConstraint<R>& D = bcs[j.bodyD];
        Vel<R> velA = A.vel, velB = B.vel, velC = C.vel, velD = D.vel;
        const R acDot = Dot(j.JvAC, velA.lin - velC.lin);
        const R bdDot = Dot(j.JvBD, velB.lin - velD.lin);
        const R Cdot = acDot + bdDot + (j.JwA * velA.ang - j.JwC * velC.ang) / R(1) +
                       (j.JwB * velB.ang - j.JwD * velD.ang) / R(1);
        const R impulse = -j.dmass * R(1) * Cdot;
        j.impulse.x += impulse;
        velA.lin = velA.lin + (A.invMass * impulse) * j.JvAC;
        velA.ang = velA.ang + A.invRotI * impulse * j.JwA / R(1);
        velB.lin = velB.lin + (B.invMass * impulse) * j.JvBD;
        velB.ang = velB.ang + B.invRotI * impulse * j.JwB / R(1);
        velC.lin = velC.lin - (C.invMass * impulse) * j.JvAC;
        velC.ang = velC.ang - C.invRotI * impulse * j.JwC / R(1);
        velD.lin = velD.lin - (D.invMass * impulse) * j.JvBD;
        velD.ang = velD.ang - D.invRotI * impulse * j.JwD / R(1);
        A.vel = velA;
        B.vel = velB;
        C.vel = velC;
        D.vel = velD;
        return impulse == R(0);
    }

    // Cap(Position, ConstraintSolverConf) (d2/Math.cpp:63-71).
    static Pos<R> CapPosition(Pos<R> p, const StepConf<R>& conf) {
        const R lsquared = MagSq(p.lin);
        if (lsquared > conf.maxLinearCorrection * conf.maxLinearCorrection) {
            const R s = conf.maxLinearCorrection / std::sqrt(lsquared);
            p.lin = V2<R>{p.lin.x * s, p.lin.y * s};
        }
        p.ang = Clamp(p.ang, -conf.maxAngularCorrection, +conf.maxAngularCorrection);
        return p;
    }

    // SolvePosition(const GearJointConf&) (GearJointConf.cpp:298-471): the
    // coordinates from the positions now (a revolute side the angle less its
    // reference angle, a prismatic one the anchor offset along the axis), one
    // capped correction of the position error on all four bodies.
    static bool GearSolvePosition(const Joint<R>& j, std::vector<BodyConstraint<R>>& bcs, const StepConf<R>& conf) {
        BodyConstraint<R>& A = bcs[j.bodyA];
        BodyConstraint<R>& B = bcs[j.bodyB];
        BodyConstraint<R>& C = bcs[j.bodyC];
        BodyConstraint<R>& D = bcs[j.bodyD];
        const Pos<R> posA = A.pos, posB = B.pos, posC = C.pos, posD = D.pos;
        if (!std::isfinite(j.constant)) return true;
        V2<R> JvAC{R(0), R(0)}, JvBD{R(0), R(0)};
        R JwA = 0, JwB = 0, JwC = 0, JwD = 0;
        R coordinateA = 0, coordinateB = 0, invMass = 0;
        if (j.gearPris[0]) {
            const V2<R> qA = UnitFromAngle(posA.ang);
            const V2<R> qC = UnitFromAngle(posC.ang);
            const V2<R> u = Rotate(j.gearAxis[0], qC);
            const V2<R> rC = Rotate(j.gearLaA[0] - C.lc, qC);
            const V2<R> rA = Rotate(j.gearLaB[0] - A.lc, qA);
            JvAC = V2<R>{u.x * R(1), u.y * R(1)};
            JwC = Cross(rC, u);
            JwA = Cross(rA, u);
            const R invLinMass = C.invMass + A.invMass;
            const R invRotMassC = C.invRotI * ((JwC * R(1) / R(1)) * (JwC * R(1) / R(1)));
            const R invRotMassA = A.invRotI * ((JwA * R(1) / R(1)) * (JwA * R(1) / R(1)));
            invMass += invLinMass + invRotMassC + invRotMassA;
            const V2<R> pC = j.gearLaA[0] - C.lc;
            const V2<R> pA = InvRotate(rA + (posA.lin - posC.lin), qC);
            coordinateA = Dot(pA - pC, j.gearAxis[0]) / R(1);
        } else {
            JwA = R(1);
            JwC = R(1);
            invMass += A.invRotI + C.invRotI;
            coordinateA = ((posA.ang - posC.ang) - j.gearRef[0]) / R(1);
        }
        if (j.gearPris[1]) {
            const V2<R> qB = UnitFromAngle(posB.ang);
            const V2<R> qD = UnitFromAngle(posD.ang);
            const V2<R> u = Rotate(j.gearAxis[1], qD);
            const V2<R> rD = Rotate(j.gearLaA[1] - D.lc, qD);
            const V2<R> rB = Rotate(j.gearLaB[1] - B.lc, qB);
            JvBD = V2<R>{j.ratio * u.x, j.ratio * u.y};
            JwD = j.ratio * Cross(rD, u);
            JwB = j.ratio * Cross(rB, u);
            const R invLinMass = (j.ratio * j.ratio) * (D.invMass + B.invMass);
            const R invRotMassD = D.invRotI * ((JwD * R(1) / R(1)) * (JwD * R(1) / R(1)));
            const R invRotMassB = B.invRotI * ((JwB * R(1) / R(1)) * (JwB * R(1) / R(1)));
            invMass += invLinMass + invRotMassD + invRotMassB;
            const V2<R> pD = j.gearLaA[1] - D.lc;
            const V2<R> pB = InvRotate(rB + (posB.lin - posD.lin), qD);
            coordinateB = Dot(pB - pD, j.gearAxis[1]) / R(1);
        } else {
            JwB = j.ratio;
            JwD = j.ratio;
            invMass += (j.ratio * j.ratio) * (B.invRotI + D.invRotI);
            coordinateB = ((posB.ang - posD.ang) - j.gearRef[1]) / R(1);
        }
        const R Cv = ((coordinateA + j.ratio * coordinateB) - j.constant) * conf.regResolutionRate;
        const R impulse = ((invMass > R(0)) ? -Cv / invMass : R(0)) * R(1) * R(1);
        auto delta = [&](const BodyConstraint<R>& b, V2<R> Jv, R Jw) {
            const R lm = b.invMass * impulse;
            return CapPosition(Pos<R>{V2<R>{lm * Jv.x, lm * Jv.y}, b.invRotI * impulse * Jw * R(1) / R(1)}, conf);
        };
        const Pos<R> dA = delta(A, JvAC, JwA);
        const Pos<R> dC = delta(C, JvAC, JwC);
        const Pos<R> dB = delta(B, JvBD, JwB);
        const Pos<R> dD = delta(D, JvBD, JwD);
        A.pos = Pos<R>{posA.lin + dA.lin, posA.ang + dA.ang};
        B.pos = Pos<R>{posB.lin + dB.lin, posB.ang + dB.ang};
        C.pos = Pos<R>{posC.lin - dC.lin, posC.ang - dC.ang};
        D.pos = Pos<R>{posD.lin - dD.lin, posD.ang - dD.ang};
        R linearError = 0, angularError = 0;
        linearError = std::max(linearError, MagSq(dA.lin));
        linearError = std::max(linearError, MagSq(dB.lin));
        linearError = std::max(linearError, MagSq(dC.lin));
        linearError = std::max(linearError, MagSq(dD.lin));
        angularError = std::max(angularError, std::abs(dA.ang));
        angularError = std::max(angularError, std::abs(dB.ang));
        angularError = std::max(angularError, std::abs(dC.ang));
        angularError = std::max(angularError, std::abs(dD.ang));
        return (linearError <= conf.linearSlop * conf.linearSlop) && (angularError <= conf.angularSlop);
    }

    // InitVelocity(RevoluteJointConf&) (RevoluteJointConf.cpp:127-226).
    static void JointInitVelocity(Joint<R>& j, std::vector<BodyConstraint<R>>& bcs, const StepConf<R>& conf) {
        if (j.kind == 1) return DistanceInitVelocity(j, bcs, conf);
        if (j.kind == 2) return WeldInitVelocity(j, bcs, conf);
        if (j.kind == 3) return RopeInitVelocity(j, bcs, conf);
        if (j.kind == 4) return FrictionInitVelocity(j, bcs, conf);
        if (j.kind == 5) return MotorInitVelocity(j, bcs, conf);
        if (j.kind == 6) return WheelInitVelocity(j, bcs, conf);
        if (j.kind == 7) return PulleyInitVelocity(j, bcs, conf);
        if (j.kind == 8) return PrismaticInitVelocity(j, bcs, conf);
        if (j.kind == 9) return TargetInitVelocity(j, bcs, conf);
        if (j.kind == 10) return GearInitVelocity(j, bcs, conf);
        BodyConstraint<R>& A = bcs[j.bodyA];
        BodyConstraint<R>& B = bcs[j.bodyB];
        const R mA = A.invMass, iA = A.invRotI, aA = A.pos.ang;
        const R mB = B.invMass, iB = B.invRotI, aB = B.pos.ang;
        Vel<R> velA = A.vel, velB = B.vel;
        const V2<R> qA = UnitFromAngle(aA);
        const V2<R> qB = UnitFromAngle(aB);
        j.rA = Rotate(j.laA - A.lc, qA);
        j.rB = Rotate(j.laB - B.lc, qB);
        const R totInvI = iA + iB;
        const bool fixedRotation = (totInvI == R(0));
        j.mass = JointMass(mA, j.rA, iA, mB, j.rB, iB);
        j.angularMass = (totInvI > R(0)) ? R(R(1) / totInvI) : R(0);
        if (!j.enableMotor || fixedRotation) j.motorImpulse = 0;
        if (j.enableLimit && !fixedRotation) {
            const R jointAngle = aB - aA - j.referenceAngle;
            if (std::abs(j.upper - j.lower) < (conf.angularSlop * R(2))) {
                j.limitState = kLimitEqual;
            } else if (jointAngle <= j.lower) {
                if (j.limitState != kLimitLower) {
                    j.limitState = kLimitLower;
                    j.impulse.z = 0;
                }
            } else if (jointAngle >= j.upper) {
                if (j.limitState != kLimitUpper) {
                    j.limitState = kLimitUpper;
                    j.impulse.z = 0;
                }
            } else {
                j.limitState = kLimitInactive;
                j.impulse.z = 0;
            }
        } else {
            j.limitState = kLimitInactive;
        }
        if (conf.warmStart) {
            j.impulse.x *= conf.dtRatio;
            j.impulse.y *= conf.dtRatio;
            j.impulse.z *= conf.dtRatio;
            j.motorImpulse *= conf.dtRatio;
            const V2<R> P{j.impulse.x, j.impulse.y};
            const R L = j.motorImpulse + (j.impulse.z * R(1) * R(1) / (R(1) * R(1)));
            const R LA = Cross(j.rA, P) / R(1) + L;
            const R LB = Cross(j.rB, P) / R(1) + L;
            velA.lin = velA.lin - mA * P;
            velA.ang = velA.ang - iA * LA;
            velB.lin = velB.lin + mB * P;
            velB.ang = velB.ang + iB * LB;
        } else {
            j.impulse = V3<R>{};
            j.motorImpulse = 0;
        }
        A.vel = velA;
        B.vel = velB;
    }

    // SolveVelocity(RevoluteJointConf&) (RevoluteJointConf.cpp:228-331).
    static bool JointSolveVelocity(Joint<R>& j, std::vector<BodyConstraint<R>>& bcs, const StepConf<R>& conf) {
        if (j.kind == 1) return DistanceSolveVelocity(j, bcs);
        if (j.kind == 2) return WeldSolveVelocity(j, bcs);
        if (j.kind == 3) return RopeSolveVelocity(j, bcs, conf);
        if (j.kind == 4) return FrictionSolveVelocity(j, bcs, conf);
        if (j.kind == 5) return MotorSolveVelocity(j, bcs, conf);
        if (j.kind == 6) return WheelSolveVelocity(j, bcs, conf);
        if (j.kind == 7) return PulleySolveVelocity(j, bcs);
        if (j.kind == 8) return PrismaticSolveVelocity(j, bcs, conf);
        if (j.kind == 9) return TargetSolveVelocity(j, bcs, conf);
        if (j.kind == 10) return GearSolveVelocity(j, bcs);
        BodyConstraint<R>& A = bcs[j.bodyA];
        BodyConstraint<R>& B = bcs[j.bodyB];
        const Vel<R> oldVelA = A.vel, oldVelB = B.vel;
        Vel<R> velA = oldVelA, velB = oldVelB;
        const R mA = A.invMass, iA = A.invRotI, mB = B.invMass, iB = B.invRotI;
        const bool fixedRotation = (iA + iB == R(0));
        if (j.enableMotor && j.limitState != kLimitEqual && !fixedRotation) {
            const R impulse = -j.angularMass * (velB.ang - velA.ang - j.motorSpeed);
            const R oldImpulse = j.motorImpulse;
            const R maxImpulse = conf.dt * j.maxMotorTorque;
            j.motorImpulse = Clamp(j.motorImpulse + impulse, -maxImpulse, maxImpulse);
            const R inc = j.motorImpulse - oldImpulse;
            velA.ang = velA.ang - iA * inc;
            velB.ang = velB.ang + iB * inc;
        }
        const V2<R> vb = velB.lin + RevPerp(j.rB) * (velB.ang / R(1));
        const V2<R> va = velA.lin + RevPerp(j.rA) * (velA.ang / R(1));
        const V2<R> vDelta = vb - va;
        V3<R> impulse;
        if (j.enableLimit && j.limitState != kLimitInactive && !fixedRotation) {
            const V3<R> cdot{vDelta.x / R(1), vDelta.y / R(1), (velB.ang - velA.ang) / R(1)};
            const V3<R> s = Solve33(j.mass, cdot);
            impulse = V3<R>{-s.x, -s.y, -s.z};
            auto update = [&](V3<R>& imp) {
                const V2<R> rhs = -V2<R>{vDelta.x / R(1), vDelta.y / R(1)} + j.impulse.z * V2<R>{j.mass.ez.x, j.mass.ez.y};
                const V2<R> reduced = Solve22(j.mass, rhs);
                imp = V3<R>{reduced.x, reduced.y, -j.impulse.z};
                return V3<R>{j.impulse.x + reduced.x, j.impulse.y + reduced.y, R(0)};
            };
            auto add = [](V3<R> a, V3<R> b) { return V3<R>{a.x + b.x, a.y + b.y, a.z + b.z}; };
            if (j.limitState == kLimitLower) {
                const R newZ = j.impulse.z + impulse.z;
                j.impulse = (newZ < R(0)) ? update(impulse) : add(j.impulse, impulse);
            } else if (j.limitState == kLimitUpper) {
                const R newZ = j.impulse.z + impulse.z;
                j.impulse = (newZ > R(0)) ? update(impulse) : add(j.impulse, impulse);
            } else {
                j.impulse = add(j.impulse, impulse);
            }
            const V2<R> P{impulse.x, impulse.y};
            const R L = impulse.z * R(1) * R(1) / (R(1) * R(1));
            const R LA = Cross(j.rA, P) / R(1) + L;
            const R LB = Cross(j.rB, P) / R(1) + L;
            velA.lin = velA.lin - mA * P;
            velA.ang = velA.ang - iA * LA;
            velB.lin = velB.lin + mB * P;
            velB.ang = velB.ang + iB * LB;
        } else {
            const V2<R> imp = Solve22(j.mass, -V2<R>{vDelta.x / R(1), vDelta.y / R(1)});
            j.impulse.x += imp.x;
            j.impulse.y += imp.y;
            const V2<R> P{imp.x, imp.y};
            const R LA = Cross(j.rA, P) / R(1);
            const R LB = Cross(j.rB, P) / R(1);
            velA.lin = velA.lin - mA * P;
            velA.ang = velA.ang - iA * LA;
            velB.lin = velB.lin + mB * P;
            velB.ang = velB.ang + iB * LB;
        }
        const bool changed = !(velA.lin == oldVelA.lin) || velA.ang != oldVelA.ang || !(velB.lin == oldVelB.lin) ||
                             velB.ang != oldVelB.ang;
        if (changed) {
            A.vel = velA;
            B.vel = velB;
            return false;
        }
        return true;
    }

    // SolvePosition(const RevoluteJointConf&) (RevoluteJointConf.cpp:333-433).
    static bool JointSolvePosition(const Joint<R>& j, std::vector<BodyConstraint<R>>& bcs, const StepConf<R>& conf) {
        if (j.kind == 1) return DistanceSolvePosition(j, bcs, conf);
        if (j.kind == 2) return WeldSolvePosition(j, bcs, conf);
        if (j.kind == 3) return RopeSolvePosition(j, bcs, conf);
        if (j.kind == 4) return true;  // FrictionJointConf.cpp:233-237
        if (j.kind == 5) return true;  // MotorJointConf.cpp:231-234
        if (j.kind == 6) return WheelSolvePosition(j, bcs, conf);
        if (j.kind == 7) return PulleySolvePosition(j, bcs, conf);
        if (j.kind == 8) return PrismaticSolvePosition(j, bcs, conf);
        if (j.kind == 9) return true;  // TargetJointConf.cpp:180-184
        if (j.kind == 10) return GearSolvePosition(j, bcs, conf);
        BodyConstraint<R>& A = bcs[j.bodyA];
        BodyConstraint<R>& B = bcs[j.bodyB];
        Pos<R> posA = A.pos, posB = B.pos;
        const R iA = A.invRotI, iB = B.invRotI;
        const bool fixedRotation = ((iA + iB) == R(0));
        R angularError = 0;
        if (j.enableLimit && j.limitState != kLimitInactive && !fixedRotation) {
            const R angle = posB.ang - posA.ang - j.referenceAngle;
            R limitImpulse = R(0) * R(1) * R(1) / R(1);
            if (j.limitState == kLimitLower) {
                R C = angle - j.lower;
                angularError = -C;
                C = Clamp(C + conf.angularSlop, -conf.maxAngularCorrection, R(0));
                limitImpulse = -j.angularMass * C;
            } else if (j.limitState == kLimitUpper) {
                R C = angle - j.upper;
                angularError = C;
                C = Clamp(C - conf.angularSlop, R(0), conf.maxAngularCorrection);
                limitImpulse = -j.angularMass * C;
            } else {
                const R C = Clamp(angle - j.lower, -conf.maxAngularCorrection, conf.maxAngularCorrection);
                limitImpulse = -j.angularMass * C;
                angularError = std::abs(C);
            }
            posA.ang = posA.ang - iA * limitImpulse;
            posB.ang = posB.ang + iB * limitImpulse;
        }
        R positionError;
        {
            const V2<R> qA = UnitFromAngle(posA.ang);
            const V2<R> qB = UnitFromAngle(posB.ang);
            const V2<R> rA = Rotate(j.laA - A.lc, qA);
            const V2<R> rB = Rotate(j.laB - B.lc, qB);
            const V2<R> C = (posB.lin + rB) - (posA.lin + rA);
            positionError = MagSq(C);
            const R mA = A.invMass, mB = B.invMass;
            const R exx = mA + (iA * (rA.y * rA.y) / R(1)) + mB + (iB * (rB.y * rB.y) / R(1));
            const R exy = (-iA * rA.x * rA.y / R(1)) + (-iB * rB.x * rB.y / R(1));
            const R eyy = mA + (iA * (rA.x * rA.x) / R(1)) + mB + (iB * (rB.x * rB.x) / R(1));
            // Solve(Matrix22, Vector2) (Math.hpp:413-426): columns (exx, exy), (exy, eyy).
            V2<R> sol{};
            const R cp = Cross(V2<R>{exx, exy}, V2<R>{exy, eyy});
            if (!AlmostZero(cp)) {
                const R inverse = R(1) / cp;
                sol = V2<R>{(eyy * C.x - exy * C.y) * inverse, (exx * C.y - exy * C.x) * inverse};
            }
            const V2<R> P = -sol;
            posA.lin = posA.lin - mA * P;
            posA.ang = posA.ang - iA * Cross(rA, P) / R(1);
            posB.lin = posB.lin + mB * P;
            posB.ang = posB.ang + iB * Cross(rB, P) / R(1);
        }
        A.pos = posA;
        B.pos = posB;
        return (positionError <= conf.linearSlop * conf.linearSlop) && (angularError <= conf.angularSlop);
    }

    // AddToIsland (AabbTreeWorld.cpp:1358-1444)
    void AddToIsland(Island& island, std::uint32_t seed) {
        std::vector<std::uint32_t> stack;
        stack.push_back(seed);
        islandedBodies[seed] = 1;
        while (!stack.empty()) {
            const std::uint32_t bid = stack.back();
            stack.pop_back();
            Body<R>& b = bodies[bid];
            island.bodies.push_back(bid);
            if (!b.speedable()) continue;
            b.SetAwakeFlag();
            for (const auto& kc : bodyContacts[bid]) {
                const std::uint32_t cid = kc.id;
                if (islandedContacts[cid]) continue;
                const Contact<R>& c = contactBuf[cid];
                if ((c.flags & kCEnabled) && (c.flags & kTouching) && !(c.flags & kSensor)) {
                    const std::uint32_t other = (c.bodyA == bid) ? c.bodyB : c.bodyA;
                    island.contacts.push_back(cid);
                    islandedContacts[cid] = 1;
                    if (!islandedBodies[other]) {
                        islandedBodies[other] = 1;
                        stack.push_back(other);
                    }
                }
            }
            // AddJointsToIsland (AabbTreeWorld.cpp:1446-1470).
            for (const auto& bj : bodyJoints[bid]) {
                const std::uint32_t jid = bj.second;
                if (islandedJoints[jid]) continue;
                const std::uint32_t other = bj.first;
                if (other == kNoBody || (bodies[other].flags & kEnabled)) {
                    islandedJoints[jid] = 1;
                    island.joints.push_back(jid);
                    if (other != kNoBody && !islandedBodies[other]) {
                        islandedBodies[other] = 1;
                        stack.push_back(other);
                    }
                }
            }
        }
    }

    // SolveRegIslandViaGS (AabbTreeWorld.cpp:1563-1682)
    IslandStats SolveIsland(const Island& island, const StepConf<R>& conf) {
        IslandStats is;
        const R h = conf.dt;
        for (std::uint32_t bid : island.bodies) bodies[bid].pos0 = bodies[bid].pos1;
        std::vector<BodyConstraint<R>> bcs(bodies.size());
        for (std::uint32_t bid : island.bodies) {
            const Body<R>& b = bodies[bid];
            bcs[bid] = BodyConstraint<R>{b.invMass, b.invRotI, b.lc, b.pos1, IntegratedVelocity(b, conf)};
        }
        std::vector<PositionConstraint<R>> pcs;
        std::vector<VelocityConstraint<R>> vcs;
        for (std::uint32_t cid : island.contacts) {
            const Contact<R>& c = contactBuf[cid];
            const R rA = shapes[c.shapeA].radius, rB = shapes[c.shapeB].radius;
            pcs.push_back(PositionConstraint<R>{manifoldBuf[cid], c.bodyA, c.bodyB, rA + rB});
        }
        for (std::uint32_t cid : island.contacts) {
            const Contact<R>& c = contactBuf[cid];
            const R rA = shapes[c.shapeA].radius, rB = shapes[c.shapeB].radius;
            const Xf<R> xfA = GetXf(bcs[c.bodyA].pos, bcs[c.bodyA].lc);
            const Xf<R> xfB = GetXf(bcs[c.bodyB].pos, bcs[c.bodyB].lc);
            const WorldManifold<R> wm = GetWorldManifold(manifoldBuf[cid], xfA, rA, xfB, rB);
            vcs.push_back(MakeVC(c, wm, bcs, conf));
        }
        if (conf.warmStart) {
            // WarmStartVelocities + CalcWarmStartVelocityDeltas (AabbTreeWorld.cpp:230-274)
            for (const auto& vc : vcs) {
                BodyConstraint<R>& A = bcs[vc.bodyA];
                BodyConstraint<R>& B = bcs[vc.bodyB];
                Vel<R> dA{}, dB{};
                for (std::uint32_t j = 0; j < vc.n; ++j) {
                    const VcPoint<R>& p = vc.p[j];
                    const V2<R> P = p.ni * vc.normal + p.ti * vc.tangent();
                    const R LA = Cross(p.relA, P) / R(1);
                    const R LB = Cross(p.relB, P) / R(1);
                    dA.lin = dA.lin - A.invMass * P;
                    dA.ang = dA.ang - A.invRotI * LA;
                    dB.lin = dB.lin + B.invMass * P;
                    dB.ang = dB.ang + B.invRotI * LB;
                }
                A.vel.lin = A.vel.lin + dA.lin;
                A.vel.ang = A.vel.ang + dA.ang;
                B.vel.lin = B.vel.lin + dB.lin;
                B.vel.ang = B.vel.ang + dB.ang;
            }
        }
        for (std::uint32_t jid : island.joints) JointInitVelocity(joints[jid], bcs, conf);
        is.velocityIters = conf.velIters;
        for (std::uint32_t i = 0; i < conf.velIters; ++i) {
            bool jointsOkay = true;
            for (std::uint32_t jid : island.joints) jointsOkay &= JointSolveVelocity(joints[jid], bcs, conf);
            R newInc = 0;
            for (auto& vc : vcs) newInc = Max(newInc, SolveVC(vc, bcs));
            is.maxIncImpulse = Max(is.maxIncImpulse, double(newInc));
            if (jointsOkay && newInc <= conf.regMinMomentum) {
                is.velocityIters = i + 1;
                break;
            }
        }
        for (std::uint32_t bid : island.bodies) {
            BodyConstraint<R>& bc = bcs[bid];
            const V2<R> tr = h * bc.vel.lin;
            const R rot = h * bc.vel.ang;
            bc.pos = bc.pos + Pos<R>{tr, rot};
        }
        bool solved = false;
        std::uint32_t posIters = conf.posIters;
        for (std::uint32_t i = 0; i < conf.posIters; ++i) {
            R minSep = std::numeric_limits<R>::infinity();
            for (const auto& pc : pcs) minSep = Min(minSep, SolvePC(pc, bcs, conf));
            is.minSeparation = Min(is.minSeparation, double(minSep));
            const bool contactsOkay = minSep >= conf.regMinSeparation;
            bool jointsOkay = true;
            for (std::uint32_t jid : island.joints) jointsOkay &= JointSolvePosition(joints[jid], bcs, conf);
            if (contactsOkay && jointsOkay) {
                solved = true;
                posIters = i + 1;
                break;
            }
        }
        for (std::size_t i = 0; i < vcs.size(); ++i) {
            Manifold<R>& m = manifoldBuf[island.contacts[i]];
            for (std::uint32_t j = 0; j < vcs[i].n; ++j) {
                m.pts[j].ni = vcs[i].p[j].ni;
                m.pts[j].ti = vcs[i].p[j].ti;
            }
        }
        for (std::uint32_t bid : island.bodies) {
            Body<R>& b = bodies[bid];
            const BodyConstraint<R>& bc = bcs[bid];
            b.vel = bc.vel;
            if (b.pos1 != bc.pos) {
                b.pos1 = bc.pos;
                b.xf = GetXf(b.pos1, b.lc);
                for (const auto& kc : bodyContacts[bid]) contactBuf[kc.id].flags |= kNeedsUpdating;
            }
        }
        // Post-solve reports (AabbTreeWorld.cpp:1671-1674): the velocity
        // constraints' impulses and the solved position iteration (255, the
        // 8-bit InvalidIteration, if not solved).
        for (std::size_t i = 0; i < vcs.size(); ++i) {
            EventHead(3, island.contacts[i]);
            events.push_back(vcs[i].n);
            events.push_back(solved ? std::uint32_t(std::uint8_t(posIters - 1)) : 255u);
            R f[4] = {};
            for (std::uint32_t j = 0; j < vcs[i].n; ++j) {
                f[j] = vcs[i].p[j].ni;
                f[2 + j] = vcs[i].p[j].ti;
            }
            for (R v : f) events.push_back(FloatBits(v));
        }
        // UpdateUnderActiveTimes (AabbTreeWorld.cpp:403-417)
        R minUnder = std::numeric_limits<R>::infinity();
        for (std::uint32_t bid : island.bodies) {
            Body<R>& b = bodies[bid];
            if (!b.speedable()) continue;
            const R linSq = MagSq(b.vel.lin);
            const R angSq = b.vel.ang * b.vel.ang;
            const bool under = (angSq <= conf.angularSleepTolerance * conf.angularSleepTolerance) &&
                               (linSq <= conf.linearSleepTolerance * conf.linearSleepTolerance);
            const bool sleepable = (b.flags & kAutoSleep) != 0;
            const R t = (sleepable && under) ? b.underActive + conf.dt : R(0);
            // Body::SetUnderActiveTime (Body.hpp:640-645)
            if (t == R(0) || b.accelerable()) b.underActive = t;
            minUnder = Min(minUnder, t);
        }
        std::uint32_t slept = 0;
        if (minUnder >= conf.minStillTimeToSleep && solved) {
            for (std::uint32_t bid : island.bodies) slept += bodies[bid].Unawaken() ? 1u : 0u;
        }
        is.solved = solved;
        is.positionIters = posIters;
        is.bodiesSlept = slept;
        return is;
    }

    // SolveReg (AabbTreeWorld.cpp:1472-1561)
    void SolveReg(const StepConf<R>& conf, StepStats& st) {
        islandedBodies.assign(bodies.size(), 0);
        islandedContacts.assign(contactBuf.size(), 0);
        islandedJoints.assign(joints.size(), 0);
        for (std::uint32_t bid : bodyOrder) {
            if (islandedBodies[bid]) continue;
            Body<R>& b = bodies[bid];
            if (b.awake() && (b.flags & kEnabled)) {
                ++st.islandsFound;
                Island island;
                AddToIsland(island, bid);
                st.maxIslandBodies = std::max(st.maxIslandBodies, std::uint32_t(island.bodies.size()));
                for (std::uint32_t ib : island.bodies) {
                    if (!bodies[ib].speedable()) islandedBodies[ib] = 0;
                }
                // Update(RegStepStats&, const IslandStats&) (StepStats.cpp:29-38)
                const IslandStats is = SolveIsland(island, conf);
                st.maxIncImpulse = Max(st.maxIncImpulse, is.maxIncImpulse);
                st.minSeparation = Min(st.minSeparation, is.minSeparation);
                st.islandsSolved += is.solved ? 1u : 0u;
                st.sumPosIters += is.positionIters;
                st.sumVelIters += is.velocityIters;
                st.bodiesSlept += is.bodiesSlept;
            }
        }
        for (std::uint32_t bid : bodyOrder) {
            if (islandedBodies[bid] && bodies[bid].speedable()) st.regProxiesMoved += Synchronize(bid, conf);
        }
        islandedBodies.assign(bodies.size(), 0);
        islandedContacts.assign(contactBuf.size(), 0);
        st.regUpdated = UpdateContacts(conf, &st.regSkipped);
        st.regAdded = FindAndAddContacts(conf);
    }

    // ---------------------------------------------------------------------
    // TOI phase (doToi = true): SolveToi (AabbTreeWorld.cpp:1752-1831) and the
    // functions it calls.

    // Sweep of a body (Sweep.hpp:48-68) and its functions.
    struct SweepR {
        Pos<R> pos0, pos1;
        V2<R> lc;
        R alpha0;
    };
    static constexpr R kPi = R(3.14159265358979323846264338327950288L);  // RealConstants.hpp:57
    static SweepR SweepOf(const Body<R>& b) { return SweepR{b.pos0, b.pos1, b.lc, b.alpha0}; }
    // GetPosition(pos0, pos1, beta) = pos0 + (pos1 - pos0) * beta (d2/Math.cpp:30-59).
    static Pos<R> Lerp(const Pos<R>& p0, const Pos<R>& p1, R beta) {
        const V2<R> dl = p1.lin - p0.lin;
        const R da = p1.ang - p0.ang;
        return Pos<R>{p0.lin + dl * beta, p0.ang + da * beta};
    }
    // Advance0(sweep, alpha) (Sweep.cpp:29-33).
    static SweepR Advance0(const SweepR& s, R alpha) {
        const R beta = (alpha - s.alpha0) / (R(1) - s.alpha0);
        return SweepR{Lerp(s.pos0, s.pos1, beta), s.pos1, s.lc, alpha};
    }
    // GetNormalized(Angle), trunc form (Math.cpp:33-91).
    static R NormalizeAngle(R a) {
        const R twoPi = R(2) * kPi;
        const R rTwoPi = R(1) / twoPi;
        a = a - std::trunc(a * rTwoPi) * twoPi;
        if (a >= kPi) a -= kPi * R(2);
        else if (a < -kPi) a += kPi * R(2);
        return a;
    }
    // GetNormalized(Sweep) (d2/Math.hpp:160-167).
    static SweepR Normalized(SweepR s) {
        const R a0 = NormalizeAngle(s.pos0.ang);
        const R d = s.pos0.ang - a0;
        s.pos0.ang = a0;
        s.pos1.ang -= d;
        return s;
    }
    // GetTransformation(sweep, beta) (d2/Math.hpp:251-254).
    static Xf<R> XfAt(const SweepR& s, R beta) { return GetXf(Lerp(s.pos0, s.pos1, beta), s.lc); }
    // Body::Advance0 / Advance / SetSweep (Body.hpp:697-727, 1004-1009).
    static void BodyAdvance0(Body<R>& b, R alpha) {
        const SweepR s = Advance0(SweepOf(b), alpha);
        b.pos0 = s.pos0;
        b.alpha0 = s.alpha0;
    }
    static void BodyAdvance(Body<R>& b, R alpha) {
        BodyAdvance0(b, alpha);
        b.pos1 = b.pos0;
        b.xf = GetXf(b.pos1, b.lc);
    }
    static void SetSweep(Body<R>& b, const SweepR& s) {
        b.pos0 = s.pos0;
        b.pos1 = s.pos1;
        b.lc = s.lc;
        b.alpha0 = s.alpha0;
        b.xf = GetXf(b.pos1, b.lc);
    }
    void FlagForUpdating(std::uint32_t bid) {
        for (const auto& kc : bodyContacts[bid]) contactBuf[kc.id].flags |= kNeedsUpdating;
    }

    // GetSupportIndex (DistanceProxy.hpp:233-249).
    static std::uint32_t SupportIndex(const Proxy<R>& p, V2<R> dir) {
        std::uint32_t index = 255;
        R maxValue = -std::numeric_limits<R>::infinity();
        for (std::uint32_t i = 0; i < p.n; ++i) {
            const R value = Dot(p.v[i], dir);
            if (maxValue < value) {
                maxValue = value;
                index = i;
            }
        }
        return index;
    }
    static bool AlmostZero(R v) { return std::abs(v) < std::numeric_limits<R>::min(); }  // Math.hpp:177-180

    // SimplexEdge / Simplex (SimplexEdge.hpp, Simplex.cpp:25-241).
    struct SEdge {
        V2<R> wA, wB;
        std::uint32_t iA = 255, iB = 255;
    };
    struct Simplex {
        SEdge e[3];
        R c[3] = {0, 0, 0};
        std::uint32_t n = 0;
    };
    static V2<R> PD(const SEdge& s) { return s.wB - s.wA; }
    static Simplex Simplex1(const SEdge& s0) {
        Simplex r;
        r.e[0] = s0;
        r.c[0] = 1;
        r.n = 1;
        return r;
    }
    static Simplex Simplex2(const SEdge& a, const SEdge& b, R ca, R cb) {
        Simplex r;
        r.e[0] = a;
        r.e[1] = b;
        r.c[0] = ca;
        r.c[1] = cb;
        r.n = 2;
        return r;
    }
    static Simplex SolveSimplex(const SEdge* e, std::uint32_t n) {
        if (n == 1) return Simplex1(e[0]);
        if (n == 2) {
            const V2<R> w1 = PD(e[0]), w2 = PD(e[1]);
            const V2<R> e12 = w2 - w1;
            const R d12_2 = -Dot(w1, e12);
            if (d12_2 <= 0) return Simplex1(e[0]);
            const R d12_1 = Dot(w2, e12);
            if (d12_1 <= 0) return Simplex1(e[1]);
            const R inv = R(1) / (d12_1 + d12_2);
            return Simplex2(e[0], e[1], d12_1 * inv, d12_2 * inv);
        }
        const V2<R> w1 = PD(e[0]), w2 = PD(e[1]), w3 = PD(e[2]);
        const V2<R> e12 = w2 - w1;
        const R d12_1 = Dot(w2, e12);
        const R d12_2 = -Dot(w1, e12);
        const V2<R> e13 = w3 - w1;
        const R d13_1 = Dot(w3, e13);
        const R d13_2 = -Dot(w1, e13);
        const V2<R> e23 = w3 - w2;
        const R d23_1 = Dot(w3, e23);
        const R d23_2 = -Dot(w2, e23);
        if (d12_2 <= 0 && d13_2 <= 0) return Simplex1(e[0]);
        if (d12_1 <= 0 && d23_2 <= 0) return Simplex1(e[1]);
        if (d13_1 <= 0 && d23_1 <= 0) return Simplex1(e[2]);
        const R n123 = Cross(e12, e13);
        const R d123_3 = n123 * Cross(w1, w2);
        if (d12_1 > 0 && d12_2 > 0 && d123_3 <= 0) {
            const R inv = R(1) / (d12_1 + d12_2);
            return Simplex2(e[0], e[1], d12_1 * inv, d12_2 * inv);
        }
        const R d123_2 = n123 * Cross(w3, w1);
        if (d13_1 > 0 && d13_2 > 0 && d123_2 <= 0) {
            const R inv = R(1) / (d13_1 + d13_2);
            return Simplex2(e[0], e[2], d13_1 * inv, d13_2 * inv);
        }
        const R d123_1 = n123 * Cross(w2, w3);
        if (d23_1 > 0 && d23_2 > 0 && d123_1 <= 0) {
            const R inv = R(1) / (d23_1 + d23_2);
            return Simplex2(e[2], e[1], d23_2 * inv, d23_1 * inv);
        }
        const R inv = R(1) / (d123_1 + d123_2 + d123_3);
        Simplex r;
        r.e[0] = e[0];
        r.e[1] = e[1];
        r.e[2] = e[2];
        r.c[0] = d123_1 * inv;
        r.c[1] = d123_2 * inv;
        r.c[2] = d123_3 * inv;
        r.n = 3;
        return r;
    }
    static R CalcMetric(const SEdge* e, std::uint32_t n) {
        if (n == 2) return std::sqrt(MagSq(PD(e[1]) - PD(e[0])));
        if (n == 3) return Cross(PD(e[1]) - PD(e[0]), PD(e[2]) - PD(e[0]));
        return R(0);
    }
    // CalcSearchDirection (Simplex.cpp:44-60).
    static V2<R> SearchDirection(const SEdge* e, std::uint32_t n) {
        if (n == 1) return -PD(e[0]);
        if (n == 2) {
            const V2<R> e12 = PD(e[1]) - PD(e[0]);
            const V2<R> e0 = PD(e[0]);
            const R sgn = Cross(e12, -e0);
            return (sgn > 0) ? RevPerp(e12) : FwdPerp(e12);
        }
        return V2<R>{0, 0};
    }
    // Simplex::Cache (Simplex.hpp:96-101, 156-159).
    struct Cache {
        R metric = std::numeric_limits<R>::quiet_NaN();
        std::uint32_t ia[3] = {255, 255, 255}, ib[3] = {255, 255, 255};
        std::uint32_t Count() const {
            std::uint32_t k = 0;
            for (int i = 0; i < 3; ++i) k += (ia[i] == 255 && ib[i] == 255) ? 0u : 1u;
            return k;
        }
    };
    static Cache CacheOf(const Simplex& s) {
        Cache c;
        c.metric = CalcMetric(s.e, s.n);
        for (std::uint32_t i = 0; i < s.n; ++i) {
            c.ia[i] = s.e[i].iA;
            c.ib[i] = s.e[i].iB;
        }
        return c;
    }
    static SEdge EdgeOf(const Proxy<R>& pA, const Xf<R>& xfA, std::uint32_t iA, const Proxy<R>& pB, const Xf<R>& xfB,
                        std::uint32_t iB) {
        return SEdge{Transform(pA.v[iA], xfA), Transform(pB.v[iB], xfB), iA, iB};
    }
    enum DistState { kDistHitMaxIters, kDistMaxPoints, kDistUnfitSearchDir, kDistDuplicateIndexPair };
    struct DistOut {
        Simplex s;
        std::uint32_t iters;
        DistState state;
    };
    // Distance (Distance.cpp:95-175).
    static DistOut Distance(const Proxy<R>& pA, const Xf<R>& xfA, const Proxy<R>& pB, const Xf<R>& xfB,
                            const Cache& cache, std::uint32_t maxIters) {
        std::uint32_t sa[3] = {cache.ia[0], cache.ia[1], cache.ia[2]};
        std::uint32_t sb[3] = {cache.ib[0], cache.ib[1], cache.ib[2]};
        SEdge e[3];
        std::uint32_t n = cache.Count();
        for (std::uint32_t i = 0; i < n; ++i) e[i] = EdgeOf(pA, xfA, sa[i], pB, xfB, sb[i]);
        if (n > 1) {
            const R metric1 = cache.metric;
            const R metric2 = CalcMetric(e, n);
            if ((metric2 < (metric1 / 2)) || (metric2 > (metric1 * 2)) || (metric2 < 0) || AlmostZero(metric2)) n = 0;
        }
        if (n == 0) {
            e[0] = EdgeOf(pA, xfA, 0, pB, xfB, 0);
            n = 1;
            sa[0] = 0;
            sb[0] = 0;
            sa[1] = sb[1] = sa[2] = sb[2] = 255;
        }
        Simplex simplex;
        DistState state = kDistHitMaxIters;
        std::uint32_t iter = 0;
        while (iter < maxIters) {
            ++iter;
            simplex = SolveSimplex(e, n);
            n = simplex.n;
            for (std::uint32_t i = 0; i < n; ++i) e[i] = simplex.e[i];
            if (n == 3) {
                state = kDistMaxPoints;
                break;
            }
            const V2<R> d = SearchDirection(e, n);
            if (AlmostZero(MagSq(d))) {
                state = kDistUnfitSearchDir;
                break;
            }
            const std::uint32_t iA = SupportIndex(pA, InvRotate(-d, xfA.q));
            const std::uint32_t iB = SupportIndex(pB, InvRotate(d, xfB.q));
            bool dup = false;
            for (int k = 0; k < 3; ++k) dup = dup || (sa[k] == iA && sb[k] == iB);
            if (dup) {
                state = kDistDuplicateIndexPair;
                break;
            }
            e[n++] = EdgeOf(pA, xfA, iA, pB, xfB, iB);
            for (std::uint32_t k = 0; k < 3; ++k) {
                sa[k] = k < n ? e[k].iA : 255u;
                sb[k] = k < n ? e[k].iB : 255u;
            }
        }
        return DistOut{simplex, iter, state};
    }
    // GetDelta(GetWitnessPoints(simplex)) (Distance.cpp:75-93, Distance.hpp:55-58).
    static V2<R> WitnessDelta(const Simplex& s) {
        V2<R> a{0, 0}, b{0, 0};
        for (std::uint32_t i = 0; i < s.n; ++i) {
            a = a + s.e[i].wA * s.c[i];
            b = b + s.e[i].wB * s.c[i];
        }
        return b - a;
    }

    // SeparationScenario (SeparationScenario.cpp:25-220).
    enum ScenarioType { kScenPoints, kScenFaceA, kScenFaceB };
    struct Scenario {
        ScenarioType type;
        V2<R> axis, localPoint;
    };
    static Scenario GetScenario(const Cache& c, const Proxy<R>& pA, const Xf<R>& xfA, const Proxy<R>& pB,
                                const Xf<R>& xfB) {
        const V2<R> zero{0, 0};
        if (c.Count() >= 2) {
            if (c.ia[0] == c.ia[1]) {
                const V2<R> b1 = pB.v[c.ib[0]], b2 = pB.v[c.ib[1]];
                const V2<R> axis = UnitVector(FwdPerp(b2 - b1), zero);
                const V2<R> normal = Rotate(axis, xfB.q);
                const V2<R> lp = (b1 + b2) / R(2);
                const V2<R> pointB = Transform(lp, xfB);
                const V2<R> pointA = Transform(pA.v[c.ia[0]], xfA);
                const V2<R> delta = pointA - pointB;
                return Scenario{kScenFaceB, (Dot(delta, normal) < 0) ? -axis : axis, lp};
            }
            const V2<R> a1 = pA.v[c.ia[0]], a2 = pA.v[c.ia[1]];
            const V2<R> axis = UnitVector(FwdPerp(a2 - a1), zero);
            const V2<R> normal = Rotate(axis, xfA.q);
            const V2<R> lp = (a1 + a2) / R(2);
            const V2<R> pointA = Transform(lp, xfA);
            const V2<R> pointB = Transform(pB.v[c.ib[0]], xfB);
            const V2<R> delta = pointB - pointA;
            return Scenario{kScenFaceA, (Dot(delta, normal) < 0) ? -axis : axis, lp};
        }
        const V2<R> pointA = Transform(pA.v[c.ia[0]], xfA);
        const V2<R> pointB = Transform(pB.v[c.ib[0]], xfB);
        return Scenario{kScenPoints, UnitVector(pointB - pointA, zero), zero};
    }
    struct SepIdx {
        R distance;
        std::uint32_t ia, ib;
    };
    static SepIdx FindMinSep(const Scenario& f, const Proxy<R>& pA, const Xf<R>& xfA, const Proxy<R>& pB,
                             const Xf<R>& xfB) {
        if (f.type == kScenPoints) {
            const std::uint32_t ia = SupportIndex(pA, InvRotate(f.axis, xfA.q));
            const std::uint32_t ib = SupportIndex(pB, InvRotate(-f.axis, xfB.q));
            const V2<R> delta = Transform(pB.v[ib], xfB) - Transform(pA.v[ia], xfA);
            return SepIdx{Dot(delta, f.axis), ia, ib};
        }
        if (f.type == kScenFaceA) {
            const V2<R> normal = Rotate(f.axis, xfA.q);
            const V2<R> pointA = Transform(f.localPoint, xfA);
            const std::uint32_t ib = SupportIndex(pB, InvRotate(-normal, xfB.q));
            const V2<R> delta = Transform(pB.v[ib], xfB) - pointA;
            return SepIdx{Dot(delta, normal), 255u, ib};
        }
        const V2<R> normal = Rotate(f.axis, xfB.q);
        const std::uint32_t ia = SupportIndex(pA, InvRotate(-normal, xfA.q));
        const V2<R> pointA = Transform(pA.v[ia], xfA);
        const V2<R> delta = pointA - Transform(f.localPoint, xfB);
        return SepIdx{Dot(delta, normal), ia, 255u};
    }
    static R Evaluate(const Scenario& f, const Proxy<R>& pA, const Xf<R>& xfA, const Proxy<R>& pB, const Xf<R>& xfB,
                      std::uint32_t ia, std::uint32_t ib) {
        if (f.type == kScenPoints) {
            const V2<R> delta = Transform(pB.v[ib], xfB) - Transform(pA.v[ia], xfA);
            return Dot(delta, f.axis);
        }
        if (f.type == kScenFaceA) {
            const V2<R> normal = Rotate(f.axis, xfA.q);
            const V2<R> pointA = Transform(f.localPoint, xfA);
            const V2<R> delta = Transform(pB.v[ib], xfB) - pointA;
            return Dot(delta, normal);
        }
        const V2<R> normal = Rotate(f.axis, xfB.q);
        const V2<R> pointB = Transform(f.localPoint, xfB);
        const V2<R> delta = Transform(pA.v[ia], xfA) - pointB;
        return Dot(delta, normal);
    }

    enum ToiState {
        kToiUnknown, kToiTouching, kToiSeparated, kToiOverlapped, kToiMaxRootIters, kToiNextAfter, kToiMaxToiIters,
        kToiBelowMinTarget, kToiMaxDistIters, kToiTargetDepthExceeds, kToiOverflow
    };
    // ToiOutput (ToiOutput.hpp:44-70): time, state and the iteration statistics.
    struct ToiOut {
        R time;
        ToiState state;
        std::uint32_t toiIters = 0, maxDistIters = 0, maxRootIters = 0;
    };
    // GetToiViaSat (TimeOfImpact.cpp:35-251) with ToiConf from the StepConf
    // (ToiConf.cpp:26-35): timeMax 1.
    static ToiOut ToiViaSat(const Proxy<R>& pA, const SweepR& sA, const Proxy<R>& pB, const SweepR& sB,
                            const StepConf<R>& conf) {
        const R totalRadius = pA.radius + pB.radius;
        if (conf.targetDepth > totalRadius) return ToiOut{0, kToiTargetDepthExceeds};
        const R target = totalRadius - conf.targetDepth;
        const R maxTarget = Max(target + conf.tolerance, R(0));
        const R minTarget = Max(target - conf.tolerance, R(0));
        const R minTargetSq = minTarget * minTarget;
        if (!std::isfinite(minTargetSq) && std::isfinite(minTarget)) return ToiOut{0, kToiOverflow};
        const R maxTargetSq = maxTarget * maxTarget;
        if (!std::isfinite(maxTargetSq) && std::isfinite(maxTarget)) return ToiOut{0, kToiOverflow};
        R timeLo = 0;
        Xf<R> loA = XfAt(sA, timeLo), loB = XfAt(sB, timeLo);
        Cache cache;
        std::uint32_t toiIters = 0, maxDist = 0, maxRoot = 0;
        auto out = [&](R t, ToiState st) {
            ToiOut o{t, st};
            o.toiIters = toiIters;
            o.maxDistIters = maxDist;
            o.maxRootIters = maxRoot;
            return o;
        };
        while (toiIters < conf.maxToiIters) {
            const DistOut di = Distance(pA, loA, pB, loB, cache, conf.maxDistanceIters);
            ++toiIters;
            maxDist = std::max(maxDist, di.iters);
            if (di.state == kDistHitMaxIters) return out(timeLo, kToiMaxDistIters);
            cache = CacheOf(di.s);
            const R distSq = MagSq(WitnessDelta(di.s));
            if (distSq < minTargetSq) return out(timeLo, kToiOverlapped);
            if (distSq <= maxTargetSq) return out(timeLo, kToiTouching);
            const Scenario fcn = GetScenario(cache, pA, loA, pB, loB);
            R timeHi = R(1);
            Xf<R> hiA = XfAt(sA, timeHi), hiB = XfAt(sB, timeHi);
            for (std::uint32_t pbIter = 0; pbIter < 254u; ++pbIter) {
                const SepIdx hiMin = FindMinSep(fcn, pA, hiA, pB, hiB);
                if (hiMin.distance > maxTarget) return out(timeHi, kToiSeparated);
                if (hiMin.distance >= minTarget) {
                    if (timeHi == timeLo) return out(timeHi, kToiSeparated);
                    timeLo = timeHi;
                    loA = hiA;
                    loB = hiB;
                    break;
                }
                const R loEval = Evaluate(fcn, pA, loA, pB, loB, hiMin.ia, hiMin.ib);
                if (loEval <= maxTarget) return out(timeLo, (loEval < minTarget) ? kToiBelowMinTarget : kToiTouching);
                R a1 = timeLo, a2 = timeHi, s1 = loEval, s2 = hiMin.distance;
                std::uint32_t roots = 0;
                R t = timeLo;
                for (;;) {
                    if (roots == conf.maxToiRootIters) {
                        maxRoot = std::max(maxRoot, roots);
                        return out(t, kToiMaxRootIters);
                    }
                    if (std::nextafter(a1, a2) >= a2) {
                        maxRoot = std::max(maxRoot, roots);
                        return out(t, kToiNextAfter);
                    }
                    // Secant / Bisect (Math.hpp:95-108).
                    t = (roots & 1u) ? a1 + (target - s1) * (a2 - a1) / (s2 - s1) : (a1 + a2) / 2;
                    ++roots;
                    const Xf<R> tA = XfAt(sA, t), tB = XfAt(sB, t);
                    const R sv = Evaluate(fcn, pA, tA, pB, tB, hiMin.ia, hiMin.ib);
                    if (std::abs(sv - target) <= conf.tolerance) {
                        timeHi = t;
                        hiA = tA;
                        hiB = tB;
                        break;
                    }
                    if (sv > target) {
                        a1 = t;
                        s1 = sv;
                    } else {
                        a2 = t;
                        s2 = sv;
                    }
                }
                maxRoot = std::max(maxRoot, roots);
            }
        }
        return out(timeLo, kToiMaxToiIters);
    }

    // UpdateContactTOIs (AabbTreeWorld.cpp:1685-1750).
    void UpdateContactTOIs(const StepConf<R>& conf, StepStats& st) {
        for (const auto& kc : contacts) {
            Contact<R>& c = contactBuf[kc.id];
            if (c.toiValid) continue;
            if (!(c.flags & kCEnabled) || (c.flags & kSensor) || !(c.flags & kCImpenetrable)) continue;
            if (c.toiCount >= conf.maxSubSteps) {
                ++st.toiAtMaxSubSteps;
                continue;
            }
            Body<R>& bA = bodies[c.bodyA];
            Body<R>& bB = bodies[c.bodyB];
            if (!bA.awake() && !bB.awake()) continue;
            const R alpha0 = Max(bA.alpha0, bB.alpha0);
            BodyAdvance0(bA, alpha0);
            BodyAdvance0(bB, alpha0);
            const SweepR sA = Normalized(SweepOf(bA));
            const SweepR sB = Normalized(SweepOf(bB));
            const ToiOut out = ToiViaSat(shapes[c.shapeA].proxy(), sA, shapes[c.shapeB].proxy(), sB, conf);
            c.toi = (out.state == kToiTouching) ? Min(alpha0 + (R(1) - alpha0) * out.time, R(1)) : R(1);
            c.toiValid = true;
            st.toiMaxDistIters = std::max(st.toiMaxDistIters, out.maxDistIters);
            st.toiMaxToiIters = std::max(st.toiMaxToiIters, out.toiIters);
            st.toiMaxRootIters = std::max(st.toiMaxRootIters, out.maxRootIters);
            ++st.toiUpdatedToi;
        }
    }

    // ProcessContactsForTOI (AabbTreeWorld.cpp:2059-2142).
    void ProcessContactsForTOI(std::uint32_t id, Island& island, R toi, const StepConf<R>& conf,
                               ToiIslandStats& is) {
        const bool bodyImpenetrable = (bodies[id].flags & kImpenetrable) != 0;
        for (const auto& kc : bodyContacts[id]) {
            const std::uint32_t cid = kc.id;
            if (islandedContacts[cid]) continue;
            if (contactBuf[cid].flags & kSensor) continue;
            const std::uint32_t otherId = (contactBuf[cid].bodyA == id) ? contactBuf[cid].bodyB : contactBuf[cid].bodyA;
            Body<R>& other = bodies[otherId];
            if (!(bodyImpenetrable || (other.flags & kImpenetrable))) continue;
            const bool otherIslanded = islandedBodies[otherId] != 0;
            const SweepR backup = SweepOf(other);
            if (!otherIslanded) {
                BodyAdvance(other, toi);
                if (other.pos0 != backup.pos0 || other.pos1 != backup.pos1) FlagForUpdating(otherId);
            }
            if (contactBuf[cid].flags & kNeedsUpdating) {
                Update(cid, conf);
                ++is.contactsUpdated;
            } else {
                ++is.contactsSkipped;
            }
            const std::uint32_t fl = contactBuf[cid].flags;
            if (!(fl & kCEnabled) || !(fl & kTouching)) {
                SetSweep(other, backup);
                continue;
            }
            island.contacts.push_back(cid);
            islandedContacts[cid] = 1;
            if (!otherIslanded) {
                if (other.speedable()) other.SetAwakeFlag();
                island.bodies.push_back(otherId);
                islandedBodies[otherId] = 1;
            }
        }
    }

    // SolveToiViaGS (AabbTreeWorld.cpp:1968-2057) with conf.dt the rest of the step.
    void SolveToiViaGS(const Island& island, const StepConf<R>& conf, IslandStats& is) {
        StepConf<R> still = conf;
        still.dt = 0;
        std::vector<BodyConstraint<R>> bcs(bodies.size());
        for (std::uint32_t bid : island.bodies) {
            const Body<R>& b = bodies[bid];
            bcs[bid] = BodyConstraint<R>{b.invMass, b.invRotI, b.lc, b.pos1, IntegratedVelocity(b, still)};
        }
        std::vector<PositionConstraint<R>> pcs;
        for (std::uint32_t cid : island.contacts) {
            const Contact<R>& c = contactBuf[cid];
            const R rA = shapes[c.shapeA].radius, rB = shapes[c.shapeB].radius;
            pcs.push_back(PositionConstraint<R>{manifoldBuf[cid], c.bodyA, c.bodyB, rA + rB});
        }
        std::uint32_t positionIters = conf.toiPosIters;
        StepConf<R> ps = conf;  // GetToiConstraintSolverConf (ConstraintSolverConf.cpp:37-45)
        ps.regResolutionRate = conf.toiResolutionRate;
        for (std::uint32_t i = 0; i < conf.toiPosIters; ++i) {
            R minSep = std::numeric_limits<R>::infinity();
            for (const auto& pc : pcs) minSep = Min(minSep, SolvePC(pc, bcs, ps));
            is.minSeparation = Min(is.minSeparation, double(minSep));
            if (minSep >= conf.toiMinSeparation) {
                positionIters = i + 1;
                is.solved = true;
                break;
            }
        }
        is.positionIters = positionIters;
        for (std::uint32_t bid : island.bodies) bodies[bid].pos0 = bcs[bid].pos;
        StepConf<R> vconf = conf;  // GetToiVelocityConstraintConf (VelocityConstraint.cpp:207-210)
        vconf.warmStart = false;
        std::vector<VelocityConstraint<R>> vcs;
        for (std::uint32_t cid : island.contacts) {
            const Contact<R>& c = contactBuf[cid];
            const R rA = shapes[c.shapeA].radius, rB = shapes[c.shapeB].radius;
            const Xf<R> xfA = GetXf(bcs[c.bodyA].pos, bcs[c.bodyA].lc);
            const Xf<R> xfB = GetXf(bcs[c.bodyB].pos, bcs[c.bodyB].lc);
            const WorldManifold<R> wm = GetWorldManifold(manifoldBuf[cid], xfA, rA, xfB, rB);
            vcs.push_back(MakeVC(c, wm, bcs, vconf));
        }
        is.velocityIters = conf.toiVelIters;
        for (std::uint32_t i = 0; i < conf.toiVelIters; ++i) {
            R newInc = 0;
            for (auto& vc : vcs) newInc = Max(newInc, SolveVC(vc, bcs));
            if (newInc <= conf.toiMinMomentum) {
                is.velocityIters = i + 1;
                break;
            }
            is.maxIncImpulse = Max(is.maxIncImpulse, double(newInc));
        }
        const R h = conf.dt;
        for (std::uint32_t bid : island.bodies) {
            BodyConstraint<R>& bc = bcs[bid];
            const V2<R> tr = h * bc.vel.lin;
            const R rot = h * bc.vel.ang;
            bc.pos = bc.pos + Pos<R>{tr, rot};
        }
        for (std::uint32_t bid : island.bodies) {
            Body<R>& b = bodies[bid];
            b.vel = bcs[bid].vel;
            if (b.pos1 != bcs[bid].pos) {
                b.pos1 = bcs[bid].pos;
                b.xf = GetXf(b.pos1, b.lc);
                FlagForUpdating(bid);
            }
        }
        // Post-solve reports with the position iterations (AabbTreeWorld.cpp:2052-2054).
        for (std::size_t i = 0; i < vcs.size(); ++i) {
            EventHead(3, island.contacts[i]);
            events.push_back(vcs[i].n);
            events.push_back(positionIters);
            R f[4] = {};
            for (std::uint32_t j = 0; j < vcs[i].n; ++j) {
                f[j] = vcs[i].p[j].ni;
                f[2 + j] = vcs[i].p[j].ti;
            }
            for (R v : f) events.push_back(FloatBits(v));
        }
    }

    // SolveToi(ContactID) (AabbTreeWorld.cpp:1833-1966).
    IslandStats SolveToiContact(std::uint32_t cid, const StepConf<R>& conf, std::uint32_t& updated,
                                std::uint32_t& skipped) {
        IslandStats is;
        updated = skipped = 0;
        const R toi = contactBuf[cid].toi;
        const std::uint32_t ia = contactBuf[cid].bodyA, ib = contactBuf[cid].bodyB;
        Body<R>& bA = bodies[ia];
        Body<R>& bB = bodies[ib];
        {
            const SweepR backupA = SweepOf(bA), backupB = SweepOf(bB);
            BodyAdvance(bA, toi);
            if (bA.pos0 != backupA.pos0 || bA.pos1 != backupA.pos1) FlagForUpdating(ia);
            BodyAdvance(bB, toi);
            if (bB.pos0 != backupB.pos0 || bB.pos1 != backupB.pos1) FlagForUpdating(ib);
            if (contactBuf[cid].flags & kNeedsUpdating) {
                Update(cid, conf);
                ++updated;
            }
            contactBuf[cid].toiValid = false;
            ++contactBuf[cid].toiCount;
            const std::uint32_t fl = contactBuf[cid].flags;
            if (!(fl & kCEnabled) || !(fl & kTouching)) {
                SetSweep(bA, backupA);
                SetSweep(bB, backupB);
                skipped = updated ? 0u : 1u;
                return is;
            }
        }
        if (bA.speedable()) bA.SetAwakeFlag();
        if (bB.speedable()) bB.SetAwakeFlag();
        Island island;
        islandedBodies[ia] = 1;
        islandedBodies[ib] = 1;
        islandedContacts[cid] = 1;
        island.bodies.push_back(ia);
        island.bodies.push_back(ib);
        island.contacts.push_back(cid);
        ToiIslandStats proc;
        if (bA.accelerable()) ProcessContactsForTOI(ia, island, toi, conf, proc);
        if (bB.accelerable()) ProcessContactsForTOI(ib, island, toi, conf, proc);
        for (std::uint32_t b : island.bodies) {
            if (!bodies[b].speedable()) islandedBodies[b] = 0;
        }
        StepConf<R> sub = conf;
        sub.dt = (R(1) - toi) * conf.dt;
        SolveToiViaGS(island, sub, is);
        updated += proc.contactsUpdated;
        skipped += proc.contactsSkipped;
        return is;
    }

    // SolveToi(const StepConf&) (AabbTreeWorld.cpp:1752-1831), no sub-stepping.
    void SolveToiPhase(const StepConf<R>& conf, StepStats& st) {
        islandedBodies.assign(bodies.size(), 0);
        islandedContacts.assign(contactBuf.size(), 0);
        for (;;) {
            UpdateContactTOIs(conf, st);
            std::uint32_t next = 0xffffffffu;
            R minToi = R(1);
            for (const auto& kc : contacts) {
                const Contact<R>& c = contactBuf[kc.id];
                if (c.toiValid && minToi > c.toi) {
                    minToi = c.toi;
                    next = kc.id;
                }
            }
            if (next == 0xffffffffu) {
                for (std::uint32_t bid : bodyOrder) bodies[bid].alpha0 = 0;
                for (const auto& kc : contacts) {
                    contactBuf[kc.id].toiValid = false;
                    contactBuf[kc.id].toiCount = 0;
                }
                break;
            }
            ++st.toiContactsFound;
            if (!islandedContacts[next]) {
                std::uint32_t updated = 0, skipped = 0;
                const IslandStats r = SolveToiContact(next, conf, updated, skipped);
                st.toiMinSeparation = Min(st.toiMinSeparation, r.minSeparation);
                st.toiMaxIncImpulse = Max(st.toiMaxIncImpulse, r.maxIncImpulse);
                st.toiIslandsSolved += r.solved ? 1u : 0u;
                st.toiSumPosIters += r.positionIters;
                st.toiSumVelIters += r.velocityIters;
                if (r.positionIters > 0 || r.velocityIters > 0) ++st.toiIslandsFound;
                st.toiUpdatedTouching += updated;
                st.toiSkippedTouching += skipped;
            }
            for (std::uint32_t bid : bodyOrder) {
                if (!islandedBodies[bid]) continue;
                islandedBodies[bid] = 0;
                if (bodies[bid].accelerable()) {
                    st.toiProxiesMoved += Synchronize(bid, conf);
                    for (const auto& kc : bodyContacts[bid]) {
                        contactBuf[kc.id].toiValid = false;
                        islandedContacts[kc.id] = 0;
                    }
                }
            }
            st.toiAdded += FindAndAddContacts(conf);
            if (islandedContacts.size() < contactBuf.size()) islandedContacts.resize(contactBuf.size(), 0);
        }
        std::uint32_t skipped = 0;
        st.toiUpdatedTouching += UpdateContacts(conf, &skipped);
        st.toiSkippedTouching += skipped;
    }

    // World edits between steps.  SetBody (AabbTreeWorld.cpp:2723-2795) for a
    // body whose sweep becomes one position: a new transformation flags the
    // body's contacts for updating and queues the body for the next step's
    // Synchronize.
    void SetSweepPosition(std::uint32_t bid, const Pos<R>& pos) {
        Body<R>& b = bodies[bid];
        const Xf<R> old = b.xf;
        b.pos0 = pos;
        b.pos1 = pos;
        b.alpha0 = 0;
        b.xf = GetXf(pos, b.lc);
        if (!(old.p == b.xf.p) || !(old.q == b.xf.q)) {
            FlagForUpdating(bid);
            bodiesForSync.push_back(bid);
        }
    }
    // SetTransformation(Body&, xfm) (Body.cpp:270-273): the angle of q.
    void SetTransformation(std::uint32_t bid, V2<R> p, V2<R> q) { SetSweepPosition(bid, Pos<R>{p, std::atan2(q.y, q.x)}); }
    // SetLocation(Body&, v) (Body.cpp:275-278).
    void SetLocation(std::uint32_t bid, V2<R> v) { SetTransformation(bid, v, bodies[bid].xf.q); }
    // SetAngle(Body&, a) (Body.cpp:285-288).
    void SetAngle(std::uint32_t bid, R a) { SetSweepPosition(bid, Pos<R>{bodies[bid].pos1.lin, a}); }
    // ShiftOrigin (AabbTreeWorld.cpp:2221-2242, DynamicTree.cpp:650-658).
    void ShiftOrigin(V2<R> o) {
        // ShiftOrigin(PulleyJointConf&) (PulleyJointConf.cpp:255-260); the
        // other joint kinds hold nothing in world coordinates.
        for (Joint<R>& j : joints) {
            if (j.kind != 7 || j.destroyed) continue;
            j.groundA = j.groundA - o;
            j.groundB = j.groundB - o;
        }
        for (std::uint32_t bid : bodyOrder) {
            Body<R>& b = bodies[bid];
            b.pos0.lin = b.pos0.lin - o;
            b.pos1.lin = b.pos1.lin - o;
            b.xf = GetXf(b.pos1, b.lc);
            FlagForUpdating(bid);
        }
        for (auto& lf : tree) {
            if (!lf.used) continue;
            for (int k = 0; k < 2; ++k) {
                const R ok = k ? o.y : o.x;
                lf.lo[k] = lf.lo[k] + (-ok);
                lf.hi[k] = lf.hi[k] + (-ok);
            }
        }
    }

    // SetVelocity(World&, BodyID, Velocity) (WorldBody.cpp:263-268) with
    // Body::SetVelocity (Body.cpp:184-194): a velocity other than zero is
    // ignored by a body that is not speedable and wakes the others (their
    // under-active time restarts).
    void SetVelocity(std::uint32_t bid, V2<R> v, R w) {
        Body<R>& b = bodies[bid];
        if (!(v.x == R(0) && v.y == R(0) && w == R(0))) {
            if (!(b.flags & kSpeedable)) return;
            b.flags |= kAwake;
            b.underActive = R(0);
        }
        b.vel = Vel<R>{v, w};
    }

    // Destroy(World&, JointID) (AabbTreeWorld.cpp:1215-1227) with
    // AabbTreeWorld::Remove(JointID) (:1177-1213): both bodies wake and the
    // joint leaves their joint lists.  Contacts between its bodies would need
    // filtering, but none exists while a joint keeps its bodies from
    // colliding (joints here are created before any contact).
    void DestroyJoint(std::uint32_t jid) {
        Joint<R>& j = joints[jid];
        if (j.destroyed) return;
        auto drop = [&](std::uint32_t body) {
            bodies[body].SetAwake();
            auto& lst = bodyJoints[body];
            for (auto it = lst.begin(); it != lst.end(); ++it) {
                if (it->second == jid) {
                    lst.erase(it);
                    break;
                }
            }
        };
        if (j.bodyA != kNoBody) drop(j.bodyA);
        if (j.bodyB != kNoBody) drop(j.bodyB);
        j.destroyed = true;
    }

    // Destroy(World&, BodyID) (AabbTreeWorld.cpp:1062-1104): the body's joints
    // (the first of its list each time), its contacts in its list's order
    // (AabbTreeWorld::Destroy(ContactID, from) :2281-2288: out of m_contacts,
    // InternalDestroy leaves the destroyed body's own list alone), its queued
    // proxy creations, its proxies in leaf order (FindProxies :617-630), then
    // AabbTreeWorld::Remove(BodyID) (:1046-1060): out of the sync queue and
    // the body list.
    void DestroyBody(std::uint32_t bid) {
        Body<R>& b = bodies[bid];
        if (b.destroyed) return;
        FreshEvents();
        while (!bodyJoints[bid].empty()) DestroyJoint(bodyJoints[bid].front().second);
        const std::vector<KeyedId> own = bodyContacts[bid];
        for (const KeyedId& kc : own) {
            for (auto it = contacts.begin(); it != contacts.end(); ++it) {
                if (it->id == kc.id) {
                    contacts.erase(it);
                    break;
                }
            }
            InternalDestroy(kc.id, bid);
        }
        bodyContacts[bid].clear();
        fixturesForProxies.erase(std::remove_if(fixturesForProxies.begin(), fixturesForProxies.end(),
                                                [bid](const std::pair<std::uint32_t, std::uint32_t>& f) {
                                                    return f.first == bid;
                                                }),
                                 fixturesForProxies.end());
        std::vector<std::uint32_t> leaves = b.proxies;
        std::sort(leaves.begin(), leaves.end());
        for (std::uint32_t id : leaves) FreeLeaf(id);
        b.proxies.clear();
        bodiesForSync.erase(std::remove(bodiesForSync.begin(), bodiesForSync.end(), bid), bodiesForSync.end());
        bodyOrder.erase(std::remove(bodyOrder.begin(), bodyOrder.end(), bid), bodyOrder.end());
        b.destroyed = true;
    }

    // Mass data of one shape (MassData.cpp:40-181): a disk (one vertex), an
    // edge with rounded ends (two), a polygon (more; the vertex radius does
    // not enter).
    struct MassDataR {
        V2<R> center;
        R mass = 0, I = 0;
    };
    static R PolarMoment(const V2<R>* v, std::uint32_t count) {  // GetPolarMoment (Math.cpp:243-274)
        R sx = 0, sy = 0;
        for (std::uint32_t i = 0; i < count; ++i) {
            const V2<R> a = v[i];
            const V2<R> b = v[(i + 1 < count) ? i + 1 : 0];
            const R fb = Cross(a, b);
            sx += ((a.y * a.y) + a.y * b.y + (b.y * b.y)) * fb;
            sy += ((a.x * a.x) + a.x * b.x + (b.x * b.x)) * fb;
        }
        return (sx + sy) * (R(1) / R(12));
    }
    static MassDataR ShapeMassData(const Shape<R>& sh) {
        const std::uint32_t count = std::uint32_t(sh.v.size());
        const R r = sh.radius, density = sh.density;
        if (count == 0) return MassDataR{};
        if (count == 1) {
            const R r2 = r * r;
            const R area = r2 * kPiR;
            const R mass = density * area;
            const R Iz = area * ((r2 / R(2)) + MagSq(sh.v[0]));
            return MassDataR{sh.v[0], mass, Iz * density / R(1)};
        }
        if (count == 2) {
            const V2<R> v0 = sh.v[0], v1 = sh.v[1];
            const R r2 = r * r;
            const R circleArea = r2 * kPiR;
            const R circleMass = density * circleArea;
            const V2<R> d = v1 - v0;
            const V2<R> offset = RevPerp(UnitVector(d, V2<R>{R(0), R(0)})) * r;
            const R b = std::sqrt(MagSq(d));
            const R h = r * R(2);
            const R rectMass = density * b * h;
            const R totalMass = circleMass + rectMass;
            const V2<R> center = (v0 + v1) / R(2);
            const R halfCircleArea = circleArea / R(2);
            const R halfR2 = r2 / R(2);
            const V2<R> verts[4] = {v0 + offset, v0 - offset, v1 - offset, v1 + offset};
            const R Iz = PolarMoment(verts, 4);
            const R I0 = halfCircleArea * (halfR2 + MagSq(v0));
            const R I1 = halfCircleArea * (halfR2 + MagSq(v1));
            return MassDataR{center, totalMass, (I0 + I1 + Iz) * density / R(1)};
        }
        V2<R> center{}, sum{};
        R area = 0, I = 0;
        for (const V2<R>& p : sh.v) sum = sum + p;
        const V2<R> sc = sum / R(count);
        for (std::uint32_t i = 0; i < count; ++i) {
            const V2<R> e1 = sh.v[i] - sc;
            const V2<R> e2 = sh.v[(i + 1 < count) ? i + 1 : 0] - sc;
            const R D = Cross(e1, e2);
            const R triangleArea = D * (R(1) / R(2));
            area += triangleArea;
            center = center + (triangleArea * (e1 + e2)) * (R(1) / R(3));
            const R intx2 = (e1.x * e1.x) + e2.x * e1.x + (e2.x * e2.x);
            const R inty2 = (e1.y * e1.y) + e2.y * e1.y + (e2.y * e2.y);
            I += D * (intx2 + inty2) * (R(1) / R(12));
        }
        const R mass = density * area;
        center = ((area > R(0)) && !AlmostZero(area)) ? center / area : V2<R>{};
        const V2<R> mc = center + sc;
        const R lever = MagSq(mc) - MagSq(center);
        return MassDataR{mc, mass, ((density * I) + (mass * lever)) / R(1)};
    }
    static constexpr R kPiR = R(3.14159265358979323846264338327950288L);  // RealConstants.hpp:57

    // SetBody's handling of a changed transformation (AabbTreeWorld.cpp:2787-2793).
    void TransformationChanged(std::uint32_t bid, const Xf<R>& old) {
        const Xf<R>& xf = bodies[bid].xf;
        if (!(old.p == xf.p) || !(old.q == xf.q)) {
            FlagForUpdating(bid);
            bodiesForSync.push_back(bid);
        }
    }

    // ResetMassData (WorldBody.cpp:360-363): ComputeMassData over the body's
    // shapes of positive density (WorldShape.cpp:160-176), then SetMassData
    // (WorldBody.cpp:365-400).
    void ResetMassData(std::uint32_t bid) {
        Body<R>& b = bodies[bid];
        R mass = 0, I = 0;
        V2<R> wc{};
        for (std::uint32_t sid : b.shapes) {
            const Shape<R>& sh = shapes[sid];
            if (sh.density > R(0)) {
                const MassDataR md = ShapeMassData(sh);
                mass += md.mass;
                wc = wc + (md.mass / R(1)) * md.center;
                I += md.I;
            }
        }
        const V2<R> center = (mass > R(0)) ? wc / (mass / R(1)) : V2<R>{};
        const Xf<R> old = b.xf;
        if (!(b.flags & kAccelerable)) {
            b.invMass = R(0);
            b.invRotI = R(0);
            if (!(b.flags & kSpeedable)) {
                // Sweep{Position{GetLocation(body), GetAngle(body)}}: local centre zero.
                const Pos<R> p{b.xf.p, b.pos1.ang};
                b.pos0 = p;
                b.pos1 = p;
                b.lc = V2<R>{};
                b.alpha0 = R(0);
                b.xf = GetXf(b.pos1, b.lc);
            }
            TransformationChanged(bid, old);
            return;
        }
        const R m = (mass > R(0)) ? mass : R(1);
        const R invMass = R(1) / m;
        R invRotI = R(0) / R(1);
        if ((I > R(0)) && !(b.flags & kFixedRotation)) {
            const R lsq = MagSq(center);
            invRotI = R(1) / (I - (m * lsq) / R(1));
        }
        b.invMass = invMass;
        b.invRotI = invRotI;
        const V2<R> oldCenter = b.pos1.lin;
        const Pos<R> p{Transform(center, b.xf), b.pos1.ang};
        b.pos0 = p;
        b.pos1 = p;
        b.lc = center;
        b.alpha0 = R(0);
        b.xf = GetXf(b.pos1, b.lc);
        const V2<R> dc = b.pos1.lin - oldCenter;
        b.vel.lin = b.vel.lin + RevPerp(dc) * (b.vel.ang / R(1));
        TransformationChanged(bid, old);
    }

    // SetType(World&, BodyID, BodyType, resetMassData = true) (WorldBody.cpp:236-246):
    // Body::SetType (Body.cpp:124-149), SetBody (AabbTreeWorld.cpp:2723-2795)
    // for the new type (the body's contacts destroyed in its list's order; a
    // static body queued for synchronisation), then ResetMassData.
    // type: 0 static, 1 kinematic, 2 dynamic.
    void SetType(std::uint32_t bid, std::uint32_t type) {
        Body<R>& b = bodies[bid];
        if (b.type == type) return;
        b.type = type;
        b.flags &= ~(kImpenetrable | kSpeedable | kAccelerable);
        b.flags |= (type == 2) ? (kSpeedable | kAccelerable) : (type == 1) ? (kImpenetrable | kSpeedable) : kImpenetrable;
        if (type == 2) {
            b.flags |= kAwake;
        } else if (type == 1) {
            b.flags |= kAwake;
            b.lacc = V2<R>{};
            b.aacc = R(0);
            b.invMass = R(0);
            b.invRotI = R(0);
        } else {
            b.flags &= ~kAwake;
            b.vel = Vel<R>{};
            b.pos0 = b.pos1;
            b.lacc = V2<R>{};
            b.aacc = R(0);
            b.invMass = R(0);
            b.invRotI = R(0);
        }
        b.underActive = R(0);
        FreshEvents();
        const std::vector<KeyedId> own = bodyContacts[bid];
        for (const KeyedId& kc : own) {
            for (auto it = contacts.begin(); it != contacts.end(); ++it) {
                if (it->id == kc.id) {
                    contacts.erase(it);
                    break;
                }
            }
            InternalDestroy(kc.id, bid);
        }
        bodyContacts[bid].clear();
        if (type == 0) bodiesForSync.push_back(bid);
        ResetMassData(bid);
    }

    // Listener calls made by edits between steps (a destroyed contact's
    // end-contact call) come before the next step's calls.
    bool eventsStale = false;
    void FreshEvents() {
        if (eventsStale) {
            events.clear();
            nevents = 0;
            eventsStale = false;
        }
    }

    // Step (AabbTreeWorld.cpp:2144-2219); the TOI phase when conf.toi.
    StepStats Step(const StepConf<R>& conf) {
        StepStats st;
        FreshEvents();
        for (const auto& fs : fixturesForProxies) {
            CreateProxies(fs.first, fs.second, conf);
            ++st.preProxiesCreated;
        }
        fixturesForProxies.clear();
        for (std::uint32_t bid : bodiesForSync) st.preProxiesMoved += Synchronize(bid, conf);
        bodiesForSync.clear();
        st.preDestroyed = DestroyContacts();
        st.preUpdated = UpdateContacts(conf, &st.preSkipped);
        st.preAdded = FindAndAddContacts(conf);
        if (conf.dt != R(0)) {
            SolveReg(conf, st);
            if (conf.toi) SolveToiPhase(conf, st);
        }
        eventsStale = true;
        return st;
    }

    // Adds a body described as in the dump format; its shapes are queued for
    // proxy creation as CreateBody does (AabbTreeWorld.cpp:1019-1044).
    std::uint32_t AddBody(const Body<R>& b) {
        const std::uint32_t id = std::uint32_t(bodies.size());
        bodies.push_back(b);
        bodies.back().proxies.clear();
        bodies.back().xf = GetXf(b.pos1, b.lc);
        bodyOrder.push_back(id);
        bodyContacts.emplace_back();
        bodyJoints.emplace_back();
        if (b.flags & kEnabled) {
            for (std::uint32_t s : b.shapes) fixturesForProxies.emplace_back(id, s);
        }
        return id;
    }

    // CreateJoint + AabbTreeWorld::Add (AabbTreeWorld.cpp:1132-1175); joints
    // are created before any contact exists, so no contact needs filtering.
    std::uint32_t AddJoint(const Joint<R>& j) {
        const std::uint32_t id = std::uint32_t(joints.size());
        joints.push_back(j);
        if (j.bodyA != kNoBody) bodyJoints[j.bodyA].emplace_back(j.bodyB, id);
        bodyJoints[j.bodyB].emplace_back(j.bodyA, id);
        return id;
    }
};

}  // namespace oracle
