// Narrowphase on the device: CollideShapes for disk / edge / polygon children
// (every child is a vertex list with forward normals and one vertex radius).
// One lane computes one contact's manifold; the arithmetic follows the
// reference line by line so that manifold points and ContactFeature ids are
// bit-identical.
//
//   CollideShapes          Manifold.cpp:393-438
//   polygon-polygon        Manifold.cpp:168-291 (clip: :61-102, :143-156;
//                          most anti-parallel edge :122-141)
//   polygon-circle         Manifold.cpp:293-372
//   circle-circle          Manifold.cpp:374-386
//   max separation         ShapeSeparation.cpp:39-128 (4x4 variant :66-104)
#pragma once

#include "devmath.hpp"

namespace prxd {

// ContactFeature packed as typeA | typeB << 1 | indexA << 8 | indexB << 16.
constexpr std::uint32_t kVertex = 0, kFace = 1, kInvalidVertex = 255;
constexpr std::uint32_t kDefaultCf = (255u << 8) | (255u << 16);
__device__ __forceinline__ std::uint32_t make_cf(std::uint32_t tA, std::uint32_t iA, std::uint32_t tB,
                                                 std::uint32_t iB) {
    return tA | (tB << 1) | ((iA & 0xffu) << 8) | ((iB & 0xffu) << 16);
}
__device__ __forceinline__ std::uint32_t flip_cf(std::uint32_t c) {
    return make_cf((c >> 1) & 1u, (c >> 16) & 0xffu, c & 1u, (c >> 8) & 0xffu);
}

enum : std::uint32_t { kUnset = 0, kCircles = 1, kFaceA = 2, kFaceB = 3 };

// Vertex counts up to which a shape's vertices and normals are loaded into
// registers in one batch before the separation loops (static indices, so no
// load waits inside the loops); larger polygons take a loop with a load per
// iteration.
constexpr std::uint32_t kStagedVerts = 8;

struct MPoint {
    V2 lp;
    std::uint32_t cf;
    float ni, ti;
};

struct Manifold {
    std::uint32_t type, n;
    V2 ln, lpt;
    MPoint p[2];
    __device__ __forceinline__ void add(V2 lp, std::uint32_t cf) {
        // Static indices only, so the points stay in registers.
        if (n == 0) {
            p[0] = MPoint{lp, cf, 0.0f, 0.0f};
        } else {
            p[1] = MPoint{lp, cf, 0.0f, 0.0f};
        }
        ++n;
    }
};

__device__ __forceinline__ Manifold empty_manifold() {
    Manifold m;
    m.type = kUnset;
    m.n = 0;
    m.ln = V2{0, 0};
    m.lpt = V2{0, 0};
    m.p[0] = MPoint{V2{0, 0}, kDefaultCf, 0, 0};
    m.p[1] = MPoint{V2{0, 0}, kDefaultCf, 0, 0};
    return m;
}
__device__ __forceinline__ Manifold for_circles(V2 vA, std::uint32_t iA, V2 vB, std::uint32_t iB) {
    Manifold m = empty_manifold();
    m.type = kCircles;
    m.lpt = vA;
    m.add(vB, make_cf(kVertex, iA, kVertex, iB));
    return m;
}
__device__ __forceinline__ Manifold for_face(bool faceB, V2 ln, V2 lp) {
    Manifold m = empty_manifold();
    m.type = faceB ? kFaceB : kFaceA;
    m.ln = ln;
    m.lpt = lp;
    return m;
}

struct Proxy {
    const float2* v;
    const float2* nrm;
    std::uint32_t n;
    float radius;
    __device__ __forceinline__ V2 vert(std::uint32_t i) const {
        const float2 t = v[i];
        return V2{t.x, t.y};
    }
    __device__ __forceinline__ V2 norm(std::uint32_t i) const {
        const float2 t = nrm[i];
        return V2{t.x, t.y};
    }
};

__device__ __forceinline__ std::uint32_t mod_next(std::uint32_t i, std::uint32_t n) { return (i + 1 < n) ? i + 1 : 0; }
__device__ __forceinline__ std::uint32_t mod_prev(std::uint32_t i, std::uint32_t n) { return i ? i - 1 : n - 1; }

struct ClipV {
    V2 v;
    std::uint32_t cf;
};
struct ClipList {
    ClipV e[2];
    std::uint32_t n;
};

__device__ __forceinline__ ClipList clip_segment(const ClipList& in, V2 normal, float offset, std::uint32_t indexA) {
    ClipList out;
    out.n = 0;
    out.e[0] = in.e[0];
    out.e[1] = in.e[1];
    if (in.n == 2) {
        const float d0 = dot(normal, in.e[0].v) - offset;
        const float d1 = dot(normal, in.e[1].v) - offset;
        const bool k0 = d0 <= 0.0f || almost_zero(d0);
        const bool k1 = d1 <= 0.0f || almost_zero(d1);
        // Kept vertices in order: e[0] if k0, then e[1] if k1.
        if (!k0 && k1) out.e[0] = in.e[1];
        out.n = (k0 ? 1u : 0u) + (k1 ? 1u : 0u);
        if (out.n < 2 && (__builtin_signbit(d0) != 0) != (__builtin_signbit(d1) != 0)) {
            const float interp = d0 / (d0 - d1);
            const V2 vertex = in.e[0].v + mul(in.e[1].v - in.e[0].v, interp);
            const ClipV cv{vertex, make_cf(kVertex, indexA, kFace, (in.e[0].cf >> 16) & 0xffu)};
            if (out.n == 0) out.e[0] = cv;
            else out.e[1] = cv;
            ++out.n;
        }
    }
    return out;
}

__device__ __forceinline__ Manifold poly_manifold(bool flipped, const Proxy& s0, const Xf& xf0, std::uint32_t idx0,
                                  const Proxy& s1, const Xf& xf1, std::uint32_t first, std::uint32_t second,
                                  float maxCirclesRatio) {
    const float r0 = s0.radius;
    const float totalRadius = r0 + s1.radius;
    const std::uint32_t idx0Next = mod_next(idx0, s0.n);
    const V2 rv0 = s0.vert(idx0);
    const V2 rv1 = s0.vert(idx0Next);
    const V2 av0 = transform(rv0, xf0);
    const V2 av1 = transform(rv1, xf0);
    const float lenEdge0 = magsq(rv1 - rv0);
    const V2 rn0 = s0.norm(idx0);
    const V2 re0dir = revperp(rn0);
    std::uint32_t e1a, e1b;
    if (second == kInvalidVertex) {
        const V2 normal = invrotate(rotate(rn0, xf0.q), xf1.q);
        const std::uint32_t prev = mod_prev(first, s1.n);
        const float prevDot = dot(normal, s1.norm(prev));
        const float currDot = dot(normal, s1.norm(first));
        if (prevDot < currDot) {
            e1a = prev;
            e1b = first;
        } else {
            e1a = first;
            e1b = mod_next(first, s1.n);
        }
    } else if ((second > first) && ((first + 1) == second)) {
        e1a = first;
        e1b = second;
    } else {
        e1a = second;
        e1b = first;
    }
    const V2 s1rv0 = s1.vert(e1a);
    const V2 s1av0 = transform(s1rv0, xf1);
    const V2 s1rv1 = s1.vert(e1b);
    const V2 s1av1 = transform(s1rv1, xf1);
    {
        const V2 ae0dir = rotate(re0dir, xf0.q);
        ClipList ie;
        ie.n = 2;
        ie.e[0] = ClipV{s1av0, make_cf(kFace, idx0, kVertex, e1a)};
        ie.e[1] = ClipV{s1av1, make_cf(kFace, idx0, kVertex, e1b)};
        const float dp0 = -dot(ae0dir, av0);
        const float dp1 = +dot(ae0dir, av1);
        const ClipList pts = clip_segment(ie, -ae0dir, dp0, idx0);
        const ClipList clip = clip_segment(pts, ae0dir, dp1, idx0Next);
        if (clip.n == 2) {
            const V2 absNormal = fwdperp(ae0dir);
            const V2 relMid = divs(rv0 + rv1, 2.0f);
            const float absOffset = dot(absNormal, av0);
            const V2 normal = fwdperp(re0dir);
            Manifold m = for_face(flipped, normal, relMid);
#pragma unroll
            for (std::uint32_t k = 0; k < 2; ++k) {
                if ((dot(absNormal, clip.e[k].v) - absOffset) <= totalRadius) {
                    m.add(invtransform(clip.e[k].v, xf1), flipped ? flip_cf(clip.e[k].cf) : clip.e[k].cf);
                }
            }
            if (m.n > 0) return m;
        }
    }
    const float trSq = totalRadius * totalRadius;
    const float cr = maxCirclesRatio * r0;
    const bool mustFace = lenEdge0 > cr * cr;
    V2 v0rel, v1rel;
    std::uint32_t i0, i1;
    if (magsq(av0 - s1av0) <= trSq) {
        v0rel = rv0; i0 = idx0; v1rel = s1rv0; i1 = e1a;
    } else if (magsq(av1 - s1av1) <= trSq) {
        v0rel = rv1; i0 = idx0Next; v1rel = s1rv1; i1 = e1b;
    } else if (magsq(av0 - s1av1) <= trSq) {
        v0rel = rv0; i0 = idx0; v1rel = s1rv1; i1 = e1b;
    } else if (magsq(av1 - s1av0) <= trSq) {
        v0rel = rv1; i0 = idx0Next; v1rel = s1rv0; i1 = e1a;
    } else {
        return empty_manifold();
    }
    if (mustFace) {
        Manifold m = for_face(flipped, fwdperp(re0dir), v0rel);
        m.add(v1rel, flipped ? make_cf(kVertex, i1, kFace, i0) : make_cf(kFace, i0, kVertex, i1));
        return m;
    }
    return !flipped ? for_circles(v0rel, i0, v1rel, i1) : for_circles(v1rel, i1, v0rel, i0);
}

__device__ __forceinline__ Manifold poly_circle_manifold(bool flipped, float totalRadius, const Proxy& s, const Xf& sxf,
                                         V2 point, const Xf& xf) {
    const std::uint32_t n = s.n;
    const V2 cLocal = invtransform(transform(point, xf), sxf);
    std::uint32_t iMax = 0;
    float maxSep = -3.40282347e+38f;
    if (n <= kStagedVerts) {
        // Vertices and normals loaded in one batch (static indices).
        V2 sv[kStagedVerts], sn[kStagedVerts];
#pragma unroll
        for (std::uint32_t i = 0; i < kStagedVerts; ++i) {
            sv[i] = (i < n) ? s.vert(i) : V2{0, 0};
            sn[i] = (i < n) ? s.norm(i) : V2{0, 0};
        }
        bool apart = false;
#pragma unroll
        for (std::uint32_t i = 0; i < kStagedVerts; ++i) {
            if (i < n && !apart) {
                const float sep = dot(sn[i], cLocal - sv[i]);
                if (sep > totalRadius) {
                    apart = true;
                } else if (maxSep < sep) {
                    maxSep = sep;
                    iMax = i;
                }
            }
        }
        if (apart) return empty_manifold();
    } else {
        for (std::uint32_t i = 0; i < n; ++i) {
            const float sep = dot(s.norm(i), cLocal - s.vert(i));
            if (sep > totalRadius) return empty_manifold();
            if (maxSep < sep) {
                maxSep = sep;
                iMax = i;
            }
        }
    }
    const std::uint32_t iMax2 = mod_next(iMax, n);
    const V2 v1 = s.vert(iMax);
    const V2 v2 = s.vert(iMax2);
    bool face = false;
    V2 fc;
    if (maxSep < 0.0f) {
        face = true;
        fc = divs(v1 + v2, 2.0f);
    } else {
        const V2 cv1 = cLocal - v1;
        if (dot(cv1, v2 - v1) <= 0.0f) {
            if (magsq(cv1) > totalRadius * totalRadius) return empty_manifold();
            return !flipped ? for_circles(v1, iMax, point, 0u) : for_circles(point, 0u, v1, iMax);
        }
        const V2 cv2 = cLocal - v2;
        if (dot(cv2, v1 - v2) <= 0.0f) {
            if (magsq(cv2) > totalRadius * totalRadius) return empty_manifold();
            return !flipped ? for_circles(v2, iMax2, point, 0u) : for_circles(point, 0u, v2, iMax2);
        }
        fc = divs(v1 + v2, 2.0f);
        if (dot(cLocal - fc, s.norm(iMax)) > totalRadius) return empty_manifold();
        face = true;
    }
    Manifold m = for_face(flipped, s.norm(iMax), fc);
    m.add(point, flipped ? make_cf(kVertex, 0, kFace, iMax) : make_cf(kFace, iMax, kVertex, 0));
    (void)face;
    return m;
}

struct SepInfo {
    float d;
    std::uint32_t first, s0, s1;
};

// Minimum separation of vertices from a face (ShapeSeparation.cpp:39-64).
__device__ __forceinline__ void min_sep_step(float s, std::uint32_t i, float& minSep, std::uint32_t& first,
                                             std::uint32_t& second) {
    if (minSep > s) {
        minSep = s;
        first = i;
        second = kInvalidVertex;
    } else if (minSep == s) {
        second = i;
    }
}

__device__ __forceinline__ SepInfo max_sep_4x4(const Proxy& p1, const Xf& xf1, const Proxy& p2, const Xf& xf2) {
    SepInfo out{-__builtin_inff(), kInvalidVertex, kInvalidVertex, kInvalidVertex};
    const Xf xf = mulT(xf1, xf2);
    V2 verts[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) verts[k] = transform(p2.vert(std::uint32_t(k)), xf);
#pragma unroll
    for (std::uint32_t i = 0; i < 4; ++i) {
        const V2 origin = p1.vert(i);
        const V2 normal = p1.norm(i);
        float d = __builtin_inff();
        std::uint32_t f = kInvalidVertex, s = kInvalidVertex;
#pragma unroll
        for (std::uint32_t k = 0; k < 4; ++k) min_sep_step(dot(normal, verts[k] - origin), k, d, f, s);
        if (out.d < d) {
            out.d = d;
            out.s0 = f;
            out.s1 = s;
            out.first = i;
        }
    }
    return out;
}

__device__ __forceinline__ SepInfo max_sep_staged(const Proxy& p1, const Xf& xf1, const Proxy& p2, const Xf& xf2) {
    SepInfo out{-__builtin_inff(), kInvalidVertex, kInvalidVertex, kInvalidVertex};
    const Xf xf = mulT(xf2, xf1);
    V2 v2[kStagedVerts], v1[kStagedVerts], n1[kStagedVerts];
#pragma unroll
    for (std::uint32_t k = 0; k < kStagedVerts; ++k) {
        v2[k] = (k < p2.n) ? p2.vert(k) : V2{0, 0};
        v1[k] = (k < p1.n) ? p1.vert(k) : V2{0, 0};
        n1[k] = (k < p1.n) ? p1.norm(k) : V2{0, 0};
    }
#pragma unroll
    for (std::uint32_t i = 0; i < kStagedVerts; ++i) {
        if (i < p1.n) {
            const V2 origin = transform(v1[i], xf);
            const V2 normal = rotate(n1[i], xf.q);
            float d = __builtin_inff();
            std::uint32_t f = kInvalidVertex, s = kInvalidVertex;
#pragma unroll
            for (std::uint32_t k = 0; k < kStagedVerts; ++k)
                if (k < p2.n) min_sep_step(dot(normal, v2[k] - origin), k, d, f, s);
            if (out.d < d) {
                out.d = d;
                out.s0 = f;
                out.s1 = s;
                out.first = i;
            }
        }
    }
    return out;
}

__device__ __forceinline__ SepInfo max_sep(const Proxy& p1, const Xf& xf1, const Proxy& p2, const Xf& xf2) {
    if (p1.n <= kStagedVerts && p2.n <= kStagedVerts) return max_sep_staged(p1, xf1, p2, xf2);
    SepInfo out{-__builtin_inff(), kInvalidVertex, kInvalidVertex, kInvalidVertex};
    const Xf xf = mulT(xf2, xf1);
    for (std::uint32_t i = 0; i < p1.n; ++i) {
        const V2 origin = transform(p1.vert(i), xf);
        const V2 normal = rotate(p1.norm(i), xf.q);
        float d = __builtin_inff();
        std::uint32_t f = kInvalidVertex, s = kInvalidVertex;
        for (std::uint32_t k = 0; k < p2.n; ++k) min_sep_step(dot(normal, p2.vert(k) - origin), k, d, f, s);
        if (out.d < d) {
            out.d = d;
            out.s0 = f;
            out.s1 = s;
            out.first = i;
        }
    }
    return out;
}

__device__ __forceinline__ Manifold collide_shapes(const Proxy& A, const Xf& xfA, const Proxy& B, const Xf& xfB, float linearSlop,
                                   float maxCirclesRatio) {
    const float totalRadius = A.radius + B.radius;
    if (A.n == 1 && B.n == 1) {
        const V2 pA = transform(A.vert(0), xfA);
        const V2 pB = transform(B.vert(0), xfB);
        const float lenSq = magsq(pB - pA);
        const float totSq = totalRadius * totalRadius;
        if (lenSq > totSq) return empty_manifold();
        return for_circles(A.vert(0), 0u, B.vert(0), 0u);
    }
    if (A.n == 1) return poly_circle_manifold(true, totalRadius, B, xfB, A.vert(0), xfA);
    if (B.n == 1) return poly_circle_manifold(false, totalRadius, A, xfA, B.vert(0), xfB);
    const bool do4x4 = (A.n == 4) && (B.n == 4);
    const SepInfo sa = do4x4 ? max_sep_4x4(A, xfA, B, xfB) : max_sep(A, xfA, B, xfB);
    if (sa.d > totalRadius) return empty_manifold();
    const SepInfo sb = do4x4 ? max_sep_4x4(B, xfB, A, xfA) : max_sep(B, xfB, A, xfA);
    if (sb.d > totalRadius) return empty_manifold();
    const float kTol = linearSlop / 10.0f;
    return (sb.d > (sa.d + kTol)) ? poly_manifold(true, B, xfB, sb.first, A, xfA, sb.s0, sb.s1, maxCirclesRatio)
                                  : poly_manifold(false, A, xfA, sa.first, B, xfB, sa.s0, sa.s1, maxCirclesRatio);
}

}  // namespace prxd
