// TOI phase of the step (doToi = true) for one world per wavefront:
// SolveToi (AabbTreeWorld.cpp:1752-1831) with UpdateContactTOIs (:1685-1750),
// SolveToi(ContactID) (:1833-1966), ProcessContactsForTOI (:2059-2142) and
// SolveToiViaGS (:1968-2057); time of impact by conservative advancement
// (GetToiViaSat, TimeOfImpact.cpp:35-251) over GJK distance (Distance.cpp:
// 95-175, Simplex.cpp:25-241) and separation functions
// (SeparationScenario.cpp:25-220).
//
// Order: the first TOI pass of a step computes every contact's time of
// impact one lane per contact (all sweeps start at alpha0 = 0, so advancing
// a body to 0 is the same from any lane); later passes run on lane 0 in the
// contact list's order, as do the event solves, because they advance bodies
// shared between contacts.  Included by step_device.hpp.
#pragma once

namespace prxd {

// ---------------------------------------------------------------------------
// Sweeps (Sweep.hpp:48-68, Sweep.cpp:29-33, d2/Math.hpp:160-167, 251-254).
struct SweepD {
    float4 p0, p1;  // x, y, angle, -
    V2 lc;
    float a0;
};

// GetPosition(pos0, pos1, beta) = pos0 + (pos1 - pos0) * beta (d2/Math.cpp:30-59).
__device__ __forceinline__ float4 lerp_pos(float4 p0, float4 p1, float beta) {
    return float4{p0.x + (p1.x - p0.x) * beta, p0.y + (p1.y - p0.y) * beta, p0.z + (p1.z - p0.z) * beta, p0.w};
}

__device__ __forceinline__ SweepD sweep_of(const DeviceState& d, int gb) {
    const float4 l = d.lcinv[gb];
    return SweepD{d.pos0[gb], d.pos1[gb], V2{l.x, l.y}, d.alpha0[gb]};
}

// GetNormalized(Angle), the trunc form (Math.cpp:33-91).
__device__ __forceinline__ float normalize_angle(float a) {
    constexpr float kPiF = 3.14159265358979323846f;
    constexpr float kTwoPi = 2.0f * kPiF;
    constexpr float kRTwoPi = 1.0f / kTwoPi;
    a = a - __builtin_truncf(a * kRTwoPi) * kTwoPi;
    if (a >= kPiF) a -= kPiF * 2.0f;
    else if (a < -kPiF) a += kPiF * 2.0f;
    return a;
}

__device__ __forceinline__ SweepD normalized(SweepD s) {
    const float a0 = normalize_angle(s.p0.z);
    const float dd = s.p0.z - a0;
    s.p0.z = a0;
    s.p1.z -= dd;
    return s;
}

__device__ __forceinline__ Xf xf_at(const SweepD& s, float beta) {
    const float4 p = lerp_pos(s.p0, s.p1, beta);
    return get_xf(p.x, p.y, p.z, s.lc);
}

// A sweep whose two positions are equal has the same transformation at every
// beta in [0, 1]: each pos1 - pos0 is a zero whose sign does not depend on
// beta, so pos0 + (pos1 - pos0) * beta is one value.  Such sweeps (static
// bodies, bodies at rest) are transformed once per time of impact.
struct SweepXf {
    SweepD s;
    bool fixed;
    Xf xf;
    __device__ __forceinline__ explicit SweepXf(const SweepD& sw)
        : s(sw), fixed(sw.p0.x == sw.p1.x && sw.p0.y == sw.p1.y && sw.p0.z == sw.p1.z), xf(xf_at(sw, 0.0f)) {}
    __device__ __forceinline__ Xf at(float beta) const { return fixed ? xf : xf_at(s, beta); }
};

// Body::Advance0 / Advance / SetSweep (Body.hpp:697-727, 1004-1009).
__device__ __forceinline__ void body_advance0(const DeviceState& d, int gb, float alpha) {
    const float a0 = d.alpha0[gb];
    const float beta = (alpha - a0) / (1.0f - a0);
    d.pos0[gb] = lerp_pos(d.pos0[gb], d.pos1[gb], beta);
    d.alpha0[gb] = alpha;
}

__device__ __forceinline__ void body_advance(const DeviceState& d, int gb, float alpha) {
    body_advance0(d, gb, alpha);
    const float4 p = d.pos0[gb];
    d.pos1[gb] = p;
    const float4 l = d.lcinv[gb];
    const Xf xf = get_xf(p.x, p.y, p.z, V2{l.x, l.y});
    d.xf[gb] = float4{xf.p.x, xf.p.y, xf.q.x, xf.q.y};
}

__device__ __forceinline__ void set_sweep(const DeviceState& d, int gb, const SweepD& s) {
    d.pos0[gb] = s.p0;
    d.pos1[gb] = s.p1;
    d.alpha0[gb] = s.a0;
    const Xf xf = get_xf(s.p1.x, s.p1.y, s.p1.z, s.lc);
    d.xf[gb] = float4{xf.p.x, xf.p.y, xf.q.x, xf.q.y};
}

__device__ __forceinline__ bool pos_differ(float4 a, float4 b) { return !(a.x == b.x && a.y == b.y) || a.z != b.z; }

// ---------------------------------------------------------------------------
// GJK distance.

// GetSupportIndex (DistanceProxy.hpp:233-249).  Where every dot product is
// NaN the reference returns InvalidVertex and then reads past the vertices;
// here vertex 0 stands in, so no access leaves the shape.
__device__ __forceinline__ std::uint32_t support_index(const Proxy& p, V2 dir) {
    std::uint32_t index = 0u;
    float maxValue = -__builtin_inff();
    for (std::uint32_t i = 0; i < p.n; ++i) {
        const float value = dot(p.vert(i), dir);
        if (maxValue < value) {
            maxValue = value;
            index = i;
        }
    }
    return index;
}

struct SEdge {
    V2 wA, wB;
    std::uint32_t iA, iB;
};
struct Simplex {
    SEdge e[3];
    float c[3];
    std::uint32_t n;
};

__device__ __forceinline__ V2 pdelta(const SEdge& s) { return s.wB - s.wA; }

__device__ __forceinline__ Simplex simplex1(const SEdge& a) {
    Simplex r;
    r.e[0] = a;
    r.c[0] = 1.0f;
    r.n = 1;
    return r;
}

__device__ __forceinline__ Simplex simplex2(const SEdge& a, const SEdge& b, float ca, float cb) {
    Simplex r;
    r.e[0] = a;
    r.e[1] = b;
    r.c[0] = ca;
    r.c[1] = cb;
    r.n = 2;
    return r;
}

// Simplex::Get (Simplex.cpp:62-175).
__device__ __forceinline__ Simplex solve_simplex(const SEdge* e, std::uint32_t n) {
    if (n == 1) return simplex1(e[0]);
    if (n == 2) {
        const V2 w1 = pdelta(e[0]), w2 = pdelta(e[1]);
        const V2 e12 = w2 - w1;
        const float d12_2 = -dot(w1, e12);
        if (d12_2 <= 0.0f) return simplex1(e[0]);
        const float d12_1 = dot(w2, e12);
        if (d12_1 <= 0.0f) return simplex1(e[1]);
        const float inv = 1.0f / (d12_1 + d12_2);
        return simplex2(e[0], e[1], d12_1 * inv, d12_2 * inv);
    }
    const V2 w1 = pdelta(e[0]), w2 = pdelta(e[1]), w3 = pdelta(e[2]);
    const V2 e12 = w2 - w1;
    const float d12_1 = dot(w2, e12);
    const float d12_2 = -dot(w1, e12);
    const V2 e13 = w3 - w1;
    const float d13_1 = dot(w3, e13);
    const float d13_2 = -dot(w1, e13);
    const V2 e23 = w3 - w2;
    const float d23_1 = dot(w3, e23);
    const float d23_2 = -dot(w2, e23);
    if (d12_2 <= 0.0f && d13_2 <= 0.0f) return simplex1(e[0]);
    if (d12_1 <= 0.0f && d23_2 <= 0.0f) return simplex1(e[1]);
    if (d13_1 <= 0.0f && d23_1 <= 0.0f) return simplex1(e[2]);
    const float n123 = cross(e12, e13);
    const float d123_3 = n123 * cross(w1, w2);
    if (d12_1 > 0.0f && d12_2 > 0.0f && d123_3 <= 0.0f) {
        const float inv = 1.0f / (d12_1 + d12_2);
        return simplex2(e[0], e[1], d12_1 * inv, d12_2 * inv);
    }
    const float d123_2 = n123 * cross(w3, w1);
    if (d13_1 > 0.0f && d13_2 > 0.0f && d123_2 <= 0.0f) {
        const float inv = 1.0f / (d13_1 + d13_2);
        return simplex2(e[0], e[2], d13_1 * inv, d13_2 * inv);
    }
    const float d123_1 = n123 * cross(w2, w3);
    if (d23_1 > 0.0f && d23_2 > 0.0f && d123_1 <= 0.0f) {
        const float inv = 1.0f / (d23_1 + d23_2);
        return simplex2(e[2], e[1], d23_2 * inv, d23_1 * inv);
    }
    const float inv = 1.0f / (d123_1 + d123_2 + d123_3);
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

// Simplex::CalcMetric (Simplex.cpp:208-228).
__device__ __forceinline__ float simplex_metric(const SEdge* e, std::uint32_t n) {
    if (n == 2) return __builtin_sqrtf(magsq(pdelta(e[1]) - pdelta(e[0])));
    if (n == 3) return cross(pdelta(e[1]) - pdelta(e[0]), pdelta(e[2]) - pdelta(e[0]));
    return 0.0f;
}

// Simplex::Cache (Simplex.hpp:96-101): metric and index pairs (255: none).
struct SimplexCache {
    float metric;
    std::uint32_t ia[3], ib[3];
    __device__ __forceinline__ std::uint32_t count() const {
        std::uint32_t k = 0;
        for (int i = 0; i < 3; ++i) k += (ia[i] == 255u && ib[i] == 255u) ? 0u : 1u;
        return k;
    }
};

__device__ __forceinline__ SimplexCache empty_cache() {
    SimplexCache c;
    c.metric = __builtin_nanf("");
    for (int i = 0; i < 3; ++i) c.ia[i] = c.ib[i] = 255u;
    return c;
}

__device__ __forceinline__ SimplexCache cache_of(const Simplex& s) {
    SimplexCache c = empty_cache();
    c.metric = simplex_metric(s.e, s.n);
    // Constant indices after unrolling keep the simplex in registers.
#pragma unroll
    for (std::uint32_t i = 0; i < 3; ++i) {
        if (i < s.n) {
            c.ia[i] = s.e[i].iA;
            c.ib[i] = s.e[i].iB;
        }
    }
    return c;
}

__device__ __forceinline__ SEdge edge_of(const Proxy& pA, const Xf& xfA, std::uint32_t iA, const Proxy& pB,
                                         const Xf& xfB, std::uint32_t iB) {
    return SEdge{transform(pA.vert(iA), xfA), transform(pB.vert(iB), xfB), iA, iB};
}

enum : int { kDistHitMaxIters = 0, kDistMaxPoints, kDistUnfitSearchDir, kDistDuplicateIndexPair };

struct DistOut {
    Simplex s;
    int state;
    std::uint32_t iters;
};

// Distance (Distance.cpp:95-175).
__device__ __forceinline__ DistOut gjk_distance(const Proxy& pA, const Xf& xfA, const Proxy& pB, const Xf& xfB,
                                             const SimplexCache& cache, std::uint32_t maxIters) {
    std::uint32_t sa[3] = {cache.ia[0], cache.ia[1], cache.ia[2]};
    std::uint32_t sb[3] = {cache.ib[0], cache.ib[1], cache.ib[2]};
    SEdge e[3];
    std::uint32_t n = cache.count();
#pragma unroll
    for (std::uint32_t i = 0; i < 3; ++i)
        if (i < n) e[i] = edge_of(pA, xfA, sa[i], pB, xfB, sb[i]);
    if (n > 1) {
        const float metric1 = cache.metric;
        const float metric2 = simplex_metric(e, n);
        if ((metric2 < (metric1 / 2.0f)) || (metric2 > (metric1 * 2.0f)) || (metric2 < 0.0f) || almost_zero(metric2))
            n = 0;
    }
    if (n == 0) {
        e[0] = edge_of(pA, xfA, 0u, pB, xfB, 0u);
        n = 1;
        sa[0] = sb[0] = 0u;
        sa[1] = sb[1] = sa[2] = sb[2] = 255u;
    }
    Simplex simplex = simplex1(e[0]);
    int state = kDistHitMaxIters;
    std::uint32_t iter = 0;
    while (iter < maxIters) {
        ++iter;
        simplex = solve_simplex(e, n);
        n = simplex.n;
#pragma unroll
        for (std::uint32_t i = 0; i < 3; ++i)
            if (i < n) e[i] = simplex.e[i];
        if (n == 3) {
            state = kDistMaxPoints;
            break;
        }
        // CalcSearchDirection (Simplex.cpp:44-60).
        V2 dir;
        if (n == 1) {
            dir = -pdelta(e[0]);
        } else {
            const V2 e12 = pdelta(e[1]) - pdelta(e[0]);
            const V2 e0 = pdelta(e[0]);
            const float sgn = cross(e12, -e0);
            dir = (sgn > 0.0f) ? revperp(e12) : fwdperp(e12);
        }
        if (almost_zero(magsq(dir))) {
            state = kDistUnfitSearchDir;
            break;
        }
        const std::uint32_t iA = support_index(pA, invrotate(-dir, xfA.q));
        const std::uint32_t iB = support_index(pB, invrotate(dir, xfB.q));
        bool dup = false;
        for (int k = 0; k < 3; ++k) dup = dup || (sa[k] == iA && sb[k] == iB);
        if (dup) {
            state = kDistDuplicateIndexPair;
            break;
        }
        const SEdge added = edge_of(pA, xfA, iA, pB, xfB, iB);
#pragma unroll
        for (std::uint32_t k = 0; k < 3; ++k)
            if (k == n) e[k] = added;
        ++n;
#pragma unroll
        for (std::uint32_t k = 0; k < 3; ++k) {
            sa[k] = k < n ? e[k].iA : 255u;
            sb[k] = k < n ? e[k].iB : 255u;
        }
    }
    return DistOut{simplex, state, iter};
}

// GetDelta(GetWitnessPoints(simplex)) (Distance.cpp:75-93, Distance.hpp:55-58).
__device__ __forceinline__ V2 witness_delta(const Simplex& s) {
    V2 a{0.0f, 0.0f}, b{0.0f, 0.0f};
#pragma unroll
    for (std::uint32_t i = 0; i < 3; ++i) {
        if (i < s.n) {
            a = a + mul(s.e[i].wA, s.c[i]);
            b = b + mul(s.e[i].wB, s.c[i]);
        }
    }
    return b - a;
}

// ---------------------------------------------------------------------------
// Separation functions (SeparationScenario.cpp:25-220).
enum : int { kScenPoints = 0, kScenFaceA, kScenFaceB };

struct Scenario {
    int type;
    V2 axis, lp;
};

__device__ __forceinline__ Scenario get_scenario(const SimplexCache& c, const Proxy& pA, const Xf& xfA,
                                                 const Proxy& pB, const Xf& xfB) {
    const V2 zero{0.0f, 0.0f};
    if (c.count() >= 2) {
        if (c.ia[0] == c.ia[1]) {
            const V2 b1 = pB.vert(c.ib[0]), b2 = pB.vert(c.ib[1]);
            const V2 axis = unit_vector(fwdperp(b2 - b1), zero);
            const V2 normal = rotate(axis, xfB.q);
            const V2 lp = divs(b1 + b2, 2.0f);
            const V2 delta = transform(pA.vert(c.ia[0]), xfA) - transform(lp, xfB);
            return Scenario{kScenFaceB, (dot(delta, normal) < 0.0f) ? -axis : axis, lp};
        }
        const V2 a1 = pA.vert(c.ia[0]), a2 = pA.vert(c.ia[1]);
        const V2 axis = unit_vector(fwdperp(a2 - a1), zero);
        const V2 normal = rotate(axis, xfA.q);
        const V2 lp = divs(a1 + a2, 2.0f);
        const V2 delta = transform(pB.vert(c.ib[0]), xfB) - transform(lp, xfA);
        return Scenario{kScenFaceA, (dot(delta, normal) < 0.0f) ? -axis : axis, lp};
    }
    const V2 pointA = transform(pA.vert(c.ia[0]), xfA);
    const V2 pointB = transform(pB.vert(c.ib[0]), xfB);
    return Scenario{kScenPoints, unit_vector(pointB - pointA, zero), zero};
}

struct SepIdx {
    float distance;
    std::uint32_t ia, ib;
};

__device__ __forceinline__ SepIdx find_min_sep(const Scenario& f, const Proxy& pA, const Xf& xfA, const Proxy& pB,
                                               const Xf& xfB) {
    if (f.type == kScenPoints) {
        const std::uint32_t ia = support_index(pA, invrotate(f.axis, xfA.q));
        const std::uint32_t ib = support_index(pB, invrotate(-f.axis, xfB.q));
        const V2 delta = transform(pB.vert(ib), xfB) - transform(pA.vert(ia), xfA);
        return SepIdx{dot(delta, f.axis), ia, ib};
    }
    if (f.type == kScenFaceA) {
        const V2 normal = rotate(f.axis, xfA.q);
        const V2 pointA = transform(f.lp, xfA);
        const std::uint32_t ib = support_index(pB, invrotate(-normal, xfB.q));
        const V2 delta = transform(pB.vert(ib), xfB) - pointA;
        return SepIdx{dot(delta, normal), 255u, ib};
    }
    const V2 normal = rotate(f.axis, xfB.q);
    const std::uint32_t ia = support_index(pA, invrotate(-normal, xfA.q));
    const V2 pointA = transform(pA.vert(ia), xfA);
    const V2 delta = pointA - transform(f.lp, xfB);
    return SepIdx{dot(delta, normal), ia, 255u};
}

__device__ __forceinline__ float evaluate_sep(const Scenario& f, const Proxy& pA, const Xf& xfA, const Proxy& pB,
                                              const Xf& xfB, std::uint32_t ia, std::uint32_t ib) {
    if (f.type == kScenPoints) {
        const V2 delta = transform(pB.vert(ib), xfB) - transform(pA.vert(ia), xfA);
        return dot(delta, f.axis);
    }
    if (f.type == kScenFaceA) {
        const V2 normal = rotate(f.axis, xfA.q);
        const V2 pointA = transform(f.lp, xfA);
        const V2 delta = transform(pB.vert(ib), xfB) - pointA;
        return dot(delta, normal);
    }
    const V2 normal = rotate(f.axis, xfB.q);
    const V2 pointB = transform(f.lp, xfB);
    const V2 delta = transform(pA.vert(ia), xfA) - pointB;
    return dot(delta, normal);
}

// ToiOutput (ToiOutput.hpp:44-70): the time when the state is e_touching (a
// negative value otherwise, for which the caller uses 1) and the iteration
// statistics.
struct ToiRes {
    float time;
    std::uint32_t toiIters, maxDist, maxRoot;
};

// GetToiViaSat (TimeOfImpact.cpp:35-251) with the step's ToiConf
// (ToiConf.cpp:26-35, timeMax 1).
__device__ __forceinline__ ToiRes toi_via_sat(const Proxy& pA, const SweepD& sA, const Proxy& pB, const SweepD& sB,
                                              const StepParams& c) {
    ToiRes r{-1.0f, 0u, 0u, 0u};
    const float totalRadius = pA.radius + pB.radius;
    if (c.targetDepth > totalRadius) return r;
    const float target = totalRadius - c.targetDepth;
    const float maxTarget = fmax_std(target + c.tolerance, 0.0f);
    const float minTarget = fmax_std(target - c.tolerance, 0.0f);
    const float minTargetSq = minTarget * minTarget;
    if (!__builtin_isfinite(minTargetSq) && __builtin_isfinite(minTarget)) return r;
    const float maxTargetSq = maxTarget * maxTarget;
    if (!__builtin_isfinite(maxTargetSq) && __builtin_isfinite(maxTarget)) return r;
    auto done = [&](float t) {
        r.time = t;
        return r;
    };
    float timeLo = 0.0f;
    const SweepXf wA(sA), wB(sB);
    Xf loA = wA.xf, loB = wB.xf;
    SimplexCache cache = empty_cache();
    for (std::uint32_t toiIters = 0; toiIters < c.maxToiIters;) {
        const DistOut di = gjk_distance(pA, loA, pB, loB, cache, c.maxDistanceIters);
        ++toiIters;
        r.toiIters = toiIters;
        r.maxDist = (r.maxDist < di.iters) ? di.iters : r.maxDist;
        if (di.state == kDistHitMaxIters) return done(-1.0f);
        cache = cache_of(di.s);
        const float distSq = magsq(witness_delta(di.s));
        if (distSq < minTargetSq) return done(-1.0f);  // e_overlapped
        if (distSq <= maxTargetSq) return done(timeLo);  // e_touching
        const Scenario fcn = get_scenario(cache, pA, loA, pB, loB);
        float timeHi = 1.0f;
        Xf hiA = wA.at(timeHi), hiB = wB.at(timeHi);
        for (std::uint32_t pbIter = 0; pbIter < 254u; ++pbIter) {
            const SepIdx hiMin = find_min_sep(fcn, pA, hiA, pB, hiB);
            if (hiMin.distance > maxTarget) return done(-1.0f);  // e_separated
            if (hiMin.distance >= minTarget) {
                if (timeHi == timeLo) return done(-1.0f);  // e_separated
                timeLo = timeHi;
                loA = hiA;
                loB = hiB;
                break;
            }
            const float loEval = evaluate_sep(fcn, pA, loA, pB, loB, hiMin.ia, hiMin.ib);
            if (loEval <= maxTarget) return done((loEval < minTarget) ? -1.0f : timeLo);  // e_belowMinTarget / e_touching
            float a1 = timeLo, a2 = timeHi, s1 = loEval, s2 = hiMin.distance;
            std::uint32_t roots = 0;
            for (;;) {
                if (roots == c.maxToiRootIters || nextafter_f(a1, a2) >= a2) {  // e_maxRootIters, e_nextAfter
                    r.maxRoot = (r.maxRoot < roots) ? roots : r.maxRoot;
                    return done(-1.0f);
                }
                // Secant / Bisect (Math.hpp:95-108).
                const float t = (roots & 1u) ? a1 + (target - s1) * (a2 - a1) / (s2 - s1) : (a1 + a2) / 2.0f;
                ++roots;
                const Xf tA = wA.at(t), tB = wB.at(t);
                const float sv = evaluate_sep(fcn, pA, tA, pB, tB, hiMin.ia, hiMin.ib);
                if (__builtin_fabsf(sv - target) <= c.tolerance) {
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
            r.maxRoot = (r.maxRoot < roots) ? roots : r.maxRoot;
        }
    }
    return done(-1.0f);  // e_maxToiIters
}

// ---------------------------------------------------------------------------
// Contact TOI state: c_toi the value, c_toiinfo bit 31 valid (e_toiFlag),
// low bits the sub-step count.
constexpr std::uint32_t kToiValid = 0x80000000u;

// The body of UpdateContactTOIs' loop for a contact that passed its filters;
// returns the time of impact's iteration statistics.
__device__ __forceinline__ ToiRes contact_toi(const DeviceState& d, const WorldCtx& x, int i, const StepParams& c) {
    const int gi = x.co + i;
    const int2 bd = d.c_body[gi];
    const int ga = x.bo + bd.x, gb = x.bo + bd.y;
    const float alpha0 = fmax_std(d.alpha0[ga], d.alpha0[gb]);
    body_advance0(d, ga, alpha0);
    body_advance0(d, gb, alpha0);
    const SweepD sA = normalized(sweep_of(d, ga));
    const SweepD sB = normalized(sweep_of(d, gb));
    const int2 sh = d.c_shape[gi];
    const ToiRes r = toi_via_sat(shape_proxy(d, sh.x), sA, shape_proxy(d, sh.y), sB, c);
    d.c_toi[gi] = (r.time >= 0.0f) ? fmin_std(alpha0 + (1.0f - alpha0) * r.time, 1.0f) : 1.0f;
    d.c_toiinfo[gi] |= kToiValid;
    return r;
}

// What one event adds (IslandStats, StepStats.hpp:149-163, with the contact counts).
struct ToiEventStats {
    float minSep, maxInc;
    bool solved;
    std::uint32_t posIters, velIters, updated, skipped;
};

// UpdateContactTOIs' filters: 1 compute, 2 skipped at maxSubSteps (counted), 0 skipped.
__device__ __forceinline__ int needs_toi(const DeviceState& d, const WorldCtx& x, int i, const StepParams& c) {
    const int gi = x.co + i;
    const std::uint32_t info = d.c_toiinfo[gi];
    if (info & kToiValid) return 0;
    const std::uint32_t f = d.c_flags[gi];
    if (!(f & prx::kCEnabled) || (f & prx::kSensor) || !(f & prx::kCImpenetrable)) return 0;
    if ((info & 0xffffu) >= c.maxSubSteps) return 2;
    const int2 bd = d.c_body[gi];
    return ((d.bflags[x.bo + bd.x] | d.bflags[x.bo + bd.y]) & prx::kAwake) ? 1 : 0;
}

// Listener calls made on lane 0 (the event solves): appended one by one.
__device__ __forceinline__ int ev_claim_one(const WorldCtx& x, Counters& k, int words, int calls) {
    const int at = k.evw;
    k.evw += words;
    k.evn += calls;
    if (at + words > x.ecap) {
        k.status |= prx::kStatusEventOverflow;
        return -1;
    }
    return x.eo + at;
}

__device__ __forceinline__ void ev_update_one(const DeviceState& d, const WorldCtx& x, Counters& k,
                                              std::uint32_t calls, std::uint32_t cid, const Manifold& old) {
    if (x.ecap <= 0 || calls == 0) return;
    const bool touch = (calls & 3u) != 0, pre = (calls & 4u) != 0;
    const int at = ev_claim_one(x, k, (touch ? 2 : 0) + (pre ? prx::kEvPreWords : 0), (touch ? 1 : 0) + (pre ? 1 : 0));
    if (at < 0) return;
    std::uint32_t* o = d.ev + at;
    if (touch) {
        o[0] = (calls & 1u) ? prx::kEvBegin : prx::kEvEnd;
        o[1] = cid;
        o += 2;
    }
    if (pre) {
        const bool face = old.type == kFaceA || old.type == kFaceB;
        const bool set = old.type != kUnset;
        o[0] = prx::kEvPreSolve;
        o[1] = cid;
        o[2] = old.type | (old.n << 8);
        o[3] = old.n > 0 ? old.p[0].cf : kDefaultCf;
        o[4] = old.n > 1 ? old.p[1].cf : kDefaultCf;
        o[5] = face ? __float_as_uint(old.ln.x) : 0u;
        o[6] = face ? __float_as_uint(old.ln.y) : 0u;
        o[7] = set ? __float_as_uint(old.lpt.x) : 0u;
        o[8] = set ? __float_as_uint(old.lpt.y) : 0u;
        for (std::uint32_t q = 0; q < 2; ++q) {
            const bool has = q < old.n;
            o[9 + 4 * q] = has ? __float_as_uint(old.p[q].lp.x) : 0u;
            o[10 + 4 * q] = has ? __float_as_uint(old.p[q].lp.y) : 0u;
            o[11 + 4 * q] = has ? __float_as_uint(old.p[q].ni) : 0u;
            o[12 + 4 * q] = has ? __float_as_uint(old.p[q].ti) : 0u;
        }
    }
}

// Update (AabbTreeWorld.cpp:2587-2721) of one contact from lane 0.
__device__ __forceinline__ void toi_update(const DeviceState& d, const WorldCtx& x, Counters& k, int i,
                                           const StepParams& c) {
    Manifold old = empty_manifold();
    const std::uint32_t calls = update_contact(d, x.bo, x.co + i, c, &old);
    ev_update_one(d, x, k, calls, d.c_id[x.co + i], old);
}

// FlagForUpdating for every contact of a body (the all-contacts adjacency).
__device__ __forceinline__ void flag_body_contacts(const DeviceState& d, const WorldCtx& x, const IslandScratch& S,
                                                   int b) {
    const int e1 = S.off[b + 1];
    for (int e = S.off[b]; e < e1; ++e) d.c_flags[x.co + S.csr[e].x] |= prx::kNeedsUpdating;
}

// Lists of the TOI island (lane 0): bodies in S.ibody, contacts in S.icont.
struct ToiIsland {
    int nb, nc;
};

// ProcessContactsForTOI (AabbTreeWorld.cpp:2059-2142).
__device__ __forceinline__ void toi_process_contacts(const DeviceState& d, const WorldCtx& x, Counters& k,
                                                     const IslandScratch& S, int b, ToiIsland& I, float toi,
                                                     const StepParams& c, ToiEventStats& es) {
    const bool bodyImp = (d.bflags[x.bo + b] & prx::kImpenetrable) != 0;
    const int e1 = S.off[b + 1];
    for (int e = S.off[b]; e < e1; ++e) {
        const int2 en = S.csr[e];
        const int i = en.x, other = en.y;
        const int gi = x.co + i;
        if (d.t_isl_c[gi]) continue;
        if (d.c_flags[gi] & prx::kSensor) continue;
        const int go = x.bo + other;
        if (!(bodyImp || (d.bflags[go] & prx::kImpenetrable))) continue;
        const bool otherIslanded = d.t_isl_b[go] != 0;
        const SweepD backup = sweep_of(d, go);
        if (!otherIslanded) {
            body_advance(d, go, toi);
            if (pos_differ(d.pos0[go], backup.p0) || pos_differ(d.pos1[go], backup.p1)) flag_body_contacts(d, x, S, other);
        }
        if (d.c_flags[gi] & prx::kNeedsUpdating) {
            toi_update(d, x, k, i, c);
            ++es.updated;
        } else {
            ++es.skipped;
        }
        const std::uint32_t fl = d.c_flags[gi];
        if (!(fl & prx::kCEnabled) || !(fl & prx::kTouching)) {
            set_sweep(d, go, backup);
            continue;
        }
        S.icont[I.nc++] = i;
        d.t_isl_c[gi] = 1;
        if (!otherIslanded) {
            if (d.bflags[go] & prx::kSpeedable) d.bflags[go] |= prx::kAwake;
            S.ibody[I.nb++] = other;
            d.t_isl_b[go] = 1;
        }
    }
}

// SolveToiViaGS (AabbTreeWorld.cpp:1968-2057) for the island in S, with
// h the rest of the step; lane 0.
__device__ __forceinline__ void toi_solve_island(const DeviceState& d, const WorldCtx& x, Counters& k,
                                                 const IslandScratch& S, const ToiIsland& I, float h,
                                                 const StepParams& c, ToiEventStats& es) {
    const Bodies<false> bs{d.s_bc_pos + x.bo, d.s_bc_vel + x.bo, x.nb};
    StepParams still = c;
    still.dt = 0.0f;
    for (int q = 0; q < I.nb; ++q) {
        const int b = S.ibody[q];
        bs.set_pos(b, d.pos1[x.bo + b]);
        bs.set_vel(b, integrated_velocity(d, x.bo + b, still));
    }
    float4* pbase = d.s_pc + std::size_t(x.co) * prx::kPcStride;
    float4* vbase = d.s_vc + std::size_t(x.co) * prx::kVcStride;
    for (int e = 0; e < I.nc; ++e) {
        const int gi = x.co + S.icont[e];
        const int2 bd = d.c_body[gi];
        const int2 ss = d.c_shape[gi];
        const BodyC A = load_bodyc(d, x.bo + bd.x), B = load_bodyc(d, x.bo + bd.y);
        store_pc(pbase + std::size_t(e) * prx::kPcStride, load_manifold(d, gi), bd.x, bd.y,
                 d.shape_f[ss.x].x + d.shape_f[ss.y].x, A, B);
    }
    // Position iterations (GetToiConstraintSolverConf, ConstraintSolverConf.cpp:37-45).
    StepParams ps = c;
    ps.regResolutionRate = c.toiResolutionRate;
    std::uint32_t positionIters = c.toiPosIters;
    for (std::uint32_t it = 0; it < c.toiPosIters; ++it) {
        float minSep = __builtin_inff();
        for (int e = 0; e < I.nc; ++e) {
            const Pc pc = load_pc(pbase + std::size_t(e) * prx::kPcStride);
            float4 pa = bs.pos(pc.bA), pb = bs.pos(pc.bB);
            minSep = fmin_std(minSep, solve_pc(pc.m, pc.totalRadius, pc.A, pc.B, pa, pb, ps));
            bs.set_pos(pc.bA, pa);
            bs.set_pos(pc.bB, pb);
        }
        es.minSep = fmin_std(es.minSep, minSep);
        if (minSep >= c.toiMinSeparation) {
            positionIters = it + 1;
            es.solved = true;
            break;
        }
    }
    es.posIters = positionIters;
    for (int q = 0; q < I.nb; ++q) {
        const int b = S.ibody[q];
        const float4 p = bs.pos(b);
        float4 p0 = d.pos0[x.bo + b];
        p0.x = p.x;
        p0.y = p.y;
        p0.z = p.z;
        d.pos0[x.bo + b] = p0;
    }
    // Velocity constraints without warm starting (GetToiVelocityConstraintConf,
    // VelocityConstraint.cpp:207-210).
    StepParams vconf = c;
    vconf.warmStart = 0u;
    for (int e = 0; e < I.nc; ++e) {
        const int gi = x.co + S.icont[e];
        const int2 bd = d.c_body[gi];
        const int2 ss = d.c_shape[gi];
        const float rA = d.shape_f[ss.x].x, rB = d.shape_f[ss.y].x;
        const BodyC A = load_bodyc(d, x.bo + bd.x), B = load_bodyc(d, x.bo + bd.y);
        const float4 pA = bs.pos(bd.x), pB = bs.pos(bd.y);
        const Xf xfA = get_xf(pA.x, pA.y, pA.z, A.lc);
        const Xf xfB = get_xf(pB.x, pB.y, pB.z, B.lc);
        const WM wm = world_manifold(load_manifold(d, gi), xfA, rA, xfB, rB);
        const float4 mt = d.c_mat[gi];
        const Vc vc = make_vc(wm, mt.x, mt.y, mt.z, bd.x, bd.y, A, pA, bs.vel(bd.x), B, pB, bs.vel(bd.y), vconf);
        store_vc(vbase + std::size_t(e) * prx::kVcStride, vc, A, B);
    }
    es.velIters = c.toiVelIters;
    for (std::uint32_t it = 0; it < c.toiVelIters; ++it) {
        float newInc = 0.0f;
        for (int e = 0; e < I.nc; ++e) {
            float4* rec = vbase + std::size_t(e) * prx::kVcStride;
            BodyC A, B;
            Vc vc = load_vc(rec, A, B);
            newInc = fmax_std(newInc, solve_velocity_one(vc, A, B, bs));
            rec[4] = float4{vc.p[0].ni, vc.p[0].ti, vc.p[0].nm, vc.p[0].tm};
            rec[7] = float4{vc.p[1].ni, vc.p[1].ti, vc.p[1].nm, vc.p[1].tm};
        }
        if (newInc <= c.toiMinMomentum) {
            es.velIters = it + 1;
            break;
        }
        es.maxInc = fmax_std(es.maxInc, newInc);
    }
    // IntegratePositions (AabbTreeWorld.cpp:176-188) over the rest of the step, writeback.
    for (int q = 0; q < I.nb; ++q) {
        const int b = S.ibody[q];
        float4 p = bs.pos(b);
        const float4 v = bs.vel(b);
        p.x = p.x + h * v.x;
        p.y = p.y + h * v.y;
        p.z = p.z + h * v.z;
        bs.set_pos(b, p);
    }
    for (int q = 0; q < I.nb; ++q) {
        const int b = S.ibody[q];
        const int gb = x.bo + b;
        const float4 bv = bs.vel(b);
        float4 v = d.vel[gb];
        v.x = bv.x;
        v.y = bv.y;
        v.z = bv.z;
        d.vel[gb] = v;
        const float4 bp = bs.pos(b);
        const float4 p1 = d.pos1[gb];
        if (!(p1.x == bp.x && p1.y == bp.y) || p1.z != bp.z) {
            d.pos1[gb] = float4{bp.x, bp.y, bp.z, p1.w};
            const float4 l = d.lcinv[gb];
            const Xf xf = get_xf(bp.x, bp.y, bp.z, V2{l.x, l.y});
            d.xf[gb] = float4{xf.p.x, xf.p.y, xf.q.x, xf.q.y};
            flag_body_contacts(d, x, S, b);
        }
    }
    // Post-solve reports with the position iterations (AabbTreeWorld.cpp:2052-2054).
    if (x.ecap > 0) {
        for (int e = 0; e < I.nc; ++e) {
            const int at = ev_claim_one(x, k, prx::kEvPostWords, 1);
            if (at < 0) continue;
            const float4* rec = vbase + std::size_t(e) * prx::kVcStride;
            const std::uint32_t n = __float_as_uint(rec[1].w);
            std::uint32_t* o = d.ev + at;
            o[0] = prx::kEvPostSolve;
            o[1] = d.c_id[x.co + S.icont[e]];
            o[2] = n;
            o[3] = positionIters;
            o[4] = n > 0 ? __float_as_uint(rec[4].x) : 0u;
            o[5] = n > 1 ? __float_as_uint(rec[7].x) : 0u;
            o[6] = n > 0 ? __float_as_uint(rec[4].y) : 0u;
            o[7] = n > 1 ? __float_as_uint(rec[7].y) : 0u;
        }
    }
}

// SolveToi(ContactID) (AabbTreeWorld.cpp:1833-1966) on lane 0.
__device__ __forceinline__ ToiEventStats toi_event(const DeviceState& d, const WorldCtx& x, Counters& k,
                                                   const IslandScratch& S, int i, const StepParams& c) {
    ToiEventStats es{__builtin_inff(), 0.0f, false, 0u, 0u, 0u, 0u};
    const int gi = x.co + i;
    const float toi = d.c_toi[gi];
    const int2 bd = d.c_body[gi];
    const int ga = x.bo + bd.x, gb = x.bo + bd.y;
    {
        const SweepD backupA = sweep_of(d, ga), backupB = sweep_of(d, gb);
        body_advance(d, ga, toi);
        if (pos_differ(d.pos0[ga], backupA.p0) || pos_differ(d.pos1[ga], backupA.p1)) flag_body_contacts(d, x, S, bd.x);
        body_advance(d, gb, toi);
        if (pos_differ(d.pos0[gb], backupB.p0) || pos_differ(d.pos1[gb], backupB.p1)) flag_body_contacts(d, x, S, bd.y);
        if (d.c_flags[gi] & prx::kNeedsUpdating) {
            toi_update(d, x, k, i, c);
            ++es.updated;
        }
        const std::uint32_t info = d.c_toiinfo[gi];
        d.c_toiinfo[gi] = ((info & 0xffffu) + 1u) & 0xffffu;  // SetToi({}), IncrementToiCount
        const std::uint32_t fl = d.c_flags[gi];
        if (!(fl & prx::kCEnabled) || !(fl & prx::kTouching)) {
            set_sweep(d, ga, backupA);
            set_sweep(d, gb, backupB);
            es.skipped = es.updated ? 0u : 1u;
            return es;
        }
    }
    if (d.bflags[ga] & prx::kSpeedable) d.bflags[ga] |= prx::kAwake;
    if (d.bflags[gb] & prx::kSpeedable) d.bflags[gb] |= prx::kAwake;
    ToiIsland I{0, 0};
    d.t_isl_b[ga] = 1;
    d.t_isl_b[gb] = 1;
    d.t_isl_c[gi] = 1;
    S.ibody[I.nb++] = bd.x;
    S.ibody[I.nb++] = bd.y;
    S.icont[I.nc++] = i;
    if (d.bflags[ga] & prx::kAccelerable) toi_process_contacts(d, x, k, S, bd.x, I, toi, c, es);
    if (d.bflags[gb] & prx::kAccelerable) toi_process_contacts(d, x, k, S, bd.y, I, toi, c, es);
    // RemoveUnspeedablesFromIslanded.
    for (int q = 0; q < I.nb; ++q) {
        const int g = x.bo + S.ibody[q];
        if (!(d.bflags[g] & prx::kSpeedable)) d.t_isl_b[g] = 0;
    }
    toi_solve_island(d, x, k, S, I, (1.0f - toi) * c.dt, c, es);
    return es;
}

// Every solved event raises one contact's sub-step count, which stops at
// maxSubSteps, so a world has at most maxSubSteps events per contact slot;
// the loop is bounded by that (reaching it ends the phase with
// kStatusToiLimit, whatever the inputs).
__device__ __forceinline__ int toi_event_bound(const WorldCtx& x, const StepParams& c) {
    return int(c.maxSubSteps + 1u) * (x.cap + 1) + 64;
}

__device__ __forceinline__ std::uint32_t wave_sum(std::uint32_t v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}
__device__ __forceinline__ std::uint32_t wave_max(std::uint32_t v) {
    for (int o = 32; o > 0; o >>= 1) {
        const std::uint32_t u = __shfl_xor(v, o);
        v = (v < u) ? u : v;
    }
    return v;
}

// The TOI phase of one world, after the regular phase; its ToiStepStats go to
// the header (lane 0).
__device__ __forceinline__ void toi_phase(const DeviceState& d, const WorldCtx& x, Counters& k,
                                          const IslandScratch& S, const StepParams& c, WorldHeader& H) {
    // The statistics accumulate in the header (lane 0), so that they hold no
    // registers across the loop.
    std::uint32_t* st = H.stats;
    if (x.lane == 0) {
        for (int i = prx::kStToiMinSeparation; i < prx::kStats; ++i) st[i] = 0u;
        st[prx::kStToiMinSeparation] = 0x7f800000u;
    }
    auto addMax = [&](int i, std::uint32_t v) { st[i] = (st[i] < v) ? v : st[i]; };
    for (int i = x.lane; i < k.nc; i += kWave) {
        d.c_toiinfo[x.co + i] = 0u;
        d.t_isl_c[x.co + i] = 0;
    }
    for (int b = x.lane; b < x.nb; b += kWave) d.t_isl_b[x.bo + b] = 0;
    __syncthreads();
    build_adjacency<true>(d, x, k.nc, S);
    for (int pass = 0;; ++pass) {
        // Per-lane maxima and counts of the times of impact this pass computes.
        std::uint32_t lMaxDist = 0, lMaxToi = 0, lMaxRoot = 0, lUpdated = 0, lAtMax = 0;
        auto addToi = [&](const ToiRes& r) {
            lMaxDist = (lMaxDist < r.maxDist) ? r.maxDist : lMaxDist;
            lMaxToi = (lMaxToi < r.toiIters) ? r.toiIters : lMaxToi;
            lMaxRoot = (lMaxRoot < r.maxRoot) ? r.maxRoot : lMaxRoot;
            ++lUpdated;
        };
        // UpdateContactTOIs.
        if (pass == 0) {
            for (int i = x.lane; i < k.nc; i += kWave) {
                const int need = needs_toi(d, x, i, c);
                if (need == 1) addToi(contact_toi(d, x, i, c));
                else if (need == 2) ++lAtMax;
            }
        } else {
            // The filters do not change during the pass (they read the
            // contact's own TOI state, its flags and the bodies' awake
            // flags), so the wave lists the candidates in order and lane 0
            // computes them one after another.
            int ncand = 0;
            for (int base = 0; base < k.nc; base += kWave) {
                const int i = base + x.lane;
                const int nd = (i < k.nc) ? needs_toi(d, x, i, c) : 0;
                const bool need = nd == 1;
                if (nd == 2) ++lAtMax;
                const unsigned long long m = __ballot(need);
                if (need) S.icont[ncand + __popcll(m & lanes_below(x.lane))] = i;
                ncand += __popcll(m);
            }
            __syncthreads();
            if (x.lane == 0)
                for (int q = 0; q < ncand; ++q) addToi(contact_toi(d, x, S.icont[q], c));
        }
        {
            const std::uint32_t upd = wave_sum(lUpdated), atm = wave_sum(lAtMax);
            const std::uint32_t md = wave_max(lMaxDist), mt = wave_max(lMaxToi), mr = wave_max(lMaxRoot);
            if (x.lane == 0) {
                st[prx::kStToiUpdatedToi] += upd;
                st[prx::kStToiAtMaxSubSteps] += atm;
                addMax(prx::kStToiMaxDistIters, md);
                addMax(prx::kStToiMaxToiIters, mt);
                addMax(prx::kStToiMaxRootIters, mr);
            }
        }
        __syncthreads();
        // GetSoonestContact (AabbTreeWorld.cpp:2870-2890): the least TOI below
        // 1, the first in list order on ties.
        float best = 1.0f;
        int bi = 0x7fffffff;
        for (int i = x.lane; i < k.nc; i += kWave) {
            const int gi = x.co + i;
            if ((d.c_toiinfo[gi] & kToiValid) && best > d.c_toi[gi]) {
                best = d.c_toi[gi];
                bi = i;
            }
        }
        for (int o = 32; o > 0; o >>= 1) {
            const float ob = __shfl_xor(best, o);
            const int oi = __shfl_xor(bi, o);
            if (ob < best || (ob == best && oi < bi)) {
                best = ob;
                bi = oi;
            }
        }
        if (bi == 0x7fffffff || pass >= toi_event_bound(x, c)) {
            if (bi != 0x7fffffff) k.status |= prx::kStatusToiLimit;
            break;
        }
        if (x.lane == 0) {
            ++st[prx::kStToiContactsFound];
            if (!d.t_isl_c[x.co + bi]) {
                const ToiEventStats es = toi_event(d, x, k, S, bi, c);
                st[prx::kStToiMinSeparation] =
                    __float_as_uint(fmin_std(__uint_as_float(st[prx::kStToiMinSeparation]), es.minSep));
                st[prx::kStToiMaxIncImpulse] =
                    __float_as_uint(fmax_std(__uint_as_float(st[prx::kStToiMaxIncImpulse]), es.maxInc));
                st[prx::kStToiIslandsSolved] += es.solved ? 1u : 0u;
                st[prx::kStToiSumPosIters] += es.posIters;
                st[prx::kStToiSumVelIters] += es.velIters;
                if (es.posIters > 0 || es.velIters > 0) ++st[prx::kStToiIslandsFound];
                st[prx::kStToiUpdatedTouching] += es.updated;
                st[prx::kStToiSkippedTouching] += es.skipped;
            }
        }
        k.evw = __shfl(k.evw, 0);
        k.evn = __shfl(k.evn, 0);
        k.status = __shfl(k.status, 0);
        __syncthreads();
        // Island bodies: synchronise the accelerable ones' proxies and
        // invalidate their contacts' TOIs, then clear the island flags.
        std::uint32_t moved = 0;
        for (int b = x.lane; b < x.nb; b += kWave) {
            const int g = x.bo + b;
            if (d.t_isl_b[g] && (d.bflags[g] & prx::kAccelerable)) moved += synchronize_body(d, x, b, c);
        }
        moved = wave_sum(moved);
        if (x.lane == 0) st[prx::kStToiProxiesMoved] += moved;
        for (int i = x.lane; i < k.nc; i += kWave) {
            const int2 bd = d.c_body[x.co + i];
            const int ga = x.bo + bd.x, gb = x.bo + bd.y;
            const bool a = d.t_isl_b[ga] && (d.bflags[ga] & prx::kAccelerable);
            const bool b2 = d.t_isl_b[gb] && (d.bflags[gb] & prx::kAccelerable);
            if (a || b2) {
                d.c_toiinfo[x.co + i] &= ~kToiValid;
                d.t_isl_c[x.co + i] = 0;
            }
        }
        __syncthreads();
        for (int b = x.lane; b < x.nb; b += kWave) d.t_isl_b[x.bo + b] = 0;
        __syncthreads();
        const int nc0 = k.nc;
        const std::uint32_t added = phase_find_add(d, x, k, c, 0);
        if (x.lane == 0) st[prx::kStToiAdded] += added;
        __syncthreads();
        if (k.nc > nc0) {
            for (int i = nc0 + x.lane; i < k.nc; i += kWave) {
                d.c_toiinfo[x.co + i] = 0u;
                d.t_isl_c[x.co + i] = 0;
            }
            __syncthreads();
            build_adjacency<true>(d, x, k.nc, S);
        }
    }
    // ResetBodiesForSolveTOI / ResetContactsForSolveTOI.
    for (int b = x.lane; b < x.nb; b += kWave) d.alpha0[x.bo + b] = 0.0f;
    for (int i = x.lane; i < k.nc; i += kWave) d.c_toiinfo[x.co + i] = 0u;
    __syncthreads();
    const std::uint32_t updated = phase_update(d, x, k, c, 0);
    if (x.lane == 0) {
        st[prx::kStToiUpdatedTouching] += updated;
        st[prx::kStToiSkippedTouching] += std::uint32_t(k.nc) - updated;
    }
    __syncthreads();
}

}  // namespace prxd
