// Forward-mode dual numbers and dual copies of the regular step's constraint
// functions, for the general adjoint (adjoint.hip).  The vector-Jacobian
// product of one solver update comes from one evaluation of its dual copy per
// input, with that input's derivative seeded to 1 (one input per lane).
//
// The value part of every dual operation is the float operation the forward
// step performs, in the same order, so a dual copy evaluated on the recorded
// inputs reproduces the forward values bit for bit and takes the same
// branches (clamps, block-solver cases, the position solver's point order).
// Each copy cites the forward function it mirrors; those cite the reference.
#pragma once

#include "step_device.hpp"

namespace prxd {
namespace dl {

struct D {
    float v, d;
};
__device__ __forceinline__ D cst(float v) { return D{v, 0.0f}; }
__device__ __forceinline__ D operator+(D a, D b) { return D{a.v + b.v, a.d + b.d}; }
__device__ __forceinline__ D operator-(D a, D b) { return D{a.v - b.v, a.d - b.d}; }
__device__ __forceinline__ D operator-(D a) { return D{-a.v, -a.d}; }
__device__ __forceinline__ D operator*(D a, D b) { return D{a.v * b.v, a.d * b.v + a.v * b.d}; }
__device__ __forceinline__ D operator*(float a, D b) { return D{a * b.v, a * b.d}; }
__device__ __forceinline__ D operator*(D a, float b) { return D{a.v * b, a.d * b}; }
__device__ __forceinline__ D operator/(D a, D b) {
    const float q = a.v / b.v;
    return D{q, (a.d - q * b.d) / b.v};
}
__device__ __forceinline__ D operator/(D a, float b) { return D{a.v / b, a.d / b}; }
__device__ __forceinline__ D operator/(float a, D b) {
    const float q = a / b.v;
    return D{q, -q * b.d / b.v};
}
__device__ __forceinline__ D sqrtd(D a) {
    const float r = __builtin_sqrtf(a.v);
    return D{r, r > 0.0f ? a.d / (2.0f * r) : 0.0f};
}
__device__ __forceinline__ D fabsd(D a) { return (a.v < 0.0f) ? -a : a; }
__device__ __forceinline__ D fmax_d(D a, D b) { return (a.v < b.v) ? b : a; }
__device__ __forceinline__ D fmin_d(D a, D b) { return (b.v < a.v) ? b : a; }
__device__ __forceinline__ D clamp_d(D v, D lo, D hi) { return (v.v < lo.v) ? lo : ((hi.v < v.v) ? hi : v); }

struct V {
    D x, y;
};
__device__ __forceinline__ V vc2(V2 a) { return V{cst(a.x), cst(a.y)}; }
__device__ __forceinline__ V operator+(V a, V b) { return V{a.x + b.x, a.y + b.y}; }
__device__ __forceinline__ V operator-(V a, V b) { return V{a.x - b.x, a.y - b.y}; }
__device__ __forceinline__ V operator-(V a) { return V{-a.x, -a.y}; }
__device__ __forceinline__ V mul(V a, D s) { return V{a.x * s, a.y * s}; }
__device__ __forceinline__ V mul(D s, V a) { return V{s * a.x, s * a.y}; }
__device__ __forceinline__ V mul(float s, V a) { return V{s * a.x, s * a.y}; }
__device__ __forceinline__ V divs(V a, D s) {
    const D inv = 1.0f / s;
    return V{a.x * inv, a.y * inv};
}
__device__ __forceinline__ V divs(V a, float s) {
    const float inv = 1.0f / s;
    return V{a.x * inv, a.y * inv};
}
__device__ __forceinline__ D dot(V a, V b) {
    D r = cst(0.0f);
    r = r + a.x * b.x;
    r = r + a.y * b.y;
    return r;
}
__device__ __forceinline__ D cross(V a, V b) {
    const D m = a.x * b.y;
    const D s = a.y * b.x;
    return m - s;
}
__device__ __forceinline__ D magsq(V v) {
    D r = cst(0.0f);
    r = r + v.x * v.x;
    r = r + v.y * v.y;
    return r;
}
__device__ __forceinline__ V revperp(V v) { return V{-v.y, v.x}; }
__device__ __forceinline__ V fwdperp(V v) { return V{v.y, -v.x}; }
__device__ __forceinline__ V rotate(V v, V q) {
    const D nx = (q.x * v.x) - (q.y * v.y);
    const D ny = (q.y * v.x) + (q.x * v.y);
    return V{nx, ny};
}
struct X {
    V p, q;
};
__device__ __forceinline__ V transform(V v, const X& xf) { return rotate(v, xf.q) + xf.p; }
// unit_from_angle: values from the forward routine, derivative (-sin, cos).
__device__ __forceinline__ V unit_from_angle(D a) {
    const V2 q = prxd::unit_from_angle(a.v);
    return V{D{q.x, -q.y * a.d}, D{q.y, q.x * a.d}};
}
__device__ __forceinline__ X get_xf(D px, D py, D pa, V2 lc) {
    const V q = unit_from_angle(pa);
    return X{V{px, py} - rotate(vc2(lc), q), q};
}
__device__ __forceinline__ V unit_vector(V v, V2 fallback) {
    const D m2 = v.x * v.x + v.y * v.y;
    if (__builtin_isnormal(m2.v)) {
        const D mag = sqrtd(m2);
        const D inv = 1.0f / mag;
        return V{v.x * inv, v.y * inv};
    }
    const float magv = hypot_f(v.x.v, v.y.v);
    if (!__builtin_isnormal(magv)) return vc2(fallback);
    const D mag = D{magv, (v.x.v * v.x.d + v.y.v * v.y.d) / magv};
    return V{v.x / mag, v.y / mag};
}
__device__ __forceinline__ V rel_vel(V vA, D wA, V relA, V vB, D wB, V relB) {
    const V velBrot = mul(revperp(relB), wB / 1.0f);
    const V velArot = mul(revperp(relA), wA / 1.0f);
    return (vB + velBrot) - (vA + velArot);
}

// ---------------------------------------------------------------------------
// Contacts.

struct WMd {
    V normal;
    std::uint32_t n;
    V p[2];
};

// world_manifold (WorldManifold.cpp:37-130); lpd (when given): the
// manifold's local points as dual values (else its constants).
__device__ __forceinline__ WMd world_manifold(const Manifold& m, const X& xfA, float rA, const X& xfB, float rB,
                                              const V* lpd = nullptr) {
    WMd w;
    w.n = 0;
    w.normal = V{cst(0), cst(0)};
    w.p[0] = w.normal;
    w.p[1] = w.normal;
    if (m.type == kCircles) {
        const V pA = transform(vc2(m.lpt), xfA);
        const V pB = transform(lpd ? lpd[0] : vc2(m.p[0].lp), xfB);
        const V normal = unit_vector(pB - pA, V2{1.0f, 0.0f});
        const V cA = pA + mul(cst(rA), normal);
        const V cB = pB - mul(cst(rB), normal);
        w.normal = normal;
        w.n = 1;
        w.p[0] = divs(cA + cB, 2.0f);
    } else if (m.type == kFaceA) {
        const V normal = rotate(vc2(m.ln), xfA.q);
        const V plane = transform(vc2(m.lpt), xfA);
        w.normal = normal;
        w.n = m.n;
#pragma unroll
        for (std::uint32_t i = 0; i < 2; ++i) {
            if (i >= m.n) continue;
            const V clip = transform(lpd ? lpd[i] : vc2(m.p[i].lp), xfB);
            const V cA = clip + mul(cst(rA) - dot(clip - plane, normal), normal);
            const V cB = clip - mul(cst(rB), normal);
            w.p[i] = divs(cA + cB, 2.0f);
        }
    } else if (m.type == kFaceB) {
        const V normal = rotate(vc2(m.ln), xfB.q);
        const V plane = transform(vc2(m.lpt), xfB);
        w.normal = -normal;
        w.n = m.n;
#pragma unroll
        for (std::uint32_t i = 0; i < 2; ++i) {
            if (i >= m.n) continue;
            const V clip = transform(lpd ? lpd[i] : vc2(m.p[i].lp), xfA);
            const V cB = clip + mul(cst(rB) - dot(clip - plane, normal), normal);
            const V cA = clip - mul(cst(rA), normal);
            w.p[i] = divs(cA + cB, 2.0f);
        }
    }
    return w;
}

struct VcPointD {
    V relA, relB;
    D ni, ti, nm, tm, bias;
};
struct Vcd {
    V normal;
    D friction;
    float tangentSpeed;
    std::uint32_t n;
    D K00, K11, K01, M00, M11, M01;
    VcPointD p[2];
};

// make_vc (VelocityConstraint.cpp:38-185); raw: the manifold impulses
// (n0, t0, n1, t1) before the dtRatio scaling.
__device__ __forceinline__ Vcd make_vc(const WMd& wm, D friction, D restitution, float tangentSpeed, const BodyC& A,
                                       const D* posA, const D* velA, const BodyC& B, const D* posB, const D* velB,
                                       const D* raw, const StepParams& c) {
    Vcd vc;
    const D z = cst(0.0f);
    vc.normal = wm.normal;
    vc.friction = friction;
    vc.tangentSpeed = tangentSpeed;
    vc.n = 0;
    vc.K00 = vc.K11 = vc.K01 = z;
    vc.M00 = vc.M11 = vc.M01 = z;
    vc.p[0] = VcPointD{V{z, z}, V{z, z}, z, z, z, z, z};
    vc.p[1] = vc.p[0];
    const float dtRatio = c.warmStart ? c.dtRatio : 0.0f;
    const float invMass = A.invMass + B.invMass;
    const V tangent = fwdperp(vc.normal);
#pragma unroll
    for (std::uint32_t j = 0; j < 2; ++j) {
        if (j >= wm.n) continue;
        VcPointD p;
        const V relA = wm.p[j] - V{posA[0], posA[1]};
        const V relB = wm.p[j] - V{posB[0], posB[1]};
        p.ni = raw[2 * j] * dtRatio;
        p.ti = raw[2 * j + 1] * dtRatio;
        {
            const V dv = rel_vel(V{velA[0], velA[1]}, velA[2], relA, V{velB[0], velB[1]}, velB[2], relB);
            const D vn = dot(dv, vc.normal);
            p.bias = (vn.v < -c.velocityThreshold) ? -restitution * vn : z;
        }
        p.relA = relA;
        p.relB = relB;
        {
            const D cA = cross(relA, vc.normal);
            const D cB = cross(relB, vc.normal);
            const D rmA = A.invRotI * (cA * cA) / 1.0f;
            const D rmB = B.invRotI * (cB * cB) / 1.0f;
            const D value = cst(invMass) + rmA + rmB;
            p.nm = (value.v != 0.0f) ? 1.0f / value : z;
        }
        {
            const D cA = cross(relA, tangent);
            const D cB = cross(relB, tangent);
            const D rmA = A.invRotI * (cA * cA) / 1.0f;
            const D rmB = B.invRotI * (cB * cB) / 1.0f;
            const D value = cst(invMass) + rmA + rmB;
            p.tm = (value.v != 0.0f) ? 1.0f / value : z;
        }
        vc.p[j] = p;
        ++vc.n;
    }
    if (c.blockSolve && vc.n == 2) {
        const D rnA0 = cross(vc.p[0].relA, vc.normal);
        const D rnB0 = cross(vc.p[0].relB, vc.normal);
        const D rnA1 = cross(vc.p[1].relA, vc.normal);
        const D rnB1 = cross(vc.p[1].relB, vc.normal);
        const D irmA0 = (A.invRotI * (rnA0 * rnA0)) / 1.0f;
        const D irmA1 = (A.invRotI * (rnA1 * rnA1)) / 1.0f;
        const D irmB0 = (B.invRotI * (rnB0 * rnB0)) / 1.0f;
        const D irmB1 = (B.invRotI * (rnB1 * rnB1)) / 1.0f;
        const D irmA = (A.invRotI * rnA0 * rnA1) / 1.0f;
        const D irmB = (B.invRotI * rnB0 * rnB1) / 1.0f;
        const D k00 = cst(invMass) + irmA0 + irmB0;
        const D k11 = cst(invMass) + irmA1 + irmB1;
        const D k01 = cst(invMass) + irmA + irmB;
        const D k00sq = k00 * k00;
        const D kdiff = k00 * k11 - k01 * k01;
        if (k00sq.v < 1000.0f * kdiff.v) {
            vc.K00 = k00;
            vc.K11 = k11;
            vc.K01 = k01;
            const D cpm = k00 * k11;
            const D cps = k01 * k01;
            const D cp = cpm - cps;
            if (!almost_zero(cp.v)) {
                const D inv = 1.0f / cp;
                vc.M00 = k11 * inv;
                vc.M11 = k00 * inv;
                vc.M01 = -k01 * inv;
            }
        } else {
            vc.n = 1;
        }
    }
    return vc;
}

// warm_start_one (AabbTreeWorld.cpp:230-274).
__device__ __forceinline__ void warm_start(const Vcd& vc, const BodyC& A, const BodyC& B, D* velA, D* velB) {
    const V tangent = fwdperp(vc.normal);
    const D z = cst(0.0f);
    V dAl{z, z}, dBl{z, z};
    D dAa = z, dBa = z;
#pragma unroll
    for (std::uint32_t j = 0; j < 2; ++j) {
        if (j >= vc.n) continue;
        const VcPointD& p = vc.p[j];
        const V P = mul(p.ni, vc.normal) + mul(p.ti, tangent);
        const D LA = cross(p.relA, P) / 1.0f;
        const D LB = cross(p.relB, P) / 1.0f;
        dAl = dAl - mul(A.invMass, P);
        dAa = dAa - A.invRotI * LA;
        dBl = dBl + mul(B.invMass, P);
        dBa = dBa + B.invRotI * LB;
    }
    velA[0] = velA[0] + dAl.x;
    velA[1] = velA[1] + dAl.y;
    velA[2] = velA[2] + dAa;
    velB[0] = velB[0] + dBl.x;
    velB[1] = velB[1] + dBl.y;
    velB[2] = velB[2] + dBa;
}

// solve_dir (ContactSolver.cpp:337-454).
__device__ __forceinline__ void solve_dir(Vcd& vc, bool tangent, const BodyC& A, const BodyC& B, D* velA, D* velB) {
    const V dir = tangent ? fwdperp(vc.normal) : vc.normal;
    V vA{velA[0], velA[1]}, vB{velB[0], velB[1]};
    D wA = velA[2], wB = velB[2];
#pragma unroll
    for (int kk = 0; kk < 2; ++kk) {
        const int k = 1 - kk;
        if (k >= int(vc.n)) continue;
        VcPointD& p = vc.p[k];
        const V closing = rel_vel(vA, wA, p.relA, vB, wB, p.relB);
        D oldI, newI;
        if (tangent) {
            const D dirVel = cst(vc.tangentSpeed) - dot(closing, dir);
            const D lambda = p.tm * dirVel;
            const D maxI = vc.friction * p.ni;
            oldI = p.ti;
            newI = clamp_d(oldI + lambda, -maxI, maxI);
        } else {
            const D dirVel = dot(closing, dir);
            const D lambda = p.nm * (p.bias - dirVel);
            oldI = p.ni;
            newI = fmax_d(oldI + lambda, cst(0.0f));
        }
        const D inc = newI - oldI;
        const V P = mul(inc, dir);
        const D LA = cross(p.relA, P) / 1.0f;
        const D LB = cross(p.relB, P) / 1.0f;
        vA = vA - mul(A.invMass, P);
        wA = wA - A.invRotI * LA;
        vB = vB + mul(B.invMass, P);
        wB = wB + B.invRotI * LB;
        if (tangent) p.ti = oldI + inc;
        else p.ni = oldI + inc;
    }
    velA[0] = vA.x;
    velA[1] = vA.y;
    velA[2] = wA;
    velB[0] = vB.x;
    velB[1] = vB.y;
    velB[2] = wB;
}

// block_solve (ContactSolver.cpp:65-335).
__device__ __forceinline__ void block_solve(Vcd& vc, const BodyC& A, const BodyC& B, D* velA, D* velB) {
    const V normal = vc.normal;
    const D z = cst(0.0f);
    V bp;
    {
        const V vA{velA[0], velA[1]}, vB{velB[0], velB[1]};
        const V dv0 = rel_vel(vA, velA[2], vc.p[0].relA, vB, velB[2], vc.p[0].relB);
        const V dv1 = rel_vel(vA, velA[2], vc.p[1].relA, vB, velB[2], vc.p[1].relB);
        const D vn0 = dot(dv0, normal);
        const D vn1 = dot(dv1, normal);
        const V b{vn0 - vc.p[0].bias, vn1 - vc.p[1].bias};
        D t0 = z;
        t0 = t0 + vc.K00 * vc.p[0].ni;
        t0 = t0 + vc.K01 * vc.p[1].ni;
        D t1 = z;
        t1 = t1 + vc.K01 * vc.p[0].ni;
        t1 = t1 + vc.K11 * vc.p[1].ni;
        bp = V{b.x - t0, b.y - t1};
    }
    D n0, n1;
    bool ok = false;
    {
        D x0 = z;
        x0 = x0 + vc.M00 * bp.x;
        x0 = x0 + vc.M01 * bp.y;
        D x1 = z;
        x1 = x1 + vc.M01 * bp.x;
        x1 = x1 + vc.M11 * bp.y;
        n0 = -x0;
        n1 = -x1;
        ok = (n0.v >= 0.0f && n1.v >= 0.0f);
    }
    if (!ok) {
        n0 = -vc.p[0].nm * bp.x;
        n1 = z;
        const D vn2 = vc.K01 * n0 + bp.y;
        ok = (n0.v >= 0.0f && vn2.v >= 0.0f);
    }
    if (!ok) {
        n0 = z;
        n1 = -vc.p[1].nm * bp.y;
        const D vn1 = vc.K01 * n1 + bp.x;
        ok = (n1.v >= 0.0f && vn1.v >= 0.0f);
    }
    if (!ok) {
        n0 = z;
        n1 = z;
        ok = (bp.x.v >= 0.0f && bp.y.v >= 0.0f);
    }
    if (!ok) return;
    const D d0 = n0 - vc.p[0].ni;
    const D d1 = n1 - vc.p[1].ni;
    const V P0 = mul(d0, normal);
    const V P1 = mul(d1, normal);
    const V P = P0 + P1;
    const D LA = (cross(vc.p[0].relA, P0) + cross(vc.p[1].relA, P1)) / 1.0f;
    const D LB = (cross(vc.p[0].relB, P0) + cross(vc.p[1].relB, P1)) / 1.0f;
    const V dAl = -mul(A.invMass, P);
    const D dAa = -(A.invRotI * LA);
    const V dBl = mul(B.invMass, P);
    const D dBa = B.invRotI * LB;
    velA[0] = velA[0] + dAl.x;
    velA[1] = velA[1] + dAl.y;
    velA[2] = velA[2] + dAa;
    velB[0] = velB[0] + dBl.x;
    velB[1] = velB[1] + dBl.y;
    velB[2] = velB[2] + dBa;
    vc.p[0].ni = n0;
    vc.p[1].ni = n1;
}

// solve_velocity_one (ContactSolver.cpp:481-492).
__device__ __forceinline__ void solve_velocity(Vcd& vc, const BodyC& A, const BodyC& B, D* velA, D* velB) {
    solve_dir(vc, true, A, B, velA, velB);
    if (vc.n == 1 || (vc.K00.v == 0.0f && vc.K11.v == 0.0f && vc.K01.v == 0.0f)) solve_dir(vc, false, A, B, velA, velB);
    else block_solve(vc, A, B, velA, velB);
}

struct PSMd {
    V normal, point;
    D sep;
};
// get_psm (PositionSolverManifold.cpp:42-121); lpd as for world_manifold.
__device__ __forceinline__ PSMd get_psm(const Manifold& m, std::uint32_t idx, const X& xfA, const X& xfB,
                                        const V* lpd = nullptr) {
    const V plp = lpd ? lpd[idx ? 1 : 0] : vc2(idx ? m.p[1].lp : m.p[0].lp);
    if (m.type == kCircles) {
        const V pA = transform(vc2(m.lpt), xfA);
        const V pB = transform(plp, xfB);
        const V delta = pB - pA;
        const V normal = unit_vector(delta, V2{0, 0});
        return PSMd{normal, divs(pA + pB, 2.0f), dot(delta, normal)};
    }
    if (m.type == kFaceA) {
        const V plane = transform(vc2(m.lpt), xfA);
        const V normal = rotate(vc2(m.ln), xfA.q);
        const V clip = transform(plp, xfB);
        return PSMd{normal, clip, dot(clip - plane, normal)};
    }
    const V plane = transform(vc2(m.lpt), xfB);
    const V normal = rotate(vc2(m.ln), xfB.q);
    const V clip = transform(plp, xfA);
    return PSMd{-normal, clip, dot(clip - plane, normal)};
}

// solve_pc (ContactSolver.cpp:494-643); pos: x, y, angle of A then B, updated.
__device__ __forceinline__ void solve_pc(const Manifold& m, float totalRadius, const BodyC& A, const BodyC& B, D* pos,
                                         const StepParams& c, const V* lpd = nullptr) {
    const float invMassTotal = A.invMass + B.invMass;
    struct Sol {
        V la;
        D aa;
        V lb;
        D ab;
    };
    auto solve = [&](const PSMd& psm, V pA, V pB) {
        const D separation = psm.sep - cst(totalRadius);
        const V rA = psm.point - pA;
        const V rB = psm.point - pB;
        const D crA = cross(rA, psm.normal) / 1.0f;
        const D crB = cross(rB, psm.normal) / 1.0f;
        const D K = cst(invMassTotal) + A.invRotI * (crA * crA) + B.invRotI * (crB * crB);
        const D C = -clamp_d(c.regResolutionRate * (separation + cst(c.linearSlop)), cst(-c.maxLinearCorrection),
                             cst(0.0f));
        const V P = divs(mul(psm.normal, C), K);
        const D LA = cross(rA, P) / 1.0f;
        const D LB = cross(rB, P) / 1.0f;
        return Sol{-mul(A.invMass, P), -(A.invRotI * LA), mul(B.invMass, P), B.invRotI * LB};
    };
    V pAl{pos[0], pos[1]}, pBl{pos[3], pos[4]};
    D pAa = pos[2], pBa = pos[5];
    auto apply = [&](const Sol& s) {
        pAl = pAl + s.la;
        pAa = pAa + s.aa;
        pBl = pBl + s.lb;
        pBa = pBa + s.ab;
    };
    if (m.n == 1) {
        const PSMd psm0 = get_psm(m, 0, get_xf(pAl.x, pAl.y, pAa, A.lc), get_xf(pBl.x, pBl.y, pBa, B.lc), lpd);
        apply(solve(psm0, pAl, pBl));
    } else if (m.n == 2) {
        const X xfA = get_xf(pAl.x, pAl.y, pAa, A.lc);
        const X xfB = get_xf(pBl.x, pBl.y, pBa, B.lc);
        const PSMd psm0 = get_psm(m, 0, xfA, xfB, lpd);
        const PSMd psm1 = get_psm(m, 1, xfA, xfB, lpd);
        if (almost_equal(psm0.sep.v, psm1.sep.v)) {
            const Sol s0 = solve(psm0, pAl, pBl);
            const Sol s1 = solve(psm1, pAl, pBl);
            pAl = pAl + s0.la + s1.la;
            pAa = pAa + s0.aa + s1.aa;
            pBl = pBl + s0.lb + s1.lb;
            pBa = pBa + s0.ab + s1.ab;
        } else if (psm0.sep.v < psm1.sep.v) {
            apply(solve(psm0, pAl, pBl));
            const PSMd p1 = get_psm(m, 1, get_xf(pAl.x, pAl.y, pAa, A.lc), get_xf(pBl.x, pBl.y, pBa, B.lc), lpd);
            apply(solve(p1, pAl, pBl));
        } else if (psm1.sep.v < psm0.sep.v) {
            apply(solve(psm1, pAl, pBl));
            const PSMd p0 = get_psm(m, 0, get_xf(pAl.x, pAl.y, pAa, A.lc), get_xf(pBl.x, pBl.y, pBa, B.lc), lpd);
            apply(solve(p0, pAl, pBl));
        }
    }
    pos[0] = pAl.x;
    pos[1] = pAl.y;
    pos[2] = pAa;
    pos[3] = pBl.x;
    pos[4] = pBl.y;
    pos[5] = pBa;
}

// A manifold point's local coordinates as a function of the two bodies'
// transformations at the update that made the manifold (poly_manifold,
// CollidePolygons in CollideShapes.cpp).  A point kept from the incident
// edge is a vertex of the incident shape: constant.  A point the clipping
// made (ContactFeature: vertex of the reference shape, face of the incident
// one) lies where the incident edge crosses the side plane through the
// reference face's end vertex; with d_i the signed distances of the incident
// vertices from that plane, lp = v_a + (v_b - v_a) d_a / (d_a - d_b).
// cf: the point's ContactFeature as stored; faceB: the manifold's reference
// face is on B (the feature is stored flipped); ln: the reference face's
// normal in its shape's frame; X0 (X1): the reference (incident) body's
// transformation.
__device__ __forceinline__ V manifold_local_point(const Proxy& s0, const X& X0, const Proxy& s1, const X& X1,
                                                  std::uint32_t cf, bool faceB, V2 ln, V2 lp) {
    const std::uint32_t u = faceB ? flip_cf(cf) : cf;
    const std::uint32_t tA = u & 1u, tB = (u >> 1) & 1u, iA = (u >> 8) & 0xffu, iB = (u >> 16) & 0xffu;
    if (!(tA == kVertex && tB == kFace) || iB >= s1.n) return vc2(lp);
    std::uint32_t idx0 = s0.n;
    for (std::uint32_t i = 0; i < s0.n; ++i) {
        const V2 nr = s0.norm(i);
        if (idx0 == s0.n && nr.x == ln.x && nr.y == ln.y) idx0 = i;
    }
    if (idx0 == s0.n) return vc2(lp);
    const std::uint32_t idx0Next = mod_next(idx0, s0.n);
    if (iA != idx0 && iA != idx0Next) return vc2(lp);
    const V2 va = s1.vert(iB), vb = s1.vert(mod_next(iB, s1.n));
    const V dir = rotate(vc2(revperp(s0.norm(idx0))), X0.q);
    const V av = transform(vc2(s0.vert(iA)), X0);
    const D da = dot(dir, transform(vc2(va), X1) - av);
    const D db = dot(dir, transform(vc2(vb), X1) - av);
    const D t = da / (da - db);
    return V{cst(va.x) + (vb.x - va.x) * t, cst(va.y) + (vb.y - va.y) * t};
}

// ---------------------------------------------------------------------------
// Revolute joints.  Jd holds the differentiable values of a joint record;
// the constants (masses, flags, limit state, motor speed, anchors, angles)
// come from the recorded record.

struct Jd {
    V rA, rB;
    D exx, eyx, ezx, eyy, ezy, ezz, angularMass;
    D ix, iy, iz, motorImp;
    D maxImpulse;
};

// build_jc (RevoluteJointConf.cpp:127-197): angles, impulses before the
// set-up (x, y, z, motor) and the maximum motor torque in; the record's
// values out.  prevLimit: the limit state before the set-up.
__device__ __forceinline__ Jd build_jc(D angA, D angB, const D* raw, D maxTorque, const float4* u,
                                       std::uint32_t prevLimit, const StepParams& c) {
    const float4 ms = u[4], la = u[6], ang = u[7];
    const std::uint32_t flags = __float_as_uint(u[2].w) & 0xffu;
    const float mA = ms.x, iA = ms.y, mB = ms.z, iB = ms.w;
    const V qA = unit_from_angle(angA);
    const V qB = unit_from_angle(angB);
    const V rA = rotate(V{cst(la.x), cst(la.y)}, qA);
    const V rB = rotate(V{cst(la.z), cst(la.w)}, qB);
    const float totInvI = iA + iB;
    const bool fixedRotation = (totInvI == 0.0f);
    Jd j;
    j.rA = rA;
    j.rB = rB;
    j.exx = cst(mA) + ((rA.y * rA.y) * iA / 1.0f) + cst(mB) + ((rB.y * rB.y) * iB / 1.0f);
    j.eyx = (-rA.y * rA.x * iA / 1.0f) + (-rB.y * rB.x * iB / 1.0f);
    j.ezx = (-rA.y * iA) + (-rB.y * iB);
    j.eyy = cst(mA) + ((rA.x * rA.x) * iA / 1.0f) + cst(mB) + ((rB.x * rB.x) * iB / 1.0f);
    j.ezy = (rA.x * iA) + (rB.x * iB);
    j.ezz = cst(totInvI);
    j.angularMass = cst((totInvI > 0.0f) ? (1.0f / totInvI) : 0.0f);
    D im[4] = {raw[0], raw[1], raw[2], raw[3]};
    std::uint32_t limit = prevLimit;
    if (!(flags & prx::kJointMotor) || fixedRotation) im[3] = cst(0.0f);
    if ((flags & prx::kJointLimit) && !fixedRotation) {
        const float jointAngle = angB.v - angA.v - ang.x;
        if (__builtin_fabsf(ang.z - ang.y) < (c.angularSlop * 2.0f)) {
            limit = 3u;
        } else if (jointAngle <= ang.y) {
            if (limit != 1u) im[2] = cst(0.0f);
        } else if (jointAngle >= ang.z) {
            if (limit != 2u) im[2] = cst(0.0f);
        } else {
            im[2] = cst(0.0f);
        }
    }
    if (c.warmStart) {
        for (int k = 0; k < 4; ++k) im[k] = im[k] * c.dtRatio;
    } else {
        for (int k = 0; k < 4; ++k) im[k] = cst(0.0f);
    }
    j.ix = im[0];
    j.iy = im[1];
    j.iz = im[2];
    j.motorImp = im[3];
    j.maxImpulse = c.dt * maxTorque;
    return j;
}

// joint_warm_start (RevoluteJointConf.cpp:199-224).
__device__ __forceinline__ void joint_warm_start(const Jd& j, const float4* u, D* velA, D* velB) {
    const float4 ms = u[4];
    const V P{j.ix, j.iy};
    const D L = j.motorImp + j.iz;
    const D LA = cross(j.rA, P) / 1.0f + L;
    const D LB = cross(j.rB, P) / 1.0f + L;
    velA[0] = velA[0] - ms.x * P.x;
    velA[1] = velA[1] - ms.x * P.y;
    velA[2] = velA[2] - ms.y * LA;
    velB[0] = velB[0] + ms.z * P.x;
    velB[1] = velB[1] + ms.z * P.y;
    velB[2] = velB[2] + ms.w * LB;
}

__device__ __forceinline__ D dot3(D ax, D ay, D az, D bx, D by, D bz) {
    D r = cst(0.0f);
    r = r + ax * bx;
    r = r + ay * by;
    r = r + az * bz;
    return r;
}
__device__ __forceinline__ void cross3(D ax, D ay, D az, D bx, D by, D bz, D& cx, D& cy, D& cz) {
    cx = ay * bz - az * by;
    cy = az * bx - ax * bz;
    cz = ax * by - ay * bx;
}
__device__ __forceinline__ void solve33(const Jd& j, D bx, D by, D bz, D& ox, D& oy, D& oz) {
    D yzx, yzy, yzz;
    cross3(j.eyx, j.eyy, j.ezy, j.ezx, j.ezy, j.ezz, yzx, yzy, yzz);
    const D dp = dot3(j.exx, j.eyx, j.ezx, yzx, yzy, yzz);
    const D det = (dp.v != 0.0f) ? (1.0f / dp) : dp;
    ox = det * dot3(bx, by, bz, yzx, yzy, yzz);
    D bzx, bzy, bzz;
    cross3(bx, by, bz, j.ezx, j.ezy, j.ezz, bzx, bzy, bzz);
    oy = det * dot3(j.exx, j.eyx, j.ezx, bzx, bzy, bzz);
    D ybx, yby, ybz;
    cross3(j.eyx, j.eyy, j.ezy, bx, by, bz, ybx, yby, ybz);
    oz = det * dot3(j.exx, j.eyx, j.ezx, ybx, yby, ybz);
}
__device__ __forceinline__ V solve22(const Jd& j, V b) {
    const D xx = j.exx, xy = j.eyx, yx = j.eyx, yy = j.eyy;
    const D cp = xx * yy - xy * yx;
    const D det = (cp.v != 0.0f) ? (1.0f / cp) : cp;
    return V{det * (yy * b.x - xy * b.y), det * (xx * b.y - yx * b.x)};
}

// joint_solve_velocity (RevoluteJointConf.cpp:228-331).
__device__ __forceinline__ void joint_solve_velocity(Jd& j, const float4* u, D* velA, D* velB) {
    const float4 n = u[2], ms = u[4], mo = u[5];
    const std::uint32_t flags = __float_as_uint(n.w) & 0xffu, limit = __float_as_uint(n.w) >> 8;
    const D oldA[3] = {velA[0], velA[1], velA[2]}, oldB[3] = {velB[0], velB[1], velB[2]};
    const float mA = ms.x, iA = ms.y, mB = ms.z, iB = ms.w;
    const float motorSpeed = mo.x;
    const bool fixedRotation = (iA + iB == 0.0f);
    if ((flags & prx::kJointMotor) && limit != 3u && !fixedRotation) {
        const D impulse = -j.angularMass * (velB[2] - velA[2] - cst(motorSpeed));
        const D oldImpulse = j.motorImp;
        j.motorImp = clamp_d(j.motorImp + impulse, -j.maxImpulse, j.maxImpulse);
        const D inc = j.motorImp - oldImpulse;
        velA[2] = velA[2] - iA * inc;
        velB[2] = velB[2] + iB * inc;
    }
    const V vb = V{velB[0], velB[1]} + mul(revperp(j.rB), velB[2] / 1.0f);
    const V va = V{velA[0], velA[1]} + mul(revperp(j.rA), velA[2] / 1.0f);
    const V vDelta = vb - va;
    V P;
    if ((flags & prx::kJointLimit) && limit != 0u && !fixedRotation) {
        D sx, sy, sz;
        solve33(j, vDelta.x / 1.0f, vDelta.y / 1.0f, (velB[2] - velA[2]) / 1.0f, sx, sy, sz);
        D ix = -sx, iy = -sy, iz = -sz;
        const bool lowerCase = limit == 1u, upperCase = limit == 2u;
        const D newZ = j.iz + iz;
        if ((lowerCase && newZ.v < 0.0f) || (upperCase && newZ.v > 0.0f)) {
            const V rhs = -V{vDelta.x / 1.0f, vDelta.y / 1.0f} + mul(j.iz, V{j.ezx, j.ezy});
            const V reduced = solve22(j, rhs);
            ix = reduced.x;
            iy = reduced.y;
            iz = -j.iz;
            j.ix = j.ix + reduced.x;
            j.iy = j.iy + reduced.y;
            j.iz = cst(0.0f);
        } else {
            j.ix = j.ix + ix;
            j.iy = j.iy + iy;
            j.iz = j.iz + iz;
        }
        P = V{ix, iy};
        const D L = iz;
        const D LA = cross(j.rA, P) / 1.0f + L;
        const D LB = cross(j.rB, P) / 1.0f + L;
        velA[0] = velA[0] - mA * P.x;
        velA[1] = velA[1] - mA * P.y;
        velA[2] = velA[2] - iA * LA;
        velB[0] = velB[0] + mB * P.x;
        velB[1] = velB[1] + mB * P.y;
        velB[2] = velB[2] + iB * LB;
    } else {
        const V imp = solve22(j, -V{vDelta.x / 1.0f, vDelta.y / 1.0f});
        j.ix = j.ix + imp.x;
        j.iy = j.iy + imp.y;
        P = imp;
        const D LA = cross(j.rA, P) / 1.0f;
        const D LB = cross(j.rB, P) / 1.0f;
        velA[0] = velA[0] - mA * P.x;
        velA[1] = velA[1] - mA * P.y;
        velA[2] = velA[2] - iA * LA;
        velB[0] = velB[0] + mB * P.x;
        velB[1] = velB[1] + mB * P.y;
        velB[2] = velB[2] + iB * LB;
    }
    const bool same = velA[0].v == oldA[0].v && velA[1].v == oldA[1].v && velA[2].v == oldA[2].v &&
                      velB[0].v == oldB[0].v && velB[1].v == oldB[1].v && velB[2].v == oldB[2].v;
    if (same) {
        for (int k = 0; k < 3; ++k) {
            velA[k] = oldA[k];
            velB[k] = oldB[k];
        }
    }
}

// joint_solve_position (RevoluteJointConf.cpp:333-433); pos: A then B.
__device__ __forceinline__ void joint_solve_position(const float4* u, D* pos, const StepParams& c) {
    const float4 n = u[2], ms = u[4], la = u[6], ang = u[7];
    const float angularMass = n.z;
    const std::uint32_t flags = __float_as_uint(n.w) & 0xffu, limit = __float_as_uint(n.w) >> 8;
    const float mA = ms.x, iA = ms.y, mB = ms.z, iB = ms.w;
    const bool fixedRotation = ((iA + iB) == 0.0f);
    if ((flags & prx::kJointLimit) && limit != 0u && !fixedRotation) {
        const D angle = pos[5] - pos[2] - cst(ang.x);
        D limitImpulse = cst(0.0f);
        if (limit == 1u) {
            D C = angle - cst(ang.y);
            C = clamp_d(C + cst(c.angularSlop), cst(-c.maxAngularCorrection), cst(0.0f));
            limitImpulse = -angularMass * C;
        } else if (limit == 2u) {
            D C = angle - cst(ang.z);
            C = clamp_d(C - cst(c.angularSlop), cst(0.0f), cst(c.maxAngularCorrection));
            limitImpulse = -angularMass * C;
        } else {
            const D C = clamp_d(angle - cst(ang.y), cst(-c.maxAngularCorrection), cst(c.maxAngularCorrection));
            limitImpulse = -angularMass * C;
        }
        pos[2] = pos[2] - iA * limitImpulse;
        pos[5] = pos[5] + iB * limitImpulse;
    }
    const V qA = unit_from_angle(pos[2]);
    const V qB = unit_from_angle(pos[5]);
    const V rA = rotate(V{cst(la.x), cst(la.y)}, qA);
    const V rB = rotate(V{cst(la.z), cst(la.w)}, qB);
    const V C = (V{pos[3], pos[4]} + rB) - (V{pos[0], pos[1]} + rA);
    const D exx = cst(mA) + (iA * (rA.y * rA.y) / 1.0f) + cst(mB) + (iB * (rB.y * rB.y) / 1.0f);
    const D exy = (-iA * rA.x * rA.y / 1.0f) + (-iB * rB.x * rB.y / 1.0f);
    const D eyy = cst(mA) + (iA * (rA.x * rA.x) / 1.0f) + cst(mB) + (iB * (rB.x * rB.x) / 1.0f);
    V sol{cst(0.0f), cst(0.0f)};
    const D cp = cross(V{exx, exy}, V{exy, eyy});
    if (!almost_zero(cp.v)) {
        const D inverse = 1.0f / cp;
        sol = V{(eyy * C.x - exy * C.y) * inverse, (exx * C.y - exy * C.x) * inverse};
    }
    const V P = -sol;
    pos[0] = pos[0] - mA * P.x;
    pos[1] = pos[1] - mA * P.y;
    pos[2] = pos[2] - iA * cross(rA, P) / 1.0f;
    pos[3] = pos[3] + mB * P.x;
    pos[4] = pos[4] + mB * P.y;
    pos[5] = pos[5] + iB * cross(rB, P) / 1.0f;
}

}  // namespace dl
}  // namespace prxd
