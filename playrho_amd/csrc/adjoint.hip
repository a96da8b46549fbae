// Hand-written adjoint of the regular step's contact solve and integration
// (SolveRegIslandViaGS, AabbTreeWorld.cpp:1563-1682) for worlds of disks
// against each other and against static edges/polygons: one-point manifolds
// of type circles or face A.  The active set of each step is held fixed as
// recorded by the forward pass (k_step_worlds_tape): which contacts touch and
// form which islands, the clamp branches of every Gauss-Seidel update, the
// iteration counts of the early exits.  Within it every differentiable
// dependence is followed:
//
//   velocity integration        v' = (v + h a) / (1 + h damping)   (Body.cpp:290-309)
//   constraint creation         normal, points, effective masses and the
//                               restitution bias from positions and v'
//                               (WorldManifold.cpp:37-130, VelocityConstraint.cpp:81-185)
//   warm start                  (AabbTreeWorld.cpp:230-274)
//   Gauss-Seidel updates        tangent (friction clamp) then normal
//                               (ContactSolver.cpp:337-454)
//   position integration        (AabbTreeWorld.cpp:176-188)
//   position solver             (ContactSolver.cpp:494-643, PositionSolverManifold.cpp:42-121)
//   impulse carry-over          manifold impulses warm-start the next step
//   material mixing             friction sqrt(fA fB), restitution max(eA, eB) (Contact.hpp:57-69)
//
// One launch reverses one recorded step for every world: one wavefront per
// world, one lane per island (islands share no dynamic body).
#include "layout.hpp"
#include "step_device.hpp"

namespace prxd {

using prx::AdjState;
using prx::Tape;

__device__ __forceinline__ V2 v2(float4 a) { return V2{a.x, a.y}; }
__device__ __forceinline__ V2 perp_grad(V2 c) { return V2{c.y, -c.x}; }  // d dot(revperp(r), c) / dr

__device__ __forceinline__ void add_body(float4* p, V2 v, float w) {
    float4 a = *p;
    a.x += v.x;
    a.y += v.y;
    a.z += w;
    *p = a;
}

__global__ __launch_bounds__(64) void k_adjoint_step(DeviceState d, Tape t, AdjState a, StepParams c) {
    const int w = int(blockIdx.x);
    const int lane = int(threadIdx.x);
    const WorldHeader H = d.hdr[w];
    const int bo = H.body_off, co = H.contact_off, nb = H.nbodies, cap = H.contact_cap;
    const int nisl = t.nisl[w];
    const int4* isl = t.isl + bo + w;
    const int nslots = nisl > 0 ? isl[nisl - 1].y : 0;
    // Impulse adjoints carried in from the next step, by contact id: the
    // step's final impulses warm-started the next step only if they reached
    // it unchanged (the contact kept touching; the manifold update carried
    // them by ContactFeature).
    for (int e = lane; e < nslots; e += kWave) {
        const float4* vc = t.vc + (std::size_t(co) + e) * prx::kTapeVc;
        const std::uint32_t id = __float_as_uint(vc[10].y);
        const float4 lb = a.lambar[co + int(id)];
        const float4 fin = vc[14];
        const bool carried = lb.z == fin.x && lb.w == fin.y;
        a.lamtmp[co + e] = carried ? float2{lb.x, lb.y} : float2{0.0f, 0.0f};
    }
    __syncthreads();
    for (int i = lane; i < cap; i += kWave) a.lambar[co + i] = float4{0.0f, 0.0f, 0.0f, 0.0f};
    __syncthreads();

    // Position solver, reversed (per island, iterations and slots in reverse).
    for (int is = lane; is < nisl; is += kWave) {
        const int4 rng = isl[is];
        for (int it = rng.w - 1; it >= 0; --it) {
            for (int e = rng.y - 1; e >= rng.x; --e) {
                const std::size_t slot = std::size_t(co) + e;
                const float4* vc = t.vc + slot * prx::kTapeVc;
                const std::uint32_t type = __float_as_uint(vc[10].z) & 0xffu;
                const std::uint32_t npts = __float_as_uint(vc[10].z) >> 8;
                if (npts != 1u || (type != 1u && type != 2u)) {
                    *a.status = 1;
                    continue;
                }
                const int bA = __float_as_int(vc[2].w), bB = __float_as_int(vc[5].z);
                const float4 ms = vc[8];
                const float4* g = t.ngs + (slot * std::size_t(t.itP) + it) * 2;
                // Circle centres: a dynamic disk's is its position (disks at
                // the body origin); a static body's point does not move.
                const bool dynA = ms.x != 0.0f || ms.y != 0.0f;
                const V2 pA = dynA ? v2(g[0]) : V2{vc[9].x, vc[9].y};
                const V2 pB = v2(g[1]);
                float4* xa = a.xbar + bo + bA;
                float4* xb = a.xbar + bo + bB;
                const V2 xbA = v2(*xa), xbB = v2(*xb);
                atomicAdd(a.diag + 2, 1.0f);
                const float s = g[0].w;  // separation the forward update used
                const float inner = c.regResolutionRate * (s + c.linearSlop);
                const float Cc = -clamp_std(inner, -c.maxLinearCorrection, 0.0f);
                const bool active = !(inner < -c.maxLinearCorrection) && !(0.0f < inner);
                const float K = ms.x + ms.z;
                if (!(K > 0.0f)) continue;
                const V2 gP = mul(ms.z, xbB) - mul(ms.x, xbA);
                if (type == 1u) {
                    const V2 dd = pB - pA;
                    const float L = __builtin_sqrtf(magsq(dd));
                    if (!(L > 0.0f)) continue;
                    const V2 n{dd.x / L, dd.y / L};
                    const float Cbar = dot(gP, n) / K;
                    const V2 nbar = mul(gP, Cc / K);
                    const float sbar = active ? -c.regResolutionRate * Cbar : 0.0f;
                    const float nn = dot(n, nbar);
                    const V2 dbar = mul(nbar - mul(n, nn), 1.0f / L) + mul(n, sbar);
                    if (ms.x != 0.0f || ms.y != 0.0f) add_body(xa, -dbar, 0.0f);
                    add_body(xb, dbar, 0.0f);
                } else {
                    // Face A: A is the polygon (static: its plane and normal
                    // are fixed), B the disk whose centre is the clip point.
                    const V2 n = v2(vc[0]);
                    const float Cbar = dot(gP, n) / K;
                    const float sbar = active ? -c.regResolutionRate * Cbar : 0.0f;
                    add_body(xb, mul(n, sbar), 0.0f);
                }
            }
        }
    }
    __syncthreads();
    // Position integration, reversed: x1 = x0 + h v.
    for (int b = lane; b < nb; b += kWave) {
        if (!t.inisl[bo + b]) continue;
        const float4 xb = a.xbar[bo + b];
        add_body(a.vbar + bo + b, V2{c.dt * xb.x, c.dt * xb.y}, c.dt * xb.z);
    }
    __syncthreads();
    // Gauss-Seidel sweeps, warm start and constraint creation, reversed.
    for (int is = lane; is < nisl; is += kWave) {
        const int4 rng = isl[is];
        for (int e = rng.x; e < rng.y; ++e) {
            float4* cb = a.cbar + (std::size_t(co) + e) * 3;
            cb[0] = float4{0, 0, 0, 0};
            cb[1] = float4{0, 0, 0, 0};
            cb[2] = float4{0, 0, 0, 0};
        }
        for (int it = rng.z - 1; it >= 0; --it) {
            for (int e = rng.y - 1; e >= rng.x; --e) {
                const std::size_t slot = std::size_t(co) + e;
                const float4* vc = t.vc + slot * prx::kTapeVc;
                if ((__float_as_uint(vc[10].z) >> 8) != 1u) continue;
                const V2 n = v2(vc[0]);
                const V2 tg = fwdperp(n);
                const float f = vc[0].z, ts = vc[0].w;
                const V2 rA{vc[3].x, vc[3].y}, rB{vc[3].z, vc[3].w};
                const float nm = vc[4].z, tm = vc[4].w, bias = vc[5].x;
                const float4 ms = vc[8];
                const float mA = ms.x, iA = ms.y, mB = ms.z, iB = ms.w;
                const int bA = __float_as_int(vc[2].w), bB = __float_as_int(vc[5].z);
                const float4* g = t.gs + (slot * std::size_t(t.itV) + it) * 2;
                const V2 vA{g[0].x, g[0].y}, vB{g[1].x, g[1].y};
                const float wA = g[0].z, wB = g[1].z, ltOld = g[0].w, lnOld = g[1].w;
                atomicAdd(a.diag + 3, 1.0f);
                // Forward recomputation of this update.
                const V2 cl0 = (vB + mul(revperp(rB), wB)) - (vA + mul(revperp(rA), wA));
                const float vt = dot(cl0, tg);
                const float maxI = f * lnOld;
                const float rawT = ltOld + tm * (ts - vt);
                const bool lowT = rawT < -maxI, highT = !lowT && (maxI < rawT);
                const float ltNew = lowT ? -maxI : (highT ? maxI : rawT);
                const float incT = ltNew - ltOld;
                const V2 PT = mul(incT, tg);
                const V2 vA1 = vA - mul(mA, PT), vB1 = vB + mul(mB, PT);
                const float wA1 = wA - iA * cross(rA, PT), wB1 = wB + iB * cross(rB, PT);
                const V2 cl1 = (vB1 + mul(revperp(rB), wB1)) - (vA1 + mul(revperp(rA), wA1));
                const float vn = dot(cl1, n);
                const float rawN = lnOld + nm * (bias - vn);
                const float lnNew = (rawN < 0.0f) ? 0.0f : rawN;
                const float incN = lnNew - lnOld;
                // Adjoints of the outputs.
                float4 va4 = a.vbar[bo + bA], vb4 = a.vbar[bo + bB];
                V2 vbA = v2(va4), vbB = v2(vb4);
                float wbA = va4.z, wbB = vb4.z;
                float2 lb = a.lamtmp[slot];  // (normal, tangent)
                float4 c0 = a.cbar[slot * 3 + 0], c1 = a.cbar[slot * 3 + 1], c2 = a.cbar[slot * 3 + 2];
                V2 nbar{c0.x, c0.y}, rAbar{c0.z, c0.w}, rBbar{c1.x, c1.y}, tbar{c1.z, c1.w};
                float nmbar = c2.x, tmbar = c2.y, biasbar = c2.z, fbar = c2.w;
                // Normal update.
                {
                    const V2 gv = mul(mB, vbB) - mul(mA, vbA);
                    const float ginc = dot(gv, n) - iA * wbA * cross(rA, n) + iB * wbB * cross(rB, n);
                    nbar = nbar + mul(gv, incN) + mul(revperp(rA), -iA * wbA * incN) + mul(revperp(rB), iB * wbB * incN);
                    rAbar = rAbar + mul(perp_grad(n), -iA * wbA * incN);
                    rBbar = rBbar + mul(perp_grad(n), iB * wbB * incN);
                    const float lnTot = lb.x + ginc;
                    float lnOldBar = -ginc;
                    const float rawBar = (rawN < 0.0f) ? 0.0f : lnTot;
                    lnOldBar += rawBar;
                    nmbar += rawBar * (bias - vn);
                    biasbar += rawBar * nm;
                    const float vnbar = -rawBar * nm;
                    const V2 cbv = mul(n, vnbar);
                    nbar = nbar + mul(cl1, vnbar);
                    vbB = vbB + cbv;
                    wbB += cross(rB, cbv);
                    rBbar = rBbar + mul(perp_grad(cbv), wB1);
                    vbA = vbA - cbv;
                    wbA -= cross(rA, cbv);
                    rAbar = rAbar - mul(perp_grad(cbv), wA1);
                    lb.x = lnOldBar;
                }
                // Tangent update.
                {
                    const V2 gv = mul(mB, vbB) - mul(mA, vbA);
                    const float ginc = dot(gv, tg) - iA * wbA * cross(rA, tg) + iB * wbB * cross(rB, tg);
                    tbar = tbar + mul(gv, incT) + mul(revperp(rA), -iA * wbA * incT) + mul(revperp(rB), iB * wbB * incT);
                    rAbar = rAbar + mul(perp_grad(tg), -iA * wbA * incT);
                    rBbar = rBbar + mul(perp_grad(tg), iB * wbB * incT);
                    const float ltTot = lb.y + ginc;
                    float ltOldBar = -ginc;
                    float rawBar = 0.0f, maxBar = 0.0f;
                    if (lowT) maxBar = -ltTot;
                    else if (highT) maxBar = ltTot;
                    else rawBar = ltTot;
                    fbar += maxBar * lnOld;
                    lb.x += maxBar * f;
                    ltOldBar += rawBar;
                    tmbar += rawBar * (ts - vt);
                    const float vtbar = -rawBar * tm;
                    const V2 cbv = mul(tg, vtbar);
                    tbar = tbar + mul(cl0, vtbar);
                    vbB = vbB + cbv;
                    wbB += cross(rB, cbv);
                    rBbar = rBbar + mul(perp_grad(cbv), wB);
                    vbA = vbA - cbv;
                    wbA -= cross(rA, cbv);
                    rAbar = rAbar - mul(perp_grad(cbv), wA);
                    lb.y = ltOldBar;
                }
                if (mA != 0.0f || iA != 0.0f) a.vbar[bo + bA] = float4{vbA.x, vbA.y, wbA, va4.w};
                a.vbar[bo + bB] = float4{vbB.x, vbB.y, wbB, vb4.w};
                a.lamtmp[slot] = lb;
                a.cbar[slot * 3 + 0] = float4{nbar.x, nbar.y, rAbar.x, rAbar.y};
                a.cbar[slot * 3 + 1] = float4{rBbar.x, rBbar.y, tbar.x, tbar.y};
                a.cbar[slot * 3 + 2] = float4{nmbar, tmbar, biasbar, fbar};
            }
        }
        // Warm start, per slot.  Its deltas do not depend on velocities, so
        // the velocity adjoints here are those after every slot's warm start
        // and before any creation-time term below is added: constraint
        // creation (the restitution bias reads the velocities) precedes the
        // whole island's warm start in the forward step.
        for (int e = rng.x; e < rng.y; ++e) {
            const std::size_t slot = std::size_t(co) + e;
            const float4* vc = t.vc + slot * prx::kTapeVc;
            if ((__float_as_uint(vc[10].z) >> 8) != 1u) continue;
            const V2 n = v2(vc[0]);
            const V2 tg = fwdperp(n);
            const V2 rA{vc[3].x, vc[3].y}, rB{vc[3].z, vc[3].w};
            const float lnInit = vc[4].x, ltInit = vc[4].y;
            const float4 ms = vc[8];
            const float mA = ms.x, iA = ms.y, mB = ms.z, iB = ms.w;
            const int bA = __float_as_int(vc[2].w), bB = __float_as_int(vc[5].z);
            float2 lb = a.lamtmp[slot];
            if (c.warmStart) {
                const float4 va4 = a.vbar[bo + bA], vb4 = a.vbar[bo + bB];
                const V2 vbA = v2(va4), vbB = v2(vb4);
                const float wbA = va4.z, wbB = vb4.z;
                const float4 c0 = a.cbar[slot * 3 + 0], c1 = a.cbar[slot * 3 + 1];
                V2 nbar{c0.x, c0.y}, rAbar{c0.z, c0.w}, rBbar{c1.x, c1.y}, tbar{c1.z, c1.w};
                const V2 P = mul(lnInit, n) + mul(ltInit, tg);
                const V2 Pbar = mul(mB, vbB) - mul(mA, vbA) + mul(revperp(rA), -iA * wbA) + mul(revperp(rB), iB * wbB);
                rAbar = rAbar + mul(perp_grad(P), -iA * wbA);
                rBbar = rBbar + mul(perp_grad(P), iB * wbB);
                lb.x += dot(Pbar, n);
                lb.y += dot(Pbar, tg);
                nbar = nbar + mul(Pbar, lnInit);
                tbar = tbar + mul(Pbar, ltInit);
                a.cbar[slot * 3 + 0] = float4{nbar.x, nbar.y, rAbar.x, rAbar.y};
                a.cbar[slot * 3 + 1] = float4{rBbar.x, rBbar.y, tbar.x, tbar.y};
            }
            // Impulses as carried from the previous step (dtRatio scaling is
            // part of the record: warm start uses them as stored).
            const std::uint32_t id = __float_as_uint(vc[10].y);
            a.lambar[co + int(id)] = float4{lb.x, lb.y, lnInit, ltInit};
        }
        // Constraint creation, per slot.
        for (int e = rng.x; e < rng.y; ++e) {
            const std::size_t slot = std::size_t(co) + e;
            const float4* vc = t.vc + slot * prx::kTapeVc;
            if ((__float_as_uint(vc[10].z) >> 8) != 1u) continue;
            const std::uint32_t type = __float_as_uint(vc[10].z) & 0xffu;
            const V2 n = v2(vc[0]);
            const V2 tg = fwdperp(n);
            const V2 rA{vc[3].x, vc[3].y}, rB{vc[3].z, vc[3].w};
            const float nm = vc[4].z, tm = vc[4].w;
            const float4 ms = vc[8];
            const float mA = ms.x, iA = ms.y, mB = ms.z, iB = ms.w;
            const int bA = __float_as_int(vc[2].w), bB = __float_as_int(vc[5].z);
            const float4 va4 = a.vbar[bo + bA], vb4 = a.vbar[bo + bB];
            V2 vbA = v2(va4), vbB = v2(vb4);
            float wbA = va4.z, wbB = vb4.z;
            const float4 c0 = a.cbar[slot * 3 + 0], c1 = a.cbar[slot * 3 + 1], c2 = a.cbar[slot * 3 + 2];
            V2 nbar{c0.x, c0.y}, rAbar{c0.z, c0.w}, rBbar{c1.x, c1.y}, tbar{c1.z, c1.w};
            const float nmbar = c2.x, tmbar = c2.y, biasbar = c2.z, fbarC = c2.w;
            // Effective masses (VelocityConstraint.cpp:136-185).
            {
                const float cnA = cross(rA, n), cnB = cross(rB, n);
                const float K = mA + mB + iA * cnA * cnA + iB * cnB * cnB;
                if (K > 0.0f) {
                    const float Kbar = -nmbar * nm * nm;
                    const float gA = 2.0f * iA * cnA * Kbar, gB = 2.0f * iB * cnB * Kbar;
                    rAbar = rAbar + mul(perp_grad(n), gA);
                    nbar = nbar + mul(revperp(rA), gA);
                    rBbar = rBbar + mul(perp_grad(n), gB);
                    nbar = nbar + mul(revperp(rB), gB);
                }
                const float ctA = cross(rA, tg), ctB = cross(rB, tg);
                const float Kt = mA + mB + iA * ctA * ctA + iB * ctB * ctB;
                if (Kt > 0.0f) {
                    const float Kbar = -tmbar * tm * tm;
                    const float gA = 2.0f * iA * ctA * Kbar, gB = 2.0f * iB * ctB * Kbar;
                    rAbar = rAbar + mul(perp_grad(tg), gA);
                    tbar = tbar + mul(revperp(rA), gA);
                    rBbar = rBbar + mul(perp_grad(tg), gB);
                    tbar = tbar + mul(revperp(rB), gB);
                }
            }
            // Restitution bias from the closing speed at creation.
            const float vn0 = vc[10].x;
            const float rest = vc[5].w;
            const float4 vA0 = vc[12], vB0 = vc[13];
            atomicAdd(a.diag + 0, 1.0f);
            if (vn0 < -c.velocityThreshold) {
                atomicAdd(a.diag + 1, 1.0f);
                const float ebarC = -vn0 * biasbar;
                const float vn0bar = -rest * biasbar;
                const V2 cl0 = (V2{vB0.x, vB0.y} + mul(revperp(rB), vB0.z)) - (V2{vA0.x, vA0.y} + mul(revperp(rA), vA0.z));
                const V2 cbv = mul(n, vn0bar);
                nbar = nbar + mul(cl0, vn0bar);
                vbB = vbB + cbv;
                wbB += cross(rB, cbv);
                rBbar = rBbar + mul(perp_grad(cbv), vB0.z);
                vbA = vbA - cbv;
                wbA -= cross(rA, cbv);
                rAbar = rAbar - mul(perp_grad(cbv), vA0.z);
                // Restitution = max(eA, eB), ties to B (Contact.hpp:66-69).
                const int sA = __float_as_int(vc[11].y), sB = __float_as_int(vc[11].z);
                const float eA = d.shape_f[sA].z, eB = d.shape_f[sB].z;
                atomicAdd(a.ebar + ((eA > eB) ? sA : sB), ebarC);
            }
            // Friction = sqrt(fA fB) (Contact.hpp:57-60).
            {
                const int sA = __float_as_int(vc[11].y), sB = __float_as_int(vc[11].z);
                const float fA = d.shape_f[sA].y, fB = d.shape_f[sB].y;
                const float fm = vc[0].z;
                if (fm > 0.0f) {
                    atomicAdd(a.fbar + sA, fbarC * fB / (2.0f * fm));
                    atomicAdd(a.fbar + sB, fbarC * fA / (2.0f * fm));
                }
            }
            if (mA != 0.0f || iA != 0.0f) a.vbar[bo + bA] = float4{vbA.x, vbA.y, wbA, va4.w};
            a.vbar[bo + bB] = float4{vbB.x, vbB.y, wbB, vb4.w};
            // Geometry (WorldManifold.cpp:37-130): tangent from the normal,
            // then the normal and points from the positions.
            nbar.x += -tbar.y;
            nbar.y += tbar.x;
            const float4 pos = vc[9];
            const float rAd = vc[10].w, rBd = vc[11].x;
            if (type == 1u) {
                const V2 dd = V2{pos.z - pos.x, pos.w - pos.y};
                const float L = __builtin_sqrtf(magsq(dd));
                if (L > 0.0f) {
                    const float k = 0.5f * (rAd - rBd);
                    V2 dbar = mul(rAbar - rBbar, 0.5f);
                    const V2 ntot = nbar + mul(rAbar + rBbar, k);
                    const float nn = dot(n, ntot);
                    dbar = dbar + mul(ntot - mul(n, nn), 1.0f / L);
                    if (mA != 0.0f || iA != 0.0f) add_body(a.xbar + bo + bA, -dbar, 0.0f);
                    add_body(a.xbar + bo + bB, dbar, 0.0f);
                }
            } else if (type == 2u) {
                // relB = ((rA - dd) - rB) / 2 n with dd = dot(cB - plane, n);
                // the normal and plane belong to the static body.
                const float ddbar = -0.5f * dot(rBbar, n);
                add_body(a.xbar + bo + bB, mul(n, ddbar), 0.0f);
            }
        }
    }
    __syncthreads();
    // Velocity integration, reversed (damping factors; caps not active).
    for (int b = lane; b < nb; b += kWave) {
        if (!t.inisl[bo + b]) continue;
        const int gb = bo + b;
        if (!(d.bflags[gb] & prx::kAccelerable)) continue;
        const float2 dm = d.damp[gb];
        const float fl = 1.0f / (1.0f + c.dt * dm.x);
        const float fa = 1.0f / (1.0f + c.dt * dm.y);
        float4 v = a.vbar[gb];
        v.x *= fl;
        v.y *= fl;
        v.z *= fa;
        a.vbar[gb] = v;
    }
}

}  // namespace prxd

extern "C" hipError_t prx_launch_adjoint_step(const prx::DeviceState* d, const prx::Tape* t, const prx::AdjState* a,
                                              const prx::StepParams* c, int nworlds, hipStream_t s) {
    hipLaunchKernelGGL(prxd::k_adjoint_step, dim3(unsigned(nworlds)), dim3(64), 0, s, *d, *t, *a, *c);
    return hipGetLastError();
}
