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
#include "adjoint_dual.hpp"
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
    // step's final impulses warm-started the next step where the manifold
    // update after the solve carried them (by ContactFeature, else from the
    // nearest point; the tape's carry map says which).  The next step wrote
    // nothing for contacts it did not solve or that started afresh.
    for (int e = lane; e < nslots; e += kWave) {
        const float4* vc = t.vc + (std::size_t(co) + e) * prx::kTapeVc;
        const std::uint32_t id = __float_as_uint(vc[10].y);
        const float4 lb = a.lambar[co + int(id)];
        const std::uint32_t cm = t.cmap[co + int(id)];
        const bool carried = ((cm >> 8) & 0xfu) >= 1u && (cm & 0xfu) == 0u;
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
                const float4* g = t.gs + (slot * std::size_t(t.itV) + it) * 3;
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
            // part of the record: warm start uses them as stored); nothing
            // for a contact created before this step's solve.
            const std::uint32_t id = __float_as_uint(vc[10].y);
            const bool fresh = (t.cmap[co + int(id)] & prx::kCarryFresh) != 0u;
            a.lambar[co + int(id)] = fresh ? float4{0.0f, 0.0f, 0.0f, 0.0f} : float4{lb.x, lb.y, lnInit, ltInit};
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

// ---------------------------------------------------------------------------
// General form: any manifold type and point count, revolute joints, bodies
// with local centres.  The active set is held fixed as recorded; each solver
// update is reversed by its vector-Jacobian product, which the wave computes
// from the update's dual copy (adjoint_dual.hpp): lane i evaluates the copy
// with input i seeded and returns sum_j ybar_j d y_j / d x_i.  One wave per
// world walks the recorded updates in reverse order.

template <int N, int M, class F>
__device__ __forceinline__ float lane_vjp(const float (&x)[N], const float (&ybar)[M], F f) {
    const int lane = int(threadIdx.x);
    dl::D in[N];
#pragma unroll
    for (int i = 0; i < N; ++i) in[i] = dl::D{x[i], (i == lane) ? 1.0f : 0.0f};
    dl::D out[M];
    f(in, out);
    float s = 0.0f;
#pragma unroll
    for (int j = 0; j < M; ++j) s += ybar[j] * out[j].d;
    return (lane < N) ? s : 0.0f;
}

// Velocity-constraint values a contact's updates read, in the order of the
// creation outputs (kAdjFields): normal x, y; friction; point 0 relA, relB,
// ni, ti, normal mass, tangent mass, bias; point 1 the same; K00, K11, K01,
// M00, M11, M01.
__device__ __forceinline__ void vc_fields(const float4* v, float (&f)[27]) {
    f[0] = v[0].x; f[1] = v[0].y; f[2] = v[0].z;
    f[3] = v[3].x; f[4] = v[3].y; f[5] = v[3].z; f[6] = v[3].w;
    f[7] = v[4].x; f[8] = v[4].y; f[9] = v[4].z; f[10] = v[4].w; f[11] = v[5].x;
    f[12] = v[6].x; f[13] = v[6].y; f[14] = v[6].z; f[15] = v[6].w;
    f[16] = v[7].x; f[17] = v[7].y; f[18] = v[7].z; f[19] = v[7].w; f[20] = v[5].y;
    f[21] = v[1].x; f[22] = v[1].y; f[23] = v[1].z; f[24] = v[2].x; f[25] = v[2].y; f[26] = v[2].z;
}
__device__ __forceinline__ dl::Vcd vcd_from(const dl::D* f, const float4* v) {
    dl::Vcd vc;
    vc.normal = dl::V{f[0], f[1]};
    vc.friction = f[2];
    vc.tangentSpeed = v[0].w;
    vc.n = __float_as_uint(v[1].w);
    vc.p[0] = dl::VcPointD{dl::V{f[3], f[4]}, dl::V{f[5], f[6]}, f[7], f[8], f[9], f[10], f[11]};
    vc.p[1] = dl::VcPointD{dl::V{f[12], f[13]}, dl::V{f[14], f[15]}, f[16], f[17], f[18], f[19], f[20]};
    vc.K00 = f[21]; vc.K11 = f[22]; vc.K01 = f[23]; vc.M00 = f[24]; vc.M11 = f[25]; vc.M01 = f[26];
    return vc;
}

// Joint record values its updates read (kAdjJointFields): rA, rB; exx, eyx,
// ezx, eyy, ezy, ezz; angular mass; impulses x, y, z, motor; max impulse.
__device__ __forceinline__ void jc_fields(const float4* u, float (&f)[16]) {
    f[0] = u[0].x; f[1] = u[0].y; f[2] = u[0].z; f[3] = u[0].w;
    f[4] = u[1].x; f[5] = u[1].y; f[6] = u[1].z; f[7] = u[1].w;
    f[8] = u[2].x; f[9] = u[2].y; f[10] = u[2].z;
    f[11] = u[3].x; f[12] = u[3].y; f[13] = u[3].z; f[14] = u[3].w;
    f[15] = u[5].y;
}
__device__ __forceinline__ dl::Jd jd_from(const dl::D* f) {
    dl::Jd j;
    j.rA = dl::V{f[0], f[1]};
    j.rB = dl::V{f[2], f[3]};
    j.exx = f[4]; j.eyx = f[5]; j.ezx = f[6]; j.eyy = f[7]; j.ezy = f[8]; j.ezz = f[9];
    j.angularMass = f[10];
    j.ix = f[11]; j.iy = f[12]; j.iz = f[13]; j.motorImp = f[14];
    j.maxImpulse = f[15];
    return j;
}

__device__ __forceinline__ void load3(const float4& v, float* o) {
    o[0] = v.x;
    o[1] = v.y;
    o[2] = v.z;
}
__device__ __forceinline__ void load4(const float4& v, float* o) {
    o[0] = v.x;
    o[1] = v.y;
    o[2] = v.z;
    o[3] = v.w;
}
// Lane k < 3 of a body's three adjoint components (x, y, angle).
__device__ __forceinline__ float& comp(float4& v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : v.z); }

__global__ __launch_bounds__(64) void k_adjoint_general(DeviceState d, Tape t, AdjState a, StepParams c) {
    const int w = int(blockIdx.x);
    const int lane = int(threadIdx.x);
    const WorldHeader H = d.hdr[w];
    const int bo = H.body_off, co = H.contact_off, nb = H.nbodies, cap = H.contact_cap, jo = H.joint_off;
    const int nisl = t.nisl[w];
    const int4* isl = t.isl + bo + w;
    const int2* islj = t.islj + bo + w;
    const int nslots = nisl > 0 ? isl[nisl - 1].y : 0;
    const int njslots = nisl > 0 ? islj[nisl - 1].y : 0;
    float4* xbar = a.xbar + bo;
    float4* vbar = a.vbar + bo;
    auto vcrec = [&](int e) { return t.vc + (std::size_t(co) + std::size_t(e)) * prx::kTapeVc; };
    auto jcrec = [&](int e) { return t.jc + (std::size_t(jo) + std::size_t(e)) * prx::kTapeJc; };

    // Impulse adjoints carried in from the next step.  Contacts: by pool id,
    // each of the next step's points to the point of this step its impulses
    // came from in the manifold update after the solve (the tape's carry
    // map: by ContactFeature, else the nearest point; two points may share
    // one source).  Joints: by joint.
    for (int e = lane; e < nslots; e += kWave) {
        const float4* vc = vcrec(e);
        const std::uint32_t id = __float_as_uint(vc[10].y);
        const float4 lb = a.lamc[(std::size_t(co) + id) * 2];
        const std::uint32_t cm = t.cmap[std::size_t(co) + id];
        const std::uint32_t nnext = (cm >> 8) & 0xfu;
        const float bn[2] = {lb.x, lb.z}, bt[2] = {lb.y, lb.w};
        float r[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        for (std::uint32_t j = 0; j < 2; ++j) {
            const std::uint32_t k = (cm >> (4 * j)) & 0xfu;
            if (j < nnext && k < 2u) {
                r[2 * k] += bn[j];
                r[2 * k + 1] += bt[j];
            }
        }
        a.lam4[std::size_t(co) + e] = float4{r[0], r[1], r[2], r[3]};
        float* fl = a.fld + (std::size_t(co) + e) * prx::kAdjFields;
        for (int k = 0; k < prx::kAdjFields; ++k) fl[k] = 0.0f;
    }
    for (int e = lane; e < njslots; e += kWave) {
        const int jg = __float_as_int(jcrec(e)[7].w);
        a.jlam[std::size_t(jo) + e] = a.jcarry[jg];
        float* fl = a.jfld + (std::size_t(jo) + e) * prx::kAdjJointFields;
        for (int k = 0; k < prx::kAdjJointFields; ++k) fl[k] = 0.0f;
    }
    __syncthreads();
    for (int i = lane; i < cap; i += kWave) {
        a.lamc[(std::size_t(co) + i) * 2] = float4{0.0f, 0.0f, 0.0f, 0.0f};
        a.lamc[(std::size_t(co) + i) * 2 + 1] = float4{0.0f, 0.0f, 0.0f, 0.0f};
    }
    __syncthreads();

    // Position sweeps, reversed: per iteration the joints (after the
    // contacts in the forward pass), then the contacts, each in reverse.
    for (int is = 0; is < nisl; ++is) {
        const int4 rng = isl[is];
        const int2 jr = islj[is];
        for (int it = rng.w - 1; it >= 0; --it) {
            for (int e = jr.y - 1; e >= jr.x; --e) {
                const float4* u = jcrec(e);
                const int bA = __float_as_int(u[5].z), bB = __float_as_int(u[5].w);
                const float4* g = t.jngs + ((std::size_t(jo) + e) * std::size_t(t.itP) + it) * 2;
                float x[6], yb[6];
                load3(g[0], x);
                load3(g[1], x + 3);
                const float4 xa = xbar[bA], xb = xbar[bB];
                load3(xa, yb);
                load3(xb, yb + 3);
                const float r = lane_vjp<6, 6>(x, yb, [&](dl::D* in, dl::D* out) {
                    dl::joint_solve_position(u, in, c);
                    for (int k = 0; k < 6; ++k) out[k] = in[k];
                });
                __syncthreads();
                if (lane < 3) comp(xbar[bA], lane) = r;
                else if (lane < 6) comp(xbar[bB], lane - 3) = r;
                __syncthreads();
            }
            for (int e = rng.y - 1; e >= rng.x; --e) {
                const std::size_t slot = std::size_t(co) + e;
                const float4* vc = vcrec(e);
                const Pc pc = load_pc(vc + 17);
                const float4* g = t.ngs + (slot * std::size_t(t.itP) + it) * 2;
                // Positions before the update, then the manifold's local
                // points (their adjoints collect for the constraint creation).
                float x[10], yb[6];
                load3(g[0], x);
                load3(g[1], x + 3);
                x[6] = pc.m.p[0].lp.x;
                x[7] = pc.m.p[0].lp.y;
                x[8] = pc.m.p[1].lp.x;
                x[9] = pc.m.p[1].lp.y;
                load3(xbar[pc.bA], yb);
                load3(xbar[pc.bB], yb + 3);
                const float r = lane_vjp<10, 6>(x, yb, [&](dl::D* in, dl::D* out) {
                    const dl::V lpd[2] = {dl::V{in[6], in[7]}, dl::V{in[8], in[9]}};
                    dl::solve_pc(pc.m, pc.totalRadius, pc.A, pc.B, in, c, lpd);
                    for (int k = 0; k < 6; ++k) out[k] = in[k];
                });
                __syncthreads();
                if (lane < 3) comp(xbar[pc.bA], lane) = r;
                else if (lane < 6) comp(xbar[pc.bB], lane - 3) = r;
                else if (lane < 10) a.fld[slot * prx::kAdjFields + 28 + (lane - 6)] += r;
                __syncthreads();
                if (lane == 0) atomicAdd(a.diag + 2, 1.0f);
            }
        }
    }
    __syncthreads();
    // Position integration, reversed: x1 = x0 + h v.
    for (int b = lane; b < nb; b += kWave) {
        if (!t.inisl[bo + b]) continue;
        const float4 xb = xbar[b];
        add_body(vbar + b, V2{c.dt * xb.x, c.dt * xb.y}, c.dt * xb.z);
    }
    __syncthreads();

    for (int is = 0; is < nisl; ++is) {
        const int4 rng = isl[is];
        const int2 jr = islj[is];
        // Velocity sweeps, reversed: per iteration the contacts, then the
        // joints (solved first in the forward pass), each in reverse.
        for (int it = rng.z - 1; it >= 0; --it) {
            for (int e = rng.y - 1; e >= rng.x; --e) {
                const std::size_t slot = std::size_t(co) + e;
                const float4* vc = vcrec(e);
                const int bA = __float_as_int(vc[2].w), bB = __float_as_int(vc[5].z);
                const BodyC A{vc[8].x, vc[8].y, V2{0, 0}}, B{vc[8].z, vc[8].w, V2{0, 0}};
                const float4* g = t.gs + (slot * std::size_t(t.itV) + it) * 3;
                float x[33], yb[10];
                load3(g[0], x);
                load3(g[1], x + 3);
                load4(g[2], x + 6);
                float f[27];
                vc_fields(vc, f);
                // Fields the update reads (creation-output indices).
                constexpr int kMap[23] = {0, 1, 2, 3, 4, 5, 6, 9, 10, 11, 12, 13, 14, 15, 18, 19, 20, 21, 22, 23, 24, 25, 26};
#pragma unroll
                for (int k = 0; k < 23; ++k) x[10 + k] = f[kMap[k]];
                load3(vbar[bA], yb);
                load3(vbar[bB], yb + 3);
                load4(a.lam4[slot], yb + 6);
                const float r = lane_vjp<33, 10>(x, yb, [&](dl::D* in, dl::D* out) {
                    dl::D fd[27];
                    for (int k = 0; k < 27; ++k) fd[k] = dl::cst(f[k]);
#pragma unroll
                    for (int k = 0; k < 23; ++k) fd[kMap[k]] = in[10 + k];
                    fd[7] = in[6];
                    fd[8] = in[7];
                    fd[16] = in[8];
                    fd[17] = in[9];
                    dl::Vcd v = vcd_from(fd, vc);
                    dl::D va[3] = {in[0], in[1], in[2]}, vb[3] = {in[3], in[4], in[5]};
                    dl::solve_velocity(v, A, B, va, vb);
                    for (int k = 0; k < 3; ++k) {
                        out[k] = va[k];
                        out[3 + k] = vb[k];
                    }
                    out[6] = v.p[0].ni;
                    out[7] = v.p[0].ti;
                    out[8] = v.p[1].ni;
                    out[9] = v.p[1].ti;
                });
                __syncthreads();
                if (lane < 3) comp(vbar[bA], lane) = r;
                else if (lane < 6) comp(vbar[bB], lane - 3) = r;
                else if (lane < 10) (&a.lam4[slot].x)[lane - 6] = r;
                else if (lane < 33) a.fld[slot * prx::kAdjFields + kMap[lane - 10]] += r;
                __syncthreads();
                if (lane == 0) atomicAdd(a.diag + 3, 1.0f);
            }
            for (int e = jr.y - 1; e >= jr.x; --e) {
                const std::size_t js = std::size_t(jo) + e;
                const float4* u = jcrec(e);
                const int bA = __float_as_int(u[5].z), bB = __float_as_int(u[5].w);
                const float4* g = t.jgs + (js * std::size_t(t.itV) + it) * 3;
                float f[16];
                jc_fields(u, f);
                float x[22], yb[10];
                load3(g[0], x);
                load3(g[1], x + 3);
                load4(g[2], x + 6);
                constexpr int kJMap[12] = {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 15};
#pragma unroll
                for (int k = 0; k < 12; ++k) x[10 + k] = f[kJMap[k]];
                load3(vbar[bA], yb);
                load3(vbar[bB], yb + 3);
                load4(a.jlam[js], yb + 6);
                const float r = lane_vjp<22, 10>(x, yb, [&](dl::D* in, dl::D* out) {
                    dl::D fd[16];
                    for (int k = 0; k < 16; ++k) fd[k] = dl::cst(f[k]);
#pragma unroll
                    for (int k = 0; k < 12; ++k) fd[kJMap[k]] = in[10 + k];
                    fd[11] = in[6];
                    fd[12] = in[7];
                    fd[13] = in[8];
                    fd[14] = in[9];
                    dl::Jd j = jd_from(fd);
                    dl::D va[3] = {in[0], in[1], in[2]}, vb[3] = {in[3], in[4], in[5]};
                    dl::joint_solve_velocity(j, u, va, vb);
                    for (int k = 0; k < 3; ++k) {
                        out[k] = va[k];
                        out[3 + k] = vb[k];
                    }
                    out[6] = j.ix;
                    out[7] = j.iy;
                    out[8] = j.iz;
                    out[9] = j.motorImp;
                });
                __syncthreads();
                if (lane < 3) comp(vbar[bA], lane) = r;
                else if (lane < 6) comp(vbar[bB], lane - 3) = r;
                else if (lane < 10) (&a.jlam[js].x)[lane - 6] = r;
                else if (lane < 22) a.jfld[js * prx::kAdjJointFields + kJMap[lane - 10]] += r;
                __syncthreads();
            }
        }
        // Joint warm start, reversed (after the contacts' in the forward
        // pass; linear in the velocities, so their values do not enter).
        for (int e = jr.y - 1; e >= jr.x; --e) {
            const std::size_t js = std::size_t(jo) + e;
            const float4* u = jcrec(e);
            const int bA = __float_as_int(u[5].z), bB = __float_as_int(u[5].w);
            if (c.warmStart) {
                float f[16];
                jc_fields(u, f);
                float x[14] = {0, 0, 0, 0, 0, 0, f[0], f[1], f[2], f[3], f[11], f[12], f[13], f[14]};
                float yb[6];
                load3(vbar[bA], yb);
                load3(vbar[bB], yb + 3);
                const float r = lane_vjp<14, 6>(x, yb, [&](dl::D* in, dl::D* out) {
                    dl::D fd[16];
                    for (int k = 0; k < 16; ++k) fd[k] = dl::cst(f[k]);
                    for (int k = 0; k < 4; ++k) fd[k] = in[6 + k];
                    for (int k = 0; k < 4; ++k) fd[11 + k] = in[10 + k];
                    const dl::Jd j = jd_from(fd);
                    dl::D va[3] = {in[0], in[1], in[2]}, vb[3] = {in[3], in[4], in[5]};
                    dl::joint_warm_start(j, u, va, vb);
                    for (int k = 0; k < 3; ++k) {
                        out[k] = va[k];
                        out[3 + k] = vb[k];
                    }
                });
                __syncthreads();
                if (lane >= 6 && lane < 10) a.jfld[js * prx::kAdjJointFields + (lane - 6)] += r;
                else if (lane >= 10 && lane < 14) a.jfld[js * prx::kAdjJointFields + 11 + (lane - 10)] += r;
                __syncthreads();
            }
            // The sweeps started from the set-up's impulses.
            if (lane < 4) a.jfld[js * prx::kAdjJointFields + 11 + lane] += (&a.jlam[js].x)[lane];
            __syncthreads();
        }
        // Contact warm start, reversed (linear in the velocities).
        for (int e = rng.y - 1; e >= rng.x; --e) {
            const std::size_t slot = std::size_t(co) + e;
            const float4* vc = vcrec(e);
            const int bA = __float_as_int(vc[2].w), bB = __float_as_int(vc[5].z);
            const BodyC A{vc[8].x, vc[8].y, V2{0, 0}}, B{vc[8].z, vc[8].w, V2{0, 0}};
            if (c.warmStart) {
                float f[27];
                vc_fields(vc, f);
                constexpr int kWMap[14] = {0, 1, 3, 4, 5, 6, 7, 8, 12, 13, 14, 15, 16, 17};
                float x[20] = {0, 0, 0, 0, 0, 0};
#pragma unroll
                for (int k = 0; k < 14; ++k) x[6 + k] = f[kWMap[k]];
                float yb[6];
                load3(vbar[bA], yb);
                load3(vbar[bB], yb + 3);
                const float r = lane_vjp<20, 6>(x, yb, [&](dl::D* in, dl::D* out) {
                    dl::D fd[27];
                    for (int k = 0; k < 27; ++k) fd[k] = dl::cst(f[k]);
#pragma unroll
                    for (int k = 0; k < 14; ++k) fd[kWMap[k]] = in[6 + k];
                    const dl::Vcd v = vcd_from(fd, vc);
                    dl::D va[3] = {in[0], in[1], in[2]}, vb[3] = {in[3], in[4], in[5]};
                    dl::warm_start(v, A, B, va, vb);
                    for (int k = 0; k < 3; ++k) {
                        out[k] = va[k];
                        out[3 + k] = vb[k];
                    }
                });
                __syncthreads();
                if (lane >= 6 && lane < 20) a.fld[slot * prx::kAdjFields + kWMap[lane - 6]] += r;
                __syncthreads();
            }
            // The sweeps started from the warm-start impulses.
            if (lane < 4) {
                constexpr int kImp[4] = {7, 8, 16, 17};
                a.fld[slot * prx::kAdjFields + kImp[lane]] += (&a.lam4[slot].x)[lane];
            }
            __syncthreads();
        }
        // Joint set-up, reversed: angles, the impulses it started from and
        // the maximum motor torque.
        for (int e = jr.y - 1; e >= jr.x; --e) {
            const std::size_t js = std::size_t(jo) + e;
            const float4* u = jcrec(e);
            const int bA = __float_as_int(u[5].z), bB = __float_as_int(u[5].w);
            const int jg = __float_as_int(u[7].w);
            const float4 in8 = u[prx::kJcStride], raw = u[prx::kJcStride + 1];
            float x[7] = {in8.x, in8.y, raw.x, raw.y, raw.z, raw.w, in8.z};
            float yb[16];
            for (int k = 0; k < 16; ++k) yb[k] = a.jfld[js * prx::kAdjJointFields + k];
            const std::uint32_t prevLimit = __float_as_uint(in8.w);
            const float r = lane_vjp<7, 16>(x, yb, [&](dl::D* in, dl::D* out) {
                const dl::Jd j = dl::build_jc(in[0], in[1], in + 2, in[6], u, prevLimit, c);
                out[0] = j.rA.x; out[1] = j.rA.y; out[2] = j.rB.x; out[3] = j.rB.y;
                out[4] = j.exx; out[5] = j.eyx; out[6] = j.ezx; out[7] = j.eyy; out[8] = j.ezy; out[9] = j.ezz;
                out[10] = j.angularMass;
                out[11] = j.ix; out[12] = j.iy; out[13] = j.iz; out[14] = j.motorImp;
                out[15] = j.maxImpulse;
            });
            __syncthreads();
            if (lane == 0) xbar[bA].z += r;
            else if (lane == 1) xbar[bB].z += r;
            else if (lane < 6) (&a.jcarry[jg].x)[lane - 2] = r;
            else if (lane == 6) a.tbar[jg] += r;
            __syncthreads();
        }
        // Contact constraint creation, reversed: positions and velocities at
        // creation, materials, and the manifold impulses it started from.
        for (int e = rng.y - 1; e >= rng.x; --e) {
            const std::size_t slot = std::size_t(co) + e;
            const float4* vc = vcrec(e);
            const Pc pc = load_pc(vc + 17);
            const std::uint32_t id = __float_as_uint(vc[10].y);
            const float rA = vc[10].w, rB = vc[11].x;
            float x[22];
            load3(vc[15], x);
            load3(vc[16], x + 3);
            load3(vc[12], x + 6);
            load3(vc[13], x + 9);
            x[12] = vc[0].z;
            x[13] = vc[5].w;
            load4(vc[22], x + 14);
            x[18] = pc.m.p[0].lp.x;
            x[19] = pc.m.p[0].lp.y;
            x[20] = pc.m.p[1].lp.x;
            x[21] = pc.m.p[1].lp.y;
            float yb[27];
            for (int k = 0; k < 27; ++k) yb[k] = a.fld[slot * prx::kAdjFields + k];
            const float r = lane_vjp<22, 27>(x, yb, [&](dl::D* in, dl::D* out) {
                const dl::X xfA = dl::get_xf(in[0], in[1], in[2], pc.A.lc);
                const dl::X xfB = dl::get_xf(in[3], in[4], in[5], pc.B.lc);
                const dl::V lpd[2] = {dl::V{in[18], in[19]}, dl::V{in[20], in[21]}};
                const dl::WMd wm = dl::world_manifold(pc.m, xfA, rA, xfB, rB, lpd);
                const dl::Vcd v = dl::make_vc(wm, in[12], in[13], vc[0].w, pc.A, in, in + 6, pc.B, in + 3, in + 9,
                                              in + 14, c);
                out[0] = v.normal.x; out[1] = v.normal.y; out[2] = v.friction;
                for (int k = 0; k < 2; ++k) {
                    const dl::VcPointD& p = v.p[k];
                    dl::D* o = out + 3 + 9 * k;
                    o[0] = p.relA.x; o[1] = p.relA.y; o[2] = p.relB.x; o[3] = p.relB.y;
                    o[4] = p.ni; o[5] = p.ti; o[6] = p.nm; o[7] = p.tm; o[8] = p.bias;
                }
                out[21] = v.K00; out[22] = v.K11; out[23] = v.K01; out[24] = v.M00; out[25] = v.M11; out[26] = v.M01;
            });
            __syncthreads();
            if (lane < 3) comp(xbar[pc.bA], lane) += r;
            else if (lane < 6) comp(xbar[pc.bB], lane - 3) += r;
            else if (lane < 9) comp(vbar[pc.bA], lane - 6) += r;
            else if (lane < 12) comp(vbar[pc.bB], lane - 9) += r;
            else if (lane == 12 || lane == 13) {
                const int sA = __float_as_int(vc[11].y), sB = __float_as_int(vc[11].z);
                if (lane == 12) {
                    // Friction = sqrt(fA fB) (Contact.hpp:57-60).
                    const float fA = d.shape_f[sA].y, fB = d.shape_f[sB].y, fm = vc[0].z;
                    if (fm > 0.0f) {
                        atomicAdd(a.fbar + sA, r * fB / (2.0f * fm));
                        atomicAdd(a.fbar + sB, r * fA / (2.0f * fm));
                    }
                } else {
                    // Restitution = max(eA, eB), ties to B (Contact.hpp:66-69).
                    const float eA = d.shape_f[sA].z, eB = d.shape_f[sB].z;
                    atomicAdd(a.ebar + ((eA > eB) ? sA : sB), r);
                }
            } else if (lane < 18) {
                // Nothing for a contact created before this step's solve.
                const bool fresh = (t.cmap[std::size_t(co) + id] & prx::kCarryFresh) != 0u;
                (&a.lamc[(std::size_t(co) + id) * 2].x)[lane - 14] = fresh ? 0.0f : r;
                (&a.lamc[(std::size_t(co) + id) * 2 + 1].x)[lane - 14] = (&vc[22].x)[lane - 14];
            } else if (lane < 22) {
                a.fld[slot * prx::kAdjFields + 28 + (lane - 18)] += r;
            }
            __syncthreads();
            // The manifold's local points as made by the update before this
            // step, at the same positions: points the clipping made move
            // with the bodies.
            const std::uint32_t mtype = pc.m.type;
            const float4 cfs = vc[23];
            if (mtype == kFaceA || mtype == kFaceB) {
                const bool faceB = mtype == kFaceB;
                const int sA = __float_as_int(vc[11].y), sB = __float_as_int(vc[11].z);
                const Proxy s0 = shape_proxy(d, faceB ? sB : sA), s1 = shape_proxy(d, faceB ? sA : sB);
                float xm[6], ybm[4];
                load3(vc[15], xm);
                load3(vc[16], xm + 3);
                for (int k = 0; k < 4; ++k) ybm[k] = a.fld[slot * prx::kAdjFields + 28 + k];
                const float rm = lane_vjp<6, 4>(xm, ybm, [&](dl::D* in, dl::D* out) {
                    const dl::X xfA = dl::get_xf(in[0], in[1], in[2], pc.A.lc);
                    const dl::X xfB = dl::get_xf(in[3], in[4], in[5], pc.B.lc);
                    for (std::uint32_t k = 0; k < 2; ++k) {
                        dl::V lp = dl::vc2(pc.m.p[k].lp);
                        if (k < pc.m.n)
                            lp = dl::manifold_local_point(s0, faceB ? xfB : xfA, s1, faceB ? xfA : xfB,
                                                          __float_as_uint(k ? cfs.y : cfs.x), faceB, pc.m.ln,
                                                          pc.m.p[k].lp);
                        out[2 * k] = lp.x;
                        out[2 * k + 1] = lp.y;
                    }
                });
                __syncthreads();
                if (lane < 3) comp(xbar[pc.bA], lane) += rm;
                else if (lane < 6) comp(xbar[pc.bB], lane - 3) += rm;
                __syncthreads();
            }
            if (lane == 0) atomicAdd(a.diag + 0, 1.0f);
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
        float4 v = vbar[b];
        v.x *= fl;
        v.y *= fl;
        v.z *= fa;
        vbar[b] = v;
    }
    // Bodies that do not move carry no adjoint (the reverse updates leave
    // scratch values there, which never reach a moving body: their inverse
    // masses are zero).
    for (int b = lane; b < nb; b += kWave) {
        if (d.bflags[bo + b] & prx::kAccelerable) continue;
        xbar[b] = float4{0.0f, 0.0f, 0.0f, 0.0f};
        vbar[b] = float4{0.0f, 0.0f, 0.0f, 0.0f};
    }
}

}  // namespace prxd

extern "C" hipError_t prx_launch_adjoint_general(const prx::DeviceState* d, const prx::Tape* t,
                                                 const prx::AdjState* a, const prx::StepParams* c, int nworlds,
                                                 hipStream_t s) {
    hipLaunchKernelGGL(prxd::k_adjoint_general, dim3(unsigned(nworlds)), dim3(64), 0, s, *d, *t, *a, *c);
    return hipGetLastError();
}

extern "C" hipError_t prx_launch_adjoint_step(const prx::DeviceState* d, const prx::Tape* t, const prx::AdjState* a,
                                              const prx::StepParams* c, int nworlds, hipStream_t s) {
    hipLaunchKernelGGL(prxd::k_adjoint_step, dim3(unsigned(nworlds)), dim3(64), 0, s, *d, *t, *a, *c);
    return hipGetLastError();
}
