// Golden-vector driver: builds a named scene through the REFERENCE PlayRho API
// (float build in oracle/_ref/libplayrho_ref.a), steps it, and dumps the
// reference's own state after every step as a raw little-endian record stream.
// This program is test infrastructure only: it is our code, linked against the
// reference library built from its sources by oracle/Makefile.ref.
//
//   ref_driver <scene> <steps> <doToi 0|1> <out.bin> [int params...]
//   ref_driver --time <scene> <warmup> <steps> <doToi> <threads> <nworlds> [params...]
//
// Record layout (all u32/f32):
//   "PRXD" 2
//   nshapes, per shape: nverts, radius, density, friction, restitution,
//                       sensor, (cat|mask<<16), group, verts[2n], normals[2n]
//   nbodies, per body:  type, flags, f32[19] (pos0 3, pos1 3, lc 2, vel 3,
//                       linAcc 2, angAcc, linDamp, angDamp, invMass, invRotI,
//                       underActive), nshape, shapeIds[]
//   njoints, per joint: bodyA, bodyB, flags (limit 1, motor 2, collideConnected 4,
//                       kind << 8), f32[12] (revolute: laA 2, laB 2,
//                       referenceAngle, lower, upper, motorSpeed,
//                       maxMotorTorque; other kinds see WriteDescription);
//                       gear joints (kind 10) then u32 bodyC, bodyD, f32
//                       ratio, constant
//   conf: f32/u32 StepConf fields used by the step (see WriteConf)
//   nrecords, per record: step, stats u32[8],
//                         B, per body f32[10] (pos0, pos1, vel, underActive), flags
//                         C, per contact u32[4] (kmin,kmax,id,flags), u32[3]
//                           (type|npts<<8, cf0, cf1), f32[14]
//                         P, per proxy u32 id, f32[4]
//                         J, per joint u32 limitState, f32[4] (impulse 3,
//                           angular motor impulse)
#include <algorithm>
#include <playrho/d2/BasicAPI.hpp>
#include <playrho/d2/World.hpp>
#include <playrho/d2/WorldBody.hpp>
#include <playrho/d2/WorldContact.hpp>
#include <playrho/d2/WorldMisc.hpp>
#include <playrho/d2/WorldShape.hpp>
#include <playrho/d2/DynamicTree.hpp>
#include <playrho/d2/PolygonShapeConf.hpp>
#include <playrho/d2/DiskShapeConf.hpp>
#include <playrho/d2/EdgeShapeConf.hpp>
#include <playrho/d2/ChainShapeConf.hpp>
#include <playrho/d2/MultiShapeConf.hpp>
#include <playrho/d2/VertexSet.hpp>
#include <playrho/d2/RevoluteJointConf.hpp>
#include <playrho/d2/DistanceJointConf.hpp>
#include <playrho/d2/WeldJointConf.hpp>
#include <playrho/d2/RopeJointConf.hpp>
#include <playrho/d2/FrictionJointConf.hpp>
#include <playrho/d2/MotorJointConf.hpp>
#include <playrho/d2/WheelJointConf.hpp>
#include <playrho/d2/PulleyJointConf.hpp>
#include <playrho/d2/PrismaticJointConf.hpp>
#include <playrho/d2/TargetJointConf.hpp>
#include <playrho/d2/GearJointConf.hpp>
#include <playrho/d2/WorldJoint.hpp>
#include <playrho/StepConf.hpp>

#include <chrono>
#include <cstdio>
#include <thread>
#include <cstdlib>
#include <cstring>
#include <sstream>
#include <string>
#include <vector>

#include "../playrho_amd/csrc/scenes.hpp"

using namespace playrho;
using namespace playrho::d2;

namespace {

struct Out {
    std::vector<unsigned char> buf;
    void u32(std::uint32_t v) { put(&v, 4); }
    void f32(float v) { put(&v, 4); }
    void put(const void* p, size_t n) {
        const auto* c = static_cast<const unsigned char*>(p);
        buf.insert(buf.end(), c, c + n);
    }
};

std::uint32_t PackCf(const ContactFeature& cf) {
    return std::uint32_t(cf.typeA) | (std::uint32_t(cf.typeB) << 1) |
           (std::uint32_t(cf.indexA) << 8) | (std::uint32_t(cf.indexB) << 16);
}

std::uint32_t BodyFlags(const Body& b) {
    std::uint32_t f = 0;
    if (IsSpeedable(b)) f |= 1u;
    if (IsAccelerable(b)) f |= 2u;
    if (IsImpenetrable(b)) f |= 4u;
    if (IsAwake(b)) f |= 8u;
    if (IsSleepingAllowed(b)) f |= 16u;
    if (IsEnabled(b)) f |= 32u;
    if (IsFixedRotation(b)) f |= 64u;
    return f;
}

ShapeID MakeShape(World& world, const prx_scene::ShapeSpec& s) {
    switch (s.kind) {
    case prx_scene::kDisk: {
        auto conf = DiskShapeConf{}.UseRadius(s.radius);
        conf.UseDensity(s.density).UseFriction(s.friction).UseRestitution(s.restitution);
        return CreateShape(world, conf);
    }
    case prx_scene::kBox: {
        auto conf = PolygonShapeConf{};
        if (s.vertexRadius >= 0) conf.UseVertexRadius(s.vertexRadius);
        conf.SetAsBox(s.hx, s.hy);
        conf.UseDensity(s.density).UseFriction(s.friction).UseRestitution(s.restitution);
        return CreateShape(world, conf);
    }
    case prx_scene::kPolygon: {
        auto conf = PolygonShapeConf{};
        if (s.vertexRadius >= 0) conf.UseVertexRadius(s.vertexRadius);
        std::vector<Length2> pts;
        for (const auto& p : s.pts) pts.push_back(Length2{p.x, p.y});
        conf.Set(Span<const Length2>(pts.data(), pts.size()));
        conf.UseDensity(s.density).UseFriction(s.friction).UseRestitution(s.restitution);
        return CreateShape(world, conf);
    }
    case prx_scene::kChain: {
        auto conf = ChainShapeConf{};
        if (s.vertexRadius >= 0) conf.UseVertexRadius(s.vertexRadius);
        std::vector<Length2> pts;
        for (const auto& p : s.pts) pts.push_back(Length2{p.x, p.y});
        conf.Set(pts);
        conf.UseDensity(s.density).UseFriction(s.friction).UseRestitution(s.restitution);
        return CreateShape(world, conf);
    }
    case prx_scene::kMulti: {
        auto conf = MultiShapeConf{};
        std::size_t at = 0;
        for (int n : s.hullSizes) {
            VertexSet set;
            for (int k = 0; k < n; ++k, ++at) set.add(Length2{s.pts[at].x, s.pts[at].y});
            if (s.vertexRadius >= 0) conf.AddConvexHull(set, s.vertexRadius);
            else conf.AddConvexHull(set);
        }
        conf.UseDensity(s.density).UseFriction(s.friction).UseRestitution(s.restitution);
        return CreateShape(world, conf);
    }
    default: {
        auto conf = EdgeShapeConf{Length2{s.pts[0].x, s.pts[0].y}, Length2{s.pts[1].x, s.pts[1].y}};
        if (s.vertexRadius >= 0) conf.UseVertexRadius(s.vertexRadius);
        conf.UseDensity(s.density).UseFriction(s.friction).UseRestitution(s.restitution);
        return CreateShape(world, conf);
    }
    }
}

// A shape with several children (a chain's segments, a multi shape's hulls)
// is written as one description shape per child, in child order, with the
// shape's material; bodies list the children.  The shape index of child 0 of
// each shape is first[shape].
std::vector<std::uint32_t> ChildBase(const World& world) {
    std::vector<std::uint32_t> first;
    std::uint32_t at = 0;
    for (auto i = 0u; i < GetShapeRange(world); ++i) {
        first.push_back(at);
        at += std::uint32_t(GetChildCount(GetShape(world, ShapeID(i))));
    }
    first.push_back(at);
    return first;
}

void WriteDescription(Out& o, const World& world) {
    const auto first = ChildBase(world);
    const auto nshapes = GetShapeRange(world);
    o.u32(first.back());
    for (auto i = 0u; i < nshapes; ++i)
    for (auto c = 0u; c < first[i + 1] - first[i]; ++c) {
        const auto shape = GetShape(world, ShapeID(i));
        const auto child = GetChild(shape, ChildCounter(c));
        const auto n = child.GetVertexCount();
        o.u32(n);
        o.f32(child.GetVertexRadius());
        o.f32(GetDensity(shape));
        o.f32(GetFriction(shape));
        o.f32(GetRestitution(shape));
        o.u32(IsSensor(shape) ? 1u : 0u);
        const auto filt = GetFilter(shape);
        o.u32(std::uint32_t(filt.categoryBits) | (std::uint32_t(filt.maskBits) << 16));
        o.u32(std::uint32_t(std::int32_t(filt.groupIndex)));
        for (auto v = 0u; v < n; ++v) {
            o.f32(GetX(child.GetVertex(v)));
            o.f32(GetY(child.GetVertex(v)));
        }
        for (auto v = 0u; v < n; ++v) {
            // A one-vertex proxy (disk) carries no normals.
            o.f32(n > 1 ? float(GetX(child.GetNormal(v))) : 0.0f);
            o.f32(n > 1 ? float(GetY(child.GetNormal(v))) : 0.0f);
        }
    }
    const auto nbodies = GetBodyRange(world);
    o.u32(nbodies);
    for (auto i = 0u; i < nbodies; ++i) {
        const auto b = GetBody(world, BodyID(i));
        o.u32(std::uint32_t(GetType(b)));
        o.u32(BodyFlags(b));
        const auto sw = GetSweep(b);
        const float vals[19] = {
            GetX(sw.pos0.linear), GetY(sw.pos0.linear), sw.pos0.angular,
            GetX(sw.pos1.linear), GetY(sw.pos1.linear), sw.pos1.angular,
            GetX(sw.localCenter), GetY(sw.localCenter),
            GetX(GetVelocity(b).linear), GetY(GetVelocity(b).linear), GetVelocity(b).angular,
            GetX(GetLinearAcceleration(b)), GetY(GetLinearAcceleration(b)),
            GetAngularAcceleration(b), GetLinearDamping(b), GetAngularDamping(b),
            GetInvMass(b), GetInvRotInertia(b), GetUnderActiveTime(b)};
        o.put(vals, sizeof(vals));
        const auto shapes = GetShapes(world, BodyID(i));
        std::uint32_t nchildren = 0;
        for (auto s : shapes) nchildren += first[to_underlying(s) + 1] - first[to_underlying(s)];
        o.u32(nchildren);
        for (auto s : shapes)
            for (auto c = first[to_underlying(s)]; c < first[to_underlying(s) + 1]; ++c) o.u32(c);
    }
    const auto njoints = GetJointRange(world);
    o.u32(njoints);
    // Joint records: flags bits 0-2 (limit, motor, collideConnected), bits 8-15
    // the kind (0 revolute, 1 distance, 2 weld), then 9 floats.
    for (auto i = 0u; i < njoints; ++i) {
        const auto joint = GetJoint(world, JointID(i));
        if (const auto* j = TypeCast<const RevoluteJointConf>(&joint)) {
            o.u32(to_underlying(GetBodyA(*j)));
            o.u32(to_underlying(GetBodyB(*j)));
            o.u32((j->enableLimit ? 1u : 0u) | (j->enableMotor ? 2u : 0u) | (j->collideConnected ? 4u : 0u));
            const float f[12] = {GetX(j->localAnchorA), GetY(j->localAnchorA), GetX(j->localAnchorB),
                                GetY(j->localAnchorB), j->referenceAngle, j->lowerAngle, j->upperAngle,
                                j->motorSpeed, j->maxMotorTorque};
            o.put(f, sizeof(f));
        } else if (const auto* j = TypeCast<const DistanceJointConf>(&joint)) {
            o.u32(to_underlying(GetBodyA(*j)));
            o.u32(to_underlying(GetBodyB(*j)));
            o.u32((j->collideConnected ? 4u : 0u) | (1u << 8));
            const float f[12] = {GetX(j->localAnchorA), GetY(j->localAnchorA), GetX(j->localAnchorB),
                                GetY(j->localAnchorB), j->length, j->frequency, j->dampingRatio, 0.0f, 0.0f};
            o.put(f, sizeof(f));
        } else if (const auto* j = TypeCast<const WeldJointConf>(&joint)) {
            o.u32(to_underlying(GetBodyA(*j)));
            o.u32(to_underlying(GetBodyB(*j)));
            o.u32((j->collideConnected ? 4u : 0u) | (2u << 8));
            const float f[12] = {GetX(j->localAnchorA), GetY(j->localAnchorA), GetX(j->localAnchorB),
                                GetY(j->localAnchorB), j->referenceAngle, j->frequency, j->dampingRatio,
                                0.0f, 0.0f};
            o.put(f, sizeof(f));
        } else if (const auto* j = TypeCast<const RopeJointConf>(&joint)) {
            o.u32(to_underlying(GetBodyA(*j)));
            o.u32(to_underlying(GetBodyB(*j)));
            o.u32((j->collideConnected ? 4u : 0u) | (3u << 8));
            const float f[12] = {GetX(j->localAnchorA), GetY(j->localAnchorA), GetX(j->localAnchorB),
                                GetY(j->localAnchorB), j->maxLength, 0.0f, 0.0f, 0.0f, 0.0f};
            o.put(f, sizeof(f));
        } else if (const auto* j = TypeCast<const FrictionJointConf>(&joint)) {
            o.u32(to_underlying(GetBodyA(*j)));
            o.u32(to_underlying(GetBodyB(*j)));
            o.u32((j->collideConnected ? 4u : 0u) | (4u << 8));
            const float f[12] = {GetX(j->localAnchorA), GetY(j->localAnchorA), GetX(j->localAnchorB),
                                GetY(j->localAnchorB), j->maxForce, j->maxTorque, 0.0f, 0.0f, 0.0f};
            o.put(f, sizeof(f));
        } else if (const auto* j = TypeCast<const PulleyJointConf>(&joint)) {
            o.u32(to_underlying(GetBodyA(*j)));
            o.u32(to_underlying(GetBodyB(*j)));
            o.u32((j->collideConnected ? 4u : 0u) | (7u << 8));
            const float f[12] = {GetX(j->localAnchorA),  GetY(j->localAnchorA),  GetX(j->localAnchorB),
                                 GetY(j->localAnchorB),  GetX(j->groundAnchorA), GetY(j->groundAnchorA),
                                 GetX(j->groundAnchorB), GetY(j->groundAnchorB), j->ratio,
                                 j->lengthA,             j->lengthB,             0.0f};
            o.put(f, sizeof(f));
        } else if (const auto* j = TypeCast<const TargetJointConf>(&joint)) {
            o.u32(to_underlying(GetBodyA(*j)));
            o.u32(to_underlying(GetBodyB(*j)));
            o.u32((j->collideConnected ? 4u : 0u) | (9u << 8));
            const float f[12] = {0.0f,          0.0f,           GetX(j->localAnchorB), GetY(j->localAnchorB),
                                 GetX(j->target), GetY(j->target), j->maxForce,          j->frequency,
                                 j->dampingRatio, 0.0f,          0.0f,                  0.0f};
            o.put(f, sizeof(f));
        } else if (const auto* j = TypeCast<const PrismaticJointConf>(&joint)) {
            o.u32(to_underlying(GetBodyA(*j)));
            o.u32(to_underlying(GetBodyB(*j)));
            o.u32((j->enableLimit ? 1u : 0u) | (j->enableMotor ? 2u : 0u) | (j->collideConnected ? 4u : 0u) |
                  (8u << 8));
            const float f[12] = {GetX(j->localAnchorA), GetY(j->localAnchorA),  GetX(j->localAnchorB),
                                 GetY(j->localAnchorB), j->referenceAngle,       j->lowerTranslation,
                                 j->upperTranslation,   j->motorSpeed,           j->maxMotorForce,
                                 GetX(j->localXAxisA),  GetY(j->localXAxisA),    0.0f};
            o.put(f, sizeof(f));
        } else if (const auto* j = TypeCast<const WheelJointConf>(&joint)) {
            o.u32(to_underlying(GetBodyA(*j)));
            o.u32(to_underlying(GetBodyB(*j)));
            o.u32((j->enableMotor ? 2u : 0u) | (j->collideConnected ? 4u : 0u) | (6u << 8));
            const float f[12] = {GetX(j->localAnchorA), GetY(j->localAnchorA), GetX(j->localAnchorB),
                                 GetY(j->localAnchorB), j->frequency, j->dampingRatio, 0.0f, j->motorSpeed,
                                 j->maxMotorTorque, GetX(j->localXAxisA), GetY(j->localXAxisA), 0.0f};
            o.put(f, sizeof(f));
        } else if (const auto* j = TypeCast<const GearJointConf>(&joint)) {
            // Gear: per side (A-C, then B-D) the prismatic data or the
            // revolute reference angle, bit 16 + side set for prismatic;
            // then bodies C, D, ratio, constant.
            float f[12] = {};
            std::uint32_t pris = 0;
            const GearJointConf::TypeData* sides[2] = {&j->typeDataAC, &j->typeDataBD};
            for (int k = 0; k < 2; ++k) {
                if (const auto* pd = std::get_if<GearJointConf::PrismaticData>(sides[k])) {
                    pris |= 1u << (16 + k);
                    f[6 * k + 0] = GetX(pd->localAnchorA);
                    f[6 * k + 1] = GetY(pd->localAnchorA);
                    f[6 * k + 2] = GetX(pd->localAnchorB);
                    f[6 * k + 3] = GetY(pd->localAnchorB);
                    f[6 * k + 4] = GetX(pd->localAxis);
                    f[6 * k + 5] = GetY(pd->localAxis);
                } else if (const auto* rd = std::get_if<GearJointConf::RevoluteData>(sides[k])) {
                    f[6 * k] = rd->referenceAngle;
                }
            }
            o.u32(to_underlying(GetBodyA(*j)));
            o.u32(to_underlying(GetBodyB(*j)));
            o.u32((j->collideConnected ? 4u : 0u) | (10u << 8) | pris);
            o.put(f, sizeof(f));
            o.u32(to_underlying(j->bodyC));
            o.u32(to_underlying(j->bodyD));
            const float g[2] = {j->ratio, j->constant};
            o.put(g, sizeof(g));
        } else if (const auto* j = TypeCast<const MotorJointConf>(&joint)) {
            o.u32(to_underlying(GetBodyA(*j)));
            o.u32(to_underlying(GetBodyB(*j)));
            o.u32((j->collideConnected ? 4u : 0u) | (5u << 8));
            const float f[12] = {GetX(j->linearOffset), GetY(j->linearOffset), 0.0f, 0.0f, j->angularOffset,
                                j->maxForce, j->maxTorque, j->correctionFactor, 0.0f};
            o.put(f, sizeof(f));
        }
    }
}

void WriteConf(Out& o, const StepConf& c) {
    const float f[] = {c.deltaTime, c.dtRatio, c.minStillTimeToSleep, c.linearSlop,
                       c.angularSlop, c.regResolutionRate, c.regMinSeparation,
                       c.regMinMomentum, c.velocityThreshold, c.maxTranslation,
                       c.maxRotation, c.maxLinearCorrection, c.maxAngularCorrection,
                       c.linearSleepTolerance, c.angularSleepTolerance,
                       c.displaceMultiplier, c.aabbExtension, c.maxCirclesRatio};
    o.u32(sizeof(f) / 4);
    o.put(f, sizeof(f));
    o.u32(c.regVelocityIters);
    o.u32(c.regPositionIters);
    o.u32((c.doWarmStart ? 1u : 0u) | (c.doToi ? 2u : 0u) | (c.doBlocksolve ? 4u : 0u));
    // TOI phase fields.
    const float t[] = {c.toiResolutionRate, c.toiMinSeparation, c.toiMinMomentum, c.targetDepth, c.tolerance};
    o.u32(sizeof(t) / 4);
    o.put(t, sizeof(t));
    const std::uint32_t ti[] = {c.toiVelocityIters, c.toiPositionIters, c.maxToiRootIters,
                                c.maxToiIters,      c.maxDistanceIters, c.maxSubSteps};
    o.put(ti, sizeof(ti));
}

// Contact listener calls of the last step, serialised in call order:
// u32 kind (0 begin, 1 end, 2 pre-solve, 3 post-solve), u32 contact id, then
// for pre-solve the old manifold (u32 type | count << 8, two packed contact
// features, 12 floats: local normal, local point, per point local point and
// normal / tangent impulse) and for post-solve u32 point count, u32 solved
// position iteration, normal impulses [2], tangent impulses [2].
struct EventLog {
    Out o;
    std::uint32_t n = 0;
    void head(std::uint32_t kind, ContactID id) {
        o.u32(kind);
        o.u32(to_underlying(id));
        ++n;
    }
};
EventLog g_events;

void PutManifold(Out& o, const Manifold& m) {
    const auto npts = m.GetPointCount();
    const auto type = m.GetType();
    std::uint32_t cf[2] = {PackCf(ContactFeature{}), PackCf(ContactFeature{})};
    float f[12] = {};
    if (type == Manifold::e_faceA || type == Manifold::e_faceB) {
        f[0] = GetX(m.GetLocalNormal());
        f[1] = GetY(m.GetLocalNormal());
    }
    if (type != Manifold::e_unset) {
        f[2] = GetX(m.GetLocalPoint());
        f[3] = GetY(m.GetLocalPoint());
    }
    for (auto p = 0u; p < npts; ++p) {
        const auto mp = m.GetPoint(p);
        cf[p] = PackCf(mp.contactFeature);
        f[4 + 4 * p + 0] = GetX(mp.localPoint);
        f[4 + 4 * p + 1] = GetY(mp.localPoint);
        f[4 + 4 * p + 2] = mp.normalImpulse;
        f[4 + 4 * p + 3] = mp.tangentImpulse;
    }
    const std::uint32_t t[3] = {std::uint32_t(type) | (std::uint32_t(npts) << 8), cf[0], cf[1]};
    o.put(t, sizeof(t));
    o.put(f, sizeof(f));
}

void InstallListeners(World& world) {
    SetBeginContactListener(world, [](ContactID id) { g_events.head(0, id); });
    SetEndContactListener(world, [](ContactID id) { g_events.head(1, id); });
    SetPreSolveContactListener(world, [](ContactID id, const Manifold& m) {
        g_events.head(2, id);
        PutManifold(g_events.o, m);
    });
    SetPostSolveContactListener(world, [](ContactID id, const ContactImpulsesList& l, unsigned iters) {
        g_events.head(3, id);
        g_events.o.u32(std::uint32_t(l.GetCount()));
        g_events.o.u32(std::uint32_t(iters));
        float f[4] = {};
        for (auto i = 0u; i < l.GetCount(); ++i) {
            f[i] = l.GetEntryNormal(i);
            f[2 + i] = l.GetEntryTanget(i);
        }
        g_events.o.put(f, sizeof(f));
    });
}

void WriteState(Out& o, const World& world, std::uint32_t step, const StepStats& st) {
    o.u32(step);
    // PreStepStats then RegStepStats, each in declaration order.
    auto bits = [](float f) {
        std::uint32_t u;
        std::memcpy(&u, &f, 4);
        return u;
    };
    const std::uint32_t stats[18] = {
        st.pre.proxiesCreated, st.pre.proxiesMoved, st.pre.contactsDestroyed, st.pre.contactsAdded,
        st.pre.contactsUpdated, st.pre.contactsSkipped, bits(float(st.reg.minSeparation)),
        bits(float(st.reg.maxIncImpulse)), std::uint32_t(st.reg.islandsFound),
        std::uint32_t(st.reg.islandsSolved), std::uint32_t(st.reg.bodiesSlept),
        std::uint32_t(st.reg.maxIslandBodies), st.reg.contactsAdded, st.reg.contactsUpdated,
        st.reg.contactsSkipped, st.reg.proxiesMoved, st.reg.sumPosIters, st.reg.sumVelIters};
    o.put(stats, sizeof(stats));
    const std::uint32_t toi[16] = {
        bits(float(st.toi.minSeparation)), bits(float(st.toi.maxIncImpulse)), st.toi.islandsFound,
        st.toi.islandsSolved, st.toi.contactsFound, st.toi.contactsAtMaxSubSteps, st.toi.contactsUpdatedToi,
        st.toi.contactsUpdatedTouching, st.toi.contactsSkippedTouching, st.toi.contactsAdded,
        st.toi.proxiesMoved, st.toi.sumPosIters, st.toi.sumVelIters, std::uint32_t(st.toi.maxDistIters),
        std::uint32_t(st.toi.maxToiIters), std::uint32_t(st.toi.maxRootIters)};
    o.put(toi, sizeof(toi));
    const auto nbodies = GetBodyRange(world);
    o.u32(nbodies);
    const auto alive = GetBodies(world);
    for (auto i = 0u; i < nbodies; ++i) {
        if (std::find(alive.begin(), alive.end(), BodyID(i)) == alive.end()) {
            // A destroyed body: zeros and flags 0x80.
            const float v[10] = {};
            o.put(v, sizeof(v));
            o.u32(0x80u);
            continue;
        }
        const auto b = GetBody(world, BodyID(i));
        const auto sw = GetSweep(b);
        const float v[10] = {GetX(sw.pos0.linear), GetY(sw.pos0.linear), sw.pos0.angular,
                             GetX(sw.pos1.linear), GetY(sw.pos1.linear), sw.pos1.angular,
                             GetX(GetVelocity(b).linear), GetY(GetVelocity(b).linear),
                             GetVelocity(b).angular, GetUnderActiveTime(b)};
        o.put(v, sizeof(v));
        o.u32(BodyFlags(b));
    }
    const auto contacts = GetContacts(world);
    o.u32(std::uint32_t(contacts.size()));
    for (const auto& kc : contacts) {
        const auto key = std::get<ContactKey>(kc);
        const auto id = std::get<ContactID>(kc);
        const auto c = GetContact(world, id);
        std::uint32_t flags = 0;
        if (IsTouching(c)) flags |= 1u;
        if (IsEnabled(c)) flags |= 2u;
        if (IsSensor(c)) flags |= 4u;
        if (IsImpenetrable(c)) flags |= 8u;
        if (NeedsUpdating(c)) flags |= 16u;
        if (NeedsFiltering(c)) flags |= 32u;
        const std::uint32_t u[4] = {key.GetMin(), key.GetMax(), to_underlying(id), flags};
        o.put(u, sizeof(u));
        const auto m = GetManifold(world, id);
        const auto npts = m.GetPointCount();
        const auto type = m.GetType();
        std::uint32_t cf[2] = {PackCf(ContactFeature{}), PackCf(ContactFeature{})};
        float f[14] = {GetFriction(c), GetRestitution(c)};
        if (type == Manifold::e_faceA || type == Manifold::e_faceB) {
            f[2] = GetX(m.GetLocalNormal());
            f[3] = GetY(m.GetLocalNormal());
        }
        if (type != Manifold::e_unset) {
            f[4] = GetX(m.GetLocalPoint());
            f[5] = GetY(m.GetLocalPoint());
        }
        for (auto p = 0u; p < npts; ++p) {
            const auto mp = m.GetPoint(p);
            cf[p] = PackCf(mp.contactFeature);
            f[6 + 4 * p + 0] = GetX(mp.localPoint);
            f[6 + 4 * p + 1] = GetY(mp.localPoint);
            f[6 + 4 * p + 2] = mp.normalImpulse;
            f[6 + 4 * p + 3] = mp.tangentImpulse;
        }
        const std::uint32_t t[3] = {std::uint32_t(type) | (std::uint32_t(npts) << 8), cf[0], cf[1]};
        o.put(t, sizeof(t));
        o.put(f, sizeof(f));
    }
    const auto& tree = GetTree(world);
    const auto cap = tree.GetNodeCapacity();
    std::vector<std::uint32_t> leaves;
    for (auto i = 0u; i < cap; ++i) {
        if (DynamicTree::IsLeaf(tree.GetHeight(i))) leaves.push_back(i);
    }
    o.u32(std::uint32_t(leaves.size()));
    for (auto id : leaves) {
        o.u32(id);
        const auto aabb = tree.GetAABB(id);
        const float a[4] = {aabb.ranges[0].GetMin(), aabb.ranges[1].GetMin(),
                            aabb.ranges[0].GetMax(), aabb.ranges[1].GetMax()};
        o.put(a, sizeof(a));
    }
    const auto njoints = GetJointRange(world);
    o.u32(njoints);
    const auto aliveJoints = GetJoints(world);
    for (auto i = 0u; i < njoints; ++i) {
        if (std::find(aliveJoints.begin(), aliveJoints.end(), JointID(i)) == aliveJoints.end()) {
            // A destroyed joint: limit state 0xffffffff and zeros.
            o.u32(0xffffffffu);
            const float f[4] = {};
            o.put(f, sizeof(f));
            continue;
        }
        const auto joint = GetJoint(world, JointID(i));
        if (const auto* j = TypeCast<const RevoluteJointConf>(&joint)) {
            o.u32(std::uint32_t(j->limitState));
            const float f[4] = {GetX(j->impulse), GetY(j->impulse), GetZ(j->impulse), j->angularMotorImpulse};
            o.put(f, sizeof(f));
        } else if (const auto* j = TypeCast<const DistanceJointConf>(&joint)) {
            o.u32(0u);
            const float f[4] = {j->impulse, 0.0f, 0.0f, 0.0f};
            o.put(f, sizeof(f));
        } else if (const auto* j = TypeCast<const WeldJointConf>(&joint)) {
            o.u32(0u);
            const float f[4] = {GetX(j->impulse), GetY(j->impulse), GetZ(j->impulse), 0.0f};
            o.put(f, sizeof(f));
        } else if (const auto* j = TypeCast<const RopeJointConf>(&joint)) {
            o.u32(std::uint32_t(j->limitState));
            const float f[4] = {j->impulse, 0.0f, 0.0f, 0.0f};
            o.put(f, sizeof(f));
        } else if (const auto* j = TypeCast<const FrictionJointConf>(&joint)) {
            o.u32(0u);
            const float f[4] = {GetX(j->linearImpulse), GetY(j->linearImpulse), j->angularImpulse, 0.0f};
            o.put(f, sizeof(f));
        } else if (const auto* j = TypeCast<const PulleyJointConf>(&joint)) {
            o.u32(0u);
            const float f[4] = {j->impulse, 0.0f, 0.0f, 0.0f};
            o.put(f, sizeof(f));
        } else if (const auto* j = TypeCast<const TargetJointConf>(&joint)) {
            o.u32(0u);
            const float f[4] = {GetX(j->impulse), GetY(j->impulse), 0.0f, 0.0f};
            o.put(f, sizeof(f));
        } else if (const auto* j = TypeCast<const PrismaticJointConf>(&joint)) {
            o.u32(std::uint32_t(j->limitState));
            const float f[4] = {GetX(j->impulse), GetY(j->impulse), GetZ(j->impulse), j->motorImpulse};
            o.put(f, sizeof(f));
        } else if (const auto* j = TypeCast<const WheelJointConf>(&joint)) {
            o.u32(0u);
            const float f[4] = {j->impulse, j->springImpulse, j->angularImpulse, 0.0f};
            o.put(f, sizeof(f));
        } else if (const auto* j = TypeCast<const GearJointConf>(&joint)) {
            o.u32(0u);
            const float f[4] = {j->impulse, 0.0f, 0.0f, 0.0f};
            o.put(f, sizeof(f));
        } else if (const auto* j = TypeCast<const MotorJointConf>(&joint)) {
            o.u32(0u);
            const float f[4] = {GetX(j->linearImpulse), GetY(j->linearImpulse), j->angularImpulse, 0.0f};
            o.put(f, sizeof(f));
        }
    }
    o.u32(g_events.n);
    o.put(g_events.o.buf.data(), g_events.o.buf.size());
    g_events.o.buf.clear();
    g_events.n = 0;
}

}  // namespace

World BuildWorld(const prx_scene::SceneSpec& spec) {
    auto world = World{};
    // Shapes are created in spec order, each immediately before the first body
    // that uses it, matching the order the product's scene builder uses.
    std::vector<ShapeID> shapeIds(spec.shapes.size(), InvalidShapeID);
    for (const auto& bs : spec.bodies) {
        auto conf = BodyConf{};
        conf.Use(bs.type == prx_scene::kDynamic ? BodyType::Dynamic
                 : bs.type == prx_scene::kKinematic ? BodyType::Kinematic
                                                    : BodyType::Static);
        conf.UseLocation(Length2{bs.x, bs.y});
        conf.UseAngle(bs.angle);
        conf.UseLinearVelocity(LinearVelocity2{bs.vx, bs.vy});
        conf.UseAngularVelocity(bs.w);
        conf.UseLinearAcceleration(LinearAcceleration2{bs.ax, bs.ay});
        for (int si : bs.shapes) {
            if (shapeIds[size_t(si)] == InvalidShapeID) {
                shapeIds[size_t(si)] = MakeShape(world, spec.shapes[size_t(si)]);
            }
            conf.Use(shapeIds[size_t(si)]);
        }
        CreateBody(world, conf);
    }
    for (const auto& js : spec.joints) {
        const auto ba = BodyID(static_cast<BodyID::underlying_type>(js.a));
        const auto bb = BodyID(static_cast<BodyID::underlying_type>(js.b));
        if (js.kind == prx_scene::kGear) {
            // a, b: the revolute or prismatic joints the gear couples.
            auto jc = GetGearJointConf(world, JointID(static_cast<JointID::underlying_type>(js.a)),
                                       JointID(static_cast<JointID::underlying_type>(js.b)), js.ratio);
            jc.UseCollideConnected(js.collideConnected);
            CreateJoint(world, jc);
        } else if (js.kind == prx_scene::kDistance) {
            auto jc = DistanceJointConf{ba, bb, Length2{js.lax, js.lay}, Length2{js.lbx, js.lby}, js.length};
            jc.UseFrequency(js.frequency).UseDampingRatio(js.dampingRatio);
            jc.UseCollideConnected(js.collideConnected);
            CreateJoint(world, jc);
        } else if (js.kind == prx_scene::kRope) {
            auto jc = RopeJointConf{ba, bb};
            jc.localAnchorA = Length2{js.lax, js.lay};
            jc.localAnchorB = Length2{js.lbx, js.lby};
            jc.UseMaxLength(js.length);
            jc.UseCollideConnected(js.collideConnected);
            CreateJoint(world, jc);
        } else if (js.kind == prx_scene::kFriction) {
            auto jc = FrictionJointConf{ba, bb, Length2{js.lax, js.lay}, Length2{js.lbx, js.lby}};
            jc.UseMaxForce(js.maxForce).UseMaxTorque(js.maxTorque);
            jc.UseCollideConnected(js.collideConnected);
            CreateJoint(world, jc);
        } else if (js.kind == prx_scene::kPulley) {
            auto jc = PulleyJointConf{ba, bb, Length2{js.gax, js.gay}, Length2{js.gbx, js.gby},
                                      Length2{js.lax, js.lay}, Length2{js.lbx, js.lby}, js.length, js.lengthB};
            jc.UseRatio(js.ratio);
            jc.UseCollideConnected(js.collideConnected);
            CreateJoint(world, jc);
        } else if (js.kind == prx_scene::kTarget) {
            auto jc = TargetJointConf{bb};
            jc.UseTarget(Length2{js.gax, js.gay}).UseAnchor(Length2{js.lbx, js.lby});
            jc.UseMaxForce(js.maxForce).UseFrequency(js.frequency).UseDampingRatio(js.dampingRatio);
            jc.UseCollideConnected(js.collideConnected);
            CreateJoint(world, jc);
        } else if (js.kind == prx_scene::kPrismatic) {
            auto jc = PrismaticJointConf{ba, bb, Length2{js.lax, js.lay}, Length2{js.lbx, js.lby},
                                         UnitVec::Get(js.axisAngle), js.referenceAngle};
            jc.UseEnableLimit(js.enableLimit).UseLowerLength(js.lower).UseUpperLength(js.upper);
            jc.UseEnableMotor(js.enableMotor).UseMotorSpeed(js.motorSpeed).UseMaxMotorForce(js.maxForce);
            jc.UseCollideConnected(js.collideConnected);
            CreateJoint(world, jc);
        } else if (js.kind == prx_scene::kWheel) {
            auto jc = WheelJointConf{ba, bb, Length2{js.lax, js.lay}, Length2{js.lbx, js.lby},
                                     UnitVec::Get(js.axisAngle)};
            jc.UseEnableMotor(js.enableMotor).UseMotorSpeed(js.motorSpeed).UseMaxMotorTorque(js.maxMotorTorque);
            jc.UseFrequency(js.frequency).UseDampingRatio(js.dampingRatio);
            jc.UseCollideConnected(js.collideConnected);
            CreateJoint(world, jc);
        } else if (js.kind == prx_scene::kMotor) {
            auto jc = MotorJointConf{ba, bb, Length2{js.lax, js.lay}, js.referenceAngle};
            jc.UseMaxForce(js.maxForce).UseMaxTorque(js.maxTorque).UseCorrectionFactor(js.correctionFactor);
            jc.UseCollideConnected(js.collideConnected);
            CreateJoint(world, jc);
        } else if (js.kind == prx_scene::kWeld) {
            auto jc = WeldJointConf{ba, bb, Length2{js.lax, js.lay}, Length2{js.lbx, js.lby}, js.referenceAngle};
            jc.UseFrequency(js.frequency).UseDampingRatio(js.dampingRatio);
            jc.UseCollideConnected(js.collideConnected);
            CreateJoint(world, jc);
        } else {
            auto jc = RevoluteJointConf{ba, bb, Length2{js.lax, js.lay}, Length2{js.lbx, js.lby}, js.referenceAngle};
            jc.UseEnableLimit(js.enableLimit).UseLowerAngle(js.lower).UseUpperAngle(js.upper);
            jc.UseEnableMotor(js.enableMotor).UseMotorSpeed(js.motorSpeed).UseMaxMotorTorque(js.maxMotorTorque);
            jc.UseCollideConnected(js.collideConnected);
            CreateJoint(world, jc);
        }
    }
    return world;
}

StepConf RegularStepConf(bool doToi) {
    auto conf = StepConf{};
    conf.deltaTime = 1.0f / 60;
    conf.regVelocityIters = 8;
    conf.regPositionIters = 3;
    conf.doToi = doToi;
    return conf;
}

// --time <scene> <warmup> <steps> <doToi> <threads> <nworlds> [params...]:
// builds nworlds copies of a scene (for "pile", "chain" and "billiards",
// param 1 is the first world index and world i uses index first + i), steps each `warmup` times untimed,
// then `steps` times timed, on `threads` std::threads (static partition
// balanced by body count; PlayRho steps one world on one thread), and prints
// the timing as one JSON line.
int TimeWorlds(int argc, char** argv) {
    if (argc < 8) return 2;
    const std::string name = argv[2];
    const int warmup = std::atoi(argv[3]);
    const int steps = std::atoi(argv[4]);
    const bool doToi = std::atoi(argv[5]) != 0;
    int threads = std::atoi(argv[6]);
    const int nworlds = std::atoi(argv[7]);
    std::vector<int> params;
    for (int i = 8; i < argc; ++i) params.push_back(std::atoi(argv[i]));
    std::vector<World> worlds;
    std::vector<double> weight;
    double dyn = 0;
    for (int k = 0; k < nworlds; ++k) {
        std::vector<int> p = params;
        if (name == "pile" || name == "chain" || name == "billiards") {
            if (p.size() < 2) p.resize(2, 0);
            p[1] += k;
        }
        prx_scene::SceneSpec spec;
        if (!prx_scene::Make(name, p.data(), int(p.size()), spec)) return 2;
        double nb = 0;
        for (const auto& b : spec.bodies) nb += (b.type == prx_scene::kDynamic) ? 1.0 : 0.0;
        dyn += nb;
        weight.push_back(nb);
        worlds.push_back(BuildWorld(spec));
    }
    if (threads < 1) threads = 1;
    std::vector<size_t> cut(size_t(threads) + 1, worlds.size());
    cut[0] = 0;
    double total = 0;
    for (double x : weight) total += x;
    double acc = 0;
    int t = 1;
    for (size_t i = 0; i < worlds.size() && t < threads; ++i) {
        acc += weight[i];
        while (t < threads && acc >= total * t / threads) cut[size_t(t++)] = i + 1;
    }
    const auto conf = RegularStepConf(doToi);
    auto run = [&](int nsteps) {
        std::vector<std::thread> pool;
        for (int k = 0; k < threads; ++k) {
            pool.emplace_back([&, k] {
                for (size_t i = cut[size_t(k)]; i < cut[size_t(k) + 1]; ++i)
                    for (int s = 0; s < nsteps; ++s) Step(worlds[i], conf);
            });
        }
        for (auto& th : pool) th.join();
    };
    run(warmup);
    const auto t0 = std::chrono::steady_clock::now();
    run(steps);
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::printf("{\"seconds\": %.6f, \"body_steps\": %.0f, \"threads\": %d, \"worlds\": %d, \"steps\": %d}\n",
                secs, dyn * steps, threads, nworlds, steps);
    return 0;
}

// World edits between steps, from the environment variable PRX_EDITS:
// "step kind body a b c; ...", applied before step `step`:
//   L  SetLocation(world, body, (a, b))
//   A  SetAngle(world, body, a)
//   T  SetTransformation(world, body, {(a, b), UnitVec::Get(c)})
//   O  ShiftOrigin(world, (a, b))
//   V  SetVelocity(world, body, {(a, b), c})
//   J  Destroy(world, JointID(body))
//   D  Destroy(world, BodyID(body))
//   S  SetType(world, body, BodyType(a)) (0 static, 1 kinematic, 2 dynamic)
struct Edit {
    int step = 0;
    char kind = 0;
    unsigned body = 0;
    float a = 0, b = 0, c = 0;
};

std::vector<Edit> ReadEdits() {
    std::vector<Edit> out;
    const char* env = std::getenv("PRX_EDITS");
    if (!env) return out;
    std::stringstream all(env);
    std::string item;
    while (std::getline(all, item, ';')) {
        std::stringstream in(item);
        Edit e;
        if (in >> e.step >> e.kind >> e.body >> e.a >> e.b >> e.c) out.push_back(e);
    }
    return out;
}

void ApplyEdit(World& world, const Edit& e) {
    const auto id = BodyID(static_cast<BodyID::underlying_type>(e.body));
    switch (e.kind) {
    case 'L':
        SetLocation(world, id, Length2{e.a * Meter, e.b * Meter});
        break;
    case 'A':
        SetAngle(world, id, e.a * Radian);
        break;
    case 'T':
        SetTransformation(world, id, Transformation{Length2{e.a * Meter, e.b * Meter}, UnitVec::Get(e.c * Radian)});
        break;
    case 'O':
        ShiftOrigin(world, Length2{e.a * Meter, e.b * Meter});
        break;
    case 'V':
        SetVelocity(world, id, Velocity{LinearVelocity2{e.a * MeterPerSecond, e.b * MeterPerSecond},
                                        e.c * RadianPerSecond});
        break;
    case 'J':
        Destroy(world, JointID(static_cast<JointID::underlying_type>(e.body)));
        break;
    case 'D':
        Destroy(world, id);
        break;
    case 'S':
        SetType(world, id, static_cast<BodyType>(int(e.a)));
        break;
    default:
        break;
    }
}

int main(int argc, char** argv) {
    if (argc > 1 && std::string(argv[1]) == "--time") return TimeWorlds(argc, argv);
    if (argc < 5) {
        std::fprintf(stderr, "usage: %s scene steps doToi out.bin [params...]\n", argv[0]);
        return 2;
    }
    const std::string name = argv[1];
    const int steps = std::atoi(argv[2]);
    const bool doToi = std::atoi(argv[3]) != 0;
    std::vector<int> params;
    for (int i = 5; i < argc; ++i) params.push_back(std::atoi(argv[i]));
    prx_scene::SceneSpec spec;
    if (!prx_scene::Make(name, params.data(), int(params.size()), spec)) {
        std::fprintf(stderr, "unknown scene %s\n", name.c_str());
        return 2;
    }

    auto world = BuildWorld(spec);
    InstallListeners(world);

    const auto conf = RegularStepConf(doToi);

    // PRX_DUMP_STEPS="a,b,...": write only the initial state and those steps
    // (large worlds); otherwise every step.
    std::vector<int> only;
    if (const char* env = std::getenv("PRX_DUMP_STEPS")) {
        std::stringstream in(env);
        std::string item;
        while (std::getline(in, item, ',')) only.push_back(std::atoi(item.c_str()));
    }
    auto wanted = [&](int s) {
        if (only.empty()) return true;
        for (int k : only)
            if (k == s) return true;
        return false;
    };
    std::uint32_t nstates = 1;
    for (int s = 1; s <= steps; ++s) nstates += wanted(s) ? 1u : 0u;
    Out o;
    o.put("PRXD", 4);
    o.u32(3);
    WriteDescription(o, world);
    WriteConf(o, conf);
    o.u32(nstates);
    WriteState(o, world, 0, StepStats{});
    const auto edits = ReadEdits();
    for (int s = 1; s <= steps; ++s) {
        for (const auto& e : edits)
            if (e.step == s) ApplyEdit(world, e);
        const auto st = Step(world, conf);
        if (wanted(s)) {
            WriteState(o, world, std::uint32_t(s), st);
        } else {
            g_events.o.buf.clear();  // the listener calls of an unwritten step
            g_events.n = 0;
        }
    }
    FILE* f = std::fopen(argv[4], "wb");
    if (!f) {
        std::perror("fopen");
        return 1;
    }
    std::fwrite(o.buf.data(), 1, o.buf.size(), f);
    std::fclose(f);
    return 0;
}
