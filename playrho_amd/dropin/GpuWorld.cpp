// GpuWorld: the world-model free functions over the engine's C ABI.  Each
// cites the reference function (detail/WorldModel.hpp forwards to it) whose
// meaning it keeps; see GpuWorld.hpp for what is supported.
#include "GpuWorld.hpp"

#include <cstring>
#include <string>

#include <playrho/InvalidArgument.hpp>
#include <playrho/LengthError.hpp>
#include <playrho/OutOfRange.hpp>
#include <playrho/WrongState.hpp>
#include <playrho/d2/ChainShapeConf.hpp>
#include <playrho/d2/DiskShapeConf.hpp>
#include <playrho/d2/DistanceJointConf.hpp>
#include <playrho/d2/EdgeShapeConf.hpp>
#include <playrho/d2/MultiShapeConf.hpp>
#include <playrho/d2/FrictionJointConf.hpp>
#include <playrho/d2/MotorJointConf.hpp>
#include <playrho/d2/PolygonShapeConf.hpp>
#include <playrho/d2/PrismaticJointConf.hpp>
#include <playrho/d2/TargetJointConf.hpp>
#include <playrho/d2/GearJointConf.hpp>
#include <playrho/d2/PulleyJointConf.hpp>
#include <playrho/d2/RevoluteJointConf.hpp>
#include <playrho/d2/RopeJointConf.hpp>
#include <playrho/d2/WeldJointConf.hpp>
#include <playrho/d2/WheelJointConf.hpp>

namespace playrho_amd {

using namespace playrho;
using namespace playrho::d2;

namespace {

// Status codes of the C ABI back to the reference's exception types.
void Check(int st) {
    if (st == PRX_OK) return;
    const std::string msg = std::string("playrho_amd: ") + prx_last_error();
    switch (st) {
    case PRX_E_INVALID_ARGUMENT: throw InvalidArgument(msg);
    case PRX_E_LENGTH: throw LengthError(msg);
    case PRX_E_WRONG_STATE: throw WrongState(msg);
    case PRX_E_OUT_OF_RANGE: throw OutOfRange<std::size_t>(0u, msg);
    case PRX_E_UNSUPPORTED: throw InvalidArgument(msg);
    default: throw WrongState(msg);
    }
}

[[noreturn]] void Unsupported(const char* what) {
    throw InvalidArgument(std::string("playrho_amd: not supported by the device engine: ") + what);
}

prx_world* NewWorld(const Interval<Positive<Length>>& vr) {
    prx_world_conf c;
    prx_world_conf_default(&c);
    c.min_vertex_radius = float(Real(vr.GetMin()) / Real(Meter));
    c.max_vertex_radius = float(Real(vr.GetMax()) / Real(Meter));
    prx_world* w = nullptr;
    Check(prx_world_create(&c, &w));
    return w;
}

// Device body state: pos0 (x, y, angle), pos1, velocity, under-active time.
struct BodyState {
    float v[10];
    std::uint32_t flags;
};
constexpr std::uint32_t kAwakeFlag = 8u;  // body flag bits of the state records

std::vector<BodyState> ReadBodies(const GpuWorld& g) {
    const std::uint32_t n = prx_world_body_count(g.m_world);
    std::vector<float> f(std::size_t(n) * 10u + 1u);
    std::vector<std::uint32_t> fl(std::size_t(n) + 1u);
    Check(prx_world_get_bodies(g.m_world, f.data(), fl.data()));
    std::vector<BodyState> out(n);
    for (std::uint32_t i = 0; i < n; ++i) {
        std::memcpy(out[i].v, f.data() + 10u * i, sizeof(out[i].v));
        out[i].flags = fl[i];
    }
    return out;
}

std::vector<prx_contact_info> ReadContacts(const GpuWorld& g) {
    std::size_t n = 0;
    Check(prx_world_get_contacts(g.m_world, nullptr, 0, &n));
    std::vector<prx_contact_info> out(n);
    if (n) Check(prx_world_get_contacts(g.m_world, out.data(), n, &n));
    return out;
}

const prx_contact_info& FindContact(const std::vector<prx_contact_info>& cs, ContactID id) {
    for (const auto& c : cs)
        if (c.id == to_underlying(id)) return c;
    throw OutOfRange<ContactID>(id, "playrho_amd: no such contact");
}

Body BodyFromState(const GpuWorld& g, BodyID id, const BodyState& s) {
    Body b = g.m_bodies.at(to_underlying(id));
    float mass[4];
    Check(prx_world_get_body_mass(g.m_world, to_underlying(id), mass));
    b.SetInvMassData(NonNegative<InvMass>{Real(mass[0]) / Kilogram},
                     NonNegative<InvRotInertia>{Real(mass[1]) * SquareRadian / (SquareMeter * Kilogram)});
    const Length2 lc{Real(mass[2]) * Meter, Real(mass[3]) * Meter};
    const Position p0{Length2{Real(s.v[0]) * Meter, Real(s.v[1]) * Meter}, Real(s.v[2]) * Radian};
    const Position p1{Length2{Real(s.v[3]) * Meter, Real(s.v[4]) * Meter}, Real(s.v[5]) * Radian};
    b.SetSweep(Sweep{p0, p1, lc});
    b.JustSetVelocity(Velocity{LinearVelocity2{Real(s.v[6]) * MeterPerSecond, Real(s.v[7]) * MeterPerSecond},
                               Real(s.v[8]) * RadianPerSecond});
    b.SetUnderActiveTime(Real(s.v[9]) * Second);
    if (s.flags & kAwakeFlag) b.SetAwakeFlag();
    else b.UnsetAwakeFlag();
    return b;
}

// Listener trampolines: the engine delivers each step's calls after it.
void OnBegin(void* user, std::uint32_t c) {
    auto* g = static_cast<GpuWorld*>(user);
    if (g->m_listeners.beginContact) g->m_listeners.beginContact(ContactID(c));
}
void OnEnd(void* user, std::uint32_t c) {
    auto* g = static_cast<GpuWorld*>(user);
    if (g->m_listeners.endContact) g->m_listeners.endContact(ContactID(c));
}
Manifold ToManifold(std::uint32_t type, std::uint32_t count, const float ln[2], const float lp[2],
                    const float pts[2][2], const std::uint32_t cfs[2], const float ni[2], const float ti[2]) {
    auto cf = [](std::uint32_t v) {
        return ContactFeature{ContactFeature::Type(v & 1u), ContactFeature::Index((v >> 8) & 0xffu),
                              ContactFeature::Type((v >> 1) & 1u), ContactFeature::Index((v >> 16) & 0xffu)};
    };
    auto point = [&](std::uint32_t k) {
        return Manifold::Point{Length2{Real(pts[k][0]) * Meter, Real(pts[k][1]) * Meter}, cf(cfs[k]),
                               Real(ni[k]) * NewtonSecond, Real(ti[k]) * NewtonSecond};
    };
    const Length2 lpt{Real(lp[0]) * Meter, Real(lp[1]) * Meter};
    if (type == 1u) {  // circles
        Manifold m = Manifold::GetForCircles(lpt, ContactFeature::Index((cfs[0] >> 8) & 0xffu),
                                             Length2{Real(pts[0][0]) * Meter, Real(pts[0][1]) * Meter},
                                             ContactFeature::Index((cfs[0] >> 16) & 0xffu));
        m.SetImpulses(0, Real(ni[0]) * NewtonSecond, Real(ti[0]) * NewtonSecond);
        return m;
    }
    if (type == 2u || type == 3u) {
        const UnitVec n = UnitVec::Get(Real(ln[0]), Real(ln[1])).first;
        if (type == 2u) {
            if (count == 0) return Manifold::GetForFaceA(n, lpt);
            if (count == 1) return Manifold::GetForFaceA(n, lpt, point(0));
            return Manifold::GetForFaceA(n, lpt, point(0), point(1));
        }
        if (count == 0) return Manifold::GetForFaceB(n, lpt);
        if (count == 1) return Manifold::GetForFaceB(n, lpt, point(0));
        return Manifold::GetForFaceB(n, lpt, point(0), point(1));
    }
    return Manifold{};
}
Manifold ToManifold(const prx_manifold& m) {
    const float pts[2][2] = {{m.points[0].local_point[0], m.points[0].local_point[1]},
                             {m.points[1].local_point[0], m.points[1].local_point[1]}};
    const std::uint32_t cfs[2] = {m.points[0].contact_feature, m.points[1].contact_feature};
    const float ni[2] = {m.points[0].normal_impulse, m.points[1].normal_impulse};
    const float ti[2] = {m.points[0].tangent_impulse, m.points[1].tangent_impulse};
    return ToManifold(m.type, m.point_count, m.local_normal, m.local_point, pts, cfs, ni, ti);
}
void OnPreSolve(void* user, std::uint32_t c, const prx_manifold* old) {
    auto* g = static_cast<GpuWorld*>(user);
    if (g->m_listeners.preSolve) g->m_listeners.preSolve(ContactID(c), ToManifold(*old));
}
void OnPostSolve(void* user, std::uint32_t c, const prx_contact_impulses* im, std::uint32_t it) {
    auto* g = static_cast<GpuWorld*>(user);
    if (!g->m_listeners.postSolve) return;
    ContactImpulsesList list;
    for (std::uint32_t k = 0; k < im->count; ++k)
        list.AddEntry(Real(im->normal[k]) * NewtonSecond, Real(im->tangent[k]) * NewtonSecond);
    g->m_listeners.postSolve(ContactID(c), list, it);
}

float F(Real v) { return float(v); }

}  // namespace

// ---------------------------------------------------------------------------
// Construction: World{WorldConf} (d2/World.cpp:28); copies as Clone_ does.

GpuWorld::GpuWorld() : GpuWorld(WorldConf{}) {}

GpuWorld::GpuWorld(const WorldConf& conf) : m_vertexRadius(conf.vertexRadius) { m_world = NewWorld(m_vertexRadius); }

GpuWorld::GpuWorld(const GpuWorld& o)
    : m_vertexRadius(o.m_vertexRadius), m_shapes(o.m_shapes), m_shapeFirst(o.m_shapeFirst),
      m_shapeCount(o.m_shapeCount), m_engineShape(o.m_engineShape), m_bodies(o.m_bodies), m_joints(o.m_joints),
      m_listeners(o.m_listeners), m_invDeltaTime(o.m_invDeltaTime), m_subStepping(o.m_subStepping) {
    m_world = NewWorld(m_vertexRadius);
    std::size_t len = 0;
    Check(prx_world_save_desc(o.m_world, nullptr, 0, &len));
    std::vector<std::uint8_t> buf(len);
    Check(prx_world_save_desc(o.m_world, buf.data(), buf.size(), &len));
    Check(prx_world_load_desc(m_world, buf.data(), len));
    InstallListeners();
}

GpuWorld::GpuWorld(GpuWorld&& o) noexcept
    : m_world(o.m_world), m_vertexRadius(o.m_vertexRadius), m_shapes(std::move(o.m_shapes)),
      m_shapeFirst(std::move(o.m_shapeFirst)), m_shapeCount(std::move(o.m_shapeCount)),
      m_engineShape(std::move(o.m_engineShape)), m_bodies(std::move(o.m_bodies)), m_joints(std::move(o.m_joints)), m_listeners(std::move(o.m_listeners)),
      m_invDeltaTime(o.m_invDeltaTime), m_subStepping(o.m_subStepping) {
    o.m_world = nullptr;
    InstallListeners();
}

GpuWorld& GpuWorld::operator=(const GpuWorld& o) {
    if (this != &o) {
        GpuWorld tmp(o);
        *this = std::move(tmp);
    }
    return *this;
}

GpuWorld& GpuWorld::operator=(GpuWorld&& o) noexcept {
    if (this != &o) {
        if (m_world) prx_world_destroy(m_world);
        m_world = o.m_world;
        o.m_world = nullptr;
        m_vertexRadius = o.m_vertexRadius;
        m_shapes = std::move(o.m_shapes);
        m_shapeFirst = std::move(o.m_shapeFirst);
        m_shapeCount = std::move(o.m_shapeCount);
        m_engineShape = std::move(o.m_engineShape);
        m_bodies = std::move(o.m_bodies);
        m_joints = std::move(o.m_joints);
        m_listeners = std::move(o.m_listeners);
        m_invDeltaTime = o.m_invDeltaTime;
        m_subStepping = o.m_subStepping;
        InstallListeners();
    }
    return *this;
}

GpuWorld::~GpuWorld() {
    if (m_world) prx_world_destroy(m_world);
}

// The engine calls back with this object as the user pointer; any contact
// listener being set turns on the device's recording of the calls.
void GpuWorld::InstallListeners() {
    if (!m_world) return;
    const Listeners& l = m_listeners;
    if (!l.beginContact && !l.endContact && !l.preSolve && !l.postSolve) {
        (void)prx_world_set_contact_listener(m_world, nullptr);
        return;
    }
    prx_contact_listener cl{};
    cl.user = this;
    cl.begin_contact = OnBegin;
    cl.end_contact = OnEnd;
    cl.pre_solve = OnPreSolve;
    cl.post_solve = OnPostSolve;
    (void)prx_world_set_contact_listener(m_world, &cl);
}

bool operator==(const GpuWorld& a, const GpuWorld& b) noexcept {
    if (&a == &b) return true;
    try {
        std::size_t la = 0, lb = 0;
        if (prx_world_state_record(a.m_world, nullptr, 0, &la) || prx_world_state_record(b.m_world, nullptr, 0, &lb))
            return false;
        if (la != lb) return false;
        std::vector<std::uint8_t> ba(la), bb(lb);
        if (prx_world_state_record(a.m_world, ba.data(), la, &la) || prx_world_state_record(b.m_world, bb.data(), lb, &lb))
            return false;
        return ba == bb;
    } catch (...) {
        return false;
    }
}
bool operator!=(const GpuWorld& a, const GpuWorld& b) noexcept { return !(a == b); }

// ---------------------------------------------------------------------------
// Listeners (d2/World.hpp:105-139).  Shape, detach and joint destruction
// listeners are kept but never called: nothing is destroyed.

void SetShapeDestructionListener(GpuWorld& g, ShapeFunction l) noexcept { g.m_listeners.shapeDestruction = std::move(l); }
void SetDetachListener(GpuWorld& g, BodyShapeFunction l) noexcept { g.m_listeners.detach = std::move(l); }
void SetJointDestructionListener(GpuWorld& g, JointFunction l) noexcept { g.m_listeners.jointDestruction = std::move(l); }
void SetBeginContactListener(GpuWorld& g, ContactFunction l) noexcept {
    g.m_listeners.beginContact = std::move(l);
    g.InstallListeners();
}
void SetEndContactListener(GpuWorld& g, ContactFunction l) noexcept {
    g.m_listeners.endContact = std::move(l);
    g.InstallListeners();
}
void SetPreSolveContactListener(GpuWorld& g, ContactManifoldFunction l) noexcept {
    g.m_listeners.preSolve = std::move(l);
    g.InstallListeners();
}
void SetPostSolveContactListener(GpuWorld& g, ContactImpulsesFunction l) noexcept {
    g.m_listeners.postSolve = std::move(l);
    g.InstallListeners();
}

// ---------------------------------------------------------------------------
// Miscellaneous (AabbTreeWorld.hpp / d2/WorldMisc.hpp).

std::optional<pmr::StatsResource::Stats> GetResourceStats(const GpuWorld&) noexcept { return std::nullopt; }

void Clear(GpuWorld& g) noexcept {
    try {
        prx_world* w = NewWorld(g.m_vertexRadius);
        if (g.m_world) prx_world_destroy(g.m_world);
        g.m_world = w;
        g.m_shapes.clear();
        g.m_shapeFirst.clear();
        g.m_shapeCount.clear();
        g.m_engineShape.clear();
        g.m_bodies.clear();
        g.m_joints.clear();
        g.InstallListeners();
    } catch (...) {
    }
}

// Step(AabbTreeWorld&, const StepConf&) (AabbTreeWorld.cpp:2144-2219).
StepStats Step(GpuWorld& g, const StepConf& c) {
    if (g.m_subStepping) throw WrongState("playrho_amd: sub-stepping is not supported by the device engine");
    prx_step_conf s;
    prx_step_conf_default(&s);
    s.delta_time = F(Real(c.deltaTime / Second));
    s.dt_ratio = F(c.dtRatio);
    s.min_still_time_to_sleep = F(Real(c.minStillTimeToSleep / Second));
    s.linear_slop = F(Real(Length{c.linearSlop} / Meter));
    s.angular_slop = F(Real(Angle{c.angularSlop} / Radian));
    s.reg_resolution_rate = F(c.regResolutionRate);
    s.reg_min_separation = F(Real(c.regMinSeparation / Meter));
    s.reg_min_momentum = F(Real(c.regMinMomentum / NewtonSecond));
    s.velocity_threshold = F(Real(c.velocityThreshold / MeterPerSecond));
    s.max_translation = F(Real(c.maxTranslation / Meter));
    s.max_rotation = F(Real(c.maxRotation / Radian));
    s.max_linear_correction = F(Real(c.maxLinearCorrection / Meter));
    s.max_angular_correction = F(Real(c.maxAngularCorrection / Radian));
    s.linear_sleep_tolerance = F(Real(c.linearSleepTolerance / MeterPerSecond));
    s.angular_sleep_tolerance = F(Real(c.angularSleepTolerance / RadianPerSecond));
    s.displace_multiplier = F(c.displaceMultiplier);
    s.aabb_extension = F(Real(c.aabbExtension / Meter));
    s.max_circles_ratio = F(c.maxCirclesRatio);
    s.reg_velocity_iters = c.regVelocityIters;
    s.reg_position_iters = c.regPositionIters;
    s.do_warm_start = c.doWarmStart ? 1u : 0u;
    s.do_toi = c.doToi ? 1u : 0u;
    s.do_blocksolve = c.doBlocksolve ? 1u : 0u;
    s.toi_resolution_rate = F(c.toiResolutionRate);
    s.toi_min_separation = F(Real(c.toiMinSeparation / Meter));
    s.toi_min_momentum = F(Real(c.toiMinMomentum / NewtonSecond));
    s.target_depth = F(Real(Length{c.targetDepth} / Meter));
    s.tolerance = F(Real(Length{c.tolerance} / Meter));
    s.toi_velocity_iters = c.toiVelocityIters;
    s.toi_position_iters = c.toiPositionIters;
    s.max_toi_root_iters = c.maxToiRootIters;
    s.max_toi_iters = c.maxToiIters;
    s.max_distance_iters = c.maxDistanceIters;
    s.max_sub_steps = c.maxSubSteps;
    prx_step_stats st{};
    Check(prx_step(g.m_world, &s, &st));
    if (c.deltaTime != 0 * Second) g.m_invDeltaTime = Real(1) / c.deltaTime;
    StepStats out{};
    out.pre.proxiesCreated = st.pre_proxies_created;
    out.pre.proxiesMoved = st.pre_proxies_moved;
    out.pre.contactsDestroyed = st.pre_contacts_destroyed;
    out.pre.contactsAdded = st.pre_contacts_added;
    out.pre.contactsUpdated = st.pre_contacts_updated;
    out.pre.contactsSkipped = st.pre_contacts_skipped;
    out.reg.minSeparation = Real(st.reg_min_separation) * Meter;
    out.reg.maxIncImpulse = Real(st.reg_max_inc_impulse) * NewtonSecond;
    out.reg.islandsFound = st.reg_islands_found;
    out.reg.islandsSolved = st.reg_islands_solved;
    out.reg.bodiesSlept = st.reg_bodies_slept;
    out.reg.maxIslandBodies = st.reg_max_island_bodies;
    out.reg.contactsAdded = st.reg_contacts_added;
    out.reg.contactsUpdated = st.reg_contacts_updated;
    out.reg.contactsSkipped = st.reg_contacts_skipped;
    out.reg.proxiesMoved = st.reg_proxies_moved;
    out.reg.sumPosIters = st.reg_sum_pos_iters;
    out.reg.sumVelIters = st.reg_sum_vel_iters;
    out.toi.minSeparation = Real(st.toi_min_separation) * Meter;
    out.toi.maxIncImpulse = Real(st.toi_max_inc_impulse) * NewtonSecond;
    out.toi.islandsFound = st.toi_islands_found;
    out.toi.islandsSolved = st.toi_islands_solved;
    out.toi.contactsFound = st.toi_contacts_found;
    out.toi.contactsAtMaxSubSteps = st.toi_contacts_at_max_sub_steps;
    out.toi.contactsUpdatedToi = st.toi_contacts_updated_toi;
    out.toi.contactsUpdatedTouching = st.toi_contacts_updated_touching;
    out.toi.contactsSkippedTouching = st.toi_contacts_skipped_touching;
    out.toi.contactsAdded = st.toi_contacts_added;
    out.toi.proxiesMoved = st.toi_proxies_moved;
    out.toi.sumPosIters = st.toi_sum_pos_iters;
    out.toi.sumVelIters = st.toi_sum_vel_iters;
    out.toi.maxDistIters = st.toi_max_dist_iters;
    out.toi.maxToiIters = st.toi_max_toi_iters;
    out.toi.maxRootIters = st.toi_max_root_iters;
    return out;
}

bool IsStepComplete(const GpuWorld&) noexcept { return true; }
bool GetSubStepping(const GpuWorld& g) noexcept { return g.m_subStepping; }
void SetSubStepping(GpuWorld& g, bool flag) noexcept { g.m_subStepping = flag; }
bool IsLocked(const GpuWorld&) noexcept { return false; }

// A DynamicTree holding the world's proxies: leaves created in proxy order,
// so leaf ids equal the proxy ids (the reference's tree numbers its leaves
// the same way when created in that order).
const DynamicTree& GetTree(const GpuWorld& g) noexcept {
    try {
        std::size_t n = 0;
        Check(prx_world_get_proxies(g.m_world, nullptr, nullptr, nullptr, nullptr, 0, &n));
        std::vector<std::uint32_t> ids(n + 1), bodies(n + 1), shapes(n + 1);
        std::vector<float> ab(4 * n + 4);
        Check(prx_world_get_proxies(g.m_world, ids.data(), bodies.data(), shapes.data(), ab.data(), n, &n));
        g.m_tree = DynamicTree{};
        for (std::size_t i = 0; i < n; ++i) {
            const AABB aabb{LengthInterval{Real(ab[4 * i]) * Meter, Real(ab[4 * i + 2]) * Meter},
                            LengthInterval{Real(ab[4 * i + 1]) * Meter, Real(ab[4 * i + 3]) * Meter}};
            const auto es = g.m_engineShape.at(shapes[i]);
            g.m_tree.CreateLeaf(aabb, Contactable{BodyID(bodies[i]), ShapeID(es.first), ChildCounter(es.second)});
        }
    } catch (...) {
        g.m_tree = DynamicTree{};
    }
    return g.m_tree;
}

// ShiftOrigin (AabbTreeWorld.cpp:2221-2242).
void ShiftOrigin(GpuWorld& g, const Length2& o) {
    const float v[2] = {F(Real(GetX(o) / Meter)), F(Real(GetY(o) / Meter))};
    Check(prx_world_shift_origin(g.m_world, v));
}

Interval<Positive<Length>> GetVertexRadiusInterval(const GpuWorld& g) noexcept { return g.m_vertexRadius; }
Frequency GetInvDeltaTime(const GpuWorld& g) noexcept { return g.m_invDeltaTime; }

// ---------------------------------------------------------------------------
// Bodies (AabbTreeWorld.cpp:1004-1060, 2723-2795; d2/WorldBody.hpp).

BodyCounter GetBodyRange(const GpuWorld& g) noexcept { return BodyCounter(g.m_bodies.size()); }

std::vector<BodyID> GetBodies(const GpuWorld& g) {
    std::vector<BodyID> out;
    for (std::size_t i = 0; i < g.m_bodies.size(); ++i)
        if (!g.m_bodies[i].IsDestroyed()) out.push_back(BodyID(static_cast<BodyID::underlying_type>(i)));
    return out;
}

// Bodies whose proxies are pending synchronisation: the engine synchronises
// edited bodies within its next step, so none are pending between calls.
std::vector<BodyID> GetBodiesForProxies(const GpuWorld&) { return {}; }

// CreateBody (AabbTreeWorld.cpp:1004-1060): the body's shapes must exist; its
// mass data is computed from them as ResetMassData does.
BodyID CreateBody(GpuWorld& g, const Body& body) {
    prx_body_conf c{};
    c.type = GetType(body) == BodyType::Static ? PRX_STATIC
             : GetType(body) == BodyType::Kinematic ? PRX_KINEMATIC : PRX_DYNAMIC;
    const Transformation xf = body.GetTransformation();
    const Sweep& sw = body.GetSweep();
    c.location[0] = F(Real(GetX(xf.p) / Meter));
    c.location[1] = F(Real(GetY(xf.p) / Meter));
    c.angle = F(Real(sw.pos1.angular / Radian));
    const Velocity v = body.GetVelocity();
    c.linear_velocity[0] = F(Real(GetX(v.linear) / MeterPerSecond));
    c.linear_velocity[1] = F(Real(GetY(v.linear) / MeterPerSecond));
    c.angular_velocity = F(Real(v.angular / RadianPerSecond));
    const LinearAcceleration2 la = body.GetLinearAcceleration();
    c.linear_acceleration[0] = F(Real(GetX(la) / MeterPerSquareSecond));
    c.linear_acceleration[1] = F(Real(GetY(la) / MeterPerSquareSecond));
    c.angular_acceleration = F(Real(body.GetAngularAcceleration() / RadianPerSquareSecond));
    c.linear_damping = F(Real(Frequency{body.GetLinearDamping()} / Hertz));
    c.angular_damping = F(Real(Frequency{body.GetAngularDamping()} / Hertz));
    c.allow_sleep = body.IsSleepingAllowed() ? 1u : 0u;
    c.awake = body.IsAwake() ? 1u : 0u;
    c.fixed_rotation = body.IsFixedRotation() ? 1u : 0u;
    c.bullet = body.IsImpenetrable() ? 1u : 0u;
    c.enabled = body.IsEnabled() ? 1u : 0u;
    std::vector<std::uint32_t> shapes;
    for (const ShapeID s : body.GetShapes()) {
        const auto i = to_underlying(s);
        if (i >= g.m_shapes.size()) throw OutOfRange<ShapeID>(s, "playrho_amd: no such shape");
        for (std::uint32_t k = 0; k < g.m_shapeCount[i]; ++k) shapes.push_back(g.m_shapeFirst[i] + k);
    }
    c.shapes = shapes.data();
    c.nshapes = std::uint32_t(shapes.size());
    std::uint32_t id = 0;
    Check(prx_body_create(g.m_world, &c, &id));
    g.m_bodies.push_back(body);
    return BodyID(static_cast<BodyID::underlying_type>(id));
}

Body GetBody(const GpuWorld& g, BodyID id) {
    if (to_underlying(id) >= g.m_bodies.size()) throw OutOfRange<BodyID>(id, "playrho_amd: no such body");
    if (g.m_bodies[to_underlying(id)].IsDestroyed()) throw WasDestroyed<BodyID>(id, "playrho_amd: the body was destroyed");
    const auto st = ReadBodies(g);
    return BodyFromState(g, id, st.at(to_underlying(id)));
}

// SetBody (AabbTreeWorld.cpp:2723-2795) for the changes the engine takes.
void SetBody(GpuWorld& g, BodyID id, const Body& value) {
    const Body cur = GetBody(g, id);
    const auto i = to_underlying(id);
    if (GetType(value) != GetType(cur)) {
        // SetType(World&, BodyID, BodyType) arrives here with Body::SetType
        // applied; the engine makes the change and resets the mass data, as
        // SetType's default resetMassData does (its ResetMassData then finds
        // nothing left to change).
        Check(prx_world_set_type(g.m_world, i, std::uint32_t(GetType(value))));
        g.m_bodies[i] = value;  // type, flags and accelerations; the state comes from the engine
        return;
    }
    if (value.GetShapes() != cur.GetShapes()) Unsupported("attaching or detaching shapes after creation");
    if (value.GetInvMass() != cur.GetInvMass() || value.GetInvRotInertia() != cur.GetInvRotInertia() ||
        value.GetSweep().localCenter != cur.GetSweep().localCenter)
        Unsupported("mass data other than the shapes' (ResetMassData)");
    if (value.IsEnabled() != cur.IsEnabled() || value.IsFixedRotation() != cur.IsFixedRotation() ||
        value.IsImpenetrable() != cur.IsImpenetrable() || value.IsSleepingAllowed() != cur.IsSleepingAllowed() ||
        value.GetLinearDamping() != cur.GetLinearDamping() || value.GetAngularDamping() != cur.GetAngularDamping())
        Unsupported("changing body flags or damping");
    const Sweep& a = value.GetSweep();
    const Sweep& b = cur.GetSweep();
    if (a.pos1 != b.pos1 || a.pos0 != b.pos0) {
        const Transformation xf = value.GetTransformation();
        const float t[4] = {F(Real(GetX(xf.p) / Meter)), F(Real(GetY(xf.p) / Meter)), F(GetX(xf.q)), F(GetY(xf.q))};
        Check(prx_world_set_transformation(g.m_world, i, t));
    }
    if (value.GetVelocity() != cur.GetVelocity()) {
        const Velocity v = value.GetVelocity();
        const float f[3] = {F(Real(GetX(v.linear) / MeterPerSecond)), F(Real(GetY(v.linear) / MeterPerSecond)),
                            F(Real(v.angular / RadianPerSecond))};
        Check(prx_world_set_velocity(g.m_world, i, f));
    }
    if (value.GetLinearAcceleration() != cur.GetLinearAcceleration() ||
        value.GetAngularAcceleration() != cur.GetAngularAcceleration()) {
        const LinearAcceleration2 la = value.GetLinearAcceleration();
        const float f[3] = {F(Real(GetX(la) / MeterPerSquareSecond)), F(Real(GetY(la) / MeterPerSquareSecond)),
                            F(Real(value.GetAngularAcceleration() / RadianPerSquareSecond))};
        Check(prx_world_set_acceleration(g.m_world, i, f));
    }
}

// Destroy (AabbTreeWorld.cpp:1062-1104): the engine destroys the body's
// joints, contacts and proxies; the body keeps its id.
void Destroy(GpuWorld& g, BodyID id) {
    const auto i = to_underlying(id);
    if (i >= g.m_bodies.size()) throw OutOfRange<BodyID>(id, "playrho_amd: no such body");
    if (g.m_bodies[i].IsDestroyed()) throw WasDestroyed<BodyID>(id, "playrho_amd: the body was destroyed");
    Check(prx_world_destroy_body(g.m_world, i));
    for (auto& j : g.m_joints)
        if (j.has_value() && (GetBodyA(j) == id || GetBodyB(j) == id)) j = Joint{};
    SetDestroyed(g.m_bodies[i]);
}

std::vector<std::pair<BodyID, JointID>> GetJoints(const GpuWorld& g, BodyID id) {
    std::vector<std::pair<BodyID, JointID>> out;
    for (std::size_t j = 0; j < g.m_joints.size(); ++j) {
        const Joint& jt = g.m_joints[j];
        const JointID jid(static_cast<JointID::underlying_type>(j));
        if (!jt.has_value()) continue;
        if (GetBodyA(jt) == id) out.emplace_back(GetBodyB(jt), jid);
        else if (GetBodyB(jt) == id) out.emplace_back(GetBodyA(jt), jid);
    }
    return out;
}

std::vector<std::tuple<ContactKey, ContactID>> GetContacts(const GpuWorld& g, BodyID id) {
    std::vector<std::tuple<ContactKey, ContactID>> out;
    for (const auto& c : ReadContacts(g)) {
        if (c.body_a == to_underlying(id) || c.body_b == to_underlying(id))
            out.emplace_back(ContactKey{c.key_min, c.key_max}, ContactID(static_cast<ContactID::underlying_type>(c.id)));
    }
    return out;
}

std::vector<ShapeID> GetShapes(const GpuWorld& g, BodyID id) { return g.m_bodies.at(to_underlying(id)).GetShapes(); }

// ---------------------------------------------------------------------------
// Joints (AabbTreeWorld.cpp:1132-1260).

JointCounter GetJointRange(const GpuWorld& g) noexcept { return JointCounter(g.m_joints.size()); }

std::vector<JointID> GetJoints(const GpuWorld& g) {
    std::vector<JointID> out;
    for (std::size_t j = 0; j < g.m_joints.size(); ++j)
        if (g.m_joints[j].has_value()) out.push_back(JointID(static_cast<JointID::underlying_type>(j)));
    return out;
}

JointID CreateJoint(GpuWorld& g, const Joint& def) {
    const std::uint32_t bA = to_underlying(GetBodyA(def)), bB = to_underlying(GetBodyB(def));
    const std::uint32_t cc = GetCollideConnected(def) ? 1u : 0u;
    auto xy = [](const Length2& v, float out[2]) {
        out[0] = F(Real(GetX(v) / Meter));
        out[1] = F(Real(GetY(v) / Meter));
    };
    std::uint32_t id = 0;
    if (const auto* j = TypeCast<const RevoluteJointConf>(&def)) {
        prx_revolute_joint_conf c{};
        c.body_a = bA;
        c.body_b = bB;
        xy(j->localAnchorA, c.local_anchor_a);
        xy(j->localAnchorB, c.local_anchor_b);
        c.reference_angle = F(Real(j->referenceAngle / Radian));
        c.enable_limit = j->enableLimit ? 1u : 0u;
        c.lower_angle = F(Real(j->lowerAngle / Radian));
        c.upper_angle = F(Real(j->upperAngle / Radian));
        c.enable_motor = j->enableMotor ? 1u : 0u;
        c.motor_speed = F(Real(j->motorSpeed / RadianPerSecond));
        c.max_motor_torque = F(Real(j->maxMotorTorque / NewtonMeter));
        c.collide_connected = cc;
        Check(prx_joint_create(g.m_world, &c, &id));
    } else if (const auto* j = TypeCast<const DistanceJointConf>(&def)) {
        prx_distance_joint_conf c{};
        c.body_a = bA;
        c.body_b = bB;
        xy(j->localAnchorA, c.local_anchor_a);
        xy(j->localAnchorB, c.local_anchor_b);
        c.length = F(Real(j->length / Meter));
        c.frequency = F(Real(Frequency{j->frequency} / Hertz));
        c.damping_ratio = F(j->dampingRatio);
        c.collide_connected = cc;
        Check(prx_distance_joint_create(g.m_world, &c, &id));
    } else if (const auto* j = TypeCast<const WeldJointConf>(&def)) {
        prx_weld_joint_conf c{};
        c.body_a = bA;
        c.body_b = bB;
        xy(j->localAnchorA, c.local_anchor_a);
        xy(j->localAnchorB, c.local_anchor_b);
        c.reference_angle = F(Real(j->referenceAngle / Radian));
        c.frequency = F(Real(Frequency{j->frequency} / Hertz));
        c.damping_ratio = F(j->dampingRatio);
        c.collide_connected = cc;
        Check(prx_weld_joint_create(g.m_world, &c, &id));
    } else if (const auto* j = TypeCast<const RopeJointConf>(&def)) {
        prx_rope_joint_conf c{};
        c.body_a = bA;
        c.body_b = bB;
        xy(j->localAnchorA, c.local_anchor_a);
        xy(j->localAnchorB, c.local_anchor_b);
        c.max_length = F(Real(j->maxLength / Meter));
        c.collide_connected = cc;
        Check(prx_rope_joint_create(g.m_world, &c, &id));
    } else if (const auto* j = TypeCast<const FrictionJointConf>(&def)) {
        prx_friction_joint_conf c{};
        c.body_a = bA;
        c.body_b = bB;
        xy(j->localAnchorA, c.local_anchor_a);
        xy(j->localAnchorB, c.local_anchor_b);
        c.max_force = F(Real(Force{j->maxForce} / Newton));
        c.max_torque = F(Real(Torque{j->maxTorque} / NewtonMeter));
        c.collide_connected = cc;
        Check(prx_friction_joint_create(g.m_world, &c, &id));
    } else if (const auto* j = TypeCast<const MotorJointConf>(&def)) {
        prx_motor_joint_conf c{};
        c.body_a = bA;
        c.body_b = bB;
        xy(j->linearOffset, c.linear_offset);
        c.angular_offset = F(Real(j->angularOffset / Radian));
        c.max_force = F(Real(Force{j->maxForce} / Newton));
        c.max_torque = F(Real(Torque{j->maxTorque} / NewtonMeter));
        c.correction_factor = F(j->correctionFactor);
        c.collide_connected = cc;
        Check(prx_motor_joint_create(g.m_world, &c, &id));
    } else if (const auto* j = TypeCast<const WheelJointConf>(&def)) {
        prx_wheel_joint_conf c{};
        c.body_a = bA;
        c.body_b = bB;
        xy(j->localAnchorA, c.local_anchor_a);
        xy(j->localAnchorB, c.local_anchor_b);
        c.local_axis_a[0] = F(GetX(j->localXAxisA));
        c.local_axis_a[1] = F(GetY(j->localXAxisA));
        c.enable_motor = j->enableMotor ? 1u : 0u;
        c.motor_speed = F(Real(j->motorSpeed / RadianPerSecond));
        c.max_motor_torque = F(Real(j->maxMotorTorque / NewtonMeter));
        c.frequency = F(Real(Frequency{j->frequency} / Hertz));
        c.damping_ratio = F(j->dampingRatio);
        c.collide_connected = cc;
        Check(prx_wheel_joint_create(g.m_world, &c, &id));
    } else if (const auto* j = TypeCast<const PulleyJointConf>(&def)) {
        prx_pulley_joint_conf c{};
        c.body_a = bA;
        c.body_b = bB;
        xy(j->groundAnchorA, c.ground_anchor_a);
        xy(j->groundAnchorB, c.ground_anchor_b);
        xy(j->localAnchorA, c.local_anchor_a);
        xy(j->localAnchorB, c.local_anchor_b);
        c.length_a = F(Real(j->lengthA / Meter));
        c.length_b = F(Real(j->lengthB / Meter));
        c.ratio = F(j->ratio);
        c.collide_connected = cc;
        Check(prx_pulley_joint_create(g.m_world, &c, &id));
    } else if (const auto* j = TypeCast<const PrismaticJointConf>(&def)) {
        prx_prismatic_joint_conf c{};
        c.body_a = bA;
        c.body_b = bB;
        xy(j->localAnchorA, c.local_anchor_a);
        xy(j->localAnchorB, c.local_anchor_b);
        c.local_axis_a[0] = F(GetX(j->localXAxisA));
        c.local_axis_a[1] = F(GetY(j->localXAxisA));
        c.reference_angle = F(Real(j->referenceAngle / Radian));
        c.enable_limit = j->enableLimit ? 1u : 0u;
        c.lower_translation = F(Real(j->lowerTranslation / Meter));
        c.upper_translation = F(Real(j->upperTranslation / Meter));
        c.enable_motor = j->enableMotor ? 1u : 0u;
        c.motor_speed = F(Real(j->motorSpeed / RadianPerSecond));
        c.max_motor_force = F(Real(j->maxMotorForce / Newton));
        c.collide_connected = cc;
        Check(prx_prismatic_joint_create(g.m_world, &c, &id));
    } else if (const auto* j = TypeCast<const TargetJointConf>(&def)) {
        prx_target_joint_conf c{};
        c.body = bB;
        xy(j->target, c.target);
        xy(j->localAnchorB, c.local_anchor);
        c.max_force = F(Real(j->maxForce / Newton));
        c.frequency = F(Real(j->frequency / Hertz));
        c.damping_ratio = F(j->dampingRatio);
        c.collide_connected = cc;
        Check(prx_target_joint_create(g.m_world, &c, &id));
    } else if (const auto* j = TypeCast<const GearJointConf>(&def)) {
        // The conf as GetGearJointConf made it (sides, ratio, constant).
        prx_gear_joint_conf c{};
        c.body_a = bA;
        c.body_b = bB;
        c.body_c = std::uint32_t(to_underlying(j->bodyC));
        c.body_d = std::uint32_t(to_underlying(j->bodyD));
        const GearJointConf::TypeData* sides[2] = {&j->typeDataAC, &j->typeDataBD};
        for (int k = 0; k < 2; ++k) {
            float* out = k ? c.bd : c.ac;
            if (const auto* pd = std::get_if<GearJointConf::PrismaticData>(sides[k])) {
                (k ? c.prismatic_bd : c.prismatic_ac) = 1u;
                xy(pd->localAnchorA, out);
                xy(pd->localAnchorB, out + 2);
                out[4] = F(GetX(pd->localAxis));
                out[5] = F(GetY(pd->localAxis));
            } else if (const auto* rd = std::get_if<GearJointConf::RevoluteData>(sides[k])) {
                out[0] = F(Real(rd->referenceAngle / Radian));
            } else {
                throw InvalidArgument("playrho_amd: a gear joint side that is neither revolute nor prismatic");
            }
        }
        c.ratio = F(j->ratio);
        c.constant = F(j->constant);
        c.collide_connected = cc;
        Check(prx_gear_joint_create(g.m_world, &c, &id));
    } else {
        Unsupported("this joint type");
    }
    // The engine reuses the id destroyed last (ObjectPool).
    if (id < g.m_joints.size()) g.m_joints[id] = def;
    else g.m_joints.push_back(def);
    return JointID(static_cast<JointID::underlying_type>(id));
}

Joint GetJoint(const GpuWorld& g, JointID id) {
    if (to_underlying(id) >= g.m_joints.size()) throw OutOfRange<JointID>(id, "playrho_amd: no such joint");
    if (!g.m_joints[to_underlying(id)].has_value())
        throw WasDestroyed<JointID>(id, "playrho_amd: the joint was destroyed");
    return g.m_joints[to_underlying(id)];
}

void SetJoint(GpuWorld& g, JointID id, const Joint& def) {
    if (!(GetJoint(g, id) == def)) Unsupported("changing a joint after creation");
}

// Destroy (AabbTreeWorld.cpp:1215-1227).
void Destroy(GpuWorld& g, JointID id) {
    const auto i = to_underlying(id);
    if (i >= g.m_joints.size()) throw OutOfRange<JointID>(id, "playrho_amd: no such joint");
    if (!g.m_joints[i].has_value()) return;
    Check(prx_world_destroy_joint(g.m_world, i));
    g.m_joints[i] = Joint{};
}

// ---------------------------------------------------------------------------
// Shapes (AabbTreeWorld.cpp:1236-1300; d2/WorldShape.hpp).

ShapeCounter GetShapeRange(const GpuWorld& g) noexcept { return ShapeCounter(g.m_shapes.size()); }

// The drop-in's record of a shape made of `count` engine shapes from `first` on.
ShapeID AddShapeRecord(GpuWorld& g, const Shape& def, std::uint32_t first, std::uint32_t count) {
    const auto id = std::uint32_t(g.m_shapes.size());
    g.m_shapes.push_back(def);
    g.m_shapeFirst.push_back(first);
    g.m_shapeCount.push_back(count);
    if (g.m_engineShape.size() < first + count) g.m_engineShape.resize(first + count);
    for (std::uint32_t k = 0; k < count; ++k) g.m_engineShape[first + k] = {id, k};
    return ShapeID(static_cast<ShapeID::underlying_type>(id));
}

ShapeID CreateShape(GpuWorld& g, const Shape& def) {
    prx_shape_conf c{};
    c.friction = F(Real(GetFriction(def)));
    c.restitution = F(Real(GetRestitution(def)));
    c.density = F(Real(AreaDensity{GetDensity(def)} / KilogramPerSquareMeter));
    c.sensor = IsSensor(def) ? 1u : 0u;
    const Filter flt = GetFilter(def);
    c.category_bits = flt.categoryBits;
    c.mask_bits = flt.maskBits;
    c.group_index = flt.groupIndex;
    std::vector<float> pts;
    if (const auto* s = TypeCast<const DiskShapeConf>(&def)) {
        c.kind = PRX_SHAPE_DISK;
        c.radius = F(Real(Length{s->vertexRadius} / Meter));
        c.vertex_radius = c.radius;
        c.location[0] = F(Real(GetX(s->location) / Meter));
        c.location[1] = F(Real(GetY(s->location) / Meter));
    } else if (const auto* s = TypeCast<const PolygonShapeConf>(&def)) {
        c.kind = PRX_SHAPE_POLYGON;
        c.vertex_radius = F(Real(Length{s->vertexRadius} / Meter));
        for (VertexCounter i = 0; i < s->GetVertexCount(); ++i) {
            pts.push_back(F(Real(GetX(s->GetVertex(i)) / Meter)));
            pts.push_back(F(Real(GetY(s->GetVertex(i)) / Meter)));
        }
    } else if (const auto* s = TypeCast<const EdgeShapeConf>(&def)) {
        c.kind = PRX_SHAPE_EDGE;
        c.vertex_radius = F(Real(Length{s->vertexRadius} / Meter));
        for (const Length2 v : {s->GetVertexA(), s->GetVertexB()}) {
            pts.push_back(F(Real(GetX(v) / Meter)));
            pts.push_back(F(Real(GetY(v) / Meter)));
        }
    } else if (const auto* s = TypeCast<const ChainShapeConf>(&def)) {
        c.kind = PRX_SHAPE_CHAIN;
        c.vertex_radius = F(Real(Length{s->vertexRadius} / Meter));
        for (std::size_t i = 0; i < s->GetVertexCount(); ++i) {
            pts.push_back(F(Real(GetX(s->GetVertex(ChildCounter(i))) / Meter)));
            pts.push_back(F(Real(GetY(s->GetVertex(ChildCounter(i))) / Meter)));
        }
    } else if (const auto* s = TypeCast<const MultiShapeConf>(&def)) {
        // Each hull's vertices (its convex hull again is itself).
        c.kind = PRX_SHAPE_MULTI;
        std::vector<std::uint32_t> sizes;
        for (const ConvexHull& ch : s->children) {
            const auto dp = ch.GetDistanceProxy();
            c.vertex_radius = F(Real(Length{ch.GetVertexRadius()} / Meter));
            sizes.push_back(std::uint32_t(dp.GetVertexCount()));
            for (VertexCounter i = 0; i < dp.GetVertexCount(); ++i) {
                pts.push_back(F(Real(GetX(dp.GetVertex(i)) / Meter)));
                pts.push_back(F(Real(GetY(dp.GetVertex(i)) / Meter)));
            }
        }
        c.points = pts.empty() ? nullptr : pts.data();
        c.npoints = std::uint32_t(pts.size() / 2);
        std::uint32_t first = 0, count = 0;
        Check(prx_compound_shape_create(g.m_world, &c, sizes.data(), std::uint32_t(sizes.size()), &first, &count));
        return AddShapeRecord(g, def, first, count);
    } else {
        Unsupported("this shape type");
    }
    c.points = pts.empty() ? nullptr : pts.data();
    c.npoints = std::uint32_t(pts.size() / 2);
    std::uint32_t first = 0, count = 1;
    if (c.kind == PRX_SHAPE_CHAIN) Check(prx_compound_shape_create(g.m_world, &c, nullptr, 0, &first, &count));
    else Check(prx_shape_create(g.m_world, &c, &first));
    return AddShapeRecord(g, def, first, count);
}

Shape GetShape(const GpuWorld& g, ShapeID id) {
    if (to_underlying(id) >= g.m_shapes.size()) throw OutOfRange<ShapeID>(id, "playrho_amd: no such shape");
    return g.m_shapes[to_underlying(id)];
}

// SetShape (WorldShape.cpp:79-96): friction and restitution changes.
void SetShape(GpuWorld& g, ShapeID id, const Shape& def) {
    Shape cur = GetShape(g, id);
    const auto i = to_underlying(id);
    if (GetFriction(def) != GetFriction(cur)) {
        for (std::uint32_t k = 0; k < g.m_shapeCount[i]; ++k)
            Check(prx_shape_set_friction(g.m_world, g.m_shapeFirst[i] + k, F(Real(GetFriction(def)))));
        SetFriction(cur, GetFriction(def));
    }
    if (GetRestitution(def) != GetRestitution(cur)) {
        for (std::uint32_t k = 0; k < g.m_shapeCount[i]; ++k)
            Check(prx_shape_set_restitution(g.m_world, g.m_shapeFirst[i] + k, F(Real(GetRestitution(def)))));
        SetRestitution(cur, GetRestitution(def));
    }
    if (!(cur == def)) Unsupported("changing a shape other than its friction and restitution");
    g.m_shapes[i] = def;
}

void Destroy(GpuWorld&, ShapeID) { throw WrongState("playrho_amd: destroying shapes is not supported by the device engine"); }

// ---------------------------------------------------------------------------
// Contacts (d2/WorldContact.hpp; AabbTreeWorld.hpp:494-527).

ContactCounter GetContactRange(const GpuWorld& g) noexcept {
    try {
        ContactCounter n = 0;
        for (const auto& c : ReadContacts(g)) n = std::max<ContactCounter>(n, ContactCounter(c.id + 1u));
        return n;
    } catch (...) {
        return 0;
    }
}

std::vector<KeyedContactID> GetContacts(const GpuWorld& g) {
    std::vector<KeyedContactID> out;
    for (const auto& c : ReadContacts(g))
        out.emplace_back(ContactKey{c.key_min, c.key_max}, ContactID(static_cast<ContactID::underlying_type>(c.id)));
    return out;
}

Contact GetContact(const GpuWorld& g, ContactID id) {
    const auto cs = ReadContacts(g);
    const prx_contact_info& c = FindContact(cs, id);
    const auto ea = g.m_engineShape.at(c.shape_a), eb = g.m_engineShape.at(c.shape_b);
    Contact out{Contactable{BodyID(static_cast<BodyID::underlying_type>(c.body_a)),
                            ShapeID(static_cast<ShapeID::underlying_type>(ea.first)), ChildCounter(ea.second)},
                Contactable{BodyID(static_cast<BodyID::underlying_type>(c.body_b)),
                            ShapeID(static_cast<ShapeID::underlying_type>(eb.first)), ChildCounter(eb.second)}};
    out.SetFriction(NonNegative<Real>{Real(c.friction)});
    out.SetRestitution(Real(c.restitution));
    if (c.flags & 1u) out.SetTouching();
    else out.UnsetTouching();
    if (c.flags & 2u) out.SetEnabled();
    else out.UnsetEnabled();
    if (c.flags & 4u) out.SetSensor();
    if (c.flags & 8u) out.SetImpenetrable();
    if (!(c.flags & 16u)) out.UnflagForUpdating();
    return out;
}

void SetContact(GpuWorld&, ContactID, const Contact&) {
    throw WrongState("playrho_amd: contacts are owned by the device engine");
}

Manifold GetManifold(const GpuWorld& g, ContactID id) {
    const auto cs = ReadContacts(g);
    const prx_contact_info& c = FindContact(cs, id);
    const float pts[2][2] = {{c.points[0].local_point[0], c.points[0].local_point[1]},
                             {c.points[1].local_point[0], c.points[1].local_point[1]}};
    const std::uint32_t cfs[2] = {c.points[0].contact_feature, c.points[1].contact_feature};
    const float ni[2] = {c.points[0].normal_impulse, c.points[1].normal_impulse};
    const float ti[2] = {c.points[0].tangent_impulse, c.points[1].tangent_impulse};
    return ToManifold(c.manifold_type, c.point_count, c.local_normal, c.local_point, pts, cfs, ni, ti);
}

void SetManifold(GpuWorld&, ContactID, const Manifold&) {
    throw WrongState("playrho_amd: manifolds are owned by the device engine");
}

}  // namespace playrho_amd
