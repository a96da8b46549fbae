// The literal drop-in: PlayRho's own API on playrho::d2::World{GpuWorld{}}
// beside the reference's World{} (AabbTreeWorld), built against the
// reference's headers and library (oracle/Makefile.ref, target dropin_check).
//
//   dropin_check cpu    construction only (no device needed): shapes, bodies
//                       and joints created through the reference API; every
//                       body's transformation, velocity and mass data and
//                       every shape equal the reference world's.
//   dropin_check hello  HelloWorld's program (HelloWorld/HelloWorld.cpp) on
//                       World{GpuWorld{}}: prints its 60 lines (the test
//                       compares them with HelloWorld/ExpectedOutput.txt).
//   dropin_check gpu    the 20-row pyramid stepped on both worlds through
//                       Step(World&, const StepConf&): after every step each
//                       body's transformation and velocity, the contact keys
//                       and every manifold equal the reference's bit for bit;
//                       then SetVelocity and SetTransformation edits and a
//                       contact listener through the same API.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <iomanip>
#include <iostream>
#include <string>
#include <vector>

#include <playrho/d2/VertexSet.hpp>
#include <playrho/d2/MultiShapeConf.hpp>
#include <playrho/d2/ChainShapeConf.hpp>
#include <playrho/WrongState.hpp>
#include <playrho/d2/BasicAPI.hpp>
#include <playrho/d2/RevoluteJointConf.hpp>
#include <playrho/d2/PrismaticJointConf.hpp>
#include <playrho/d2/GearJointConf.hpp>
#include <playrho/d2/WorldContact.hpp>

#include "../../playrho_amd/dropin/GpuWorld.hpp"

using namespace playrho;
using namespace playrho::d2;

namespace {

int g_fail = 0;

template <class T>
bool Same(const T& a, const T& b) {
    return std::memcmp(&a, &b, sizeof(T)) == 0;
}

void Expect(bool ok, const std::string& what) {
    if (!ok) {
        ++g_fail;
        if (g_fail < 20) std::fprintf(stderr, "MISMATCH: %s\n", what.c_str());
    }
}

// The pyramid of the first benchmark workload (20 rows of 0.5 half-extent
// boxes on a ground edge), built through the reference API on any World.
void BuildPyramid(World& world, int rows) {
    const auto ground = CreateShape(world, EdgeShapeConf{Length2{-40_m, 0_m}, Length2{40_m, 0_m}});
    CreateBody(world, BodyConf{}.Use(ground));
    const auto box = CreateShape(world, PolygonShapeConf{}.UseDensity(1_kgpm2).UseFriction(Real(0.6)).SetAsBox(0.5_m, 0.5_m));
    auto x = Length2{-7_m, 0.75_m};
    for (auto i = 0; i < rows; ++i) {
        auto y = x;
        for (auto j = i; j < rows; ++j) {
            CreateBody(world, BodyConf{}
                                  .Use(BodyType::Dynamic)
                                  .Use(box)
                                  .UseLocation(y)
                                  .UseLinearAcceleration(LinearAcceleration2{0, -10 * MeterPerSquareSecond}));
            y += Length2{1.125_m, 0_m};
        }
        x += Length2{0.5625_m, 1.25_m};
    }
}

void BuildChain(World& world) {
    const auto ground = CreateShape(world, EdgeShapeConf{Length2{-40_m, 0_m}, Length2{40_m, 0_m}});
    const auto g = CreateBody(world, BodyConf{}.Use(ground));
    const auto link = CreateShape(world, PolygonShapeConf{}.UseDensity(1_kgpm2).SetAsBox(0.5_m, 0.125_m));
    auto prev = g;
    for (int i = 0; i < 8; ++i) {
        const auto b = CreateBody(world, BodyConf{}
                                             .Use(BodyType::Dynamic)
                                             .Use(link)
                                             .UseLocation(Length2{Real(0.5f + i) * Meter, 12_m})
                                             .UseLinearAcceleration(LinearAcceleration2{0, -10 * MeterPerSquareSecond}));
        auto jc = RevoluteJointConf{}.UseBodyA(prev).UseBodyB(b);
        jc.localAnchorA = (i == 0) ? Length2{0_m, 12_m} : Length2{0.5_m, 0_m};
        jc.localAnchorB = Length2{-0.5_m, 0_m};
        CreateJoint(world, Joint{jc});
        prev = b;
    }
}

// Two wheels on revolute joints to the ground, the first motorised, meshed
// by a gear, and a rack on a horizontal prismatic joint geared to the second
// wheel (GetGearJointConf on the world being built), beside two falling boxes.
// Returns the gear joints.
std::vector<JointID> BuildGears(World& world) {
    const auto ground = CreateShape(world, EdgeShapeConf{Length2{-40_m, 0_m}, Length2{40_m, 0_m}});
    const auto g = CreateBody(world, BodyConf{}.Use(ground));
    const auto gravity = LinearAcceleration2{0, -10 * MeterPerSquareSecond};
    std::vector<JointID> revs, gears;
    for (int i = 0; i < 2; ++i) {
        const auto disk = CreateShape(world, DiskShapeConf{}.UseRadius(Real(1.0f + 0.5f * i) * Meter).UseDensity(1_kgpm2));
        const auto at = Length2{Real(-4.0f + 5.0f * i) * Meter, 6_m};
        const auto b = CreateBody(world, BodyConf{}.Use(BodyType::Dynamic).Use(disk).UseLocation(at)
                                             .UseAngle(Real(0.3f * i) * Radian).UseLinearAcceleration(gravity));
        auto jc = RevoluteJointConf{}.UseBodyA(g).UseBodyB(b);
        jc.localAnchorA = at;
        jc.referenceAngle = Real(0.3f * i) * Radian;
        if (i == 0) jc.UseEnableMotor(true).UseMotorSpeed(1.5_rad / 1_s).UseMaxMotorTorque(200_Nm);
        revs.push_back(CreateJoint(world, Joint{jc}));
    }
    gears.push_back(CreateJoint(world, Joint{GetGearJointConf(world, revs[0], revs[1], 1.0f / 1.5f)}));
    const auto bar = CreateShape(world, PolygonShapeConf{}.UseDensity(1_kgpm2).SetAsBox(3_m, 0.25_m));
    const auto rack = CreateBody(world, BodyConf{}.Use(BodyType::Dynamic).Use(bar).UseLocation(Length2{1_m, 4.2_m})
                                            .UseLinearAcceleration(gravity));
    const auto pj = CreateJoint(world, Joint{PrismaticJointConf{g, rack, Length2{1_m, 4.2_m}, Length2{}, UnitVec::GetRight()}});
    gears.push_back(CreateJoint(world, Joint{GetGearJointConf(world, revs[1], pj, -1.0f / 1.5f)}));
    const auto box = CreateShape(world, PolygonShapeConf{}.UseDensity(1_kgpm2).SetAsBox(0.4_m, 0.4_m));
    for (int i = 0; i < 2; ++i)
        CreateBody(world, BodyConf{}.Use(BodyType::Dynamic).Use(box).UseLocation(Length2{Real(-4.5f + 6.0f * i) * Meter, 12_m})
                              .UseLinearAcceleration(gravity));
    return gears;
}

void CompareBodies(const World& a, const World& b, const std::string& when) {
    Expect(GetBodyRange(a) == GetBodyRange(b), when + ": body count");
    for (const auto id : GetBodies(a)) {
        const std::string w = when + ": body " + std::to_string(to_underlying(id));
        const auto ta = GetTransformation(a, id), tb = GetTransformation(b, id);
        Expect(Same(ta.p, tb.p) && Same(GetX(ta.q), GetX(tb.q)) && Same(GetY(ta.q), GetY(tb.q)), w + " transformation");
        const auto va = GetVelocity(a, id), vb = GetVelocity(b, id);
        Expect(Same(va.linear, vb.linear) && Same(va.angular, vb.angular), w + " velocity");
        Expect(Same(GetInvMass(a, id), GetInvMass(b, id)), w + " inverse mass");
        Expect(Same(GetInvRotInertia(a, id), GetInvRotInertia(b, id)), w + " inverse rotational inertia");
        Expect(Same(GetLocalCenter(a, id), GetLocalCenter(b, id)), w + " local centre");
        Expect(IsAwake(a, id) == IsAwake(b, id), w + " awake");
    }
}

void CompareContacts(const World& a, const World& b, const std::string& when) {
    const auto ca = GetContacts(a), cb = GetContacts(b);
    Expect(ca.size() == cb.size(), when + ": contact count " + std::to_string(ca.size()) + " vs " + std::to_string(cb.size()));
    for (std::size_t i = 0; i < ca.size() && i < cb.size(); ++i) {
        const std::string w = when + ": contact " + std::to_string(i);
        Expect(std::get<ContactKey>(ca[i]) == std::get<ContactKey>(cb[i]), w + " key");
        Expect(std::get<ContactID>(ca[i]) == std::get<ContactID>(cb[i]), w + " id");
        const auto ida = std::get<ContactID>(ca[i]);
        Expect(IsTouching(a, ida) == IsTouching(b, ida), w + " touching");
        const auto ma = GetManifold(a, ida), mb = GetManifold(b, ida);
        Expect(ma.GetType() == mb.GetType() && ma.GetPointCount() == mb.GetPointCount(), w + " manifold type");
        if (ma.GetType() == Manifold::e_faceA || ma.GetType() == Manifold::e_faceB)
            Expect(Same(GetX(ma.GetLocalNormal()), GetX(mb.GetLocalNormal())) &&
                       Same(GetY(ma.GetLocalNormal()), GetY(mb.GetLocalNormal())),
                   w + " local normal");
        Expect(Same(ma.GetLocalPoint(), mb.GetLocalPoint()), w + " local point");
        for (auto k = 0u; k < ma.GetPointCount() && k < mb.GetPointCount(); ++k) {
            const auto& pa = ma.GetPoint(k);
            const auto& pb = mb.GetPoint(k);
            Expect(Same(pa.localPoint, pb.localPoint) && pa.contactFeature == pb.contactFeature &&
                       Same(pa.normalImpulse, pb.normalImpulse) && Same(pa.tangentImpulse, pb.tangentImpulse),
                   w + " point " + std::to_string(k));
        }
    }
}

int Cpu() {
    auto ref = World{};
    auto gpu = World{playrho_amd::GpuWorld{}};
    BuildPyramid(ref, 20);
    BuildPyramid(gpu, 20);
    CompareBodies(ref, gpu, "pyramid built");
    auto ref2 = World{};
    auto gpu2 = World{playrho_amd::GpuWorld{}};
    BuildChain(ref2);
    BuildChain(gpu2);
    CompareBodies(ref2, gpu2, "chain built");
    Expect(GetJointRange(ref2) == GetJointRange(gpu2), "joint count");
    Expect(GetShapeRange(ref) == GetShapeRange(gpu), "shape count");
    for (auto s = 0u; s < GetShapeRange(ref); ++s)
        Expect(GetShape(ref, ShapeID(s)) == GetShape(gpu, ShapeID(s)), "shape " + std::to_string(s));
    // The type-erased copy (Clone_) of a device world and its equality.
    const World copy = gpu;
    Expect(copy == gpu, "copy equals");
    // Destroy(JointID) and Destroy(BodyID) (the last body: its proxies are
    // still queued), then a new joint, through both worlds.
    Destroy(ref2, JointID(3));
    Destroy(gpu2, JointID(3));
    Destroy(ref2, BodyID(8));
    Destroy(gpu2, BodyID(8));
    Expect(GetBodies(ref2) == GetBodies(gpu2), "bodies after destroying");
    Expect(GetJoints(ref2) == GetJoints(gpu2), "joints after destroying");
    CompareBodies(ref2, gpu2, "chain after destroying");
    // SetType: a link made static and then dynamic again (mass data reset).
    SetType(ref2, BodyID(4), BodyType::Static);
    SetType(gpu2, BodyID(4), BodyType::Static);
    CompareBodies(ref2, gpu2, "chain link made static");
    SetType(ref2, BodyID(4), BodyType::Dynamic);
    SetType(gpu2, BodyID(4), BodyType::Dynamic);
    CompareBodies(ref2, gpu2, "chain link made dynamic");
    const auto jc = Joint{RevoluteJointConf{}.UseBodyA(BodyID(2)).UseBodyB(BodyID(5))};
    Expect(CreateJoint(ref2, jc) == CreateJoint(gpu2, jc), "joint id after destroying");
    Expect(GetJoints(ref2) == GetJoints(gpu2), "joints after creating");
    // Chain and multi shapes: a ground line and an L-shaped dynamic body.
    auto ref3 = World{};
    auto gpu3 = World{playrho_amd::GpuWorld{}};
    for (World* wp : {&ref3, &gpu3}) {
        auto line = ChainShapeConf{};
        line.Set({Length2{-8_m, 2_m}, Length2{-4_m, 0_m}, Length2{0_m, 1_m}, Length2{4_m, 0.5_m}, Length2{8_m, 3_m}});
        const auto ground = CreateShape(*wp, line);
        CreateBody(*wp, BodyConf{}.Use(ground));
        auto ell = MultiShapeConf{};
        VertexSet a, b;
        for (const auto& p : {Length2{-1_m, -0.2_m}, Length2{1_m, -0.2_m}, Length2{1_m, 0.2_m}, Length2{-1_m, 0.2_m}}) a.add(p);
        for (const auto& p : {Length2{-1_m, 0.2_m}, Length2{-0.6_m, 0.2_m}, Length2{-0.6_m, 1_m}, Length2{-1_m, 1_m}}) b.add(p);
        ell.AddConvexHull(a).AddConvexHull(b);
        ell.UseDensity(1_kgpm2);
        const auto body = CreateShape(*wp, ell);
        CreateBody(*wp, BodyConf{}.Use(BodyType::Dynamic).Use(body).UseLocation(Length2{0_m, 6_m}));
    }
    CompareBodies(ref3, gpu3, "chain and multi shapes");
    Expect(GetShapeRange(ref3) == GetShapeRange(gpu3), "compound shape count");
    for (auto s = 0u; s < GetShapeRange(ref3); ++s)
        Expect(GetShape(ref3, ShapeID(s)) == GetShape(gpu3, ShapeID(s)), "compound shape " + std::to_string(s));
    // Gear joints made by GetGearJointConf through each world's queries.
    auto ref4 = World{};
    auto gpu4 = World{playrho_amd::GpuWorld{}};
    const auto refGears = BuildGears(ref4);
    const auto gpuGears = BuildGears(gpu4);
    Expect(refGears == gpuGears, "gear joint ids");
    for (const auto id : refGears) {
        Expect(GetJoint(ref4, id) == GetJoint(gpu4, id), "gear joint " + std::to_string(to_underlying(id)));
        Expect(GetType(gpu4, id) == GetTypeID<GearJointConf>(), "gear joint type");
    }
    CompareBodies(ref4, gpu4, "gear world");
    std::printf("cpu: %s\n", g_fail ? "FAIL" : "OK");
    return g_fail ? 1 : 0;
}

int Hello() {
    // HelloWorld/HelloWorld.cpp's program, with the world on the device.
    constexpr auto boxX = +0_m;
    constexpr auto boxY = -10_m;
    constexpr auto boxWidth = 50_m;
    constexpr auto bodyHeight = 10_m;
    constexpr auto ballStart = Length2{0_m, 4_m};
    auto world = World{playrho_amd::GpuWorld{}};
    const auto box = CreateShape(world, PolygonShapeConf{}.SetAsBox(boxWidth, bodyHeight));
    CreateBody(world, BodyConf{}.Use(box).UseLocation(Length2{boxX, boxY}));
    const auto diskShape = CreateShape(world, DiskShapeConf{}.UseRadius(1_m));
    const auto ball = CreateBody(world, BodyConf{}
                                            .Use(BodyType::Dynamic)
                                            .Use(diskShape)
                                            .UseLocation(ballStart)
                                            .UseLinearAcceleration(EarthlyGravity));
    std::cout << std::fixed << std::setprecision(2);
    for (auto i = 0; i < 60; ++i) {
        Step(world);
        const auto transformation = GetTransformation(world, ball);
        const auto location = GetLocation(transformation);
        const auto angle = GetAngle(transformation);
        std::cout << std::setw(5) << GetX(location);
        std::cout << std::setw(5) << GetY(location);
        std::cout << std::setw(5) << angle;
        std::cout << '\n';
    }
    return 0;
}

int Gpu() {
    auto ref = World{};
    auto gpu = World{playrho_amd::GpuWorld{}};
    BuildPyramid(ref, 20);
    BuildPyramid(gpu, 20);
    std::vector<std::string> refCalls, gpuCalls;
    auto listen = [](World& w, std::vector<std::string>& out) {
        SetBeginContactListener(w, [&out](ContactID c) { out.push_back("b" + std::to_string(to_underlying(c))); });
        SetEndContactListener(w, [&out](ContactID c) { out.push_back("e" + std::to_string(to_underlying(c))); });
    };
    listen(ref, refCalls);
    listen(gpu, gpuCalls);
    auto conf = StepConf{};
    conf.doToi = false;
    for (int s = 1; s <= 150; ++s) {
        if (s == 40) {
            for (auto* w : {&ref, &gpu}) SetVelocity(*w, BodyID(200), Velocity{LinearVelocity2{3_mps, 1_mps}, 2_rpm});
        }
        if (s == 80) {
            for (auto* w : {&ref, &gpu})
                SetTransformation(*w, BodyID(150), Transformation{Length2{2_m, 30_m}, UnitVec::Get(0.5_rad)});
        }
        const auto sa = Step(ref, conf);
        const auto sb = Step(gpu, conf);
        const std::string when = "step " + std::to_string(s);
        Expect(sa.pre.contactsAdded == sb.pre.contactsAdded && sa.reg.islandsFound == sb.reg.islandsFound &&
                   sa.reg.sumVelIters == sb.reg.sumVelIters && Same(sa.reg.minSeparation, sb.reg.minSeparation),
               when + " stats");
        CompareBodies(ref, gpu, when);
        CompareContacts(ref, gpu, when);
        Expect(refCalls == gpuCalls, when + " listener calls");
    }
    // Geared bodies stepped on the device against the reference.
    auto ref4 = World{};
    auto gpu4 = World{playrho_amd::GpuWorld{}};
    BuildGears(ref4);
    BuildGears(gpu4);
    for (int s = 1; s <= 120; ++s) {
        Step(ref4, conf);
        Step(gpu4, conf);
        CompareBodies(ref4, gpu4, "gear world step " + std::to_string(s));
    }
    std::printf("gpu: %s (%zu listener calls)\n", g_fail ? "FAIL" : "OK", refCalls.size());
    return g_fail ? 1 : 0;
}

}  // namespace

int main(int argc, char** argv) {
    const std::string mode = argc > 1 ? argv[1] : "cpu";
    try {
        if (mode == "cpu") return Cpu();
        if (mode == "hello") return Hello();
        if (mode == "gpu") return Gpu();
    } catch (const std::exception& e) {
        std::fprintf(stderr, "exception: %s\n", e.what());
        return 3;
    }
    std::fprintf(stderr, "usage: %s cpu|hello|gpu\n", argv[0]);
    return 2;
}
