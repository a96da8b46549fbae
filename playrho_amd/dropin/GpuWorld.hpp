// GpuWorld: a world model for PlayRho's type-erased playrho::d2::World over
// this engine's C ABI (include/playrho_amd.h), so that
//
//     auto world = playrho::d2::World{playrho_amd::GpuWorld{}};
//
// runs the reference's API (CreateShape, CreateBody, CreateJoint, Step,
// GetTransformation, SetVelocity, GetContacts, ...) with every Step on the
// GPU.  World(T&&) (d2/World.hpp:739-745) wraps any copyable T for which the
// free functions that detail::WorldModel<T> calls (detail/WorldModel.hpp,
// one per detail::WorldConcept virtual) are found by argument-dependent
// lookup; they are declared here, in T's namespace.
//
// What maps to the engine and what does not (the unsupported operations throw
// the reference's exception types, InvalidArgument or WrongState):
// * shapes: DiskShapeConf, PolygonShapeConf (up to 16 vertices) and
//   EdgeShapeConf; ChainShapeConf and MultiShapeConf throw InvalidArgument.
//   SetShape accepts friction and restitution changes only.
// * bodies: creation from a Body (shapes given at creation, as BodyConf::Use
//   does); SetBody accepts changes of the transformation (SetTransformation,
//   SetLocation, SetAngle), the velocity and the accelerations, and the mass
//   data that ResetMassData computes for the body's shapes; other changes
//   (type, shapes, flags, custom mass data) throw InvalidArgument.
// * joints: revolute, distance, weld, rope, friction, motor, wheel and pulley
//   confs; the others throw InvalidArgument.  GetJoint returns the conf the
//   joint was created with.
// * contacts: GetContacts, GetContact and GetManifold read the device state;
//   SetContact and SetManifold throw WrongState (the engine owns contacts).
// * destruction (Destroy of a body, joint or shape) throws WrongState.
// * GetTree: the engine has no DynamicTree (its broadphase runs on the
//   device).  GetTree returns a DynamicTree rebuilt from the world's proxies
//   (CreateLeaf in proxy order, so leaf ids equal the reference's proxy ids
//   and queries return the same leaves; the inner nodes differ).
// * sub-stepping (SetSubStepping(true)) makes Step throw WrongState.
// * contact listeners are called after each Step completes, in the
//   reference's call order (the engine records them on the device).
#pragma once

#include <functional>
#include <optional>
#include <tuple>
#include <utility>
#include <vector>

#include <playrho/d2/World.hpp>
#include <playrho/d2/DynamicTree.hpp>

#include "../../include/playrho_amd.h"

namespace playrho_amd {

class GpuWorld {
public:
    GpuWorld();
    explicit GpuWorld(const playrho::d2::WorldConf& conf);
    GpuWorld(const GpuWorld& other);
    GpuWorld(GpuWorld&& other) noexcept;
    GpuWorld& operator=(const GpuWorld& other);
    GpuWorld& operator=(GpuWorld&& other) noexcept;
    ~GpuWorld();

    prx_world* handle() const noexcept { return m_world; }

    // State kept on the host beside the engine's world (internal).
    struct Listeners {
        playrho::ShapeFunction shapeDestruction;
        playrho::BodyShapeFunction detach;
        playrho::JointFunction jointDestruction;
        playrho::ContactFunction beginContact, endContact;
        playrho::d2::ContactManifoldFunction preSolve;
        playrho::d2::ContactImpulsesFunction postSolve;
    };
    prx_world* m_world = nullptr;
    playrho::Interval<playrho::Positive<playrho::Length>> m_vertexRadius;
    std::vector<playrho::d2::Shape> m_shapes;
    // Engine shape ids: a chain or multi shape is one engine shape per child.
    std::vector<std::uint32_t> m_shapeFirst, m_shapeCount;
    std::vector<std::pair<std::uint32_t, std::uint32_t>> m_engineShape;  // engine id -> (ShapeID, child)
    std::vector<playrho::d2::Body> m_bodies;  // as created (type, flags, shapes, damping)
    std::vector<playrho::d2::Joint> m_joints;
    Listeners m_listeners;
    playrho::Frequency m_invDeltaTime{};
    bool m_subStepping = false;
    mutable playrho::d2::DynamicTree m_tree;

    void InstallListeners();
};

bool operator==(const GpuWorld& lhs, const GpuWorld& rhs) noexcept;
bool operator!=(const GpuWorld& lhs, const GpuWorld& rhs) noexcept;

void SetShapeDestructionListener(GpuWorld& world, playrho::ShapeFunction listener) noexcept;
void SetDetachListener(GpuWorld& world, playrho::BodyShapeFunction listener) noexcept;
void SetJointDestructionListener(GpuWorld& world, playrho::JointFunction listener) noexcept;
void SetBeginContactListener(GpuWorld& world, playrho::ContactFunction listener) noexcept;
void SetEndContactListener(GpuWorld& world, playrho::ContactFunction listener) noexcept;
void SetPreSolveContactListener(GpuWorld& world, playrho::d2::ContactManifoldFunction listener) noexcept;
void SetPostSolveContactListener(GpuWorld& world, playrho::d2::ContactImpulsesFunction listener) noexcept;

std::optional<playrho::pmr::StatsResource::Stats> GetResourceStats(const GpuWorld& world) noexcept;
void Clear(GpuWorld& world) noexcept;
playrho::StepStats Step(GpuWorld& world, const playrho::StepConf& conf);
bool IsStepComplete(const GpuWorld& world) noexcept;
bool GetSubStepping(const GpuWorld& world) noexcept;
void SetSubStepping(GpuWorld& world, bool flag) noexcept;
const playrho::d2::DynamicTree& GetTree(const GpuWorld& world) noexcept;
bool IsLocked(const GpuWorld& world) noexcept;
void ShiftOrigin(GpuWorld& world, const playrho::Length2& newOrigin);
playrho::Interval<playrho::Positive<playrho::Length>> GetVertexRadiusInterval(const GpuWorld& world) noexcept;
playrho::Frequency GetInvDeltaTime(const GpuWorld& world) noexcept;

playrho::BodyCounter GetBodyRange(const GpuWorld& world) noexcept;
std::vector<playrho::BodyID> GetBodies(const GpuWorld& world);
std::vector<playrho::BodyID> GetBodiesForProxies(const GpuWorld& world);
playrho::BodyID CreateBody(GpuWorld& world, const playrho::d2::Body& body);
playrho::d2::Body GetBody(const GpuWorld& world, playrho::BodyID id);
void SetBody(GpuWorld& world, playrho::BodyID id, const playrho::d2::Body& value);
void Destroy(GpuWorld& world, playrho::BodyID id);
std::vector<std::pair<playrho::BodyID, playrho::JointID>> GetJoints(const GpuWorld& world, playrho::BodyID id);
std::vector<std::tuple<playrho::ContactKey, playrho::ContactID>> GetContacts(const GpuWorld& world,
                                                                             playrho::BodyID id);
std::vector<playrho::ShapeID> GetShapes(const GpuWorld& world, playrho::BodyID id);

playrho::JointCounter GetJointRange(const GpuWorld& world) noexcept;
std::vector<playrho::JointID> GetJoints(const GpuWorld& world);
playrho::JointID CreateJoint(GpuWorld& world, const playrho::d2::Joint& def);
playrho::d2::Joint GetJoint(const GpuWorld& world, playrho::JointID id);
void SetJoint(GpuWorld& world, playrho::JointID id, const playrho::d2::Joint& def);
void Destroy(GpuWorld& world, playrho::JointID id);

playrho::ShapeCounter GetShapeRange(const GpuWorld& world) noexcept;
playrho::ShapeID CreateShape(GpuWorld& world, const playrho::d2::Shape& def);
playrho::d2::Shape GetShape(const GpuWorld& world, playrho::ShapeID id);
void SetShape(GpuWorld& world, playrho::ShapeID id, const playrho::d2::Shape& def);
void Destroy(GpuWorld& world, playrho::ShapeID id);

playrho::ContactCounter GetContactRange(const GpuWorld& world) noexcept;
std::vector<playrho::KeyedContactID> GetContacts(const GpuWorld& world);
playrho::Contact GetContact(const GpuWorld& world, playrho::ContactID id);
void SetContact(GpuWorld& world, playrho::ContactID id, const playrho::Contact& value);
playrho::d2::Manifold GetManifold(const GpuWorld& world, playrho::ContactID id);
void SetManifold(GpuWorld& world, playrho::ContactID id, const playrho::d2::Manifold& value);

}  // namespace playrho_amd
