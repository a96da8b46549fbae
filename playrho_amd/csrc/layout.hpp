// Device-resident state of a batch of worlds (one world = one batch entry).
//
// Layout principles (MI355X): structure-of-arrays per record kind, 16-byte
// vectors where a record field group is always read together (one dwordx4
// load), world-contiguous segments so that the workgroup stepping a world
// touches one contiguous range of every array, and all indices world-local
// (int32) so a world can be moved between batches without rewriting them.
//
// Per world w: bodies [body_off, body_off + nbodies), proxies
// [proxy_off, proxy_off + nproxies), contacts [contact_off, contact_off + cap)
// with the live ones first in the world's contact order (creation order, as
// PlayRho's m_contacts: stable erase, append at back).  Shapes and their
// vertices are batch-global.
#pragma once

#include <cstdint>

#ifdef __HIPCC__
#include <hip/hip_runtime.h>
#else
struct alignas(16) float4 {
    float x, y, z, w;
};
struct alignas(8) float2 {
    float x, y;
};
struct alignas(16) uint4 {
    unsigned x, y, z, w;
};
struct alignas(8) int2 {
    int x, y;
};
struct alignas(16) int4 {
    int x, y, z, w;
};
#endif

namespace prx {

// Body flags (same bits as the state record).
enum : std::uint32_t {
    kSpeedable = 1u, kAccelerable = 2u, kImpenetrable = 4u, kAwake = 8u, kAutoSleep = 16u,
    kEnabled = 32u, kFixedRotation = 64u,
    // Host edits: the body's proxies are synchronised in the next step's
    // pre-phase (m_bodiesForSync, AabbTreeWorld.cpp:2162-2173).
    kNeedsSync = 0x200u
};
// Contact flags (same bits as the state record).
enum : std::uint32_t {
    kTouching = 1u, kCEnabled = 2u, kSensor = 4u, kCImpenetrable = 8u, kNeedsUpdating = 16u
};
// Proxy flags.
enum : std::uint32_t { kProxyPending = 1u, kProxyMoved = 2u };
// World status bits.
enum : std::uint32_t {
    kStatusContactOverflow = 1u, kStatusScratchOverflow = 2u, kStatusEventOverflow = 4u, kStatusToiLimit = 8u
};

// Per-world header (AoS: read once per phase by the world's workgroup).
// Words of WorldHeader::stats: PreStepStats, RegStepStats and ToiStepStats,
// each in the reference's declaration order (StepStats.hpp:38-46, 69-82,
// 117-151); the real fields hold float bits.
enum StatIndex : int {
    kStPreProxiesCreated = 0, kStPreProxiesMoved, kStPreDestroyed, kStPreAdded, kStPreUpdated, kStPreSkipped,
    kStRegMinSeparation, kStRegMaxIncImpulse, kStRegIslandsFound, kStRegIslandsSolved, kStRegBodiesSlept,
    kStRegMaxIslandBodies, kStRegAdded, kStRegUpdated, kStRegSkipped, kStRegProxiesMoved, kStRegSumPosIters,
    kStRegSumVelIters, kStToiMinSeparation, kStToiMaxIncImpulse, kStToiIslandsFound, kStToiIslandsSolved,
    kStToiContactsFound, kStToiAtMaxSubSteps, kStToiUpdatedToi, kStToiUpdatedTouching, kStToiSkippedTouching,
    kStToiAdded, kStToiProxiesMoved, kStToiSumPosIters, kStToiSumVelIters, kStToiMaxDistIters, kStToiMaxToiIters,
    kStToiMaxRootIters, kStats
};

struct WorldHeader {
    std::int32_t body_off, nbodies, proxy_off, nproxies;
    std::int32_t contact_off, contact_cap, ncontacts, pool_size;
    std::int32_t nfree, status, joint_off, njoints;
    std::uint32_t stats[34];  // StatIndex
    // Contact listener calls of the last step: word segment [ev_off, ev_off +
    // ev_cap) of DeviceState::ev (ev_cap 0: not recorded), words and calls
    // written (words may exceed ev_cap: then the status has kStatusEventOverflow).
    std::int32_t ev_off, ev_cap, ev_words, ev_count;
};

struct StepParams {
    float dt, dtRatio, minStillTimeToSleep, linearSlop, angularSlop, regResolutionRate,
        regMinSeparation, regMinMomentum, velocityThreshold, maxTranslation, maxRotation,
        maxLinearCorrection, maxAngularCorrection, linearSleepTolerance, angularSleepTolerance,
        displaceMultiplier, aabbExtension, maxCirclesRatio;
    std::uint32_t velIters, posIters, warmStart, blockSolve;
    std::uint32_t ldsBytes;  // dynamic LDS per workgroup (island solve state + record stage)
    std::uint32_t coopMin;   // islands with at least this many contacts are swept by the whole wave
    // TOI phase (doToi).
    std::uint32_t doToi;
    std::uint32_t jointKinds;  // 1: some joint of the batch is not revolute (kind dispatch in the joint sweeps)
    float toiResolutionRate, toiMinSeparation, toiMinMomentum, targetDepth, tolerance;
    std::uint32_t toiVelIters, toiPosIters, maxToiRootIters, maxToiIters, maxDistanceIters, maxSubSteps;
};

// Record of one step's solve for the adjoint (one-point manifolds): per
// island-order constraint slot the velocity-constraint record as built plus
// its geometry inputs, the inputs of every Gauss-Seidel update and of every
// position-solver update, and the island table.
constexpr int kTapeVc = 24;
constexpr int kJcStride = 10;
constexpr int kTapeJc = kJcStride + 2;
struct Tape {
    float4* vc;           // slot * kTapeVc: VelocityConstraint record [0..8], body positions [9],
                          // (vn0, contact id, manifold type, radius A) [10], (radius B, shape A, shape B, -) [11],
                          // velocities of A [12] and B [13] at creation, final impulses (n0, t0, n1, t1) [14],
                          // positions of A [15] and B [16] at creation, position-constraint record [17..21],
                          // manifold impulses before the dtRatio scaling (n0, t0, n1, t1) [22],
                          // ContactFeatures of the manifold's points (bits) [23]
    float4* gs;           // (slot * itV + it) * 3: (vA, wA, point-0 tangent impulse), (vB, wB, point-0 normal
                          // impulse), impulses (n0, t0, n1, t1) before the update
    float4* ngs;          // (slot * itP + it) * 2: position A (w: separation), position B before the update
    int4* isl;            // per island: first slot, end slot, velocity iterations run, position iterations run
    int2* islj;           // per island: first and end joint slot
    int* nisl;            // per world
    std::uint8_t* inisl;  // per body: part of an island
    float4* jc;           // joint slot * kTapeJc: joint record at set-up [0..kJcStride-1], (angle A, angle B,
                          // maximum motor torque, previous limit state) [kJcStride], impulses before the
                          // set-up [kJcStride + 1]
    float4* jgs;          // (joint slot * itV + it) * 3: velocities of A, B and the impulses before the update
    float4* jngs;         // (joint slot * itP + it) * 2: positions of A and B before the update
    std::uint32_t* cmap;  // per world contact pool id: how the next step's manifold took this step's final
                          // impulses: bits 0-3 (4-7) the point point 0 (1) took them from (kNoSource: none),
                          // bits 8-11 the next step's point count; bit 16 (kCarryFresh): the contact was
                          // created before this step's solve (its impulses owe nothing to the step before)
    int itV, itP;
};
constexpr std::uint32_t kNoSource = 0xfu;
constexpr std::uint32_t kCarryFresh = 0x10000u;

// Adjoint state of a batch (one backward step per launch).
struct AdjState {
    float4* xbar;   // per body: d loss / d (x, y, angle)
    float4* vbar;   // per body: d loss / d (vx, vy, w)
    float4* lambar; // per world contact pool id (segment of contact capacity): d loss / d (normal, tangent)
                    // impulse at the start of the later step, and those impulses' values
    float2* lamtmp; // per island-order slot
    float4* cbar;   // per slot, 3 float4: accumulated adjoints of the constraint's constants
    float* fbar;    // per shape: d loss / d friction
    float* ebar;    // per shape: d loss / d restitution
    int* status;    // 1: a record the adjoint does not model (two-point manifold, face B, joints)
    float* diag;    // counters of the last run: slots reversed, of them with a restitution bias,
                    // position-solver updates reversed, velocity updates reversed
    // General form (polygons, two-point manifolds, joints):
    float* fld;     // per slot kAdjFields: adjoints of the velocity-constraint values [0..26] and of
                    // the manifold's local points (point 0 x, y, point 1 x, y) [28..31]
    float4* lam4;   // per slot: adjoint of the running impulses (n0, t0, n1, t1)
    float4* lamc;   // per contact pool id, 2 float4: adjoint of the raw impulses at the later step's start
                    // and those impulses' values
    float* jfld;    // per joint slot kAdjJointFields: adjoints of the joint record's values
    float4* jlam;   // per joint slot: adjoint of the running impulses (x, y, z, motor)
    float4* jcarry; // per joint: adjoint of the impulses the later step started from
    float* tbar;    // per joint: d loss / d maximum motor torque
};
constexpr int kAdjFields = 32;
constexpr int kAdjJointFields = 16;

struct DeviceState {
    int nworlds;
    WorldHeader* hdr;
    const int* schedule;  // launch order of worlds (largest first)

    // Bodies.
    std::uint32_t* bflags;
    float4* pos0;    // x, y, angle, -
    float4* pos1;    // x, y, angle, -
    float4* vel;     // vx, vy, w, underActiveTime
    float4* xf;      // transform of pos1: px, py, cos, sin
    float4* lcinv;   // localCenter x, y, invMass, invRotI
    float4* acc;     // linear accel x, y, angular accel, -
    float2* damp;    // linear, angular damping
    int2* bproxy;    // first entry in bproxy_list, count (world-local)
    int* bproxy_list;

    // Proxies (broadphase leaves).
    int* pbody;          // world-local body
    int* pshape;         // batch-global shape
    std::uint32_t* pid;  // tree leaf id (contact keys use it)
    float4* paabb;       // fat AABB: lo x, lo y, hi x, hi y
    std::uint32_t* pflags;

    // Shapes.
    int4* shape_i;       // nverts, vertex offset, sensor, group
    float4* shape_f;     // vertex radius, friction, restitution, density
    std::uint32_t* shape_catmask;
    float2* verts;
    float2* norms;

    // Contacts (world order).  Keys are (pid[pa], pid[pb]) with pa < pb by id.
    int2* c_prox;        // world-local proxy of key min, key max
    int2* c_body;        // body A, body B (world-local)
    int2* c_shape;       // shape A, shape B (batch-global)
    std::uint32_t* c_id; // contact pool id
    std::uint32_t* c_flags;
    float4* c_mat;       // friction, restitution, tangent speed, -
    uint4* m_hdr;        // type | count << 8, cf0, cf1, -
    float4* m_geo;       // local normal x, y, local point x, y
    float4* m_p0;        // point 0: local x, y, normal impulse, tangent impulse
    float4* m_p1;
    std::uint32_t* c_free;  // pool free list (LIFO), per world segment of cap

    // Revolute joints (world order = creation order).
    int2* j_body;        // body A, body B (world-local)
    float4* j_anchor;    // local anchor A x, y, local anchor B x, y
    float4* j_cfg;       // revolute: reference angle, lower angle, upper angle, motor speed;
                         // distance: length, frequency, damping ratio, -; weld: reference
                         // angle, frequency, damping ratio, -; rope: maximum length, -, -, -;
                         // friction: maximum force, maximum torque, -, -; motor: angular
                         // offset, maximum force, maximum torque, correction factor (the
                         // linear offset in j_anchor.xy); wheel: frequency, damping ratio,
                         // -, motor speed; pulley: ground anchors A, B (j_cfg2: ratio, flags,
                         // lengths A, B)
    float4* j_cfg2;      // max motor torque, flags (limit 1, motor 2, collideConnected 4,
                         // kind << 8), wheel: local x axis of A (x, y)
    float4* j_gear;      // gear joints, 2 per joint: (ratio, constant, body C, body D (world-local int
                         // bits)), (local x axis of the A-C side, of the B-D side); the sides' anchors in
                         // j_anchor (A-C) and j_cfg (B-D), a revolute side's reference angle in .x
    float4* j_imp;       // impulse x, y, z, angular motor impulse
    std::uint32_t* j_limit;  // limit state (0 inactive, 1 lower, 2 upper, 3 equal)
    int2* bjoint;        // per body: first entry in bjoint_list, count (world-local)
    int2* bjoint_list;   // per body, in joint creation order: joint, other body (2 * njoints per world)

    // Scratch (per world, same segments as above).
    std::uint32_t* s_keys;   // candidate pair keys (cap)
    std::uint32_t* s_keys2;  // sorted candidate keys (cap)
    int* s_csr_off;          // per-body offsets (nbodies + 1 per world: body_off + w)
    int* s_csr;              // per-body touching-contact lists: (contact, other body) pairs (4 * cap ints)
    int* s_stack;            // DFS stack / adjacency cursors, then island of each body (2 * nbodies)
    std::uint8_t* s_isl_b;   // islanded bodies (nbodies)
    std::uint8_t* s_isl_c;   // islanded contacts (cap)
    int* s_isl_body;         // island body lists (nbodies + cap)
    int* s_isl_contact;      // island contact lists (cap)
    int2* s_isl_start;       // island: first body entry, first contact entry (nbodies + 1)
    float4* s_bc_pos;        // body constraint position x, y, a (nbodies)
    float4* s_bc_vel;        // body constraint velocity x, y, w (nbodies)
    float4* s_vc;            // velocity constraints in island order, kVcStride float4 each
    float4* s_pc;            // position constraints in island order, kPcStride float4 each
    float4* s_vc2;           // level form: velocity constraints of an island in level order
    float4* s_pc2;           // level form: position constraints of an island in level order
    int* s_lvl;              // level form: kLevelInts per slot (levels, order, level starts, rounds)
    float4* s_jc;            // joint constraints in island order, kJcStride float4 each (njoints)
    int* s_isl_joint;        // island joint lists (njoints)
    int* s_isl_jstart;       // island i: first joint entry (nbodies + 1, offset body_off + w)
    std::uint8_t* s_isl_j;   // islanded joints (njoints)

    // Optional per-world phase timings of the last step (shader clock cycles),
    // kPhases entries per world; null when not collected.
    unsigned long long* phase_cycles;
    // Contact listener calls of the last step, per world segment (see
    // WorldHeader::ev_off); null when no world records them.  Each call is
    // u32 kind (kEvBegin, kEvEnd, kEvPreSolve, kEvPostSolve), u32 contact id,
    // then for pre-solve the contact's manifold before the update (u32 type |
    // count << 8, two contact features, 12 floats: local normal, local point,
    // per point local point, normal and tangent impulse; zero where the type
    // or count has none), for post-solve u32 point count, u32 solved position
    // iteration (255 when not solved) and 4 floats (normal impulses [2],
    // tangent impulses [2]).
    std::uint32_t* ev;

    // TOI phase: per body the sweep's alpha0 (0 between steps) and a TOI
    // island flag; per contact slot the TOI value, its state (bit 31 valid,
    // low bits the sub-step count) and a TOI island flag.
    float* alpha0;
    std::uint8_t* t_isl_b;
    float* c_toi;
    std::uint32_t* c_toiinfo;
    std::uint8_t* t_isl_c;
    // Solve record for the adjoint (only read by the recording kernel).
    Tape tape;
};

// Phases timed by phase_cycles.
enum Phase : int {
    kPhProxies = 0, kPhDestroy, kPhUpdate, kPhFindAdd, kPhAdjacency, kPhIslands, kPhConstraints,
    kPhVelocity, kPhPosition, kPhWriteback, kPhSync, kPhUpdate2, kPhFindAdd2, kPhTotal, kPhases = 16
};

enum : std::uint32_t { kEvBegin = 0u, kEvEnd = 1u, kEvPreSolve = 2u, kEvPostSolve = 3u };
constexpr int kEvPreWords = 17;
constexpr int kEvPostWords = 8;
// Most event words one contact slot can produce in a step: end on destroy (2),
// update or creation before the solve (2 + 17), post-solve (8), update or
// creation after the solve (2 + 17).
constexpr int kEvWordsPerSlot = 48;

constexpr int kVcStride = 9;
constexpr int kPcStride = 5;
enum : std::uint32_t { kJointLimit = 1u, kJointMotor = 2u, kJointCollideConnected = 4u };
// Pi as the float build's Real (RealConstants.hpp:57).
constexpr float kPiF = 3.14159265358979323846f;
// Joint kinds, in bits 8-15 of a joint's flags.
enum : std::uint32_t {
    kJointRevolute = 0u,
    kJointDistance = 1u,
    kJointWeld = 2u,
    kJointRope = 3u,
    kJointFriction = 4u,
    kJointMotorKind = 5u,  // MotorJointConf (kJointMotor is the motor-enable flag)
    kJointWheel = 6u,
    kJointPulley = 7u,
    kJointPrismatic = 8u,
    kJointTarget = 9u,
    kJointGear = 10u
};
// Gear joints: flag bits set when the A-C (B-D) side is a prismatic joint.
enum : std::uint32_t { kGearPrisAC = 8u, kGearPrisBD = 16u };
// Constraint records staged through LDS per chunk by the wave-wide sweeps;
// the stage sits at the end of the dynamic LDS.
constexpr int kStageRecords = 16;
// Level form of the wave-wide sweeps: scratch ints per constraint slot, and
// the most bodies a world may have (one int per body in the record stage).
constexpr int kLevelInts = 8;
constexpr int kStageFloat4 = kStageRecords * kVcStride;

}  // namespace prx
