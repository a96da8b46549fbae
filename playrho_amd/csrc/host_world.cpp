// Host side of the C ABI: the world model (shapes, bodies, proxies) built
// through the playrho::d2 construction calls, packing into the device layout,
// batch management and stepping.  Shape geometry and mass properties are
// computed here exactly as the reference computes them, in float:
//
//   PolygonShapeConf::SetAsBox / Set (weld + gift wrap)  PolygonShapeConf.cpp:45-134,
//                                                         DistanceProxy.cpp:42-105
//   forward normals                                       d2/Math.cpp:106-124
//   disk / edge / polygon mass data                       MassData.cpp:40-190
//   ComputeMassData over a body's shapes                  WorldShape.cpp:160-176
//   SetMassData                                           WorldBody.cpp:365-399
//   Body flags from BodyConf                              Body.cpp:32-120
//   tree leaf (proxy) ids of a growing tree               DynamicTree.cpp:456-537
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstddef>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <memory>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/playrho_amd.h"
#include "big_world.hpp"
#include "layout.hpp"
#include "scenes.hpp"
#include "sincosf.hpp"

extern "C" hipError_t prx_launch_step(const prx::DeviceState* d, const prx::StepParams* c, int first, int count,
                                      hipStream_t stream);
extern "C" hipError_t prx_big_launch_step(const prx::DeviceState* d, const prx::BigState* g,
                                          const prx::StepParams* c, hipStream_t s);
extern "C" hipError_t prx_big_launch_pairs(const prx::DeviceState* d, const prx::BigState* g, hipStream_t s);
extern "C" int prx_lbvh_pairs_noalloc_impl(const float* aabbs, const std::int32_t* bodies, const std::uint8_t* moved,
                                           int n, std::int32_t* pairs, int cap, int* count, void* stream,
                                           const char** err);
extern "C" hipError_t prx_big_launch_metrics(const prx::DeviceState* d, const prx::BigState* g, hipStream_t s);
extern "C" hipError_t prx_launch_step_tape(const prx::DeviceState* d, const prx::StepParams* c, int count,
                                           hipStream_t stream);
extern "C" hipError_t prx_launch_adjoint_step(const prx::DeviceState* d, const prx::Tape* t, const prx::AdjState* a,
                                              const prx::StepParams* c, int nworlds, hipStream_t s);
extern "C" hipError_t prx_launch_adjoint_general(const prx::DeviceState* d, const prx::Tape* t,
                                                 const prx::AdjState* a, const prx::StepParams* c, int nworlds,
                                                 hipStream_t s);

namespace {

thread_local std::string g_error;

int Fail(int code, const std::string& msg) {
    g_error = msg;
    return code;
}

#define PRX_HIP(call)                                                                          \
    do {                                                                                       \
        const hipError_t e_ = (call);                                                          \
        if (e_ != hipSuccess) return Fail(PRX_E_DEVICE, std::string(#call ": ") + hipGetErrorString(e_)); \
    } while (0)

struct V2 {
    float x = 0, y = 0;
};
inline V2 operator+(V2 a, V2 b) { return {a.x + b.x, a.y + b.y}; }
inline V2 operator-(V2 a, V2 b) { return {a.x - b.x, a.y - b.y}; }
inline V2 Mul(V2 a, float s) { return {a.x * s, a.y * s}; }
inline V2 Mul(float s, V2 a) { return {s * a.x, s * a.y}; }
inline V2 DivS(V2 a, float s) {
    const float inv = 1.0f / s;
    return {a.x * inv, a.y * inv};
}
inline float Cross(V2 a, V2 b) {
    const float m = a.x * b.y;
    const float s = a.y * b.x;
    return m - s;
}
inline float MagSq(V2 v) {
    float r = 0.0f;
    r += v.x * v.x;
    r += v.y * v.y;
    return r;
}
inline V2 Rotate(V2 v, V2 q) { return {(q.x * v.x) - (q.y * v.y), (q.y * v.x) + (q.x * v.y)}; }
inline V2 InvRotate(V2 v, V2 q) { return {(q.x * v.x) + (q.y * v.y), (q.x * v.y) - (q.y * v.x)}; }
inline float Dot(V2 a, V2 b) {
    float r = 0.0f;
    r += a.x * b.x;
    r += a.y * b.y;
    return r;
}
inline V2 UnitVector(V2 v, V2 fallback) {
    const float m2 = v.x * v.x + v.y * v.y;
    if (std::isnormal(m2)) {
        const float mag = std::sqrt(m2);
        const float inv = 1.0f / mag;
        return {v.x * inv, v.y * inv};
    }
    const float mag = float(std::sqrt(double(v.x) * v.x + double(v.y) * v.y));
    if (!std::isnormal(mag)) return fallback;
    return {v.x / mag, v.y / mag};
}
inline V2 UnitFromAngle(float a) { return {prx_math::CosF<true>(a), prx_math::SinF<true>(a)}; }

constexpr float kPi = 3.14159265358979323846f;
constexpr float kDefaultLinearSlop = 0.005f;
constexpr int kMaxShapeVertices = 16;

struct HostShape {
    std::vector<V2> v, n;
    float radius = 0.01f, density = 0, friction = 0.2f, restitution = 0;
    std::uint32_t sensor = 0, catmask = 0xffff0001u;
    std::int32_t group = 0;
    // A child of a chain (1) or multi (2) shape: the children are consecutive
    // shapes from `first` on, `count` of them, one proxy each.
    std::uint32_t compound = 0, first = 0, count = 1;
};

struct MassData {
    V2 center;
    float mass = 0, I = 0;
};

std::vector<V2> FwdNormals(const std::vector<V2>& v) {
    std::vector<V2> n;
    const std::size_t count = v.size();
    if (count > 1) {
        for (std::size_t i = 0; i < count; ++i) {
            const V2 e = v[(i + 1 < count) ? i + 1 : 0] - v[i];
            n.push_back(UnitVector(V2{e.y, -e.x}, V2{0, 0}));
        }
    } else if (count == 1) {
        n.push_back(V2{0, 0});
    }
    return n;
}

std::vector<V2> ConvexHull(const std::vector<V2>& pts) {
    std::vector<V2> out;
    const std::size_t nv = pts.size();
    if (nv == 0) return out;
    std::size_t i0 = 0;
    float maxx = pts[0].x;
    for (std::size_t i = 1; i < nv; ++i) {
        const float x = pts[i].x;
        if ((maxx < x) || ((maxx == x) && (pts[i].y < pts[i0].y))) {
            maxx = x;
            i0 = i;
        }
    }
    std::vector<std::size_t> hull;
    std::size_t ih = i0;
    for (;;) {
        hull.push_back(ih);
        std::size_t ie = 0;
        for (std::size_t j = 1; j < nv; ++j) {
            if (ie == ih) {
                ie = j;
                continue;
            }
            const V2 r = pts[ie] - pts[ih];
            const V2 v = pts[j] - pts[ih];
            const float c = Cross(r, v);
            if ((c < 0.0f) || ((c == 0.0f) && (MagSq(v) > MagSq(r)))) ie = j;
        }
        ih = ie;
        if (ie == i0) break;
        if (hull.size() > nv) break;
    }
    for (std::size_t k : hull) out.push_back(pts[k]);
    return out;
}

MassData DiskMass(float r, float density, V2 loc) {
    const float r2 = r * r;
    const float area = r2 * kPi;
    const float mass = density * area;
    const float Iz = area * ((r2 / 2.0f) + MagSq(loc));
    const float I = Iz * density / 1.0f;
    return MassData{loc, mass, I};
}

float PolarMoment(const V2* v, std::size_t count) {
    float sx = 0.0f, sy = 0.0f;
    for (std::size_t i = 0; i < count; ++i) {
        const V2 a = v[i];
        const V2 b = v[(i + 1 < count) ? i + 1 : 0];
        const float fb = Cross(a, b);
        sx += ((a.y * a.y) + a.y * b.y + (b.y * b.y)) * fb;
        sy += ((a.x * a.x) + a.x * b.x + (b.x * b.x)) * fb;
    }
    return (sx + sy) * (1.0f / 12.0f);
}

MassData EdgeMass(float r, float density, V2 v0, V2 v1) {
    const float r2 = r * r;
    const float circleArea = r2 * kPi;
    const float circleMass = density * circleArea;
    const V2 d = v1 - v0;
    const V2 u = UnitVector(d, V2{0, 0});
    const V2 offset = Mul(V2{-u.y, u.x}, r);
    const float b = std::sqrt(MagSq(d));
    const float h = r * 2.0f;
    const float rectMass = density * b * h;
    const float totalMass = circleMass + rectMass;
    const V2 center = DivS(v0 + v1, 2.0f);
    const float halfCircleArea = circleArea / 2.0f;
    const float halfR2 = r2 / 2.0f;
    const V2 verts[4] = {v0 + offset, v0 - offset, v1 - offset, v1 + offset};
    const float Iz = PolarMoment(verts, 4);
    const float I0 = halfCircleArea * (halfR2 + MagSq(v0));
    const float I1 = halfCircleArea * (halfR2 + MagSq(v1));
    const float I = (I0 + I1 + Iz) * density / 1.0f;
    return MassData{center, totalMass, I};
}

MassData PolygonMass(float r, float density, const std::vector<V2>& v) {
    const std::size_t count = v.size();
    if (count == 0) return MassData{};
    if (count == 1) return DiskMass(r, density, v[0]);
    if (count == 2) return EdgeMass(r, density, v[0], v[1]);
    V2 center{0, 0};
    float area = 0.0f, I = 0.0f;
    V2 sum{0, 0};
    for (const V2& p : v) sum = sum + p;
    const V2 s = DivS(sum, float(count));
    for (std::size_t i = 0; i < count; ++i) {
        const V2 e1 = v[i] - s;
        const V2 e2 = v[(i + 1 < count) ? i + 1 : 0] - s;
        const float D = Cross(e1, e2);
        const float triangleArea = D * (1.0f / 2.0f);
        area += triangleArea;
        center = center + Mul(Mul(triangleArea, e1 + e2), 1.0f / 3.0f);
        const float intx2 = (e1.x * e1.x) + e2.x * e1.x + (e2.x * e2.x);
        const float inty2 = (e1.y * e1.y) + e2.y * e1.y + (e2.y * e2.y);
        const float triangleI = D * (intx2 + inty2) * (1.0f / 12.0f);
        I += triangleI;
    }
    const float mass = density * area;
    center = ((area > 0.0f) && !(std::abs(area) < 1.17549435e-38f)) ? DivS(center, area) : V2{0, 0};
    const V2 mcenter = center + s;
    const float lever = MagSq(mcenter) - MagSq(center);
    const float massI = ((density * I) + (mass * lever)) / 1.0f;
    return MassData{mcenter, mass, massI};
}

MassData ShapeMass(const HostShape& s) { return PolygonMass(s.radius, s.density, s.v); }

struct HostBody {
    std::uint32_t type = PRX_STATIC, flags = 0;
    float pos0[3] = {0, 0, 0}, pos1[3] = {0, 0, 0};
    V2 lc;
    float vel[3] = {0, 0, 0};
    float under = 0;
    float acc[3] = {0, 0, 0};
    float damp[2] = {0, 0};
    float invMass = 0, invRotI = 0;
    std::vector<std::uint32_t> shapes;
    std::vector<int> proxies;  // world-local proxy indices
    bool destroyed = false;    // Destroy(World&, BodyID): state frozen, out of every list
};

struct HostProxy {
    int body = 0, shape = 0;
    std::uint32_t id = 0, flags = prx::kProxyPending;
    float aabb[4] = {0, 0, 0, 0};
    bool dead = false;  // its body was destroyed (uploaded with an empty box, never moved)
};

struct HostJoint {
    // prx::kJointLimit / kJointMotor / kJointCollideConnected, the kind
    // (prx::kJointRevolute / kJointDistance / kJointWeld) in bits 8-15.
    std::uint32_t a = 0, b = 0, flags = 0;
    float la[4] = {0, 0, 0, 0};             // local anchor A x, y, B x, y
    // Revolute: reference angle, lower, upper, motor speed.  Distance: length,
    // frequency, damping ratio, 0.  Weld: reference angle, frequency, damping
    // ratio, 0.  Rope: maximum length, 0, 0, 0.  Friction: maximum force,
    // maximum torque, 0, 0.  Motor: angular offset, maximum force, maximum
    // torque, correction factor (linear offset in la[0..1]).  Wheel: frequency,
    // damping ratio, 0, motor speed (maximum motor torque in maxTorque, axis in
    // ext).  Pulley: ground anchors A, B (ratio in maxTorque, lengths in ext).
    float cfg[4] = {0, 0, 0, 0};
    float maxTorque = 0;
    float ext[3] = {0, 0, 0};               // wheel: local x axis of A (x, y), -
    float imp[4] = {0, 0, 0, 0};            // impulse x, y, z, angular motor impulse
    // Gear (kind 10, prismatic sides in flag bits kGearPrisAC / kGearPrisBD):
    // bodies C, D, the ratio, the constant, and per side (A-C in gear[0..5],
    // B-D in gear[6..11]) the prismatic joint's local anchors A, B and local
    // x axis or the revolute joint's reference angle (first float).
    std::uint32_t c = 0, d = 0;
    float ratio = 1, constant = 0;
    float gear[12] = {};
    std::uint32_t limit = 0;
    std::uint64_t seq = 0;   // creation order (the bodies' joint lists follow it)
    bool destroyed = false;  // Destroy(World&, JointID)
};

// Contact records as held on the host between device residencies.
struct HostContacts {
    std::vector<int2> prox, body, shape;
    std::vector<std::uint32_t> id, flags;
    std::vector<float4> mat, geo, p0, p1;
    std::vector<uint4> hdr;
    std::vector<std::uint32_t> freelist;
    int pool = 0;
    std::size_t size() const { return id.size(); }
    void resize(std::size_t n) {
        prox.resize(n);
        body.resize(n);
        shape.resize(n);
        id.resize(n);
        flags.resize(n);
        mat.resize(n);
        geo.resize(n);
        p0.resize(n);
        p1.resize(n);
        hdr.resize(n);
    }
};

}  // namespace

struct prx_batch;

namespace {
prx_step_stats DefaultStats() {
    prx_step_stats s{};
    s.reg_min_separation = std::numeric_limits<float>::infinity();
    s.toi_min_separation = std::numeric_limits<float>::infinity();
    return s;
}
}  // namespace

struct prx_world {
    prx_world_conf conf{};
    std::vector<HostShape> shapes;
    std::vector<HostBody> bodies;
    std::vector<HostProxy> proxies;
    std::vector<HostJoint> joints;
    HostContacts contacts;
    std::uint32_t leafCount = 0;
    prx_batch* owner = nullptr;  // batch holding the authoritative device state
    int index_in_owner = -1;
    bool host_stale = false;     // device state newer than this mirror
    prx_batch* solo = nullptr;   // single-world batch used by prx_step
    prx_step_stats last_stats = DefaultStats();  // StepStats{}
    bool record_events = false;  // the device records the contact listener calls of each step
    std::uint32_t event_words = 0;  // initial words of the call record (0: kEvWordsPerSlot per contact slot)
    bool has_listener = false;
    prx_contact_listener listener{};
    std::vector<std::uint32_t> events;  // listener calls of the last step (device word format)
    std::uint32_t nevents = 0;
    // Listener calls of edits since the last step (a destroyed body's touching
    // contacts end): they come before the next step's calls.
    std::vector<std::uint32_t> pre_events;
    std::uint32_t pre_nevents = 0;
    std::vector<std::uint32_t> joint_free;  // destroyed joint ids, reused last first (ObjectPool)
    std::uint64_t joint_seq = 0;
    std::uint32_t destroyed_bodies = 0;
};

namespace {

// Device arrays of one batch.
struct DeviceArena {
    std::vector<void*> ptrs;
    ~DeviceArena() {
        for (void* p : ptrs) (void)hipFree(p);
    }
    template <class T>
    hipError_t alloc(T*& out, std::size_t n) {
        void* p = nullptr;
        const hipError_t e = hipMalloc(&p, std::max<std::size_t>(n, 1) * sizeof(T));
        if (e == hipSuccess) ptrs.push_back(p);
        out = static_cast<T*>(p);
        return e;
    }
};

}  // namespace

struct prx_batch {
    std::vector<prx_world*> worlds;
    prx_batch_conf conf{};
    DeviceArena arena;
    prx::DeviceState d{};
    std::vector<prx::WorldHeader> hdr;
    std::vector<int> shape_base, vert_base;
    int* d_schedule = nullptr;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    float last_ms = 0;
    int last_launches = 0;
    hipGraphExec_t graph_exec = nullptr;
    hipGraph_t graph = nullptr;
    prx::StepParams graph_params{};
    bool graph_valid = false;
    bool joint_kinds = false;  // some joint is not revolute
    std::size_t totB = 0, totP = 0, totS = 0, totV = 0, totC = 0, totJ = 0, totE = 0;
    int lds_bytes = 16 * 1024;
    int lds_min = 16 * 1024;
    std::vector<void*> snapshot;  // state before a recorded step (listener record growth)
    ~prx_batch() {
        if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

// The device's StepStats words in the C ABI's struct (same order, same size).
static_assert(sizeof(prx_step_stats) == sizeof(std::uint32_t) * prx::kStats, "StepStats layout");
prx_step_stats StatsOf(const prx::WorldHeader& h) {
    prx_step_stats s;
    std::memcpy(&s, h.stats, sizeof(s));
    return s;
}

int CheckWorld(prx_world* w) {
    if (!w) return Fail(PRX_E_INVALID_ARGUMENT, "null world");
    return PRX_OK;
}

// Copies the batch's device state for every member world back to the host.
int DownloadBatch(prx_batch* b);

// Brings a world's host mirror up to date (device -> host).
int Refresh(prx_world* w) {
    if (w->owner && w->host_stale) return DownloadBatch(w->owner);
    return PRX_OK;
}

// Leaves the current batch (device state copied back first).
int Detach(prx_world* w) {
    if (!w->owner) return PRX_OK;
    int st = Refresh(w);
    if (st != PRX_OK) return st;
    prx_batch* b = w->owner;
    for (auto& m : b->worlds)
        if (m == w) m = nullptr;
    w->owner = nullptr;
    w->index_in_owner = -1;
    if (w->solo == b) {
        w->solo = nullptr;
        delete b;
    }
    return PRX_OK;
}

int AddShape(prx_world* w, const prx_shape_conf& c, std::uint32_t* out) {
    HostShape s;
    const float defaultRadius = kDefaultLinearSlop * 2.0f;
    switch (c.kind) {
    case PRX_SHAPE_DISK: {
        s.radius = (c.vertex_radius >= 0) ? c.vertex_radius : c.radius;
        s.v = {V2{c.location[0], c.location[1]}};
        break;
    }
    case PRX_SHAPE_BOX: {
        s.radius = (c.vertex_radius >= 0) ? c.vertex_radius : defaultRadius;
        const float hx = c.hx, hy = c.hy;
        s.v = {V2{+hx, -hy}, V2{+hx, +hy}, V2{-hx, +hy}, V2{-hx, -hy}};
        break;
    }
    case PRX_SHAPE_POLYGON: {
        s.radius = (c.vertex_radius >= 0) ? c.vertex_radius : defaultRadius;
        if (!c.points && c.npoints) return Fail(PRX_E_INVALID_ARGUMENT, "polygon points missing");
        // VertexSet weld at Square(DefaultLinearSlop) (VertexSet.hpp:67-108).
        const float minSepSq = kDefaultLinearSlop * kDefaultLinearSlop;
        std::vector<V2> set;
        for (std::uint32_t i = 0; i < c.npoints; ++i) {
            const V2 p{c.points[2 * i], c.points[2 * i + 1]};
            bool dup = false;
            for (const V2& e : set) {
                if (MagSq(p - e) <= minSepSq) {
                    dup = true;
                    break;
                }
            }
            if (!dup) set.push_back(p);
        }
        s.v = ConvexHull(set);
        break;
    }
    case PRX_SHAPE_EDGE: {
        s.radius = (c.vertex_radius >= 0) ? c.vertex_radius : defaultRadius;
        if (!c.points || c.npoints < 2) return Fail(PRX_E_INVALID_ARGUMENT, "edge needs two points");
        s.v = {V2{c.points[0], c.points[1]}, V2{c.points[2], c.points[3]}};
        break;
    }
    default:
        return Fail(PRX_E_INVALID_ARGUMENT, "unknown shape kind");
    }
    if (s.v.empty()) return Fail(PRX_E_INVALID_ARGUMENT, "shape has no vertices");
    if (int(s.v.size()) > kMaxShapeVertices) return Fail(PRX_E_UNSUPPORTED, "more than 16 vertices");
    // Vertex radius bounds (AabbTreeWorld.cpp:1243-1248).
    if (!(s.radius >= w->conf.min_vertex_radius) || !(s.radius <= w->conf.max_vertex_radius))
        return Fail(PRX_E_INVALID_ARGUMENT, "CreateShape: vertex radius out of range");
    s.n = FwdNormals(s.v);
    s.density = c.density;
    s.friction = c.friction;
    s.restitution = c.restitution;
    s.sensor = c.sensor ? 1u : 0u;
    s.catmask = std::uint32_t(c.category_bits) | (std::uint32_t(c.mask_bits) << 16);
    s.group = c.group_index;
    w->shapes.push_back(std::move(s));
    if (out) *out = std::uint32_t(w->shapes.size() - 1);
    return PRX_OK;
}

// A chain or multi shape (ChainShapeConf, MultiShapeConf) as its children,
// consecutive shapes sharing the conf's material and filter: a chain of n > 1
// vertices has n - 1 segments (ChainShapeConf::GetChild, ChainShapeConf.cpp:
// 149-160; each a two-vertex proxy with normals n, -n where
// n = unit(fwdperp(v[i+1] - v[i])), ChainShapeConf.cpp:38-59), of one vertex a
// single point; a multi shape has one convex hull per child
// (MultiShapeConf::AddConvexHull, MultiShapeConf.cpp:57-62; ConvexHull::Get,
// ConvexHull.cpp:30-36: points welded at the VertexSet default, then the hull
// and its forward normals).  hull_sizes: points per hull (multi shapes).
int AddCompound(prx_world* w, const prx_shape_conf& c, const std::uint32_t* hull_sizes, std::uint32_t nhulls,
                std::uint32_t* first_out, std::uint32_t* count_out) {
    if (c.kind != PRX_SHAPE_CHAIN && c.kind != PRX_SHAPE_MULTI) return Fail(PRX_E_INVALID_ARGUMENT, "not a chain or multi shape");
    if (!c.points && c.npoints) return Fail(PRX_E_INVALID_ARGUMENT, "points missing");
    const float radius = (c.vertex_radius >= 0) ? c.vertex_radius : kDefaultLinearSlop * 2.0f;
    if (!(radius >= w->conf.min_vertex_radius) || !(radius <= w->conf.max_vertex_radius))
        return Fail(PRX_E_INVALID_ARGUMENT, "CreateShape: vertex radius out of range");
    std::vector<std::vector<V2>> children;
    if (c.kind == PRX_SHAPE_CHAIN) {
        if (c.npoints == 0) return Fail(PRX_E_INVALID_ARGUMENT, "a chain needs a vertex");
        if (c.npoints == 1) children.push_back({V2{c.points[0], c.points[1]}});
        for (std::uint32_t i = 0; i + 1 < c.npoints; ++i)
            children.push_back({V2{c.points[2 * i], c.points[2 * i + 1]}, V2{c.points[2 * i + 2], c.points[2 * i + 3]}});
    } else {
        if (!hull_sizes || nhulls == 0) return Fail(PRX_E_INVALID_ARGUMENT, "a multi shape needs hulls");
        const float minSepSq = std::sqrt(std::numeric_limits<float>::min());
        std::uint32_t at = 0;
        for (std::uint32_t h = 0; h < nhulls; ++h) {
            if (at + hull_sizes[h] > c.npoints) return Fail(PRX_E_INVALID_ARGUMENT, "hull sizes exceed the points");
            std::vector<V2> set;
            for (std::uint32_t k = 0; k < hull_sizes[h]; ++k, ++at) {
                const V2 p{c.points[2 * at], c.points[2 * at + 1]};
                bool dup = false;
                for (const V2& e : set) {
                    if (MagSq(p - e) <= minSepSq) {
                        dup = true;
                        break;
                    }
                }
                if (!dup) set.push_back(p);
            }
            std::vector<V2> hull = ConvexHull(set);
            if (hull.empty()) return Fail(PRX_E_INVALID_ARGUMENT, "empty hull");
            if (int(hull.size()) > kMaxShapeVertices) return Fail(PRX_E_UNSUPPORTED, "more than 16 vertices");
            children.push_back(std::move(hull));
        }
    }
    const std::uint32_t first = std::uint32_t(w->shapes.size());
    for (auto& v : children) {
        HostShape s;
        s.radius = radius;
        s.n = (c.kind == PRX_SHAPE_CHAIN && v.size() == 2) ? std::vector<V2>{} : FwdNormals(v);
        if (c.kind == PRX_SHAPE_CHAIN && v.size() == 2) {
            const V2 e = v[1] - v[0];
            const V2 n = UnitVector(V2{e.y, -e.x}, V2{0, 0});
            s.n = {n, V2{-n.x, -n.y}};
        }
        s.v = std::move(v);
        s.density = c.density;
        s.friction = c.friction;
        s.restitution = c.restitution;
        s.sensor = c.sensor ? 1u : 0u;
        s.catmask = std::uint32_t(c.category_bits) | (std::uint32_t(c.mask_bits) << 16);
        s.group = c.group_index;
        s.compound = (c.kind == PRX_SHAPE_CHAIN) ? 1u : 2u;
        s.first = first;
        s.count = std::uint32_t(children.size());
        w->shapes.push_back(std::move(s));
    }
    if (first_out) *first_out = first;
    if (count_out) *count_out = std::uint32_t(children.size());
    return PRX_OK;
}

// Mass data of a whole chain or multi shape from its children: a chain
// (ChainShapeConf::GetMassData, ChainShapeConf.cpp:113-147) sums its
// segments' capsule masses and divides their weighted centres by the summed
// capsule areas; a multi shape (GetMassData(MultiShapeConf),
// MultiShapeConf.cpp:35-55) sums its hulls' masses, centres weighted by mass
// and inertias.
MassData CompoundMass(const prx_world* w, const HostShape& first) {
    const HostShape& c0 = w->shapes[first.first];
    if (c0.compound == 1) {
        if (c0.count == 1 && c0.v.size() == 1) return DiskMass(c0.radius, c0.density, c0.v[0]);
        float mass = 0, I = 0, area = 0;
        V2 center{0, 0};
        const float circleArea = (c0.radius * c0.radius) * kPi;
        for (std::uint32_t k = 0; k < c0.count; ++k) {
            const HostShape& e = w->shapes[first.first + k];
            const MassData md = EdgeMass(e.radius, e.density, e.v[0], e.v[1]);
            mass += md.mass;
            center = center + Mul(md.mass / 1.0f, md.center);
            I += md.I;
            const V2 d = e.v[1] - e.v[0];
            const float b = std::sqrt(MagSq(d));
            const float h = e.radius * 2.0f;
            area += b * h + circleArea;
        }
        center = V2{center.x / area, center.y / area};
        return MassData{center, mass, I};
    }
    float mass = 0, I = 0;
    V2 wc{0, 0};
    for (std::uint32_t k = 0; k < c0.count; ++k) {
        const HostShape& h = w->shapes[first.first + k];
        const MassData md = PolygonMass(h.radius, h.density, h.v);
        mass += md.mass;
        wc = wc + Mul(md.mass, md.center);
        I += md.I;
    }
    const V2 center = (mass > 0) ? V2{wc.x / mass, wc.y / mass} : V2{0, 0};
    return MassData{center, mass, I};
}

// ResetMassData (WorldBody.cpp:360-363): ComputeMassData over the body's
// shapes of positive density (WorldShape.cpp:160-176), then SetMassData
// (WorldBody.cpp:365-400): a body that is not accelerable gets no mass (a
// static one a sweep at its origin with local centre zero); an accelerable
// one its mass (1 kg if none), rotational inertia about the centre of mass
// (none with fixed rotation), the centre of mass moved there and the linear
// velocity changed by the angular one about the move.  Returns whether the
// transformation changed (SetBody then flags the contacts and syncs).
bool ResetMass(const prx_world* w, HostBody& b) {
    const V2 q = UnitFromAngle(b.pos1[2]);
    const V2 xfp = V2{b.pos1[0], b.pos1[1]} - Rotate(b.lc, q);
    if (!(b.flags & prx::kAccelerable)) {
        b.invMass = 0;
        b.invRotI = 0;
        if (b.flags & prx::kSpeedable) return false;
        b.pos0[0] = b.pos1[0] = xfp.x;
        b.pos0[1] = b.pos1[1] = xfp.y;
        b.pos0[2] = b.pos1[2];
        b.lc = V2{0, 0};
        const V2 p1 = V2{b.pos1[0], b.pos1[1]} - Rotate(b.lc, q);
        return !(p1.x == xfp.x && p1.y == xfp.y);
    }
    float mass = 0, I = 0;
    V2 wc{0, 0};
    for (std::uint32_t sid : b.shapes) {
        const HostShape& s = w->shapes[sid];
        // A chain or multi shape counts once, at its first child.
        if (s.compound && sid != s.first) continue;
        if (s.density > 0) {
            const MassData md = s.compound ? CompoundMass(w, s) : ShapeMass(s);
            mass += md.mass;
            wc = wc + Mul(md.mass / 1.0f, md.center);
            I += md.I;
        }
    }
    const V2 center = (mass > 0) ? DivS(wc, mass / 1.0f) : V2{0, 0};
    const float m = (mass > 0) ? mass : 1.0f;
    b.invMass = 1.0f / m;
    b.invRotI = 0.0f / 1.0f;
    if ((I > 0) && !(b.flags & prx::kFixedRotation)) {
        const float lsq = MagSq(center);
        const float Ic = I - (m * lsq) / 1.0f;
        b.invRotI = 1.0f / Ic;
    }
    // Move the center of mass (sweep) and update the velocity.
    const V2 oldCenter{b.pos1[0], b.pos1[1]};
    const V2 newc = Rotate(center, q) + xfp;
    b.pos0[0] = b.pos1[0] = newc.x;
    b.pos0[1] = b.pos1[1] = newc.y;
    b.pos0[2] = b.pos1[2];
    b.lc = center;
    const V2 dc = newc - oldCenter;
    const float wv = b.vel[2] / 1.0f;
    b.vel[0] += -dc.y * wv;
    b.vel[1] += dc.x * wv;
    const V2 p1 = newc - Rotate(b.lc, q);
    return !(p1.x == xfp.x && p1.y == xfp.y);
}

int AddBody(prx_world* w, const prx_body_conf& c, std::uint32_t* out) {
    // The reference would give the new body the id freed last while putting
    // it at the end of the body list, and its proxies the tree nodes freed
    // last; the device keeps bodies in id order.
    if (w->destroyed_bodies)
        return Fail(PRX_E_UNSUPPORTED, "creating a body in a world where a body was destroyed");
    if (w->bodies.size() >= 65534) return Fail(PRX_E_LENGTH, "CreateBody: operation would exceed MaxBodies");
    for (std::uint32_t i = 0; i < c.nshapes; ++i)
        if (c.shapes[i] >= w->shapes.size()) return Fail(PRX_E_OUT_OF_RANGE, "no such shape");
    HostBody b;
    b.type = c.type;
    // Body::GetFlags(BodyConf) (Body.cpp:32-87).
    std::uint32_t f = 0;
    if (c.type == PRX_DYNAMIC) f |= prx::kSpeedable | prx::kAccelerable;
    else if (c.type == PRX_KINEMATIC) f |= prx::kImpenetrable | prx::kSpeedable;
    else f |= prx::kImpenetrable;
    if (c.bullet) f |= prx::kImpenetrable;
    if (c.fixed_rotation) f |= prx::kFixedRotation;
    if (c.allow_sleep) f |= prx::kAutoSleep;
    if (c.awake) {
        if (f & prx::kSpeedable) f |= prx::kAwake;
    } else if (!c.allow_sleep && (f & prx::kSpeedable)) {
        f |= prx::kAwake;
    }
    if (c.enabled) f |= prx::kEnabled;
    b.flags = f;
    b.pos0[0] = b.pos1[0] = c.location[0];
    b.pos0[1] = b.pos1[1] = c.location[1];
    b.pos0[2] = b.pos1[2] = c.angle;
    b.damp[0] = c.linear_damping;
    b.damp[1] = c.angular_damping;
    // Velocity (Body::SetVelocity, Body.cpp:184-194): non-speedable bodies
    // keep zero velocity.
    const bool nonzeroVel = c.linear_velocity[0] != 0 || c.linear_velocity[1] != 0 || c.angular_velocity != 0;
    if (!nonzeroVel || (f & prx::kSpeedable)) {
        b.vel[0] = c.linear_velocity[0];
        b.vel[1] = c.linear_velocity[1];
        b.vel[2] = c.angular_velocity;
    }
    // Acceleration (Body::SetAcceleration, Body.cpp:203-234).
    const bool nonzeroAcc = c.linear_acceleration[0] != 0 || c.linear_acceleration[1] != 0 ||
                            c.angular_acceleration != 0;
    if (!nonzeroAcc || (f & prx::kAccelerable)) {
        b.acc[0] = c.linear_acceleration[0];
        b.acc[1] = c.linear_acceleration[1];
        b.acc[2] = c.angular_acceleration;
    }
    b.shapes.assign(c.shapes, c.shapes + c.nshapes);
    ResetMass(w, b);
    const int bid = int(w->bodies.size());
    if (f & prx::kEnabled) {
        for (std::uint32_t sid : b.shapes) {
            HostProxy p;
            p.body = bid;
            p.shape = int(sid);
            p.id = (w->leafCount == 0) ? 0u : 2u * w->leafCount - 1u;
            ++w->leafCount;
            p.flags = prx::kProxyPending;
            b.proxies.push_back(int(w->proxies.size()));
            w->proxies.push_back(p);
        }
    }
    w->bodies.push_back(std::move(b));
    if (out) *out = std::uint32_t(bid);
    return PRX_OK;
}

prx::StepParams ToParams(const prx_step_conf& c) {
    prx::StepParams p{};
    p.dt = c.delta_time;
    p.dtRatio = c.dt_ratio;
    p.minStillTimeToSleep = c.min_still_time_to_sleep;
    p.linearSlop = c.linear_slop;
    p.angularSlop = c.angular_slop;
    p.regResolutionRate = c.reg_resolution_rate;
    p.regMinSeparation = c.reg_min_separation;
    p.regMinMomentum = c.reg_min_momentum;
    p.velocityThreshold = c.velocity_threshold;
    p.maxTranslation = c.max_translation;
    p.maxRotation = c.max_rotation;
    p.maxLinearCorrection = c.max_linear_correction;
    p.maxAngularCorrection = c.max_angular_correction;
    p.linearSleepTolerance = c.linear_sleep_tolerance;
    p.angularSleepTolerance = c.angular_sleep_tolerance;
    p.displaceMultiplier = c.displace_multiplier;
    p.aabbExtension = c.aabb_extension;
    p.maxCirclesRatio = c.max_circles_ratio;
    p.velIters = c.reg_velocity_iters;
    p.posIters = c.reg_position_iters;
    p.warmStart = c.do_warm_start;
    p.blockSolve = c.do_blocksolve;
    p.doToi = c.do_toi;
    p.toiResolutionRate = c.toi_resolution_rate;
    p.toiMinSeparation = c.toi_min_separation;
    p.toiMinMomentum = c.toi_min_momentum;
    p.targetDepth = c.target_depth;
    p.tolerance = c.tolerance;
    p.toiVelIters = c.toi_velocity_iters;
    p.toiPosIters = c.toi_position_iters;
    p.maxToiRootIters = c.max_toi_root_iters;
    p.maxToiIters = c.max_toi_iters;
    p.maxDistanceIters = c.max_distance_iters;
    p.maxSubSteps = c.max_sub_steps;
    return p;
}

// CreateJoint (AabbTreeWorld.cpp:1132-1175) for any joint kind.
// A target joint has no body A: its a holds InvalidBodyID's index (BodyID.hpp:50).
constexpr std::uint32_t kNoBody = 0xffffu;

bool JointBodyValid(const prx_world* w, const HostJoint& j, bool bodyA) {
    const std::uint32_t id = bodyA ? j.a : j.b;
    if (bodyA && id == kNoBody && (j.flags >> 8) == prx::kJointTarget) return true;
    return id < w->bodies.size() && !w->bodies[id].destroyed;
}

// A body id the world holds and has not destroyed (the reference throws
// WasDestroyed for a destroyed one).
int CheckBody(const prx_world* w, std::uint32_t body) {
    if (body >= w->bodies.size()) return Fail(PRX_E_OUT_OF_RANGE, "no such body");
    if (w->bodies[body].destroyed) return Fail(PRX_E_WRONG_STATE, "the body was destroyed");
    return PRX_OK;
}

// Body::SetAwake (Body.cpp:163-171).
void WakeBody(HostBody& b) {
    if (b.flags & prx::kSpeedable) {
        b.flags |= prx::kAwake;
        b.under = 0.0f;
    }
}

// The contacts of a body being destroyed or changing type, in creation
// order (AabbTreeWorld::Destroy(ContactID, from) :2281-2288 and
// InternalDestroy :2245-2279): an end-contact call for a touching one (with
// contact events on; they come before the next step's calls), both bodies
// awake if the manifold has points, the id to the contact pool's free list.
void DestroyBodyContacts(prx_world* w, std::uint32_t body) {
    HostContacts& hc = w->contacts;
    std::size_t out = 0;
    for (std::size_t i = 0; i < hc.size(); ++i) {
        const int2 bd = hc.body[i];
        if (bd.x != int(body) && bd.y != int(body)) {
            if (out != i) {
                hc.prox[out] = hc.prox[i];
                hc.body[out] = hc.body[i];
                hc.shape[out] = hc.shape[i];
                hc.id[out] = hc.id[i];
                hc.flags[out] = hc.flags[i];
                hc.mat[out] = hc.mat[i];
                hc.geo[out] = hc.geo[i];
                hc.p0[out] = hc.p0[i];
                hc.p1[out] = hc.p1[i];
                hc.hdr[out] = hc.hdr[i];
            }
            ++out;
            continue;
        }
        if (w->record_events && (hc.flags[i] & prx::kTouching)) {
            w->pre_events.push_back(prx::kEvEnd);
            w->pre_events.push_back(hc.id[i]);
            ++w->pre_nevents;
        }
        if (((hc.hdr[i].x >> 8) & 0xffu) > 0 && !(hc.flags[i] & prx::kSensor)) {
            WakeBody(w->bodies[std::size_t(bd.x)]);
            WakeBody(w->bodies[std::size_t(bd.y)]);
        }
        hc.freelist.push_back(hc.id[i]);
    }
    hc.resize(out);
}

// AabbTreeWorld::Remove(JointID) (AabbTreeWorld.cpp:1177-1213) and the rest
// of Destroy(World&, JointID) (:1215-1227): both bodies wake, the joint
// leaves their joint lists (rebuilt at upload from the live joints in
// creation order) and its id goes to the pool's free list.  Contacts between
// its bodies would need filtering, but none exists while a joint keeps its
// bodies from colliding (AddJointRecord refuses such a joint between bodies
// that have a contact).
void DestroyJointRecord(prx_world* w, std::uint32_t joint) {
    HostJoint& j = w->joints[joint];
    if (j.a != kNoBody) WakeBody(w->bodies[j.a]);
    WakeBody(w->bodies[j.b]);
    j.destroyed = true;
    w->joint_free.push_back(joint);
}

int AddJointRecord(prx_world* w, const HostJoint& j, std::uint32_t* out) {
    if (!JointBodyValid(w, j, true) || !JointBodyValid(w, j, false)) return Fail(PRX_E_OUT_OF_RANGE, "no such body");
    if (!(j.flags & prx::kJointCollideConnected) && j.a != kNoBody) {
        for (std::size_t i = 0; i < w->contacts.size(); ++i) {
            const int2 bd = w->contacts.body[i];
            if ((bd.x == int(j.a) && bd.y == int(j.b)) || (bd.x == int(j.b) && bd.y == int(j.a)))
                return Fail(PRX_E_UNSUPPORTED, "joint between bodies that already have a contact");
        }
    }
    HostJoint k = j;
    k.seq = w->joint_seq++;
    k.destroyed = false;
    // The joint pool reuses the id freed last (ObjectPool::Allocate).
    if (!w->joint_free.empty()) {
        const std::uint32_t id = w->joint_free.back();
        w->joint_free.pop_back();
        w->joints[id] = k;
        if (out) *out = id;
        return PRX_OK;
    }
    if (out) *out = std::uint32_t(w->joints.size());
    w->joints.push_back(k);
    return PRX_OK;
}

int AddJoint(prx_world* w, const prx_revolute_joint_conf& c, std::uint32_t* out) {
    HostJoint j;
    j.a = c.body_a;
    j.b = c.body_b;
    j.flags = (c.enable_limit ? prx::kJointLimit : 0u) | (c.enable_motor ? prx::kJointMotor : 0u) |
              (c.collide_connected ? prx::kJointCollideConnected : 0u);
    j.la[0] = c.local_anchor_a[0];
    j.la[1] = c.local_anchor_a[1];
    j.la[2] = c.local_anchor_b[0];
    j.la[3] = c.local_anchor_b[1];
    j.cfg[0] = c.reference_angle;
    j.cfg[1] = c.lower_angle;
    j.cfg[2] = c.upper_angle;
    j.cfg[3] = c.motor_speed;
    j.maxTorque = c.max_motor_torque;
    return AddJointRecord(w, j, out);
}

// PulleyJointConf: anchors, ground anchors in cfg, the ratio in the maximum
// torque slot, the lengths in ext.
int AddPulleyJoint(prx_world* w, const prx_pulley_joint_conf& c, std::uint32_t* out) {
    HostJoint j;
    j.a = c.body_a;
    j.b = c.body_b;
    j.flags = (c.collide_connected ? prx::kJointCollideConnected : 0u) | (prx::kJointPulley << 8);
    j.la[0] = c.local_anchor_a[0];
    j.la[1] = c.local_anchor_a[1];
    j.la[2] = c.local_anchor_b[0];
    j.la[3] = c.local_anchor_b[1];
    j.cfg[0] = c.ground_anchor_a[0];
    j.cfg[1] = c.ground_anchor_a[1];
    j.cfg[2] = c.ground_anchor_b[0];
    j.cfg[3] = c.ground_anchor_b[1];
    j.maxTorque = c.ratio;
    j.ext[0] = c.length_a;
    j.ext[1] = c.length_b;
    return AddJointRecord(w, j, out);
}

// WheelJointConf: anchors, (frequency, damping ratio, -, motor speed), the
// maximum motor torque, and the local x axis of A.
int AddWheelJoint(prx_world* w, const prx_wheel_joint_conf& c, std::uint32_t* out) {
    if (!(c.frequency >= 0.0f)) return Fail(PRX_E_INVALID_ARGUMENT, "frequency must be non-negative");
    HostJoint j;
    j.a = c.body_a;
    j.b = c.body_b;
    j.flags = (c.enable_motor ? prx::kJointMotor : 0u) | (c.collide_connected ? prx::kJointCollideConnected : 0u) |
              (prx::kJointWheel << 8);
    j.la[0] = c.local_anchor_a[0];
    j.la[1] = c.local_anchor_a[1];
    j.la[2] = c.local_anchor_b[0];
    j.la[3] = c.local_anchor_b[1];
    j.cfg[0] = c.frequency;
    j.cfg[1] = c.damping_ratio;
    j.cfg[3] = c.motor_speed;
    j.maxTorque = c.max_motor_torque;
    j.ext[0] = c.local_axis_a[0];
    j.ext[1] = c.local_axis_a[1];
    return AddJointRecord(w, j, out);
}

// PrismaticJointConf: anchors, (reference angle, lower, upper, motor speed),
// the maximum motor force, and the local x axis of A (the revolute joint's
// slots plus the wheel joint's axis).
int AddPrismaticJoint(prx_world* w, const prx_prismatic_joint_conf& c, std::uint32_t* out) {
    HostJoint j;
    j.a = c.body_a;
    j.b = c.body_b;
    j.flags = (c.enable_limit ? prx::kJointLimit : 0u) | (c.enable_motor ? prx::kJointMotor : 0u) |
              (c.collide_connected ? prx::kJointCollideConnected : 0u) | (prx::kJointPrismatic << 8);
    j.la[0] = c.local_anchor_a[0];
    j.la[1] = c.local_anchor_a[1];
    j.la[2] = c.local_anchor_b[0];
    j.la[3] = c.local_anchor_b[1];
    j.cfg[0] = c.reference_angle;
    j.cfg[1] = c.lower_translation;
    j.cfg[2] = c.upper_translation;
    j.cfg[3] = c.motor_speed;
    j.maxTorque = c.max_motor_force;
    j.ext[0] = c.local_axis_a[0];
    j.ext[1] = c.local_axis_a[1];
    return AddJointRecord(w, j, out);
}

// GearJointConf: bodies A, B of the record plus C, D; the sides' data, the
// ratio and the constant as given (CreateJoint validates bodies A and B only,
// AabbTreeWorld.cpp:1143-1149; C and D must be bodies of the world too, the
// solver indexes them).
int AddGearJoint(prx_world* w, const prx_gear_joint_conf& c, std::uint32_t* out) {
    if (c.body_c >= w->bodies.size() || c.body_d >= w->bodies.size() || w->bodies[c.body_c].destroyed ||
        w->bodies[c.body_d].destroyed)
        return Fail(PRX_E_OUT_OF_RANGE, "no such body");
    HostJoint j;
    j.a = c.body_a;
    j.b = c.body_b;
    j.c = c.body_c;
    j.d = c.body_d;
    j.flags = (c.collide_connected ? prx::kJointCollideConnected : 0u) | (c.prismatic_ac ? prx::kGearPrisAC : 0u) |
              (c.prismatic_bd ? prx::kGearPrisBD : 0u) | (prx::kJointGear << 8);
    for (int k = 0; k < 6; ++k) {
        j.gear[k] = c.ac[k];
        j.gear[6 + k] = c.bd[k];
    }
    j.ratio = c.ratio;
    j.constant = c.constant;
    return AddJointRecord(w, j, out);
}

// GetGearJointConf(const World&, JointID, JointID, Real) (GearJointConf.cpp:
// 89-169): bodies A, B the joints' second bodies, C, D their first; a revolute
// side's scalar the angle difference less the reference angle, a prismatic
// side's the offset of the second body's anchor from the first's along the
// axis, in the first body's frame; constant = scalar1 + ratio * scalar2.  A
// body's transformation is derived from its sweep (position less the rotated
// local centre), as after any step.
int GearConf(const prx_world* w, std::uint32_t j1, std::uint32_t j2, float ratio, prx_gear_joint_conf& out) {
    const std::uint32_t ids[2] = {j1, j2};
    for (std::uint32_t id : ids) {
        if (id >= w->joints.size() || w->joints[id].destroyed) return Fail(PRX_E_OUT_OF_RANGE, "no such joint");
        const std::uint32_t kind = w->joints[id].flags >> 8;
        if (kind != prx::kJointRevolute && kind != prx::kJointPrismatic)
            return Fail(PRX_E_INVALID_ARGUMENT, "gear joints couple revolute or prismatic joints only");
    }
    out = prx_gear_joint_conf{};
    out.body_a = w->joints[j1].b;
    out.body_b = w->joints[j2].b;
    out.body_c = w->joints[j1].a;
    out.body_d = w->joints[j2].a;
    auto xf = [&](std::uint32_t body, V2& p, V2& q) {
        const HostBody& hb = w->bodies[body];
        q = UnitFromAngle(hb.pos1[2]);
        p = V2{hb.pos1[0], hb.pos1[1]} - Rotate(hb.lc, q);
    };
    float scalar[2];
    for (int k = 0; k < 2; ++k) {
        const HostJoint& jt = w->joints[ids[k]];
        float* side = k ? out.bd : out.ac;
        const std::uint32_t self = k ? out.body_b : out.body_a;
        const std::uint32_t other = k ? out.body_d : out.body_c;
        if ((jt.flags >> 8) == prx::kJointRevolute) {
            const float referenceAngle = jt.cfg[0];
            side[0] = referenceAngle;
            scalar[k] = ((w->bodies[self].pos1[2] - w->bodies[other].pos1[2]) - referenceAngle) / 1.0f;
        } else {
            V2 ps, qs, po, qo;
            xf(self, ps, qs);
            xf(other, po, qo);
            const V2 laA{jt.la[0], jt.la[1]}, laB{jt.la[2], jt.la[3]}, lxA{jt.ext[0], jt.ext[1]};
            side[0] = laA.x;
            side[1] = laA.y;
            side[2] = laB.x;
            side[3] = laB.y;
            side[4] = lxA.x;
            side[5] = lxA.y;
            const V2 pOther = laA;
            const V2 pSelf = InvRotate(Rotate(laB, qs) + (ps - po), qo);
            scalar[k] = Dot(pSelf - pOther, lxA) / 1.0f;
            (k ? out.prismatic_bd : out.prismatic_ac) = 1u;
        }
    }
    out.ratio = ratio;
    out.constant = scalar[0] + out.ratio * scalar[1];
    return PRX_OK;
}

// TargetJointConf: no body A; the anchor on B in the anchor B slot, (target x,
// y, maximum force, frequency), the damping ratio in the maximum torque slot.
int AddTargetJoint(prx_world* w, const prx_target_joint_conf& c, std::uint32_t* out) {
    if (!(c.max_force >= 0.0f) || !(c.frequency >= 0.0f) || !(c.damping_ratio >= 0.0f))
        return Fail(PRX_E_INVALID_ARGUMENT, "maximum force, frequency and damping ratio must be non-negative");
    HostJoint j;
    j.a = kNoBody;
    j.b = c.body;
    j.flags = (c.collide_connected ? prx::kJointCollideConnected : 0u) | (prx::kJointTarget << 8);
    j.la[2] = c.local_anchor[0];
    j.la[3] = c.local_anchor[1];
    j.cfg[0] = c.target[0];
    j.cfg[1] = c.target[1];
    j.cfg[2] = c.max_force;
    j.cfg[3] = c.frequency;
    j.maxTorque = c.damping_ratio;
    return AddJointRecord(w, j, out);
}

// MotorJointConf: the linear offset takes the anchor A slot, then (angular
// offset, maximum force, maximum torque, correction factor).
int AddMotorJoint(prx_world* w, const prx_motor_joint_conf& c, std::uint32_t* out) {
    if (!(c.max_force >= 0.0f) || !(c.max_torque >= 0.0f))
        return Fail(PRX_E_INVALID_ARGUMENT, "maximum force and torque must be non-negative");
    HostJoint j;
    j.a = c.body_a;
    j.b = c.body_b;
    j.flags = (c.collide_connected ? prx::kJointCollideConnected : 0u) | (prx::kJointMotorKind << 8);
    j.la[0] = c.linear_offset[0];
    j.la[1] = c.linear_offset[1];
    j.cfg[0] = c.angular_offset;
    j.cfg[1] = c.max_force;
    j.cfg[2] = c.max_torque;
    j.cfg[3] = c.correction_factor;
    return AddJointRecord(w, j, out);
}

// A distance or weld joint: kind, anchors, then (length | reference angle,
// frequency, damping ratio).
int AddSpringJoint(prx_world* w, std::uint32_t kind, std::uint32_t a, std::uint32_t b, const float la[2],
                   const float lb[2], float first, float frequency, float dampingRatio, bool collideConnected,
                   std::uint32_t* out) {
    if (!(frequency >= 0.0f)) return Fail(PRX_E_INVALID_ARGUMENT, "frequency must be non-negative");
    HostJoint j;
    j.a = a;
    j.b = b;
    j.flags = (collideConnected ? prx::kJointCollideConnected : 0u) | (kind << 8);
    j.la[0] = la[0];
    j.la[1] = la[1];
    j.la[2] = lb[0];
    j.la[3] = lb[1];
    j.cfg[0] = first;
    j.cfg[1] = frequency;
    j.cfg[2] = dampingRatio;
    return AddJointRecord(w, j, out);
}

int ContactCap(const prx_world* w) {
    const int per = w->conf.contact_capacity_per_proxy ? int(w->conf.contact_capacity_per_proxy) : 8;
    const int c = std::max<int>(int(w->contacts.size()) + 64, per * int(w->proxies.size()) + 64);
    return std::max<int>(c, int(w->bodies.size()) + 64);
}

template <class T>
int Upload(T* dst, const std::vector<T>& src) {
    if (!src.empty()) PRX_HIP(hipMemcpy(dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice));
    return PRX_OK;
}

// Launch schedule from the batch conf: dynamic LDS per workgroup (at least
// what the largest world needs) and the launch order of the worlds.
int ApplySchedule(prx_batch* b) {
    int lds = b->lds_min;
    if (b->conf.lds_bytes) lds = std::max<int>(lds, int(std::min<std::uint32_t>(b->conf.lds_bytes, 64u * 1024u)));
    b->lds_bytes = ((lds + 255) / 256) * 256;
    // Launch order: largest worlds first (stable), so the long worlds start
    // early and the small ones fill in behind them.
    const int nw = int(b->worlds.size());
    std::vector<int> order(static_cast<std::size_t>(nw));
    std::iota(order.begin(), order.end(), 0);
    if (b->conf.order == 0) {
        std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
            return b->hdr[std::size_t(x)].nbodies + b->hdr[std::size_t(x)].nproxies >
                   b->hdr[std::size_t(y)].nbodies + b->hdr[std::size_t(y)].nproxies;
        });
    }
    PRX_HIP(hipMemcpy(b->d_schedule, order.data(), order.size() * sizeof(int), hipMemcpyHostToDevice));
    b->graph_valid = false;
    return PRX_OK;
}

int BuildBatch(prx_batch* b) {
    const int nw = int(b->worlds.size());
    std::size_t B = 0, P = 0, S = 0, V = 0, C = 0, J = 0;
    b->hdr.assign(std::size_t(nw), prx::WorldHeader{});
    b->shape_base.assign(std::size_t(nw), 0);
    b->vert_base.assign(std::size_t(nw), 0);
    for (int i = 0; i < nw; ++i) {
        prx_world* w = b->worlds[std::size_t(i)];
        if (w->proxies.size() >= 65535) return Fail(PRX_E_LENGTH, "more than 65534 proxies in one world");
        auto& h = b->hdr[std::size_t(i)];
        h.body_off = int(B);
        h.nbodies = int(w->bodies.size());
        h.proxy_off = int(P);
        h.nproxies = int(w->proxies.size());
        h.contact_off = int(C);
        h.contact_cap = ContactCap(w);
        h.ncontacts = int(w->contacts.size());
        h.pool_size = w->contacts.pool;
        h.nfree = int(w->contacts.freelist.size());
        h.status = 0;
        h.joint_off = int(J);
        h.njoints = int(w->joints.size());
        J += w->joints.size();
        b->shape_base[std::size_t(i)] = int(S);
        b->vert_base[std::size_t(i)] = int(V);
        B += w->bodies.size();
        P += w->proxies.size();
        S += w->shapes.size();
        for (const auto& s : w->shapes) V += s.v.size();
        C += std::size_t(h.contact_cap);
    }
    // Listener-call segments of the worlds that record them (the most words
    // a step can write for the world's contact capacity).
    std::size_t E = 0;
    for (int i = 0; i < nw; ++i) {
        auto& h = b->hdr[std::size_t(i)];
        h.ev_off = int(E);
        const prx_world* wi = b->worlds[std::size_t(i)];
        h.ev_cap = !wi->record_events ? 0
                   : wi->event_words  ? int(wi->event_words)
                                      : prx::kEvWordsPerSlot * h.contact_cap + 16;
        E += std::size_t(h.ev_cap);
        if (E > std::size_t(0x7fffffff)) return Fail(PRX_E_LENGTH, "contact listener records exceed 2^31 words");
    }
    int maxNb = 0;
    for (const prx_world* w : b->worlds) maxNb = std::max<int>(maxNb, int(w->bodies.size()));
    // Body constraint state (32 B per body) plus the record stage; worlds whose
    // bodies do not fit keep them in global scratch.
    const int stageBytes = prx::kStageFloat4 * 16;
    b->lds_min = std::min<int>(64 * 1024, std::max<int>(1024 + stageBytes, ((maxNb * 32 + stageBytes + 255) / 256) * 256));
    b->totB = B;
    b->totP = P;
    b->totS = S;
    b->totV = V;
    b->totC = C;
    b->totJ = J;
    b->totE = E;
    auto& a = b->arena;
    auto& d = b->d;
    d.nworlds = nw;
    PRX_HIP(a.alloc(d.hdr, std::size_t(nw)));
    PRX_HIP(a.alloc(b->d_schedule, std::size_t(nw)));
    PRX_HIP(a.alloc(d.bflags, B));
    PRX_HIP(a.alloc(d.pos0, B));
    PRX_HIP(a.alloc(d.pos1, B));
    PRX_HIP(a.alloc(d.vel, B));
    PRX_HIP(a.alloc(d.xf, B));
    PRX_HIP(a.alloc(d.lcinv, B));
    PRX_HIP(a.alloc(d.acc, B));
    PRX_HIP(a.alloc(d.damp, B));
    PRX_HIP(a.alloc(d.bproxy, B));
    PRX_HIP(a.alloc(d.bproxy_list, P));
    PRX_HIP(a.alloc(d.pbody, P));
    PRX_HIP(a.alloc(d.pshape, P));
    PRX_HIP(a.alloc(d.pid, P));
    PRX_HIP(a.alloc(d.paabb, P));
    PRX_HIP(a.alloc(d.pflags, P));
    PRX_HIP(a.alloc(d.shape_i, S));
    PRX_HIP(a.alloc(d.shape_f, S));
    PRX_HIP(a.alloc(d.shape_catmask, S));
    PRX_HIP(a.alloc(d.verts, V));
    PRX_HIP(a.alloc(d.norms, V));
    PRX_HIP(a.alloc(d.c_prox, C));
    PRX_HIP(a.alloc(d.c_body, C));
    PRX_HIP(a.alloc(d.c_shape, C));
    PRX_HIP(a.alloc(d.c_id, C));
    PRX_HIP(a.alloc(d.c_flags, C));
    PRX_HIP(a.alloc(d.c_mat, C));
    PRX_HIP(a.alloc(d.m_hdr, C));
    PRX_HIP(a.alloc(d.m_geo, C));
    PRX_HIP(a.alloc(d.m_p0, C));
    PRX_HIP(a.alloc(d.m_p1, C));
    PRX_HIP(a.alloc(d.c_free, C));
    PRX_HIP(a.alloc(d.s_keys, C));
    PRX_HIP(a.alloc(d.s_keys2, C));
    PRX_HIP(a.alloc(d.s_csr_off, B + std::size_t(nw)));
    PRX_HIP(a.alloc(d.s_csr, 4 * C));
    PRX_HIP(a.alloc(d.s_stack, 2 * B));
    PRX_HIP(a.alloc(d.s_isl_b, B));
    PRX_HIP(a.alloc(d.s_isl_c, C));
    PRX_HIP(a.alloc(d.s_isl_body, B + C));
    PRX_HIP(a.alloc(d.s_isl_contact, C));
    PRX_HIP(a.alloc(d.s_isl_start, B + std::size_t(nw)));
    PRX_HIP(a.alloc(d.s_bc_pos, B));
    PRX_HIP(a.alloc(d.s_bc_vel, B));
    PRX_HIP(a.alloc(d.s_vc, C * std::size_t(prx::kVcStride)));
    PRX_HIP(a.alloc(d.s_pc, C * std::size_t(prx::kPcStride)));
    PRX_HIP(a.alloc(d.s_vc2, C * std::size_t(prx::kVcStride)));
    PRX_HIP(a.alloc(d.s_pc2, C * std::size_t(prx::kPcStride)));
    PRX_HIP(a.alloc(d.s_lvl, C * std::size_t(prx::kLevelInts)));
    PRX_HIP(a.alloc(d.j_body, J));
    PRX_HIP(a.alloc(d.j_anchor, J));
    PRX_HIP(a.alloc(d.j_cfg, J));
    PRX_HIP(a.alloc(d.j_cfg2, J));
    PRX_HIP(a.alloc(d.j_gear, 2 * J));
    PRX_HIP(a.alloc(d.j_imp, J));
    PRX_HIP(a.alloc(d.j_limit, J));
    PRX_HIP(a.alloc(d.bjoint, B));
    PRX_HIP(a.alloc(d.bjoint_list, 2 * J));
    PRX_HIP(a.alloc(d.s_jc, J * std::size_t(prx::kJcStride)));
    PRX_HIP(a.alloc(d.s_isl_joint, J));
    PRX_HIP(a.alloc(d.s_isl_jstart, B + std::size_t(nw)));
    PRX_HIP(a.alloc(d.s_isl_j, J));
    d.ev = nullptr;
    if (E) PRX_HIP(a.alloc(d.ev, E));
    PRX_HIP(a.alloc(d.alpha0, B));
    PRX_HIP(a.alloc(d.t_isl_b, B));
    PRX_HIP(a.alloc(d.c_toi, C));
    PRX_HIP(a.alloc(d.c_toiinfo, C));
    PRX_HIP(a.alloc(d.t_isl_c, C));
    PRX_HIP(hipMemset(d.alpha0, 0, std::max<std::size_t>(B, 1) * sizeof(float)));
    PRX_HIP(hipMemset(d.c_toiinfo, 0, std::max<std::size_t>(C, 1) * sizeof(std::uint32_t)));
    d.schedule = b->d_schedule;
    return ApplySchedule(b);
}

// Host mirror -> device for all member worlds.
int UploadBatch(prx_batch* b) {
    const std::size_t B = b->totB, P = b->totP, S = b->totS, V = b->totV, C = b->totC;
    std::vector<std::uint32_t> bflags(B), pid(P), pflags(P), catmask(S), cid(C), cflags(C), cfree(C);
    std::vector<float4> pos0(B), pos1(B), vel(B), xf(B), lcinv(B), acc(B), paabb(P), shape_f(S);
    std::vector<float2> damp(B), verts(V), norms(V);
    std::vector<int2> bproxy(B), cprox(C), cbody(C), cshape(C);
    std::vector<int> bplist(P), pbody(P), pshape(P);
    std::vector<int4> shape_i(S);
    std::vector<float4> cmat(C), mgeo(C), mp0(C), mp1(C);
    std::vector<uint4> mhdr(C);
    const std::size_t J = b->totJ;
    std::vector<int2> jbody(J), bjoint(B), bjlist(2 * J);
    std::vector<float4> janchor(J), jcfg(J), jcfg2(J), jimp(J), jgear(2 * J);
    std::vector<std::uint32_t> jlimit(J);
    b->joint_kinds = false;
    for (std::size_t wi = 0; wi < b->worlds.size(); ++wi) {
        const prx_world* w = b->worlds[wi];
        const auto& h = b->hdr[wi];
        // Joints and per-body joint lists in creation order (m_bodyJoints).
        {
            std::vector<std::vector<int2>> lists(w->bodies.size());
            std::vector<std::size_t> order(w->joints.size());
            for (std::size_t ji = 0; ji < order.size(); ++ji) order[ji] = ji;
            std::stable_sort(order.begin(), order.end(),
                             [w](std::size_t a, std::size_t c) { return w->joints[a].seq < w->joints[c].seq; });
            for (std::size_t ji : order) {
                const HostJoint& j = w->joints[ji];
                const std::size_t g = std::size_t(h.joint_off) + ji;
                // A target joint's record names body B twice (its solver
                // leaves the body-A copy unchanged, written back first).
                const bool noA = j.a == kNoBody;
                jbody[g] = int2{noA ? int(j.b) : int(j.a), int(j.b)};
                janchor[g] = float4{j.la[0], j.la[1], j.la[2], j.la[3]};
                jcfg[g] = float4{j.cfg[0], j.cfg[1], j.cfg[2], j.cfg[3]};
                float fl;
                std::memcpy(&fl, &j.flags, 4);
                jcfg2[g] = float4{j.maxTorque, fl, j.ext[0], j.ext[1]};
                if ((j.flags >> 8) == prx::kJointGear) {
                    janchor[g] = float4{j.gear[0], j.gear[1], j.gear[2], j.gear[3]};
                    jcfg[g] = float4{j.gear[6], j.gear[7], j.gear[8], j.gear[9]};
                    jcfg2[g] = float4{0.0f, fl, 0.0f, 0.0f};
                    float cbits, dbits;
                    const int ci = int(j.c), di = int(j.d);
                    std::memcpy(&cbits, &ci, 4);
                    std::memcpy(&dbits, &di, 4);
                    jgear[2 * g] = float4{j.ratio, j.constant, cbits, dbits};
                    jgear[2 * g + 1] = float4{j.gear[4], j.gear[5], j.gear[10], j.gear[11]};
                }
                if (j.flags >> 8) b->joint_kinds = true;
                jimp[g] = float4{j.imp[0], j.imp[1], j.imp[2], j.imp[3]};
                jlimit[g] = j.limit;
                if (j.destroyed) continue;
                if (!noA) lists[j.a].push_back(int2{int(ji), int(j.b)});
                lists[j.b].push_back(int2{int(ji), noA ? -1 : int(j.a)});
            }
            int at = 0;
            for (std::size_t bi = 0; bi < w->bodies.size(); ++bi) {
                bjoint[std::size_t(h.body_off) + bi] = int2{at, int(lists[bi].size())};
                for (const int2& e : lists[bi]) bjlist[2 * std::size_t(h.joint_off) + std::size_t(at++)] = e;
            }
        }
        const int sb = b->shape_base[wi];
        int vb = b->vert_base[wi];
        for (std::size_t si = 0; si < w->shapes.size(); ++si) {
            const HostShape& s = w->shapes[si];
            const std::size_t g = std::size_t(sb) + si;
            shape_i[g] = int4{int(s.v.size()), vb, int(s.sensor), s.group};
            shape_f[g] = float4{s.radius, s.friction, s.restitution, s.density};
            catmask[g] = s.catmask;
            for (std::size_t k = 0; k < s.v.size(); ++k) {
                verts[std::size_t(vb) + k] = float2{s.v[k].x, s.v[k].y};
                norms[std::size_t(vb) + k] = float2{s.n[k].x, s.n[k].y};
            }
            vb += int(s.v.size());
        }
        int plist = 0;
        for (std::size_t bi = 0; bi < w->bodies.size(); ++bi) {
            const HostBody& hb = w->bodies[bi];
            const std::size_t g = std::size_t(h.body_off) + bi;
            bflags[g] = hb.flags;
            pos0[g] = float4{hb.pos0[0], hb.pos0[1], hb.pos0[2], 0};
            pos1[g] = float4{hb.pos1[0], hb.pos1[1], hb.pos1[2], 0};
            vel[g] = float4{hb.vel[0], hb.vel[1], hb.vel[2], hb.under};
            const V2 q = UnitFromAngle(hb.pos1[2]);
            const V2 p = V2{hb.pos1[0], hb.pos1[1]} - Rotate(hb.lc, q);
            xf[g] = float4{p.x, p.y, q.x, q.y};
            lcinv[g] = float4{hb.lc.x, hb.lc.y, hb.invMass, hb.invRotI};
            acc[g] = float4{hb.acc[0], hb.acc[1], hb.acc[2], 0};
            damp[g] = float2{hb.damp[0], hb.damp[1]};
            bproxy[g] = int2{plist, int(hb.proxies.size())};
            for (int p2 : hb.proxies) bplist[std::size_t(h.proxy_off) + std::size_t(plist++)] = p2;
        }
        for (std::size_t pi = 0; pi < w->proxies.size(); ++pi) {
            const HostProxy& hp = w->proxies[pi];
            const std::size_t g = std::size_t(h.proxy_off) + pi;
            pbody[g] = hp.body;
            pshape[g] = sb + hp.shape;
            pid[g] = hp.id;
            pflags[g] = hp.dead ? 0u : hp.flags;
            // A dead proxy's box is empty: it overlaps nothing.
            const float inf = std::numeric_limits<float>::infinity();
            paabb[g] = hp.dead ? float4{inf, inf, -inf, -inf} : float4{hp.aabb[0], hp.aabb[1], hp.aabb[2], hp.aabb[3]};
        }
        const HostContacts& hc = w->contacts;
        for (std::size_t ci = 0; ci < hc.size(); ++ci) {
            const std::size_t g = std::size_t(h.contact_off) + ci;
            cprox[g] = hc.prox[ci];
            cbody[g] = hc.body[ci];
            cshape[g] = int2{sb + hc.shape[ci].x, sb + hc.shape[ci].y};
            cid[g] = hc.id[ci];
            cflags[g] = hc.flags[ci];
            cmat[g] = hc.mat[ci];
            mhdr[g] = hc.hdr[ci];
            mgeo[g] = hc.geo[ci];
            mp0[g] = hc.p0[ci];
            mp1[g] = hc.p1[ci];
        }
        for (std::size_t k = 0; k < hc.freelist.size(); ++k) cfree[std::size_t(h.contact_off) + k] = hc.freelist[k];
    }
    auto& d = b->d;
    // Listener calls of the last step, as the host mirror holds them.
    for (std::size_t wi = 0; wi < b->worlds.size(); ++wi) {
        const prx_world* w = b->worlds[wi];
        auto& h = b->hdr[wi];
        h.ev_words = 0;
        h.ev_count = 0;
        if (h.ev_cap == 0 || w->events.size() > std::size_t(h.ev_cap)) continue;
        h.ev_words = int(w->events.size());
        h.ev_count = int(w->nevents);
        if (!w->events.empty())
            PRX_HIP(hipMemcpy(d.ev + h.ev_off, w->events.data(), w->events.size() * 4, hipMemcpyHostToDevice));
    }
    int st;
    if ((st = Upload(d.hdr, b->hdr)) || (st = Upload(d.bflags, bflags)) || (st = Upload(d.pos0, pos0)) ||
        (st = Upload(d.pos1, pos1)) || (st = Upload(d.vel, vel)) || (st = Upload(d.xf, xf)) ||
        (st = Upload(d.lcinv, lcinv)) || (st = Upload(d.acc, acc)) || (st = Upload(d.damp, damp)) ||
        (st = Upload(d.bproxy, bproxy)) || (st = Upload(d.bproxy_list, bplist)) || (st = Upload(d.pbody, pbody)) ||
        (st = Upload(d.pshape, pshape)) || (st = Upload(d.pid, pid)) || (st = Upload(d.paabb, paabb)) ||
        (st = Upload(d.pflags, pflags)) || (st = Upload(d.shape_i, shape_i)) || (st = Upload(d.shape_f, shape_f)) ||
        (st = Upload(d.shape_catmask, catmask)) || (st = Upload(d.verts, verts)) || (st = Upload(d.norms, norms)) ||
        (st = Upload(d.c_prox, cprox)) || (st = Upload(d.c_body, cbody)) || (st = Upload(d.c_shape, cshape)) ||
        (st = Upload(d.c_id, cid)) || (st = Upload(d.c_flags, cflags)) || (st = Upload(d.c_mat, cmat)) ||
        (st = Upload(d.m_hdr, mhdr)) || (st = Upload(d.m_geo, mgeo)) || (st = Upload(d.m_p0, mp0)) ||
        (st = Upload(d.m_p1, mp1)) || (st = Upload(d.c_free, cfree)) || (st = Upload(d.j_body, jbody)) ||
        (st = Upload(d.j_anchor, janchor)) || (st = Upload(d.j_cfg, jcfg)) || (st = Upload(d.j_cfg2, jcfg2)) || (st = Upload(d.j_gear, jgear)) ||
        (st = Upload(d.j_imp, jimp)) || (st = Upload(d.j_limit, jlimit)) || (st = Upload(d.bjoint, bjoint)) ||
        (st = Upload(d.bjoint_list, bjlist)))
        return st;
    return PRX_OK;
}

template <class T>
int Download(std::vector<T>& dst, const T* src, std::size_t n) {
    dst.resize(n);
    if (n) PRX_HIP(hipMemcpy(dst.data(), src, n * sizeof(T), hipMemcpyDeviceToHost));
    return PRX_OK;
}

// Copies the listener calls of a world's last step to its host mirror.
int FetchEvents(prx_batch* b, int wi, const prx::WorldHeader& h) {
    prx_world* w = b->worlds[std::size_t(wi)];
    if (h.ev_cap == 0) {
        w->events.clear();
        w->nevents = 0;
        w->pre_events.clear();
        w->pre_nevents = 0;
        return PRX_OK;
    }
    const std::size_t words = std::size_t(std::min(h.ev_words, h.ev_cap));
    if (int st = Download(w->events, b->d.ev + h.ev_off, words)) return st;
    w->nevents = h.ev_words <= h.ev_cap ? std::uint32_t(h.ev_count) : 0u;
    if (!w->pre_events.empty()) {
        w->events.insert(w->events.begin(), w->pre_events.begin(), w->pre_events.end());
        w->nevents += w->pre_nevents;
        w->pre_events.clear();
        w->pre_nevents = 0;
    }
    return PRX_OK;
}

int DownloadBatch(prx_batch* b) {
    PRX_HIP(hipStreamSynchronize(b->stream));
    const std::size_t B = b->totB, P = b->totP, C = b->totC;
    auto& d = b->d;
    std::vector<prx::WorldHeader> hdr;
    std::vector<std::uint32_t> bflags, pflags, cid, cflags, cfree;
    std::vector<float4> pos0, pos1, vel, paabb, cmat, mgeo, mp0, mp1;
    std::vector<int2> cprox, cbody, cshape;
    std::vector<uint4> mhdr;
    int st;
    if ((st = Download(hdr, d.hdr, b->worlds.size())) || (st = Download(bflags, d.bflags, B)) ||
        (st = Download(pos0, d.pos0, B)) || (st = Download(pos1, d.pos1, B)) || (st = Download(vel, d.vel, B)) ||
        (st = Download(paabb, d.paabb, P)) || (st = Download(pflags, d.pflags, P)) ||
        (st = Download(cprox, d.c_prox, C)) || (st = Download(cbody, d.c_body, C)) ||
        (st = Download(cshape, d.c_shape, C)) || (st = Download(cid, d.c_id, C)) ||
        (st = Download(cflags, d.c_flags, C)) || (st = Download(cmat, d.c_mat, C)) ||
        (st = Download(mhdr, d.m_hdr, C)) || (st = Download(mgeo, d.m_geo, C)) || (st = Download(mp0, d.m_p0, C)) ||
        (st = Download(mp1, d.m_p1, C)) || (st = Download(cfree, d.c_free, C)))
        return st;
    std::vector<float4> jimp;
    std::vector<std::uint32_t> jlimit;
    if ((st = Download(jimp, d.j_imp, b->totJ)) || (st = Download(jlimit, d.j_limit, b->totJ))) return st;
    b->hdr = hdr;
    for (std::size_t wi = 0; wi < b->worlds.size(); ++wi) {
        prx_world* w = b->worlds[wi];
        if (!w) continue;
        const auto& h = hdr[wi];
        const int sb = b->shape_base[wi];
        for (std::size_t bi = 0; bi < w->bodies.size(); ++bi) {
            HostBody& hb = w->bodies[bi];
            const std::size_t g = std::size_t(h.body_off) + bi;
            hb.flags = bflags[g];
            hb.pos0[0] = pos0[g].x;
            hb.pos0[1] = pos0[g].y;
            hb.pos0[2] = pos0[g].z;
            hb.pos1[0] = pos1[g].x;
            hb.pos1[1] = pos1[g].y;
            hb.pos1[2] = pos1[g].z;
            hb.vel[0] = vel[g].x;
            hb.vel[1] = vel[g].y;
            hb.vel[2] = vel[g].z;
            hb.under = vel[g].w;
        }
        for (std::size_t pi = 0; pi < w->proxies.size(); ++pi) {
            HostProxy& hp = w->proxies[pi];
            const std::size_t g = std::size_t(h.proxy_off) + pi;
            hp.flags = pflags[g];
            hp.aabb[0] = paabb[g].x;
            hp.aabb[1] = paabb[g].y;
            hp.aabb[2] = paabb[g].z;
            hp.aabb[3] = paabb[g].w;
        }
        HostContacts& hc = w->contacts;
        hc.resize(std::size_t(h.ncontacts));
        for (int ci = 0; ci < h.ncontacts; ++ci) {
            const std::size_t g = std::size_t(h.contact_off) + std::size_t(ci);
            hc.prox[std::size_t(ci)] = cprox[g];
            hc.body[std::size_t(ci)] = cbody[g];
            hc.shape[std::size_t(ci)] = int2{cshape[g].x - sb, cshape[g].y - sb};
            hc.id[std::size_t(ci)] = cid[g];
            hc.flags[std::size_t(ci)] = cflags[g];
            hc.mat[std::size_t(ci)] = cmat[g];
            hc.hdr[std::size_t(ci)] = mhdr[g];
            hc.geo[std::size_t(ci)] = mgeo[g];
            hc.p0[std::size_t(ci)] = mp0[g];
            hc.p1[std::size_t(ci)] = mp1[g];
        }
        for (std::size_t ji = 0; ji < w->joints.size(); ++ji) {
            HostJoint& j = w->joints[ji];
            const std::size_t g = std::size_t(h.joint_off) + ji;
            j.imp[0] = jimp[g].x;
            j.imp[1] = jimp[g].y;
            j.imp[2] = jimp[g].z;
            j.imp[3] = jimp[g].w;
            j.limit = jlimit[g];
        }
        hc.freelist.assign(cfree.begin() + h.contact_off, cfree.begin() + h.contact_off + h.nfree);
        hc.pool = h.pool_size;
        if (int st2 = FetchEvents(b, int(wi), h)) return st2;
        w->last_stats = StatsOf(h);
        w->host_stale = false;
    }
    return PRX_OK;
}

int CreateBatch(prx_world* const* worlds, int n, const prx_batch_conf* conf, prx_batch** out) {
    if (!worlds || n <= 0 || !out) return Fail(PRX_E_INVALID_ARGUMENT, "bad batch arguments");
    for (int i = 0; i < n; ++i) {
        if (!worlds[i]) return Fail(PRX_E_INVALID_ARGUMENT, "null world in batch");
        const int st = Detach(worlds[i]);
        if (st != PRX_OK) return st;
    }
    auto b = std::make_unique<prx_batch>();
    if (conf) b->conf = *conf;
    else prx_batch_conf_default(&b->conf);
    b->worlds.assign(worlds, worlds + n);
    PRX_HIP(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
    PRX_HIP(hipEventCreate(&b->ev0));
    PRX_HIP(hipEventCreate(&b->ev1));
    int st = BuildBatch(b.get());
    if (st != PRX_OK) return st;
    st = UploadBatch(b.get());
    if (st != PRX_OK) return st;
    for (int i = 0; i < n; ++i) {
        worlds[i]->owner = b.get();
        worlds[i]->index_in_owner = i;
        worlds[i]->host_stale = false;
    }
    *out = b.release();
    return PRX_OK;
}

int CheckStatus(prx_batch* b) {
    // Status words are checked after each synchronising call.  A lost
    // listener call or a TOI phase cut short affects the steps since the last
    // check only: those bits are reported once, then cleared, so the worlds
    // step on.  Contact and scratch overflows leave a world's contact set
    // incomplete and stay set.
    std::vector<prx::WorldHeader> hdr(b->worlds.size());
    PRX_HIP(hipMemcpy(hdr.data(), b->d.hdr, hdr.size() * sizeof(prx::WorldHeader), hipMemcpyDeviceToHost));
    constexpr std::uint32_t kRecoverable = prx::kStatusEventOverflow | prx::kStatusToiLimit;
    for (std::size_t i = 0; i < hdr.size(); ++i) {
        if (std::uint32_t(hdr[i].status) & kRecoverable) {
            const int cleared = int(std::uint32_t(hdr[i].status) & ~kRecoverable);
            PRX_HIP(hipMemcpy(reinterpret_cast<char*>(b->d.hdr + i) + offsetof(prx::WorldHeader, status), &cleared,
                              sizeof(int), hipMemcpyHostToDevice));
        }
    }
    for (const auto& h : hdr) {
        if (h.status & prx::kStatusContactOverflow)
            return Fail(PRX_E_LENGTH, "contact capacity of a world exceeded (raise contact_capacity_per_proxy)");
        if (h.status & prx::kStatusScratchOverflow)
            return Fail(PRX_E_LENGTH, "candidate pair scratch of a world exceeded");
        if (h.status & prx::kStatusEventOverflow)
            return Fail(PRX_E_LENGTH, "contact listener record of a world exceeded");
        if (h.status & prx::kStatusToiLimit)
            return Fail(PRX_E_LENGTH, "TOI events of one step of a world exceeded (maxSubSteps + 1) per contact slot");
    }
    return PRX_OK;
}

int StepBatch(prx_batch* b, const prx_step_conf* conf, int nsteps) {
    if (!conf) return Fail(PRX_E_INVALID_ARGUMENT, "null step conf");
    if (nsteps < 0) return Fail(PRX_E_INVALID_ARGUMENT, "negative step count");
    prx::StepParams p = ToParams(*conf);
    p.ldsBytes = std::uint32_t(b->lds_bytes);
    p.coopMin = b->conf.coop_min_contacts;
    p.jointKinds = b->joint_kinds ? 1u : 0u;
    const int nw = int(b->worlds.size());
    PRX_HIP(hipEventRecord(b->ev0, b->stream));
    if (b->conf.use_graph && nsteps > 1) {
        if (!b->graph_valid || std::memcmp(&p, &b->graph_params, sizeof(p)) != 0) {
            if (b->graph_exec) (void)hipGraphExecDestroy(b->graph_exec);
            if (b->graph) (void)hipGraphDestroy(b->graph);
            b->graph_exec = nullptr;
            b->graph = nullptr;
            PRX_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
            const hipError_t le = prx_launch_step(&b->d, &p, 0, nw, b->stream);
            hipGraph_t g = nullptr;
            const hipError_t ce = hipStreamEndCapture(b->stream, &g);
            if (le != hipSuccess) return Fail(PRX_E_DEVICE, std::string("launch: ") + hipGetErrorString(le));
            if (ce != hipSuccess) return Fail(PRX_E_DEVICE, std::string("capture: ") + hipGetErrorString(ce));
            b->graph = g;
            PRX_HIP(hipGraphInstantiate(&b->graph_exec, g, nullptr, nullptr, 0));
            b->graph_params = p;
            b->graph_valid = true;
        }
        for (int s = 0; s < nsteps; ++s) PRX_HIP(hipGraphLaunch(b->graph_exec, b->stream));
    } else {
        for (int s = 0; s < nsteps; ++s) {
            const hipError_t e = prx_launch_step(&b->d, &p, 0, nw, b->stream);
            if (e != hipSuccess) return Fail(PRX_E_DEVICE, std::string("launch: ") + hipGetErrorString(e));
        }
    }
    PRX_HIP(hipEventRecord(b->ev1, b->stream));
    b->last_launches = nsteps;
    for (prx_world* w : b->worlds)
        if (w) w->host_stale = true;
    return PRX_OK;
}

int SetSweepPosition(prx_world* w, std::uint32_t body, float px, float py, float angle);

// Calls the world's listeners with the recorded calls of its last step, in
// the order the step made them.
int DispatchEvents(prx_world* w) {
    const prx_contact_listener& L = w->listener;
    const std::vector<std::uint32_t>& e = w->events;
    std::size_t o = 0;
    auto f32 = [&](std::size_t i) {
        float v;
        std::memcpy(&v, &e[i], 4);
        return v;
    };
    for (std::uint32_t i = 0; i < w->nevents; ++i) {
        if (o + 2 > e.size()) return Fail(PRX_E_DEVICE, "truncated contact listener record");
        const std::uint32_t kind = e[o], cid = e[o + 1];
        const std::size_t n = kind == prx::kEvPreSolve ? prx::kEvPreWords : kind == prx::kEvPostSolve ? prx::kEvPostWords : 2;
        if (kind > prx::kEvPostSolve || o + n > e.size()) return Fail(PRX_E_DEVICE, "malformed contact listener record");
        if (kind == prx::kEvBegin) {
            if (L.begin_contact) L.begin_contact(L.user, cid);
        } else if (kind == prx::kEvEnd) {
            if (L.end_contact) L.end_contact(L.user, cid);
        } else if (kind == prx::kEvPreSolve) {
            prx_manifold m{};
            m.type = e[o + 2] & 0xffu;
            m.point_count = (e[o + 2] >> 8) & 0xffu;
            m.local_normal[0] = f32(o + 5);
            m.local_normal[1] = f32(o + 6);
            m.local_point[0] = f32(o + 7);
            m.local_point[1] = f32(o + 8);
            for (int q = 0; q < 2; ++q) {
                m.points[q].contact_feature = e[o + 3 + std::size_t(q)];
                m.points[q].local_point[0] = f32(o + 9 + 4 * std::size_t(q));
                m.points[q].local_point[1] = f32(o + 10 + 4 * std::size_t(q));
                m.points[q].normal_impulse = f32(o + 11 + 4 * std::size_t(q));
                m.points[q].tangent_impulse = f32(o + 12 + 4 * std::size_t(q));
            }
            if (L.pre_solve) L.pre_solve(L.user, cid, &m);
        } else {
            prx_contact_impulses im{};
            im.count = e[o + 2];
            im.normal[0] = f32(o + 4);
            im.normal[1] = f32(o + 5);
            im.tangent[0] = f32(o + 6);
            im.tangent[1] = f32(o + 7);
            if (L.post_solve) L.post_solve(L.user, cid, &im, e[o + 3]);
        }
        o += n;
    }
    return PRX_OK;
}

// After a synchronised step: the listener calls of every world of the batch
// that has listeners, delivered world by world.
int DeliverEvents(prx_batch* b) {
    bool any = false;
    for (const prx_world* w : b->worlds) any = any || (w && w->has_listener);
    if (!any) return PRX_OK;
    std::vector<prx::WorldHeader> hdr;
    if (int st = Download(hdr, b->d.hdr, b->worlds.size())) return st;
    for (std::size_t wi = 0; wi < b->worlds.size(); ++wi) {
        prx_world* w = b->worlds[wi];
        if (!w || !w->has_listener) continue;
        if (int st = FetchEvents(b, int(wi), hdr[wi])) return st;
        if (int st = DispatchEvents(w)) return st;
    }
    return PRX_OK;
}

bool HasListener(const prx_batch* b) {
    for (const prx_world* w : b->worlds)
        if (w && w->has_listener) return true;
    return false;
}

std::vector<std::pair<void*, std::size_t>> DynamicArrays(prx_batch* b);

// One step of a batch whose worlds record listener calls.  A step whose calls
// exceed a world's call record (the TOI phase's calls have no fixed bound per
// contact slot) is stepped again from a device copy of the state taken before
// it, with that world's record grown to what the step needed; the calls
// delivered are then the complete list.
int StepRecorded(prx_batch* b, const prx_step_conf* conf) {
    if (b->totE == 0) {
        if (int st = StepBatch(b, conf, 1)) return st;
        PRX_HIP(hipStreamSynchronize(b->stream));
        return CheckStatus(b);
    }
    std::vector<std::pair<void*, std::size_t>> arrays = DynamicArrays(b);
    arrays.pop_back();  // the previous step's calls (the last entry) need no copy
    if (b->snapshot.empty()) {
        for (const auto& a : arrays) {
            unsigned char* p = nullptr;
            PRX_HIP(b->arena.alloc(p, a.second));
            b->snapshot.push_back(p);
        }
    }
    for (int attempt = 0;; ++attempt) {
        for (std::size_t i = 0; i < arrays.size(); ++i)
            PRX_HIP(hipMemcpyAsync(b->snapshot[i], arrays[i].first, arrays[i].second, hipMemcpyDeviceToDevice,
                                   b->stream));
        if (int st = StepBatch(b, conf, 1)) return st;
        PRX_HIP(hipStreamSynchronize(b->stream));
        std::vector<prx::WorldHeader> hdr;
        if (int st = Download(hdr, b->d.hdr, b->worlds.size())) return st;
        bool over = false;
        for (const auto& h : hdr) over = over || (h.status & prx::kStatusEventOverflow) != 0;
        if (!over || attempt >= 8) return CheckStatus(b);
        // Back to the state before the step, then a larger record for every
        // world that ran out (new offsets for all: the segments are contiguous).
        for (std::size_t i = 0; i < arrays.size(); ++i)
            PRX_HIP(hipMemcpyAsync(arrays[i].first, b->snapshot[i], arrays[i].second, hipMemcpyDeviceToDevice,
                                   b->stream));
        PRX_HIP(hipStreamSynchronize(b->stream));
        std::vector<prx::WorldHeader> cur;
        if (int st = Download(cur, b->d.hdr, b->worlds.size())) return st;
        std::size_t E = 0;
        for (std::size_t i = 0; i < cur.size(); ++i) {
            int cap = cur[i].ev_cap;
            if (cap > 0 && (hdr[i].status & prx::kStatusEventOverflow))
                cap = std::max(2 * cap, hdr[i].ev_words + 16);
            cur[i].ev_off = int(E);
            cur[i].ev_cap = cap;
            b->hdr[i].ev_off = cur[i].ev_off;
            b->hdr[i].ev_cap = cur[i].ev_cap;
            E += std::size_t(cap);
            if (E > std::size_t(0x7fffffff)) return Fail(PRX_E_LENGTH, "contact listener records exceed 2^31 words");
        }
        PRX_HIP(b->arena.alloc(b->d.ev, E));
        b->totE = E;
        if (int st = Upload(b->d.hdr, cur)) return st;
        b->graph_valid = false;  // the graph holds the old record's address
    }
}

}  // namespace

extern "C" {

int prx_abi_version(void) { return PRX_ABI_VERSION; }
const char* prx_last_error(void) { return g_error.c_str(); }

void prx_step_conf_default(prx_step_conf* c) {
    // Bit-exact values of a default StepConf in the reference float build.
    c->delta_time = 1.0f / 60.0f;
    c->dt_ratio = 1.0f;
    c->min_still_time_to_sleep = 0.5f;
    c->linear_slop = 0.005f;
    c->angular_slop = (3.14159265358979323846f * 2.0f) / 180.0f;
    c->reg_resolution_rate = 2.0f / 10.0f;
    c->reg_min_separation = -0.005f * 3.0f;
    c->reg_min_momentum = 0.0f / 100.0f;
    c->velocity_threshold = 1.0f;
    c->max_translation = 2.0f;
    c->max_rotation = float(179.0L * (3.14159265358979323846264338327950288L / 180.0L));
    c->max_linear_correction = 0.2f;
    c->max_angular_correction = (8.0f / 180.0f) * 3.14159265358979323846f;
    c->linear_sleep_tolerance = 0.01f;
    c->angular_sleep_tolerance = (3.14159265358979323846f * 2.0f) / 180.0f;
    c->displace_multiplier = 2.0f;
    c->aabb_extension = 0.005f * 20.0f;
    c->max_circles_ratio = 10.0f;
    c->reg_velocity_iters = 8;
    c->reg_position_iters = 3;
    c->do_warm_start = 1;
    c->do_toi = 1;
    c->do_blocksolve = 1;
    c->toi_resolution_rate = 75.0f / 100.0f;
    c->toi_min_separation = -0.005f * 1.5f;
    c->toi_min_momentum = 0.0f / 100.0f;
    c->target_depth = 0.005f * 3.0f;
    c->tolerance = 0.005f / 4.0f;
    c->toi_velocity_iters = 8;
    c->toi_position_iters = 20;
    c->max_toi_root_iters = 30;
    c->max_toi_iters = 20;
    c->max_distance_iters = 20;
    c->max_sub_steps = 8;
}

void prx_world_conf_default(prx_world_conf* c) {
    c->abi_version = PRX_ABI_VERSION;
    c->min_vertex_radius = 0.01f;
    c->max_vertex_radius = 255.0f;
    c->contact_capacity_per_proxy = 8;
}

void prx_batch_conf_default(prx_batch_conf* c) {
    c->threads_per_world = 64;
    c->use_graph = 1;
    c->order = 0;
    c->coop_min_contacts = 32;
    c->lds_bytes = 0;
}

int prx_world_create(const prx_world_conf* conf, prx_world** out) {
    if (!out) return Fail(PRX_E_INVALID_ARGUMENT, "null out");
    auto* w = new prx_world;
    if (conf) w->conf = *conf;
    else prx_world_conf_default(&w->conf);
    *out = w;
    return PRX_OK;
}

void prx_world_destroy(prx_world* w) {
    if (!w) return;
    if (w->owner) {
        prx_batch* b = w->owner;
        for (auto& m : b->worlds)
            if (m == w) m = nullptr;
        if (w->solo == b) delete b;
    }
    delete w;
}

int prx_shape_create(prx_world* w, const prx_shape_conf* conf, std::uint32_t* id) {
    if (int st = CheckWorld(w)) return st;
    if (!conf) return Fail(PRX_E_INVALID_ARGUMENT, "null shape conf");
    if (int st = Detach(w)) return st;
    return AddShape(w, *conf, id);
}

int prx_compound_shape_create(prx_world* w, const prx_shape_conf* conf, const std::uint32_t* hull_sizes,
                              std::uint32_t nhulls, std::uint32_t* first_shape, std::uint32_t* child_count) {
    if (int st = CheckWorld(w)) return st;
    if (!conf) return Fail(PRX_E_INVALID_ARGUMENT, "null shape conf");
    if (int st = Detach(w)) return st;
    return AddCompound(w, *conf, hull_sizes, nhulls, first_shape, child_count);
}

int prx_body_create(prx_world* w, const prx_body_conf* conf, std::uint32_t* id) {
    if (int st = CheckWorld(w)) return st;
    if (!conf) return Fail(PRX_E_INVALID_ARGUMENT, "null body conf");
    if (int st = Detach(w)) return st;
    return AddBody(w, *conf, id);
}

int prx_scene_build(prx_world* w, const char* name, const std::int32_t* params, int nparams) {
    if (int st = CheckWorld(w)) return st;
    if (int st = Detach(w)) return st;
    prx_scene::SceneSpec spec;
    if (!name || !prx_scene::Make(name, params, nparams, spec)) return Fail(PRX_E_INVALID_ARGUMENT, "unknown scene");
    // Shapes are created right before the first body using them, matching the
    // order of the golden-vector driver.
    std::vector<std::uint32_t> ids(spec.shapes.size(), 0xffffffffu), counts(spec.shapes.size(), 1u);
    for (const auto& bs : spec.bodies) {
        std::vector<std::uint32_t> bsh;
        for (int si : bs.shapes) {
            if (ids[std::size_t(si)] == 0xffffffffu) {
                const auto& s = spec.shapes[std::size_t(si)];
                prx_shape_conf c{};
                c.kind = std::uint32_t(s.kind);
                c.vertex_radius = s.vertexRadius;
                c.radius = s.radius;
                c.hx = s.hx;
                c.hy = s.hy;
                std::vector<float> pts;
                for (const auto& p : s.pts) {
                    pts.push_back(p.x);
                    pts.push_back(p.y);
                }
                c.points = pts.data();
                c.npoints = std::uint32_t(s.pts.size());
                c.density = s.density;
                c.friction = s.friction;
                c.restitution = s.restitution;
                c.category_bits = 1;
                c.mask_bits = 0xffff;
                if (s.kind == prx_scene::kChain || s.kind == prx_scene::kMulti) {
                    std::vector<std::uint32_t> hs(s.hullSizes.begin(), s.hullSizes.end());
                    if (int st = AddCompound(w, c, hs.data(), std::uint32_t(hs.size()), &ids[std::size_t(si)],
                                             &counts[std::size_t(si)]))
                        return st;
                } else if (int st = AddShape(w, c, &ids[std::size_t(si)])) {
                    return st;
                }
            }
            for (std::uint32_t k = 0; k < counts[std::size_t(si)]; ++k) bsh.push_back(ids[std::size_t(si)] + k);
        }
        prx_body_conf c{};
        c.type = std::uint32_t(bs.type);
        c.location[0] = bs.x;
        c.location[1] = bs.y;
        c.angle = bs.angle;
        c.linear_velocity[0] = bs.vx;
        c.linear_velocity[1] = bs.vy;
        c.angular_velocity = bs.w;
        c.linear_acceleration[0] = bs.ax;
        c.linear_acceleration[1] = bs.ay;
        c.allow_sleep = 1;
        c.awake = 1;
        c.enabled = 1;
        c.shapes = bsh.data();
        c.nshapes = std::uint32_t(bsh.size());
        if (int st = AddBody(w, c, nullptr)) return st;
    }
    for (const auto& js : spec.joints) {
        if (js.kind == prx_scene::kPulley) {
            prx_pulley_joint_conf pc{};
            pc.body_a = std::uint32_t(js.a);
            pc.body_b = std::uint32_t(js.b);
            pc.ground_anchor_a[0] = js.gax;
            pc.ground_anchor_a[1] = js.gay;
            pc.ground_anchor_b[0] = js.gbx;
            pc.ground_anchor_b[1] = js.gby;
            pc.local_anchor_a[0] = js.lax;
            pc.local_anchor_a[1] = js.lay;
            pc.local_anchor_b[0] = js.lbx;
            pc.local_anchor_b[1] = js.lby;
            pc.length_a = js.length;
            pc.length_b = js.lengthB;
            pc.ratio = js.ratio;
            pc.collide_connected = js.collideConnected ? 1u : 0u;
            if (int st = AddPulleyJoint(w, pc, nullptr)) return st;
            continue;
        }
        if (js.kind == prx_scene::kGear) {
            prx_gear_joint_conf gc{};
            if (int st = GearConf(w, std::uint32_t(js.a), std::uint32_t(js.b), js.ratio, gc)) return st;
            gc.collide_connected = js.collideConnected ? 1u : 0u;
            if (int st = AddGearJoint(w, gc, nullptr)) return st;
            continue;
        }
        if (js.kind == prx_scene::kTarget) {
            prx_target_joint_conf tc{};
            tc.body = std::uint32_t(js.b);
            tc.target[0] = js.gax;
            tc.target[1] = js.gay;
            tc.local_anchor[0] = js.lbx;
            tc.local_anchor[1] = js.lby;
            tc.max_force = js.maxForce;
            tc.frequency = js.frequency;
            tc.damping_ratio = js.dampingRatio;
            tc.collide_connected = js.collideConnected ? 1u : 0u;
            if (int st = AddTargetJoint(w, tc, nullptr)) return st;
            continue;
        }
        if (js.kind == prx_scene::kPrismatic) {
            prx_prismatic_joint_conf pc{};
            pc.body_a = std::uint32_t(js.a);
            pc.body_b = std::uint32_t(js.b);
            pc.local_anchor_a[0] = js.lax;
            pc.local_anchor_a[1] = js.lay;
            pc.local_anchor_b[0] = js.lbx;
            pc.local_anchor_b[1] = js.lby;
            const V2 axis = UnitFromAngle(js.axisAngle);
            pc.local_axis_a[0] = axis.x;
            pc.local_axis_a[1] = axis.y;
            pc.reference_angle = js.referenceAngle;
            pc.enable_limit = js.enableLimit ? 1u : 0u;
            pc.lower_translation = js.lower;
            pc.upper_translation = js.upper;
            pc.enable_motor = js.enableMotor ? 1u : 0u;
            pc.motor_speed = js.motorSpeed;
            pc.max_motor_force = js.maxForce;
            pc.collide_connected = js.collideConnected ? 1u : 0u;
            if (int st = AddPrismaticJoint(w, pc, nullptr)) return st;
            continue;
        }
        if (js.kind == prx_scene::kWheel) {
            prx_wheel_joint_conf wc{};
            wc.body_a = std::uint32_t(js.a);
            wc.body_b = std::uint32_t(js.b);
            wc.local_anchor_a[0] = js.lax;
            wc.local_anchor_a[1] = js.lay;
            wc.local_anchor_b[0] = js.lbx;
            wc.local_anchor_b[1] = js.lby;
            const V2 axis = UnitFromAngle(js.axisAngle);
            wc.local_axis_a[0] = axis.x;
            wc.local_axis_a[1] = axis.y;
            wc.enable_motor = js.enableMotor ? 1u : 0u;
            wc.motor_speed = js.motorSpeed;
            wc.max_motor_torque = js.maxMotorTorque;
            wc.frequency = js.frequency;
            wc.damping_ratio = js.dampingRatio;
            wc.collide_connected = js.collideConnected ? 1u : 0u;
            if (int st = AddWheelJoint(w, wc, nullptr)) return st;
            continue;
        }
        if (js.kind == prx_scene::kMotor) {
            prx_motor_joint_conf mc{};
            mc.body_a = std::uint32_t(js.a);
            mc.body_b = std::uint32_t(js.b);
            mc.linear_offset[0] = js.lax;
            mc.linear_offset[1] = js.lay;
            mc.angular_offset = js.referenceAngle;
            mc.max_force = js.maxForce;
            mc.max_torque = js.maxTorque;
            mc.correction_factor = js.correctionFactor;
            mc.collide_connected = js.collideConnected ? 1u : 0u;
            if (int st = AddMotorJoint(w, mc, nullptr)) return st;
            continue;
        }
        if (js.kind == prx_scene::kRope || js.kind == prx_scene::kFriction) {
            HostJoint j;
            j.a = std::uint32_t(js.a);
            j.b = std::uint32_t(js.b);
            const bool rope = js.kind == prx_scene::kRope;
            j.flags = (js.collideConnected ? prx::kJointCollideConnected : 0u) |
                      ((rope ? prx::kJointRope : prx::kJointFriction) << 8);
            j.la[0] = js.lax;
            j.la[1] = js.lay;
            j.la[2] = js.lbx;
            j.la[3] = js.lby;
            j.cfg[0] = rope ? js.length : js.maxForce;
            j.cfg[1] = rope ? 0.0f : js.maxTorque;
            if (int st = AddJointRecord(w, j, nullptr)) return st;
            continue;
        }
        if (js.kind != prx_scene::kRevolute) {
            const float la[2] = {js.lax, js.lay}, lb[2] = {js.lbx, js.lby};
            const bool dist = js.kind == prx_scene::kDistance;
            if (int st = AddSpringJoint(w, dist ? prx::kJointDistance : prx::kJointWeld, std::uint32_t(js.a),
                                        std::uint32_t(js.b), la, lb, dist ? js.length : js.referenceAngle,
                                        js.frequency, js.dampingRatio, js.collideConnected, nullptr))
                return st;
            continue;
        }
        prx_revolute_joint_conf jc{};
        jc.body_a = std::uint32_t(js.a);
        jc.body_b = std::uint32_t(js.b);
        jc.local_anchor_a[0] = js.lax;
        jc.local_anchor_a[1] = js.lay;
        jc.local_anchor_b[0] = js.lbx;
        jc.local_anchor_b[1] = js.lby;
        jc.reference_angle = js.referenceAngle;
        jc.enable_limit = js.enableLimit ? 1u : 0u;
        jc.lower_angle = js.lower;
        jc.upper_angle = js.upper;
        jc.enable_motor = js.enableMotor ? 1u : 0u;
        jc.motor_speed = js.motorSpeed;
        jc.max_motor_torque = js.maxMotorTorque;
        jc.collide_connected = js.collideConnected ? 1u : 0u;
        if (int st = AddJoint(w, jc, nullptr)) return st;
    }
    return PRX_OK;
}

int prx_joint_create(prx_world* w, const prx_revolute_joint_conf* conf, std::uint32_t* id) {
    if (int st = CheckWorld(w)) return st;
    if (!conf) return Fail(PRX_E_INVALID_ARGUMENT, "null joint conf");
    if (int st = Detach(w)) return st;
    return AddJoint(w, *conf, id);
}

int prx_distance_joint_create(prx_world* w, const prx_distance_joint_conf* conf, std::uint32_t* id) {
    if (int st = CheckWorld(w)) return st;
    if (!conf) return Fail(PRX_E_INVALID_ARGUMENT, "null joint conf");
    if (int st = Detach(w)) return st;
    return AddSpringJoint(w, prx::kJointDistance, conf->body_a, conf->body_b, conf->local_anchor_a,
                          conf->local_anchor_b, conf->length, conf->frequency, conf->damping_ratio,
                          conf->collide_connected != 0, id);
}

int prx_pulley_joint_create(prx_world* w, const prx_pulley_joint_conf* conf, std::uint32_t* id) {
    if (int st = CheckWorld(w)) return st;
    if (!conf) return Fail(PRX_E_INVALID_ARGUMENT, "null joint conf");
    if (int st = Detach(w)) return st;
    return AddPulleyJoint(w, *conf, id);
}

int prx_target_joint_create(prx_world* w, const prx_target_joint_conf* conf, std::uint32_t* id) {
    if (int st = CheckWorld(w)) return st;
    if (!conf) return Fail(PRX_E_INVALID_ARGUMENT, "null joint conf");
    if (int st = Detach(w)) return st;
    return AddTargetJoint(w, *conf, id);
}

int prx_gear_joint_conf_get(prx_world* w, std::uint32_t joint1, std::uint32_t joint2, float ratio,
                            prx_gear_joint_conf* out) {
    if (int st = CheckWorld(w)) return st;
    if (!out) return Fail(PRX_E_INVALID_ARGUMENT, "null conf");
    if (int st = Refresh(w)) return st;
    return GearConf(w, joint1, joint2, ratio, *out);
}

int prx_gear_joint_create(prx_world* w, const prx_gear_joint_conf* conf, std::uint32_t* id) {
    if (int st = CheckWorld(w)) return st;
    if (!conf) return Fail(PRX_E_INVALID_ARGUMENT, "null conf");
    if (int st = Detach(w)) return st;
    return AddGearJoint(w, *conf, id);
}

int prx_prismatic_joint_create(prx_world* w, const prx_prismatic_joint_conf* conf, std::uint32_t* id) {
    if (int st = CheckWorld(w)) return st;
    if (!conf) return Fail(PRX_E_INVALID_ARGUMENT, "null joint conf");
    if (int st = Detach(w)) return st;
    return AddPrismaticJoint(w, *conf, id);
}

int prx_wheel_joint_create(prx_world* w, const prx_wheel_joint_conf* conf, std::uint32_t* id) {
    if (int st = CheckWorld(w)) return st;
    if (!conf) return Fail(PRX_E_INVALID_ARGUMENT, "null joint conf");
    if (int st = Detach(w)) return st;
    return AddWheelJoint(w, *conf, id);
}

int prx_motor_joint_create(prx_world* w, const prx_motor_joint_conf* conf, std::uint32_t* id) {
    if (int st = CheckWorld(w)) return st;
    if (!conf) return Fail(PRX_E_INVALID_ARGUMENT, "null joint conf");
    if (int st = Detach(w)) return st;
    return AddMotorJoint(w, *conf, id);
}

int prx_rope_joint_create(prx_world* w, const prx_rope_joint_conf* conf, std::uint32_t* id) {
    if (int st = CheckWorld(w)) return st;
    if (!conf) return Fail(PRX_E_INVALID_ARGUMENT, "null joint conf");
    if (int st = Detach(w)) return st;
    HostJoint j;
    j.a = conf->body_a;
    j.b = conf->body_b;
    j.flags = (conf->collide_connected ? prx::kJointCollideConnected : 0u) | (prx::kJointRope << 8);
    j.la[0] = conf->local_anchor_a[0];
    j.la[1] = conf->local_anchor_a[1];
    j.la[2] = conf->local_anchor_b[0];
    j.la[3] = conf->local_anchor_b[1];
    j.cfg[0] = conf->max_length;
    return AddJointRecord(w, j, id);
}

int prx_friction_joint_create(prx_world* w, const prx_friction_joint_conf* conf, std::uint32_t* id) {
    if (int st = CheckWorld(w)) return st;
    if (!conf) return Fail(PRX_E_INVALID_ARGUMENT, "null joint conf");
    if (!(conf->max_force >= 0.0f) || !(conf->max_torque >= 0.0f))
        return Fail(PRX_E_INVALID_ARGUMENT, "maximum force and torque must be non-negative");
    if (int st = Detach(w)) return st;
    HostJoint j;
    j.a = conf->body_a;
    j.b = conf->body_b;
    j.flags = (conf->collide_connected ? prx::kJointCollideConnected : 0u) | (prx::kJointFriction << 8);
    j.la[0] = conf->local_anchor_a[0];
    j.la[1] = conf->local_anchor_a[1];
    j.la[2] = conf->local_anchor_b[0];
    j.la[3] = conf->local_anchor_b[1];
    j.cfg[0] = conf->max_force;
    j.cfg[1] = conf->max_torque;
    return AddJointRecord(w, j, id);
}

int prx_weld_joint_create(prx_world* w, const prx_weld_joint_conf* conf, std::uint32_t* id) {
    if (int st = CheckWorld(w)) return st;
    if (!conf) return Fail(PRX_E_INVALID_ARGUMENT, "null joint conf");
    if (int st = Detach(w)) return st;
    return AddSpringJoint(w, prx::kJointWeld, conf->body_a, conf->body_b, conf->local_anchor_a, conf->local_anchor_b,
                          conf->reference_angle, conf->frequency, conf->damping_ratio, conf->collide_connected != 0,
                          id);
}

std::uint32_t prx_world_joint_count(const prx_world* w) { return w ? std::uint32_t(w->joints.size()) : 0u; }

int prx_world_load_desc(prx_world* w, const std::uint8_t* desc, std::size_t len) {
    if (int st = CheckWorld(w)) return st;
    if (!desc) return Fail(PRX_E_INVALID_ARGUMENT, "null description");
    if (int st = Detach(w)) return st;
    std::size_t o = 0;
    bool bad = false;
    auto u32 = [&]() {
        std::uint32_t v = 0;
        if (o + 4 > len) {
            bad = true;
            return v;
        }
        std::memcpy(&v, desc + o, 4);
        o += 4;
        return v;
    };
    auto f32 = [&]() {
        float v = 0;
        if (o + 4 > len) {
            bad = true;
            return v;
        }
        std::memcpy(&v, desc + o, 4);
        o += 4;
        return v;
    };
    prx_world fresh;
    fresh.conf = w->conf;
    const std::uint32_t ns = u32();
    for (std::uint32_t i = 0; i < ns && !bad; ++i) {
        HostShape s;
        const std::uint32_t nv = u32();
        if (nv == 0 || nv > std::uint32_t(kMaxShapeVertices)) return Fail(PRX_E_UNSUPPORTED, "bad vertex count");
        s.radius = f32();
        s.density = f32();
        s.friction = f32();
        s.restitution = f32();
        s.sensor = u32();
        s.catmask = u32();
        s.group = std::int32_t(u32());
        for (std::uint32_t k = 0; k < nv; ++k) {
            const float x = f32();
            const float y = f32();
            s.v.push_back(V2{x, y});
        }
        for (std::uint32_t k = 0; k < nv; ++k) {
            const float x = f32();
            const float y = f32();
            s.n.push_back(V2{x, y});
        }
        fresh.shapes.push_back(std::move(s));
    }
    const std::uint32_t nb = u32();
    for (std::uint32_t i = 0; i < nb && !bad; ++i) {
        HostBody b;
        b.type = u32();
        b.flags = u32();
        float f[19];
        for (auto& x : f) x = f32();
        b.pos0[0] = f[0];
        b.pos0[1] = f[1];
        b.pos0[2] = f[2];
        b.pos1[0] = f[3];
        b.pos1[1] = f[4];
        b.pos1[2] = f[5];
        b.lc = V2{f[6], f[7]};
        b.vel[0] = f[8];
        b.vel[1] = f[9];
        b.vel[2] = f[10];
        b.acc[0] = f[11];
        b.acc[1] = f[12];
        b.acc[2] = f[13];
        b.damp[0] = f[14];
        b.damp[1] = f[15];
        b.invMass = f[16];
        b.invRotI = f[17];
        b.under = f[18];
        const std::uint32_t nsh = u32();
        for (std::uint32_t k = 0; k < nsh && !bad; ++k) {
            const std::uint32_t sid = u32();
            if (sid >= fresh.shapes.size()) return Fail(PRX_E_OUT_OF_RANGE, "no such shape");
            b.shapes.push_back(sid);
        }
        const int bid = int(fresh.bodies.size());
        if (b.flags & prx::kEnabled) {
            for (std::uint32_t sid : b.shapes) {
                HostProxy p;
                p.body = bid;
                p.shape = int(sid);
                p.id = (fresh.leafCount == 0) ? 0u : 2u * fresh.leafCount - 1u;
                ++fresh.leafCount;
                b.proxies.push_back(int(fresh.proxies.size()));
                fresh.proxies.push_back(p);
            }
        }
        fresh.bodies.push_back(std::move(b));
    }
    const std::uint32_t nj = u32();
    for (std::uint32_t i = 0; i < nj && !bad; ++i) {
        HostJoint j;
        j.a = u32();
        j.b = u32();
        const std::uint32_t rawFlags = u32();
        j.flags = rawFlags & 0xff07u;
        if ((j.flags >> 8) > prx::kJointGear) return Fail(PRX_E_INVALID_ARGUMENT, "unknown joint kind");
        if ((j.a >= fresh.bodies.size() && !(j.a == kNoBody && (j.flags >> 8) == prx::kJointTarget)) ||
            j.b >= fresh.bodies.size())
            return Fail(PRX_E_OUT_OF_RANGE, "no such body");
        // Limits belong to revolute and prismatic joints, motors to those and wheels.
        const std::uint32_t kind = j.flags >> 8;
        if (kind && kind != prx::kJointPrismatic)
            j.flags &= ~(prx::kJointLimit | (kind == prx::kJointWheel ? 0u : prx::kJointMotor));
        if (kind == prx::kJointGear) {
            // Gear: the sides' 12 floats, then bodies C, D, ratio, constant;
            // prismatic sides in bits 16 (A-C) and 17 (B-D) of the flags.
            for (float& v : j.gear) v = f32();
            j.c = u32();
            j.d = u32();
            j.ratio = f32();
            j.constant = f32();
            if (j.c >= fresh.bodies.size() || j.d >= fresh.bodies.size())
                return Fail(PRX_E_OUT_OF_RANGE, "no such body");
            j.flags |= ((rawFlags >> 16) & 1u ? prx::kGearPrisAC : 0u) | ((rawFlags >> 17) & 1u ? prx::kGearPrisBD : 0u);
        } else {
            for (float& v : j.la) v = f32();
            for (float& v : j.cfg) v = f32();
            j.maxTorque = f32();
            for (float& v : j.ext) v = f32();
        }
        j.seq = fresh.joints.size();
        fresh.joints.push_back(j);
    }
    if (bad) return Fail(PRX_E_INVALID_ARGUMENT, "truncated description");
    w->shapes = std::move(fresh.shapes);
    w->bodies = std::move(fresh.bodies);
    w->joints = std::move(fresh.joints);
    w->proxies = std::move(fresh.proxies);
    w->contacts = HostContacts{};
    w->leafCount = fresh.leafCount;
    w->joint_free.clear();
    w->joint_seq = w->joints.size();
    w->destroyed_bodies = 0;
    w->pre_events.clear();
    w->pre_nevents = 0;
    return PRX_OK;
}

int prx_world_save_desc(prx_world* w, std::uint8_t* buf, std::size_t cap, std::size_t* len) {
    if (int st = CheckWorld(w)) return st;
    if (int st = Refresh(w)) return st;
    // The description has no place for destroyed bodies and joints.
    if (w->destroyed_bodies || !w->joint_free.empty())
        return Fail(PRX_E_UNSUPPORTED, "saving the description of a world with destroyed bodies or joints");
    std::vector<std::uint8_t> o;
    auto put = [&](const void* p, std::size_t n) {
        const auto* c = static_cast<const std::uint8_t*>(p);
        o.insert(o.end(), c, c + n);
    };
    auto u32 = [&](std::uint32_t v) { put(&v, 4); };
    auto f32 = [&](float v) { put(&v, 4); };
    u32(std::uint32_t(w->shapes.size()));
    for (const auto& s : w->shapes) {
        u32(std::uint32_t(s.v.size()));
        f32(s.radius);
        f32(s.density);
        f32(s.friction);
        f32(s.restitution);
        u32(s.sensor);
        u32(s.catmask);
        u32(std::uint32_t(s.group));
        for (const auto& v : s.v) {
            f32(v.x);
            f32(v.y);
        }
        for (const auto& v : s.n) {
            f32(v.x);
            f32(v.y);
        }
    }
    u32(std::uint32_t(w->bodies.size()));
    for (const auto& b : w->bodies) {
        u32(b.type);
        u32(b.flags & 0x7fu);
        const float f[19] = {b.pos0[0], b.pos0[1], b.pos0[2], b.pos1[0], b.pos1[1], b.pos1[2], b.lc.x,
                             b.lc.y,    b.vel[0],  b.vel[1],  b.vel[2],  b.acc[0],  b.acc[1],  b.acc[2],
                             b.damp[0], b.damp[1], b.invMass, b.invRotI, b.under};
        put(f, sizeof(f));
        u32(std::uint32_t(b.shapes.size()));
        for (auto s : b.shapes) u32(s);
    }
    u32(std::uint32_t(w->joints.size()));
    for (const auto& j : w->joints) {
        u32(j.a);
        u32(j.b);
        if ((j.flags >> 8) == prx::kJointGear) {
            u32((j.flags & 0xff07u) | ((j.flags & prx::kGearPrisAC) ? 1u << 16 : 0u) |
                ((j.flags & prx::kGearPrisBD) ? 1u << 17 : 0u));
            put(j.gear, sizeof(j.gear));
            u32(j.c);
            u32(j.d);
            f32(j.ratio);
            f32(j.constant);
            continue;
        }
        u32(j.flags);
        put(j.la, sizeof(j.la));
        put(j.cfg, sizeof(j.cfg));
        f32(j.maxTorque);
        put(j.ext, sizeof(j.ext));
    }
    if (len) *len = o.size();
    if (buf && cap >= o.size()) std::memcpy(buf, o.data(), o.size());
    else if (buf) return Fail(PRX_E_LENGTH, "buffer too small");
    return PRX_OK;
}

std::uint32_t prx_world_body_count(const prx_world* w) { return w ? std::uint32_t(w->bodies.size()) : 0u; }
std::uint32_t prx_world_shape_count(const prx_world* w) { return w ? std::uint32_t(w->shapes.size()) : 0u; }

int prx_step(prx_world* w, const prx_step_conf* conf, prx_step_stats* stats) {
    if (int st = CheckWorld(w)) return st;
    if (!conf) return Fail(PRX_E_INVALID_ARGUMENT, "null step conf");
    if (w->owner && w->owner != w->solo) {
        if (int st = Detach(w)) return st;
    }
    if (!w->solo) {
        prx_batch_conf bc;
        prx_batch_conf_default(&bc);
        bc.use_graph = 0;
        prx_batch* b = nullptr;
        prx_world* one[1] = {w};
        if (int st = CreateBatch(one, 1, &bc, &b)) return st;
        w->solo = b;
    }
    if (int st = StepRecorded(w->solo, conf)) return st;
    if (int st = DeliverEvents(w->solo)) return st;
    if (stats) {
        prx::WorldHeader h;
        PRX_HIP(hipMemcpy(&h, w->solo->d.hdr, sizeof(h), hipMemcpyDeviceToHost));
        *stats = StatsOf(h);
    }
    return PRX_OK;
}

int prx_world_get_bodies(prx_world* w, float* out, std::uint32_t* flags) {
    if (int st = CheckWorld(w)) return st;
    if (int st = Refresh(w)) return st;
    for (std::size_t i = 0; i < w->bodies.size(); ++i) {
        const HostBody& b = w->bodies[i];
        if (out) {
            float* o = out + 10 * i;
            o[0] = b.pos0[0];
            o[1] = b.pos0[1];
            o[2] = b.pos0[2];
            o[3] = b.pos1[0];
            o[4] = b.pos1[1];
            o[5] = b.pos1[2];
            o[6] = b.vel[0];
            o[7] = b.vel[1];
            o[8] = b.vel[2];
            o[9] = b.under;
        }
        if (flags) flags[i] = b.flags & 0x7fu;
    }
    return PRX_OK;
}

int prx_world_get_transformation(prx_world* w, std::uint32_t body, float out[4]) {
    if (int st = CheckWorld(w)) return st;
    if (int st = CheckBody(w, body)) return st;
    if (int st = Refresh(w)) return st;
    const HostBody& b = w->bodies[body];
    const V2 q = UnitFromAngle(b.pos1[2]);
    const V2 p = V2{b.pos1[0], b.pos1[1]} - Rotate(b.lc, q);
    out[0] = p.x;
    out[1] = p.y;
    out[2] = q.x;
    out[3] = q.y;
    return PRX_OK;
}

int prx_world_set_transformation(prx_world* w, std::uint32_t body, const float xf[4]) {
    if (int st = CheckWorld(w)) return st;
    if (!xf) return Fail(PRX_E_INVALID_ARGUMENT, "null transformation");
    if (int st = CheckBody(w, body)) return st;
    if (int st = Detach(w)) return st;
    // SetTransformation(Body&, xfm) (Body.cpp:270-273): the angle of q.
    return SetSweepPosition(w, body, xf[0], xf[1], std::atan2(xf[3], xf[2]));
}

int prx_world_set_location(prx_world* w, std::uint32_t body, const float v[2]) {
    if (int st = CheckWorld(w)) return st;
    if (!v) return Fail(PRX_E_INVALID_ARGUMENT, "null location");
    if (int st = CheckBody(w, body)) return st;
    if (int st = Detach(w)) return st;
    // SetLocation(Body&, v) (Body.cpp:275-278): the current rotation, re-read as an angle.
    const V2 q = UnitFromAngle(w->bodies[body].pos1[2]);
    return SetSweepPosition(w, body, v[0], v[1], std::atan2(q.y, q.x));
}

int prx_world_set_angle(prx_world* w, std::uint32_t body, float angle) {
    if (int st = CheckWorld(w)) return st;
    if (int st = CheckBody(w, body)) return st;
    if (int st = Detach(w)) return st;
    // SetAngle(Body&, a) (Body.cpp:285-288): the sweep's linear position kept.
    const HostBody& b = w->bodies[body];
    return SetSweepPosition(w, body, b.pos1[0], b.pos1[1], angle);
}

int prx_world_shift_origin(prx_world* w, const float origin[2]) {
    if (int st = CheckWorld(w)) return st;
    if (!origin) return Fail(PRX_E_INVALID_ARGUMENT, "null origin");
    if (int st = Detach(w)) return st;
    // ShiftOrigin (AabbTreeWorld.cpp:2221-2242): sweeps, every contact flagged
    // for updating, and the tree's AABBs (DynamicTree.cpp:650-658).  Pulleys
    // move their ground anchors (PulleyJointConf.cpp:255-260).
    for (HostJoint& j : w->joints) {
        if ((j.flags >> 8) != prx::kJointPulley) continue;
        for (int k = 0; k < 4; ++k) j.cfg[k] = j.cfg[k] - origin[k & 1];
    }
    for (HostBody& b : w->bodies) {
        for (int k = 0; k < 2; ++k) {
            b.pos0[k] = b.pos0[k] - origin[k];
            b.pos1[k] = b.pos1[k] - origin[k];
        }
    }
    HostContacts& hc = w->contacts;
    for (std::size_t i = 0; i < hc.size(); ++i) hc.flags[i] |= prx::kNeedsUpdating;
    for (HostProxy& p : w->proxies) {
        if (p.flags & prx::kProxyPending) continue;
        for (int k = 0; k < 2; ++k) {
            p.aabb[k] = p.aabb[k] + (-origin[k]);
            p.aabb[2 + k] = p.aabb[2 + k] + (-origin[k]);
        }
    }
    return PRX_OK;
}

int prx_world_set_velocity(prx_world* w, std::uint32_t body, const float v[3]) {
    if (int st = CheckWorld(w)) return st;
    if (!v) return Fail(PRX_E_INVALID_ARGUMENT, "null velocity");
    if (int st = CheckBody(w, body)) return st;
    if (int st = Detach(w)) return st;
    HostBody& b = w->bodies[body];
    const bool nonzero = v[0] != 0 || v[1] != 0 || v[2] != 0;
    if (nonzero) {
        if (!(b.flags & prx::kSpeedable)) return PRX_OK;
        b.flags |= prx::kAwake;
        b.under = 0;
    }
    b.vel[0] = v[0];
    b.vel[1] = v[1];
    b.vel[2] = v[2];
    return PRX_OK;
}

// Destroy(World&, JointID) (AabbTreeWorld.cpp:1215-1227); an id already
// destroyed is left alone, as the reference leaves an id not in its joint list.
int prx_world_destroy_joint(prx_world* w, std::uint32_t joint) {
    if (int st = CheckWorld(w)) return st;
    if (joint >= w->joints.size()) return Fail(PRX_E_OUT_OF_RANGE, "no such joint");
    if (w->joints[joint].destroyed) return PRX_OK;
    if (int st = Detach(w)) return st;
    DestroyJointRecord(w, joint);
    return PRX_OK;
}

// Destroy(World&, BodyID) (AabbTreeWorld.cpp:1062-1104): the body's joints
// (the first of its list each time), then its contacts in creation order
// (AabbTreeWorld::Destroy(ContactID, from) :2281-2288 and InternalDestroy
// :2245-2279: an end-contact call for a touching one, both bodies awake if
// the manifold has points, the id to the contact pool's free list), then its
// proxies (their boxes emptied), then AabbTreeWorld::Remove (:1046-1060).
// The body keeps its id; its state is no longer stepped.  Proxies queued
// since the last step were never in the reference's tree: they give their
// leaf ids back when they were the last ones handed out.
int prx_world_destroy_body(prx_world* w, std::uint32_t body) {
    if (int st = CheckWorld(w)) return st;
    if (int st = CheckBody(w, body)) return st;
    if (int st = Detach(w)) return st;
    for (const HostJoint& j : w->joints)
        if (!j.destroyed && (j.flags >> 8) == prx::kJointGear && (j.c == body || j.d == body) && j.a != body &&
            j.b != body)
            return Fail(PRX_E_UNSUPPORTED, "destroying body C or D of a gear joint (destroy the gear joint first)");
    HostBody& hb = w->bodies[body];
    std::vector<std::uint32_t> pending;
    for (int pi : hb.proxies)
        if (w->proxies[std::size_t(pi)].flags & prx::kProxyPending) pending.push_back(w->proxies[std::size_t(pi)].id);
    if (!pending.empty()) {
        std::sort(pending.begin(), pending.end());
        const std::uint32_t k = std::uint32_t(pending.size());
        for (std::uint32_t i = 0; i < k; ++i) {
            const std::uint32_t n = w->leafCount - k + i;
            const std::uint32_t want = (n == 0) ? 0u : 2u * n - 1u;
            if (pending[i] != want)
                return Fail(PRX_E_UNSUPPORTED, "destroying a body whose queued proxies precede other queued proxies");
        }
        w->leafCount -= k;
    }
    for (;;) {
        std::size_t first = w->joints.size();
        for (std::size_t ji = 0; ji < w->joints.size(); ++ji) {
            const HostJoint& j = w->joints[ji];
            if (j.destroyed || (j.a != body && j.b != body)) continue;
            if (first == w->joints.size() || j.seq < w->joints[first].seq) first = ji;
        }
        if (first == w->joints.size()) break;
        DestroyJointRecord(w, std::uint32_t(first));
    }
    DestroyBodyContacts(w, body);
    for (int pi : hb.proxies) {
        HostProxy& p = w->proxies[std::size_t(pi)];
        p.dead = true;
        p.flags = 0;
    }
    hb.proxies.clear();
    hb.flags = 0;
    hb.destroyed = true;
    ++w->destroyed_bodies;
    return PRX_OK;
}

// SetType(World&, BodyID, BodyType, resetMassData = true) (WorldBody.cpp:236-246):
// Body::SetType (Body.cpp:124-149): the type's flags, awake unless static;
// kinematic and static bodies lose their accelerations and mass, a static
// one its velocity (position 0 becomes position 1).  SetBody
// (AabbTreeWorld.cpp:2739-2759) destroys the body's contacts and queues a
// static body for synchronisation; ResetMassData follows.
int prx_world_set_type(prx_world* w, std::uint32_t body, std::uint32_t type) {
    if (int st = CheckWorld(w)) return st;
    if (int st = CheckBody(w, body)) return st;
    if (type > PRX_DYNAMIC) return Fail(PRX_E_INVALID_ARGUMENT, "no such body type");
    if (int st = Detach(w)) return st;
    HostBody& b = w->bodies[body];
    if (b.type == type) return PRX_OK;
    b.type = type;
    b.flags &= ~(prx::kImpenetrable | prx::kSpeedable | prx::kAccelerable);
    if (type == PRX_DYNAMIC) {
        b.flags |= prx::kSpeedable | prx::kAccelerable | prx::kAwake;
    } else if (type == PRX_KINEMATIC) {
        b.flags |= prx::kImpenetrable | prx::kSpeedable | prx::kAwake;
        b.acc[0] = b.acc[1] = b.acc[2] = 0.0f;
        b.invMass = b.invRotI = 0.0f;
    } else {
        b.flags = (b.flags | prx::kImpenetrable) & ~prx::kAwake;
        b.vel[0] = b.vel[1] = b.vel[2] = 0.0f;
        for (int k = 0; k < 3; ++k) b.pos0[k] = b.pos1[k];
        b.acc[0] = b.acc[1] = b.acc[2] = 0.0f;
        b.invMass = b.invRotI = 0.0f;
    }
    b.under = 0.0f;
    DestroyBodyContacts(w, body);
    if (type == PRX_STATIC) b.flags |= prx::kNeedsSync;
    if (ResetMass(w, b)) {
        HostContacts& hc = w->contacts;
        for (std::size_t i = 0; i < hc.size(); ++i)
            if (hc.body[i].x == int(body) || hc.body[i].y == int(body)) hc.flags[i] |= prx::kNeedsUpdating;
        b.flags |= prx::kNeedsSync;
    }
    return PRX_OK;
}

// SetAcceleration(World&, BodyID, ...) through Body::SetAcceleration
// (Body.cpp:203-238): no change, nothing; a body that is not accelerable
// takes zero only; a larger linear or angular acceleration, a new direction
// or a new sign of the angular one wakes the body (under-active time reset).
int prx_world_set_acceleration(prx_world* w, std::uint32_t body, const float a[3]) {
    if (int st = CheckWorld(w)) return st;
    if (!a) return Fail(PRX_E_INVALID_ARGUMENT, "null acceleration");
    if (int st = CheckBody(w, body)) return st;
    if (int st = Detach(w)) return st;
    HostBody& b = w->bodies[body];
    if (b.acc[0] == a[0] && b.acc[1] == a[1] && b.acc[2] == a[2]) return PRX_OK;
    if (!(b.flags & prx::kAccelerable)) {
        if (a[0] != 0.0f || a[1] != 0.0f || a[2] != 0.0f) return PRX_OK;
    } else {
        const float m0 = b.acc[0] * b.acc[0] + b.acc[1] * b.acc[1];
        const float m1 = a[0] * a[0] + a[1] * a[1];
        if (b.acc[2] < a[2] || m0 < m1 || std::atan2(b.acc[1], b.acc[0]) != std::atan2(a[1], a[0]) ||
            std::signbit(b.acc[2]) != std::signbit(a[2])) {
            b.flags |= prx::kAwake;
            b.under = 0;
        }
    }
    b.acc[0] = a[0];
    b.acc[1] = a[1];
    b.acc[2] = a[2];
    return PRX_OK;
}

namespace {

// SetShape (AabbTreeWorld.cpp:1284-1357) for a change of friction or
// restitution only: every body holding the shape is woken (Body::SetAwake,
// Body.cpp:163-171: speedable bodies only, under-active time reset), and, if
// the value changed, the shape's contacts are flagged for updating with
// their bodies woken.
// SetBody (AabbTreeWorld.cpp:2723-2795) for a body whose sweep becomes one
// position (Sweep{Position}, alpha0 0): a new transformation flags the body's
// contacts for updating and queues the body for the next step's Synchronize.
int SetSweepPosition(prx_world* w, std::uint32_t body, float px, float py, float angle) {
    HostBody& b = w->bodies[body];
    const V2 q0 = UnitFromAngle(b.pos1[2]);
    const V2 p0 = V2{b.pos1[0], b.pos1[1]} - Rotate(b.lc, q0);
    b.pos0[0] = b.pos1[0] = px;
    b.pos0[1] = b.pos1[1] = py;
    b.pos0[2] = b.pos1[2] = angle;
    const V2 q1 = UnitFromAngle(angle);
    const V2 p1 = V2{px, py} - Rotate(b.lc, q1);
    if (!(p0.x == p1.x && p0.y == p1.y) || !(q0.x == q1.x && q0.y == q1.y)) {
        HostContacts& hc = w->contacts;
        for (std::size_t i = 0; i < hc.size(); ++i)
            if (hc.body[i].x == int(body) || hc.body[i].y == int(body)) hc.flags[i] |= prx::kNeedsUpdating;
        b.flags |= prx::kNeedsSync;
    }
    return PRX_OK;
}

int SetShapeMaterial(prx_world* w, std::uint32_t shape, bool friction, float value) {
    if (int st = CheckWorld(w)) return st;
    if (shape >= w->shapes.size()) return Fail(PRX_E_OUT_OF_RANGE, "no such shape");
    if (value != value || (friction && value < 0.0f))
        return Fail(PRX_E_INVALID_ARGUMENT, friction ? "friction must be non-negative" : "restitution is NaN");
    if (int st = Detach(w)) return st;
    HostShape& sh = w->shapes[shape];
    float& field = friction ? sh.friction : sh.restitution;
    auto wake = [&](int bi) {
        HostBody& b = w->bodies[std::size_t(bi)];
        if (b.flags & prx::kSpeedable) {
            b.flags |= prx::kAwake;
            b.under = 0;
        }
    };
    for (std::size_t bi = 0; bi < w->bodies.size(); ++bi)
        for (std::uint32_t s : w->bodies[bi].shapes)
            if (s == shape) wake(int(bi));
    if (field == value) return PRX_OK;
    HostContacts& c = w->contacts;
    for (std::size_t i = 0; i < c.size(); ++i) {
        if (std::uint32_t(c.shape[i].x) != shape && std::uint32_t(c.shape[i].y) != shape) continue;
        c.flags[i] |= prx::kNeedsUpdating;
        wake(c.body[i].x);
        wake(c.body[i].y);
    }
    field = value;
    return PRX_OK;
}

}  // namespace

int prx_shape_set_friction(prx_world* w, std::uint32_t shape, float friction) {
    return SetShapeMaterial(w, shape, true, friction);
}

int prx_shape_set_restitution(prx_world* w, std::uint32_t shape, float restitution) {
    return SetShapeMaterial(w, shape, false, restitution);
}

std::uint32_t prx_world_contact_count(prx_world* w) {
    if (!w || Refresh(w) != PRX_OK) return 0;
    return std::uint32_t(w->contacts.size());
}

int prx_world_record_contact_events(prx_world* w, int on) {
    if (int st = CheckWorld(w)) return st;
    const bool want = on != 0;
    if (want == w->record_events) return PRX_OK;
    if (w->owner && w->owner != w->solo)
        return Fail(PRX_E_WRONG_STATE, "the world is in a batch: switch contact event recording before batching it");
    if (int st = Detach(w)) return st;  // prx_step rebuilds its single-world batch with the new segment
    w->record_events = want;
    if (!want) {
        w->events.clear();
        w->nevents = 0;
    }
    return PRX_OK;
}

int prx_world_set_event_record_words(prx_world* w, std::uint32_t words) {
    if (int st = CheckWorld(w)) return st;
    if (words > 0x40000000u) return Fail(PRX_E_INVALID_ARGUMENT, "record size above 2^30 words");
    if (w->owner && w->owner != w->solo)
        return Fail(PRX_E_WRONG_STATE, "the world is in a batch: size its call record before batching it");
    if (int st = Detach(w)) return st;
    w->event_words = words;
    return PRX_OK;
}

int prx_world_set_contact_listener(prx_world* w, const prx_contact_listener* listener) {
    if (int st = CheckWorld(w)) return st;
    if (listener) {
        if (int st = prx_world_record_contact_events(w, 1)) return st;
        w->listener = *listener;
        w->has_listener = true;
    } else {
        w->listener = prx_contact_listener{};
        w->has_listener = false;
    }
    return PRX_OK;
}

int prx_world_contact_events(prx_world* w, std::uint32_t* words, std::size_t cap, std::size_t* nwords,
                             std::uint32_t* ncalls) {
    if (int st = CheckWorld(w)) return st;
    if (int st = Refresh(w)) return st;
    if (nwords) *nwords = w->events.size();
    if (ncalls) *ncalls = w->nevents;
    if (words) {
        if (cap < w->events.size()) return Fail(PRX_E_LENGTH, "buffer too small");
        if (!w->events.empty()) std::memcpy(words, w->events.data(), w->events.size() * 4);
    }
    return PRX_OK;
}

int prx_world_get_body_mass(prx_world* w, std::uint32_t body, float out[4]) {
    if (int st = CheckWorld(w)) return st;
    if (!out) return Fail(PRX_E_INVALID_ARGUMENT, "null output");
    if (int st = CheckBody(w, body)) return st;
    const HostBody& b = w->bodies[body];
    out[0] = b.invMass;
    out[1] = b.invRotI;
    out[2] = b.lc.x;
    out[3] = b.lc.y;
    return PRX_OK;
}

int prx_world_get_contacts(prx_world* w, prx_contact_info* out, std::size_t cap, std::size_t* count) {
    if (int st = CheckWorld(w)) return st;
    if (!count || (cap > 0 && !out)) return Fail(PRX_E_INVALID_ARGUMENT, "null output");
    if (int st = Refresh(w)) return st;
    const HostContacts& hc = w->contacts;
    *count = hc.size();
    for (std::size_t i = 0; i < hc.size() && i < cap; ++i) {
        prx_contact_info c{};
        c.key_min = w->proxies[std::size_t(hc.prox[i].x)].id;
        c.key_max = w->proxies[std::size_t(hc.prox[i].y)].id;
        c.id = hc.id[i];
        c.flags = hc.flags[i];
        c.body_a = std::uint32_t(hc.body[i].x);
        c.body_b = std::uint32_t(hc.body[i].y);
        c.shape_a = std::uint32_t(hc.shape[i].x);
        c.shape_b = std::uint32_t(hc.shape[i].y);
        c.friction = hc.mat[i].x;
        c.restitution = hc.mat[i].y;
        const uint4 h = hc.hdr[i];
        c.manifold_type = h.x & 0xffu;
        c.point_count = (h.x >> 8) & 0xffu;
        if (c.manifold_type == 2 || c.manifold_type == 3) {
            c.local_normal[0] = hc.geo[i].x;
            c.local_normal[1] = hc.geo[i].y;
        }
        if (c.manifold_type != 0) {
            c.local_point[0] = hc.geo[i].z;
            c.local_point[1] = hc.geo[i].w;
        }
        const float4 pts[2] = {hc.p0[i], hc.p1[i]};
        const std::uint32_t cfs[2] = {h.y, h.z};
        for (std::uint32_t k = 0; k < 2; ++k) {
            c.points[k].contact_feature = cfs[k];
            if (k < c.point_count) {
                c.points[k].local_point[0] = pts[k].x;
                c.points[k].local_point[1] = pts[k].y;
                c.points[k].normal_impulse = pts[k].z;
                c.points[k].tangent_impulse = pts[k].w;
            }
        }
        out[i] = c;
    }
    return PRX_OK;
}

int prx_world_get_proxies(prx_world* w, std::uint32_t* ids, std::uint32_t* bodies, std::uint32_t* shapes,
                          float* aabbs, std::size_t cap, std::size_t* count) {
    if (int st = CheckWorld(w)) return st;
    if (!count || (cap > 0 && (!ids || !bodies || !shapes || !aabbs)))
        return Fail(PRX_E_INVALID_ARGUMENT, "null output");
    if (int st = Refresh(w)) return st;
    std::size_t k = 0;
    for (const auto& p : w->proxies) {
        if (p.flags & prx::kProxyPending) continue;
        if (k < cap) {
            ids[k] = p.id;
            bodies[k] = std::uint32_t(p.body);
            shapes[k] = std::uint32_t(p.shape);
            for (int j = 0; j < 4; ++j) aabbs[4 * k + j] = p.aabb[j];
        }
        ++k;
    }
    *count = k;
    return PRX_OK;
}

int prx_world_state_record(prx_world* w, std::uint8_t* buf, std::size_t cap, std::size_t* len) {
    if (int st = CheckWorld(w)) return st;
    if (int st = Refresh(w)) return st;
    std::vector<std::uint8_t> o;
    auto put = [&](const void* p, std::size_t n) {
        const auto* c = static_cast<const std::uint8_t*>(p);
        o.insert(o.end(), c, c + n);
    };
    auto u32 = [&](std::uint32_t v) { put(&v, 4); };
    u32(0);
    put(&w->last_stats, sizeof(prx_step_stats));  // the 34 StepStats words, in order
    u32(std::uint32_t(w->bodies.size()));
    for (const auto& b : w->bodies) {
        if (b.destroyed) {
            // A destroyed body: zeros and flags 0x80.
            const float z[10] = {};
            put(z, sizeof(z));
            u32(0x80u);
            continue;
        }
        const float v[10] = {b.pos0[0], b.pos0[1], b.pos0[2], b.pos1[0], b.pos1[1],
                             b.pos1[2], b.vel[0],  b.vel[1],  b.vel[2],  b.under};
        put(v, sizeof(v));
        u32(b.flags & 0x7fu);
    }
    const HostContacts& hc = w->contacts;
    u32(std::uint32_t(hc.size()));
    for (std::size_t i = 0; i < hc.size(); ++i) {
        const std::uint32_t u[4] = {w->proxies[std::size_t(hc.prox[i].x)].id, w->proxies[std::size_t(hc.prox[i].y)].id,
                                    hc.id[i], hc.flags[i]};
        put(u, sizeof(u));
        const uint4 h = hc.hdr[i];
        const std::uint32_t type = h.x & 0xffu, n = (h.x >> 8) & 0xffu;
        const std::uint32_t t[3] = {type | (n << 8), h.y, h.z};
        put(t, sizeof(t));
        float f[14] = {hc.mat[i].x, hc.mat[i].y};
        if (type == 2 || type == 3) {
            f[2] = hc.geo[i].x;
            f[3] = hc.geo[i].y;
        }
        if (type != 0) {
            f[4] = hc.geo[i].z;
            f[5] = hc.geo[i].w;
        }
        if (n > 0) {
            f[6] = hc.p0[i].x;
            f[7] = hc.p0[i].y;
            f[8] = hc.p0[i].z;
            f[9] = hc.p0[i].w;
        }
        if (n > 1) {
            f[10] = hc.p1[i].x;
            f[11] = hc.p1[i].y;
            f[12] = hc.p1[i].z;
            f[13] = hc.p1[i].w;
        }
        put(f, sizeof(f));
    }
    std::uint32_t np = 0;
    for (const auto& p : w->proxies) np += ((p.flags & prx::kProxyPending) || p.dead) ? 0u : 1u;
    u32(np);
    for (const auto& p : w->proxies) {
        if ((p.flags & prx::kProxyPending) || p.dead) continue;
        u32(p.id);
        put(p.aabb, sizeof(p.aabb));
    }
    u32(std::uint32_t(w->joints.size()));
    for (const auto& j : w->joints) {
        if (j.destroyed) {
            // A destroyed joint: limit state 0xffffffff and zeros.
            const float z[4] = {};
            u32(0xffffffffu);
            put(z, sizeof(z));
            continue;
        }
        u32(j.limit);
        put(j.imp, sizeof(j.imp));
    }
    u32(w->nevents);
    if (!w->events.empty()) put(w->events.data(), w->events.size() * 4);
    if (len) *len = o.size();
    if (buf && cap >= o.size()) std::memcpy(buf, o.data(), o.size());
    else if (buf) return Fail(PRX_E_LENGTH, "buffer too small");
    return PRX_OK;
}

int prx_batch_create(prx_world* const* worlds, int nworlds, const prx_batch_conf* conf, prx_batch** out) {
    return CreateBatch(worlds, nworlds, conf, out);
}

void prx_batch_destroy(prx_batch* b) {
    if (!b) return;
    (void)DownloadBatch(b);
    for (prx_world* w : b->worlds) {
        if (w && w->owner == b) {
            w->owner = nullptr;
            w->index_in_owner = -1;
        }
    }
    delete b;
}

int prx_batch_step(prx_batch* b, const prx_step_conf* conf, int nsteps) {
    if (!b) return Fail(PRX_E_INVALID_ARGUMENT, "null batch");
    if (!HasListener(b)) return StepBatch(b, conf, nsteps);
    // Worlds with listeners: one step at a time, each step's calls delivered
    // before the next step.
    for (int s = 0; s < nsteps; ++s) {
        if (int st = StepRecorded(b, conf)) return st;
        if (int st = DeliverEvents(b)) return st;
    }
    return PRX_OK;
}

int prx_batch_sync(prx_batch* b) {
    if (!b) return Fail(PRX_E_INVALID_ARGUMENT, "null batch");
    PRX_HIP(hipStreamSynchronize(b->stream));
    return CheckStatus(b);
}

int prx_batch_download(prx_batch* b) {
    if (!b) return Fail(PRX_E_INVALID_ARGUMENT, "null batch");
    return DownloadBatch(b);
}

int prx_batch_last_timing(prx_batch* b, float* ms, int* launches) {
    if (!b) return Fail(PRX_E_INVALID_ARGUMENT, "null batch");
    PRX_HIP(hipEventSynchronize(b->ev1));
    float t = 0;
    PRX_HIP(hipEventElapsedTime(&t, b->ev0, b->ev1));
    if (ms) *ms = t;
    if (launches) *launches = b->last_launches;
    return PRX_OK;
}

std::uint64_t prx_batch_dynamic_bodies(const prx_batch* b) {
    if (!b) return 0;
    std::uint64_t n = 0;
    for (const prx_world* w : b->worlds)
        if (w)
            for (const auto& hb : w->bodies) n += (hb.flags & prx::kAccelerable) ? 1u : 0u;
    return n;
}

int prx_batch_counts(prx_batch* b, std::uint64_t out[5]) {
    if (!b || !out) return Fail(PRX_E_INVALID_ARGUMENT, "null argument");
    for (int i = 0; i < 5; ++i) out[i] = 0;
    for (const prx_world* w : b->worlds) {
        if (!w) continue;
        for (const auto& hb : w->bodies)
            out[0] += ((hb.flags & prx::kAccelerable) && (hb.flags & prx::kAwake)) ? 1u : 0u;
        for (const auto& p : w->proxies) {
            ++out[1];
            out[2] += 8u + 16u * std::uint64_t(w->shapes[std::size_t(p.shape)].v.size());
        }
        out[3] += w->contacts.size();
        for (std::size_t i = 0; i < w->contacts.size(); ++i) out[4] += (w->contacts.flags[i] & prx::kTouching) ? 1u : 0u;
    }
    return PRX_OK;
}

int prx_batch_phase_cycles(prx_batch* b, int enable, std::uint64_t* out, int max_worlds) {
    if (!b) return Fail(PRX_E_INVALID_ARGUMENT, "null batch");
    const std::size_t n = b->worlds.size() * std::size_t(prx::kPhases);
    if (b->d.phase_cycles) {
        PRX_HIP(hipStreamSynchronize(b->stream));
        if (out) {
            const std::size_t m = std::min<std::size_t>(n, std::size_t(std::max(0, max_worlds)) * prx::kPhases);
            PRX_HIP(hipMemcpy(out, b->d.phase_cycles, m * sizeof(std::uint64_t), hipMemcpyDeviceToHost));
        }
    } else if (out) {
        std::memset(out, 0, std::size_t(std::max(0, max_worlds)) * prx::kPhases * sizeof(std::uint64_t));
    }
    if (enable && !b->d.phase_cycles) {
        PRX_HIP(b->arena.alloc(b->d.phase_cycles, n));
        b->graph_valid = false;
    }
    if (!enable && b->d.phase_cycles) {
        b->d.phase_cycles = nullptr;
        b->graph_valid = false;
    }
    if (b->d.phase_cycles) PRX_HIP(hipMemset(b->d.phase_cycles, 0, n * sizeof(std::uint64_t)));
    return PRX_OK;
}

int prx_batch_autotune(prx_batch* b, const prx_step_conf* step, int nsteps, prx_batch_conf* best) {
    if (!b || !step) return Fail(PRX_E_INVALID_ARGUMENT, "null argument");
    if (nsteps < 1) return Fail(PRX_E_INVALID_ARGUMENT, "nsteps < 1");
    for (prx_world* w : b->worlds)
        if (!w) return Fail(PRX_E_WRONG_STATE, "a member world left the batch");
    // Start state: the host mirror after a download; every trial starts from
    // it and it is restored at the end.
    if (int st = DownloadBatch(b)) return st;
    const prx_batch_conf start = b->conf;
    auto trial = [&](const prx_batch_conf& c, float* ms) -> int {
        b->conf = c;
        if (int st = ApplySchedule(b)) return st;
        if (int st = UploadBatch(b)) return st;
        // Two untimed steps capture the graph for this schedule.
        if (int st = StepBatch(b, step, 2)) return st;
        PRX_HIP(hipStreamSynchronize(b->stream));
        if (int st = UploadBatch(b)) return st;
        if (int st = StepBatch(b, step, nsteps)) return st;
        PRX_HIP(hipEventSynchronize(b->ev1));
        PRX_HIP(hipEventElapsedTime(ms, b->ev0, b->ev1));
        return CheckStatus(b);
    };
    // Coordinate search: workgroups per CU (LDS per workgroup), then the
    // wave-wide sweep threshold, then the launch order.  A candidate replaces
    // the current schedule only when at least 2 % faster: schedules that tie
    // over the trial steps can part ways as the worlds evolve, and the one in
    // hand is kept then.
    constexpr float kSwitchMargin = 0.98f;
    prx_batch_conf cur = start;
    float bestMs = 0;
    if (int st = trial(cur, &bestMs)) return st;
    const std::uint32_t ldsCands[] = {0u, 20u * 1024u, 26u * 1024u + 512u, 32u * 1024u + 256u, 40u * 1024u};
    for (std::uint32_t l : ldsCands) {
        prx_batch_conf c = cur;
        c.lds_bytes = l;
        if (c.lds_bytes == cur.lds_bytes) continue;
        float ms = 0;
        if (int st = trial(c, &ms)) return st;
        if (ms < bestMs * kSwitchMargin) {
            bestMs = ms;
            cur = c;
        }
    }
    const std::uint32_t coopCands[] = {8u, 16u, 32u, 64u, 128u, 0xffffffffu};
    for (std::uint32_t k : coopCands) {
        prx_batch_conf c = cur;
        c.coop_min_contacts = k;
        if (c.coop_min_contacts == cur.coop_min_contacts) continue;
        float ms = 0;
        if (int st = trial(c, &ms)) return st;
        if (ms < bestMs * kSwitchMargin) {
            bestMs = ms;
            cur = c;
        }
    }
    {
        prx_batch_conf c = cur;
        c.order = cur.order ? 0u : 1u;
        float ms = 0;
        if (int st = trial(c, &ms)) return st;
        if (ms < bestMs * kSwitchMargin) {
            bestMs = ms;
            cur = c;
        }
    }
    b->conf = cur;
    if (int st = ApplySchedule(b)) return st;
    if (int st = UploadBatch(b)) return st;
    for (prx_world* w : b->worlds) w->host_stale = false;
    b->last_launches = 0;
    if (best) *best = cur;
    return PRX_OK;
}


// ---------------------------------------------------------------------------
// Single large world (graph-coloured mode) and the LBVH pair finder.

}  // extern "C"

namespace {

std::size_t NextPow2(std::size_t v) {
    std::size_t p = 1;
    while (p < v) p <<= 1;
    return p;
}

int AllocBig(DeviceArena& a, prx::BigState& g, int P, int B, int cap) {
    g.P = std::max(P, 1);
    g.B = std::max(B, 1);
    g.cap = std::max(cap, 1);
    g.pairCap = std::max(16 * g.P, 1024);
    g.hashMask = int(NextPow2(std::size_t(2 * g.cap))) - 1;
    const std::size_t n = std::size_t(g.P), nodes = 2 * n;
    PRX_HIP(a.alloc(g.morton, n));
    PRX_HIP(a.alloc(g.morton2, n));
    PRX_HIP(a.alloc(g.order, n));
    PRX_HIP(a.alloc(g.order2, n));
    PRX_HIP(a.alloc(g.left, n));
    PRX_HIP(a.alloc(g.right, n));
    PRX_HIP(a.alloc(g.parent, nodes));
    PRX_HIP(a.alloc(g.escape, nodes));
    PRX_HIP(a.alloc(g.visit, n));
    PRX_HIP(a.alloc(g.naabb, nodes));
    PRX_HIP(a.alloc(g.bounds, 1));
    PRX_HIP(a.alloc(g.pcount, n + 1));
    PRX_HIP(a.alloc(g.pairs, std::size_t(g.pairCap)));
    PRX_HIP(a.alloc(g.paccept, std::size_t(g.pairCap)));
    PRX_HIP(a.alloc(g.hash, std::size_t(g.hashMask) + 1));
    PRX_HIP(a.alloc(g.list, std::size_t(g.cap)));
    PRX_HIP(a.alloc(g.list2, std::size_t(g.cap)));
    PRX_HIP(a.alloc(g.color, std::size_t(g.cap)));
    PRX_HIP(a.alloc(g.claim, std::size_t(g.B) * prx::kBigMaxColors));
    PRX_HIP(hipMemset(g.claim, 0xff, sizeof(unsigned long long) * std::size_t(g.B) * prx::kBigMaxColors));
    PRX_HIP(a.alloc(g.used, std::size_t(g.B)));
    PRX_HIP(a.alloc(g.tent, std::size_t(g.cap)));
    PRX_HIP(a.alloc(g.win, std::size_t(g.cap)));
    PRX_HIP(a.alloc(g.colorStart, std::size_t(prx::kBigMaxColors + 2)));
    PRX_HIP(a.alloc(g.toiKey, std::size_t(g.B)));
    PRX_HIP(hipMemset(g.toiKey, 0xff, sizeof(unsigned long long) * std::size_t(g.B)));
    PRX_HIP(a.alloc(g.toiLeft, std::size_t(g.B)));
    PRX_HIP(a.alloc(g.sweep, std::size_t(prx::kBigSweepSlots)));
    PRX_HIP(a.alloc(g.counters, std::size_t(prx::kBigCounters)));
    PRX_HIP(a.alloc(g.metrics, std::size_t(prx::kBigMetrics)));
    PRX_HIP(hipMemset(g.counters, 0, sizeof(int) * prx::kBigCounters));
    PRX_HIP(hipMemset(g.metrics, 0, sizeof(float) * prx::kBigMetrics));
    return PRX_OK;
}

}  // namespace

struct prx_big {
    prx_world* world = nullptr;
    prx_batch* batch = nullptr;  // the world's one-world device state
    DeviceArena arena;
    prx::BigState g{};
    hipGraphExec_t exec = nullptr;
    hipGraph_t graph = nullptr;
    prx::StepParams gparams{};
    bool graph_valid = false;
    int last_launches = 0;
    ~prx_big() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
    }
};

extern "C" {

int prx_big_create(prx_world* w, prx_big** out) {
    if (int st = CheckWorld(w)) return st;
    if (!out) return Fail(PRX_E_INVALID_ARGUMENT, "null out");
    if (!w->joints.empty()) return Fail(PRX_E_UNSUPPORTED, "joints are not stepped by the coloured large-world mode");
    if (w->record_events)
        return Fail(PRX_E_UNSUPPORTED, "contact listener calls are recorded by the batched step only");
    if (w->proxies.empty()) return Fail(PRX_E_INVALID_ARGUMENT, "world has no shapes");
    prx_batch_conf bc;
    prx_batch_conf_default(&bc);
    bc.use_graph = 0;
    prx_batch* b = nullptr;
    prx_world* one[1] = {w};
    if (int st = CreateBatch(one, 1, &bc, &b)) return st;
    auto big = std::make_unique<prx_big>();
    big->world = w;
    big->batch = b;
    if (int st = AllocBig(big->arena, big->g, int(b->totP), int(b->totB), int(b->totC))) {
        prx_batch_destroy(b);
        return st;
    }
    *out = big.release();
    return PRX_OK;
}

void prx_big_destroy(prx_big* big) {
    if (!big) return;
    prx_batch_destroy(big->batch);
    delete big;
}

int prx_big_step(prx_big* big, const prx_step_conf* conf, int nsteps) {
    if (!big || !conf) return Fail(PRX_E_INVALID_ARGUMENT, "null argument");
    if (nsteps < 0) return Fail(PRX_E_INVALID_ARGUMENT, "negative step count");
    prx_batch* b = big->batch;
    if (b->worlds[0] != big->world || big->world->owner != b)
        return Fail(PRX_E_WRONG_STATE, "the world left its large-world state");
    if (conf->reg_velocity_iters > std::uint32_t(prx::kBigPosSlot - prx::kBigVelSlot) ||
        conf->reg_position_iters > std::uint32_t(prx::kBigSweepSlots - prx::kBigPosSlot))
        return Fail(PRX_E_UNSUPPORTED, "the large-world mode runs at most 16 velocity and 16 position iterations");
    prx::StepParams p = ToParams(*conf);
    p.ldsBytes = 0;
    p.coopMin = 0;
    PRX_HIP(hipEventRecord(b->ev0, b->stream));
    if (!big->graph_valid || std::memcmp(&p, &big->gparams, sizeof(p)) != 0) {
        if (big->exec) (void)hipGraphExecDestroy(big->exec);
        if (big->graph) (void)hipGraphDestroy(big->graph);
        big->exec = nullptr;
        big->graph = nullptr;
        PRX_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        const hipError_t le = prx_big_launch_step(&b->d, &big->g, &p, b->stream);
        hipGraph_t gr = nullptr;
        const hipError_t ce = hipStreamEndCapture(b->stream, &gr);
        if (le != hipSuccess) return Fail(PRX_E_DEVICE, std::string("launch: ") + hipGetErrorString(le));
        if (ce != hipSuccess) return Fail(PRX_E_DEVICE, std::string("capture: ") + hipGetErrorString(ce));
        big->graph = gr;
        PRX_HIP(hipGraphInstantiate(&big->exec, gr, nullptr, nullptr, 0));
        big->gparams = p;
        big->graph_valid = true;
        PRX_HIP(hipEventRecord(b->ev0, b->stream));
    }
    for (int s = 0; s < nsteps; ++s) PRX_HIP(hipGraphLaunch(big->exec, b->stream));
    PRX_HIP(hipEventRecord(b->ev1, b->stream));
    b->last_launches = nsteps;
    big->world->host_stale = true;
    return PRX_OK;
}

int prx_big_sync(prx_big* big) {
    if (!big) return Fail(PRX_E_INVALID_ARGUMENT, "null argument");
    PRX_HIP(hipStreamSynchronize(big->batch->stream));
    return CheckStatus(big->batch);
}

int prx_big_last_timing(prx_big* big, float* ms, int* launches) {
    if (!big) return Fail(PRX_E_INVALID_ARGUMENT, "null argument");
    return prx_batch_last_timing(big->batch, ms, launches);
}

int prx_big_metrics(prx_big* big, float out[4]) {
    if (!big || !out) return Fail(PRX_E_INVALID_ARGUMENT, "null argument");
    prx_batch* b = big->batch;
    const int init[1] = {0x7fffffff};
    PRX_HIP(hipMemcpyAsync(big->g.counters + prx::kBigMinSepBits, init, sizeof(int), hipMemcpyHostToDevice,
                           b->stream));
    PRX_HIP(prx_big_launch_metrics(&b->d, &big->g, b->stream));
    PRX_HIP(hipStreamSynchronize(b->stream));
    int counters[prx::kBigCounters];
    PRX_HIP(hipMemcpy(counters, big->g.counters, sizeof(counters), hipMemcpyDeviceToHost));
    prx::WorldHeader h;
    PRX_HIP(hipMemcpy(&h, b->d.hdr, sizeof(h), hipMemcpyDeviceToHost));
    const int key = counters[prx::kBigMinSepBits];
    float minSep = __builtin_inff();
    if (key != 0x7fffffff) {
        const int bits = key >= 0 ? key : (key ^ 0x7fffffff);
        std::memcpy(&minSep, &bits, 4);
    }
    out[0] = minSep;
    out[1] = float(counters[prx::kBigTouching]);
    out[2] = float(counters[prx::kBigColors]);
    out[3] = float(h.ncontacts);
    return PRX_OK;
}

int prx_lbvh_pairs(const float* aabbs, const std::int32_t* bodies, const std::uint8_t* moved, int n,
                   std::int32_t* pairs, int cap, int* count) {
    if (!aabbs || !bodies || !moved || n <= 0 || !count || (cap > 0 && !pairs))
        return Fail(PRX_E_INVALID_ARGUMENT, "bad argument");
    DeviceArena a;
    prx::DeviceState d{};
    prx::BigState g{};
    d.nworlds = 1;
    PRX_HIP(a.alloc(d.hdr, 1));
    PRX_HIP(a.alloc(d.paabb, std::size_t(n)));
    PRX_HIP(a.alloc(d.pbody, std::size_t(n)));
    PRX_HIP(a.alloc(d.pflags, std::size_t(n)));
    if (int st = AllocBig(a, g, n, 1, 1)) return st;
    prx::WorldHeader h{};
    h.nproxies = n;
    h.nbodies = 0;
    std::vector<float4> ab(static_cast<std::size_t>(n));
    std::vector<std::uint32_t> fl(static_cast<std::size_t>(n));
    for (int i = 0; i < n; ++i) {
        ab[std::size_t(i)] = float4{aabbs[4 * i], aabbs[4 * i + 1], aabbs[4 * i + 2], aabbs[4 * i + 3]};
        fl[std::size_t(i)] = moved[i] ? prx::kProxyMoved : 0u;
    }
    PRX_HIP(hipMemcpy(d.hdr, &h, sizeof(h), hipMemcpyHostToDevice));
    PRX_HIP(hipMemcpy(d.paabb, ab.data(), ab.size() * sizeof(float4), hipMemcpyHostToDevice));
    PRX_HIP(hipMemcpy(d.pbody, bodies, std::size_t(n) * sizeof(int), hipMemcpyHostToDevice));
    PRX_HIP(hipMemcpy(d.pflags, fl.data(), fl.size() * sizeof(std::uint32_t), hipMemcpyHostToDevice));
    PRX_HIP(prx_big_launch_pairs(&d, &g, nullptr));
    PRX_HIP(hipDeviceSynchronize());
    int total = 0;
    PRX_HIP(hipMemcpy(&total, g.pcount + n, sizeof(int), hipMemcpyDeviceToHost));
    *count = total;
    const int m = std::min(std::min(total, cap), g.pairCap);
    if (m > 0) PRX_HIP(hipMemcpy(pairs, g.pairs, std::size_t(m) * sizeof(int2), hipMemcpyDeviceToHost));
    if (total > g.pairCap) return Fail(PRX_E_LENGTH, "more pairs than the pair capacity");
    return PRX_OK;
}

int prx_lbvh_pairs_noalloc(const float* aabbs, const std::int32_t* bodies, const std::uint8_t* moved, int n,
                           std::int32_t* pairs, int cap, int* count, void* stream) {
    const char* err = "";
    const int st = prx_lbvh_pairs_noalloc_impl(aabbs, bodies, moved, n, pairs, cap, count, stream, &err);
    return st ? Fail(st, err) : PRX_OK;
}


// ---------------------------------------------------------------------------
// Adjoint of K steps of a batch (loss = sum of squared distances of the
// dynamic bodies' final positions to targets).

}  // extern "C"

namespace {

// Arrays a step changes (the checkpoint of a batch).
std::vector<std::pair<void*, std::size_t>> DynamicArrays(prx_batch* b) {
    const std::size_t B = b->totB, P = b->totP, C = b->totC, J = b->totJ, nw = b->worlds.size();
    auto& d = b->d;
    std::vector<std::pair<void*, std::size_t>> v = {
        {d.hdr, nw * sizeof(prx::WorldHeader)}, {d.bflags, B * 4}, {d.pos0, B * 16}, {d.pos1, B * 16},
        {d.vel, B * 16}, {d.xf, B * 16}, {d.paabb, P * 16}, {d.pflags, P * 4}, {d.c_prox, C * 8},
        {d.c_body, C * 8}, {d.c_shape, C * 8}, {d.c_id, C * 4}, {d.c_flags, C * 4}, {d.c_mat, C * 16},
        {d.m_hdr, C * 16}, {d.m_geo, C * 16}, {d.m_p0, C * 16}, {d.m_p1, C * 16}, {d.c_free, C * 4},
        {d.j_imp, J * 16}, {d.j_limit, J * 4}};
    if (b->totE) v.push_back({d.ev, b->totE * 4});  // listener calls of the last step
    return v;
}

}  // namespace

struct prx_adjoint {
    prx_batch* b = nullptr;
    int K = 0, S = 0, nseg = 0, itV = 0, itP = 0;
    bool general = false;  // joints, polygons or off-centre disks: the general reverse kernel
    DeviceArena arena;
    std::vector<prx::Tape> tapes;
    prx::AdjState a{};
    std::vector<std::vector<void*>> ckpt;
    bool forwarded = false;  // checkpoints of a forward pass not yet reversed
};

namespace {

int AdjointParams(prx_adjoint* adj, const prx_step_conf* conf, prx::StepParams& p) {
    if (conf->do_toi) return Fail(PRX_E_UNSUPPORTED, "the adjoint does not differentiate the TOI phase (doToi = true)");
    if (int(conf->reg_velocity_iters) > adj->itV || int(conf->reg_position_iters) > adj->itP)
        return Fail(PRX_E_INVALID_ARGUMENT, "iteration counts above those the adjoint was created for");
    for (prx_world* w : adj->b->worlds)
        if (!w) return Fail(PRX_E_WRONG_STATE, "a member world left the batch");
    p = ToParams(*conf);
    p.ldsBytes = std::uint32_t(adj->b->lds_bytes);
    p.coopMin = 0xffffffffu;  // every island on one lane (the recorded form)
    return PRX_OK;
}

int AdjointCopy(prx_adjoint* adj, int k, bool save) {
    prx_batch* b = adj->b;
    const auto arrays = DynamicArrays(b);
    for (std::size_t i = 0; i < arrays.size(); ++i) {
        if (!arrays[i].second) continue;
        void* ck = adj->ckpt[std::size_t(k)][i];
        PRX_HIP(hipMemcpyAsync(save ? ck : arrays[i].first, save ? arrays[i].first : ck, arrays[i].second,
                               hipMemcpyDeviceToDevice, b->stream));
    }
    return PRX_OK;
}

// K steps from the batch's state with a checkpoint every S steps; the batch
// is left at the final state.
int AdjointForward(prx_adjoint* adj, const prx::StepParams& p) {
    prx_batch* b = adj->b;
    const int nw = int(b->worlds.size());
    hipStream_t st = b->stream;
    PRX_HIP(hipEventRecord(b->ev0, st));
    for (int k = 0; k < adj->K; ++k) {
        if (k % adj->S == 0)
            if (int e = AdjointCopy(adj, k / adj->S, true)) return e;
        const hipError_t le = prx_launch_step(&b->d, &p, 0, nw, st);
        if (le != hipSuccess) return Fail(PRX_E_DEVICE, std::string("launch: ") + hipGetErrorString(le));
    }
    PRX_HIP(hipStreamSynchronize(st));
    if (int e = CheckStatus(b)) return e;
    adj->forwarded = true;
    return PRX_OK;
}

// The reversed steps from the adjoint of the final positions (x, y, angle)
// and velocities (vx, vy, w), segment by segment from the checkpoints; the
// batch is left at its starting state.
int AdjointBackward(prx_adjoint* adj, const prx::StepParams& p, const std::vector<float4>& xbar,
                    const std::vector<float4>& vbar, float* grad_velocity, float* grad_friction,
                    float* grad_restitution) {
    prx_batch* b = adj->b;
    const int nw = int(b->worlds.size());
    const std::size_t B = b->totB, S = b->totS;
    hipStream_t st = b->stream;
    PRX_HIP(hipMemcpy(adj->a.xbar, xbar.data(), B * sizeof(float4), hipMemcpyHostToDevice));
    PRX_HIP(hipMemcpy(adj->a.vbar, vbar.data(), B * sizeof(float4), hipMemcpyHostToDevice));
    PRX_HIP(hipMemset(adj->a.lambar, 0, b->totC * sizeof(float4)));
    PRX_HIP(hipMemset(adj->a.fbar, 0, S * sizeof(float)));
    PRX_HIP(hipMemset(adj->a.ebar, 0, S * sizeof(float)));
    PRX_HIP(hipMemset(adj->a.status, 0, sizeof(int)));
    PRX_HIP(hipMemset(adj->a.diag, 0, 4 * sizeof(float)));
    {
        const std::size_t J = std::max<std::size_t>(b->totJ, 1);
        PRX_HIP(hipMemset(adj->a.lamc, 0, 2 * b->totC * sizeof(float4)));
        PRX_HIP(hipMemset(adj->a.jcarry, 0, J * sizeof(float4)));
        PRX_HIP(hipMemset(adj->a.tbar, 0, J * sizeof(float)));
    }
    for (int seg = adj->nseg - 1; seg >= 0; --seg) {
        if (int e = AdjointCopy(adj, seg, false)) return e;
        const int n = std::min(adj->S, adj->K - seg * adj->S);
        for (int s = 0; s < n; ++s) {
            prx::DeviceState dd = b->d;
            dd.tape = adj->tapes[std::size_t(s)];
            const hipError_t le = prx_launch_step_tape(&dd, &p, nw, st);
            if (le != hipSuccess) return Fail(PRX_E_DEVICE, std::string("launch: ") + hipGetErrorString(le));
        }
        for (int s = n - 1; s >= 0; --s) {
            const hipError_t le =
                adj->general ? prx_launch_adjoint_general(&b->d, &adj->tapes[std::size_t(s)], &adj->a, &p, nw, st)
                             : prx_launch_adjoint_step(&b->d, &adj->tapes[std::size_t(s)], &adj->a, &p, nw, st);
            if (le != hipSuccess) return Fail(PRX_E_DEVICE, std::string("launch: ") + hipGetErrorString(le));
        }
    }
    PRX_HIP(hipEventRecord(b->ev1, st));
    b->last_launches = 2 * adj->K;
    PRX_HIP(hipStreamSynchronize(st));
    adj->forwarded = false;
    int status = 0;
    PRX_HIP(hipMemcpy(&status, adj->a.status, sizeof(int), hipMemcpyDeviceToHost));
    if (grad_velocity) {
        std::vector<float4> vb(B);
        PRX_HIP(hipMemcpy(vb.data(), adj->a.vbar, B * sizeof(float4), hipMemcpyDeviceToHost));
        for (std::size_t i = 0; i < B; ++i) {
            grad_velocity[3 * i] = vb[i].x;
            grad_velocity[3 * i + 1] = vb[i].y;
            grad_velocity[3 * i + 2] = vb[i].z;
        }
    }
    if (grad_friction) PRX_HIP(hipMemcpy(grad_friction, adj->a.fbar, S * sizeof(float), hipMemcpyDeviceToHost));
    if (grad_restitution)
        PRX_HIP(hipMemcpy(grad_restitution, adj->a.ebar, S * sizeof(float), hipMemcpyDeviceToHost));
    if (int e = AdjointCopy(adj, 0, false)) return e;
    PRX_HIP(hipStreamSynchronize(st));
    for (prx_world* w : b->worlds) w->host_stale = true;
    if (status) return Fail(PRX_E_UNSUPPORTED, "a contact the adjoint does not model (two-point manifold or face B)");
    return PRX_OK;
}

}  // namespace

extern "C" {

int prx_adjoint_create(prx_batch* b, int nsteps, int every, const prx_step_conf* conf, prx_adjoint** out) {
    if (!b || !out || !conf || nsteps < 1 || every < 1) return Fail(PRX_E_INVALID_ARGUMENT, "bad argument");
    bool general = false;
    for (prx_world* w : b->worlds) {
        if (!w) return Fail(PRX_E_WRONG_STATE, "a member world left the batch");
        if (!w->joints.empty()) general = true;
        for (const auto& j : w->joints)
            if (j.flags >> 8) return Fail(PRX_E_UNSUPPORTED, "the adjoint differentiates revolute joints only");
        for (const auto& hb : w->bodies) {
            if (!(hb.flags & prx::kAccelerable)) continue;
            if (hb.lc.x != 0.0f || hb.lc.y != 0.0f) general = true;
            for (std::uint32_t sid : hb.shapes) {
                const HostShape& sh = w->shapes[sid];
                if (sh.v.size() != 1 || sh.v[0].x != 0.0f || sh.v[0].y != 0.0f) general = true;
            }
        }
    }
    auto adj = std::make_unique<prx_adjoint>();
    adj->general = general;
    adj->b = b;
    adj->K = nsteps;
    adj->S = std::min(every, nsteps);
    adj->nseg = (nsteps + adj->S - 1) / adj->S;
    adj->itV = int(std::max<std::uint32_t>(1u, conf->reg_velocity_iters));
    adj->itP = int(std::max<std::uint32_t>(1u, conf->reg_position_iters));
    auto& A = adj->arena;
    const std::size_t B = b->totB, C = b->totC, S = b->totS, nw = b->worlds.size();
    const std::size_t J = std::max<std::size_t>(b->totJ, 1);
    adj->tapes.resize(std::size_t(adj->S));
    for (auto& t : adj->tapes) {
        PRX_HIP(A.alloc(t.vc, C * prx::kTapeVc));
        PRX_HIP(A.alloc(t.gs, C * std::size_t(adj->itV) * 3));
        PRX_HIP(A.alloc(t.ngs, C * std::size_t(adj->itP) * 2));
        PRX_HIP(A.alloc(t.isl, B + nw));
        PRX_HIP(A.alloc(t.islj, B + nw));
        PRX_HIP(A.alloc(t.nisl, nw));
        PRX_HIP(A.alloc(t.inisl, B));
        PRX_HIP(A.alloc(t.jc, J * prx::kTapeJc));
        PRX_HIP(A.alloc(t.jgs, J * std::size_t(adj->itV) * 3));
        PRX_HIP(A.alloc(t.jngs, J * std::size_t(adj->itP) * 2));
        PRX_HIP(A.alloc(t.cmap, std::max<std::size_t>(C, 1)));
        t.itV = adj->itV;
        t.itP = adj->itP;
    }
    PRX_HIP(A.alloc(adj->a.xbar, B));
    PRX_HIP(A.alloc(adj->a.vbar, B));
    PRX_HIP(A.alloc(adj->a.lambar, C));
    PRX_HIP(A.alloc(adj->a.lamtmp, C));
    PRX_HIP(A.alloc(adj->a.cbar, 3 * C));
    PRX_HIP(A.alloc(adj->a.fbar, S));
    PRX_HIP(A.alloc(adj->a.ebar, S));
    PRX_HIP(A.alloc(adj->a.status, 1));
    PRX_HIP(A.alloc(adj->a.diag, 4));
    PRX_HIP(A.alloc(adj->a.fld, C * prx::kAdjFields));
    PRX_HIP(A.alloc(adj->a.lam4, C));
    PRX_HIP(A.alloc(adj->a.lamc, 2 * C));
    PRX_HIP(A.alloc(adj->a.jfld, J * prx::kAdjJointFields));
    PRX_HIP(A.alloc(adj->a.jlam, J));
    PRX_HIP(A.alloc(adj->a.jcarry, J));
    PRX_HIP(A.alloc(adj->a.tbar, J));
    const auto arrays = DynamicArrays(b);
    adj->ckpt.resize(std::size_t(adj->nseg));
    for (auto& ck : adj->ckpt) {
        for (const auto& ar : arrays) {
            std::uint8_t* p = nullptr;
            PRX_HIP(A.alloc(p, std::max<std::size_t>(ar.second, 1)));
            ck.push_back(p);
        }
    }
    *out = adj.release();
    return PRX_OK;
}

void prx_adjoint_destroy(prx_adjoint* adj) { delete adj; }

int prx_adjoint_run(prx_adjoint* adj, const prx_step_conf* conf, const float* targets, float* loss,
                    float* grad_velocity, float* grad_friction, float* grad_restitution) {
    if (!adj || !conf || !targets) return Fail(PRX_E_INVALID_ARGUMENT, "null argument");
    prx::StepParams p;
    if (int e = AdjointParams(adj, conf, p)) return e;
    prx_batch* b = adj->b;
    const int nw = int(b->worlds.size());
    const std::size_t B = b->totB;
    if (int e = AdjointForward(adj, p)) return e;
    // Loss and its gradient with respect to the final positions.
    std::vector<float4> pos(B);
    std::vector<std::uint32_t> flags(B);
    PRX_HIP(hipMemcpy(pos.data(), b->d.pos1, B * sizeof(float4), hipMemcpyDeviceToHost));
    PRX_HIP(hipMemcpy(flags.data(), b->d.bflags, B * sizeof(std::uint32_t), hipMemcpyDeviceToHost));
    std::vector<float4> xbar(B, float4{0, 0, 0, 0}), vbar(B, float4{0, 0, 0, 0});
    for (int w = 0; w < nw; ++w) {
        const auto& h = b->hdr[std::size_t(w)];
        double l = 0.0;
        for (int i = 0; i < h.nbodies; ++i) {
            const std::size_t g = std::size_t(h.body_off + i);
            if (!(flags[g] & prx::kAccelerable)) continue;
            const float dx = pos[g].x - targets[2 * g], dy = pos[g].y - targets[2 * g + 1];
            l += double(dx) * dx + double(dy) * dy;
            xbar[g] = float4{2.0f * dx, 2.0f * dy, 0.0f, 0.0f};
        }
        if (loss) loss[w] = float(l);
    }
    return AdjointBackward(adj, p, xbar, vbar, grad_velocity, grad_friction, grad_restitution);
}

int prx_adjoint_forward(prx_adjoint* adj, const prx_step_conf* conf, float* final_state) {
    if (!adj || !conf) return Fail(PRX_E_INVALID_ARGUMENT, "null argument");
    prx::StepParams p;
    if (int e = AdjointParams(adj, conf, p)) return e;
    if (int e = AdjointForward(adj, p)) return e;
    if (final_state) {
        prx_batch* b = adj->b;
        const std::size_t B = b->totB;
        std::vector<float4> pos(B), vel(B);
        PRX_HIP(hipMemcpy(pos.data(), b->d.pos1, B * sizeof(float4), hipMemcpyDeviceToHost));
        PRX_HIP(hipMemcpy(vel.data(), b->d.vel, B * sizeof(float4), hipMemcpyDeviceToHost));
        for (std::size_t i = 0; i < B; ++i) {
            float* o = final_state + 6 * i;
            o[0] = pos[i].x;
            o[1] = pos[i].y;
            o[2] = pos[i].z;
            o[3] = vel[i].x;
            o[4] = vel[i].y;
            o[5] = vel[i].z;
        }
    }
    return PRX_OK;
}

int prx_adjoint_backward(prx_adjoint* adj, const prx_step_conf* conf, const float* state_bar, float* grad_velocity,
                         float* grad_friction, float* grad_restitution) {
    if (!adj || !conf || !state_bar) return Fail(PRX_E_INVALID_ARGUMENT, "null argument");
    if (!adj->forwarded) return Fail(PRX_E_WRONG_STATE, "no forward pass to reverse");
    prx::StepParams p;
    if (int e = AdjointParams(adj, conf, p)) return e;
    const std::size_t B = adj->b->totB;
    std::vector<float4> xbar(B), vbar(B);
    std::vector<std::uint32_t> flags(B);
    PRX_HIP(hipMemcpy(flags.data(), adj->b->d.bflags, B * sizeof(std::uint32_t), hipMemcpyDeviceToHost));
    for (std::size_t i = 0; i < B; ++i) {
        const float* s = state_bar + 6 * i;
        // Bodies that do not move carry no adjoint (their state is fixed).
        const bool dyn = (flags[i] & prx::kAccelerable) != 0;
        xbar[i] = dyn ? float4{s[0], s[1], s[2], 0.0f} : float4{0, 0, 0, 0};
        vbar[i] = dyn ? float4{s[3], s[4], s[5], 0.0f} : float4{0, 0, 0, 0};
    }
    return AdjointBackward(adj, p, xbar, vbar, grad_velocity, grad_friction, grad_restitution);
}

int prx_adjoint_use_general(prx_adjoint* adj, int on) {
    if (!adj) return Fail(PRX_E_INVALID_ARGUMENT, "null argument");
    bool needed = false;
    for (prx_world* w : adj->b->worlds) {
        if (!w) continue;
        if (!w->joints.empty()) needed = true;
        for (const auto& hb : w->bodies) {
            if (!(hb.flags & prx::kAccelerable)) continue;
            if (hb.lc.x != 0.0f || hb.lc.y != 0.0f) needed = true;
            for (std::uint32_t sid : hb.shapes) {
                const HostShape& sh = w->shapes[sid];
                if (sh.v.size() != 1 || sh.v[0].x != 0.0f || sh.v[0].y != 0.0f) needed = true;
            }
        }
    }
    if (!on && needed) return Fail(PRX_E_UNSUPPORTED, "this batch needs the general reverse kernel");
    adj->general = on != 0;
    return PRX_OK;
}

int prx_adjoint_joint_gradients(prx_adjoint* adj, float* grad_max_motor_torque) {
    if (!adj || !grad_max_motor_torque) return Fail(PRX_E_INVALID_ARGUMENT, "null argument");
    const std::size_t J = adj->b->totJ;
    if (J) PRX_HIP(hipMemcpy(grad_max_motor_torque, adj->a.tbar, J * sizeof(float), hipMemcpyDeviceToHost));
    return PRX_OK;
}

int prx_adjoint_diagnostics(prx_adjoint* adj, float out[4]) {
    if (!adj || !out) return Fail(PRX_E_INVALID_ARGUMENT, "null argument");
    PRX_HIP(hipMemcpy(out, adj->a.diag, 4 * sizeof(float), hipMemcpyDeviceToHost));
    return PRX_OK;
}

int prx_adjoint_timing(prx_adjoint* adj, float* ms) {
    if (!adj || !ms) return Fail(PRX_E_INVALID_ARGUMENT, "null argument");
    PRX_HIP(hipEventElapsedTime(ms, adj->b->ev0, adj->b->ev1));
    return PRX_OK;
}

}  // extern "C"
