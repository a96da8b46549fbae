// Scene specifications for the step workloads: plain data (shapes, bodies) built
// from a portable random stream, so that the same scene can be applied through
// any World API that follows playrho::d2 (this library's, or the reference's in
// the golden-vector driver) and come out identical.
//
// Geometry sources:
//  - "hello":     HelloWorld/HelloWorld.cpp:31-92 (ground box 50x10 at (0,-10),
//                 disk r=1 at (0,4), gravity (0,-9.8)).
//  - "pyramid":   Testbed/Tests/Pyramid.cpp:30-56 with 0.5 half-extent boxes,
//                 density 1, friction 0.6, ground edge (-40,0)-(40,0).
//  - "drop":      N bodies (half disks, half 3..8-gons) dropped into a walled box.
//  - "pile":      one world of the batched-uneven-worlds workload: n bodies
//                 (log-uniform in [nmin,nmax]) of boxes / disks / 3..8-gons in
//                 stacks or piles on a ground edge.
//  - "billiards": 64 disks r=0.5 in a walled 20x20 box with seeded velocities,
//                 friction and restitution.
#pragma once

#include <cmath>
#include <cstdint>
#include <random>
#include <string>
#include <vector>

namespace prx_scene {

enum ShapeKind : int { kDisk = 0, kPolygon = 1, kBox = 2, kEdge = 3, kChain = 4, kMulti = 5 };
enum BodyKind : int { kStatic = 0, kKinematic = 1, kDynamic = 2 };

struct Pt {
    float x = 0, y = 0;
};

struct ShapeSpec {
    int kind = kDisk;
    float radius = 0.01f;        // disk radius (vertex radius of a disk)
    float hx = 0, hy = 0;        // box half extents
    std::vector<Pt> pts;         // polygon points (hull taken), the 2 edge ends, a chain's vertices,
                                 // or a multi shape's hull points one hull after another
    std::vector<int> hullSizes;  // multi shape: points per hull
    float vertexRadius = -1.0f;  // <0: the conf's default (0.01 m)
    float density = 0, friction = 0.2f, restitution = 0;
};

struct BodySpec {
    int type = kStatic;
    float x = 0, y = 0, angle = 0;
    float vx = 0, vy = 0, w = 0;
    float ax = 0, ay = 0;
    std::vector<int> shapes;
};

// Joint between bodies a and b with local anchors: kind 0 RevoluteJointConf,
// 1 DistanceJointConf (length, frequency, dampingRatio), 2 WeldJointConf
// (referenceAngle, frequency, dampingRatio), 3 RopeJointConf (length is the
// maximum length), 4 FrictionJointConf (maxForce, maxTorque), 5 MotorJointConf
// (linear offset lax, lay; angular offset referenceAngle; maxForce,
// maxTorque, correctionFactor), 6 WheelJointConf (anchors, axis at axisAngle,
// frequency, dampingRatio, enableMotor, motorSpeed, maxMotorTorque), 7
// PulleyJointConf (anchors, ground anchors gax, gay, gbx, gby, lengths
// length, lengthB, ratio), 8 PrismaticJointConf (anchors, axis at axisAngle,
// referenceAngle, enableLimit, lower, upper, enableMotor, motorSpeed,
// maxForce as the maximum motor force), 9 TargetJointConf (body b only; the
// target gax, gay, the anchor on B lbx, lby, maxForce, frequency,
// dampingRatio), 10 GearJointConf (a, b: indices of two earlier revolute or
// prismatic joints, made by GetGearJointConf with `ratio`).
enum JointKind : int {
    kRevolute = 0, kDistance = 1, kWeld = 2, kRope = 3, kFriction = 4, kMotor = 5, kWheel = 6, kPulley = 7,
    kPrismatic = 8, kTarget = 9, kGear = 10
};
struct JointSpec {
    int kind = kRevolute;
    int a = 0, b = 0;
    float length = 1.0f, frequency = 0.0f, dampingRatio = 0.0f;
    float maxForce = 0.0f, maxTorque = 0.0f, correctionFactor = 0.3f, axisAngle = 0.0f;
    float gax = 0, gay = 0, gbx = 0, gby = 0, lengthB = 0, ratio = 1;
    float lax = 0, lay = 0, lbx = 0, lby = 0;
    float referenceAngle = 0;
    bool enableLimit = false;
    float lower = 0, upper = 0;
    bool enableMotor = false;
    float motorSpeed = 0, maxMotorTorque = 0;
    bool collideConnected = false;
};

struct SceneSpec {
    std::vector<ShapeSpec> shapes;
    std::vector<BodySpec> bodies;
    std::vector<JointSpec> joints;  // created after all bodies, in order
    // Order in which the scene was authored: shape ids are allocated in
    // ShapeSpec order and body ids in BodySpec order; when interleave is true
    // shape i+1... are created right before body i uses them (see Apply*).
};

// Deterministic uniform floats from std::mt19937 (whose output sequence the C++
// standard fixes); the mapping to [a,b) is done here so no library-specific
// distribution is involved.
struct Rng {
    std::mt19937 gen;
    explicit Rng(std::uint32_t seed) : gen(seed) {}
    double u01() { return double(gen() >> 8) * (1.0 / 16777216.0); }
    float uniform(double a, double b) { return float(a + (b - a) * u01()); }
    int below(int n) { return int(gen() % std::uint32_t(n)); }
    double normal() {
        // Box-Muller on two draws; deterministic given the draws.
        double u1 = u01();
        double u2 = u01();
        if (u1 < 1e-12) u1 = 1e-12;
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
};

inline ShapeSpec Edge(float x0, float y0, float x1, float y1) {
    ShapeSpec s;
    s.kind = kEdge;
    s.pts = {Pt{x0, y0}, Pt{x1, y1}};
    return s;
}

inline ShapeSpec Box(float hx, float hy, float density, float friction, float restitution) {
    ShapeSpec s;
    s.kind = kBox;
    s.hx = hx;
    s.hy = hy;
    s.density = density;
    s.friction = friction;
    s.restitution = restitution;
    return s;
}

inline ShapeSpec Disk(float r, float density, float friction, float restitution) {
    ShapeSpec s;
    s.kind = kDisk;
    s.radius = r;
    s.density = density;
    s.friction = friction;
    s.restitution = restitution;
    return s;
}

inline ShapeSpec Ngon(Rng& rng, int n, float R, float density, float friction) {
    std::vector<double> ang(n);
    for (auto& a : ang) a = rng.u01() * 6.283185307179586;
    // Sort so the points go round the circle; the hull keeps all of them unless
    // two coincide after rounding.
    for (int i = 1; i < n; ++i)
        for (int j = i; j > 0 && ang[j] < ang[j - 1]; --j) std::swap(ang[j], ang[j - 1]);
    ShapeSpec s;
    s.kind = kPolygon;
    s.density = density;
    s.friction = friction;
    for (double a : ang) s.pts.push_back(Pt{float(R * std::cos(a)), float(R * std::sin(a))});
    return s;
}

inline SceneSpec Hello() {
    SceneSpec sc;
    sc.shapes.push_back(Box(50.0f, 10.0f, 0.0f, 0.2f, 0.0f));
    sc.shapes.push_back(Disk(1.0f, 0.0f, 0.2f, 0.0f));
    BodySpec ground;
    ground.type = kStatic;
    ground.x = 0.0f;
    ground.y = -10.0f;
    ground.shapes = {0};
    BodySpec ball;
    ball.type = kDynamic;
    ball.x = 0.0f;
    ball.y = 4.0f;
    ball.ay = -9.8f;
    ball.shapes = {1};
    sc.bodies = {ground, ball};
    return sc;
}

inline SceneSpec Pyramid(int rows = 20) {
    SceneSpec sc;
    sc.shapes.push_back(Edge(-40.0f, 0.0f, 40.0f, 0.0f));
    sc.shapes.push_back(Box(0.5f, 0.5f, 1.0f, 0.6f, 0.0f));
    BodySpec ground;
    ground.shapes = {0};
    sc.bodies.push_back(ground);
    float xx = -7.0f, xy = 0.75f;
    const float dxx = 0.5625f, dxy = 1.25f;
    const float dyx = 1.125f, dyy = 0.0f;
    for (int i = 0; i < rows; ++i) {
        float yx = xx, yy = xy;
        for (int j = i; j < rows; ++j) {
            BodySpec b;
            b.type = kDynamic;
            b.x = yx;
            b.y = yy;
            b.ay = -10.0f;
            b.shapes = {1};
            sc.bodies.push_back(b);
            yx += dyx;
            yy += dyy;
        }
        xx += dxx;
        xy += dxy;
    }
    return sc;
}

// Walled container: one static body with four edges around [-half,half]x[0,height].
inline void AddContainer(SceneSpec& sc, float half, float height) {
    const int s0 = int(sc.shapes.size());
    sc.shapes.push_back(Edge(-half, 0.0f, half, 0.0f));
    sc.shapes.push_back(Edge(half, 0.0f, half, height));
    sc.shapes.push_back(Edge(half, height, -half, height));
    sc.shapes.push_back(Edge(-half, height, -half, 0.0f));
    BodySpec c;
    c.shapes = {s0, s0 + 1, s0 + 2, s0 + 3};
    sc.bodies.push_back(c);
}

inline SceneSpec Drop(int n, float half, float height, std::uint32_t seed) {
    SceneSpec sc;
    Rng rng(seed);
    AddContainer(sc, half, height);
    for (int i = 0; i < n; ++i) {
        BodySpec b;
        b.type = kDynamic;
        b.x = rng.uniform(-half + 1.0, half - 1.0);
        b.y = rng.uniform(1.0, height - 1.0);
        b.ay = -10.0f;
        const float friction = rng.uniform(0.2, 0.8);
        if ((i % 2) == 0) {
            sc.shapes.push_back(Disk(rng.uniform(0.25, 0.75), 1.0f, friction, 0.0f));
        } else {
            const int nv = 3 + rng.below(6);
            const float R = rng.uniform(0.3, 0.8);
            sc.shapes.push_back(Ngon(rng, nv, R, 1.0f, friction));
        }
        b.shapes = {int(sc.shapes.size()) - 1};
        sc.bodies.push_back(b);
    }
    return sc;
}

// Body count of world w of the batched workload: round(exp(U[ln nmin, ln nmax])).
inline int PileBodyCount(std::uint32_t seed, int w, int nmin, int nmax) {
    Rng rng(seed * 1000003u + std::uint32_t(w) * 7919u + 17u);
    const double lo = std::log(double(nmin)), hi = std::log(double(nmax));
    return int(std::lround(std::exp(lo + (hi - lo) * rng.u01())));
}

inline SceneSpec Pile(std::uint32_t seed, int w, int nmin = 8, int nmax = 512) {
    SceneSpec sc;
    const int n = PileBodyCount(seed, w, nmin, nmax);
    Rng rng(seed * 2654435761u + std::uint32_t(w) * 40503u + 3u);
    const bool pile = rng.below(2) == 1;  // else: stacks
    int ncols = int(std::lround(std::sqrt(double(n)) * 0.8));
    if (ncols < 1) ncols = 1;
    const float spacing = 1.25f;
    const float halfw = float(ncols) * spacing * 0.5f + 5.0f;
    sc.shapes.push_back(Edge(-halfw, 0.0f, halfw, 0.0f));
    BodySpec ground;
    ground.shapes = {0};
    sc.bodies.push_back(ground);
    for (int k = 0; k < n; ++k) {
        const int col = k % ncols, row = k / ncols;
        BodySpec b;
        b.type = kDynamic;
        b.x = float((double(col) - double(ncols - 1) * 0.5) * spacing);
        b.y = float(0.6 + row * 1.15);
        if (pile) {
            b.x += rng.uniform(-0.15, 0.15);
            b.y += rng.uniform(0.0, 0.1);
            b.angle = rng.uniform(-0.3, 0.3);
        }
        b.ay = -10.0f;
        const float friction = rng.uniform(0.2, 0.8);
        const int kind = rng.below(3);
        if (kind == 0) {
            // Draws are sequenced explicitly: argument evaluation order is
            // unspecified in C++ and differs between compilers.
            const float hx = rng.uniform(0.3, 0.5);
            const float hy = rng.uniform(0.3, 0.5);
            sc.shapes.push_back(Box(hx, hy, 1.0f, friction, 0.0f));
        } else if (kind == 1) {
            sc.shapes.push_back(Disk(rng.uniform(0.3, 0.5), 1.0f, friction, 0.0f));
        } else {
            const int nv = 3 + rng.below(6);
            const float R = rng.uniform(0.35, 0.55);
            sc.shapes.push_back(Ngon(rng, nv, R, 1.0f, friction));
        }
        b.shapes = {int(sc.shapes.size()) - 1};
        sc.bodies.push_back(b);
    }
    return sc;
}

inline SceneSpec Billiards(std::uint32_t seed, int w) {
    SceneSpec sc;
    Rng rng(seed * 69069u + std::uint32_t(w) * 1664525u + 1013904223u);
    AddContainer(sc, 10.0f, 20.0f);
    for (int k = 0; k < 64; ++k) {
        const int col = k % 8, row = k / 8;
        BodySpec b;
        b.type = kDynamic;
        b.x = float(-7.7 + col * 2.2) + rng.uniform(-0.2, 0.2);
        b.y = float(1.5 + row * 2.2) + rng.uniform(-0.2, 0.2);
        b.vx = float(3.0 * rng.normal());
        b.vy = float(3.0 * rng.normal());
        b.ay = -10.0f;
        const float friction = rng.uniform(0.0, 1.0);
        const float restitution = rng.uniform(0.0, 1.0);
        sc.shapes.push_back(Disk(0.5f, 1.0f, friction, restitution));
        b.shapes = {int(sc.shapes.size()) - 1};
        sc.bodies.push_back(b);
    }
    return sc;
}

// Articulated world: a static anchor at (0, 40) and a chain of `links` boxes
// (half extents 1 x 0.125, density 1) joined end to end by revolute joints
// (anchor -> link 0, link i -> link i+1, collideConnected false), laid out
// horizontally so it swings down, plus `falling` seeded boxes and disks
// dropped onto it and onto a ground edge.  Joint motors act as joint friction
// (motor speed 0, maximum torque `torque` N m); every third joint has a
// +-60 degree limit.
inline SceneSpec Chain(std::uint32_t seed, int w, int links = 32, int falling = 64, float torque = 2.0f) {
    SceneSpec sc;
    Rng rng(seed * 362437u + std::uint32_t(w) * 521288629u + 7u);
    sc.shapes.push_back(Edge(-80.0f, 0.0f, 80.0f, 0.0f));
    BodySpec ground;
    ground.shapes = {0};
    sc.bodies.push_back(ground);
    BodySpec anchor;
    anchor.x = 0.0f;
    anchor.y = 40.0f;
    sc.bodies.push_back(anchor);
    sc.shapes.push_back(Box(1.0f, 0.125f, 1.0f, 0.4f, 0.0f));
    const int linkShape = int(sc.shapes.size()) - 1;
    for (int i = 0; i < links; ++i) {
        BodySpec b;
        b.type = kDynamic;
        b.x = float(2 * i + 1);
        b.y = 40.0f;
        b.ay = -10.0f;
        b.shapes = {linkShape};
        sc.bodies.push_back(b);
        JointSpec j;
        j.a = (i == 0) ? 1 : 1 + i;
        j.b = 2 + i;
        j.lax = (i == 0) ? 0.0f : 1.0f;
        j.lbx = -1.0f;
        j.enableMotor = true;
        j.motorSpeed = 0.0f;
        j.maxMotorTorque = torque;
        if (i % 3 == 2) {
            j.enableLimit = true;
            j.lower = -1.04719758f;
            j.upper = 1.04719758f;
        }
        sc.joints.push_back(j);
    }
    for (int k = 0; k < falling; ++k) {
        BodySpec b;
        b.type = kDynamic;
        b.x = rng.uniform(-30.0, 60.0);
        b.y = rng.uniform(45.0, 90.0);
        b.ay = -10.0f;
        const float friction = rng.uniform(0.2, 0.8);
        if (rng.below(2) == 0) {
            const float hx = rng.uniform(0.3, 0.6);
            const float hy = rng.uniform(0.3, 0.6);
            sc.shapes.push_back(Box(hx, hy, 1.0f, friction, 0.0f));
        } else {
            sc.shapes.push_back(Disk(rng.uniform(0.3, 0.6), 1.0f, friction, 0.0f));
        }
        b.shapes = {int(sc.shapes.size()) - 1};
        sc.bodies.push_back(b);
    }
    return sc;
}

// Distance and weld joints: `n` pendulums hang from a static bar at (0, 30)
// by distance joints (every other one a spring: frequency 2-4 Hz, damping
// 0.1-0.7), each pendulum bob welded to a second box (every third weld soft:
// 3 Hz, damping 0.5); a welded 4-box beam rests on the ground; `falling`
// seeded boxes and disks drop onto them and onto a ground edge.
inline SceneSpec Joints(std::uint32_t seed, int w, int n = 8, int falling = 24) {
    SceneSpec sc;
    Rng rng(seed * 2654435761u + std::uint32_t(w) * 40503u + 11u);
    sc.shapes.push_back(Edge(-60.0f, 0.0f, 60.0f, 0.0f));
    BodySpec ground;
    ground.shapes = {0};
    sc.bodies.push_back(ground);
    BodySpec bar;
    bar.x = 0.0f;
    bar.y = 30.0f;
    sc.bodies.push_back(bar);
    sc.shapes.push_back(Box(0.5f, 0.5f, 1.0f, 0.5f, 0.0f));
    const int bob = int(sc.shapes.size()) - 1;
    sc.shapes.push_back(Box(0.4f, 0.25f, 1.0f, 0.5f, 0.0f));
    const int tail = int(sc.shapes.size()) - 1;
    for (int i = 0; i < n; ++i) {
        const float x = float(3 * i) - 1.5f * float(n);
        BodySpec b;
        b.type = kDynamic;
        b.x = x + rng.uniform(-2.0, 2.0);
        b.y = 30.0f - rng.uniform(4.0, 9.0);
        b.ay = -10.0f;
        b.shapes = {bob};
        sc.bodies.push_back(b);
        const int ib = int(sc.bodies.size()) - 1;
        JointSpec d;
        d.kind = kDistance;
        d.a = 1;
        d.b = ib;
        d.lax = x;
        d.length = rng.uniform(5.0, 9.0);
        if (i % 2 == 1) {
            d.frequency = rng.uniform(2.0, 4.0);
            d.dampingRatio = rng.uniform(0.1, 0.7);
        }
        sc.joints.push_back(d);
        BodySpec t;
        t.type = kDynamic;
        t.x = b.x + 0.9f;
        t.y = b.y;
        t.ay = -10.0f;
        t.shapes = {tail};
        sc.bodies.push_back(t);
        JointSpec wj;
        wj.kind = kWeld;
        wj.a = ib;
        wj.b = int(sc.bodies.size()) - 1;
        wj.lax = 0.5f;
        wj.lbx = -0.4f;
        if (i % 3 == 2) {
            wj.frequency = 3.0f;
            wj.dampingRatio = 0.5f;
        }
        sc.joints.push_back(wj);
    }
    // A cart: a box chassis on two disk wheels (wheel joints along a tilted
    // axis, soft or rigid, one with a motor).
    {
        sc.shapes.push_back(Box(1.2f, 0.25f, 1.0f, 0.6f, 0.0f));
        const int chassisShape = int(sc.shapes.size()) - 1;
        sc.shapes.push_back(Disk(0.35f, 1.0f, 0.9f, 0.0f));
        const int wheelShape = int(sc.shapes.size()) - 1;
        BodySpec ch;
        ch.type = kDynamic;
        ch.x = -30.0f;
        ch.y = 1.4f;
        ch.ay = -10.0f;
        ch.shapes = {chassisShape};
        sc.bodies.push_back(ch);
        const int chassis = int(sc.bodies.size()) - 1;
        for (int k = 0; k < 2; ++k) {
            BodySpec wb;
            wb.type = kDynamic;
            wb.x = -30.0f + (k ? 0.9f : -0.9f);
            wb.y = 0.5f;
            wb.ay = -10.0f;
            wb.shapes = {wheelShape};
            sc.bodies.push_back(wb);
            JointSpec wj;
            wj.kind = kWheel;
            wj.a = chassis;
            wj.b = int(sc.bodies.size()) - 1;
            wj.lax = k ? 0.9f : -0.9f;
            wj.lay = -0.9f;
            wj.axisAngle = 1.5707963f + (k ? 0.2f : -0.1f);
            wj.frequency = k ? 4.0f : 0.0f;
            wj.dampingRatio = 0.7f;
            wj.enableMotor = (k == 1);
            wj.motorSpeed = -6.0f;
            wj.maxMotorTorque = 8.0f;
            sc.joints.push_back(wj);
        }
    }
    // Two boxes on a pulley (ratio 1.5) from ground anchors at y = 28.
    {
        sc.shapes.push_back(Box(0.5f, 0.5f, 1.0f, 0.5f, 0.0f));
        const int boxShape = int(sc.shapes.size()) - 1;
        const float xs[2] = {35.0f, 40.0f}, ys[2] = {20.0f, 22.0f};
        for (int k = 0; k < 2; ++k) {
            BodySpec pb;
            pb.type = kDynamic;
            pb.x = xs[k];
            pb.y = ys[k];
            pb.ay = -10.0f;
            pb.shapes = {boxShape};
            sc.bodies.push_back(pb);
        }
        JointSpec pj;
        pj.kind = kPulley;
        pj.a = int(sc.bodies.size()) - 2;
        pj.b = int(sc.bodies.size()) - 1;
        pj.gax = xs[0];
        pj.gay = 28.0f;
        pj.gbx = xs[1];
        pj.gby = 28.0f;
        pj.length = 28.0f - ys[0];
        pj.lengthB = 28.0f - ys[1];
        pj.ratio = 1.5f;
        pj.collideConnected = true;
        sc.joints.push_back(pj);
    }
    for (int k = 0; k < 4; ++k) {
        BodySpec b;
        b.type = kDynamic;
        b.x = 20.0f + 0.8f * float(k);
        b.y = 0.3f;
        b.ay = -10.0f;
        b.shapes = {tail};
        sc.bodies.push_back(b);
        if (k > 0) {
            JointSpec wj;
            wj.kind = kWeld;
            wj.a = int(sc.bodies.size()) - 2;
            wj.b = int(sc.bodies.size()) - 1;
            wj.lax = 0.4f;
            wj.lbx = -0.4f;
            sc.joints.push_back(wj);
        }
        // Top-down friction against the ground (which the box also touches).
        JointSpec fj;
        fj.kind = kFriction;
        fj.a = 0;
        fj.b = int(sc.bodies.size()) - 1;
        fj.lax = b.x;
        fj.lay = b.y;
        fj.maxForce = 1.5f + 0.5f * float(k);
        fj.maxTorque = 0.25f * float(k);
        fj.collideConnected = true;
        sc.joints.push_back(fj);
    }
    for (int k = 0; k < falling; ++k) {
        BodySpec b;
        b.type = kDynamic;
        b.x = rng.uniform(-25.0, 25.0);
        b.y = rng.uniform(32.0, 60.0);
        b.ay = -10.0f;
        const float friction = rng.uniform(0.2, 0.8);
        if (rng.below(2) == 0) {
            // Drawn in sequence: argument evaluation order differs between compilers.
            const float hx = rng.uniform(0.3, 0.6);
            const float hy = rng.uniform(0.3, 0.6);
            sc.shapes.push_back(Box(hx, hy, 1.0f, friction, 0.0f));
        } else {
            sc.shapes.push_back(Disk(rng.uniform(0.3, 0.6), 1.0f, friction, 0.0f));
        }
        b.shapes = {int(sc.shapes.size()) - 1};
        sc.bodies.push_back(b);
        if (k == 1 || k == 2) {
            // Motor joints from the ground pulling the body towards an offset.
            JointSpec mj;
            mj.kind = kMotor;
            mj.a = 0;
            mj.b = int(sc.bodies.size()) - 1;
            mj.lax = -10.0f * float(k);
            mj.lay = 8.0f;
            mj.referenceAngle = 0.5f * float(k);
            mj.maxForce = 40.0f * float(k);
            mj.maxTorque = 3.0f;
            mj.correctionFactor = 0.2f + 0.1f * float(k);
            sc.joints.push_back(mj);
        }
        if (k % 3 == 0) {
            // A rope from the bar, slack at first, taut once the body has
            // fallen 2 to 6 m below the bar.
            JointSpec rj;
            rj.kind = kRope;
            rj.a = 1;
            rj.b = int(sc.bodies.size()) - 1;
            rj.lax = b.x;
            rj.length = (b.y - 30.0f) + 2.0f + float(k % 5);
            sc.joints.push_back(rj);
        }
    }
    return sc;
}

// Prismatic joints: `n` slabs slide on rails from the ground body (axes at
// seeded angles; cycling through a translation limit, a motor, and both; the
// last with equal limits), a carriage on a limited horizontal rail carries a
// motorised vertical slider (a joint between two dynamic bodies), and
// `falling` seeded boxes and disks drop onto them and onto a ground edge.
inline SceneSpec Sliders(std::uint32_t seed, int w, int n = 6, int falling = 24) {
    SceneSpec sc;
    Rng rng(seed * 2654435761u + std::uint32_t(w) * 40503u + 29u);
    sc.shapes.push_back(Edge(-50.0f, 0.0f, 50.0f, 0.0f));
    BodySpec ground;
    ground.shapes = {0};
    sc.bodies.push_back(ground);
    sc.shapes.push_back(Box(0.6f, 0.3f, 1.0f, 0.5f, 0.0f));
    const int slab = int(sc.shapes.size()) - 1;
    for (int i = 0; i < n; ++i) {
        BodySpec b;
        b.type = kDynamic;
        b.x = float(6 * i) - 3.0f * float(n);
        b.y = 4.0f + rng.uniform(0.0, 3.0);
        b.angle = rng.uniform(-0.2, 0.2);
        b.ay = -10.0f;
        b.shapes = {slab};
        sc.bodies.push_back(b);
        JointSpec pj;
        pj.kind = kPrismatic;
        pj.a = 0;
        pj.b = int(sc.bodies.size()) - 1;
        pj.lax = b.x;
        pj.lay = b.y;
        pj.axisAngle = rng.uniform(0.0, 3.14159);
        pj.referenceAngle = b.angle;
        const int m = i % 3;
        if (m != 1) {
            pj.enableLimit = true;
            pj.lower = -rng.uniform(0.5, 2.0);
            pj.upper = rng.uniform(0.5, 2.0);
        }
        if (m != 0) {
            pj.enableMotor = true;
            pj.motorSpeed = rng.uniform(-3.0, 3.0);
            pj.maxForce = rng.uniform(20.0, 200.0);
        }
        if (i == n - 1) {
            pj.enableLimit = true;
            pj.lower = 0.25f;
            pj.upper = 0.25f;
        }
        sc.joints.push_back(pj);
    }
    {
        sc.shapes.push_back(Box(1.5f, 0.3f, 1.0f, 0.6f, 0.0f));
        const int carriageShape = int(sc.shapes.size()) - 1;
        sc.shapes.push_back(Box(0.4f, 0.4f, 1.0f, 0.6f, 0.0f));
        const int sliderShape = int(sc.shapes.size()) - 1;
        BodySpec cb;
        cb.type = kDynamic;
        cb.x = 0.0f;
        cb.y = 12.0f;
        cb.ay = -10.0f;
        cb.shapes = {carriageShape};
        sc.bodies.push_back(cb);
        const int carriage = int(sc.bodies.size()) - 1;
        JointSpec rail;
        rail.kind = kPrismatic;
        rail.a = 0;
        rail.b = carriage;
        rail.lax = 0.0f;
        rail.lay = 12.0f;
        rail.enableLimit = true;
        rail.lower = -8.0f;
        rail.upper = 8.0f;
        sc.joints.push_back(rail);
        BodySpec sb;
        sb.type = kDynamic;
        sb.x = 0.5f;
        sb.y = 13.0f;
        sb.ay = -10.0f;
        sb.shapes = {sliderShape};
        sc.bodies.push_back(sb);
        JointSpec up;
        up.kind = kPrismatic;
        up.a = carriage;
        up.b = int(sc.bodies.size()) - 1;
        up.lax = 0.5f;
        up.lay = 1.0f;
        up.axisAngle = 1.5707963f;
        up.enableLimit = true;
        up.lower = -0.5f;
        up.upper = 2.5f;
        up.enableMotor = true;
        up.motorSpeed = 1.5f;
        up.maxForce = 30.0f;
        sc.joints.push_back(up);
    }
    for (int k = 0; k < falling; ++k) {
        BodySpec b;
        b.type = kDynamic;
        b.x = rng.uniform(-20.0, 20.0);
        b.y = rng.uniform(15.0, 40.0);
        b.ay = -10.0f;
        const float friction = rng.uniform(0.2, 0.8);
        if (rng.below(2) == 0) {
            const float hx = rng.uniform(0.3, 0.6);
            const float hy = rng.uniform(0.3, 0.6);
            sc.shapes.push_back(Box(hx, hy, 1.0f, friction, 0.0f));
        } else {
            sc.shapes.push_back(Disk(rng.uniform(0.3, 0.6), 1.0f, friction, 0.0f));
        }
        b.shapes = {int(sc.shapes.size()) - 1};
        sc.bodies.push_back(b);
    }
    return sc;
}

// Target joints: `n` bodies (boxes and disks) are pulled towards seeded
// targets by target joints with seeded anchors, maximum forces (some small
// enough to clamp), frequencies and damping ratios (one with frequency 0, one
// with maximum force 0), beside `falling` seeded bodies dropped onto them and
// onto a ground edge.
inline SceneSpec Targets(std::uint32_t seed, int w, int n = 8, int falling = 24) {
    SceneSpec sc;
    Rng rng(seed * 2654435761u + std::uint32_t(w) * 40503u + 31u);
    sc.shapes.push_back(Edge(-50.0f, 0.0f, 50.0f, 0.0f));
    BodySpec ground;
    ground.shapes = {0};
    sc.bodies.push_back(ground);
    auto addBody = [&](float x, float y) {
        BodySpec b;
        b.type = kDynamic;
        b.x = x;
        b.y = y;
        b.angle = rng.uniform(-0.5, 0.5);
        b.ay = -10.0f;
        const float friction = rng.uniform(0.2, 0.8);
        if (rng.below(2) == 0) {
            const float hx = rng.uniform(0.3, 0.6);
            const float hy = rng.uniform(0.3, 0.6);
            sc.shapes.push_back(Box(hx, hy, 1.0f, friction, 0.0f));
        } else {
            sc.shapes.push_back(Disk(rng.uniform(0.3, 0.6), 1.0f, friction, 0.0f));
        }
        b.shapes = {int(sc.shapes.size()) - 1};
        sc.bodies.push_back(b);
        return int(sc.bodies.size()) - 1;
    };
    for (int i = 0; i < n; ++i) {
        const float x = float(5 * i) - 2.5f * float(n);
        const float y = rng.uniform(1.0, 6.0);
        JointSpec tj;
        tj.kind = kTarget;
        tj.b = addBody(x, y);
        tj.a = -1;
        tj.gax = x + rng.uniform(-3.0, 3.0);
        tj.gay = rng.uniform(2.0, 12.0);
        tj.lbx = rng.uniform(-0.3, 0.3);
        tj.lby = rng.uniform(-0.3, 0.3);
        tj.maxForce = rng.uniform(5.0, 400.0);
        tj.frequency = rng.uniform(1.0, 6.0);
        tj.dampingRatio = rng.uniform(0.2, 1.0);
        if (i == 1) tj.frequency = 0.0f;
        if (i == 2) tj.maxForce = 0.0f;
        sc.joints.push_back(tj);
    }
    for (int k = 0; k < falling; ++k) {
        // Drawn in sequence: argument evaluation order differs between compilers.
        const float fx = rng.uniform(-20.0, 20.0);
        const float fy = rng.uniform(15.0, 40.0);
        addBody(fx, fy);
    }
    return sc;
}

// Terrain: a static ground holding a chain shape (a seeded zig-zag line of
// 16 segments) and a closed chain (a square outline), dynamic bodies made of
// multi shapes (L, T and cross outlines of two or three box hulls), and
// `falling` seeded boxes and disks.
inline SceneSpec Terrain(std::uint32_t seed, int w, int falling = 40) {
    SceneSpec sc;
    Rng rng(seed * 2654435761u + std::uint32_t(w) * 40503u + 37u);
    ShapeSpec line;
    line.kind = kChain;
    for (int i = 0; i <= 16; ++i) {
        const float y = (i == 0 || i == 16) ? 12.0f : rng.uniform(0.0, 3.0);
        line.pts.push_back(Pt{-32.0f + 4.0f * float(i), y});
    }
    sc.shapes.push_back(line);
    ShapeSpec loop;
    loop.kind = kChain;
    loop.pts = {Pt{8.0f, 6.0f}, Pt{11.0f, 6.0f}, Pt{11.0f, 9.0f}, Pt{8.0f, 9.0f}, Pt{8.0f, 6.0f}};
    sc.shapes.push_back(loop);
    BodySpec ground;
    ground.shapes = {0, 1};
    sc.bodies.push_back(ground);
    auto hullBox = [](ShapeSpec& m, float cx, float cy, float hx, float hy) {
        m.pts.push_back(Pt{cx - hx, cy - hy});
        m.pts.push_back(Pt{cx + hx, cy - hy});
        m.pts.push_back(Pt{cx + hx, cy + hy});
        m.pts.push_back(Pt{cx - hx, cy + hy});
        m.hullSizes.push_back(4);
    };
    for (int k = 0; k < 6; ++k) {
        ShapeSpec m;
        m.kind = kMulti;
        m.density = 1.0f;
        m.friction = rng.uniform(0.2, 0.8);
        const int form = k % 3;
        if (form == 0) {  // L
            hullBox(m, 0.0f, 0.0f, 1.0f, 0.2f);
            hullBox(m, -0.8f, 0.6f, 0.2f, 0.4f);
        } else if (form == 1) {  // T
            hullBox(m, 0.0f, 0.5f, 0.9f, 0.2f);
            hullBox(m, 0.0f, -0.1f, 0.2f, 0.4f);
        } else {  // cross
            hullBox(m, 0.0f, 0.0f, 0.8f, 0.15f);
            hullBox(m, 0.0f, 0.0f, 0.15f, 0.8f);
            hullBox(m, 0.5f, 0.5f, 0.2f, 0.2f);
        }
        sc.shapes.push_back(m);
        BodySpec b;
        b.type = kDynamic;
        b.x = -24.0f + 8.0f * float(k);
        b.y = rng.uniform(8.0, 14.0);
        b.angle = rng.uniform(-1.0, 1.0);
        b.ay = -10.0f;
        b.shapes = {int(sc.shapes.size()) - 1};
        sc.bodies.push_back(b);
    }
    for (int k = 0; k < falling; ++k) {
        BodySpec b;
        b.type = kDynamic;
        // Drawn in sequence: argument evaluation order differs between compilers.
        b.x = rng.uniform(-28.0, 28.0);
        b.y = rng.uniform(14.0, 40.0);
        b.angle = rng.uniform(-0.5, 0.5);
        b.ay = -10.0f;
        const float friction = rng.uniform(0.2, 0.8);
        if (rng.below(2) == 0) {
            const float hx = rng.uniform(0.2, 0.6);
            const float hy = rng.uniform(0.2, 0.6);
            sc.shapes.push_back(Box(hx, hy, 1.0f, friction, 0.0f));
        } else {
            sc.shapes.push_back(Disk(rng.uniform(0.2, 0.6), 1.0f, friction, 0.0f));
        }
        b.shapes = {int(sc.shapes.size()) - 1};
        sc.bodies.push_back(b);
    }
    return sc;
}

// Gear joints: a train of three seeded wheels on revolute joints to the
// ground (seeded angles and reference angles, the first motorised) meshed by
// gears of seeded ratios, the last wheel driving a rack on a horizontal
// prismatic joint (revolute-prismatic gear), two slabs on prismatic rails
// coupled to each other (prismatic-prismatic gear, one rail limited), a
// pinion on a revolute joint to a carriage that slides on a limited rail
// geared to the pinion (the gear's bodies B and C are the same carriage), and
// `falling` seeded boxes and disks dropped onto them.
inline SceneSpec Gears(std::uint32_t seed, int w, int falling = 24) {
    SceneSpec sc;
    Rng rng(seed * 2654435761u + std::uint32_t(w) * 40503u + 41u);
    sc.shapes.push_back(Edge(-50.0f, 0.0f, 50.0f, 0.0f));
    BodySpec ground;
    ground.shapes = {0};
    sc.bodies.push_back(ground);
    auto dynamic = [&](float x, float y, float angle, int shape) {
        BodySpec b;
        b.type = kDynamic;
        b.x = x;
        b.y = y;
        b.angle = angle;
        b.ay = -10.0f;
        b.shapes = {shape};
        sc.bodies.push_back(b);
        return int(sc.bodies.size()) - 1;
    };
    auto revolute = [&](int a, int b, float ax, float ay, float ref) {
        JointSpec j;
        j.kind = kRevolute;
        j.a = a;
        j.b = b;
        j.lax = ax;
        j.lay = ay;
        j.referenceAngle = ref;
        sc.joints.push_back(j);
        return int(sc.joints.size()) - 1;
    };
    auto prismatic = [&](int a, int b, float ax, float ay, float axisAngle) {
        JointSpec j;
        j.kind = kPrismatic;
        j.a = a;
        j.b = b;
        j.lax = ax;
        j.lay = ay;
        j.axisAngle = axisAngle;
        sc.joints.push_back(j);
        return int(sc.joints.size()) - 1;
    };
    auto gear = [&](int j1, int j2, float ratio) {
        JointSpec j;
        j.kind = kGear;
        j.a = j1;
        j.b = j2;
        j.ratio = ratio;
        sc.joints.push_back(j);
    };
    // Wheel train.
    float radius[3];
    int wheelJoint[3];
    for (int i = 0; i < 3; ++i) {
        radius[i] = rng.uniform(0.8, 2.0);
        sc.shapes.push_back(Disk(radius[i], 1.0f, 0.4f, 0.0f));
        const float x = -8.0f + 5.0f * float(i);
        const float angle = rng.uniform(-1.0, 1.0);
        const int b = dynamic(x, 8.0f, angle, int(sc.shapes.size()) - 1);
        wheelJoint[i] = revolute(0, b, x, 8.0f, angle + rng.uniform(-0.5, 0.5));
        if (i == 0) {
            sc.joints.back().enableMotor = true;
            sc.joints.back().motorSpeed = rng.uniform(-2.0, 2.0);
            sc.joints.back().maxMotorTorque = rng.uniform(50.0, 400.0);
        }
        if (i > 0) gear(wheelJoint[i - 1], wheelJoint[i], radius[i - 1] / radius[i]);
    }
    // Rack under the last wheel.
    sc.shapes.push_back(Box(3.0f, 0.25f, 1.0f, 0.5f, 0.0f));
    const int rack = dynamic(2.0f, 8.0f - radius[2] - 0.3f, 0.0f, int(sc.shapes.size()) - 1);
    const int rackJoint = prismatic(0, rack, 2.0f, 8.0f - radius[2] - 0.3f, 0.0f);
    gear(wheelJoint[2], rackJoint, -1.0f / radius[2]);
    // Two slabs on coupled rails.
    sc.shapes.push_back(Box(0.7f, 0.3f, 1.0f, 0.5f, 0.0f));
    const int slabShape = int(sc.shapes.size()) - 1;
    const int s1 = dynamic(14.0f, 5.0f, 0.0f, slabShape);
    const int p1 = prismatic(0, s1, 14.0f, 5.0f, 1.5707963f);
    sc.joints.back().enableLimit = true;
    sc.joints.back().lower = -2.0f;
    sc.joints.back().upper = 2.0f;
    const float a2 = rng.uniform(0.3, 2.8);
    const int s2 = dynamic(20.0f, 6.0f, rng.uniform(-0.3, 0.3), slabShape);
    const int p2 = prismatic(0, s2, 20.0f, 6.0f, a2);
    gear(p1, p2, rng.uniform(0.5, 2.0));
    // Pinion on a carriage.
    sc.shapes.push_back(Box(1.5f, 0.3f, 1.0f, 0.6f, 0.0f));
    const int carriage = dynamic(-20.0f, 3.0f, 0.0f, int(sc.shapes.size()) - 1);
    const int railJoint = prismatic(0, carriage, -20.0f, 3.0f, 0.0f);
    sc.joints.back().enableLimit = true;
    sc.joints.back().lower = -5.0f;
    sc.joints.back().upper = 5.0f;
    sc.shapes.push_back(Disk(0.6f, 1.0f, 0.5f, 0.0f));
    const int pinion = dynamic(-20.0f, 4.2f, 0.0f, int(sc.shapes.size()) - 1);
    JointSpec pj;
    pj.kind = kRevolute;
    pj.a = carriage;
    pj.b = pinion;
    pj.lax = 0.0f;
    pj.lay = 1.2f;
    sc.joints.push_back(pj);
    const int pinionJoint = int(sc.joints.size()) - 1;
    gear(pinionJoint, railJoint, rng.uniform(0.5, 1.5));
    for (int k = 0; k < falling; ++k) {
        const float x = rng.uniform(-25.0, 25.0);
        const float y = rng.uniform(12.0, 40.0);
        const float friction = rng.uniform(0.2, 0.8);
        if (rng.below(2) == 0) {
            const float hx = rng.uniform(0.3, 0.6);
            const float hy = rng.uniform(0.3, 0.6);
            sc.shapes.push_back(Box(hx, hy, 1.0f, friction, 0.0f));
        } else {
            sc.shapes.push_back(Disk(rng.uniform(0.3, 0.6), 1.0f, friction, 0.0f));
        }
        dynamic(x, y, 0.0f, int(sc.shapes.size()) - 1);
    }
    return sc;
}

// Named scenes with integer parameters, as used by tests and benchmarks.
//   "hello", "pyramid"(rows), "drop"(n, half, height, seed),
//   "pile"(seed, w, nmin, nmax), "billiards"(seed, w),
//   "chain"(seed, w, links, falling, torque), "joints"(seed, w, n, falling),
//   "sliders"(seed, w, n, falling), "targets"(seed, w, n, falling),
//   "terrain"(seed, w, falling), "gears"(seed, w, falling)
inline bool Make(const std::string& name, const int* p, int np, SceneSpec& out) {
    auto arg = [&](int i, int dflt) { return (i < np) ? p[i] : dflt; };
    if (name == "hello") {
        out = Hello();
    } else if (name == "pyramid") {
        out = Pyramid(arg(0, 20));
    } else if (name == "drop") {
        out = Drop(arg(0, 20000), float(arg(1, 200)), float(arg(2, 400)), std::uint32_t(arg(3, 1)));
    } else if (name == "pile") {
        out = Pile(std::uint32_t(arg(0, 1)), arg(1, 0), arg(2, 8), arg(3, 512));
    } else if (name == "billiards") {
        out = Billiards(std::uint32_t(arg(0, 1)), arg(1, 0));
    } else if (name == "chain") {
        out = Chain(std::uint32_t(arg(0, 1)), arg(1, 0), arg(2, 32), arg(3, 64), float(arg(4, 2)));
    } else if (name == "joints") {
        out = Joints(std::uint32_t(arg(0, 1)), arg(1, 0), arg(2, 8), arg(3, 24));
    } else if (name == "sliders") {
        out = Sliders(std::uint32_t(arg(0, 1)), arg(1, 0), arg(2, 6), arg(3, 24));
    } else if (name == "targets") {
        out = Targets(std::uint32_t(arg(0, 1)), arg(1, 0), arg(2, 8), arg(3, 24));
    } else if (name == "terrain") {
        out = Terrain(std::uint32_t(arg(0, 1)), arg(1, 0), arg(2, 40));
    } else if (name == "gears") {
        out = Gears(std::uint32_t(arg(0, 1)), arg(1, 0), arg(2, 24));
    } else {
        return false;
    }
    return true;
}

}  // namespace prx_scene
