// ORACLE — test infrastructure only (see cpu_ref.hpp).  C-ABI over the CPU
// restatement so the Python tests and the CPU-baseline leg of bench.py can run
// it through ctypes.  The world description and the state records use the same
// little-endian layout as oracle/ref_driver.cpp writes, so the reference's own
// description can be fed in and the outputs digested identically.
// tangent.hpp first: its std:: overloads must be visible where the
// restatement's templates name std::sqrt and the like.
#include "tangent.hpp"
#include "cpu_ref.hpp"

#include <chrono>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <thread>

namespace {

struct In {
    const unsigned char* p;
    size_t n, o = 0;
    std::uint32_t u32() {
        std::uint32_t v;
        std::memcpy(&v, p + o, 4);
        o += 4;
        return v;
    }
    float f32() {
        float v;
        std::memcpy(&v, p + o, 4);
        o += 4;
        return v;
    }
};

struct Out {
    std::vector<unsigned char> buf;
    void u32(std::uint32_t v) { put(&v, 4); }
    void f32(float v) { put(&v, 4); }
    void put(const void* q, size_t k) {
        const auto* c = static_cast<const unsigned char*>(q);
        buf.insert(buf.end(), c, c + k);
    }
};

template <class R>
struct Handle {
    oracle::World<R> world;
    oracle::StepConf<R> conf{};
    oracle::StepStats last{};
    std::uint32_t steps = 0;
};

// Reads the description (shapes, bodies, joints) and the StepConf record.
template <class R>
void Load(Handle<R>& h, In& in) {
    using namespace oracle;
    const std::uint32_t ns = in.u32();
    for (std::uint32_t i = 0; i < ns; ++i) {
        Shape<R> s;
        const std::uint32_t nv = in.u32();
        s.radius = R(in.f32());
        s.density = R(in.f32());
        s.friction = R(in.f32());
        s.restitution = R(in.f32());
        s.sensor = in.u32();
        s.catmask = in.u32();
        s.group = std::int32_t(in.u32());
        for (std::uint32_t k = 0; k < nv; ++k) {
            const R x = R(in.f32());
            const R y = R(in.f32());
            s.v.push_back(V2<R>{x, y});
        }
        for (std::uint32_t k = 0; k < nv; ++k) {
            const R x = R(in.f32());
            const R y = R(in.f32());
            s.nrm.push_back(V2<R>{x, y});
        }
        h.world.shapes.push_back(std::move(s));
    }
    const std::uint32_t nb = in.u32();
    for (std::uint32_t i = 0; i < nb; ++i) {
        Body<R> b;
        b.type = in.u32();
        b.flags = in.u32();
        R f[19];
        for (auto& x : f) x = R(in.f32());
        b.pos0 = Pos<R>{V2<R>{f[0], f[1]}, f[2]};
        b.pos1 = Pos<R>{V2<R>{f[3], f[4]}, f[5]};
        b.lc = V2<R>{f[6], f[7]};
        b.vel = Vel<R>{V2<R>{f[8], f[9]}, f[10]};
        b.lacc = V2<R>{f[11], f[12]};
        b.aacc = f[13];
        b.ldamp = f[14];
        b.adamp = f[15];
        b.invMass = f[16];
        b.invRotI = f[17];
        b.underActive = f[18];
        const std::uint32_t nsh = in.u32();
        for (std::uint32_t k = 0; k < nsh; ++k) b.shapes.push_back(in.u32());
        h.world.AddBody(b);
    }
    const std::uint32_t nj = in.u32();
    for (std::uint32_t i = 0; i < nj; ++i) {
        Joint<R> j;
        j.bodyA = in.u32();
        j.bodyB = in.u32();
        const std::uint32_t fl = in.u32();
        j.enableLimit = (fl & 1u) != 0;
        j.enableMotor = (fl & 2u) != 0;
        j.collideConnected = (fl & 4u) != 0;
        R f[12];
        for (auto& x : f) x = R(in.f32());
        j.laA = V2<R>{f[0], f[1]};
        j.laB = V2<R>{f[2], f[3]};
        j.referenceAngle = f[4];
        j.lower = f[5];
        j.upper = f[6];
        j.motorSpeed = f[7];
        j.maxMotorTorque = f[8];
        j.kind = (fl >> 8) & 0xffu;
        if (j.kind == 10) {
            // Gear: per side (the first joint's A-C in f[0..5], the second's
            // B-D in f[6..11]) the prismatic data (local anchors A, B, local
            // x axis; bit 16 + side of the flags) or the revolute reference
            // angle (first float), then bodies C, D, the ratio and the
            // constant.
            for (int k = 0; k < 2; ++k) {
                const R* g = f + 6 * k;
                j.gearPris[k] = (fl >> (16 + k)) & 1u;
                j.gearLaA[k] = V2<R>{g[0], g[1]};
                j.gearLaB[k] = V2<R>{g[2], g[3]};
                j.gearAxis[k] = V2<R>{g[4], g[5]};
                j.gearRef[k] = g[0];
            }
            j.bodyC = in.u32();
            j.bodyD = in.u32();
            j.ratio = R(in.f32());
            j.constant = R(in.f32());
            j.enableLimit = j.enableMotor = false;
            j.laA = j.laB = V2<R>{R(0), R(0)};
            j.referenceAngle = j.lower = j.upper = j.motorSpeed = j.maxMotorTorque = 0;
        } else if (j.kind == 8) {
            // Prismatic: reference angle, lower and upper translation, motor
            // speed, maximum motor force (in maxMotorTorque), local x axis of A.
            j.localXAxis = V2<R>{f[9], f[10]};
        } else if (j.kind == 9) {
            // Target: no body A; anchor on B, target, maximum force,
            // frequency, damping ratio.
            j.enableLimit = j.enableMotor = false;
            j.groundA = V2<R>{f[4], f[5]};
            j.maxForce = f[6];
            j.frequency = f[7];
            j.dampingRatio = f[8];
            j.referenceAngle = j.lower = j.upper = j.motorSpeed = j.maxMotorTorque = 0;
        } else if (j.kind != 0) {
            // Distance: length, frequency, damping ratio.  Weld: reference
            // angle, frequency, damping ratio.  Rope: maximum length.
            // Friction: maximum force, maximum torque.  Motor: linear offset,
            // -, angular offset, maximum force, maximum torque, correction
            // factor.
            j.enableLimit = false;
            if (j.kind != 6) j.enableMotor = false;
            if (j.kind == 1 || j.kind == 3) j.length = f[4];
            if (j.kind != 2 && j.kind != 5) j.referenceAngle = 0;
            if (j.kind == 7) {
                // Pulley: ground anchors, ratio, lengths.
                j.groundA = V2<R>{f[4], f[5]};
                j.groundB = V2<R>{f[6], f[7]};
                j.ratio = f[8];
                j.length = f[9];
                j.lengthB = f[10];
                j.frequency = j.dampingRatio = 0;
            } else if (j.kind == 6) {
                // Wheel: frequency, damping ratio, -, motor speed, maximum
                // motor torque, local x axis of A.
                j.frequency = f[4];
                j.dampingRatio = f[5];
                j.localXAxis = V2<R>{f[9], f[10]};
            } else if (j.kind <= 2) {
                j.frequency = f[5];
                j.dampingRatio = f[6];
            } else if (j.kind == 4) {
                j.maxForce = f[4];
                j.maxTorque = f[5];
            } else if (j.kind == 5) {
                // Motor: linear offset in laA, angular offset in referenceAngle.
                j.laB = V2<R>{R(0), R(0)};
                j.maxForce = f[5];
                j.maxTorque = f[6];
                j.correctionFactor = f[7];
            }
            j.lower = j.upper = 0;
            if (j.kind != 6) j.motorSpeed = j.maxMotorTorque = 0;
        }
        h.world.AddJoint(j);
    }
    const std::uint32_t nf = in.u32();
    R cf[32] = {};
    for (std::uint32_t k = 0; k < nf && k < 32; ++k) cf[k] = R(in.f32());
    auto& c = h.conf;
    c.dt = cf[0];
    c.dtRatio = cf[1];
    c.minStillTimeToSleep = cf[2];
    c.linearSlop = cf[3];
    c.angularSlop = cf[4];
    c.regResolutionRate = cf[5];
    c.regMinSeparation = cf[6];
    c.regMinMomentum = cf[7];
    c.velocityThreshold = cf[8];
    c.maxTranslation = cf[9];
    c.maxRotation = cf[10];
    c.maxLinearCorrection = cf[11];
    c.maxAngularCorrection = cf[12];
    c.linearSleepTolerance = cf[13];
    c.angularSleepTolerance = cf[14];
    c.displaceMultiplier = cf[15];
    c.aabbExtension = cf[16];
    c.maxCirclesRatio = cf[17];
    c.velIters = in.u32();
    c.posIters = in.u32();
    const std::uint32_t fl = in.u32();
    c.warmStart = (fl & 1u) != 0;
    c.toi = (fl & 2u) != 0;
    c.blockSolve = (fl & 4u) != 0;
    const std::uint32_t nt = in.u32();
    R tf[8] = {};
    for (std::uint32_t k = 0; k < nt; ++k) {
        const R v = R(in.f32());
        if (k < 8) tf[k] = v;
    }
    c.toiResolutionRate = tf[0];
    c.toiMinSeparation = tf[1];
    c.toiMinMomentum = tf[2];
    c.targetDepth = tf[3];
    c.tolerance = tf[4];
    c.toiVelIters = in.u32();
    c.toiPosIters = in.u32();
    c.maxToiRootIters = in.u32();
    c.maxToiIters = in.u32();
    c.maxDistanceIters = in.u32();
    c.maxSubSteps = in.u32();
}

template <class R>
void WriteState(const Handle<R>& h, Out& o) {
    using namespace oracle;
    const auto& w = h.world;
    o.u32(h.steps);
    const auto& s = h.last;
    // PreStepStats then RegStepStats, each in declaration order (StepStats.hpp:38-46, 69-82).
    auto bits = [](double v) {
        const float f = float(v);
        std::uint32_t u;
        std::memcpy(&u, &f, 4);
        return u;
    };
    const std::uint32_t st[18] = {s.preProxiesCreated, s.preProxiesMoved, s.preDestroyed, s.preAdded,
                                  s.preUpdated,        s.preSkipped,      bits(s.minSeparation),
                                  bits(s.maxIncImpulse), s.islandsFound,  s.islandsSolved,
                                  s.bodiesSlept,       s.maxIslandBodies, s.regAdded,
                                  s.regUpdated,        s.regSkipped,      s.regProxiesMoved,
                                  s.sumPosIters,       s.sumVelIters};
    o.put(st, sizeof(st));
    // ToiStepStats in declaration order (StepStats.hpp:117-151).
    const std::uint32_t ts[16] = {bits(s.toiMinSeparation), bits(s.toiMaxIncImpulse), s.toiIslandsFound,
                                  s.toiIslandsSolved,       s.toiContactsFound,       s.toiAtMaxSubSteps,
                                  s.toiUpdatedToi,          s.toiUpdatedTouching,     s.toiSkippedTouching,
                                  s.toiAdded,               s.toiProxiesMoved,        s.toiSumPosIters,
                                  s.toiSumVelIters,         s.toiMaxDistIters,        s.toiMaxToiIters,
                                  s.toiMaxRootIters};
    o.put(ts, sizeof(ts));
    o.u32(std::uint32_t(w.bodies.size()));
    for (const auto& b : w.bodies) {
        if (b.destroyed) {
            // A destroyed body: zeros and flags 0x80.
            const float z[10] = {};
            o.put(z, sizeof(z));
            o.u32(0x80u);
            continue;
        }
        const float v[10] = {float(b.pos0.lin.x), float(b.pos0.lin.y), float(b.pos0.ang),
                             float(b.pos1.lin.x), float(b.pos1.lin.y), float(b.pos1.ang),
                             float(b.vel.lin.x),  float(b.vel.lin.y),  float(b.vel.ang),
                             float(b.underActive)};
        o.put(v, sizeof(v));
        o.u32(b.flags);
    }
    o.u32(std::uint32_t(w.contacts.size()));
    for (const auto& kc : w.contacts) {
        const auto& c = w.contactBuf[kc.id];
        const auto& m = w.manifoldBuf[kc.id];
        const std::uint32_t u[4] = {kc.kmin, kc.kmax, kc.id, c.flags};
        o.put(u, sizeof(u));
        const std::uint32_t t[3] = {m.type | (m.n << 8), m.n > 0 ? m.pts[0].cf : kDefaultCf,
                                    m.n > 1 ? m.pts[1].cf : kDefaultCf};
        o.put(t, sizeof(t));
        float f[14] = {float(c.friction), float(c.restitution)};
        if (m.type == kFaceA || m.type == kFaceB) {
            f[2] = float(m.ln.x);
            f[3] = float(m.ln.y);
        }
        if (m.type != kUnset) {
            f[4] = float(m.lpt.x);
            f[5] = float(m.lpt.y);
        }
        for (std::uint32_t p = 0; p < m.n; ++p) {
            f[6 + 4 * p + 0] = float(m.pts[p].lp.x);
            f[6 + 4 * p + 1] = float(m.pts[p].lp.y);
            f[6 + 4 * p + 2] = float(m.pts[p].ni);
            f[6 + 4 * p + 3] = float(m.pts[p].ti);
        }
        o.put(f, sizeof(f));
    }
    std::uint32_t np = 0;
    for (const auto& lf : w.tree) np += lf.used ? 1u : 0u;
    o.u32(np);
    for (std::uint32_t id = 0; id < w.tree.size(); ++id) {
        const auto& lf = w.tree[id];
        if (!lf.used) continue;
        o.u32(id);
        const float a[4] = {float(lf.lo[0]), float(lf.lo[1]), float(lf.hi[0]), float(lf.hi[1])};
        o.put(a, sizeof(a));
    }
    o.u32(std::uint32_t(w.joints.size()));
    for (const auto& j : w.joints) {
        if (j.destroyed) {
            // A destroyed joint: limit state 0xffffffff and zeros.
            o.u32(0xffffffffu);
            const float z[4] = {};
            o.put(z, sizeof(z));
            continue;
        }
        o.u32(j.limitState);
        const float f[4] = {float(j.impulse.x), float(j.impulse.y), float(j.impulse.z), float(j.motorImpulse)};
        o.put(f, sizeof(f));
    }
    o.u32(w.nevents);
    if (!w.events.empty()) o.put(w.events.data(), w.events.size() * 4);
}

// precision 32: float, 64: double; 33 and 65: the same with forward-mode
// derivatives in double (oracle::Tf, oracle::Td); 34: float with derivatives
// in float (oracle::Tff).
constexpr int kTangent = 33, kTangent64 = 65, kTangentF = 34;

bool IsTangent(int precision) {
    return precision == kTangent || precision == kTangent64 || precision == kTangentF;
}

struct AnyHandle {
    int precision = 32;
    void* h = nullptr;
};

template <class F>
void Visit(AnyHandle* a, F&& f) {
    if (a->precision == 64) f(*static_cast<Handle<double>*>(a->h));
    else if (a->precision == kTangent) f(*static_cast<Handle<oracle::Tf>*>(a->h));
    else if (a->precision == kTangent64) f(*static_cast<Handle<oracle::Td>*>(a->h));
    else if (a->precision == kTangentF) f(*static_cast<Handle<oracle::Tff>*>(a->h));
    else f(*static_cast<Handle<float>*>(a->h));
}

template <class R>
void DoStep(Handle<R>& h) {
    h.last = h.world.Step(h.conf);
    ++h.steps;
}

unsigned char* Copy(const std::vector<unsigned char>& v, size_t* n) {
    auto* p = static_cast<unsigned char*>(std::malloc(v.size() ? v.size() : 1));
    if (!v.empty()) std::memcpy(p, v.data(), v.size());
    *n = v.size();
    return p;
}

// Widening of a float world's whole state to the double instantiation, for
// single-step comparisons from the same input state: every real converts
// exactly; each body's cached transformation is recomputed in double from its
// position (Body::SetTransformation with UnitVec::Get(angle), Body.cpp), as a
// double build holding these positions would have it.
template <class D, class S> oracle::V2<D> Wd(const oracle::V2<S>& v) { return {D(v.x), D(v.y)}; }
template <class D, class S> oracle::V3<D> Wd(const oracle::V3<S>& v) { return {D(v.x), D(v.y), D(v.z)}; }
template <class D, class S> oracle::Pos<D> Wd(const oracle::Pos<S>& v) { return {Wd<D>(v.lin), D(v.ang)}; }
template <class D, class S> oracle::Vel<D> Wd(const oracle::Vel<S>& v) { return {Wd<D>(v.lin), D(v.ang)}; }

template <class D, class S>
void Widen(const Handle<S>& s, Handle<D>& d) {
    using namespace oracle;
    const World<S>& a = s.world;
    World<D>& b = d.world;
    b.joints.clear();
    for (const auto& j : a.joints) {
        Joint<D> k;
        k.bodyA = j.bodyA;
        k.bodyB = j.bodyB;
        k.enableLimit = j.enableLimit;
        k.enableMotor = j.enableMotor;
        k.collideConnected = j.collideConnected;
        k.laA = Wd<D>(j.laA);
        k.laB = Wd<D>(j.laB);
        k.referenceAngle = D(j.referenceAngle);
        k.lower = D(j.lower);
        k.upper = D(j.upper);
        k.motorSpeed = D(j.motorSpeed);
        k.maxMotorTorque = D(j.maxMotorTorque);
        k.impulse = Wd<D>(j.impulse);
        k.motorImpulse = D(j.motorImpulse);
        k.limitState = j.limitState;
        k.rA = Wd<D>(j.rA);
        k.rB = Wd<D>(j.rB);
        k.mass = M33<D>{Wd<D>(j.mass.ex), Wd<D>(j.mass.ey), Wd<D>(j.mass.ez)};
        k.angularMass = D(j.angularMass);
        k.kind = j.kind;
        k.length = D(j.length);
        k.frequency = D(j.frequency);
        k.dampingRatio = D(j.dampingRatio);
        k.u = Wd<D>(j.u);
        k.dmass = D(j.dmass);
        k.invGamma = D(j.invGamma);
        k.bias = D(j.bias);
        k.gamma = D(j.gamma);
        k.curLength = D(j.curLength);
        k.maxForce = D(j.maxForce);
        k.maxTorque = D(j.maxTorque);
        k.lm00 = D(j.lm00);
        k.lm01 = D(j.lm01);
        k.lm10 = D(j.lm10);
        k.lm11 = D(j.lm11);
        k.correctionFactor = D(j.correctionFactor);
        k.angularError = D(j.angularError);
        k.linearError = Wd<D>(j.linearError);
        k.localXAxis = Wd<D>(j.localXAxis);
        k.ax = Wd<D>(j.ax);
        k.ay = Wd<D>(j.ay);
        k.sAx = D(j.sAx);
        k.sAy = D(j.sAy);
        k.sBx = D(j.sBx);
        k.sBy = D(j.sBy);
        k.springMass = D(j.springMass);
        k.groundA = Wd<D>(j.groundA);
        k.groundB = Wd<D>(j.groundB);
        k.lengthB = D(j.lengthB);
        k.ratio = D(j.ratio);
        k.s1 = D(j.s1);
        k.s2 = D(j.s2);
        k.a1 = D(j.a1);
        k.a2 = D(j.a2);
        k.motorMass = D(j.motorMass);
        k.bodyC = j.bodyC;
        k.bodyD = j.bodyD;
        for (int e = 0; e < 2; ++e) {
            k.gearPris[e] = j.gearPris[e];
            k.gearLaA[e] = Wd<D>(j.gearLaA[e]);
            k.gearLaB[e] = Wd<D>(j.gearLaB[e]);
            k.gearAxis[e] = Wd<D>(j.gearAxis[e]);
            k.gearRef[e] = D(j.gearRef[e]);
        }
        k.constant = D(j.constant);
        k.JvAC = Wd<D>(j.JvAC);
        k.JvBD = Wd<D>(j.JvBD);
        k.JwA = D(j.JwA);
        k.JwB = D(j.JwB);
        k.JwC = D(j.JwC);
        k.JwD = D(j.JwD);
        k.destroyed = j.destroyed;
        b.joints.push_back(k);
    }
    b.bodyJoints = a.bodyJoints;
    b.islandedJoints = a.islandedJoints;
    b.shapes.clear();
    for (const auto& sh : a.shapes) {
        Shape<D> t;
        for (const auto& v : sh.v) t.v.push_back(Wd<D>(v));
        for (const auto& v : sh.nrm) t.nrm.push_back(Wd<D>(v));
        t.radius = D(sh.radius);
        t.density = D(sh.density);
        t.friction = D(sh.friction);
        t.restitution = D(sh.restitution);
        t.sensor = sh.sensor;
        t.catmask = sh.catmask;
        t.group = sh.group;
        b.shapes.push_back(t);
    }
    b.bodies.clear();
    for (const auto& bo : a.bodies) {
        Body<D> t;
        t.type = bo.type;
        t.flags = bo.flags;
        t.pos0 = Wd<D>(bo.pos0);
        t.pos1 = Wd<D>(bo.pos1);
        t.lc = Wd<D>(bo.lc);
        t.vel = Wd<D>(bo.vel);
        t.lacc = Wd<D>(bo.lacc);
        t.aacc = D(bo.aacc);
        t.ldamp = D(bo.ldamp);
        t.adamp = D(bo.adamp);
        t.invMass = D(bo.invMass);
        t.invRotI = D(bo.invRotI);
        t.underActive = D(bo.underActive);
        t.alpha0 = D(bo.alpha0);
        t.xf = GetXf(t.pos1, t.lc);
        t.shapes = bo.shapes;
        t.proxies = bo.proxies;
        t.destroyed = bo.destroyed;
        b.bodies.push_back(t);
    }
    b.bodyOrder = a.bodyOrder;
    b.tree.clear();
    for (const auto& lf : a.tree) {
        Leaf<D> t;
        t.used = lf.used;
        t.body = lf.body;
        t.shape = lf.shape;
        t.child = lf.child;
        for (int i = 0; i < 2; ++i) {
            t.lo[i] = D(lf.lo[i]);
            t.hi[i] = D(lf.hi[i]);
        }
        b.tree.push_back(t);
    }
    b.leafCount = a.leafCount;
    b.leafFree = a.leafFree;
    b.fixturesForProxies = a.fixturesForProxies;
    b.bodiesForSync = a.bodiesForSync;
    b.proxiesForContacts = a.proxiesForContacts;
    b.contactBuf.clear();
    for (const auto& c : a.contactBuf) {
        Contact<D> t;
        t.kmin = c.kmin;
        t.kmax = c.kmax;
        t.bodyA = c.bodyA;
        t.bodyB = c.bodyB;
        t.shapeA = c.shapeA;
        t.shapeB = c.shapeB;
        t.flags = c.flags;
        t.friction = D(c.friction);
        t.restitution = D(c.restitution);
        t.tangentSpeed = D(c.tangentSpeed);
        t.toi = D(c.toi);
        t.toiValid = c.toiValid;
        t.toiCount = c.toiCount;
        b.contactBuf.push_back(t);
    }
    b.manifoldBuf.clear();
    for (const auto& m : a.manifoldBuf) {
        Manifold<D> t;
        t.type = m.type;
        t.n = m.n;
        t.ln = Wd<D>(m.ln);
        t.lpt = Wd<D>(m.lpt);
        for (int i = 0; i < 2; ++i) {
            t.pts[i].lp = Wd<D>(m.pts[i].lp);
            t.pts[i].cf = m.pts[i].cf;
            t.pts[i].ni = D(m.pts[i].ni);
            t.pts[i].ti = D(m.pts[i].ti);
        }
        b.manifoldBuf.push_back(t);
    }
    b.contactFree = a.contactFree;
    b.contacts.clear();
    for (const auto& k : a.contacts) b.contacts.push_back({k.kmin, k.kmax, k.id});
    b.bodyContacts.clear();
    for (const auto& lst : a.bodyContacts) {
        std::vector<typename World<D>::KeyedId> t;
        for (const auto& k : lst) t.push_back({k.kmin, k.kmax, k.id});
        b.bodyContacts.push_back(t);
    }
    b.islandedBodies = a.islandedBodies;
    b.islandedContacts = a.islandedContacts;
    b.events.clear();
    b.nevents = 0;
    const StepConf<S>& c = s.conf;
    StepConf<D>& e = d.conf;
    e.dt = D(c.dt);
    e.dtRatio = D(c.dtRatio);
    e.minStillTimeToSleep = D(c.minStillTimeToSleep);
    e.linearSlop = D(c.linearSlop);
    e.angularSlop = D(c.angularSlop);
    e.regResolutionRate = D(c.regResolutionRate);
    e.regMinSeparation = D(c.regMinSeparation);
    e.regMinMomentum = D(c.regMinMomentum);
    e.velocityThreshold = D(c.velocityThreshold);
    e.maxTranslation = D(c.maxTranslation);
    e.maxRotation = D(c.maxRotation);
    e.maxLinearCorrection = D(c.maxLinearCorrection);
    e.maxAngularCorrection = D(c.maxAngularCorrection);
    e.linearSleepTolerance = D(c.linearSleepTolerance);
    e.angularSleepTolerance = D(c.angularSleepTolerance);
    e.displaceMultiplier = D(c.displaceMultiplier);
    e.aabbExtension = D(c.aabbExtension);
    e.maxCirclesRatio = D(c.maxCirclesRatio);
    e.velIters = c.velIters;
    e.posIters = c.posIters;
    e.warmStart = c.warmStart;
    e.toi = c.toi;
    e.blockSolve = c.blockSolve;
    e.toiResolutionRate = D(c.toiResolutionRate);
    e.toiMinSeparation = D(c.toiMinSeparation);
    e.toiMinMomentum = D(c.toiMinMomentum);
    e.targetDepth = D(c.targetDepth);
    e.tolerance = D(c.tolerance);
    e.toiVelIters = c.toiVelIters;
    e.toiPosIters = c.toiPosIters;
    e.maxToiRootIters = c.maxToiRootIters;
    e.maxToiIters = c.maxToiIters;
    e.maxDistanceIters = c.maxDistanceIters;
    e.maxSubSteps = c.maxSubSteps;
    d.steps = s.steps;
}

// Sets the derivative of one parameter of a dual-number world (the kinds of
// oracle_tangent_seed).
template <class W>
int SeedTangent(W& w, int kind, std::uint32_t index, int comp, double value) {
    switch (kind) {
        case 0:
            if (index >= w.bodies.size() || comp < 0 || comp > 2) return -1;
            (comp == 0 ? w.bodies[index].vel.lin.x : comp == 1 ? w.bodies[index].vel.lin.y : w.bodies[index].vel.ang)
                .d = decltype(w.bodies[index].vel.ang.d)(value);
            return 0;
        case 1:
            if (index >= w.shapes.size()) return -1;
            w.shapes[index].friction.d = decltype(w.shapes[index].friction.d)(value);
            return 0;
        case 2:
            if (index >= w.shapes.size()) return -1;
            w.shapes[index].restitution.d = decltype(w.shapes[index].restitution.d)(value);
            return 0;
        case 3:
            if (index >= w.joints.size()) return -1;
            w.joints[index].maxMotorTorque.d = decltype(w.joints[index].maxMotorTorque.d)(value);
            return 0;
        default:
            return -1;
    }
}

}  // namespace

// A world edit between steps: kind 'L' SetLocation(body, (a, b)), 'A'
// SetAngle(body, a), 'T' SetTransformation(body, {(a, b), UnitVec::Get(c)}),
// 'O' ShiftOrigin((a, b)).
template <class R>
void Edit(Handle<R>& h, int kind, std::uint32_t body, float a, float b, float c) {
    using namespace oracle;
    auto& w = h.world;
    switch (kind) {
    case 'L': w.SetLocation(body, V2<R>{R(a), R(b)}); break;
    case 'A': w.SetAngle(body, R(a)); break;
    case 'T': w.SetTransformation(body, V2<R>{R(a), R(b)}, UnitFromAngle(R(c))); break;
    case 'O': w.ShiftOrigin(V2<R>{R(a), R(b)}); break;
    case 'V': w.SetVelocity(body, V2<R>{R(a), R(b)}, R(c)); break;
    case 'J': w.DestroyJoint(body); break;
    case 'D': w.DestroyBody(body); break;
    case 'S': w.SetType(body, std::uint32_t(a)); break;
    default: break;
    }
}

extern "C" {

// Creates an oracle world from a description + conf record (layout of
// ref_driver.cpp's WriteDescription/WriteConf).  precision: 32, 64 or 33
// (float with forward-mode derivatives, seeded by oracle_tangent_seed).
void* oracle_create(const unsigned char* desc, size_t len, int precision) {
    In in{desc, len};
    auto* a = new AnyHandle;
    a->precision = precision;
    if (precision == 64) {
        auto* h = new Handle<double>;
        Load(*h, in);
        a->h = h;
    } else if (precision == kTangent) {
        auto* h = new Handle<oracle::Tf>;
        Load(*h, in);
        a->h = h;
    } else if (precision == kTangent64) {
        auto* h = new Handle<oracle::Td>;
        Load(*h, in);
        a->h = h;
    } else if (precision == kTangentF) {
        auto* h = new Handle<oracle::Tff>;
        Load(*h, in);
        a->h = h;
    } else {
        a->precision = 32;
        auto* h = new Handle<float>;
        Load(*h, in);
        a->h = h;
    }
    return a;
}

// A double-precision copy of a float world's whole state (the FP64 oracle
// stepping from the float world's current state); null for a double world.
void* oracle_widen(void* p) {
    auto* a = static_cast<AnyHandle*>(p);
    if (!a || a->precision != 32) return nullptr;
    auto* h = new Handle<double>;
    Widen(*static_cast<Handle<float>*>(a->h), *h);
    auto* out = new AnyHandle;
    out->precision = 64;
    out->h = h;
    return out;
}

void oracle_destroy(void* p) {
    auto* a = static_cast<AnyHandle*>(p);
    if (!a) return;
    Visit(a, [](auto& h) { delete &h; });
    delete a;
}

int oracle_edit(void* p, int kind, std::uint32_t body, float a, float b, float c) {
    auto* an = static_cast<AnyHandle*>(p);
    Visit(an, [&](auto& h) { Edit(h, kind, body, a, b, c); });
    return 0;
}

// Steps n times; when `records` is non-null, appends one state record per step
// (and returns the malloc'd buffer, free with oracle_free).
unsigned char* oracle_step(void* p, int n, int record, size_t* outlen) {
    auto* a = static_cast<AnyHandle*>(p);
    Out o;
    for (int i = 0; i < n; ++i) {
        Visit(a, [&](auto& h) {
            DoStep(h);
            if (record) WriteState(h, o);
        });
    }
    return Copy(o.buf, outlen);
}

// Current state record (same layout as a per-step record).
unsigned char* oracle_state(void* p, size_t* outlen) {
    auto* a = static_cast<AnyHandle*>(p);
    Out o;
    Visit(a, [&](auto& h) { WriteState(h, o); });
    return Copy(o.buf, outlen);
}

// Full body positions/velocities in the oracle's own precision, as doubles:
// out[b*6 + {x, y, angle, vx, vy, w}].
void oracle_bodies_f64(void* p, double* out) {
    auto* a = static_cast<AnyHandle*>(p);
    auto fill = [&](const auto& w) {
        size_t k = 0;
        for (const auto& b : w.bodies) {
            out[k++] = double(b.pos1.lin.x);
            out[k++] = double(b.pos1.lin.y);
            out[k++] = double(b.pos1.ang);
            out[k++] = double(b.vel.lin.x);
            out[k++] = double(b.vel.lin.y);
            out[k++] = double(b.vel.ang);
        }
    };
    Visit(a, [&](auto& h) { fill(h.world); });
}

// Forward-mode seed of a world created with precision 33, 34 or 65: sets the derivative
// of one parameter (kind 0: body `index`'s velocity component `comp` (0 vx,
// 1 vy, 2 w); 1: shape `index`'s friction; 2: its restitution; 3: joint
// `index`'s maximum motor torque) to `value`.  Returns 0, or -1 for a bad
// handle or index.
int oracle_tangent_seed(void* p, int kind, std::uint32_t index, int comp, double value) {
    auto* a = static_cast<AnyHandle*>(p);
    if (!a || !IsTangent(a->precision)) return -1;
    int rc = -1;
    Visit(a, [&](auto& h) {
        if constexpr (!std::is_arithmetic_v<std::decay_t<decltype(h.world.bodies[0].vel.ang)>>)
            rc = SeedTangent(h.world, kind, index, comp, value);
    });
    return rc;
}

// Derivatives of the body state of a world created with precision 33, 34 or 65:
// out[b*6 + {x, y, angle, vx, vy, w}] (widened exactly).
void oracle_tangent_bodies(void* p, double* out) {
    auto* a = static_cast<AnyHandle*>(p);
    if (!a || !IsTangent(a->precision)) return;
    Visit(a, [&](auto& h) {
        if constexpr (!std::is_arithmetic_v<std::decay_t<decltype(h.world.bodies[0].vel.ang)>>) {
            size_t k = 0;
            for (const auto& b : h.world.bodies) {
                out[k++] = double(b.pos1.lin.x.d);
                out[k++] = double(b.pos1.lin.y.d);
                out[k++] = double(b.pos1.ang.d);
                out[k++] = double(b.vel.lin.x.d);
                out[k++] = double(b.vel.lin.y.d);
                out[k++] = double(b.vel.ang.d);
            }
        }
    });
}

void oracle_free(void* p) { std::free(p); }

// The C library's float cosine and sine (what UnitVec::Get(angle) calls on the
// host), for comparison with the device routine.
void oracle_sincosf(const float* in, float* out, int n) {
    for (int i = 0; i < n; ++i) {
        out[2 * i] = std::cos(in[i]);
        out[2 * i + 1] = std::sin(in[i]);
    }
}

// CPU-baseline helper: steps many independent worlds (descriptions packed
// back-to-back, each prefixed by its byte length) with `threads` std::threads,
// each owning a static contiguous partition balanced by body count.  Returns
// wall seconds for `steps` steps of every world; *body_steps receives
// sum(dynamic bodies) * steps.
double oracle_time_batch(const unsigned char* packed, size_t len, int nworlds, int steps, int threads,
                         double* body_steps) {
    std::vector<std::unique_ptr<Handle<float>>> ws;
    std::vector<double> weight;
    size_t off = 0;
    double dyn = 0;
    for (int i = 0; i < nworlds && off + 8 <= len; ++i) {
        std::uint64_t n;
        std::memcpy(&n, packed + off, 8);
        off += 8;
        In in{packed + off, size_t(n)};
        auto h = std::make_unique<Handle<float>>();
        Load(*h, in);
        off += size_t(n);
        double nb = 0;
        for (const auto& b : h->world.bodies) nb += (b.flags & oracle::kAccelerable) ? 1.0 : 0.0;
        dyn += nb;
        weight.push_back(nb);
        ws.push_back(std::move(h));
    }
    if (threads < 1) threads = 1;
    // Static partition by cumulative body count.
    std::vector<size_t> cut(size_t(threads) + 1, ws.size());
    cut[0] = 0;
    double total = 0;
    for (double x : weight) total += x;
    double acc = 0;
    int t = 1;
    for (size_t i = 0; i < ws.size() && t < threads; ++i) {
        acc += weight[i];
        while (t < threads && acc >= total * t / threads) cut[size_t(t++)] = i + 1;
    }
    const auto t0 = std::chrono::steady_clock::now();
    std::vector<std::thread> pool;
    for (int k = 0; k < threads; ++k) {
        pool.emplace_back([&, k] {
            for (size_t i = cut[size_t(k)]; i < cut[size_t(k) + 1]; ++i)
                for (int s = 0; s < steps; ++s) DoStep(*ws[i]);
        });
    }
    for (auto& th : pool) th.join();
    const auto t1 = std::chrono::steady_clock::now();
    *body_steps = dyn * steps;
    return std::chrono::duration<double>(t1 - t0).count();
}

}  // extern "C"
