/* playrho_amd — C ABI of the MI355X step engine for PlayRho worlds.
 *
 * The boundary is the one PlayRho's type-erased World forwards through:
 * every call here corresponds to a WorldConcept virtual or a free function of
 * the playrho::d2 API that the step path needs (Library/include/playrho/d2/
 * World.hpp:694-841, detail/WorldConcept.hpp, detail/WorldModel.hpp:144-147).
 * Plain pointers and sizes only; no C++ or torch types cross it.  Exceptions of
 * the reference (WrongState, LengthError, InvalidArgument, OutOfRange) map to
 * status codes; prx_last_error() gives the message of the last failure on the
 * calling thread.
 *
 * Stepping runs hand-written gfx950 HIP kernels on the current HIP device.  A
 * world's step state lives in device memory between calls; host getters
 * download it lazily.  Many worlds can be stepped together as a batch.
 */
#ifndef PLAYRHO_AMD_H
#define PLAYRHO_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PRX_ABI_VERSION 4

/* Status codes (reference exception types in parentheses). */
enum {
    PRX_OK = 0,
    PRX_E_INVALID_ARGUMENT = 1, /* InvalidArgument (AabbTreeWorld.cpp:1243-1248) */
    PRX_E_LENGTH = 2,           /* LengthError (AabbTreeWorld.cpp:1024-1026) */
    PRX_E_WRONG_STATE = 3,      /* WrongState (AabbTreeWorld.cpp:2149-2151) */
    PRX_E_OUT_OF_RANGE = 4,     /* OutOfRange (AabbTreeWorld.cpp:164-174) */
    PRX_E_DEVICE = 5,           /* HIP runtime failure */
    PRX_E_UNSUPPORTED = 6       /* a configuration this engine does not step */
};

typedef struct prx_world prx_world;
typedef struct prx_batch prx_batch;

/* Body types (BodyType.hpp). */
enum { PRX_STATIC = 0, PRX_KINEMATIC = 1, PRX_DYNAMIC = 2 };
/* Shape kinds: DiskShapeConf, PolygonShapeConf::Set (hull of points),
 * PolygonShapeConf::SetAsBox, EdgeShapeConf. */
enum { PRX_SHAPE_DISK = 0, PRX_SHAPE_POLYGON = 1, PRX_SHAPE_BOX = 2, PRX_SHAPE_EDGE = 3, PRX_SHAPE_CHAIN = 4, PRX_SHAPE_MULTI = 5 };

/* WorldConf (d2/WorldConf.hpp:40-90): vertex-radius bounds; capacity hint. */
typedef struct prx_world_conf {
    uint32_t abi_version;
    float min_vertex_radius; /* default 0.01 m */
    float max_vertex_radius; /* default 255 m */
    uint32_t contact_capacity_per_proxy; /* device contact slots per proxy (0 = 8) */
} prx_world_conf;

/* Union of the shape confs used on the step path
 * (d2/PolygonShapeConf.hpp, DiskShapeConf.hpp, EdgeShapeConf.hpp, ShapeConf.hpp). */
typedef struct prx_shape_conf {
    uint32_t kind;
    float vertex_radius; /* < 0: the conf's default (0.01 m; for a disk: its radius) */
    float radius;        /* disk radius */
    float hx, hy;        /* box half extents */
    float location[2];   /* disk center */
    const float* points; /* polygon points (x, y pairs) or edge ends */
    uint32_t npoints;
    float density, friction, restitution;
    uint32_t sensor;
    uint16_t category_bits, mask_bits; /* Filter.hpp defaults: 1, 0xffff */
    int16_t group_index;
} prx_shape_conf;

/* BodyConf (d2/BodyConf.hpp). */
typedef struct prx_body_conf {
    uint32_t type;
    float location[2];
    float angle;
    float linear_velocity[2];
    float angular_velocity;
    float linear_acceleration[2];
    float angular_acceleration;
    float linear_damping, angular_damping;
    uint32_t allow_sleep, awake, fixed_rotation, bullet, enabled;
    const uint32_t* shapes;
    uint32_t nshapes;
} prx_body_conf;

/* RevoluteJointConf (d2/RevoluteJointConf.hpp:27-123) with its JointConf base
 * (bodyA, bodyB, collideConnected; d2/JointConf.hpp). */
typedef struct prx_revolute_joint_conf {
    uint32_t body_a, body_b;
    float local_anchor_a[2], local_anchor_b[2];
    float reference_angle;
    uint32_t enable_limit;
    float lower_angle, upper_angle;
    uint32_t enable_motor;
    float motor_speed, max_motor_torque;
    uint32_t collide_connected;
} prx_revolute_joint_conf;

/* DistanceJointConf (d2/DistanceJointConf.hpp:62-130): rest length, and a
 * soft (spring) constraint when frequency > 0.  The reference's defaults are
 * length 1, frequency 0, damping ratio 0; a negative or NaN frequency is
 * PRX_E_INVALID_ARGUMENT (NonNegative<Frequency> there). */
typedef struct prx_distance_joint_conf {
    uint32_t body_a, body_b;
    float local_anchor_a[2], local_anchor_b[2];
    float length, frequency, damping_ratio;
    uint32_t collide_connected;
} prx_distance_joint_conf;

/* WeldJointConf (d2/WeldJointConf.hpp:62-130): rigid, or with a soft angular
 * part when frequency > 0.  The reference's defaults are reference angle 0,
 * frequency 0, damping ratio 0; frequency as for the distance joint. */
typedef struct prx_weld_joint_conf {
    uint32_t body_a, body_b;
    float local_anchor_a[2], local_anchor_b[2];
    float reference_angle, frequency, damping_ratio;
    uint32_t collide_connected;
} prx_weld_joint_conf;

/* RopeJointConf (d2/RopeJointConf.hpp:67-106): a maximum distance between the
 * anchors.  The reference's defaults are local anchors (-1, 0) and (1, 0),
 * maximum length 0. */
typedef struct prx_rope_joint_conf {
    uint32_t body_a, body_b;
    float local_anchor_a[2], local_anchor_b[2];
    float max_length;
    uint32_t collide_connected;
} prx_rope_joint_conf;

/* FrictionJointConf (d2/FrictionJointConf.hpp:58-117): top-down friction with
 * maximum force and torque (both default 0; negative or NaN is
 * PRX_E_INVALID_ARGUMENT, NonNegative there). */
typedef struct prx_friction_joint_conf {
    uint32_t body_a, body_b;
    float local_anchor_a[2], local_anchor_b[2];
    float max_force, max_torque;
    uint32_t collide_connected;
} prx_friction_joint_conf;

/* MotorJointConf (d2/MotorJointConf.hpp:58-138): drives body B towards the
 * linear and angular offsets from body A.  The reference's defaults are
 * offsets 0, maximum force 1, maximum torque 1, correction factor 0.3;
 * negative or NaN maxima are PRX_E_INVALID_ARGUMENT. */
typedef struct prx_motor_joint_conf {
    uint32_t body_a, body_b;
    float linear_offset[2];
    float angular_offset, max_force, max_torque, correction_factor;
    uint32_t collide_connected;
} prx_motor_joint_conf;

/* WheelJointConf (d2/WheelJointConf.hpp:71-170): point-to-line along the local
 * axis of A (a unit vector; the y axis is its reverse perpendicular), a soft
 * spring along it when frequency > 0, and an optional motor.  The
 * reference's defaults are axis (1, 0), frequency 2, damping ratio 0.7, motor
 * off; a negative or NaN frequency is PRX_E_INVALID_ARGUMENT. */
typedef struct prx_wheel_joint_conf {
    uint32_t body_a, body_b;
    float local_anchor_a[2], local_anchor_b[2];
    float local_axis_a[2];
    uint32_t enable_motor;
    float motor_speed, max_motor_torque;
    float frequency, damping_ratio;
    uint32_t collide_connected;
} prx_wheel_joint_conf;

/* PulleyJointConf (d2/PulleyJointConf.hpp:65-137): lengthA + ratio * lengthB
 * held at its initial value.  The reference's defaults are ground anchors
 * (-1, 1) and (1, 1), local anchors (-1, 0) and (1, 0), lengths 0, ratio 1,
 * collide_connected 1. */
typedef struct prx_pulley_joint_conf {
    uint32_t body_a, body_b;
    float ground_anchor_a[2], ground_anchor_b[2];
    float local_anchor_a[2], local_anchor_b[2];
    float length_a, length_b, ratio;
    uint32_t collide_connected;
} prx_pulley_joint_conf;

/* PrismaticJointConf (d2/PrismaticJointConf.hpp:61-217): body B slides along
 * the local x axis of A (a unit vector) with the relative rotation held at the
 * reference angle, with an optional translation limit [lower, upper] and an
 * optional motor (speed, maximum force).  The reference's defaults are axis
 * (1, 0), reference angle 0, limit and motor off. */
typedef struct prx_prismatic_joint_conf {
    uint32_t body_a, body_b;
    float local_anchor_a[2], local_anchor_b[2];
    float local_axis_a[2];
    float reference_angle;
    uint32_t enable_limit;
    float lower_translation, upper_translation;
    uint32_t enable_motor;
    float motor_speed, max_motor_force;
    uint32_t collide_connected;
} prx_prismatic_joint_conf;

/* TargetJointConf (d2/TargetJointConf.hpp:56-161): pulls a point of one body
 * (the local anchor) towards a world target with a soft spring (frequency,
 * damping ratio) whose impulse is limited by dt * max_force; the joint has no
 * body A.  The reference's defaults are frequency 5, damping ratio 0.7,
 * maximum force 0; negative or NaN values are PRX_E_INVALID_ARGUMENT. */
typedef struct prx_target_joint_conf {
    uint32_t body;
    float target[2], local_anchor[2];
    float max_force, frequency, damping_ratio;
    uint32_t collide_connected;
} prx_target_joint_conf;

/* GearJointConf (d2/GearJointConf.hpp:60-126): keeps coordinate1 + ratio *
 * coordinate2 equal to `constant`, the coordinates being those of two
 * revolute (angle) or prismatic (translation) joints.  Bodies A and B are the
 * two joints' second bodies, C and D their first bodies.  Per side (A-C from
 * the first joint in ac[], B-D from the second in bd[]): a prismatic side
 * (prismatic_ac / prismatic_bd set) holds the joint's local anchors A, B and
 * local x axis in [0..5]; a revolute one its reference angle in [0].  The
 * reference's defaults are ratio 1, constant 0. */
typedef struct prx_gear_joint_conf {
    uint32_t body_a, body_b, body_c, body_d;
    uint32_t prismatic_ac, prismatic_bd;
    float ac[6], bd[6];
    float ratio, constant;
    uint32_t collide_connected;
} prx_gear_joint_conf;

/* StepConf fields the regular step reads (StepConf.hpp:49-392). */
typedef struct prx_step_conf {
    float delta_time, dt_ratio, min_still_time_to_sleep, linear_slop, angular_slop;
    float reg_resolution_rate, reg_min_separation, reg_min_momentum, velocity_threshold;
    float max_translation, max_rotation, max_linear_correction, max_angular_correction;
    float linear_sleep_tolerance, angular_sleep_tolerance, displace_multiplier;
    float aabb_extension, max_circles_ratio;
    uint32_t reg_velocity_iters, reg_position_iters;
    uint32_t do_warm_start, do_toi, do_blocksolve;
    /* TOI phase (StepConf.hpp:84-158, 216-374). */
    float toi_resolution_rate, toi_min_separation, toi_min_momentum, target_depth, tolerance;
    uint32_t toi_velocity_iters, toi_position_iters, max_toi_root_iters, max_toi_iters;
    uint32_t max_distance_iters, max_sub_steps;
} prx_step_conf;

/* StepStats (StepStats.hpp:30-201): PreStepStats (:38-46), RegStepStats
 * (:69-82) and ToiStepStats (:117-151), fields in the reference's order.
 * The min separations are +infinity and the max incremental impulses 0 when
 * no island was solved in the phase; the toi_ fields stay so with doToi =
 * false. */
typedef struct prx_step_stats {
    uint32_t pre_proxies_created, pre_proxies_moved, pre_contacts_destroyed, pre_contacts_added;
    uint32_t pre_contacts_updated, pre_contacts_skipped;
    float reg_min_separation, reg_max_inc_impulse;
    uint32_t reg_islands_found, reg_islands_solved, reg_bodies_slept, reg_max_island_bodies;
    uint32_t reg_contacts_added, reg_contacts_updated, reg_contacts_skipped, reg_proxies_moved;
    uint32_t reg_sum_pos_iters, reg_sum_vel_iters;
    float toi_min_separation, toi_max_inc_impulse;
    uint32_t toi_islands_found, toi_islands_solved, toi_contacts_found, toi_contacts_at_max_sub_steps;
    uint32_t toi_contacts_updated_toi, toi_contacts_updated_touching, toi_contacts_skipped_touching;
    uint32_t toi_contacts_added, toi_proxies_moved, toi_sum_pos_iters, toi_sum_vel_iters;
    uint32_t toi_max_dist_iters, toi_max_toi_iters, toi_max_root_iters;
} prx_step_stats;

/* Launch schedule of the batched step (chosen by prx_batch_autotune). */
typedef struct prx_batch_conf {
    uint32_t threads_per_world; /* 64 (one wavefront) by default */
    uint32_t use_graph;         /* replay steps from a captured hipGraph */
    uint32_t order;             /* 0: largest worlds first, 1: as given */
    uint32_t coop_min_contacts; /* islands with at least this many constraints are
                                   swept by the whole wavefront (default 32) */
    uint32_t lds_bytes;         /* dynamic LDS per workgroup; 0: the minimum the
                                   batch needs.  Larger values leave fewer
                                   workgroups per CU (160 KB of LDS per CU), so
                                   the long worlds share their SIMD with fewer
                                   waves */
} prx_batch_conf;

int prx_abi_version(void);
const char* prx_last_error(void);

/* Reference StepConf defaults (values of a default-constructed StepConf in the
 * reference float build, with dt = 1/60 and 8/3 iterations). */
void prx_step_conf_default(prx_step_conf* conf);
void prx_world_conf_default(prx_world_conf* conf);
void prx_batch_conf_default(prx_batch_conf* conf);

/* World{WorldConf} / ~World. */
int prx_world_create(const prx_world_conf* conf, prx_world** out);
void prx_world_destroy(prx_world* world);

/* CreateShape(World&, conf) (World.hpp:512,522-526; AabbTreeWorld.cpp:1234-1257). */
int prx_shape_create(prx_world* world, const prx_shape_conf* conf, uint32_t* shape_id);
/* CreateShape(World&, ChainShapeConf / MultiShapeConf) (WorldShape.hpp;
 * ChainShapeConf.cpp:38-160, MultiShapeConf.cpp:35-62, ConvexHull.cpp:30-36):
 * conf->kind PRX_SHAPE_CHAIN (points: the chain's vertices; one child per
 * segment) or PRX_SHAPE_MULTI (points: the hulls' points one hull after
 * another, hull_sizes[nhulls] points each; one child per convex hull).  The
 * children become consecutive shape ids from *first_shape on (*child_count of
 * them), one proxy each, with the conf's material and filter; a body takes
 * the shape by listing all of them, and its mass data counts the shape once
 * (the chain's or multi shape's own mass formula). */
int prx_compound_shape_create(prx_world* world, const prx_shape_conf* conf, const uint32_t* hull_sizes,
                              uint32_t nhulls, uint32_t* first_shape, uint32_t* child_count);
/* CreateBody(World&, BodyConf, resetMassData = true) (WorldBody.cpp:55;
 * AabbTreeWorld.cpp:1019-1044; mass from WorldBody.cpp:360-392). */
int prx_body_create(prx_world* world, const prx_body_conf* conf, uint32_t* body_id);
/* CreateJoint(World&, const RevoluteJointConf&) (World.hpp:448-458;
 * AabbTreeWorld.cpp:1132-1175).  Joints must be created before the world's
 * first step that finds contacts between the two bodies (joints added to a
 * world with contacts would need contact re-filtering, which is not
 * modelled): PRX_E_UNSUPPORTED otherwise. */
int prx_joint_create(prx_world* world, const prx_revolute_joint_conf* conf, uint32_t* joint_id);
/* CreateJoint(World&, const DistanceJointConf&) and (…, const WeldJointConf&)
 * (World.hpp:448-458; solvers DistanceJointConf.cpp:81-245,
 * WeldJointConf.cpp:115-375), with the same restriction as prx_joint_create.
 * Joint ids are shared by all kinds, in creation order. */
int prx_distance_joint_create(prx_world* world, const prx_distance_joint_conf* conf, uint32_t* joint_id);
int prx_weld_joint_create(prx_world* world, const prx_weld_joint_conf* conf, uint32_t* joint_id);
/* CreateJoint(World&, const RopeJointConf&) and (…, const FrictionJointConf&)
 * (solvers RopeJointConf.cpp:59-212, FrictionJointConf.cpp:76-237). */
int prx_rope_joint_create(prx_world* world, const prx_rope_joint_conf* conf, uint32_t* joint_id);
int prx_friction_joint_create(prx_world* world, const prx_friction_joint_conf* conf, uint32_t* joint_id);
/* CreateJoint(World&, const MotorJointConf&) (solver MotorJointConf.cpp:79-234). */
int prx_motor_joint_create(prx_world* world, const prx_motor_joint_conf* conf, uint32_t* joint_id);
/* CreateJoint(World&, const WheelJointConf&) (solver WheelJointConf.cpp:105-359). */
int prx_wheel_joint_create(prx_world* world, const prx_wheel_joint_conf* conf, uint32_t* joint_id);
/* CreateJoint(World&, const PulleyJointConf&) (solver PulleyJointConf.cpp:92-253;
 * prx_world_shift_origin moves its ground anchors, PulleyJointConf.cpp:255-260). */
int prx_pulley_joint_create(prx_world* world, const prx_pulley_joint_conf* conf, uint32_t* joint_id);
/* CreateJoint(World&, const TargetJointConf&) (solver TargetJointConf.cpp:60-184). */
int prx_target_joint_create(prx_world* world, const prx_target_joint_conf* conf, uint32_t* joint_id);
/* CreateJoint(World&, const PrismaticJointConf&) (solver PrismaticJointConf.cpp:155-531). */
int prx_prismatic_joint_create(prx_world* world, const prx_prismatic_joint_conf* conf, uint32_t* joint_id);
/* GetGearJointConf(const World&, JointID, JointID, Real) (GearJointConf.cpp:89-169):
 * the gear conf for two existing joints at the bodies' present positions (the
 * world's state is brought back from the device first when it is stale);
 * PRX_E_INVALID_ARGUMENT unless both are revolute or prismatic. */
int prx_gear_joint_conf_get(prx_world* world, uint32_t joint1, uint32_t joint2, float ratio,
                            prx_gear_joint_conf* conf);
/* CreateJoint(World&, const GearJointConf&) (solver GearJointConf.cpp:171-471). */
int prx_gear_joint_create(prx_world* world, const prx_gear_joint_conf* conf, uint32_t* joint_id);
uint32_t prx_world_joint_count(const prx_world* world);
/* Builds a named workload scene through the calls above. */
int prx_scene_build(prx_world* world, const char* name, const int32_t* params, int nparams);

/* Replaces the world with a description in the engine's record layout
 * (shapes with hull vertices and normals, bodies with sweep, velocity, mass).
 * Used to start from a state produced elsewhere (e.g. the reference). */
int prx_world_load_desc(prx_world* world, const uint8_t* desc, size_t len);
/* Writes the current description (same layout) into buf. */
int prx_world_save_desc(prx_world* world, uint8_t* buf, size_t cap, size_t* len);

uint32_t prx_world_body_count(const prx_world* world);
uint32_t prx_world_shape_count(const prx_world* world);

/* Step(World&, const StepConf&) (World.hpp:917-920 -> AabbTreeWorld.cpp:2144-2219),
 * plain per-world form on the device.  doToi must be 0 (no TOI phase here). */
int prx_step(prx_world* world, const prx_step_conf* conf, prx_step_stats* stats);

/* Body getters: 10 floats per body (pos0 x,y,a; pos1 x,y,a; velocity x,y,w;
 * under-active time) and the body flags. */
int prx_world_get_bodies(prx_world* world, float* out, uint32_t* flags);
/* GetTransformation(World&, BodyID) (WorldBody.hpp): {px, py, cos, sin}. */
int prx_world_get_transformation(prx_world* world, uint32_t body, float out[4]);
/* SetVelocity(World&, BodyID, Velocity) (WorldBody.cpp:263-268; Body::SetVelocity,
 * Body.cpp:184-194): v = linear x, y, angular; a velocity other than zero is
 * ignored by a body that is not speedable and wakes the others. */
int prx_world_set_velocity(prx_world* world, uint32_t body, const float v[3]);
/* Destroy(World&, JointID) (World.hpp:469; AabbTreeWorld.cpp:1215-1227): both
 * bodies wake; the id is reused by the next joint created.  An id already
 * destroyed is left alone. */
int prx_world_destroy_joint(prx_world* world, uint32_t joint);
/* Destroy(World&, BodyID) (World.hpp:373; AabbTreeWorld.cpp:1062-1104): the
 * body's joints, then its contacts (touching ones end: with contact events on,
 * their end-contact calls come first in the next step's record; bodies they
 * held wake), then its proxies.  The body keeps its id and is no longer
 * stepped; edits and getters on it return PRX_E_WRONG_STATE, and the state
 * record writes it as zeros with flags 0x80.  Creating a body afterwards and
 * saving the world's description are PRX_E_UNSUPPORTED (the reference would
 * reuse the id at the end of its body list). */
int prx_world_destroy_body(prx_world* world, uint32_t body);
/* SetType(World&, BodyID, BodyType, resetMassData = true) (WorldBody.hpp:289;
 * WorldBody.cpp:236-246; Body::SetType, Body.cpp:124-149; SetBody,
 * AabbTreeWorld.cpp:2739-2759; ResetMassData, WorldBody.cpp:360-400): type
 * PRX_STATIC, PRX_KINEMATIC or PRX_DYNAMIC.  The body's contacts are destroyed
 * (as by prx_world_destroy_body) and its mass data recomputed from its shapes. */
int prx_world_set_type(prx_world* world, uint32_t body, uint32_t type);
/* SetAcceleration(World&, BodyID, LinearAcceleration2, AngularAcceleration)
 * (d2/WorldBody.hpp; Body::SetAcceleration, Body.cpp:203-238): a = linear x,
 * y, angular. */
int prx_world_set_acceleration(prx_world* world, uint32_t body, const float a[3]);
/* SetTransformation(World&, BodyID, Transformation) (WorldBody.hpp;
 * Body.cpp:270-273 through SetBody, AabbTreeWorld.cpp:2723-2795): xf = p.x,
 * p.y, q.x, q.y.  The body's sweep becomes the single position (p, the angle
 * of q); a changed transformation flags its contacts for updating and
 * synchronises its proxies at the next step (pre-step proxies moved). */
int prx_world_set_transformation(prx_world* world, uint32_t body, const float xf[4]);
/* SetLocation(World&, BodyID, Length2) (WorldBody.cpp:158-163, Body.cpp:275-278). */
int prx_world_set_location(prx_world* world, uint32_t body, const float v[2]);
/* SetAngle(World&, BodyID, Angle) (WorldBody.cpp:165-170, Body.cpp:285-288). */
int prx_world_set_angle(prx_world* world, uint32_t body, float angle);
/* ShiftOrigin(World&, Length2) (AabbTreeWorld.cpp:2221-2242): every body's
 * sweep and every proxy's AABB move by -origin; every contact is updated at
 * the next step. */
int prx_world_shift_origin(prx_world* world, const float origin[2]);
/* SetFriction(World&, ShapeID, NonNegative<Real>) and SetRestitution(World&,
 * ShapeID, Real) (WorldShape.hpp:114,124; WorldShape.cpp:79-96, through
 * SetShape, AabbTreeWorld.cpp:1284-1357): bodies holding the shape are
 * woken; its existing contacts are flagged for updating and keep the values
 * mixed when they were created.  Negative or NaN friction, or NaN
 * restitution: PRX_E_INVALID_ARGUMENT. */
int prx_shape_set_friction(prx_world* world, uint32_t shape, float friction);
int prx_shape_set_restitution(prx_world* world, uint32_t shape, float restitution);
/* One state record (bodies, contact list in world order with manifolds,
 * proxies, joints, then the last step's contact listener calls as
 * prx_world_contact_events gives them, none unless recorded) in the layout of
 * the golden vectors. */
int prx_world_state_record(prx_world* world, uint8_t* buf, size_t cap, size_t* len);
uint32_t prx_world_contact_count(prx_world* world);

/* Getters for a world model over this engine (GetBody's mass data,
 * GetContacts / GetContact / GetManifold, GetTree's leaves).
 *
 * A body's mass data as the world holds it (ResetMassData, WorldBody.cpp:
 * 360-399): out = inverse mass, inverse rotational inertia, local centre x, y. */
int prx_world_get_body_mass(prx_world* world, uint32_t body, float out[4]);
/* One entry of GetContacts(World&) with its Contact and Manifold values
 * (d2/WorldContact.hpp; Contact.hpp, Manifold.hpp).  key_min / key_max: the
 * ContactKey's proxy (tree leaf) ids; flags: bit 0 touching, 1 enabled,
 * 2 sensor, 3 impenetrable, 4 needs updating. */
typedef struct prx_contact_info {
    uint32_t key_min, key_max, id, flags;
    uint32_t body_a, body_b, shape_a, shape_b;
    float friction, restitution;
    uint32_t manifold_type, point_count;
    float local_normal[2], local_point[2];
    struct {
        float local_point[2];
        uint32_t contact_feature;
        float normal_impulse, tangent_impulse;
    } points[2];
} prx_contact_info;
/* The world's contacts in its contact-list order: up to cap into out (null
 * when cap is 0); *count receives the number of contacts. */
int prx_world_get_contacts(prx_world* world, prx_contact_info* out, size_t cap, size_t* count);
/* The world's proxies (DynamicTree leaves: a body's shape, AabbTreeWorld.cpp:
 * 557-579) in creation order: leaf id, body, shape and fat AABB (lo x, lo y,
 * hi x, hi y); up to cap (arrays may be null when cap is 0), *count receives
 * the number of proxies. */
int prx_world_get_proxies(prx_world* world, uint32_t* ids, uint32_t* bodies, uint32_t* shapes, float* aabbs,
                          size_t cap, size_t* count);

/* Contact listeners.
 * A contact's manifold as a pre-solve listener receives it: the manifold the
 * contact had before the step's update (Manifold.hpp:36-120).  type: 0 unset,
 * 1 circles, 2 face A, 3 face B; fields that the type or point count leaves
 * unused are zero. */
typedef struct prx_manifold {
    uint32_t type, point_count;
    float local_normal[2], local_point[2];
    struct {
        float local_point[2];
        uint32_t contact_feature; /* type A | type B << 1 | index A << 8 | index B << 16 */
        float normal_impulse, tangent_impulse;
    } points[2];
} prx_manifold;
/* ContactImpulsesList (ContactImpulsesList.hpp:47-100): the impulses the
 * solve applied at each point of the contact's velocity constraint. */
typedef struct prx_contact_impulses {
    uint32_t count;
    float normal[2], tangent[2];
} prx_contact_impulses;
/* The four contact listeners of the reference (World.hpp:105-139:
 * SetBeginContactListener, SetEndContactListener, SetPreSolveContactListener,
 * SetPostSolveContactListener); null members are not called.  post_solve's
 * solved_iteration is the position iteration that solved the contact's
 * island, 255 (the 8-bit InvalidIteration) when none did
 * (AabbTreeWorld.cpp:1671-1674).
 *
 * The calls, their order and their arguments are the reference's
 * (AabbTreeWorld.cpp:2249-2252, 2703-2719, 1671-1674); they are recorded on
 * the device during the step and delivered once it has completed, in call
 * order, before prx_step / prx_batch_step return (a batch with listeners is
 * stepped one step at a time).  A listener sees the world as it is after the
 * step, and must not step or edit the world. */
typedef struct prx_contact_listener {
    void* user;
    void (*begin_contact)(void* user, uint32_t contact);
    void (*end_contact)(void* user, uint32_t contact);
    void (*pre_solve)(void* user, uint32_t contact, const prx_manifold* old_manifold);
    void (*post_solve)(void* user, uint32_t contact, const prx_contact_impulses* impulses,
                       uint32_t solved_iteration);
} prx_contact_listener;
/* Sets (copies) or clears (null) the world's listeners; setting them turns
 * on recording of the contact listener calls. */
int prx_world_set_contact_listener(prx_world* world, const prx_contact_listener* listener);
/* Turns recording of each step's contact listener calls on or off (the
 * recorded calls also end each state record).  Not for a world inside a
 * multi-world batch (PRX_E_WRONG_STATE): switch it before batching. */
int prx_world_record_contact_events(prx_world* world, int on);
/* Initial size in 32-bit words of the world's record of one step's listener
 * calls (0: the default, 48 words per contact slot, which covers every call
 * of the regular phase).  No reference counterpart: the reference calls its
 * listeners directly.  A step whose calls need more words (the TOI phase can
 * update a contact once per TOI event) is stepped again from the state before
 * it with the record grown to fit, so no call is lost.  Not for a world
 * inside a multi-world batch (PRX_E_WRONG_STATE). */
int prx_world_set_event_record_words(prx_world* world, uint32_t words);
/* The recorded calls of the world's last step: ncalls calls in nwords 32-bit
 * words, each u32 kind (0 begin, 1 end, 2 pre-solve, 3 post-solve), u32
 * contact id, then for pre-solve u32 type | point count << 8, two contact
 * features and 12 floats (local normal, local point, per point local point,
 * normal impulse, tangent impulse), for post-solve u32 point count, u32
 * solved iteration and 4 floats (normal impulses [2], tangent impulses
 * [2]).  words may be null to query the sizes. */
int prx_world_contact_events(prx_world* world, uint32_t* words, size_t cap, size_t* nwords,
                             uint32_t* ncalls);

/* Batched worlds: all resident on the current device, stepped by one
 * persistent launch per step. */
int prx_batch_create(prx_world* const* worlds, int nworlds, const prx_batch_conf* conf,
                     prx_batch** out);
void prx_batch_destroy(prx_batch* batch);
int prx_batch_step(prx_batch* batch, const prx_step_conf* conf, int nsteps);
int prx_batch_sync(prx_batch* batch);
/* Copies device state back into the member worlds (for getters/records). */
int prx_batch_download(prx_batch* batch);
/* Measures the last prx_batch_step call's kernel time (ms, HIP events on the
 * launch stream) and the number of step launches it covered. */
int prx_batch_last_timing(prx_batch* batch, float* ms, int* launches);
/* Sum over worlds of dynamic bodies, and of contacts (after download). */
uint64_t prx_batch_dynamic_bodies(const prx_batch* batch);
/* Totals over the batch's worlds from the last download: out[0] awake
 * dynamic bodies, out[1] proxies, out[2] sum over proxies of (8 + 16 x vertex
 * count) bytes, out[3] contacts, out[4] touching contacts. */
int prx_batch_counts(prx_batch* batch, uint64_t out[5]);
/* Per-world phase timings: enable != 0 turns collection on (subsequent steps
 * accumulate shader-clock cycles per phase); out receives up to max_worlds x
 * 16 counters (proxies, destroy, update, find/add, adjacency, islands,
 * constraints, velocity, position, writeback, sync, update, find/add, total)
 * and the counters are then cleared.  enable = 0 turns collection off. */
int prx_batch_phase_cycles(prx_batch* batch, int enable, uint64_t* out, int max_worlds);
/* Times candidate launch schedules (LDS per workgroup, i.e. workgroups per CU;
 * the wave-wide sweep threshold; launch order) on `nsteps` steps each,
 * starting every trial from the batch's current state, keeps the fastest in
 * the batch (and in *best when non-null) and restores the state: the
 * simulation is not advanced.  A candidate must beat the schedule in hand by
 * 2 % to replace it. */
int prx_batch_autotune(prx_batch* batch, const prx_step_conf* step, int nsteps,
                       prx_batch_conf* best);

/* One very large world on the whole device (graph-coloured mode).  No
 * reference counterpart as an API: it is Step(World&, const StepConf&)
 * (World.hpp:917-920) for one world too large for a single wavefront, with
 * the DynamicTree broadphase (DynamicTree.cpp:673-700, AabbTreeWorld.cpp:
 * 709-751) replaced by a linear BVH giving the same pair sets, and the
 * island Gauss-Seidel solve (AabbTreeWorld.cpp:1563-1682) replaced by a
 * colour-parallel Gauss-Seidel over all awake bodies (no sleeping).  The
 * world's state moves to the device (its getters read it back). */
typedef struct prx_big prx_big;
int prx_big_create(prx_world* world, prx_big** out);
void prx_big_destroy(prx_big* big);
/* At most 16 velocity and 16 position iterations (PRX_E_UNSUPPORTED otherwise).
 * With doToi, a reduced TOI phase (AabbTreeWorld.cpp:1947-2063 restricted to
 * static shapes): up to two time-of-impact events per awake dynamic body and
 * step against static shapes, each solved as the island of the body and the
 * static body at toiResolutionRate with toiPositionIters / toiVelocityIters,
 * then the rest of the step's motion. */
int prx_big_step(prx_big* big, const prx_step_conf* conf, int nsteps);
int prx_big_sync(prx_big* big);
int prx_big_last_timing(prx_big* big, float* ms, int* launches);
/* out[0] smallest separation over touching contacts at the current poses
 * (negative: penetration), out[1] touching contacts and out[2] colours of the
 * last solve, out[3] contacts. */
int prx_big_metrics(prx_big* big, float out[4]);
/* Broadphase pair finding alone (FindContacts, AabbTreeWorld.cpp:709-751, with
 * the LBVH): n proxies with fat AABBs (lo x, lo y, hi x, hi y), owning body
 * and moved flag; writes up to cap pairs (lower index, higher index) of
 * proxies of different bodies whose AABBs overlap, at least one moved;
 * *count receives the number found. */
int prx_lbvh_pairs(const float* aabbs, const int32_t* bodies, const uint8_t* moved, int n, int32_t* pairs,
                   int cap, int* count);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops): prx_lbvh_pairs
 * on DEVICE buffers, with no hipMalloc, hipFree or other allocation.  The
 * LBVH's working arrays live in a fixed per-device array of about 6.6 MB
 * (held until the process ends) sized for at most 65536 proxies; more is
 * PRX_E_LENGTH.  aabbs (4 floats per proxy), bodies, moved, pairs (8-byte
 * aligned) and count are device pointers; the work is queued on `stream` (a
 * hipStream_t, null for the null stream) and not waited for.  *count receives
 * the number of pairs found (above cap, only the first cap were written), or
 * -1 if the tree exceeded its depth bound.  The pairs and their order equal
 * prx_lbvh_pairs's.  Calls that overlap on one device race on that array. */
int prx_lbvh_pairs_noalloc(const float* aabbs, const int32_t* bodies, const uint8_t* moved, int n, int32_t* pairs,
                           int cap, int* count, void* stream);

/* Adjoint of K regular steps of a batch.  No reference counterpart: it
 * differentiates Step (AabbTreeWorld.cpp:1563-1682; contact solves
 * ContactSolver.cpp, revolute joints RevoluteJointConf.cpp:127-433) with each
 * step's active set (touching contacts, islands, clamp branches, block-solver
 * cases, early exits) held as the forward pass took it.  Worlds of disks
 * centred on their bodies without joints use a hand-derived reverse kernel;
 * any other world (polygons, edges, two-point manifolds, revolute joints)
 * the general one, which takes each update's vector-Jacobian product from
 * dual-number copies of the solver functions.  Loss per world: sum over
 * dynamic bodies of |final position - target|^2.  The forward pass keeps a
 * checkpoint every `every` steps; the backward pass recomputes each segment
 * from its checkpoint, recording the solves, and reverses them.  Sleeping
 * is not differentiated (a body put to sleep has its velocity zeroed). */
typedef struct prx_adjoint prx_adjoint;
int prx_adjoint_create(prx_batch* batch, int nsteps, int every, const prx_step_conf* conf, prx_adjoint** out);
void prx_adjoint_destroy(prx_adjoint* adj);
/* targets: 2 floats per batch body (batch body order: world by world).
 * Outputs (each may be null): loss per world; d loss / d initial velocity,
 * 3 floats (vx, vy, w) per batch body; d loss / d friction and d loss / d
 * restitution per batch shape.  The batch is left at its starting state. */
int prx_adjoint_run(prx_adjoint* adj, const prx_step_conf* conf, const float* targets, float* loss,
                    float* grad_velocity, float* grad_friction, float* grad_restitution);
/* The two halves of prx_adjoint_run for a loss of the caller's choosing.
 * prx_adjoint_forward steps the batch nsteps steps with checkpoints and leaves
 * it at the final state; final_state (may be null) receives 6 floats per batch
 * body: x, y, angle (the sweep's position 1), vx, vy, w.
 * prx_adjoint_backward takes the gradient of the loss with respect to that
 * final state (6 floats per batch body, in the same order; entries of bodies
 * that do not move are ignored) and returns the gradients with respect to
 * initial velocities, friction and restitution as prx_adjoint_run does
 * (joint torques by prx_adjoint_joint_gradients); it reverses the last
 * forward pass (PRX_E_WRONG_STATE without one) and leaves the batch at its
 * starting state.  No reference counterpart: the reference has no adjoint. */
int prx_adjoint_forward(prx_adjoint* adj, const prx_step_conf* conf, float* final_state);
int prx_adjoint_backward(prx_adjoint* adj, const prx_step_conf* conf, const float* state_bar, float* grad_velocity,
                         float* grad_friction, float* grad_restitution);
/* Diagnostic: on != 0 runs a disk-only batch through the general reverse
 * kernel (to check the two against each other); on = 0 returns to the
 * hand-derived one where the batch allows it (else PRX_E_UNSUPPORTED). */
int prx_adjoint_use_general(prx_adjoint* adj, int on);
/* d loss / d maximum motor torque (RevoluteJointConf::maxMotorTorque) of the
 * last run, one float per batch joint (world by world, creation order). */
int prx_adjoint_joint_gradients(prx_adjoint* adj, float* grad_max_motor_torque);
/* Counters of the last run: constraint slots reversed, those with a
 * restitution bias, position-solver updates and velocity updates reversed. */
int prx_adjoint_diagnostics(prx_adjoint* adj, float out[4]);
/* Device time (ms) of the last run: K forward steps, K recorded steps and K
 * reversed steps. */
int prx_adjoint_timing(prx_adjoint* adj, float* ms);

/* Diagnostic: UnitVec::Get(angle) = {cos, sin} (UnitVec.cpp:27-30) evaluated
 * by the device routine for n angles; out receives n (cos, sin) pairs.  Used
 * by the tests to compare the device rotation with the host C library. */
int prx_selftest_sincos(const float* angles, float* out, int n);

#ifdef __cplusplus
}
#endif

#endif /* PLAYRHO_AMD_H */
