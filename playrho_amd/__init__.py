"""playrho_amd — MI355X step engine for PlayRho worlds.

Python binding over the C ABI in include/playrho_amd.h (libplayrho_amd.so,
built in-tree by playrho_amd._build).  The stepping path runs only on the HIP
device: there is no CPU fallback, and every call fails loudly when the native
library is missing or no device is present.

    import playrho_amd as pa
    w = pa.World()
    w.build_scene("pyramid", [20])
    conf = pa.StepConf()          # the reference's StepConf defaults (doToi = 1)
    w.step(conf)
    xf = w.transformation(5)
"""
import ctypes
import os

import numpy as np

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "libplayrho_amd.so")

PRX_OK = 0
STATUS = {1: "InvalidArgument", 2: "LengthError", 3: "WrongState", 4: "OutOfRange",
          5: "DeviceError", 6: "Unsupported"}

STATIC, KINEMATIC, DYNAMIC = 0, 1, 2
SHAPE_DISK, SHAPE_POLYGON, SHAPE_BOX, SHAPE_EDGE = 0, 1, 2, 3


class PrxError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s: %s" % (STATUS.get(code, "Error%d" % code), msg))
        self.code = code


class StepConf(ctypes.Structure):
    """StepConf fields of the regular step (StepConf.hpp:49-392)."""
    _fields_ = [(n, ctypes.c_float) for n in (
        "delta_time", "dt_ratio", "min_still_time_to_sleep", "linear_slop", "angular_slop",
        "reg_resolution_rate", "reg_min_separation", "reg_min_momentum", "velocity_threshold",
        "max_translation", "max_rotation", "max_linear_correction", "max_angular_correction",
        "linear_sleep_tolerance", "angular_sleep_tolerance", "displace_multiplier",
        "aabb_extension", "max_circles_ratio")] + [
        (n, ctypes.c_uint32) for n in ("reg_velocity_iters", "reg_position_iters",
                                       "do_warm_start", "do_toi", "do_blocksolve")] + [
        (n, ctypes.c_float) for n in ("toi_resolution_rate", "toi_min_separation", "toi_min_momentum",
                                      "target_depth", "tolerance")] + [
        (n, ctypes.c_uint32) for n in ("toi_velocity_iters", "toi_position_iters", "max_toi_root_iters",
                                       "max_toi_iters", "max_distance_iters", "max_sub_steps")]

    def __init__(self, **kw):
        super().__init__()
        lib().prx_step_conf_default(ctypes.byref(self))
        for k, v in kw.items():
            setattr(self, k, v)


class ManifoldPoint(ctypes.Structure):
    _fields_ = [("local_point", ctypes.c_float * 2), ("contact_feature", ctypes.c_uint32),
                ("normal_impulse", ctypes.c_float), ("tangent_impulse", ctypes.c_float)]


class Manifold(ctypes.Structure):
    """The manifold a pre-solve listener receives (Manifold.hpp:36-120)."""
    _fields_ = [("type", ctypes.c_uint32), ("point_count", ctypes.c_uint32),
                ("local_normal", ctypes.c_float * 2), ("local_point", ctypes.c_float * 2),
                ("points", ManifoldPoint * 2)]


class ContactInfo(ctypes.Structure):
    """One entry of GetContacts with its Contact and Manifold values
    (d2/WorldContact.hpp; Contact.hpp, Manifold.hpp).  flags: bit 0 touching,
    1 enabled, 2 sensor, 3 impenetrable, 4 needs updating."""
    _fields_ = [("key_min", ctypes.c_uint32), ("key_max", ctypes.c_uint32), ("id", ctypes.c_uint32),
                ("flags", ctypes.c_uint32), ("body_a", ctypes.c_uint32), ("body_b", ctypes.c_uint32),
                ("shape_a", ctypes.c_uint32), ("shape_b", ctypes.c_uint32), ("friction", ctypes.c_float),
                ("restitution", ctypes.c_float), ("manifold_type", ctypes.c_uint32),
                ("point_count", ctypes.c_uint32), ("local_normal", ctypes.c_float * 2),
                ("local_point", ctypes.c_float * 2), ("points", ManifoldPoint * 2)]


class ContactImpulses(ctypes.Structure):
    """ContactImpulsesList (ContactImpulsesList.hpp:47-100)."""
    _fields_ = [("count", ctypes.c_uint32), ("normal", ctypes.c_float * 2), ("tangent", ctypes.c_float * 2)]


_BeginEndFn = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_uint32)
_PreSolveFn = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_uint32, ctypes.POINTER(Manifold))
_PostSolveFn = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_uint32, ctypes.POINTER(ContactImpulses),
                                ctypes.c_uint32)


class ContactListener(ctypes.Structure):
    _fields_ = [("user", ctypes.c_void_p), ("begin_contact", _BeginEndFn), ("end_contact", _BeginEndFn),
                ("pre_solve", _PreSolveFn), ("post_solve", _PostSolveFn)]


# Listener call kinds and their record sizes in 32-bit words.
EV_BEGIN, EV_END, EV_PRE_SOLVE, EV_POST_SOLVE = 0, 1, 2, 3
_EV_WORDS = (2, 2, 17, 8)


def parse_contact_events(words, ncalls):
    """Listener calls from their recorded words, in call order:
    ("begin", id), ("end", id), ("pre_solve", id, manifold dict),
    ("post_solve", id, (count, normal impulses, tangent impulses), iteration)."""
    w = np.asarray(words, dtype=np.uint32)
    f = w.view(np.float32)
    out, o = [], 0
    for _ in range(ncalls):
        kind, cid = int(w[o]), int(w[o + 1])
        if kind == EV_BEGIN:
            out.append(("begin", cid))
        elif kind == EV_END:
            out.append(("end", cid))
        elif kind == EV_PRE_SOLVE:
            n = int((w[o + 2] >> 8) & 0xff)
            pts = [dict(local_point=(float(f[o + 9 + 4 * q]), float(f[o + 10 + 4 * q])),
                        contact_feature=int(w[o + 3 + q]), normal_impulse=float(f[o + 11 + 4 * q]),
                        tangent_impulse=float(f[o + 12 + 4 * q])) for q in range(n)]
            out.append(("pre_solve", cid, dict(type=int(w[o + 2] & 0xff), point_count=n,
                                                local_normal=(float(f[o + 5]), float(f[o + 6])),
                                                local_point=(float(f[o + 7]), float(f[o + 8])), points=pts)))
        elif kind == EV_POST_SOLVE:
            n = int(w[o + 2])
            out.append(("post_solve", cid, (n, tuple(float(x) for x in f[o + 4:o + 4 + n]),
                                            tuple(float(x) for x in f[o + 6:o + 6 + n])), int(w[o + 3])))
        else:
            raise ValueError("unknown listener call kind %d" % kind)
        o += _EV_WORDS[kind]
    return out


class WorldConf(ctypes.Structure):
    _fields_ = [("abi_version", ctypes.c_uint32), ("min_vertex_radius", ctypes.c_float),
                ("max_vertex_radius", ctypes.c_float), ("contact_capacity_per_proxy", ctypes.c_uint32)]


class ShapeConf(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_uint32), ("vertex_radius", ctypes.c_float), ("radius", ctypes.c_float),
                ("hx", ctypes.c_float), ("hy", ctypes.c_float), ("location", ctypes.c_float * 2),
                ("points", ctypes.POINTER(ctypes.c_float)), ("npoints", ctypes.c_uint32),
                ("density", ctypes.c_float), ("friction", ctypes.c_float), ("restitution", ctypes.c_float),
                ("sensor", ctypes.c_uint32), ("category_bits", ctypes.c_uint16),
                ("mask_bits", ctypes.c_uint16), ("group_index", ctypes.c_int16)]


class BodyConf(ctypes.Structure):
    _fields_ = [("type", ctypes.c_uint32), ("location", ctypes.c_float * 2), ("angle", ctypes.c_float),
                ("linear_velocity", ctypes.c_float * 2), ("angular_velocity", ctypes.c_float),
                ("linear_acceleration", ctypes.c_float * 2), ("angular_acceleration", ctypes.c_float),
                ("linear_damping", ctypes.c_float), ("angular_damping", ctypes.c_float),
                ("allow_sleep", ctypes.c_uint32), ("awake", ctypes.c_uint32),
                ("fixed_rotation", ctypes.c_uint32), ("bullet", ctypes.c_uint32),
                ("enabled", ctypes.c_uint32), ("shapes", ctypes.POINTER(ctypes.c_uint32)),
                ("nshapes", ctypes.c_uint32)]


class RevoluteJointConf(ctypes.Structure):
    _fields_ = [("body_a", ctypes.c_uint32), ("body_b", ctypes.c_uint32),
                ("local_anchor_a", ctypes.c_float * 2), ("local_anchor_b", ctypes.c_float * 2),
                ("reference_angle", ctypes.c_float), ("enable_limit", ctypes.c_uint32),
                ("lower_angle", ctypes.c_float), ("upper_angle", ctypes.c_float),
                ("enable_motor", ctypes.c_uint32), ("motor_speed", ctypes.c_float),
                ("max_motor_torque", ctypes.c_float), ("collide_connected", ctypes.c_uint32)]


class DistanceJointConf(ctypes.Structure):
    """DistanceJointConf (d2/DistanceJointConf.hpp:62-130); the reference's
    defaults: length 1, frequency 0 (rigid), damping ratio 0."""
    _fields_ = [("body_a", ctypes.c_uint32), ("body_b", ctypes.c_uint32),
                ("local_anchor_a", ctypes.c_float * 2), ("local_anchor_b", ctypes.c_float * 2),
                ("length", ctypes.c_float), ("frequency", ctypes.c_float), ("damping_ratio", ctypes.c_float),
                ("collide_connected", ctypes.c_uint32)]

    def __init__(self, **kw):
        kw.setdefault("length", 1.0)
        super().__init__(**kw)


class RopeJointConf(ctypes.Structure):
    """RopeJointConf (d2/RopeJointConf.hpp:67-106); the reference's defaults:
    local anchors (-1, 0) and (1, 0), maximum length 0."""
    _fields_ = [("body_a", ctypes.c_uint32), ("body_b", ctypes.c_uint32),
                ("local_anchor_a", ctypes.c_float * 2), ("local_anchor_b", ctypes.c_float * 2),
                ("max_length", ctypes.c_float), ("collide_connected", ctypes.c_uint32)]

    def __init__(self, **kw):
        kw.setdefault("local_anchor_a", (-1.0, 0.0))
        kw.setdefault("local_anchor_b", (1.0, 0.0))
        super().__init__(**kw)


class FrictionJointConf(ctypes.Structure):
    """FrictionJointConf (d2/FrictionJointConf.hpp:58-117); maximum force and
    torque default to 0."""
    _fields_ = [("body_a", ctypes.c_uint32), ("body_b", ctypes.c_uint32),
                ("local_anchor_a", ctypes.c_float * 2), ("local_anchor_b", ctypes.c_float * 2),
                ("max_force", ctypes.c_float), ("max_torque", ctypes.c_float),
                ("collide_connected", ctypes.c_uint32)]


class MotorJointConf(ctypes.Structure):
    """MotorJointConf (d2/MotorJointConf.hpp:58-138); the reference's defaults:
    offsets 0, maximum force 1, maximum torque 1, correction factor 0.3."""
    _fields_ = [("body_a", ctypes.c_uint32), ("body_b", ctypes.c_uint32),
                ("linear_offset", ctypes.c_float * 2), ("angular_offset", ctypes.c_float),
                ("max_force", ctypes.c_float), ("max_torque", ctypes.c_float),
                ("correction_factor", ctypes.c_float), ("collide_connected", ctypes.c_uint32)]

    def __init__(self, **kw):
        kw.setdefault("max_force", 1.0)
        kw.setdefault("max_torque", 1.0)
        kw.setdefault("correction_factor", 0.3)
        super().__init__(**kw)


class WheelJointConf(ctypes.Structure):
    """WheelJointConf (d2/WheelJointConf.hpp:71-170); the reference's defaults:
    axis (1, 0), frequency 2, damping ratio 0.7, motor off."""
    _fields_ = [("body_a", ctypes.c_uint32), ("body_b", ctypes.c_uint32),
                ("local_anchor_a", ctypes.c_float * 2), ("local_anchor_b", ctypes.c_float * 2),
                ("local_axis_a", ctypes.c_float * 2), ("enable_motor", ctypes.c_uint32),
                ("motor_speed", ctypes.c_float), ("max_motor_torque", ctypes.c_float),
                ("frequency", ctypes.c_float), ("damping_ratio", ctypes.c_float),
                ("collide_connected", ctypes.c_uint32)]

    def __init__(self, **kw):
        kw.setdefault("local_axis_a", (1.0, 0.0))
        kw.setdefault("frequency", 2.0)
        kw.setdefault("damping_ratio", 0.7)
        super().__init__(**kw)


class PrismaticJointConf(ctypes.Structure):
    """PrismaticJointConf (d2/PrismaticJointConf.hpp:61-217); the reference's
    defaults: axis (1, 0), reference angle 0, limit and motor off."""
    _fields_ = [("body_a", ctypes.c_uint32), ("body_b", ctypes.c_uint32),
                ("local_anchor_a", ctypes.c_float * 2), ("local_anchor_b", ctypes.c_float * 2),
                ("local_axis_a", ctypes.c_float * 2), ("reference_angle", ctypes.c_float),
                ("enable_limit", ctypes.c_uint32), ("lower_translation", ctypes.c_float),
                ("upper_translation", ctypes.c_float), ("enable_motor", ctypes.c_uint32),
                ("motor_speed", ctypes.c_float), ("max_motor_force", ctypes.c_float),
                ("collide_connected", ctypes.c_uint32)]

    def __init__(self, **kw):
        kw.setdefault("local_axis_a", (1.0, 0.0))
        super().__init__(**kw)


class GearJointConf(ctypes.Structure):
    """GearJointConf (d2/GearJointConf.hpp:60-126): coordinate1 + ratio *
    coordinate2 = constant over two revolute or prismatic joints; bodies A, B
    the joints' second bodies, C, D their first; per side (ac, bd) a
    prismatic joint's local anchors A, B and local x axis, or a revolute
    joint's reference angle in [0].  World.gear_joint_conf makes one from two
    joints (GetGearJointConf)."""
    _fields_ = [("body_a", ctypes.c_uint32), ("body_b", ctypes.c_uint32), ("body_c", ctypes.c_uint32),
                ("body_d", ctypes.c_uint32), ("prismatic_ac", ctypes.c_uint32), ("prismatic_bd", ctypes.c_uint32),
                ("ac", ctypes.c_float * 6), ("bd", ctypes.c_float * 6), ("ratio", ctypes.c_float),
                ("constant", ctypes.c_float), ("collide_connected", ctypes.c_uint32)]

    def __init__(self, **kw):
        kw.setdefault("ratio", 1.0)
        super().__init__(**kw)


class TargetJointConf(ctypes.Structure):
    """TargetJointConf (d2/TargetJointConf.hpp:56-161), a joint with body B
    only; the reference's defaults: frequency 5, damping ratio 0.7, maximum
    force 0."""
    _fields_ = [("body", ctypes.c_uint32), ("target", ctypes.c_float * 2), ("local_anchor", ctypes.c_float * 2),
                ("max_force", ctypes.c_float), ("frequency", ctypes.c_float), ("damping_ratio", ctypes.c_float),
                ("collide_connected", ctypes.c_uint32)]

    def __init__(self, **kw):
        kw.setdefault("frequency", 5.0)
        kw.setdefault("damping_ratio", 0.7)
        super().__init__(**kw)


class PulleyJointConf(ctypes.Structure):
    """PulleyJointConf (d2/PulleyJointConf.hpp:65-137); the reference's defaults:
    ground anchors (-1, 1), (1, 1), local anchors (-1, 0), (1, 0), lengths 0,
    ratio 1, collide_connected 1."""
    _fields_ = [("body_a", ctypes.c_uint32), ("body_b", ctypes.c_uint32),
                ("ground_anchor_a", ctypes.c_float * 2), ("ground_anchor_b", ctypes.c_float * 2),
                ("local_anchor_a", ctypes.c_float * 2), ("local_anchor_b", ctypes.c_float * 2),
                ("length_a", ctypes.c_float), ("length_b", ctypes.c_float), ("ratio", ctypes.c_float),
                ("collide_connected", ctypes.c_uint32)]

    def __init__(self, **kw):
        kw.setdefault("ground_anchor_a", (-1.0, 1.0))
        kw.setdefault("ground_anchor_b", (1.0, 1.0))
        kw.setdefault("local_anchor_a", (-1.0, 0.0))
        kw.setdefault("local_anchor_b", (1.0, 0.0))
        kw.setdefault("ratio", 1.0)
        kw.setdefault("collide_connected", 1)
        super().__init__(**kw)


class WeldJointConf(ctypes.Structure):
    """WeldJointConf (d2/WeldJointConf.hpp:62-130); frequency 0 is rigid."""
    _fields_ = [("body_a", ctypes.c_uint32), ("body_b", ctypes.c_uint32),
                ("local_anchor_a", ctypes.c_float * 2), ("local_anchor_b", ctypes.c_float * 2),
                ("reference_angle", ctypes.c_float), ("frequency", ctypes.c_float), ("damping_ratio", ctypes.c_float),
                ("collide_connected", ctypes.c_uint32)]


class StepStats(ctypes.Structure):
    """StepStats (StepStats.hpp:30-201): PreStepStats, RegStepStats and ToiStepStats."""
    _fields_ = [(n, ctypes.c_uint32) for n in (
        "pre_proxies_created", "pre_proxies_moved", "pre_contacts_destroyed", "pre_contacts_added",
        "pre_contacts_updated", "pre_contacts_skipped")] + [
        ("reg_min_separation", ctypes.c_float), ("reg_max_inc_impulse", ctypes.c_float)] + [
        (n, ctypes.c_uint32) for n in (
            "reg_islands_found", "reg_islands_solved", "reg_bodies_slept", "reg_max_island_bodies",
            "reg_contacts_added", "reg_contacts_updated", "reg_contacts_skipped", "reg_proxies_moved",
            "reg_sum_pos_iters", "reg_sum_vel_iters")] + [
        ("toi_min_separation", ctypes.c_float), ("toi_max_inc_impulse", ctypes.c_float)] + [
        (n, ctypes.c_uint32) for n in (
            "toi_islands_found", "toi_islands_solved", "toi_contacts_found", "toi_contacts_at_max_sub_steps",
            "toi_contacts_updated_toi", "toi_contacts_updated_touching", "toi_contacts_skipped_touching",
            "toi_contacts_added", "toi_proxies_moved", "toi_sum_pos_iters", "toi_sum_vel_iters",
            "toi_max_dist_iters", "toi_max_toi_iters", "toi_max_root_iters")]


class BatchConf(ctypes.Structure):
    _fields_ = [("threads_per_world", ctypes.c_uint32), ("use_graph", ctypes.c_uint32),
                ("order", ctypes.c_uint32), ("coop_min_contacts", ctypes.c_uint32),
                ("lds_bytes", ctypes.c_uint32)]


_lib = None

_SIGS = {
    "prx_abi_version": (ctypes.c_int, []),
    "prx_last_error": (ctypes.c_char_p, []),
    "prx_step_conf_default": (None, [ctypes.c_void_p]),
    "prx_world_conf_default": (None, [ctypes.c_void_p]),
    "prx_batch_conf_default": (None, [ctypes.c_void_p]),
    "prx_world_create": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p)]),
    "prx_world_destroy": (None, [ctypes.c_void_p]),
    "prx_shape_create": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint32)]),
    "prx_compound_shape_create": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint32),
                                                 ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint32),
                                                 ctypes.POINTER(ctypes.c_uint32)]),
    "prx_body_create": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint32)]),
    "prx_scene_build": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p, ctypes.POINTER(ctypes.c_int32),
                                       ctypes.c_int]),
    "prx_world_load_desc": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_size_t]),
    "prx_world_save_desc": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                                           ctypes.POINTER(ctypes.c_size_t)]),
    "prx_world_body_count": (ctypes.c_uint32, [ctypes.c_void_p]),
    "prx_joint_create": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint32)]),
    "prx_distance_joint_create": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint32)]),
    "prx_weld_joint_create": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint32)]),
    "prx_rope_joint_create": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint32)]),
    "prx_friction_joint_create": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint32)]),
    "prx_motor_joint_create": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint32)]),
    "prx_wheel_joint_create": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint32)]),
    "prx_pulley_joint_create": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint32)]),
    "prx_prismatic_joint_create": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p,
                                                  ctypes.POINTER(ctypes.c_uint32)]),
    "prx_target_joint_create": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint32)]),
    "prx_gear_joint_create": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint32)]),
    "prx_gear_joint_conf_get": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_float,
                                               ctypes.c_void_p]),
    "prx_world_joint_count": (ctypes.c_uint32, [ctypes.c_void_p]),
    "prx_world_shape_count": (ctypes.c_uint32, [ctypes.c_void_p]),
    "prx_step": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "prx_world_get_bodies": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "prx_world_get_transformation": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32,
                                                    ctypes.POINTER(ctypes.c_float)]),
    "prx_shape_set_friction": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_float]),
    "prx_shape_set_restitution": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_float]),
    "prx_world_set_velocity": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32,
                                              ctypes.POINTER(ctypes.c_float)]),
    "prx_world_destroy_body": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32]),
    "prx_world_destroy_joint": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32]),
    "prx_world_set_type": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32]),
    "prx_world_state_record": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                                              ctypes.POINTER(ctypes.c_size_t)]),
    "prx_world_contact_count": (ctypes.c_uint32, [ctypes.c_void_p]),
    "prx_world_get_body_mass": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, ctypes.POINTER(ctypes.c_float)]),
    "prx_world_get_contacts": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                                              ctypes.POINTER(ctypes.c_size_t)]),
    "prx_world_get_proxies": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                             ctypes.c_void_p, ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t)]),
    "prx_world_set_acceleration": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32,
                                                  ctypes.POINTER(ctypes.c_float)]),
    "prx_batch_create": (ctypes.c_int, [ctypes.POINTER(ctypes.c_void_p), ctypes.c_int, ctypes.c_void_p,
                                        ctypes.POINTER(ctypes.c_void_p)]),
    "prx_batch_destroy": (None, [ctypes.c_void_p]),
    "prx_batch_step": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]),
    "prx_batch_sync": (ctypes.c_int, [ctypes.c_void_p]),
    "prx_batch_download": (ctypes.c_int, [ctypes.c_void_p]),
    "prx_batch_last_timing": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_float),
                                             ctypes.POINTER(ctypes.c_int)]),
    "prx_batch_dynamic_bodies": (ctypes.c_uint64, [ctypes.c_void_p]),
    "prx_batch_autotune": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]),
    "prx_batch_counts": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "prx_batch_phase_cycles": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int]),
    "prx_selftest_sincos": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]),
    "prx_adjoint_create": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                          ctypes.POINTER(ctypes.c_void_p)]),
    "prx_adjoint_destroy": (None, [ctypes.c_void_p]),
    "prx_adjoint_run": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                       ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "prx_adjoint_timing": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_float)]),
    "prx_adjoint_forward": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "prx_adjoint_backward": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                            ctypes.c_void_p, ctypes.c_void_p]),
    "prx_adjoint_use_general": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "prx_adjoint_joint_gradients": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "prx_adjoint_diagnostics": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "prx_world_record_contact_events": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "prx_world_set_transformation": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p]),
    "prx_world_set_location": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p]),
    "prx_world_set_angle": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_float]),
    "prx_world_shift_origin": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "prx_world_set_contact_listener": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "prx_world_contact_events": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                                                ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_uint32)]),
    "prx_big_create": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p)]),
    "prx_big_destroy": (None, [ctypes.c_void_p]),
    "prx_big_step": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]),
    "prx_big_sync": (ctypes.c_int, [ctypes.c_void_p]),
    "prx_big_last_timing": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_float),
                                           ctypes.POINTER(ctypes.c_int)]),
    "prx_big_metrics": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "prx_lbvh_pairs": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                      ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]),
    "prx_world_set_event_record_words": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32]),
    "prx_lbvh_pairs_noalloc": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                              ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]),
}


def lib():
    """Loads libplayrho_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libplayrho_amd.so is not built: run __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def exported_symbols():
    return sorted(_SIGS)


def _check(code):
    if code != PRX_OK:
        raise PrxError(code, lib().prx_last_error().decode())


def device_available():
    """True when a HIP device is visible (counting devices does not start the
    device runtime in this process)."""
    try:
        import torch
        return torch.cuda.device_count() > 0
    except Exception:  # pragma: no cover
        return False


def require_device():
    if not device_available():
        raise RuntimeError("playrho_amd needs a HIP device: the step path has no CPU fallback")


class World:
    """A PlayRho world whose steps run on the device (plain per-world form)."""

    def __init__(self, conf=None):
        self.h = ctypes.c_void_p()
        _check(lib().prx_world_create(ctypes.byref(conf) if conf is not None else None, ctypes.byref(self.h)))

    @classmethod
    def from_desc(cls, desc: bytes, conf=None):
        w = cls(conf)
        _check(lib().prx_world_load_desc(w.h, desc, len(desc)))
        return w

    @classmethod
    def scene(cls, name, params=(), conf=None):
        w = cls(conf)
        w.build_scene(name, params)
        return w

    def __del__(self):
        h = getattr(self, "h", None)
        if h and h.value and _lib is not None:
            _lib.prx_world_destroy(h)
            self.h = None

    def build_scene(self, name, params=()):
        arr = (ctypes.c_int32 * max(1, len(params)))(*params)
        _check(lib().prx_scene_build(self.h, name.encode(), arr, len(params)))

    def create_shape(self, conf: ShapeConf):
        sid = ctypes.c_uint32()
        _check(lib().prx_shape_create(self.h, ctypes.byref(conf), ctypes.byref(sid)))
        return sid.value

    def create_compound_shape(self, conf: ShapeConf, hull_sizes=None):
        """A chain (conf.kind 4) or multi shape (conf.kind 5, hull_sizes: points
        per hull) as its children's shape ids (prx_compound_shape_create); a
        body takes the shape by listing all of them."""
        hs = (ctypes.c_uint32 * max(1, len(hull_sizes or [])))(*(hull_sizes or []))
        first, count = ctypes.c_uint32(), ctypes.c_uint32()
        _check(lib().prx_compound_shape_create(self.h, ctypes.byref(conf), hs, len(hull_sizes or []),
                                               ctypes.byref(first), ctypes.byref(count)))
        return list(range(first.value, first.value + count.value))

    def create_body(self, conf: BodyConf):
        bid = ctypes.c_uint32()
        _check(lib().prx_body_create(self.h, ctypes.byref(conf), ctypes.byref(bid)))
        return bid.value

    def create_joint(self, conf):
        """CreateJoint with a RevoluteJointConf, DistanceJointConf,
        WeldJointConf, RopeJointConf, FrictionJointConf, MotorJointConf,
        WheelJointConf, PulleyJointConf, PrismaticJointConf, TargetJointConf or GearJointConf; returns
        the joint id (shared by all kinds)."""
        fn = {RevoluteJointConf: lib().prx_joint_create, DistanceJointConf: lib().prx_distance_joint_create,
              WeldJointConf: lib().prx_weld_joint_create, RopeJointConf: lib().prx_rope_joint_create,
              FrictionJointConf: lib().prx_friction_joint_create,
              MotorJointConf: lib().prx_motor_joint_create,
              WheelJointConf: lib().prx_wheel_joint_create,
              PulleyJointConf: lib().prx_pulley_joint_create,
              PrismaticJointConf: lib().prx_prismatic_joint_create,
              TargetJointConf: lib().prx_target_joint_create,
              GearJointConf: lib().prx_gear_joint_create}.get(type(conf))
        if fn is None:
            raise TypeError("not a joint conf: %r" % type(conf))
        jid = ctypes.c_uint32()
        _check(fn(self.h, ctypes.byref(conf), ctypes.byref(jid)))
        return jid.value

    def gear_joint_conf(self, joint1, joint2, ratio=1.0):
        """GetGearJointConf(world, joint1, joint2, ratio) at the bodies'
        present positions."""
        conf = GearJointConf()
        _check(lib().prx_gear_joint_conf_get(self.h, joint1, joint2, ratio, ctypes.byref(conf)))
        return conf

    @property
    def joint_count(self):
        return lib().prx_world_joint_count(self.h)

    @property
    def body_count(self):
        return lib().prx_world_body_count(self.h)

    def step(self, conf: StepConf = None):
        conf = conf or StepConf()
        st = StepStats()
        _check(lib().prx_step(self.h, ctypes.byref(conf), ctypes.byref(st)))
        return st

    def bodies(self):
        n = self.body_count
        out = np.zeros((n, 10), dtype=np.float32)
        flags = np.zeros(n, dtype=np.uint32)
        _check(lib().prx_world_get_bodies(self.h, out.ctypes.data, flags.ctypes.data))
        return out, flags

    def transformation(self, body):
        out = (ctypes.c_float * 4)()
        _check(lib().prx_world_get_transformation(self.h, body, out))
        return tuple(out)

    def set_velocity(self, body, v):
        arr = (ctypes.c_float * 3)(*v)
        _check(lib().prx_world_set_velocity(self.h, body, arr))

    def destroy_body(self, body):
        """Destroy(World&, BodyID) (AabbTreeWorld.cpp:1062-1104)."""
        _check(lib().prx_world_destroy_body(self.h, int(body)))

    def set_type(self, body, body_type):
        """SetType(World&, BodyID, BodyType) (WorldBody.cpp:236-246): 0
        static, 1 kinematic, 2 dynamic; the mass data is reset."""
        _check(lib().prx_world_set_type(self.h, int(body), int(body_type)))

    def destroy_joint(self, joint):
        """Destroy(World&, JointID) (AabbTreeWorld.cpp:1215-1227)."""
        _check(lib().prx_world_destroy_joint(self.h, int(joint)))

    def set_acceleration(self, body, a):
        """SetAcceleration (Body.cpp:203-238): a = (linear x, y, angular)."""
        arr = (ctypes.c_float * 3)(*a)
        _check(lib().prx_world_set_acceleration(self.h, body, arr))

    def body_mass(self, body):
        """(inverse mass, inverse rotational inertia, local centre x, y)
        (ResetMassData, WorldBody.cpp:360-399)."""
        out = (ctypes.c_float * 4)()
        _check(lib().prx_world_get_body_mass(self.h, body, out))
        return tuple(out)

    def contacts(self):
        """GetContacts in the world's contact-list order, as ContactInfo."""
        n = ctypes.c_size_t()
        _check(lib().prx_world_get_contacts(self.h, None, 0, ctypes.byref(n)))
        out = (ContactInfo * max(n.value, 1))()
        _check(lib().prx_world_get_contacts(self.h, out, n.value, ctypes.byref(n)))
        return list(out[:n.value])

    def proxies(self):
        """The tree's leaves in creation order: (ids, bodies, shapes, fat AABBs)."""
        n = ctypes.c_size_t()
        _check(lib().prx_world_get_proxies(self.h, None, None, None, None, 0, ctypes.byref(n)))
        k = n.value
        ids = np.zeros(k, dtype=np.uint32)
        bodies = np.zeros(k, dtype=np.uint32)
        shapes = np.zeros(k, dtype=np.uint32)
        aabbs = np.zeros((k, 4), dtype=np.float32)
        _check(lib().prx_world_get_proxies(self.h, ids.ctypes.data, bodies.ctypes.data, shapes.ctypes.data,
                                           aabbs.ctypes.data, k, ctypes.byref(n)))
        return ids, bodies, shapes, aabbs

    def set_transformation(self, body, xf):
        """SetTransformation (Body.cpp:270-273): xf = (p.x, p.y, q.x, q.y)."""
        arr = (ctypes.c_float * 4)(*xf)
        _check(lib().prx_world_set_transformation(self.h, body, arr))

    def set_location(self, body, v):
        """SetLocation (WorldBody.cpp:158-163)."""
        arr = (ctypes.c_float * 2)(*v)
        _check(lib().prx_world_set_location(self.h, body, arr))

    def set_angle(self, body, angle):
        """SetAngle (WorldBody.cpp:165-170)."""
        _check(lib().prx_world_set_angle(self.h, body, angle))

    def shift_origin(self, origin):
        """ShiftOrigin (AabbTreeWorld.cpp:2221-2242)."""
        arr = (ctypes.c_float * 2)(*origin)
        _check(lib().prx_world_shift_origin(self.h, arr))

    def set_friction(self, shape, value):
        """SetFriction(World&, ShapeID, NonNegative<Real>) (WorldShape.cpp:79-84)."""
        _check(lib().prx_shape_set_friction(self.h, shape, value))

    def set_restitution(self, shape, value):
        """SetRestitution(World&, ShapeID, Real) (WorldShape.cpp:91-96)."""
        _check(lib().prx_shape_set_restitution(self.h, shape, value))

    def state_record(self) -> bytes:
        n = ctypes.c_size_t()
        _check(lib().prx_world_state_record(self.h, None, 0, ctypes.byref(n)))
        buf = ctypes.create_string_buffer(n.value)
        _check(lib().prx_world_state_record(self.h, buf, n.value, ctypes.byref(n)))
        return buf.raw[:n.value]

    def save_desc(self) -> bytes:
        n = ctypes.c_size_t()
        _check(lib().prx_world_save_desc(self.h, None, 0, ctypes.byref(n)))
        buf = ctypes.create_string_buffer(n.value)
        _check(lib().prx_world_save_desc(self.h, buf, n.value, ctypes.byref(n)))
        return buf.raw[:n.value]

    def contact_count(self):
        return lib().prx_world_contact_count(self.h)

    def record_contact_events(self, on=True):
        """Records each step's contact listener calls (before batching the world)."""
        _check(lib().prx_world_record_contact_events(self.h, 1 if on else 0))

    def set_event_record_words(self, words):
        """Initial size of the per-step listener-call record in 32-bit words
        (0: the default); a step needing more grows it and is stepped again."""
        _check(lib().prx_world_set_event_record_words(self.h, int(words)))

    def contact_event_words(self):
        """Raw recorded listener calls of the last step: (uint32 words, calls)."""
        nw, nc = ctypes.c_size_t(), ctypes.c_uint32()
        _check(lib().prx_world_contact_events(self.h, None, 0, ctypes.byref(nw), ctypes.byref(nc)))
        out = np.zeros(max(1, nw.value), dtype=np.uint32)
        _check(lib().prx_world_contact_events(self.h, out.ctypes.data, out.size, ctypes.byref(nw),
                                              ctypes.byref(nc)))
        return out[:nw.value], nc.value

    def contact_events(self):
        """The last step's listener calls, parsed (see parse_contact_events)."""
        return parse_contact_events(*self.contact_event_words())

    def set_contact_listener(self, begin=None, end=None, pre_solve=None, post_solve=None):
        """SetBeginContactListener / SetEndContactListener /
        SetPreSolveContactListener / SetPostSolveContactListener (World.hpp:105-139):
        begin(id), end(id), pre_solve(id, Manifold), post_solve(id, ContactImpulses,
        solved_iteration), called through the C ABI after each step in the
        order the step made them.  All None clears the listeners."""
        if begin is None and end is None and pre_solve is None and post_solve is None:
            _check(lib().prx_world_set_contact_listener(self.h, None))
            self._listener = None
            return
        L = ContactListener()  # members left unset are null
        if begin:
            L.begin_contact = _BeginEndFn(lambda u, c: begin(c))
        if end:
            L.end_contact = _BeginEndFn(lambda u, c: end(c))
        if pre_solve:
            L.pre_solve = _PreSolveFn(lambda u, c, m: pre_solve(c, m.contents))
        if post_solve:
            L.post_solve = _PostSolveFn(lambda u, c, im, it: post_solve(c, im.contents, it))
        self._listener = L  # the callbacks live as long as the world uses them
        _check(lib().prx_world_set_contact_listener(self.h, ctypes.byref(L)))


class Batch:
    """Many worlds resident on the device, stepped by one launch per step."""

    def __init__(self, worlds, use_graph=True, order=0, coop_min=None, lds_bytes=None):
        self.worlds = list(worlds)
        conf = BatchConf()
        lib().prx_batch_conf_default(ctypes.byref(conf))
        conf.use_graph = 1 if use_graph else 0
        conf.order = order
        if coop_min is not None:
            conf.coop_min_contacts = int(coop_min)
        if lds_bytes is not None:
            conf.lds_bytes = int(lds_bytes)
        arr = (ctypes.c_void_p * len(self.worlds))(*[w.h.value for w in self.worlds])
        self.h = ctypes.c_void_p()
        _check(lib().prx_batch_create(arr, len(self.worlds), ctypes.byref(conf), ctypes.byref(self.h)))

    def __del__(self):
        h = getattr(self, "h", None)
        if h and h.value and _lib is not None:
            _lib.prx_batch_destroy(h)
            self.h = None

    def step(self, conf: StepConf = None, n=1):
        conf = conf or StepConf()
        _check(lib().prx_batch_step(self.h, ctypes.byref(conf), n))

    def sync(self):
        _check(lib().prx_batch_sync(self.h))

    def download(self):
        _check(lib().prx_batch_download(self.h))

    def last_timing(self):
        ms = ctypes.c_float()
        n = ctypes.c_int()
        _check(lib().prx_batch_last_timing(self.h, ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value

    def dynamic_bodies(self):
        return lib().prx_batch_dynamic_bodies(self.h)

    PHASES = ("proxies", "destroy", "update", "find_add", "adjacency", "islands", "constraints", "velocity",
              "position", "writeback", "sync", "update2", "find_add2", "total")

    def phase_cycles(self, enable=True):
        """Per-world phase cycle counters accumulated since the last call
        (array [worlds, 16]); enable turns collection on or off afterwards."""
        n = len(self.worlds)
        out = np.zeros((n, 16), dtype=np.uint64)
        _check(lib().prx_batch_phase_cycles(self.h, 1 if enable else 0, out.ctypes.data, n))
        return out

    def autotune(self, conf=None, nsteps=10):
        """Chooses the launch schedule (prx_batch_autotune) by timing nsteps
        steps per candidate from the current state, which is restored."""
        conf = conf or StepConf()
        best = BatchConf()
        _check(lib().prx_batch_autotune(self.h, ctypes.byref(conf), int(nsteps), ctypes.byref(best)))
        return {k: getattr(best, k) for k, _ in BatchConf._fields_}

    def counts(self):
        """(awake dynamic bodies, proxies, shape bytes, contacts, touching) after download()."""
        out = np.zeros(5, dtype=np.uint64)
        _check(lib().prx_batch_counts(self.h, out.ctypes.data))
        return tuple(int(x) for x in out)


def device_sincos(angles):
    """Device evaluation of UnitVec::Get(angle) for an array of float angles."""
    a = np.ascontiguousarray(angles, dtype=np.float32)
    out = np.zeros((a.size, 2), dtype=np.float32)
    code = lib().prx_selftest_sincos(a.ctypes.data, out.ctypes.data, a.size)
    if code:
        raise PrxError(code, "device sincos self-test failed")
    return out


class BigWorld:
    """One large world stepped on the whole device (graph-coloured mode):
    LBVH broadphase with the reference's pair sets, colour-parallel
    Gauss-Seidel solve.  The World's getters read the device state back."""

    def __init__(self, world):
        self.world = world
        self.h = ctypes.c_void_p()
        _check(lib().prx_big_create(world.h, ctypes.byref(self.h)))

    def __del__(self):
        h = getattr(self, "h", None)
        if h and h.value and _lib is not None:
            _lib.prx_big_destroy(h)
            self.h = None

    def step(self, conf=None, nsteps=1):
        conf = conf or StepConf()
        _check(lib().prx_big_step(self.h, ctypes.byref(conf), int(nsteps)))

    def sync(self):
        _check(lib().prx_big_sync(self.h))

    def last_timing(self):
        ms = ctypes.c_float()
        n = ctypes.c_int()
        _check(lib().prx_big_last_timing(self.h, ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value

    def metrics(self):
        """(min separation over touching contacts, touching contacts and
        colours of the last solve, contacts)."""
        out = np.zeros(4, dtype=np.float32)
        _check(lib().prx_big_metrics(self.h, out.ctypes.data))
        return dict(min_separation=float(out[0]), touching=int(out[1]), colors=int(out[2]), contacts=int(out[3]))


def lbvh_pairs(aabbs, bodies, moved, cap=None):
    """Broadphase pairs by the device LBVH: (k, 2) int32 array of (lower, higher)
    proxy indices of different bodies with overlapping AABBs, one of them moved."""
    aabbs = np.ascontiguousarray(aabbs, dtype=np.float32).reshape(-1, 4)
    n = aabbs.shape[0]
    bodies = np.ascontiguousarray(bodies, dtype=np.int32)
    moved = np.ascontiguousarray(moved, dtype=np.uint8)
    cap = cap if cap is not None else max(16 * n, 1024)
    out = np.zeros((cap, 2), dtype=np.int32)
    count = ctypes.c_int()
    _check(lib().prx_lbvh_pairs(aabbs.ctypes.data, bodies.ctypes.data, moved.ctypes.data, n,
                                out.ctypes.data, cap, ctypes.byref(count)))
    return out[:min(count.value, cap)].copy()


class Adjoint:
    """Gradients of sum over dynamic bodies of |x_K - target|^2 after K steps
    with respect to the bodies' initial velocities and the shapes' friction and
    restitution (prx_adjoint_*).  Arrays are in batch order (worlds in the
    order given to the Batch, bodies and shapes within each)."""

    def __init__(self, batch, nsteps, every=16, conf=None):
        self.batch = batch
        self.conf = conf or StepConf(do_toi=0)  # the regular step is differentiated
        self.h = ctypes.c_void_p()
        _check(lib().prx_adjoint_create(batch.h, int(nsteps), int(every), ctypes.byref(self.conf),
                                        ctypes.byref(self.h)))
        self.nbodies = sum(w.body_count for w in batch.worlds)
        self.nshapes = sum(lib().prx_world_shape_count(w.h) for w in batch.worlds)
        self.njoints = sum(w.joint_count for w in batch.worlds)

    def __del__(self):
        h = getattr(self, "h", None)
        if h and h.value and _lib is not None:
            _lib.prx_adjoint_destroy(h)
            self.h = None

    def run(self, targets):
        t = np.ascontiguousarray(targets, dtype=np.float32).reshape(self.nbodies, 2)
        loss = np.zeros(len(self.batch.worlds), dtype=np.float32)
        gv = np.zeros((self.nbodies, 3), dtype=np.float32)
        gf = np.zeros(self.nshapes, dtype=np.float32)
        ge = np.zeros(self.nshapes, dtype=np.float32)
        _check(lib().prx_adjoint_run(self.h, ctypes.byref(self.conf), t.ctypes.data, loss.ctypes.data,
                                     gv.ctypes.data, gf.ctypes.data, ge.ctypes.data))
        out = dict(loss=loss, grad_velocity=gv, grad_friction=gf, grad_restitution=ge)
        if self.njoints:
            gt = np.zeros(self.njoints, dtype=np.float32)
            _check(lib().prx_adjoint_joint_gradients(self.h, gt.ctypes.data))
            out["grad_motor_torque"] = gt
        return out

    def forward(self):
        """Steps the batch K steps with checkpoints; returns the final state,
        (bodies, 6): x, y, angle, vx, vy, w (prx_adjoint_forward)."""
        out = np.zeros((self.nbodies, 6), dtype=np.float32)
        _check(lib().prx_adjoint_forward(self.h, ctypes.byref(self.conf), out.ctypes.data))
        return out

    def backward(self, state_bar):
        """Gradients of a loss whose gradient with respect to the final state
        of the last forward() is state_bar (bodies, 6) (prx_adjoint_backward)."""
        sb = np.ascontiguousarray(state_bar, dtype=np.float32).reshape(self.nbodies, 6)
        gv = np.zeros((self.nbodies, 3), dtype=np.float32)
        gf = np.zeros(self.nshapes, dtype=np.float32)
        ge = np.zeros(self.nshapes, dtype=np.float32)
        _check(lib().prx_adjoint_backward(self.h, ctypes.byref(self.conf), sb.ctypes.data, gv.ctypes.data,
                                          gf.ctypes.data, ge.ctypes.data))
        out = dict(grad_velocity=gv, grad_friction=gf, grad_restitution=ge)
        if self.njoints:
            gt = np.zeros(self.njoints, dtype=np.float32)
            _check(lib().prx_adjoint_joint_gradients(self.h, gt.ctypes.data))
            out["grad_motor_torque"] = gt
        return out

    def use_general(self, on=True):
        """Diagnostic: reverse a disk-only batch with the general kernel."""
        _check(lib().prx_adjoint_use_general(self.h, 1 if on else 0))

    def diagnostics(self):
        out = np.zeros(4, dtype=np.float32)
        _check(lib().prx_adjoint_diagnostics(self.h, out.ctypes.data))
        return dict(slots=int(out[0]), biased=int(out[1]), position_updates=int(out[2]),
                    velocity_updates=int(out[3]))

    def timing_ms(self):
        ms = ctypes.c_float()
        _check(lib().prx_adjoint_timing(self.h, ctypes.byref(ms)))
        return ms.value


def desc_offsets(desc):
    """Byte offsets of each shape record and each body record in a world
    description (layout of prx_world_save_desc)."""
    import struct
    u32 = lambda o: struct.unpack_from("<I", desc, o)[0]
    ns = u32(0)
    off = 4
    shapes = []
    for _ in range(ns):
        n = u32(off)
        shapes.append(off)
        off += 4 + 16 + 12 + 16 * n
    nb = u32(off)
    off += 4
    bodies = []
    for _ in range(nb):
        bodies.append(off)
        nsh = u32(off + 8 + 76)
        off += 8 + 76 + 4 + 4 * nsh
    return shapes, bodies


def desc_joint_offsets(desc):
    """Byte offsets of each joint record in a world description."""
    import struct
    shapes, bodies = desc_offsets(desc)
    u32 = lambda o: struct.unpack_from("<I", desc, o)[0]
    if bodies:
        off = bodies[-1] + 8 + 76 + 4 + 4 * u32(bodies[-1] + 8 + 76)
    else:
        off = shapes[-1] + 4 + 16 + 12 + 16 * u32(shapes[-1]) + 4 if shapes else 8
    nj = u32(off)
    return [off + 4 + 60 * k for k in range(nj)]


def desc_with(desc, velocities=None, friction=None, restitution=None, motor_torque=None):
    """A copy of a world description with body velocities (dict body ->
    (vx, vy, w)), shape friction / restitution (dict shape -> value) and
    joint maximum motor torques (dict joint -> value) replaced."""
    import struct
    buf = bytearray(desc)
    shapes, bodies = desc_offsets(desc)
    for j, v in (motor_torque or {}).items():
        struct.pack_into("<f", buf, desc_joint_offsets(desc)[j] + 44, float(v))
    for b, v in (velocities or {}).items():
        struct.pack_into("<3f", buf, bodies[b] + 8 + 32, *[float(x) for x in v])
    for sidx, f in (friction or {}).items():
        struct.pack_into("<f", buf, shapes[sidx] + 12, float(f))
    for sidx, e in (restitution or {}).items():
        struct.pack_into("<f", buf, shapes[sidx] + 16, float(e))
    return bytes(buf)
