"""ctypes wrapper over the oracle's C-ABI (oracle/cpu_ref.cpp) for tests and
the CPU-baseline leg of bench.py.  The oracle is the checker, never the thing
measured on the GPU path."""
import ctypes
import os
import subprocess

import numpy as np

from . import dumpfmt  # noqa: F401  (re-exported for callers)

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "oracle", "_build", "liboracle.so")
REF_DRIVER = os.path.join(ROOT, "oracle", "_ref", "ref_driver")

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError("oracle library not built: run __graft_entry__.build()")
        L = ctypes.CDLL(LIB)
        L.oracle_create.restype = ctypes.c_void_p
        L.oracle_create.argtypes = [ctypes.c_char_p, ctypes.c_size_t, ctypes.c_int]
        L.oracle_destroy.argtypes = [ctypes.c_void_p]
        L.oracle_widen.restype = ctypes.c_void_p
        L.oracle_widen.argtypes = [ctypes.c_void_p]
        L.oracle_step.restype = ctypes.c_void_p
        L.oracle_step.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                  ctypes.POINTER(ctypes.c_size_t)]
        L.oracle_state.restype = ctypes.c_void_p
        L.oracle_state.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_size_t)]
        L.oracle_bodies_f64.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)]
        L.oracle_free.argtypes = [ctypes.c_void_p]
        L.oracle_sincosf.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]
        L.oracle_edit.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_uint32, ctypes.c_float,
                                  ctypes.c_float, ctypes.c_float]
        L.oracle_tangent_seed.restype = ctypes.c_int
        L.oracle_tangent_seed.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_uint32, ctypes.c_int,
                                          ctypes.c_double]
        L.oracle_tangent_bodies.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)]
        L.oracle_time_batch.restype = ctypes.c_double
        L.oracle_time_batch.argtypes = [ctypes.c_char_p, ctypes.c_size_t, ctypes.c_int,
                                        ctypes.c_int, ctypes.c_int,
                                        ctypes.POINTER(ctypes.c_double)]
        _lib = L
    return _lib


def _take(ptr, n):
    buf = ctypes.string_at(ptr, n.value)
    lib().oracle_free(ptr)
    return buf


class OracleWorld:
    """One oracle world built from a description+conf blob."""

    def __init__(self, desc_conf: bytes, precision=32, handle=None):
        self.h = handle if handle is not None else lib().oracle_create(desc_conf, len(desc_conf), precision)
        self.precision = precision
        self.nbodies = None

    def widen(self):
        """The FP64 oracle stepping on from this float world's current state
        (every real widened exactly; cached transformations recomputed in
        double)."""
        assert self.precision == 32
        return OracleWorld(b"", 64, handle=lib().oracle_widen(self.h))

    def __del__(self):
        if getattr(self, "h", None):
            lib().oracle_destroy(self.h)
            self.h = None

    def edit(self, kind, body, a=0.0, b=0.0, c=0.0):
        """A world edit between steps: 'L' SetLocation, 'A' SetAngle, 'T'
        SetTransformation (c: the angle of the rotation), 'O' ShiftOrigin."""
        lib().oracle_edit(self.h, ord(kind), body, a, b, c)

    def step_with_edits(self, steps, edits, first=1):
        """Steps `steps` steps (numbered from `first`), applying each edit
        (step, kind, body, a, b, c) before its step; returns the states."""
        out = []
        for s in range(first, first + steps):
            for e in edits:
                if int(e[0]) == s:
                    self.edit(chr(int(e[1])), int(e[2]), float(e[3]), float(e[4]), float(e[5]))
            out += self.step(1)
        return out

    def step(self, n=1, record=True):
        ln = ctypes.c_size_t()
        p = lib().oracle_step(self.h, n, 1 if record else 0, ctypes.byref(ln))
        buf = _take(p, ln)
        if not record:
            return []
        r = dumpfmt.Reader(buf)
        out = []
        while r.o < len(buf):
            out.append(dumpfmt.read_state(r))
        return out

    def state(self):
        ln = ctypes.c_size_t()
        p = lib().oracle_state(self.h, ctypes.byref(ln))
        buf = _take(p, ln)
        return dumpfmt.read_state(dumpfmt.Reader(buf))

    def bodies_f64(self, nbodies):
        out = np.zeros(nbodies * 6, dtype=np.float64)
        lib().oracle_bodies_f64(self.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        return out.reshape(nbodies, 6)

    # Forward mode (precision TANGENT: the float step as executed, TANGENT64:
    # the double step), with the derivative of every value along its path in
    # double.
    SEED_KINDS = {"velocity": 0, "friction": 1, "restitution": 2, "motor_torque": 3}

    def seed(self, kind, index, comp=0, value=1.0):
        """Sets d(parameter) for one parameter: kind 'velocity' (comp 0 vx,
        1 vy, 2 w), 'friction', 'restitution' or 'motor_torque'."""
        assert self.precision in (TANGENT, TANGENT64, TANGENT_F)
        rc = lib().oracle_tangent_seed(self.h, self.SEED_KINDS[kind], int(index), int(comp), float(value))
        if rc != 0:
            raise IndexError("bad tangent seed %s %d %d" % (kind, index, comp))

    def tangent_f64(self, nbodies):
        """d(x, y, angle, vx, vy, w) of every body, (nbodies, 6)."""
        assert self.precision in (TANGENT, TANGENT64, TANGENT_F)
        out = np.zeros(nbodies * 6, dtype=np.float64)
        lib().oracle_tangent_bodies(self.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        return out.reshape(nbodies, 6)


TANGENT = 33  # oracle_create precision codes of the forward-mode worlds
TANGENT64 = 65
TANGENT_F = 34  # float values and float derivatives


def split_dump(buf):
    """Returns (desc_conf_bytes, parsed_dump) for a ref_driver dump."""
    r = dumpfmt.Reader(buf)
    r.o = 8
    dumpfmt.read_description(r)
    dumpfmt.read_conf(r)
    desc_conf = bytes(buf[8:r.o])
    return desc_conf, dumpfmt.read_dump(buf)


def run_ref_driver(scene, steps, do_toi, params, path, edits=()):
    """Runs the reference build's driver (present only where oracle/_ref was
    built) and returns the raw dump; edits (step, kind, body, a, b, c) are
    applied before their steps."""
    cmd = [REF_DRIVER, scene, str(steps), "1" if do_toi else "0", path] + [str(p) for p in params]
    env = dict(os.environ)
    if edits:
        env["PRX_EDITS"] = ";".join("%d %s %d %r %r %r" % (s, k, b, a, bb, c) for s, k, b, a, bb, c in edits)
    subprocess.run(cmd, check=True, env=env)
    with open(path, "rb") as f:
        return f.read()


def time_batch(descs, steps, threads):
    """Times the oracle on independent worlds with `threads` threads.
    Returns (seconds, body_steps)."""
    packed = b"".join(len(d).to_bytes(8, "little") + d for d in descs)
    bs = ctypes.c_double()
    secs = lib().oracle_time_batch(packed, len(packed), len(descs), steps, threads, ctypes.byref(bs))
    return secs, bs.value


def host_sincosf(angles):
    """(cos, sin) pairs from the host C library's cosf/sinf."""
    a = np.ascontiguousarray(angles, dtype=np.float32)
    out = np.zeros((a.size, 2), dtype=np.float32)
    lib().oracle_sincosf(a.ctypes.data, out.ctypes.data, a.size)
    return out
