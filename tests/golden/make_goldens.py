"""Writes the committed golden fixtures from the reference float build.

Run from the repository root, in a container where the reference checkout is
present and oracle/_ref has been built (make -f oracle/Makefile.ref):

    python tests/golden/make_goldens.py

For each case it runs oracle/_ref/ref_driver (our driver linked against the
reference library), and stores in tests/golden/<case>.npz:
  desc_conf  the world description and StepConf as the reference built them
  digests    per step, SHA-256 of bodies / contact list / proxies / joints /
             contact listener calls / StepStats
  full_*     complete state at a few steps (for diffs when a digest differs)
HelloWorld/ExpectedOutput.txt (the reference's own golden output) is copied
as data to tests/golden/hello_expected.txt.
"""
import os
import shutil
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from tests import dumpfmt, golden_cases, oracle_lib  # noqa: E402

CASES = [
    # name, scene, steps, params, full-state steps
    ("hello", "hello", 60, [], [1, 30, 60]),
    ("pyramid", "pyramid", 300, [20], [1, 2, 50, 150, 300]),
    ("drop300", "drop", 200, [300, 20, 40, 7], [1, 100, 200]),
    ("billiards0", "billiards", 120, [1, 0], [1, 60, 120]),
    ("billiards1", "billiards", 120, [1, 1], [1, 120]),
] + [("pile%d" % w, "pile", 150, [1, w], [1, 150]) for w in range(8)] + [
    ("chain0", "chain", 240, [1, 0], [1, 120, 240]),
    ("chain1", "chain", 240, [1, 1, 12, 24, 40], [1, 240]),
    # Distance and weld joints, rigid and soft, friction joints on the ground,
    # ropes from a bar, motor joints, a cart on wheel joints and a pulley, beside
    # falling bodies.
    ("joints0", "joints", 240, [1, 0], [1, 120, 240]),
    ("joints1", "joints", 240, [2, 1, 12, 40], [1, 240]),
    # The two largest worlds of the 4096-world pile batch (seed 1): 513 and
    # 511 bodies, through piling and settling (their settled islands take the
    # level form of the wave-wide sweeps on the device).
    ("pile3301", "pile", 300, [1, 3301], [1, 300]),
    ("pile859", "pile", 300, [1, 859], [1, 300]),
    # Prismatic joints: limits, motors, both, equal limits, a carriage with a
    # motorised slider between two dynamic bodies, beside falling bodies.
    ("sliders0", "sliders", 240, [1, 0], [1, 120, 240]),
    ("sliders1", "sliders", 240, [2, 1, 9, 40], [1, 240]),
    # Target joints (no body A): clamped and unclamped forces, frequency 0,
    # maximum force 0, beside falling bodies.
    ("targets0", "targets", 240, [1, 0], [1, 120, 240]),
    # Chain shapes (a zig-zag ground line, a closed square outline) and bodies
    # of multi shapes (two or three box hulls), beside falling bodies.
    ("terrain0", "terrain", 200, [1, 0], [1, 100, 200]),
    ("terrain1", "terrain", 200, [2, 1, 60], [1, 200]),
    # Gear joints: revolute-revolute (a motorised wheel train), revolute-
    # prismatic (a rack), prismatic-prismatic (one rail limited) and a pinion
    # geared to its own carriage's rail (bodies B and C the same), beside
    # falling bodies.
    ("gears0", "gears", 240, [1, 0], [1, 120, 240]),
    ("gears1", "gears", 240, [2, 1, 40], [1, 240]),
]
# World edits between steps (step, kind, body, a, b, c): SetLocation 'L',
# SetAngle 'A', SetTransformation 'T' (c the rotation's angle), ShiftOrigin 'O',
# SetVelocity 'V', Destroy(JointID) 'J', Destroy(BodyID) 'D', SetType 'S'.
EDIT_CASES = [
    ("pile3_edit", "pile", 120, [1, 3], [1, 120], False,
     [(20, "L", 5, 1.25, 6.5, 0), (35, "A", 9, 0.7, 0, 0), (50, "T", 12, -1.5, 8.0, 2.5), (70, "O", 0, 3.0, -0.5, 0),
      (90, "L", 0, 0.25, 0.0, 0)]),
    ("drop300_edit_toi", "drop", 120, [300, 20, 40, 7], [1, 120], True,
     [(15, "T", 40, 0.0, 12.0, -1.0), (30, "O", 0, -2.0, 1.5, 0), (45, "A", 77, 3.0, 0, 0), (60, "L", 120, 5.0, 9.0, 0)]),
    # ShiftOrigin with every joint kind present (the pulley's ground anchors move).
    ("joints0_edit", "joints", 120, [1, 0], [1, 120], False,
     [(40, "O", 0, 3.0, -2.0, 0), (80, "L", 30, -5.0, 20.0, 0)]),
    # SetVelocity 'V' (c the angular velocity): resting and falling bodies,
    # the static ground (ignored), a zero velocity (no wake-up).
    ("pile3_vel_edit", "pile", 120, [1, 3], [1, 120], False,
     [(30, "V", 7, 2.0, 3.0, 1.5), (55, "V", 0, 1.0, 1.0, 0), (80, "V", 12, 0, 0, 0), (100, "V", 20, -4.0, 6.0, -3.0),
      (100, "V", 21, 0.5, 0, 0)]),
    # Destroy(JointID) 'J' and Destroy(BodyID) 'D' (body: the id): a weld, the
    # pulley, a body held by friction and weld joints and resting on the
    # ground, a wheel of the cart, a falling body among others.
    # Gear joints under edits: a wheel's spin set (the gear train carries it),
    # a wheel-wheel gear and the pinion's gear destroyed, then a slab that is
    # body B of a prismatic-prismatic gear destroyed with its joints.
    ("gears0_edit", "gears", 200, [1, 0], [1, 200], False,
     [(50, "V", 2, 0.0, 0.0, 4.0), (80, "J", 4, 0, 0, 0), (110, "J", 12, 0, 0, 0), (140, "D", 6, 0, 0, 0)]),
    ("joints0_destroy_edit", "joints", 160, [1, 0], [1, 160], False,
     [(40, "J", 3, 0, 0, 0), (60, "J", 18, 0, 0, 0), (90, "D", 24, 0, 0, 0), (110, "D", 19, 0, 0, 0),
      (130, "D", 50, 0, 0, 0), (130, "D", 45, 0, 0, 0)]),
    # SetType 'S' (a: 0 static, 1 kinematic, 2 dynamic): a resting body made
    # static and dynamic again, one made kinematic (then given a velocity)
    # and dynamic again; the mass data reset each time.
    ("pile3_type_edit", "pile", 120, [1, 3], [1, 120], False,
     [(40, "S", 30, 0, 0, 0), (60, "S", 50, 1, 0, 0), (80, "S", 30, 2, 0, 0), (90, "V", 50, 2.0, 0, 1.0),
      (100, "S", 50, 2, 0, 0)]),
    ("pile3_type_edit_toi", "pile", 120, [1, 3], [1, 120], True,
     [(40, "S", 30, 0, 0, 0), (60, "S", 50, 1, 0, 0), (80, "S", 30, 2, 0, 0), (90, "V", 50, 2.0, 0, 1.0),
      (100, "S", 50, 2, 0, 0)]),
    # Bodies destroyed out of a resting pile (their touching contacts end,
    # their neighbours wake), with the TOI phase.
    ("pile3_destroy_edit_toi", "pile", 120, [1, 3], [1, 120], True,
     [(60, "D", 40, 0, 0, 0), (61, "D", 41, 0, 0, 0), (61, "D", 90, 0, 0, 0), (90, "V", 60, 0, 8.0, 0),
      (90, "D", 100, 0, 0, 0)]),
]
# The same scenes with the TOI phase (the reference's default doToi = true).
TOI_CASES = [
    ("pyramid_toi", "pyramid", 300, [20], [1, 150, 300]),
    ("drop300_toi", "drop", 200, [300, 20, 40, 7], [1, 100, 200]),
    ("billiards0_toi", "billiards", 120, [1, 0], [1, 60, 120]),
    ("pile3_toi", "pile", 150, [1, 3], [1, 150]),
    ("chain0_toi", "chain", 240, [1, 0], [1, 240]),
    ("joints0_toi", "joints", 240, [1, 0], [1, 240]),
    ("pile3301_toi", "pile", 300, [1, 3301], [1, 300]),
    ("pile859_toi", "pile", 300, [1, 859], [1, 300]),
    ("sliders0_toi", "sliders", 240, [1, 0], [1, 240]),
    ("targets0_toi", "targets", 240, [1, 0], [1, 240]),
    ("terrain0_toi", "terrain", 200, [1, 0], [1, 200]),
    ("gears0_toi", "gears", 240, [1, 0], [1, 240]),
]


def main(only=()):
    """Writes every case, or only the named ones."""
    tmp = tempfile.mkdtemp()
    try:
        for name, scene, steps, params, full, do_toi, edits in ([c + (False, []) for c in CASES] +
                                                                [c + (True, []) for c in TOI_CASES] + EDIT_CASES):
            if only and name not in only:
                continue
            path = os.path.join(tmp, name + ".bin")
            buf = oracle_lib.run_ref_driver(scene, steps, do_toi, params, path, edits)
            desc_conf, d = oracle_lib.split_dump(buf)
            states = d["states"]
            dig = np.zeros((len(states), len(golden_cases.PARTS), 32), dtype=np.uint8)
            for i, s in enumerate(states):
                for k, part in enumerate(golden_cases.PARTS):
                    dig[i, k] = np.frombuffer(dumpfmt.digest(s, (part,)), dtype=np.uint8)
            arrays = dict(desc_conf=np.frombuffer(desc_conf, dtype=np.uint8),
                          digests=dig, scene=np.array(scene), params=np.array(params, dtype=np.int32),
                          steps=np.array(steps), full_steps=np.array(full, dtype=np.int32),
                          do_toi=np.array(do_toi),
                          edits=np.array([(s, ord(k), b, a, bb, c) for s, k, b, a, bb, c in edits],
                                         dtype=np.float64).reshape(-1, 6))
            for st in full:
                s = states[st]
                for key in ("stats", "bodies", "bflags", "cu", "ct", "cf", "pid", "paabb", "jlimit", "jimp", "events"):
                    arrays["full%d_%s" % (st, key)] = s[key]
            out = os.path.join(HERE, name + ".npz")
            np.savez_compressed(out, **arrays)
            print("%-12s %4d steps  %6d bytes" % (name, steps, os.path.getsize(out)))
    finally:
        shutil.rmtree(tmp)
    ref_hello = "/root/reference/HelloWorld/ExpectedOutput.txt"
    if os.path.exists(ref_hello):
        shutil.copyfile(ref_hello, os.path.join(HERE, "hello_expected.txt"))


if __name__ == "__main__":
    main(tuple(sys.argv[1:]))
