"""Single-step error against the FP64 oracle from the same input state.

The FP64 oracle is the oracle's double instantiation (oracle/cpu_ref.hpp)
started from a float world's whole state by exact widening (oracle_widen).
For one step from the same state, the error of a float result (the device's,
bit-identical to the float reference) is reported per body and component:

    position  |x - x64| / max(|x64|, 1 m)      (x and y of the body origin)
    angle     |a - a64| / max(|a64|, 1 rad)
    velocity  |v - v64| / max(|v64|, 1 m/s)    (vx and vy)
    spin      |w - w64| / max(|w64|, 1 rad/s)

each the maximum over dynamic bodies, with the absolute maxima beside them.

Bounds asserted by the tests (stated here once): positions within 1e-5,
angles within 1e-3, velocities and spins within 1e-2.  Positions differ by
rounding only (at most 6.5e-6 over 300 steps of the pyramid, a 497-body pile
and a chain).  Angles and velocities also carry the solver's discrete choices
that the two precisions can make differently from the same state: the
position solver's simultaneous-or-deepest-first rule for two points
(AlmostEqual, ContactSolver.cpp:598) and the block solver's case order
(ContactSolver.cpp:113-243).  Measured maxima over the same 300 steps: angle
5.3e-4 rad (pyramid), spin 1.6e-3 rad/s (pile); medians 1.4e-6 and 3e-5.
"""
import numpy as np

POS_BOUND = 1e-5
ANGLE_BOUND = 1e-3
VEL_BOUND = 1e-2


def errors(got, want, dynamic):
    """got, want: (n, 6) arrays of (x, y, angle, vx, vy, w); dynamic: mask."""
    g = np.asarray(got, dtype=np.float64)[dynamic]
    w = np.asarray(want, dtype=np.float64)[dynamic]
    if g.size == 0:
        return dict(position=0.0, angle=0.0, velocity=0.0, spin=0.0, abs_position=0.0, abs_velocity=0.0)
    d = np.abs(g - w)
    scale = np.maximum(np.abs(w), 1.0)
    r = d / scale
    return dict(position=float(r[:, :2].max()), angle=float(r[:, 2].max()), velocity=float(r[:, 3:5].max()),
                spin=float(r[:, 5].max()), abs_position=float(d[:, :2].max()), abs_angle=float(d[:, 2].max()),
                abs_velocity=float(d[:, 3:5].max()), abs_spin=float(d[:, 5].max()))


def within_bounds(e):
    return e["position"] <= POS_BOUND and e["angle"] <= ANGLE_BOUND and max(e["velocity"], e["spin"]) <= VEL_BOUND


def state_bodies(state):
    """(n, 6) float64 (x, y, angle, vx, vy, w) of a parsed state record."""
    b = state["bodies"]
    return np.concatenate([b[:, 3:6], b[:, 6:9]], axis=1).astype(np.float64)


def dynamic_mask(state):
    return (state["bflags"] & 2) != 0  # accelerable
