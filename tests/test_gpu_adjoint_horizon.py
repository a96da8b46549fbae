"""The adjoint at the configured horizons (120 billiards steps, 240 chain
steps) against the forward-mode derivative of the same float step.

Finite differences stop being derivatives at these horizons: the worlds are
chaotic (perturbations grow by orders of magnitude), so within any
perturbation large enough to rise above float rounding some contact starts,
ends or switches a clamp, and float and double parted long before.  The
float dual world (oracle/tangent.hpp) carries the derivative along the float
step's own discrete path, which the device reproduces bit for bit, so it
differentiates exactly the function the adjoint reverses, at any horizon.
Each checked parameter is one forward-mode run; its directional derivative
of the loss is compared with the adjoint's gradient entry."""
import numpy as np
import pytest

import playrho_amd as pa
from tests import dumpfmt, oracle_lib
from tests.test_gpu_adjoint import _setup
from tests.test_gpu_adjoint_general import _chain

pytestmark = pytest.mark.gpu


def _tangent_check(desc, conf, steps, targets, dyn, out, params, bound):
    dc = desc + dumpfmt.conf_record(conf)
    nb = len(dyn)
    got, want = [], []
    x_final = None
    for name, kind, idx, comp, g in params:
        w = oracle_lib.OracleWorld(dc, oracle_lib.TANGENT)
        w.seed(kind, idx, comp, 1.0)
        w.step(steps, record=False)
        x = w.bodies_f64(nb)
        if x_final is None:
            x_final = x
        assert np.array_equal(x, x_final)
        d = w.tangent_f64(nb)
        r = (x[:, :2] - targets) * dyn[:, None]
        tan = float(np.sum(2.0 * r * d[:, :2]))
        print("%-6s adjoint %+.7e  tangent %+.7e" % (name, g, tan), flush=True)
        got.append(float(g))
        want.append(tan)
    got, want = np.array(got), np.array(want)
    rel = np.linalg.norm(got - want) / np.linalg.norm(want)
    print("relative error over %d parameters %.3e (gradient norm over them %.3e)" % (len(got), rel,
                                                                                   np.linalg.norm(want)), flush=True)
    assert rel < bound
    return x_final


def test_billiards_120_steps_against_forward_mode():
    conf = pa.StepConf(do_toi=0)
    steps, every = 120, 16
    w, desc, bodies, dyn, targets = _setup(1, 0, steps)
    adj = pa.Adjoint(pa.Batch([w], use_graph=False), steps, every, conf)
    print("billiards world 0, %d steps, checkpoints every %d" % (steps, every), flush=True)
    out = adj.run(targets.astype(np.float32))
    shapes, _ = pa.desc_offsets(desc)
    params = []
    for b in (1, 9, 30, 63):
        for k in (0, 1):
            params.append(("v%d%s" % (b, "xy"[k]), "velocity", b, k, out["grad_velocity"][b, k]))
    for s in (4, 12, 40, 66):
        params.append(("f%d" % s, "friction", s, 0, out["grad_friction"][s]))
        params.append(("e%d" % s, "restitution", s, 0, out["grad_restitution"][s]))
    x = _tangent_check(desc, conf, steps, targets, dyn, out, params, 1e-3)
    # The device's loss is the float path's (the forward-mode world's values).
    r = (x[:, :2] - targets) * dyn[:, None]
    assert abs(float(out["loss"][0]) - float(np.sum(r * r))) <= 1e-5 * float(np.sum(r * r))


@pytest.mark.parametrize("params_,steps", [([1, 0], 240)])
def test_chain_240_steps_against_forward_mode(params_, steps):
    conf = pa.StepConf(do_toi=0)
    w, desc, bodies, joints, dyn, targets = _chain(params_)
    adj = pa.Adjoint(pa.Batch([w], use_graph=False), steps, 16, conf)
    print("chain world", params_, "%d steps, checkpoints every 16" % steps, flush=True)
    out = adj.run(targets.astype(np.float32))
    params = []
    for j in (0, 3, 9, 21):
        params.append(("T%d" % j, "motor_torque", j, 0, out["grad_motor_torque"][j]))
    for b in (3, 17, 30, 40, 70):
        for k in range(3):
            params.append(("v%d%s" % (b, "xyw"[k]), "velocity", b, k, out["grad_velocity"][b, k]))
    sdesc, _, _ = dumpfmt.read_description(dumpfmt.Reader(desc))
    for s in (1, len(sdesc) - 1):
        params.append(("f%d" % s, "friction", s, 0, out["grad_friction"][s]))
    _tangent_check(desc, conf, steps, targets, dyn, out, params, 1e-3)
