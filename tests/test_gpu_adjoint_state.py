"""The adjoint split into a forward pass that returns the final state and a
backward pass seeded with any gradient of a loss with respect to that state
(prx_adjoint_forward / prx_adjoint_backward, torch.ops.playrho_amd.simulate).

The final state is checked bit for bit against the float32 restatement; the
gradients of linear losses sum c * state (every component: x, y, angle, vx,
vy, w) against central differences of the FP64 restatement, on the billiards
world (hand-derived disk kernel) and on an articulated chain world (general
kernel); the split against the fused prx_adjoint_run; the torch operator
against the C-ABI."""
import numpy as np
import pytest

import playrho_amd as pa
from tests import dumpfmt, oracle_lib
from tests.test_gpu_adjoint import _setup
from tests.test_gpu_adjoint_general import _chain

pytestmark = pytest.mark.gpu


def _f32(x):
    return float(np.float32(x))


def _state(desc, conf, steps, precision):
    ow = oracle_lib.OracleWorld(desc + dumpfmt.conf_record(conf), precision=precision)
    ow.step(steps, record=False)
    nb = pa.World.from_desc(desc).body_count
    return ow.bodies_f64(nb)


def _seed(nbodies, dyn, rng):
    c = rng.normal(size=(nbodies, 6)).astype(np.float32)
    c[~dyn] = 0.0
    return c


def _fd_check(out, desc, conf, steps, c, checks):
    def loss64(d):
        return float(np.sum(_state(d, conf, steps, 64) * c.astype(np.float64)))

    def central(make, x0, eps):
        xp, xm = _f32(x0 + eps), _f32(x0 - eps)
        return (loss64(make(xp)) - loss64(make(xm))) / (xp - xm)

    got, want, skipped = [], [], []
    for name, g, make, x0, eps in checks:
        fd, fd2 = central(make, x0, eps), central(make, x0, eps / 10)
        ok = abs(fd - fd2) <= 1e-4 * max(abs(fd), 1.0)
        print("%-6s adjoint %+.6e  fd %+.6e  fd/10 %+.6e%s" % (name, g, fd, fd2, "" if ok else "  (event)"),
              flush=True)
        if not ok:
            skipped.append(name)
            continue
        got.append(float(g))
        want.append(fd)
    got, want = np.array(got), np.array(want)
    rel = np.linalg.norm(got - want) / np.linalg.norm(want)
    print("relative error over %d parameters %.3e, gradient norm %.3e (left out: %s)"
          % (len(got), rel, np.linalg.norm(want), skipped), flush=True)
    assert len(skipped) <= len(checks) // 4
    assert rel < 1e-3


def _velocity_checks(out, desc, bodies, bsel, comps, eps):
    checks = []
    for b in bsel:
        v = bodies[b]["vals"][8:11].astype(np.float64)
        for k in comps:
            def make(x, b=b, k=k, v=v):
                vv = v.copy()
                vv[k] = x
                return pa.desc_with(desc, velocities={b: vv})
            checks.append(("v%d%s" % (b, "xyw"[k]), out["grad_velocity"][b, k], make, v[k], eps))
    return checks


@pytest.mark.parametrize("steps,every", [(12, 4), (30, 16)])
def test_final_state_and_linear_loss_billiards(steps, every):
    conf = pa.StepConf(do_toi=0)
    w, desc, bodies, dyn, _ = _setup(1, 0, steps)
    adj = pa.Adjoint(pa.Batch([w], use_graph=False), steps, every, conf)
    print("billiards steps", steps, "every", every, flush=True)
    state = adj.forward()
    ref = _state(desc, conf, steps, 32)
    assert np.array_equal(state.astype(np.float64), ref)
    c = _seed(len(bodies), dyn, np.random.default_rng(steps))
    out = adj.backward(c)
    checks = _velocity_checks(out, desc, bodies, (1, 9, 30, 63), (0, 1), 1e-3)
    sdesc, _, _ = dumpfmt.read_description(dumpfmt.Reader(desc))
    for s in (4, 12, 40, 66):
        fr = float(sdesc[s]["friction"])
        re = float(sdesc[s]["restitution"])
        checks.append(("f%d" % s, out["grad_friction"][s], lambda x, s=s: pa.desc_with(desc, friction={s: x}), fr,
                       1e-3))
        checks.append(("e%d" % s, out["grad_restitution"][s],
                       lambda x, s=s: pa.desc_with(desc, restitution={s: x}), re, 1e-3))
    _fd_check(out, desc, conf, steps, c, checks)


@pytest.mark.parametrize("params,steps,bsel", [([1, 0, 6, 4, 2], 30, (2, 4, 7, 9)), ([1, 0], 30, (3, 17, 40, 70))])
def test_final_state_and_linear_loss_chain(params, steps, bsel):
    conf = pa.StepConf(do_toi=0)
    w, desc, bodies, joints, dyn, _ = _chain(params)
    adj = pa.Adjoint(pa.Batch([w], use_graph=False), steps, 8, conf)
    print("chain", params, "steps", steps, flush=True)
    state = adj.forward()
    ref = _state(desc, conf, steps, 32)
    assert np.array_equal(state.astype(np.float64), ref)
    c = _seed(len(bodies), dyn, np.random.default_rng(len(bodies)))
    out = adj.backward(c)
    checks = _velocity_checks(out, desc, bodies, bsel, (0, 1, 2), 1e-4)
    for j in range(min(len(joints), 4)):
        t0 = float(joints[j]["vals"][8])
        checks.append(("T%d" % j, out["grad_motor_torque"][j],
                       lambda x, j=j: pa.desc_with(desc, motor_torque={j: x}), t0, 1e-3))
    _fd_check(out, desc, conf, steps, c, checks)


def test_split_equals_fused_run():
    """forward + backward seeded with 2 (x - target) on x and y gives the
    fused prx_adjoint_run's gradients bit for bit, and a backward without a
    forward is refused."""
    conf = pa.StepConf(do_toi=0)
    steps = 20
    _, desc, bodies, dyn, targets = _setup(1, 1, steps)
    fused = pa.Adjoint(pa.Batch([pa.World.from_desc(desc)], use_graph=False), steps, 8, conf).run(
        targets.astype(np.float32))
    adj = pa.Adjoint(pa.Batch([pa.World.from_desc(desc)], use_graph=False), steps, 8, conf)
    with pytest.raises(pa.PrxError):
        adj.backward(np.zeros((len(bodies), 6), dtype=np.float32))
    state = adj.forward()
    sb = np.zeros((len(bodies), 6), dtype=np.float32)
    sb[:, :2] = 2.0 * (state[:, :2] - targets.astype(np.float32))
    sb[~dyn] = 0.0
    split = adj.backward(sb)
    for k in ("grad_velocity", "grad_friction", "grad_restitution"):
        assert np.array_equal(split[k], fused[k]), k
    # The backward pass restores the starting state: a second forward gives
    # the same final state.
    assert np.array_equal(adj.forward(), state)


def test_simulate_torch_operator():
    """torch.ops.playrho_amd.simulate: the final state equals the adjoint's
    forward pass, and any loss of it back-propagates the adjoint's gradients
    (two chain worlds in one call, motor torques included)."""
    import torch
    from playrho_amd.autograd import simulate

    conf = pa.StepConf(do_toi=0)
    steps, every = 20, 8
    descs, expect = [], []
    rng = np.random.default_rng(5)
    for wi in range(2):
        w, desc, bodies, joints, dyn, _ = _chain([3, wi, 6, 4, 2], seed_targets=wi)
        c = _seed(len(bodies), dyn, rng)
        adj = pa.Adjoint(pa.Batch([pa.World.from_desc(desc)], use_graph=False), steps, every, conf)
        st = adj.forward()
        expect.append((st, c, adj.backward(c)))
        descs.append(desc)
    nb = pa.World.from_desc(descs[0]).body_count
    sh = [dumpfmt.read_description(dumpfmt.Reader(d))[0] for d in descs]
    v0 = torch.tensor(np.stack([[b["vals"][8:11] for b in dumpfmt.read_description(dumpfmt.Reader(d))[1]]
                                for d in descs]), dtype=torch.float32, requires_grad=True)
    fr = torch.tensor([[float(s["friction"]) for s in x] for x in sh], dtype=torch.float32, requires_grad=True)
    re = torch.tensor([[float(s["restitution"]) for s in x] for x in sh], dtype=torch.float32, requires_grad=True)
    tq = torch.tensor([[float(j["vals"][8]) for j in dumpfmt.read_description(dumpfmt.Reader(d))[2]] for d in descs],
                      dtype=torch.float32, requires_grad=True)
    state = simulate(v0, fr, re, descs, steps, every, torque=tq)
    assert state.shape == (2, nb, 6)
    for wi in range(2):
        assert np.array_equal(state[wi].detach().numpy(), expect[wi][0])
    cs = torch.tensor(np.stack([e[1] for e in expect]))
    (state * cs).sum().backward()
    for wi in range(2):
        g = expect[wi][2]
        assert np.array_equal(v0.grad[wi].numpy(), g["grad_velocity"])
        assert np.array_equal(fr.grad[wi].numpy(), g["grad_friction"])
        assert np.array_equal(re.grad[wi].numpy(), g["grad_restitution"])
        assert np.array_equal(tq.grad[wi].numpy(), g["grad_motor_torque"])
