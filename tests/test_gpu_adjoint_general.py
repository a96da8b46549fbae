"""The general reverse kernel (polygons, edges, two-point manifolds, revolute
joints; each update's vector-Jacobian product from dual-number copies of the
solver functions) against finite differences of the FP64 restatement on
articulated chain worlds, and against the hand-derived disk kernel on the
billiards world."""
import numpy as np
import pytest

import playrho_amd as pa
from tests import dumpfmt
from tests.test_gpu_adjoint import _loss_f64, _setup

pytestmark = pytest.mark.gpu


def _f32(x):
    return float(np.float32(x))


@pytest.mark.parametrize("steps", [12, 30])
def test_general_kernel_equals_disk_kernel(steps):
    """Both reverse kernels on the same recorded billiards steps."""
    conf = pa.StepConf(do_toi=0)
    _, desc, _, _, targets = _setup(1, 0, steps)
    disk = pa.Adjoint(pa.Batch([pa.World.from_desc(desc)], use_graph=False), steps, 8, conf)
    a = disk.run(targets.astype(np.float32))
    gen = pa.Adjoint(pa.Batch([pa.World.from_desc(desc)], use_graph=False), steps, 8, conf)
    gen.use_general(True)
    b = gen.run(targets.astype(np.float32))
    assert a["loss"][0] == b["loss"][0]
    for k in ("grad_velocity", "grad_friction", "grad_restitution"):
        x, y = a[k].ravel(), b[k].ravel()
        rel = np.linalg.norm(x - y) / max(np.linalg.norm(x), 1e-30)
        print(steps, k, "relative difference %.3e" % rel, flush=True)
        assert rel < 1e-5


def _chain(params, seed_targets=7):
    w = pa.World.scene("chain", params)
    desc = w.save_desc()
    _, bodies, joints = dumpfmt.read_description(dumpfmt.Reader(desc))
    dyn = np.array([bool(b["flags"] & 2) for b in bodies])
    rng = np.random.default_rng(seed_targets)
    targets = np.stack([rng.uniform(-5, 40, len(bodies)), rng.uniform(5, 40, len(bodies))], axis=1)
    return w, desc, bodies, joints, dyn, targets.astype(np.float32).astype(np.float64)


# (scene parameters: seed, world, links, falling bodies, motor torque), steps,
# joints and bodies whose gradients are checked.
CASES = [
    ([1, 0, 6, 4, 2], 30, (0, 1, 2, 3, 4, 5), (2, 4, 7, 9)),
    ([1, 0], 40, (0, 3, 9, 21), (3, 17, 30, 40, 70)),
    ([2, 1], 60, (0, 1, 2, 5, 8), (5, 33, 50, 90)),
]


@pytest.mark.parametrize("params,steps,jsel,bsel", CASES)
def test_chain_adjoint_matches_fp64_finite_differences(params, steps, jsel, bsel):
    """Gradients with respect to joint maximum motor torques, initial
    velocities and friction through revolute joints (motor, limits) and
    polygon / disk / edge contacts, against central differences of the FP64
    step (two step sizes; parameters where they disagree have an event inside
    the perturbation and are left out)."""
    conf = pa.StepConf(do_toi=0)
    w, desc, bodies, joints, dyn, targets = _chain(params)
    print("chain", params, "steps", steps, flush=True)
    adj = pa.Adjoint(pa.Batch([w], use_graph=False), steps, 8, conf)
    out = adj.run(targets.astype(np.float32))
    loss64 = _loss_f64(desc, conf, steps, targets, dyn)
    print("loss device %.5f fp64 %.5f" % (out["loss"][0], loss64), flush=True)
    assert abs(out["loss"][0] - loss64) < 1e-4 * loss64

    def central(make, x0, eps):
        xp, xm = _f32(x0 + eps), _f32(x0 - eps)
        return ((_loss_f64(make(xp), conf, steps, targets, dyn) - _loss_f64(make(xm), conf, steps, targets, dyn)) /
                (xp - xm))

    checks = []
    for j in jsel:
        t0 = float(joints[j]["vals"][8])
        checks.append(("T%d" % j, out["grad_motor_torque"][j], lambda x, j=j: pa.desc_with(desc, motor_torque={j: x}),
                       t0, 1e-3))
    for b in bsel:
        v = bodies[b]["vals"][8:11].astype(np.float64)
        for k in (0, 1, 2):
            def make(x, b=b, k=k, v=v):
                vv = v.copy()
                vv[k] = x
                return pa.desc_with(desc, velocities={b: vv})
            checks.append(("v%d%s" % (b, "xyw"[k]), out["grad_velocity"][b, k], make, v[k], 1e-4))
    sdesc, _, _ = dumpfmt.read_description(dumpfmt.Reader(desc))
    for s in (1, len(sdesc) - 1):
        fr = float(sdesc[s]["friction"])
        checks.append(("f%d" % s, out["grad_friction"][s], lambda x, s=s: pa.desc_with(desc, friction={s: x}), fr, 1e-4))
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
    print("relative error over %d parameters %.3e (left out: %s)" % (len(got), rel, skipped), flush=True)
    assert len(skipped) <= len(checks) // 4
    assert rel < 1e-3
    # Motor torques act (some gradient is not zero).
    assert np.any(out["grad_motor_torque"] != 0.0)


def test_chain_batched_equals_single():
    conf = pa.StepConf(do_toi=0)
    steps = 20
    singles = []
    for wi in range(3):
        w, desc, bodies, joints, dyn, targets = _chain([3, wi, 8, 12, 2], seed_targets=wi)
        out = pa.Adjoint(pa.Batch([w], use_graph=False), steps, 8, conf).run(targets.astype(np.float32))
        singles.append((desc, targets, out, len(bodies), len(joints)))
    worlds = [pa.World.from_desc(d) for d, _, _, _, _ in singles]
    tg = np.concatenate([t for _, t, _, _, _ in singles]).astype(np.float32)
    out = pa.Adjoint(pa.Batch(worlds, use_graph=False), steps, 8, conf).run(tg)
    ob = oj = 0
    for k, (_, _, o1, nb, nj) in enumerate(singles):
        assert out["loss"][k] == o1["loss"][0]
        assert np.array_equal(out["grad_velocity"][ob:ob + nb], o1["grad_velocity"])
        assert np.array_equal(out["grad_motor_torque"][oj:oj + nj], o1["grad_motor_torque"])
        ob += nb
        oj += nj


def test_articulated_torch_operator():
    """torch.ops.playrho_amd.articulated_loss: its gradients (motor torques,
    velocities, friction) equal the adjoint's on worlds built with the same
    parameters, two worlds in one call."""
    import torch
    from playrho_amd.autograd import articulated_loss
    steps = 20
    descs, tgs, v0s, frs, res, tqs = [], [], [], [], [], []
    for wi in range(2):
        w, desc, bodies, joints, dyn, targets = _chain([4, wi, 8, 10, 2], seed_targets=wi)
        sdesc, _, _ = dumpfmt.read_description(dumpfmt.Reader(desc))
        descs.append(desc)
        tgs.append(targets)
        v0s.append([b["vals"][8:11] for b in bodies])
        frs.append([s["friction"] for s in sdesc])
        res.append([s["restitution"] for s in sdesc])
        tqs.append([j["vals"][8] for j in joints])
    v0 = torch.tensor(np.array(v0s), dtype=torch.float64, requires_grad=True)
    fr = torch.tensor(np.array(frs), dtype=torch.float64, requires_grad=True)
    re = torch.tensor(np.array(res), dtype=torch.float64, requires_grad=True)
    tq = (torch.tensor(np.array(tqs), dtype=torch.float64) * 1.5).requires_grad_(True)
    loss = articulated_loss(v0, fr, re, tq, descs, np.array(tgs), steps, 8)
    loss.sum().backward()
    worlds = [pa.World.from_desc(pa.desc_with(d, motor_torque={j: float(np.float32(tq[wi, j].item()))
                                                               for j in range(tq.shape[1])}))
              for wi, d in enumerate(descs)]
    direct = pa.Adjoint(pa.Batch(worlds, use_graph=False), steps, 8).run(np.concatenate(tgs).astype(np.float32))
    assert np.array_equal(loss.detach().numpy().astype(np.float32), direct["loss"])
    assert np.array_equal(tq.grad.numpy().reshape(-1).astype(np.float32), direct["grad_motor_torque"])
    assert np.array_equal(v0.grad.numpy().reshape(-1, 3).astype(np.float32), direct["grad_velocity"])
    assert np.array_equal(fr.grad.numpy().reshape(-1).astype(np.float32), direct["grad_friction"])
    assert np.any(tq.grad.numpy() != 0.0)
