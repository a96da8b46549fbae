"""The adjoint kernel against finite differences of the reference restatement
in double precision (the FP64 build of the same step), on the differentiable
billiards workload: 64 disks in a walled box with a frictional floor,
loss = sum |x_K - target|^2, gradients with respect to initial velocities,
friction and restitution."""
import numpy as np
import pytest

import playrho_amd as pa
from tests import dumpfmt, oracle_lib

pytestmark = pytest.mark.gpu


def _conf_bytes(c):
    return dumpfmt.conf_record(c)


def _loss_f64(desc, conf, steps, targets, dyn):
    ow = oracle_lib.OracleWorld(desc + _conf_bytes(conf), precision=64)
    ow.step(steps, record=False)
    b = ow.bodies_f64(len(dyn))
    d = b[dyn, :2] - targets[dyn]
    return float(np.sum(d * d))


def _setup(seed, world, steps):
    w = pa.World.scene("billiards", [seed, world])
    desc = w.save_desc()
    _, bodies, _ = dumpfmt.read_description(dumpfmt.Reader(desc))
    dyn = np.array([bool(b["flags"] & 2) for b in bodies])
    rng = np.random.default_rng(100 + world)
    targets = np.zeros((len(bodies), 2), dtype=np.float64)
    targets[:, 0] = rng.uniform(-8, 8, len(bodies))
    targets[:, 1] = rng.uniform(1, 10, len(bodies))
    targets = targets.astype(np.float32).astype(np.float64)
    return w, desc, bodies, dyn, targets


# The gradient is that of the float32 step as executed, and finite differences
# of the FP64 step differentiate the same function only while both precisions
# take the same discrete path.  The velocity loop's early exit
# (newInc <= regMinMomentum, default 0) is the first to part: float32 keeps
# iterating on rounding-level increments where float64 reaches exactly zero
# (on this world from step 9 on), and gradients of different iteration counts
# differ.  The long horizon therefore runs every velocity iteration
# (reg_min_momentum < 0); with it, this world's float32 and float64 runs agree
# in every iteration count and impulse clamp over 60 steps (they part at
# step 72, on a normal impulse clamp).
@pytest.mark.parametrize("steps,every,full_iters", [(12, 4, False), (30, 16, False), (60, 16, True)])
def test_adjoint_matches_fp64_finite_differences(steps, every, full_iters):
    conf = pa.StepConf(do_toi=0)
    if full_iters:
        conf.reg_min_momentum = -1.0
    w, desc, bodies, dyn, targets = _setup(1, 0, steps)
    batch = pa.Batch([w], use_graph=False)
    adj = pa.Adjoint(batch, steps, every, conf)
    print("adjoint steps", steps, "every", every, flush=True)
    out = adj.run(targets.astype(np.float32))
    print("diagnostics", adj.diagnostics(), flush=True)
    loss64 = _loss_f64(desc, conf, steps, targets, dyn)
    print("loss device %.6f fp64 %.6f" % (out["loss"][0], loss64), flush=True)
    assert abs(out["loss"][0] - loss64) < 1e-3 * loss64

    shapes, _ = pa.desc_offsets(desc)
    sdesc, _, _ = dumpfmt.read_description(dumpfmt.Reader(desc))
    # Parameters enter the description as float32: a perturbation is the
    # difference of the two float32 values actually stored.
    def f32(x):
        return float(np.float32(x))

    checks = []
    for b in (1, 9, 30, 63):
        v = bodies[b]["vals"][8:11].astype(np.float64)
        for k in (0, 1):
            def f(eps, b=b, k=k, v=v):
                vv = v.copy()
                vv[k] = f32(v[k] + eps)
                return _loss_f64(pa.desc_with(desc, velocities={b: vv}), conf, steps, targets, dyn), vv[k]
            checks.append(("v%d%s" % (b, "xy"[k]), out["grad_velocity"][b, k], f, 1e-3))
    for s in (4, 12, 40, 66):
        fr = float(sdesc[s]["friction"])
        re = float(sdesc[s]["restitution"])
        checks.append(("f%d" % s, out["grad_friction"][s],
                       lambda eps, s=s, fr=fr: (_loss_f64(pa.desc_with(desc, friction={s: f32(fr + eps)}), conf,
                                                          steps, targets, dyn), f32(fr + eps)), 1e-3))
        checks.append(("e%d" % s, out["grad_restitution"][s],
                       lambda eps, s=s, re=re: (_loss_f64(pa.desc_with(desc, restitution={s: f32(re + eps)}),
                                                          conf, steps, targets, dyn), f32(re + eps)), 1e-3))

    def central(f, eps):
        (lp, xp), (lm, xm) = f(eps), f(-eps)
        return (lp - lm) / (xp - xm)

    # Central differences at two step sizes: where they disagree, a contact
    # event (touching onset, a clamp switching, an early exit) lies within the
    # perturbation and the finite difference is not a derivative; such
    # parameters are reported and left out.
    got, want, skipped = [], [], []
    for name, g, f, eps in checks:
        fd = central(f, eps)
        fd2 = central(f, eps / 10)
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


def test_adjoint_batched_equals_single():
    """Every world of a batch gets the gradients it gets alone."""
    conf = pa.StepConf(do_toi=0)
    steps = 10
    singles = []
    for wi in range(3):
        w, desc, bodies, dyn, targets = _setup(1, wi, steps)
        out = pa.Adjoint(pa.Batch([w], use_graph=False), steps, 4, conf).run(targets.astype(np.float32))
        singles.append((desc, targets, out))
    worlds = [pa.World.from_desc(d) for d, _, _ in singles]
    tg = np.concatenate([t for _, t, _ in singles]).astype(np.float32)
    out = pa.Adjoint(pa.Batch(worlds, use_graph=False), steps, 4, conf).run(tg)
    nb = [len(t) for _, t, _ in singles]
    ns = [pa.World.from_desc(d).body_count for d, _, _ in singles]
    ob = 0
    osh = 0
    for k, (desc, _, o1) in enumerate(singles):
        assert out["loss"][k] == o1["loss"][0]
        assert np.array_equal(out["grad_velocity"][ob:ob + nb[k]], o1["grad_velocity"])
        nsh = len(pa.desc_offsets(desc)[0])
        assert np.array_equal(out["grad_friction"][osh:osh + nsh], o1["grad_friction"])
        ob += nb[k]
        osh += nsh
    del ns


def test_autograd_function_matches_adjoint():
    """The libtorch operator (a C++ torch::autograd::Function): gradients
    equal the adjoint's on the same worlds, two worlds in one call, with the
    parameters set through the operator's inputs; a gradient step on the
    initial velocities lowers the loss."""
    import torch
    from playrho_amd.autograd import billiards_loss
    steps = 16
    descs, tgs, v0s, frs, res = [], [], [], [], []
    for wi in range(2):
        w, desc, bodies, dyn, targets = _setup(2, wi, steps)
        sdesc, _, _ = dumpfmt.read_description(dumpfmt.Reader(desc))
        descs.append(desc)
        tgs.append(targets)
        v0s.append([b["vals"][8:11] for b in bodies])
        frs.append([s["friction"] for s in sdesc])
        res.append([s["restitution"] for s in sdesc])
    # Perturbed materials and velocities enter through the operator only.
    v0 = torch.tensor(np.array(v0s), dtype=torch.float64) * 0.9
    fr = torch.tensor(np.array(frs), dtype=torch.float64) * 0.8
    re = torch.tensor(np.array(res), dtype=torch.float64) * 0.7
    v0.requires_grad_(True)
    fr.requires_grad_(True)
    re.requires_grad_(True)
    loss = billiards_loss(v0, fr, re, descs, np.array(tgs), steps, 8)
    loss.sum().backward()
    f32 = lambda x: x.detach().numpy().astype(np.float32)
    worlds = []
    for wi, desc in enumerate(descs):
        nb, ns = v0.shape[1], fr.shape[1]
        worlds.append(pa.World.from_desc(pa.desc_with(
            desc, velocities={b: f32(v0[wi, b]) for b in range(nb)},
            friction={s: float(f32(fr[wi, s])) for s in range(ns)},
            restitution={s: float(f32(re[wi, s])) for s in range(ns)})))
    direct = pa.Adjoint(pa.Batch(worlds, use_graph=False), steps, 8).run(
        np.concatenate(tgs).astype(np.float32))
    assert np.array_equal(loss.detach().numpy().astype(np.float32), direct["loss"])
    assert np.array_equal(v0.grad.numpy().reshape(-1, 3).astype(np.float32), direct["grad_velocity"])
    assert np.array_equal(fr.grad.numpy().reshape(-1).astype(np.float32), direct["grad_friction"])
    assert np.array_equal(re.grad.numpy().reshape(-1).astype(np.float32), direct["grad_restitution"])
    with torch.no_grad():
        v1 = v0 - 1e-3 * v0.grad
    loss1 = billiards_loss(v1, fr.detach(), re.detach(), descs, np.array(tgs), steps, 8)
    print("loss %.4f -> %.4f" % (float(loss.sum()), float(loss1.sum())), flush=True)
    assert float(loss1.sum()) < float(loss.sum())
