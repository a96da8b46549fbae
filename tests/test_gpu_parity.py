"""Parity of the HIP step with the reference, through the C ABI, on the GPU.

Each golden case starts the device from the reference's own world
description and compares every step's full state (bodies, contact list with
manifolds and ContactFeature ids, proxies, joints, and the step's contact
listener calls with their arguments) against the per-step digests of the
reference float build: bit-exact.  Larger cases compare against the
oracle (CPU restatement) step by step."""
import numpy as np
import pytest

import playrho_amd as pa
from tests import dumpfmt, golden_cases, oracle_lib

pytestmark = pytest.mark.gpu


def _state(world):
    return dumpfmt.read_state(dumpfmt.Reader(world.state_record()))


def _recording(world):
    world.record_contact_events(True)
    return world


def test_device_sincos_matches_host_library():
    rng = np.random.default_rng(7)
    a = np.concatenate([
        rng.uniform(-8, 8, 200000), rng.uniform(-1e-3, 1e-3, 100000), rng.uniform(-500, 500, 100000),
        np.array([0.0, -0.0, 1e-30, 0.78539816, -0.78539816, 119.9, 120.0, 1e5])]).astype(np.float32)
    got = pa.device_sincos(a)
    want = oracle_lib.host_sincosf(a)
    bad = np.nonzero(got.view(np.uint32) != want.view(np.uint32))[0]
    assert bad.size == 0, (a[bad[:5] // 2], got.reshape(-1)[bad[:5]], want.reshape(-1)[bad[:5]])


def _conf(g):
    """The golden case's StepConf: the reference defaults, doToi as recorded."""
    return pa.StepConf(do_toi=1 if g["do_toi"] else 0)


@pytest.mark.parametrize("name", golden_cases.regular_names() + golden_cases.toi_names())
def test_plain_step_matches_reference_every_step(name):
    g = golden_cases.load(name)
    print("case", name, "steps", g["steps"], "doToi", g["do_toi"], flush=True)
    w = _recording(pa.World.from_desc(g["desc_conf"]))
    conf = _conf(g)
    for step in range(1, g["steps"] + 1):
        w.step(conf)
        s = _state(w)
        for k, part in enumerate(golden_cases.PARTS):
            if dumpfmt.digest(s, (part,)) != g["digests"][step, k].tobytes():
                full = g["full"].get(step)
                detail = ""
                if full is not None and part == "bodies":
                    d = np.nonzero(np.any(s["bodies"].view("<u4") != full["bodies"].view("<u4"), axis=1))[0]
                    detail = " bodies differ: %s" % d[:10]
                pytest.fail("%s: step %d differs from the reference in %s%s" % (name, step, part, detail))


@pytest.mark.parametrize("coop_min,toi", [(None, False), (1, False), (1 << 30, False), (None, True)])
def test_batched_step_matches_reference(coop_min, toi):
    """All golden worlds of one doToi setting in one batch, one launch per
    step.  coop_min 1 sweeps every island with contacts by the whole wavefront
    (LDS-staged records), 1 << 30 sweeps every island on one lane; all equal
    the reference."""
    names = golden_cases.toi_names() if toi else golden_cases.regular_names()
    gs = [golden_cases.load(n) for n in names]
    worlds = [_recording(pa.World.from_desc(g["desc_conf"])) for g in gs]
    print("coop_min", coop_min, "doToi", toi, flush=True)
    batch = pa.Batch(worlds, use_graph=False, coop_min=coop_min)
    conf = pa.StepConf(do_toi=1 if toi else 0)
    steps = max(g["steps"] for g in gs)
    done = {}
    for step in range(1, steps + 1):
        batch.step(conf, 1)
        batch.download()
        for n, g, w in zip(names, gs, worlds):
            if step > g["steps"] or n in done:
                continue
            s = _state(w)
            for k, part in enumerate(golden_cases.PARTS):
                if dumpfmt.digest(s, (part,)) != g["digests"][step, k].tobytes():
                    done[n] = (step, part)
                    break
    assert not done, done


@pytest.mark.parametrize("toi", [False, True])
def test_graph_replay_matches_eager(toi):
    """Steps replayed from a captured hipGraph equal eager launches."""
    descs = [golden_cases.load("pile%d" % i)["desc_conf"] for i in range(4)]
    eager = [_recording(pa.World.from_desc(d)) for d in descs]
    graph = [_recording(pa.World.from_desc(d)) for d in descs]
    be = pa.Batch(eager, use_graph=False)
    bg = pa.Batch(graph, use_graph=True)
    conf = pa.StepConf(do_toi=1 if toi else 0)
    for rep in range(3):
        be.step(conf, 10)
        bg.step(conf, 10)
        be.download()
        bg.download()
        for a, b in zip(eager, graph):
            assert a.state_record() == b.state_record(), rep


def test_autotune_restores_state_and_keeps_results():
    """prx_batch_autotune times schedules from the current state, restores it,
    and the chosen schedule steps to the same bits as the default one."""
    descs = [golden_cases.load(n)["desc_conf"] for n in ("pile3", "chain0", "pyramid", "billiards0")]
    a = [_recording(pa.World.from_desc(d)) for d in descs]
    b = [_recording(pa.World.from_desc(d)) for d in descs]
    ba = pa.Batch(a, use_graph=True)
    bb = pa.Batch(b, use_graph=True)
    conf = pa.StepConf(do_toi=0)
    ba.step(conf, 5)
    bb.step(conf, 5)
    ba.sync()
    print("autotune", flush=True)
    chosen = ba.autotune(conf, nsteps=3)
    print("chosen", chosen, flush=True)
    ba.download()
    bb.download()
    for x, y in zip(a, b):
        assert x.state_record() == y.state_record()
    ba.step(conf, 20)
    bb.step(conf, 20)
    ba.download()
    bb.download()
    for x, y in zip(a, b):
        assert x.state_record() == y.state_record()


@pytest.mark.parametrize("lds_bytes", [26 * 1024, 40 * 1024])
def test_lds_schedule_matches_reference(lds_bytes):
    """Padding the workgroup's LDS (fewer workgroups per CU) changes nothing in
    the results."""
    names = ["pile1", "pile3", "chain1", "drop300"]
    gs = [golden_cases.load(n) for n in names]
    worlds = [_recording(pa.World.from_desc(g["desc_conf"])) for g in gs]
    batch = pa.Batch(worlds, use_graph=True, lds_bytes=lds_bytes)
    conf = pa.StepConf(do_toi=0)
    steps = 100
    batch.step(conf, steps)
    batch.download()
    for n, g, w in zip(names, gs, worlds):
        s = _state(w)
        for k, part in enumerate(golden_cases.PARTS):
            assert dumpfmt.digest(s, (part,)) == g["digests"][steps, k].tobytes(), (n, part)


@pytest.mark.parametrize("scene,params,steps,toi", [
    ("pile", [3, 11], 60, False), ("pile", [5, 40], 60, False), ("drop", [1000, 40, 80, 3], 40, False),
    ("chain", [3, 5, 32, 64, 5], 150, False), ("pile", [3, 11], 60, True), ("drop", [1000, 40, 80, 3], 40, True),
    ("billiards", [2, 1], 120, True)])
def test_device_matches_oracle(scene, params, steps, toi):
    w = _recording(pa.World.scene(scene, params))
    conf = pa.StepConf(do_toi=1 if toi else 0)
    ow = oracle_lib.OracleWorld(w.save_desc() + dumpfmt.conf_record(conf))
    for step in range(1, steps + 1):
        w.step(conf)
        got = _state(w)
        want = ow.step(1)[0]
        for part in golden_cases.PARTS:
            assert dumpfmt.digest(got, (part,)) == dumpfmt.digest(want, (part,)), (scene, params, step, part)


@pytest.mark.parametrize("scene,params,coop_min", [
    ("targets", [1, 0], None), ("targets", [2, 3, 10, 30], 1), ("sliders", [1, 0], None), ("joints", [1, 0], 1),
    ("gears", [1, 0], None), ("gears", [2, 1, 40], 1)])
def test_joints_without_warm_start_match_oracle(scene, params, coop_min):
    """doWarmStart = false: every joint's impulses restart from zero and a
    target joint's InitVelocity still damps its body's spin
    (TargetJointConf.cpp:125-126); the lane form (coop_min None) and the
    wave form (coop_min 1) against the oracle, every step."""
    w = _recording(pa.World.scene(scene, params))
    conf = pa.StepConf(do_toi=0, do_warm_start=0)
    ow = oracle_lib.OracleWorld(w.save_desc() + dumpfmt.conf_record(conf))
    batch = pa.Batch([w], use_graph=False, coop_min=coop_min)
    print("no warm start", scene, params, coop_min, flush=True)
    for step in range(1, 121):
        batch.step(conf, 1)
        batch.download()
        got = _state(w)
        want = ow.step(1)[0]
        for part in golden_cases.PARTS:
            assert dumpfmt.digest(got, (part,)) == dumpfmt.digest(want, (part,)), (scene, params, step, part)


def _conf_bytes():
    """The StepConf record (dump layout) of the default conf."""
    return dumpfmt.conf_record(pa.StepConf(do_toi=0))


@pytest.mark.parametrize("name", golden_cases.edit_names())
def test_world_edits_match_reference_every_step(name):
    """SetLocation / SetAngle / SetTransformation / ShiftOrigin /
    SetVelocity / Destroy(JointID) / Destroy(BodyID) / SetType between steps on a
    device-resident world (the reference driver applied the same edits before
    the same steps; a destroyed body's touching contacts end before the next
    step's listener calls)."""
    g = golden_cases.load(name)
    print("case", name, "edits", len(g["edits"]), flush=True)
    w = _recording(pa.World.from_desc(g["desc_conf"]))
    conf = _conf(g)
    for step in range(1, g["steps"] + 1):
        for e in g["edits"]:
            if int(e[0]) != step:
                continue
            kind, body, a, b, c = chr(int(e[1])), int(e[2]), float(e[3]), float(e[4]), float(e[5])
            if kind == "L":
                w.set_location(body, (a, b))
            elif kind == "A":
                w.set_angle(body, a)
            elif kind == "T":
                q = oracle_lib.host_sincosf(np.array([c], dtype=np.float32))[0]  # UnitVec::Get(c)
                w.set_transformation(body, (a, b, float(q[0]), float(q[1])))
            elif kind == "O":
                w.shift_origin((a, b))
            elif kind == "V":
                w.set_velocity(body, (a, b, c))
            elif kind == "J":
                w.destroy_joint(body)
            elif kind == "D":
                w.destroy_body(body)
            elif kind == "S":
                w.set_type(body, int(a))
        w.step(conf)
        s = _state(w)
        for k, part in enumerate(golden_cases.PARTS):
            if dumpfmt.digest(s, (part,)) != g["digests"][step, k].tobytes():
                pytest.fail("%s: step %d differs from the reference in %s" % (name, step, part))


@pytest.mark.parametrize("toi", [False, True])
def test_largest_worlds_inside_autotuned_4096_world_batch(toi):
    """The worlds that set the batched step's time: the 4096-world pile batch
    of the bench (seed 1) stepped 500 steps by one hipGraph launch per step,
    with the schedule autotuned on its state at step 200.  The two largest
    worlds (513 and 511 bodies) and eight others of all sizes equal the
    oracle every 50 steps, and the reference goldens of the two largest at
    every golden step reached (steps 1..300, checked at the 50-step marks)."""
    conf = pa.StepConf(do_toi=1 if toi else 0)
    worlds = [pa.World.scene("pile", [1, w]) for w in range(4096)]
    sizes = [w.body_count for w in worlds]
    big = sorted(range(4096), key=lambda i: -sizes[i])[:2]
    assert sorted(sizes[i] for i in big) == [511, 513]
    picks = big + [7, 100, 777, 1500, 2048, 2999, 3500, 4095]
    for i in picks:
        _recording(worlds[i])
    oracles = {i: oracle_lib.OracleWorld(worlds[i].save_desc() + dumpfmt.conf_record(conf)) for i in picks}
    golden = {3301: golden_cases.load("pile3301" + ("_toi" if toi else "")),
              859: golden_cases.load("pile859" + ("_toi" if toi else ""))}
    batch = pa.Batch(worlds, use_graph=True)
    done = 0
    for mark in range(50, 501, 50):
        batch.step(conf, mark - done)
        if mark == 200:
            batch.sync()
            print("autotune at step 200:", batch.autotune(conf, nsteps=10), flush=True)
        done = mark
        batch.download()
        print("step", mark, "toi", toi, flush=True)
        for i in picks:
            got = _state(worlds[i])
            want = oracles[i].step(50)[-1]
            for part in golden_cases.PARTS:
                assert dumpfmt.digest(got, (part,)) == dumpfmt.digest(want, (part,)), (i, mark, part)
            if i in golden and mark <= 300:
                for k, part in enumerate(golden_cases.PARTS):
                    assert dumpfmt.digest(got, (part,)) == golden[i]["digests"][mark, k].tobytes(), (i, mark, part)
