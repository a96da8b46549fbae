"""Parity of the HIP step with the reference, through the C ABI, on the GPU.

Each golden case starts the device from the reference's own world
description and compares every step's full state (bodies, contact list with
manifolds and ContactFeature ids, proxies) against the per-step digests of
the reference float build: bit-exact.  Larger cases compare against the
oracle (CPU restatement) step by step."""
import numpy as np
import pytest

import playrho_amd as pa
from tests import dumpfmt, golden_cases, oracle_lib

pytestmark = pytest.mark.gpu


def _state(world):
    return dumpfmt.read_state(dumpfmt.Reader(world.state_record()))


def test_device_sincos_matches_host_library():
    rng = np.random.default_rng(7)
    a = np.concatenate([
        rng.uniform(-8, 8, 200000), rng.uniform(-1e-3, 1e-3, 100000), rng.uniform(-500, 500, 100000),
        np.array([0.0, -0.0, 1e-30, 0.78539816, -0.78539816, 119.9, 120.0, 1e5])]).astype(np.float32)
    got = pa.device_sincos(a)
    want = oracle_lib.host_sincosf(a)
    bad = np.nonzero(got.view(np.uint32) != want.view(np.uint32))[0]
    assert bad.size == 0, (a[bad[:5] // 2], got.reshape(-1)[bad[:5]], want.reshape(-1)[bad[:5]])


@pytest.mark.parametrize("name", golden_cases.names())
def test_plain_step_matches_reference_every_step(name):
    g = golden_cases.load(name)
    print("case", name, "steps", g["steps"], flush=True)
    w = pa.World.from_desc(g["desc_conf"])
    conf = pa.StepConf()
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


def test_batched_step_matches_reference():
    """All golden worlds in one batch, one launch per step."""
    names = golden_cases.names()
    gs = [golden_cases.load(n) for n in names]
    worlds = [pa.World.from_desc(g["desc_conf"]) for g in gs]
    batch = pa.Batch(worlds, use_graph=False)
    conf = pa.StepConf()
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


def test_graph_replay_matches_eager():
    """Steps replayed from a captured hipGraph equal eager launches."""
    descs = [golden_cases.load("pile%d" % i)["desc_conf"] for i in range(4)]
    eager = [pa.World.from_desc(d) for d in descs]
    graph = [pa.World.from_desc(d) for d in descs]
    be = pa.Batch(eager, use_graph=False)
    bg = pa.Batch(graph, use_graph=True)
    conf = pa.StepConf()
    for rep in range(3):
        be.step(conf, 10)
        bg.step(conf, 10)
        be.download()
        bg.download()
        for a, b in zip(eager, graph):
            assert a.state_record() == b.state_record(), rep


@pytest.mark.parametrize("scene,params,steps", [
    ("pile", [3, 11], 60), ("pile", [5, 40], 60), ("drop", [1000, 40, 80, 3], 40)])
def test_device_matches_oracle(scene, params, steps):
    w = pa.World.scene(scene, params)
    ow = oracle_lib.OracleWorld(w.save_desc() + _conf_bytes())
    conf = pa.StepConf()
    for step in range(1, steps + 1):
        w.step(conf)
        got = _state(w)
        want = ow.step(1)[0]
        for part in golden_cases.PARTS:
            assert dumpfmt.digest(got, (part,)) == dumpfmt.digest(want, (part,)), (scene, params, step, part)


def _conf_bytes():
    """The StepConf record (dump layout) of the default conf."""
    c = pa.StepConf()
    import struct
    floats = [getattr(c, f) for f, _ in pa.StepConf._fields_[:18]]
    return (struct.pack("<I", 18) + struct.pack("<18f", *floats) +
            struct.pack("<III", c.reg_velocity_iters, c.reg_position_iters,
                        (1 if c.do_warm_start else 0) | (2 if c.do_toi else 0) | (4 if c.do_blocksolve else 0)))
