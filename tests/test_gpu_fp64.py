"""Single-step error of the device step against the FP64 oracle, from the
same input state (the error bounds and their reasons are stated in
tests/fp64.py).  Before each compared step the device state equals the
float reference's bit for bit (the oracle's float instantiation, pinned to the
reference build by the golden digests); the FP64 oracle starts from that
state widened exactly and takes the same step."""
import numpy as np
import pytest

import playrho_amd as pa
from tests import dumpfmt, fp64, oracle_lib

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("scene,params,every", [("pyramid", [20], 3), ("pile", [1, 1970], 5), ("chain", [1, 0], 4)])
def test_device_single_step_error_against_fp64(scene, params, every):
    steps = 240
    print("case", scene, params, "steps", steps, flush=True)
    w = pa.World.scene(scene, params)
    conf = pa.StepConf(do_toi=0)
    f = oracle_lib.OracleWorld(w.save_desc() + dumpfmt.conf_record(conf), 32)
    worst = None
    compared = 0
    for n in range(1, steps + 1):
        d = f.widen() if n % every == 0 else None
        w.step(conf)
        want = f.step(1)[0]
        got = dumpfmt.read_state(dumpfmt.Reader(w.state_record()))
        for part in ("bodies", "contacts", "proxies", "joints"):
            assert dumpfmt.digest(got, (part,)) == dumpfmt.digest(want, (part,)), (n, part)
        if d is None:
            continue
        d.step(1, record=False)
        e = fp64.errors(fp64.state_bodies(got), d.bodies_f64(w.body_count), fp64.dynamic_mask(got))
        assert fp64.within_bounds(e), (n, e)
        worst = e if worst is None else {k: max(worst[k], e[k]) for k in e}
        compared += 1
    print(scene, "steps compared", compared, "worst", {k: "%.2e" % v for k, v in worst.items()}, flush=True)
    assert worst["abs_position"] > 0
