"""Contact listeners on the GPU.  The step records the begin / end /
pre-solve / post-solve calls in the reference's call order
(AabbTreeWorld.cpp:2249-2252, 2703-2719, 1671-1674), and the C ABI delivers
them to the registered callbacks once the step is done.  The calls each
callback receives (kind, contact id, manifold or impulses bit for bit,
solved iteration) equal the oracle's calls.  The oracle's calls are pinned to
the reference's own listener calls by the golden digests."""
import struct

import numpy as np
import pytest

import playrho_amd as pa
from tests import dumpfmt, golden_cases, oracle_lib

pytestmark = pytest.mark.gpu


def _bits(x):
    return struct.unpack("<I", struct.pack("<f", x))[0]


def _oracle_calls(state):
    return [(k, c, tuple(int(v) for v in p)) for k, c, p in dumpfmt.iter_events(state)]


class Collector:
    """Listener callbacks that keep every call in the record's word form."""

    def __init__(self):
        self.calls = []

    def begin(self, c):
        self.calls.append((0, c, ()))

    def end(self, c):
        self.calls.append((1, c, ()))

    def pre_solve(self, c, m):
        w = [m.type | (m.point_count << 8), m.points[0].contact_feature, m.points[1].contact_feature,
             _bits(m.local_normal[0]), _bits(m.local_normal[1]), _bits(m.local_point[0]), _bits(m.local_point[1])]
        for q in range(2):
            p = m.points[q]
            w += [_bits(p.local_point[0]), _bits(p.local_point[1]), _bits(p.normal_impulse),
                  _bits(p.tangent_impulse)]
        self.calls.append((2, c, tuple(w)))

    def post_solve(self, c, im, it):
        self.calls.append((3, c, (im.count, it, _bits(im.normal[0]), _bits(im.normal[1]), _bits(im.tangent[0]),
                                  _bits(im.tangent[1]))))

    def attach(self, world):
        world.set_contact_listener(begin=self.begin, end=self.end, pre_solve=self.pre_solve,
                                   post_solve=self.post_solve)

    def take(self):
        c, self.calls = self.calls, []
        return c


@pytest.mark.parametrize("name,steps", [("hello", 60), ("pile3", 150), ("billiards0", 120), ("chain0", 120),
                                        ("drop300", 80), ("pile3_toi", 150), ("billiards0_toi", 120),
                                        ("joints0_toi", 120)])
def test_listener_calls_match_oracle(name, steps):
    """The TOI cases add the calls of the TOI phase: its updates and its
    post-solve reports with the position iterations."""
    g = golden_cases.load(name)
    conf = pa.StepConf(do_toi=1 if g["do_toi"] else 0)
    print("case", name, "steps", steps, "doToi", g["do_toi"], flush=True)
    w = pa.World.from_desc(g["desc_conf"])
    col = Collector()
    col.attach(w)
    ow = oracle_lib.OracleWorld(g["desc_conf"])
    kinds = np.zeros(4, dtype=np.int64)
    for step in range(1, steps + 1):
        w.step(conf)
        want = _oracle_calls(ow.step(1)[0])
        got = col.take()
        if got != want:
            k = next(i for i in range(min(len(got), len(want))) if got[i] != want[i]) \
                if any(a != b for a, b in zip(got, want)) else min(len(got), len(want))
            pytest.fail("%s step %d: %d calls, oracle %d; first difference at call %d: %s vs %s" %
                        (name, step, len(got), len(want), k, got[k:k + 1], want[k:k + 1]))
        words, n = w.contact_event_words()
        assert n == len(want)
        for kind, _, _ in want:
            kinds[kind] += 1
    print("calls by kind (begin, end, pre-solve, post-solve):", kinds.tolist(), flush=True)
    assert kinds[0] > 0 and kinds[2] > 0 and kinds[3] > 0


def test_batch_with_listeners_delivers_every_step():
    """A batch whose worlds have listeners steps one step at a time and
    delivers each step's calls before the next; a world without listeners in
    the same batch records nothing and steps as before."""
    names = ["pile1", "chain1", "pile5"]
    gs = [golden_cases.load(n) for n in names]
    worlds = [pa.World.from_desc(g["desc_conf"]) for g in gs]
    cols = [Collector(), None, Collector()]
    for c, w in zip(cols, worlds):
        if c:
            c.attach(w)
    batch = pa.Batch(worlds, use_graph=True)
    oracles = [oracle_lib.OracleWorld(g["desc_conf"]) for g in gs]
    steps = 40
    batch.step(pa.StepConf(do_toi=0), steps)
    batch.download()
    for c, o, w, g in zip(cols, oracles, worlds, gs):
        states = o.step(steps)
        if c is None:
            assert w.contact_event_words()[1] == 0
            continue
        want = [call for s in states for call in _oracle_calls(s)]
        assert c.take() == want
    # The world without listeners still matches the reference.
    s = dumpfmt.read_state(dumpfmt.Reader(worlds[1].state_record()))
    for k, part in enumerate(golden_cases.PARTS[:4]):
        assert dumpfmt.digest(s, (part,)) == gs[1]["digests"][steps, k].tobytes(), part


def test_recording_switch_inside_batch_is_wrong_state():
    ws = [pa.World.from_desc(golden_cases.load(n)["desc_conf"]) for n in ("pile0", "pile2")]
    batch = pa.Batch(ws, use_graph=False)
    with pytest.raises(pa.PrxError) as e:
        ws[0].record_contact_events(True)
    assert e.value.code == 3
    del batch


@pytest.mark.parametrize("name,steps", [("pile3_toi", 150), ("billiards0_toi", 120)])
def test_call_record_grows_without_losing_calls(name, steps):
    """A world whose call record starts at 24 words: every step with more
    calls (with doToi, TOI events add updates and post-solve reports per
    event) is stepped again from the state before it with a larger record.
    The calls delivered and the state after every step still equal the
    oracle's and the reference's."""
    g = golden_cases.load(name)
    conf = pa.StepConf(do_toi=1 if g["do_toi"] else 0)
    print("case", name, "steps", steps, "record 24 words", flush=True)
    w = pa.World.from_desc(g["desc_conf"])
    w.set_event_record_words(24)
    col = Collector()
    col.attach(w)
    ow = oracle_lib.OracleWorld(g["desc_conf"])
    most = 0
    for step in range(1, steps + 1):
        w.step(conf)
        want = _oracle_calls(ow.step(1)[0])
        got = col.take()
        assert got == want, "%s step %d: %d calls, oracle %d" % (name, step, len(got), len(want))
        most = max(most, w.contact_event_words()[0].size)
        s = dumpfmt.read_state(dumpfmt.Reader(w.state_record()))
        for k, part in enumerate(golden_cases.PARTS[:4]):
            assert dumpfmt.digest(s, (part,)) == g["digests"][step, k].tobytes(), (step, part)
    print("largest step record: %d words" % most, flush=True)
    assert most > 24
