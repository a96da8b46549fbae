"""Loads the committed golden fixtures (tests/golden/*.npz)."""
import glob
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")
PARTS = ("bodies", "contacts", "proxies", "joints", "events", "stats")


def names():
    return sorted(os.path.basename(p)[:-4] for p in glob.glob(os.path.join(GOLDEN, "*.npz")))


def regular_names():
    """Cases stepped with doToi = false and no world edits."""
    return [n for n in names() if not n.endswith("_toi") and "_edit" not in n]


def toi_names():
    """Cases stepped with the TOI phase (doToi = true), no world edits."""
    return [n for n in names() if n.endswith("_toi") and "_edit" not in n]


def edit_names():
    """Cases with world edits between steps."""
    return [n for n in names() if "_edit" in n]


def load(name):
    z = np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False)
    g = dict(desc_conf=z["desc_conf"].tobytes(), digests=z["digests"], scene=str(z["scene"]),
             params=[int(x) for x in z["params"]], steps=int(z["steps"]), full={},
             do_toi=bool(z["do_toi"]) if "do_toi" in z.files else False,
             edits=z["edits"] if "edits" in z.files else np.zeros((0, 6)))
    for st in z["full_steps"]:
        st = int(st)
        g["full"][st] = {k: z["full%d_%s" % (st, k)] for k in
                         ("stats", "bodies", "bflags", "cu", "ct", "cf", "pid", "paabb", "jlimit", "jimp", "events")}
    return g


def first_mismatch(digests, states, dumpfmt, start=1):
    """Compares per-step digests of `states` (steps start..) against the golden
    digests; returns (step, part) of the first difference or None."""
    for i, s in enumerate(states):
        step = start + i
        for k, part in enumerate(PARTS):
            if dumpfmt.digest(s, (part,)) != digests[step, k].tobytes():
                return step, part
    return None
