"""Writes tests/golden/drop20k.json: quality measures (tests/bigmetrics.py) of
the reference build stepping the one-large-world workload (20,000 bodies, half
disks r U[0.25, 0.75] and half 3..8-gons R U[0.3, 0.8], dropped into a 400 x
400 box of four edges, seed 1) at steps 100, 300 and 600, with doToi false
and true (the reference's default).

Run from the repository root where the reference checkout is present and
oracle/_ref is built (make -f oracle/Makefile.ref); each run takes minutes.

    python tests/golden/make_big_fixture.py
"""
import json
import os
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from tests import bigmetrics, dumpfmt, oracle_lib  # noqa: E402

PARAMS = [20000, 200, 400, 1]
STEPS = (100, 300, 600)


def run(do_toi, path):
    env = dict(os.environ, PRX_DUMP_STEPS=",".join(str(s) for s in STEPS))
    subprocess.run([oracle_lib.REF_DRIVER, "drop", str(STEPS[-1]), "1" if do_toi else "0", path] +
                   [str(p) for p in PARAMS], check=True, env=env)
    with open(path, "rb") as f:
        return f.read()


def measure(buf):
    desc_conf, d = oracle_lib.split_dump(buf)
    info = bigmetrics.WorldInfo(desc_conf)
    return {str(s["step"]): bigmetrics.measures(info, s) for s in d["states"][1:]}


def main(paths=None):
    out = {"scene": "drop", "params": PARAMS, "steps": list(STEPS), "reference": {}}
    with tempfile.TemporaryDirectory() as tmp:
        for toi in (False, True):
            p = paths[toi] if paths else os.path.join(tmp, "drop20k_%d.bin" % toi)
            buf = open(p, "rb").read() if paths else run(toi, p)
            out["reference"]["doToi_%s" % ("true" if toi else "false")] = measure(buf)
    with open(os.path.join(HERE, "drop20k.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main({False: sys.argv[1], True: sys.argv[2]} if len(sys.argv) == 3 else None)
