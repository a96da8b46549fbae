"""Benchmark of the batched step: body-steps/s at 4096 uneven worlds on one
MI355X, with the plain per-world GPU form and the reference CPU build timed
in the same run.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--worlds 4096] [--seed 1]

Workload: 4096 independent "pile" worlds, world w with a seeded body count
drawn log-uniform in [8, 512] (boxes, disks and 3..8-gons in stacks or piles
on a ground edge), PlayRho's regular step (dt = 1/60, 8 velocity and 3
position iterations, doToi = false), all worlds advanced by one hipGraph
launch per step, for a run of 500 steps.

Timed window: the K timed steps are the last K steps of the 500-step run.
The run's first 500 - W - K steps (the settle phase: bodies fall and piles
form) and the W warmup steps are stepped untimed, then the launch schedule is
autotuned on that settled state, then K steps are timed.  `full_run` times
all 500 steps of a fresh batch under the same schedule.  Inputs are resident
in device memory when a timed region starts.

Prints one JSON line (rank 0, one GPU).
"""
import argparse
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "body-steps/s at 4096 batched worlds vs plain GPU and CPU; adjoint steps/s"
HBM_PEAK_GBS = 8000.0
RUN_STEPS = 500  # steps of the batched pile run
TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "k_step_worlds_pile4096_traffic.json")


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=200)
    p.add_argument("--warmup", type=int, default=30)
    p.add_argument("--worlds", type=int, default=4096)
    p.add_argument("--seed", type=int, default=1)
    p.add_argument("--settle", type=int, default=-1,
                   help="untimed steps before the warmup (-1: 500 - warmup - steps, so that the timed steps "
                        "end at step 500 of the run)")
    p.add_argument("--full-run", type=int, default=1, help="also time all 500 steps of a fresh batch (0: skip)")
    p.add_argument("--phase-table", action="store_true",
                   help="only print the per-phase cycle table of the 16 longest worlds over the timed window "
                        "(phase timers on; a separate run, since the timers cost time)")
    p.add_argument("--cpu-threads", type=int, default=0, help="0: min(16, cores)")
    p.add_argument("--cpu-worlds", type=int, default=1024, help="worlds in the CPU sample")
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--plain-worlds", type=int, default=64)
    p.add_argument("--plain-steps", type=int, default=10)
    p.add_argument("--coop-min", type=int, default=0, help="0: library default")
    p.add_argument("--lds-bytes", type=int, default=0, help="0: library default")
    p.add_argument("--autotune", type=int, default=40, metavar="K",
                   help="steps per candidate schedule for prx_batch_autotune after warmup (0: off)")
    p.add_argument("--big-bodies", type=int, default=20000, help="large single world (0: skip)")
    p.add_argument("--big-steps", type=int, default=600)
    p.add_argument("--big-cpu-steps", type=int, default=30,
                   help="leading steps of the large world timed on the CPU reference and the GPU alike")
    p.add_argument("--adjoint-worlds", type=int, default=1024, help="differentiable billiards worlds (0: skip)")
    p.add_argument("--adjoint-steps", type=int, default=120)
    p.add_argument("--adjoint-every", type=int, default=16)
    p.add_argument("--toi-worlds", type=int, default=4096,
                   help="the batched pile worlds again with doToi = true (0: skip)")
    p.add_argument("--chain-worlds", type=int, default=8192, help="articulated chain worlds (0: skip)")
    p.add_argument("--chain-steps", type=int, default=240)
    p.add_argument("--chain-cpu-worlds", type=int, default=1024, help="worlds in the CPU sample")
    p.add_argument("--chain-adjoint-worlds", type=int, default=8192,
                   help="chain worlds differentiated with respect to joint motor torques (0: skip)")
    p.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                   help="only run a plain device copy of MIB MiB five times (counter calibration)")
    return p.parse_args()


def copy_calibration(mib):
    """A plain device-to-device copy of known size: the HBM counters read for
    it scale the counters read for the step kernel."""
    import torch
    n = mib * (1 << 20) // 4
    a = torch.arange(n, dtype=torch.int32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": n * 4, "copies": 5, "ok": bool(torch.equal(a, b))}))


def settle_steps(args):
    return args.settle if args.settle >= 0 else max(0, RUN_STEPS - args.warmup - args.steps)


def timed_window(args):
    first = settle_steps(args) + args.warmup + 1
    return [first, first + args.steps - 1]


def measured_traffic(args):
    """HBM bytes per step launch from the committed counter summary (two
    rocprofv3 --pmc runs, FETCH_SIZE and WRITE_SIZE, scaled by a plain copy's
    counters), when that summary was taken over this run's timed window of
    the same workload; otherwise None."""
    if not os.path.exists(TRAFFIC_PROFILE):
        return None
    with open(TRAFFIC_PROFILE) as f:
        tr = json.load(f)
    if tr.get("window") != timed_window(args) or tr.get("worlds") != args.worlds or tr.get("seed") != args.seed:
        return None
    return tr


def cpu_threads(args):
    return args.cpu_threads or min(16, os.cpu_count() or 1)


def algorithmic_bytes(awake, proxies, shape_bytes, contacts, joints=0):
    # B_step = 116 B_awake + 48 P + sum(8 + 16 n_v) + 136 C + 88 J
    return 116 * awake + 48 * proxies + shape_bytes + 136 * contacts + 88 * joints


def cpu_baseline(args, threads, do_toi=False):
    ref = os.path.join(ROOT, "oracle", "_ref", "ref_driver")
    n = min(args.cpu_worlds, args.worlds)
    untimed = settle_steps(args) + args.warmup
    first, last = timed_window(args)
    if os.path.exists(ref):
        out = subprocess.run([ref, "--time", "pile", str(untimed), str(args.steps), "1" if do_toi else "0",
                              str(threads), str(n), str(args.seed), "0"],
                             check=True, capture_output=True, text=True).stdout
        r = json.loads(out.strip().splitlines()[-1])
        return dict(value=r["body_steps"] / r["seconds"], unit="body-steps/s", cores=threads, kind="reference",
                    sample="reference PlayRho float build (oracle/_ref, g++ -O3), worlds 0..%d of the same "
                           "workload, steps %d..%d timed after %d untimed steps, %d std::threads, one world per "
                           "thread at a time, doToi=%s" % (n - 1, first, last, untimed, threads,
                                                           "true" if do_toi else "false"),
                    seconds=r["seconds"])
    from tests import oracle_lib
    import playrho_amd as pa
    descs = []
    for w in range(n):
        wd = pa.World.scene("pile", [args.seed, w])
        descs.append(wd.save_desc() + conf_record(pa, do_toi))
    secs, bs = oracle_lib.time_batch(descs, args.steps, threads)
    return dict(value=bs / secs, unit="body-steps/s", cores=threads, kind="port",
                sample="oracle restatement, worlds 0..%d, %d steps from step 1" % (n - 1, args.steps),
                seconds=secs)


def fp64_report(args, pa, conf, body_counts):
    """Device single-step error against the FP64 oracle from the same state,
    at the timed window's first step, on the two largest worlds and a
    median-sized one: each is stepped on the device to the step before, its
    state checked bit for bit against the float reference restatement, then
    one device step is compared with the FP64 oracle's step from that state
    widened exactly (tests/fp64.py states the metric and its bounds)."""
    from tests import dumpfmt, fp64, oracle_lib
    first, _ = timed_window(args)
    order = sorted(range(len(body_counts)), key=lambda w: -body_counts[w])
    ids = [order[0], order[1], order[len(order) // 2]]
    worlds = [pa.World.scene("pile", [args.seed, w]) for w in ids]
    oracles = [oracle_lib.OracleWorld(w.save_desc() + dumpfmt.conf_record(conf), 32) for w in worlds]
    sb = pa.Batch(worlds, use_graph=True)
    sb.step(conf, first - 1)
    sb.sync()
    del sb
    worst, exact = None, True
    for w, o in zip(worlds, oracles):
        o.step(first - 1, record=False)
        got = dumpfmt.read_state(dumpfmt.Reader(w.state_record()))
        want = o.state()
        exact = exact and all(dumpfmt.digest(got, (p,)) == dumpfmt.digest(want, (p,))
                              for p in ("bodies", "contacts", "proxies"))
        d = o.widen()
        w.step(conf)
        d.step(1, record=False)
        after = dumpfmt.read_state(dumpfmt.Reader(w.state_record()))
        e = fp64.errors(fp64.state_bodies(after), d.bodies_f64(w.body_count), fp64.dynamic_mask(after))
        worst = e if worst is None else {k: max(worst[k], e[k]) for k in e}
    return {"worlds": ids, "bodies": [body_counts[i] for i in ids], "step": first,
            "device_equals_float_reference_before_the_step": exact, "max_error": worst,
            "bounds": {"position": fp64.POS_BOUND, "angle": fp64.ANGLE_BOUND, "velocity": fp64.VEL_BOUND},
            "within_bounds": fp64.within_bounds(worst),
            "metric": "per body |x - x64| / max(|x64|, 1) for position (m), angle (rad), velocity (m/s) and "
                      "spin (rad/s), max over dynamic bodies; x64 from the FP64 oracle's step from the same state"}


def toi_bench(args, pa, threads, regular_value):
    """The batched pile workload stepped with the TOI phase (doToi = true, the
    reference's default StepConf): the regular step kernel then the TOI kernel
    per step, both in the captured graph."""
    conf = pa.StepConf(do_toi=1)
    n = args.toi_worlds
    worlds = [pa.World.scene("pile", [args.seed, w]) for w in range(n)]
    batch = pa.Batch(worlds, use_graph=True, coop_min=args.coop_min or None, lds_bytes=args.lds_bytes or None)
    dyn = batch.dynamic_bodies()
    batch.step(conf, settle_steps(args) + args.warmup)
    batch.sync()
    t1 = time.perf_counter()
    batch.step(conf, args.steps)
    batch.sync()
    wall = time.perf_counter() - t1
    kernel_ms, launches = batch.last_timing()
    first, last = timed_window(args)
    out = {"value": dyn * args.steps / wall, "unit": "body-steps/s", "ms_per_step": 1e3 * wall / args.steps,
           "workload": "%d batched uneven pile worlds (seed %d), dt=1/60 8/3 iterations, doToi=true "
                       "(TOI defaults of StepConf), steps %d..%d of the 500-step run timed" % (n, args.seed, first,
                                                                                             last),
           "kernel_ms_per_step": kernel_ms / max(1, launches),
           "ratio_vs_doToi_false": (dyn * args.steps / wall) / regular_value}
    if not args.no_cpu:
        cb = cpu_baseline(args, threads, do_toi=True)
        out["cpu_baseline"] = cb
        out["ratio_vs_cpu"] = out["value"] / cb["value"]
    return out


def large_world(args, pa, conf):
    """One 20,000-body drop world (disks and 3..8-gons into a 400 x 400 box),
    graph-coloured mode: body-steps/s over the whole run, and over its first
    steps next to the reference CPU build (one thread: PlayRho steps a world
    on one thread) on the same steps."""
    params = [args.big_bodies, 200, 400, args.seed]
    conf = pa.StepConf()  # the reference's defaults: doToi = true
    w = pa.World.scene("drop", params)
    dyn = sum(1 for f in w.bodies()[1] if f & 2)
    big = pa.BigWorld(w)
    k = min(args.big_cpu_steps, args.big_steps)
    big.sync()
    t0 = time.perf_counter()
    big.step(conf, k)
    big.sync()
    t_first = time.perf_counter() - t0
    t1 = time.perf_counter()
    big.step(conf, args.big_steps - k)
    big.sync()
    t_rest = time.perf_counter() - t1
    wall = t_first + t_rest
    met = big.metrics()
    pos, flags = w.bodies()
    d = (flags & 2) != 0
    inside = d & (abs(pos[:, 3]) < 201.0) & (pos[:, 4] > -1.0) & (pos[:, 4] < 401.0)
    met["escaped"] = int(d.sum() - inside.sum())
    met["mean_height"] = float(pos[inside, 4].mean()) if inside.any() else None
    out = {
        "workload": "one drop world: %d bodies (half disks r U[0.25,0.75], half 3..8-gons R U[0.3,0.8], "
                    "friction U[0.2,0.8]) into a 400 x 400 static box of 4 edges, seed %d, dt=1/60 8/3 "
                    "iterations doToi=true (TOI against static shapes), LBVH broadphase, graph-coloured "
                    "Gauss-Seidel" % (args.big_bodies, args.seed),
        "steps": args.big_steps,
        "value": dyn * args.big_steps / wall,
        "unit": "body-steps/s",
        "ms_per_step": 1e3 * wall / max(1, args.big_steps),
        "first_steps": {"steps": k, "value": dyn * k / t_first if k else None},
        "end_state": met,
    }
    ref = os.path.join(ROOT, "oracle", "_ref", "ref_driver")
    if k and not args.no_cpu and os.path.exists(ref):
        r = subprocess.run([ref, "--time", "drop", "0", str(k), "1", "1", "1"] + [str(x) for x in params],
                           check=True, capture_output=True, text=True).stdout
        r = json.loads(r.strip().splitlines()[-1])
        cpu = r["body_steps"] / r["seconds"]
        out["cpu_baseline"] = {"value": cpu, "unit": "body-steps/s", "cores": 1, "kind": "reference",
                               "sample": "reference float build, doToi=true, steps 1..%d of the same world, one "
                                         "thread (the reference steps one world on one thread)" % k,
                               "seconds": r["seconds"]}
        out["ratio_vs_cpu_first_steps"] = out["first_steps"]["value"] / cpu
    return out


def adjoint_bench(args, pa, conf):
    """Differentiable batched billiards: 1024 worlds of 64 disks (radius 0.5)
    on a frictional floor in a walled 20 x 20 box, initial velocities
    N(0, 3^2), friction and restitution U[0, 1]; loss = sum of squared
    distances of the final positions to seeded targets after 120 steps;
    gradients with respect to initial velocities, friction and restitution
    through the adjoint, checkpointing every 16 steps."""
    import numpy as np
    worlds = [pa.World.scene("billiards", [args.seed, w]) for w in range(args.adjoint_worlds)]
    batch = pa.Batch(worlds, use_graph=False)
    adj = pa.Adjoint(batch, args.adjoint_steps, args.adjoint_every, conf)
    rng = np.random.default_rng(args.seed)
    nb = sum(w.body_count for w in worlds)
    targets = np.stack([rng.uniform(-8, 8, nb), rng.uniform(1, 10, nb)], axis=1).astype(np.float32)
    adj.run(targets)  # warm-up (allocations, code paths)
    t0 = time.perf_counter()
    out = adj.run(targets)
    wall = time.perf_counter() - t0
    ws = args.adjoint_worlds * args.adjoint_steps
    return {
        "workload": "%d billiards worlds x 64 disks, %d steps, checkpoint every %d, loss sum |x_T - target|^2, "
                    "d loss / d (initial velocity, friction, restitution)" % (args.adjoint_worlds,
                                                                           args.adjoint_steps,
                                                                           args.adjoint_every),
        "value": ws / wall,
        "unit": "adjoint world-steps/s",
        "note": "one run = the forward pass with checkpoints, the recomputation of every segment with the "
                "solve recorded, and the reversed steps; value = worlds x steps / wall time of the run",
        "seconds": wall,
        "device_ms": adj.timing_ms(),
        "mean_loss": float(np.mean(out["loss"])),
        "grad_norm": float(np.linalg.norm(out["grad_velocity"])),
    }


def articulated(args, pa, conf, threads):
    """Batched articulated worlds: 8192 worlds, each a 32-link revolute-joint
    chain of 1 x 0.125 half-extent boxes hanging from a static anchor (joint
    friction by a zero-speed motor of 2 N m, every third joint limited to
    +-60 degrees) with 64 seeded falling boxes and disks and a ground edge;
    body-steps/s over the configured steps, next to the reference CPU build on
    a sample of the same worlds."""
    worlds = [pa.World.scene("chain", [args.seed, w]) for w in range(args.chain_worlds)]
    joints = sum(w.joint_count for w in worlds)
    batch = pa.Batch(worlds, use_graph=True)
    dyn = batch.dynamic_bodies()
    warm = 10
    batch.step(conf, warm)
    batch.sync()
    t0 = time.perf_counter()
    batch.step(conf, args.chain_steps)
    batch.sync()
    wall = time.perf_counter() - t0
    kernel_ms, launches = batch.last_timing()
    batch.download()
    c1 = batch.counts()
    out = {
        "workload": "%d chain worlds (32 revolute-jointed 1 x 0.125 boxes from a static anchor, 64 falling "
                    "boxes and disks, ground edge; seed %d), dt=1/60 8/3 iterations doToi=false, one hipGraph "
                    "launch per step; steps %d..%d timed" % (args.chain_worlds, args.seed, warm + 1,
                                                              warm + args.chain_steps),
        "worlds": args.chain_worlds,
        "dynamic_bodies": int(dyn),
        "joints": int(joints),
        "steps": args.chain_steps,
        "value": dyn * args.chain_steps / wall,
        "unit": "body-steps/s",
        "ms_per_step": 1e3 * wall / args.chain_steps,
        "kernel_ms_per_launch": kernel_ms / max(1, launches),
        "touching_contacts_end": c1[4],
    }
    if args.chain_adjoint_worlds > 0:
        import numpy as np
        n = min(args.chain_adjoint_worlds, args.chain_worlds)
        aw = [pa.World.scene("chain", [args.seed, w]) for w in range(n)]
        adj = pa.Adjoint(pa.Batch(aw, use_graph=False), args.chain_steps, args.adjoint_every, conf)
        rng = np.random.default_rng(args.seed)
        nb = sum(w.body_count for w in aw)
        targets = np.stack([rng.uniform(-5, 40, nb), rng.uniform(5, 40, nb)], axis=1).astype(np.float32)
        adj.run(targets)  # warm-up (allocations, code paths)
        t0 = time.perf_counter()
        res = adj.run(targets)
        wall_a = time.perf_counter() - t0
        out["adjoint"] = {
            "workload": "%d of the chain worlds, %d steps from their initial state, checkpoint every %d, loss sum "
                        "|x_T - target|^2, d loss / d (joint maximum motor torques, initial velocities, friction, "
                        "restitution)" % (n, args.chain_steps, args.adjoint_every),
            "value": n * args.chain_steps / wall_a,
            "unit": "adjoint world-steps/s",
            "seconds": wall_a,
            "device_ms": adj.timing_ms(),
            "note": "one run = the forward pass with checkpoints, the recomputation of every segment with the "
                    "solve recorded, and the reversed steps (general reverse kernel)",
            "grad_motor_torque_norm": float(np.linalg.norm(res["grad_motor_torque"])),
        }
    ref = os.path.join(ROOT, "oracle", "_ref", "ref_driver")
    if not args.no_cpu and os.path.exists(ref):
        n = min(args.chain_cpu_worlds, args.chain_worlds)
        r = subprocess.run([ref, "--time", "chain", str(warm), str(args.chain_steps), "0", str(threads), str(n),
                            str(args.seed), "0"], check=True, capture_output=True, text=True).stdout
        r = json.loads(r.strip().splitlines()[-1])
        cpu = r["body_steps"] / r["seconds"]
        out["cpu_baseline"] = {"value": cpu, "unit": "body-steps/s", "cores": threads, "kind": "reference",
                               "sample": "reference float build, chain worlds 0..%d, steps %d..%d timed, %d "
                                         "std::threads" % (n - 1, warm + 1, warm + args.chain_steps, threads),
                               "seconds": r["seconds"]}
        out["ratio_vs_cpu"] = out["value"] / cpu
    return out


def conf_record(pa, do_toi=False):
    """The StepConf record the oracle reads (the default conf, doToi as given)."""
    import struct
    c = pa.StepConf(do_toi=1 if do_toi else 0)
    floats = [getattr(c, f) for f, _ in pa.StepConf._fields_[:18]]
    toi_f = [c.toi_resolution_rate, c.toi_min_separation, c.toi_min_momentum, c.target_depth, c.tolerance]
    toi_i = [c.toi_velocity_iters, c.toi_position_iters, c.max_toi_root_iters, c.max_toi_iters,
             c.max_distance_iters, c.max_sub_steps]
    return (struct.pack("<I", 18) + struct.pack("<18f", *floats) +
            struct.pack("<III", c.reg_velocity_iters, c.reg_position_iters,
                        (1 if c.do_warm_start else 0) | (2 if c.do_toi else 0) | (4 if c.do_blocksolve else 0)) +
            struct.pack("<I", 5) + struct.pack("<5f", *toi_f) + struct.pack("<6I", *toi_i))


def full_run(args, pa, conf, schedule, dyn):
    """All 500 steps of a fresh batch of the same worlds, timed from step 1,
    under the schedule the autotuner chose."""
    worlds = [pa.World.scene("pile", [args.seed, w]) for w in range(args.worlds)]
    kw = {}
    if schedule:
        kw = dict(order=schedule["order"], coop_min=schedule["coop_min_contacts"],
                  lds_bytes=schedule["lds_bytes"] or None)
    batch = pa.Batch(worlds, use_graph=True, **kw)
    batch.sync()
    t0 = time.perf_counter()
    batch.step(conf, RUN_STEPS)
    batch.sync()
    wall = time.perf_counter() - t0
    kernel_ms, launches = batch.last_timing()
    return {"steps": [1, RUN_STEPS], "value": dyn * RUN_STEPS / wall, "unit": "body-steps/s",
            "ms_per_step": 1e3 * wall / RUN_STEPS, "kernel_ms_per_launch": kernel_ms / max(1, launches),
            "note": "the whole 500-step run (free fall, piling, settling) of a fresh batch, timed from step 1"}


def phase_table(args, pa, batch, conf, first, last):
    """Per-phase shader-clock cycles of the 16 longest worlds over the timed
    window (phase timers on), as one JSON line."""
    import numpy as np
    batch.phase_cycles(True)
    batch.sync()
    t0 = time.perf_counter()
    batch.step(conf, args.steps)
    batch.sync()
    wall = time.perf_counter() - t0
    pc = batch.phase_cycles(False).astype(np.float64) / args.steps
    nb = [w.body_count for w in batch.worlds]
    tot = pc[:, 13]
    order = np.argsort(-tot)[:16]
    names = pa.Batch.PHASES
    rows = [dict(world=int(w), bodies=int(nb[w]), **{names[i]: int(pc[w, i]) for i in range(14)}) for w in order]
    print(json.dumps({"window": [first, last], "ms_per_step_with_timers": 1e3 * wall / args.steps,
                      "cycles_per_step": rows,
                      "sum_over_worlds": {names[i]: int(pc[:, i].sum()) for i in range(14)}}))


def main():
    args = parse()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.gpus != 1:
        print("bench.py: this engine runs on one GPU; worlds would shard across GPUs without "
              "communication, but no figure at several GPUs is produced here (use --gpus 1)",
              file=sys.stderr)
        sys.exit(2)
    import playrho_amd as pa
    pa.require_device()
    conf = pa.StepConf(do_toi=0)

    t0 = time.time()
    worlds = [pa.World.scene("pile", [args.seed, w]) for w in range(args.worlds)]
    build_s = time.time() - t0
    total_bodies = sum(w.body_count for w in worlds)
    t_up = time.perf_counter()
    batch = pa.Batch(worlds, use_graph=True, coop_min=args.coop_min or None, lds_bytes=args.lds_bytes or None)
    batch.sync()
    upload_s = time.perf_counter() - t_up
    dyn = batch.dynamic_bodies()

    settle = settle_steps(args)
    first, last = timed_window(args)
    batch.step(conf, settle + args.warmup)
    batch.sync()
    schedule = None
    if args.autotune > 0:
        t_at = time.time()
        schedule = batch.autotune(conf, args.autotune)
        schedule["autotune_s"] = round(time.time() - t_at, 3)
    if args.phase_table:
        phase_table(args, pa, batch, conf, first, last)
        return
    batch.download()
    c0 = batch.counts()

    batch.sync()
    t1 = time.perf_counter()
    batch.step(conf, args.steps)
    batch.sync()
    wall = time.perf_counter() - t1
    kernel_ms, launches = batch.last_timing()

    t_down = time.perf_counter()
    batch.download()
    download_s = time.perf_counter() - t_down
    c1 = batch.counts()
    contacts_avg = 0.5 * (c0[4] + c1[4])
    bytes_step = 0.5 * (algorithmic_bytes(*c0[:4]) + algorithmic_bytes(*c1[:4]))
    ms_launch = kernel_ms / max(1, launches)
    achieved = bytes_step / (ms_launch * 1e-3) / 1e9

    value = dyn * args.steps / wall
    result = {
        "metric": METRIC,
        "value": value,
        "unit": "body-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * wall / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "%d batched uneven pile worlds (bodies log-uniform in [8,512], seed %d), regular "
                        "step dt=1/60 8/3 iterations doToi=false, one hipGraph launch per step; a 500-step "
                        "run whose steps %d..%d are timed (steps 1..%d untimed: %d settle + %d warmup)"
                        % (args.worlds, args.seed, first, last, first - 1, settle, args.warmup),
            "worlds": args.worlds,
            "bodies": total_bodies,
            "dynamic_bodies": int(dyn),
            "run_steps": RUN_STEPS,
            "settle_steps": settle,
            "timed_steps": [first, last],
            "schedule": schedule,
        },
        "contacts_per_s": contacts_avg * args.steps / wall,
        "host_transfer": {"upload_s": upload_s, "download_s": download_s,
                          "value_with_download_every_step": dyn / (wall / args.steps + download_s),
                          "note": "upload = batch creation (allocation + host->device copy of every world); "
                                  "download = the whole batch state device->host once; the last figure is "
                                  "body-steps/s if a caller read the full state back after every step"},
        "state": {"awake_dynamic_bodies": [c0[0], c1[0]], "proxies": c1[1], "contacts": [c0[3], c1[3]],
                  "touching_contacts": [c0[4], c1[4]]},
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": None,
            "kernel": "k_step_worlds",
            "window": [first, last],
            "algorithmic_bytes_per_launch": bytes_step,
            "kernel_ms_per_launch": ms_launch,
            "note": "achieved = algorithmic bytes per launch (mean of the formula at the window's first and "
                    "last state) / mean k_step_worlds duration over the timed window (HIP events on the "
                    "launch stream)",
        },
    }

    tr = measured_traffic(args)
    if tr is not None:
        result["roofline"]["traffic"] = tr["bytes_per_launch"]
        result["roofline"]["traffic_note"] = tr["note"]
    else:
        result["roofline"]["traffic_note"] = ("no HBM counter profile of steps %d..%d of this workload is "
                                              "committed (profiles/k_step_worlds_pile4096_traffic.json holds "
                                              "one window)" % (first, last))

    if args.full_run:
        result["full_run"] = full_run(args, pa, conf, schedule, dyn)

    if args.toi_worlds > 0:
        result["toi"] = toi_bench(args, pa, cpu_threads(args), value)
    if args.big_bodies > 0 and args.big_steps > 0:
        result["large_world"] = large_world(args, pa, conf)
    if args.adjoint_worlds > 0 and args.adjoint_steps > 0:
        result["adjoint"] = adjoint_bench(args, pa, conf)
    if args.chain_worlds > 0 and args.chain_steps > 0:
        result["articulated"] = articulated(args, pa, conf, cpu_threads(args))

    # Plain per-world form: one world per launch, synchronised, on a sample.
    if args.plain_worlds <= 0:
        result["setup_s"] = build_s
        if not args.no_cpu:
            threads = cpu_threads(args)
            cb = cpu_baseline(args, threads)
            result["cpu_baseline"] = cb
            result["ratio_vs_cpu"] = value / cb["value"]
        print(json.dumps(result))
        return
    step_idx = max(1, args.worlds // max(1, args.plain_worlds))
    sample = [pa.World.scene("pile", [args.seed, w]) for w in range(0, args.worlds, step_idx)][:args.plain_worlds]
    # The sample is brought to the timed window's first step by a batch of
    # its own (released afterwards), then stepped world by world.
    sb = pa.Batch(sample, use_graph=True)
    sb.step(conf, first - 1)
    sb.sync()
    del sb
    nb = sum(sum(1 for f in w.bodies()[1] if f & 2) for w in sample)
    t2 = time.perf_counter()
    for w in sample:
        for _ in range(args.plain_steps):
            w.step(conf)
    plain_s = time.perf_counter() - t2
    plain = nb * args.plain_steps / plain_s
    result["plain_gpu"] = {"value": plain, "unit": "body-steps/s",
                           "sample": "%d worlds (every %dth), one synchronised launch per world per step, "
                                     "steps %d..%d" % (len(sample), step_idx, first, first + args.plain_steps - 1)}
    result["ratio_vs_plain_gpu"] = value / plain

    if not args.no_cpu:
        threads = cpu_threads(args)
        cb = cpu_baseline(args, threads)
        cb["fp64_single_step_error"] = fp64_report(args, pa, conf, [w.body_count for w in worlds])
        result["cpu_baseline"] = cb
        result["ratio_vs_cpu"] = value / cb["value"]
    result["setup_s"] = build_s
    print(json.dumps(result))


if __name__ == "__main__":
    main()
