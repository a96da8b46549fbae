"""Reader for the per-step state record stream written by oracle/ref_driver.cpp
and by the oracle restatement (same layout), plus canonical per-step digests.

A record holds, after each step: bodies (pos0, pos1, velocity, under-active
time, flags), the contact list in the world's contact order (keys, ids, flags,
manifold with ContactFeatures and impulses), the broadphase proxies (id and
fat AABB), the joints (limit state, impulses) and the step's contact listener
calls (begin, end, pre-solve with the old manifold, post-solve with the
impulses).
"""
import hashlib
import struct

import numpy as np

BODY_F = 10
CONTACT_U = 4
CONTACT_T = 3
CONTACT_F = 14
JOINT_DESC_F = 12
GEAR_KIND = 10
JOINT_F = 4
# Listener kinds: 0 begin, 1 end, 2 pre-solve (old manifold: type | count << 8,
# two contact features, 12 floats), 3 post-solve (count, solved position
# iteration, normal impulses [2], tangent impulses [2]).
EVENT_PAYLOAD = (0, 0, 15, 6)
# StepStats words: PreStepStats (proxies created, proxies moved, contacts
# destroyed, added, updated, skipped), then RegStepStats (min separation and
# max incremental impulse as float bits, islands found, solved, bodies slept,
# max island bodies, contacts added, updated, skipped, proxies moved, sums of
# position and velocity iterations), then ToiStepStats (min separation and
# max incremental impulse as float bits, islands found, solved, contacts
# found, at max sub-steps, TOIs updated, updated touching, skipped touching,
# added, proxies moved, sums of position and velocity iterations, max
# distance, TOI and root iterations).
STATS = 34


class Reader:
    def __init__(self, buf):
        self.b = memoryview(buf)
        self.o = 0

    def u32(self):
        v = struct.unpack_from("<I", self.b, self.o)[0]
        self.o += 4
        return v

    def arr(self, dtype, n):
        a = np.frombuffer(self.b, dtype=dtype, count=n, offset=self.o).copy()
        self.o += 4 * n
        return a


def read_description(r):
    shapes = []
    for _ in range(r.u32()):
        n = r.u32()
        hdr = r.arr("<f4", 4)
        sensor = r.u32()
        catmask = r.u32()
        group = r.u32()
        verts = r.arr("<f4", 2 * n).reshape(n, 2)
        normals = r.arr("<f4", 2 * n).reshape(n, 2)
        shapes.append(dict(radius=hdr[0], density=hdr[1], friction=hdr[2],
                           restitution=hdr[3], sensor=sensor, catmask=catmask,
                           group=group, verts=verts, normals=normals))
    bodies = []
    for _ in range(r.u32()):
        btype = r.u32()
        flags = r.u32()
        vals = r.arr("<f4", 19)
        ns = r.u32()
        sids = r.arr("<u4", ns)
        bodies.append(dict(type=btype, flags=flags, vals=vals, shapes=sids))
    joints = []
    for _ in range(r.u32()):
        ab = r.arr("<u4", 3)
        vals = r.arr("<f4", JOINT_DESC_F)
        j = dict(a=int(ab[0]), b=int(ab[1]), flags=int(ab[2]), vals=vals)
        if (int(ab[2]) >> 8) & 0xff == GEAR_KIND:
            # Gear joints: bodies C, D, then ratio and constant.
            cd = r.arr("<u4", 2)
            j.update(c=int(cd[0]), d=int(cd[1]), extra=r.arr("<f4", 2))
        joints.append(j)
    return shapes, bodies, joints


def read_conf(r):
    nf = r.u32()
    f = r.arr("<f4", nf)
    vi = r.u32()
    pi = r.u32()
    flags = r.u32()
    nt = r.u32()
    tf = r.arr("<f4", nt)
    ti = r.arr("<u4", len(TOI_ITERS))
    return dict(floats=f, vel_iters=vi, pos_iters=pi, flags=flags, toi_floats=tf, toi_iters=ti)


# StepConf fields of the TOI phase, in record order.
TOI_FLOATS = ("toi_resolution_rate", "toi_min_separation", "toi_min_momentum", "target_depth", "tolerance")
TOI_ITERS = ("toi_velocity_iters", "toi_position_iters", "max_toi_root_iters", "max_toi_iters",
             "max_distance_iters", "max_sub_steps")


def conf_record(c):
    """The StepConf record (dump layout) of a playrho_amd.StepConf."""
    import struct
    floats = [getattr(c, f) for f, _ in c._fields_[:18]]
    flags = (1 if c.do_warm_start else 0) | (2 if c.do_toi else 0) | (4 if c.do_blocksolve else 0)
    return (struct.pack("<I", 18) + struct.pack("<18f", *floats) +
            struct.pack("<III", c.reg_velocity_iters, c.reg_position_iters, flags) +
            struct.pack("<I", len(TOI_FLOATS)) + struct.pack("<5f", *[getattr(c, f) for f in TOI_FLOATS]) +
            struct.pack("<6I", *[getattr(c, f) for f in TOI_ITERS]))


def read_state(r):
    step = r.u32()
    stats = r.arr("<u4", STATS)
    nb = r.u32()
    raw = r.arr("<u4", nb * (BODY_F + 1)).reshape(nb, BODY_F + 1)
    bodies = raw[:, :BODY_F].copy().view("<f4")
    bflags = raw[:, BODY_F].copy()
    nc = r.u32()
    per = CONTACT_U + CONTACT_T + CONTACT_F
    raw = r.arr("<u4", nc * per).reshape(nc, per)
    cu = raw[:, :CONTACT_U].copy()
    ct = raw[:, CONTACT_U:CONTACT_U + CONTACT_T].copy()
    cf = raw[:, CONTACT_U + CONTACT_T:].copy().view("<f4")
    npx = r.u32()
    raw = r.arr("<u4", npx * 5).reshape(npx, 5)
    pid = raw[:, 0].copy()
    paabb = raw[:, 1:].copy().view("<f4")
    nj = r.u32()
    raw = r.arr("<u4", nj * (1 + JOINT_F)).reshape(nj, 1 + JOINT_F)
    jlimit = raw[:, 0].copy()
    jimp = raw[:, 1:].copy().view("<f4")
    # Contact listener calls of the step, in call order: kind, contact id,
    # then EVENT_PAYLOAD[kind] words.
    ne = r.u32()
    words = []
    for _ in range(ne):
        kind = r.u32()
        cid = r.u32()
        words += [kind, cid] + list(r.arr("<u4", EVENT_PAYLOAD[kind]))
    events = np.array(words, dtype=np.uint32)
    return dict(step=step, stats=stats, bodies=bodies, bflags=bflags,
                cu=cu, ct=ct, cf=cf, pid=pid, paabb=paabb, jlimit=jlimit, jimp=jimp,
                nevents=ne, events=events)


def iter_events(state):
    """(kind, contact id, payload words) of each listener call in a state."""
    w = state["events"]
    o = 0
    for _ in range(state["nevents"]):
        kind = int(w[o])
        n = EVENT_PAYLOAD[kind]
        yield kind, int(w[o + 1]), w[o + 2:o + 2 + n]
        o += 2 + n


def read_dump(buf):
    r = Reader(buf)
    assert bytes(r.b[0:4]) == b"PRXD", "bad magic"
    r.o = 4
    assert r.u32() == 3
    shapes, bodies, joints = read_description(r)
    conf = read_conf(r)
    states = [read_state(r) for _ in range(r.u32())]
    return dict(shapes=shapes, bodies=bodies, joints=joints, conf=conf, states=states)


def digest(state, parts=("bodies", "contacts", "proxies", "joints", "events", "stats")):
    """SHA-256 over the canonical bytes of one state (selected parts)."""
    h = hashlib.sha256()
    if "bodies" in parts:
        h.update(np.ascontiguousarray(state["bodies"]).view("<u4").tobytes())
        h.update(np.ascontiguousarray(state["bflags"]).tobytes())
    if "contacts" in parts:
        h.update(np.ascontiguousarray(state["cu"]).tobytes())
        h.update(np.ascontiguousarray(state["ct"]).tobytes())
        h.update(np.ascontiguousarray(state["cf"]).view("<u4").tobytes())
    if "proxies" in parts:
        h.update(np.ascontiguousarray(state["pid"]).tobytes())
        h.update(np.ascontiguousarray(state["paabb"]).view("<u4").tobytes())
    if "joints" in parts:
        h.update(np.ascontiguousarray(state["jlimit"]).tobytes())
        h.update(np.ascontiguousarray(state["jimp"]).view("<u4").tobytes())
    if "events" in parts:
        h.update(np.uint32(state["nevents"]).tobytes())
        h.update(np.ascontiguousarray(state["events"]).tobytes())
    if "stats" in parts:
        h.update(np.ascontiguousarray(state["stats"]).tobytes())
    return h.digest()
