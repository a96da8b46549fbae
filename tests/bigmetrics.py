"""Quality measures of a large world's state, for judging the graph-coloured
solve (which does not keep the reference's constraint order, so not its bits)
against the reference's exact solve on the same world.

From a world description (dumpfmt.read_description of the save_desc bytes)
and a parsed state record:
* escaped: dynamic bodies outside the container (with a 1 m margin);
* separations of every point of every touching contact at the state's poses
  (PositionSolverManifold.cpp:42-121 in float64): minimum, 1st and 5th
  percentiles, median, mean of the ten deepest, count below -5 cm;
* linear momentum sum(m v), angular momentum about the origin
  sum(m (c x v) + I w) with c the body's centre of mass, kinetic energy;
* mean height of the contained dynamic bodies' centres.
"""
import numpy as np

from . import dumpfmt


class WorldInfo:
    def __init__(self, desc_bytes):
        shapes, bodies, _ = dumpfmt.read_description(dumpfmt.Reader(desc_bytes))
        self.radius = np.array([float(s["radius"]) for s in shapes])
        self.dyn = np.array([(b["flags"] & 2) != 0 for b in bodies])
        vals = np.array([b["vals"] for b in bodies], dtype=np.float64)
        self.lc = vals[:, 6:8]
        inv_m, inv_i = vals[:, 16], vals[:, 17]
        self.mass = np.where(inv_m > 0, 1.0 / np.where(inv_m > 0, inv_m, 1.0), 0.0)
        self.inertia = np.where(inv_i > 0, 1.0 / np.where(inv_i > 0, inv_i, 1.0), 0.0)
        # Proxy (tree leaf) ids in creation order: body order, shape order.
        lb, ls = [], []
        for bi, b in enumerate(bodies):
            if not (b["flags"] & 32):
                continue
            for sid in b["shapes"]:
                lb.append(bi)
                ls.append(int(sid))
        self.leaf_body = {}
        self.leaf_shape = {}
        for k, (bi, si) in enumerate(zip(lb, ls)):
            leaf = 0 if k == 0 else 2 * k - 1
            self.leaf_body[leaf] = bi
            self.leaf_shape[leaf] = si


def separations(info, state):
    pos = state["bodies"][:, 3:6].astype(np.float64)
    c, s = np.cos(pos[:, 2]), np.sin(pos[:, 2])
    cu, ct, cf = state["cu"], state["ct"], state["cf"].astype(np.float64)
    keep = ((cu[:, 3] & 1) != 0) & ((cu[:, 3] & 4) == 0)
    cu, ct, cf = cu[keep], ct[keep], cf[keep]
    if len(cu) == 0:
        return np.zeros(0)
    ba = np.array([info.leaf_body[int(k)] for k in cu[:, 0]])
    bb = np.array([info.leaf_body[int(k)] for k in cu[:, 1]])
    tr = np.array([info.radius[info.leaf_shape[int(a)]] + info.radius[info.leaf_shape[int(b)]]
                   for a, b in zip(cu[:, 0], cu[:, 1])])
    mtype, npts = ct[:, 0] & 0xff, ct[:, 0] >> 8

    def rot(bi, v):
        return np.stack([c[bi] * v[:, 0] - s[bi] * v[:, 1], s[bi] * v[:, 0] + c[bi] * v[:, 1]], axis=1)

    def xf(bi, v):
        return pos[bi, :2] - rot(bi, info.lc[bi]) + rot(bi, v)

    out = []
    ln, lpt = cf[:, 2:4], cf[:, 4:6]
    for i in range(2):
        has = npts > i
        lp = cf[:, 6 + 4 * i:8 + 4 * i]
        circ = has & (mtype == 1)
        fa = has & (mtype == 2)
        fb = has & (mtype == 3)
        if circ.any():
            d = xf(bb[circ], lp[circ]) - xf(ba[circ], lpt[circ])
            out.append(np.linalg.norm(d, axis=1) - tr[circ])
        if fa.any():
            n = rot(ba[fa], ln[fa])
            out.append(np.sum((xf(bb[fa], lp[fa]) - xf(ba[fa], lpt[fa])) * n, axis=1) - tr[fa])
        if fb.any():
            n = rot(bb[fb], ln[fb])
            out.append(np.sum((xf(ba[fb], lp[fb]) - xf(bb[fb], lpt[fb])) * n, axis=1) - tr[fb])
    return np.sort(np.concatenate(out)) if out else np.zeros(0)


def measures(info, state, half_width=200.0, height=400.0):
    b = state["bodies"].astype(np.float64)
    cx, cy, vx, vy, w = b[:, 3], b[:, 4], b[:, 6], b[:, 7], b[:, 8]
    dyn = info.dyn
    inside = dyn & (np.abs(cx) < half_width + 1.0) & (cy > -1.0) & (cy < height + 1.0)
    m, inertia = info.mass, info.inertia
    sep = separations(info, state)
    out = {
        "step": int(state["step"]),
        "dynamic": int(dyn.sum()),
        "escaped": int(dyn.sum() - inside.sum()),
        "touching_points": int(sep.size),
        "sep_min": float(sep[0]) if sep.size else 0.0,
        "sep_p01": float(np.percentile(sep, 1)) if sep.size else 0.0,
        "sep_p05": float(np.percentile(sep, 5)) if sep.size else 0.0,
        "sep_p50": float(np.percentile(sep, 50)) if sep.size else 0.0,
        "sep_deepest10": float(sep[:10].mean()) if sep.size >= 10 else 0.0,
        "sep_below_5cm": int(np.sum(sep < -0.05)),
        "momentum": [float(np.sum(m[dyn] * vx[dyn])), float(np.sum(m[dyn] * vy[dyn]))],
        "angular_momentum": float(np.sum(m[dyn] * (cx[dyn] * vy[dyn] - cy[dyn] * vx[dyn]) + inertia[dyn] * w[dyn])),
        "kinetic_energy": float(np.sum(0.5 * m[dyn] * (vx[dyn] ** 2 + vy[dyn] ** 2) + 0.5 * inertia[dyn] * w[dyn] ** 2)),
        "mean_height": float(cy[inside].mean()) if inside.any() else 0.0,
        "total_mass": float(m[dyn].sum()),
    }
    return out
