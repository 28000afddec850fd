"""Bitwise comparison of the soft step's outputs against the reference's fixtures.

Compares only what the reference's soft step writes: manifold columns on the solved
rows, per-point columns on the points each manifold has, the K block on manifolds of
two or more points, rolling masses on manifolds of at most two, and on the pair list
the entries store() keeps. Unused slots of the reference's arrays hold whatever an
earlier step left there, which nothing reads.
"""
import numpy as np

PER_POINT = {"normalImpulse": 1, "maxNormalImpulse": 1, "relativeVelocity": 1, "now": 1, "adjusted": 1,
             "normalMass": 1, "lever": 1, "ra": 3, "rb": 3, "la": 3, "lb": 3, "jac": 12}
PER_MANIFOLD = {"gap": 1, "tangentImpulse1": 1, "tangentImpulse2": 1, "spinImpulse": 1, "drift1": 1, "drift2": 1,
                "sliding": 1, "anchorA": 3, "anchorB": 3, "fjac": 24, "tangentMass": 3, "spinMass": 1}
ROLL = ["rollImpulse1", "rollImpulse2", "rollMass1", "rollMass2"]
BODY = ["position", "rotation", "velocity", "angular", "tumbled", "ceiling", "invI", "start", "startQ"]


def bits(a):
    a = np.ascontiguousarray(a)
    return a.view(np.uint64) if a.dtype == np.float64 else a


def first_diff(name, got, want, index=None):
    g, w = bits(got), bits(want)
    if g.shape != w.shape:
        return f"{name}: shape {g.shape} != {w.shape}"
    bad = np.nonzero(g != w)[0]
    if len(bad) == 0:
        return None
    i = bad[0]
    where = index[i] if index is not None else i
    return f"{name}: {len(bad)} differ, first at {where}: got {got[i]!r} want {want[i]!r}"


def point_index(points, count, per, solved):
    idx = []
    for m in range(solved):
        for k in range(points[m]):
            base = (4 * m + k) * per
            idx.extend(range(base, base + per))
    return np.array(idx, dtype=np.int64)


def compare(fx, got):
    """Differences between `got` (columns by fixture name: b.*, c.*, p.*) and the fixture's out.*; [] if none."""
    meta = fx.meta
    solved, count = meta["solved"], meta["count"]
    points = fx["in.c.points"].astype(np.int64)
    errs = []
    for name in BODY:
        e = first_diff("body." + name, got["b." + name], fx["out." + name])
        if e:
            errs.append(e)
    for name, per in PER_POINT.items():
        idx = point_index(points, count, per, solved)
        if len(idx):
            e = first_diff("c." + name, got["c." + name][idx], fx["out.c." + name][idx], idx)
            if e:
                errs.append(e)
    for name, per in PER_MANIFOLD.items():
        n = solved * per
        e = first_diff("c." + name, got["c." + name][:n], fx["out.c." + name][:n])
        if e:
            errs.append(e)
    rows = [m for m in range(solved) if points[m] <= 2]
    for name in ROLL:
        e = first_diff("c." + name, got["c." + name][rows], fx["out.c." + name][rows], rows)
        if e:
            errs.append(e)
    bidx = np.array([16 * m + 4 * k + l for m in range(solved) if points[m] >= 2
                     for k in range(points[m]) for l in range(points[m])], dtype=np.int64)
    if len(bidx):
        e = first_diff("c.block", got["c.block"][bidx], fx["out.c.block"][bidx], bidx)
        if e:
            errs.append(e)
    pair = fx["in.c.pair"]
    pidx = [pair[m] for m in range(solved)]
    e = first_diff("pairs.points", got["p.points"][pidx], fx["out.pairs.points"][pidx], pidx)
    if e:
        errs.append(e)
    for name in ("twist", "friction", "slip"):
        idx = np.array([3 * p + k for p in pidx for k in range(3)], dtype=np.int64)
        if len(idx):
            e = first_diff("pairs." + name, got["p." + name][idx], fx["out.pairs." + name][idx], idx)
            if e:
                errs.append(e)
    idx = np.array([4 * pair[m] + k for m in range(solved) for k in range(points[m])], dtype=np.int64)
    for name in ("pid", "normalImpulse"):
        if len(idx):
            e = first_diff("pairs." + name, got["p." + name][idx], fx["out.pairs." + name][idx], idx)
            if e:
                errs.append(e)
    return errs


def snake(name):
    out = []
    for ch in name:
        if ch.isupper():
            out.append("_" + ch.lower())
        else:
            out.append(ch)
    return "".join(out)


def world_from_columns(fx, cols):
    """A bm_columns world dict (bearing_amd.native.Batch) from oracle-style columns of a fixture."""
    meta = fx.meta
    w = {}
    for key, arr in cols.items():
        if key.startswith("b."):
            w[snake(key[2:])] = arr
        elif key.startswith("c."):
            w[snake(key[2:])] = arr
        elif key.startswith("p."):
            w["pair_" + snake(key[2:])] = arr
    w["colour"] = fx["in.c.colour"].astype(np.int32).copy()
    w["order"] = fx["in.c.order"].astype(np.int32).copy()
    w["awake_list"] = cols["awake_list"]
    w["starts"] = cols["starts"]
    w["gravity"] = cols["gravity"]
    w["solved"] = np.int32(meta["solved"])
    w["substeps"] = np.int32(meta["substeps"])
    w["dt"] = np.float64(meta["dt"])
    return w


def columns_from_world(w):
    """The inverse of world_from_columns for the output columns: names as compare() reads them."""
    from oracle.binding import BODY, MANIFOLD, PAIR
    out = {}
    for _, _, name, _ in BODY:
        out["b." + name] = w[snake(name)]
    for _, _, name, _ in MANIFOLD:
        out["c." + name] = w[snake(name)]
    for _, _, name, _ in PAIR:
        out["p." + name] = w["pair_" + snake(name)]
    return out


# The columns narrow() writes, which colour() moves to their rows (src/graph.js:83-91).
ARRANGED = ["a", "b", "pair", "sub", "points", "colour", "pid", "normal", "t1", "t2", "friction", "restitution",
            "tolerance", "patch", "roll_impulse1", "roll_impulse2", "spin_impulse", "tangent_impulse1",
            "tangent_impulse2", "drift1", "drift2", "separation", "normal_impulse", "point", "shift"]


def narrow_order(world, seed=0):
    """A laid-out world dict back in the order narrow() found its manifolds (row m of the
    result is found manifold m), with the layout columns and rest()'s outputs scrambled."""
    rng = np.random.default_rng(seed)
    row = np.asarray(world["row"], np.int64)
    n = len(row)
    out = dict(world)
    for name in ARRANGED:
        col = np.asarray(world[name])
        per = col.size // n if n else 1
        out[name] = col.reshape(n, per)[row].reshape(-1).copy()
    for name in ("row", "order"):
        out[name] = rng.integers(-5, n + 5, n).astype(np.int32)
    out["starts"] = rng.integers(0, n + 1, 35).astype(np.int32)
    out["solved"] = np.int32(rng.integers(0, n + 1))
    for name in ("max_normal_impulse", "relative_velocity", "gap"):
        out[name] = rng.normal(size=np.asarray(world[name]).size)
    return out
