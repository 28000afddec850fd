"""Comparison of narrowphase results against the reference's collide() vectors, bit for bit."""
import numpy as np


def compare_manifolds(fx, count, normal, points, separations, ids):
    """Differences from the fixture, [] if none. Normals and points are compared on the
    pairs and point slots the reference filled."""
    errs = []
    want = fx["count"]
    bad = np.nonzero(count != want)[0]
    if len(bad):
        errs.append(f"count differs on {len(bad)} pairs, first {bad[0]}: {count[bad[0]]} vs {want[bad[0]]}")
    hit = np.nonzero((count == want) & (want > 0))[0]
    b = lambda x: np.ascontiguousarray(x).view(np.uint64)
    nbad = [k for k in hit if not np.array_equal(b(normal[3 * k:3 * k + 3]), b(fx["normal"][3 * k:3 * k + 3]))]
    if nbad:
        errs.append(f"normal differs on {len(nbad)} pairs, first {nbad[0]}")
    for name, got, per in (("points", points, 3), ("separations", separations, 1), ("ids", ids, 1)):
        wantv = fx[name]
        bad = []
        for k in hit:
            n = int(want[k])
            g, w = got[4 * per * k:4 * per * k + per * n], wantv[4 * per * k:4 * per * k + per * n]
            same = np.array_equal(b(g), b(w)) if g.dtype == np.float64 else np.array_equal(g, w)
            if not same:
                bad.append(k)
        if bad:
            errs.append(f"{name} differ on {len(bad)} pairs, first {bad[0]}")
    return errs
