"""Compares touchLevel results (manifolds a shape, then count, normal, points, separations,
ids in `cap` slots a shape) with the reference's, bit for bit."""
import numpy as np


def _bits(x):
    return np.ascontiguousarray(x, np.float64).view(np.uint64)


def compare_level(fx, manifolds, count, normal, points, seps, ids, limit=10):
    """The first `limit` differences from the fixture, as messages."""
    errs = []
    cap = fx.meta["cap"]
    want_m = fx["manifolds"]
    for k in np.nonzero(manifolds != want_m)[0][:limit]:
        errs.append(f"case {k}: {manifolds[k]} manifolds, reference {want_m[k]}")
    checks = [("count", count, fx["count"], 1, False), ("normal", normal, fx["normal"], 3, True),
              ("points", points, fx["points"], 12, True), ("separations", seps, fx["separations"], 4, True),
              ("ids", ids, fx["ids"], 4, False)]
    for name, got, want, per, real in checks:
        g = _bits(got) if real else np.asarray(got)
        w = _bits(want) if real else np.asarray(want)
        bad = np.nonzero((g != w).reshape(-1, per * cap).any(axis=1))[0]
        for k in bad[:limit]:
            errs.append(f"case {k}: {name} differs")
        if len(errs) >= limit:
            break
    return errs[:limit]
