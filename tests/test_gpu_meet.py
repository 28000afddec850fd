"""The device meet (bm_meet_host: collide, time of impact, collide again, carry back)
against the oracle's restatement of src/contact.js meet, bit for bit, on the reference's
time-of-impact pairs and on fresh seeded ones. The oracle's meet is itself pinned by the
reference's time-of-impact vectors (test_oracle_impact.py) and by whole trajectories with
thousands of impacts (test_host_world.py)."""
import os

import numpy as np
import pytest

import fixtures
from bearing_amd import native
from oracle import binding

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = native.Context(0)
    yield c
    c.close()


def bits(a):
    a = np.asarray(a)
    return a.view(np.uint64) if a.dtype == np.float64 else a


def compare(got, want):
    names = ["count", "normal", "points", "separations", "ids", "shift", "flags", "loss"]
    count = want[0]
    mask = {"points": 12, "separations": 4, "ids": 4}
    for name, g, w in zip(names, got, want):
        g, w = np.asarray(g), np.asarray(w)
        if name in ("normal", "shift"):
            # Only a pair with points has a normal or a shift to read.
            keep = np.repeat(count > 0, g.size // count.size)
            g, w = g[keep], w[keep]
        if name in mask:
            per = mask[name] // 4
            keep = (np.arange(g.size) % mask[name]) < np.repeat(np.maximum(count, 0) * per, mask[name])
            g, w = g[keep], w[keep]
        assert np.array_equal(bits(g), bits(w)), f"{name} differs at {np.flatnonzero(bits(g) != bits(w))[:5]}"


def pairs_from_fixture(fx, seed):
    rng = np.random.default_rng(seed)
    b = np.asarray(fx["bodies"]).reshape(-1, 2, 32)
    n = len(b)
    margin = np.asarray(fx["tolerance"]) + rng.uniform(0, 0.3, n) * rng.integers(0, 2, n)
    pair = np.column_stack([margin, fx["tolerance"], fx["dt"], np.asarray(fx["gravity"]).reshape(-1, 3)])
    return b[:, 0].copy(), b[:, 1].copy(), pair


def test_gpu_meet_matches_oracle_on_reference_pairs(ctx):
    fx = fixtures.load(os.path.join(fixtures.DATA, "impact.2026.bmfx.gz"))
    a, b, pair = pairs_from_fixture(fx, 1)
    want = binding.meet(a, b, pair)
    got = native.meet(ctx, a, b, pair)
    compare(got, want)
    flags = want[6]
    assert (flags & 1).sum() > 50 and (flags & 2).sum() > 200, "too few impacts to test"


@pytest.mark.parametrize("seed", [3, 4])
def test_gpu_meet_matches_oracle_on_fresh_pairs(ctx, seed):
    fx = fixtures.load(os.path.join(fixtures.DATA, "impact.2026.bmfx.gz"))
    a, b, pair = pairs_from_fixture(fx, seed)
    rng = np.random.default_rng(seed)
    # Jiggle velocities and spins so the pairs are new.
    for r in (a, b):
        r[:, 23:29] *= rng.uniform(0.5, 1.5, (len(r), 6))
    want = binding.meet(a, b, pair)
    got = native.meet(ctx, a, b, pair)
    compare(got, want)


def test_gpu_meet_packed_is_the_found_pairs_of_the_oracle(ctx):
    # bm_meet_packed on the reference's time-of-impact pairs, each body given once in a table
    # (pair k: rows 2k and 2k + 1), shuffled so a row is met out of order: every pair's count
    # and flags, and the pairs with points packed in pair order, against the oracle's meet.
    fx = fixtures.load(os.path.join(fixtures.DATA, "impact.2026.bmfx.gz"))
    a, b, pair = pairs_from_fixture(fx, 5)
    n = len(pair)
    perm = np.random.default_rng(5).permutation(2 * n)
    table = np.empty((2 * n, 32))
    table[perm[0::2]] = a
    table[perm[1::2]] = b
    count, flags, row, normal, points, seps, ids, shift, loss = native.meet_packed(
        ctx, table, perm[0::2], perm[1::2], pair)
    want = binding.meet(a, b, pair)
    assert np.array_equal(count, want[0]) and np.array_equal(flags, want[6])
    has = np.flatnonzero(count > 0)
    assert has.size > 100 and has.size < n, "too few or too many pairs found to test"
    assert np.array_equal(row[has], np.arange(has.size)) and (row[count <= 0] == -1).all()
    for got, w, width in ((normal, want[1], 3), (shift, want[5], 12), (loss, want[7], 1)):
        assert np.array_equal(bits(got), bits(np.asarray(w).reshape(n, width)[has].ravel()))
    for got, w, per in ((points, want[2], 3), (seps, want[3], 1), (ids, want[4], 1)):
        g = np.asarray(got).reshape(has.size, 4 * per)
        ww = np.asarray(w).reshape(n, 4 * per)[has]
        for r, k in enumerate(has):
            used = min(count[k], 4) * per
            assert np.array_equal(bits(g[r, :used]), bits(ww[r, :used])), (k, r)
