"""The oracle's colouring against colour() as the reference ran it (tests/golden colour.*),
and against the invariants the reference's own tests hold (test/contact.test.js:186-216):
no colour shares a dynamic body, and the rows of a colour go by b's slot."""
import numpy as np
import pytest

import fixtures
from oracle import binding

COLOUR = fixtures.find("colour")


@pytest.mark.parametrize("path", COLOUR, ids=[p.split("/")[-1].replace(".bmfx.gz", "") for p in COLOUR])
def test_colour_matches_reference(path):
    fx = fixtures.load(path)
    n = fx.meta["count"]
    col, order, row, starts, solved = binding.colour(fx["in.a"], fx["in.b"], fx.meta["top"], fx["in.type"],
                                                     fx["in.frozen"], fx.meta["frozenCount"])
    # arrange() moves the colour column with the rows: row o holds manifold order[o].
    assert np.array_equal(col[order], fx["out.colour"][:n])
    assert np.array_equal(order, fx["out.order"][:n])
    assert np.array_equal(row, fx["out.row"][:n])
    assert np.array_equal(starts, fx["out.starts"])
    assert solved == fx.meta["solved"]


def random_graph(seed, bodies=300, manifolds=1500, statics=3, frozen_share=0.0):
    rng = np.random.default_rng(seed)
    type_ = np.zeros(bodies, np.uint8)
    type_[:statics] = 1
    frozen = (rng.random(bodies) < frozen_share).astype(np.uint8)
    frozen[:statics] = 0
    a = rng.integers(0, bodies, manifolds)
    b = rng.integers(0, bodies, manifolds)
    lo, hi = np.minimum(a, b), np.maximum(a, b)
    keep = lo != hi
    return lo[keep].astype(np.int32), hi[keep].astype(np.int32), type_, frozen


@pytest.mark.parametrize("seed,frozen_share", [(1, 0.0), (2, 0.2), (3, 0.0)])
def test_colour_invariants(seed, frozen_share):
    a, b, type_, frozen = random_graph(seed, frozen_share=frozen_share)
    col, order, row, starts, solved = binding.colour(a, b, len(type_), type_, frozen, int(frozen.sum()))
    moves = (type_ == 0) & (frozen == 0)
    for c in range(32):
        ms = np.nonzero(col == c)[0]
        ends = np.concatenate([a[ms][moves[a[ms]]], b[ms][moves[b[ms]]]])
        assert len(ends) == len(np.unique(ends)), f"colour {c} shares a moving body"
        rows = order[starts[c]:starts[c + 1]]
        assert np.all(np.diff(b[rows]) >= 0)
    assert np.array_equal(row[order], np.arange(len(a)))
    resting = (~moves[a]) & (~moves[b])
    assert np.array_equal(col == 33, resting if frozen_share > 0 else np.zeros_like(resting))


def test_overflow_when_a_body_holds_33_colours():
    # One dynamic body touching 40 others, each of which first touches 31 of its own:
    # a hub's manifolds take the lowest colour neither end holds, until all 32 are held.
    a, b = [], []
    hub, nxt = 0, 1
    for k in range(40):
        a.append(hub); b.append(nxt); nxt += 1
    a, b = np.array(a, np.int32), np.array(b, np.int32)
    type_ = np.zeros(nxt, np.uint8)
    col, order, row, starts, solved = binding.colour(a, b, nxt, type_, np.zeros(nxt, np.uint8), 0)
    assert list(col[:32]) == list(range(32))
    assert np.all(col[32:] == 32)
    assert starts[33] - starts[32] == 8 and solved == 40
