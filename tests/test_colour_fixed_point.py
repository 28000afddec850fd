"""The device colouring's formulation, checked on the CPU: rounds in which every manifold
whose earlier neighbours (sharing a moving body) all have colours takes the lowest colour
none of them holds give the same colours as the reference's serial greedy pass
(src/graph.js:114-136, restated by the oracle), for random graphs with frozen and static
bodies and overflowing hubs. The rows' order is then the sort of the unique key
(colour, b, m), which is what the reference's stable counting sorts produce."""
import numpy as np
import pytest

from oracle import binding


def fixed_point(a, b, type_, frozen):
    n = len(a)
    moving = (type_ == 0) & (frozen == 0)
    still = (frozen != 0) | (type_ == 1)
    colour = np.full(n, -1, np.int64)
    resting = still[a] & still[b]
    colour[resting] = 33
    lists = {}
    for m in range(n):
        if resting[m]:
            continue
        for s in (a[m], b[m]):
            if moving[s]:
                lists.setdefault(int(s), []).append(m)
    rounds = 0
    while (colour == -1).any():
        rounds += 1
        snapshot = colour.copy()
        for m in np.nonzero(snapshot == -1)[0]:
            held, ready = 0, True
            for s in (a[m], b[m]):
                if not moving[s]:
                    continue
                for o in lists[int(s)]:
                    if o >= m:
                        continue
                    if snapshot[o] == -1:
                        ready = False
                        break
                    if snapshot[o] < 32:
                        held |= 1 << int(snapshot[o])
                if not ready:
                    break
            if ready:
                colour[m] = 32 if held == 0xFFFFFFFF else (~held & (held + 1)).bit_length() - 1
    keys = [(int(colour[m]), int(b[m]), m) for m in range(n)]
    order = np.array([k[2] for k in sorted(keys)], np.int32)
    return colour.astype(np.int32), order, rounds


@pytest.mark.parametrize("seed", range(6))
def test_fixed_point_equals_serial_greedy(seed):
    rng = np.random.default_rng(seed)
    bodies = int(rng.integers(20, 200))
    n = int(rng.integers(10, 800))
    type_ = np.zeros(bodies, np.uint8)
    type_[: int(rng.integers(1, 4))] = 1
    frozen = ((rng.random(bodies) < 0.15) & (type_ == 0)).astype(np.uint8)
    hub = int(rng.integers(5, bodies))
    a = rng.integers(0, bodies, n)
    b = np.where(rng.random(n) < 0.2, hub, rng.integers(0, bodies, n))
    lo, hi = np.minimum(a, b).astype(np.int32), np.maximum(a, b).astype(np.int32)
    keep = lo != hi
    lo, hi = lo[keep], hi[keep]
    col, order, _, _, _ = binding.colour(lo, hi, bodies, type_, frozen, int(frozen.sum()))
    fcol, forder, rounds = fixed_point(lo, hi, type_, frozen)
    assert np.array_equal(fcol, col)
    assert np.array_equal(forder, order)
    assert rounds <= len(lo)
