"""The oracle's soft step against the reference's own (tests/golden fixtures): bit for bit.

Each fixture holds the reference's world and manifolds as contact.js inertias() received
them and as store() left them (src/world.js:961-1031); the oracle runs the same phase
from the same state and must leave every written column with the same bits.
"""
import numpy as np
import pytest

import fixtures
from oracle import binding
from softcheck import compare

SOFT = fixtures.find("soft")


@pytest.mark.parametrize("path", SOFT, ids=[p.split("/")[-1].replace(".bmfx.gz", "") for p in SOFT])
def test_soft_step_matches_reference(path):
    fx = fixtures.load(path)
    got = binding.soft_step(fx)
    errs = compare(fx, got)
    assert not errs, "\n".join(errs)


def test_fixtures_present():
    assert len(SOFT) >= 15


@pytest.mark.parametrize("column,index", [("c.separation", 0), ("c.normal", 4), ("b.velocity", 7)])
def test_a_planted_mistake_is_caught(column, index):
    """An input the phase always reads, changed by a part in 2^30, must show up in the comparison."""
    fx = fixtures.load(fixtures.find("soft", "pyramid")[1])
    cols = binding.columns(fx)
    a = cols[column]
    a[index] = a[index] * (1 + 2.0 ** -30) if a[index] != 0 else 2.0 ** -30
    got = binding.soft_step(fx, cols)
    assert compare(fx, got), f"an ulp in {column}[{index}] went unnoticed"


@pytest.mark.parametrize("workers", [1, 3, 7])
@pytest.mark.parametrize("path", SOFT, ids=[p.split("/")[-1].replace(".bmfx.gz", "") for p in SOFT])
def test_soft_step_with_workers_matches_reference(path, workers):
    """The colour-parallel soft step (oracle_soft_step_workers: prepare and each colour's warm
    start and solve cut across threads, as the reference's workers cut them) leaves the
    reference's bits, and the single thread's, whatever the number of workers."""
    fx = fixtures.load(path)
    one = binding.soft_step(fx)
    got = binding.soft_step(fx, workers=workers)
    errs = compare(fx, got)
    assert not errs, "\n".join(errs)
    for key in one:
        assert np.array_equal(one[key], got[key], equal_nan=True), key
