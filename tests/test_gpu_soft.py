"""The soft step on the MI355X against the reference's fixtures and the oracle: bit for bit.

Each fixture is the reference's own state before inertias() and after store()
(tests/golden); the HIP kernel runs the phase from the same state, one world per
workgroup, and every column it writes must have the reference's bits. The same worlds
are also run as one batch, to check the offset tables, and against the oracle on
copies of the fixtures' worlds with their velocities scrambled (seeded), states the
reference was never asked about.
"""
import numpy as np
import pytest

import fixtures
from bearing_amd import native
from oracle import binding
from softcheck import ARRANGED, columns_from_world, compare, narrow_order, world_from_columns

pytestmark = pytest.mark.gpu

SOFT = fixtures.find("soft")


@pytest.fixture(scope="module")
def ctx():
    c = native.Context(0)
    yield c
    c.close()


def run_world(ctx, fx, cols):
    batch = native.Batch([world_from_columns(fx, cols)])
    batch.step_host(ctx)
    return columns_from_world(batch.world(0))


@pytest.mark.parametrize("path", SOFT, ids=[p.split("/")[-1].replace(".bmfx.gz", "") for p in SOFT])
def test_gpu_soft_step_matches_reference(ctx, path):
    print("fixture", path, flush=True)
    fx = fixtures.load(path)
    got = run_world(ctx, fx, binding.columns(fx))
    errs = compare(fx, got)
    assert not errs, "\n".join(errs)


@pytest.mark.parametrize("path", SOFT, ids=[p.split("/")[-1].replace(".bmfx.gz", "") for p in SOFT])
def test_gpu_grid_wide_soft_step_matches_reference(ctx, path):
    # Every fixture's world stepped by the whole grid (bm_set_wide), not one workgroup: the
    # same bits.
    print("fixture, grid-wide", path, flush=True)
    fx = fixtures.load(path)
    ctx.set_wide(1)
    try:
        got = run_world(ctx, fx, binding.columns(fx))
    finally:
        ctx.set_wide(16384)
    errs = compare(fx, got)
    assert not errs, "\n".join(errs)


def test_gpu_batch_of_every_fixture(ctx):
    fxs = [fixtures.load(p) for p in SOFT]
    batch = native.Batch([world_from_columns(fx, binding.columns(fx)) for fx in fxs])
    batch.step_host(ctx)
    for i, fx in enumerate(fxs):
        errs = compare(fx, columns_from_world(batch.world(i)))
        assert not errs, f"{SOFT[i]}: " + "\n".join(errs)


@pytest.mark.parametrize("shape,resident", [(s, r) for s in (0, 1, 2) for r in (True, False)])
def test_gpu_every_launch_shape_gives_the_same_bits(ctx, shape, resident):
    fxs = [fixtures.load(p) for p in SOFT]
    batch = native.Batch([world_from_columns(fx, binding.columns(fx)) for fx in fxs])
    ctx.set_launch(shape, resident)
    try:
        batch.step_host(ctx)
    finally:
        ctx.set_launch(0, True)
    for i, fx in enumerate(fxs):
        errs = compare(fx, columns_from_world(batch.world(i)))
        assert not errs, f"shape {shape} resident {resident} {SOFT[i]}: " + "\n".join(errs)


def scrambled(fx, seed):
    """The fixture's input with every dynamic body's velocities changed by seeded noise."""
    rng = np.random.default_rng(seed)
    cols = binding.columns(fx)
    dyn = np.nonzero(cols["b.invMass"] > 0)[0]
    for name, scale in (("b.velocity", 0.5), ("b.angular", 1.0)):
        v = cols[name].reshape(-1, 3)
        v[dyn] += rng.normal(0, scale, size=(len(dyn), 3))
    return cols


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_gpu_matches_oracle_on_scrambled_worlds(ctx, seed):
    fxs = [fixtures.load(p) for p in SOFT]
    inputs = [scrambled(fx, seed * 100 + i) for i, fx in enumerate(fxs)]
    want = [binding.soft_step(fx, {k: v.copy() for k, v in c.items()}) for fx, c in zip(fxs, inputs)]
    batch = native.Batch([world_from_columns(fx, c) for fx, c in zip(fxs, inputs)])
    batch.step_host(ctx)
    for i, fx in enumerate(fxs):
        got = columns_from_world(batch.world(i))
        for key in ("b.position", "b.rotation", "b.velocity", "b.angular", "c.normalImpulse", "c.tangentImpulse1",
                    "c.spinImpulse", "p.twist", "p.slip"):
            g, w = got[key], want[i][key]
            same = np.array_equal(g.view(np.uint64) if g.dtype == np.float64 else g,
                                  w.view(np.uint64) if w.dtype == np.float64 else w)
            assert same, f"{SOFT[i]} seed {seed}: {key} differs from the oracle"


@pytest.mark.parametrize("path", SOFT, ids=[p.split("/")[-1].replace(".bmfx.gz", "") for p in SOFT])
@pytest.mark.parametrize("in_lds", [True, False, "stream"], ids=["lds", "hbm", "stream"])
def test_gpu_colour_then_soft_step_matches_reference(ctx, path, in_lds):
    """From the manifolds in the order narrow() found them: bm_colour must lay the rows out
    as the reference's colour() / arrange() / rest() did, and the soft step after it must
    end where the reference's did; both ways of colouring."""
    print("fixture", path, in_lds, flush=True)
    ctx.set_colour_lds(in_lds)
    fx = fixtures.load(path)
    laid = world_from_columns(fx, binding.columns(fx))
    batch = native.Batch([narrow_order(laid, seed=len(path))])
    batch.upload(ctx)
    batch.colour(ctx)
    batch.download(ctx)
    got = batch.world(0)
    n = fx.meta["count"]
    for name in ARRANGED + ["row", "order"]:
        g, w = np.asarray(got[name]), np.asarray(laid[name])
        assert np.array_equal(g.view(np.uint64) if g.dtype == np.float64 else g,
                              w.view(np.uint64) if w.dtype == np.float64 else w), f"{name} laid out differently"
    assert np.array_equal(batch.arrays["starts"], fx["in.starts"])
    assert int(batch.arrays["solved"][0]) == fx.meta["solved"]
    solved = fx.meta["solved"]
    pts = laid["points"]
    for name in ("max_normal_impulse", "relative_velocity"):
        idx = [4 * m + k for m in range(solved, n) for k in range(pts[m])]
        assert np.array_equal(np.asarray(got[name])[idx].view(np.uint64), np.asarray(laid[name])[idx].view(np.uint64))
    assert np.array_equal(np.asarray(got["gap"])[solved:n].view(np.uint64), np.asarray(laid["gap"])[solved:n].view(np.uint64))
    batch.step(ctx)
    batch.download(ctx)
    errs = compare(fx, columns_from_world(batch.world(0)))
    assert not errs, "\n".join(errs)


@pytest.mark.parametrize("in_lds", [True, False], ids=["lds", "hbm"])
def test_gpu_colour_of_a_batch(ctx, in_lds):
    ctx.set_colour_lds(in_lds)
    fxs = [fixtures.load(p) for p in SOFT]
    laid = [world_from_columns(fx, binding.columns(fx)) for fx in fxs]
    batch = native.Batch([narrow_order(w, seed=i) for i, w in enumerate(laid)])
    batch.upload(ctx)
    batch.colour(ctx)
    batch.step(ctx)
    batch.download(ctx)
    for i, fx in enumerate(fxs):
        errs = compare(fx, columns_from_world(batch.world(i)))
        assert not errs, f"{SOFT[i]}: " + "\n".join(errs)
