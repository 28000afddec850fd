"""The device narrowphase against the reference's collide() vectors (bit for bit) and against
the oracle on fresh seeded pairs the reference was never asked about."""
import glob

import numpy as np
import pytest

import fixtures
from bearing_amd import native
from collidecheck import compare_manifolds
from oracle import binding

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = native.Context(0)
    yield c
    c.close()


def test_gpu_collide_matches_reference(ctx):
    paths = sorted(glob.glob(fixtures.DATA + "/collide.*.bmfx.gz"))
    assert paths
    for path in paths:
        print("fixture", path, flush=True)
        fx = fixtures.load(path)
        sh = fx["shapes"].reshape(-1, 2, 19)
        got = native.collide(ctx, sh[:, 0].copy(), sh[:, 1].copy(), fx["margin"], fx["tolerance"])
        errs = compare_manifolds(fx, *got)
        assert not errs, "\n".join(errs)


@pytest.mark.parametrize("seed", [7, 8])
def test_gpu_collide_matches_oracle_on_fresh_pairs(ctx, seed):
    from make_fixtures import _axes, collide_pairs
    recs = collide_pairs(seed, 20000)
    shapes = np.zeros((len(recs), 2, 19))
    for t in range(2):
        r = recs[:, 14 * t:14 * (t + 1)]
        shapes[:, t, 0] = r[:, 0]
        shapes[:, t, 1:4] = r[:, 1:4]
        shapes[:, t, 4:13] = np.array([_axes(q) for q in r[:, 4:8]])
        shapes[:, t, 13:16] = r[:, 8:11]
        shapes[:, t, 16:19] = r[:, 11:14]
    margin, tol = recs[:, 28].copy(), recs[:, 29].copy()
    want = binding.collide(shapes[:, 0].copy(), shapes[:, 1].copy(), margin, tol)
    got = native.collide(ctx, shapes[:, 0].copy(), shapes[:, 1].copy(), margin, tol)
    fx = {"count": want[0], "normal": want[1], "points": want[2], "separations": want[3], "ids": want[4]}
    errs = compare_manifolds(fx, *got)
    assert not errs, "\n".join(errs)
    assert (want[0] > 0).sum() > 5000


def test_gpu_capsules_match_reference(ctx):
    paths = sorted(glob.glob(fixtures.DATA + "/collide_capsule.*.bmfx.gz"))
    assert paths
    for path in paths:
        print("fixture", path, flush=True)
        fx = fixtures.load(path)
        sh = fx["shapes"].reshape(-1, 2, 19)
        got = native.collide(ctx, sh[:, 0].copy(), sh[:, 1].copy(), fx["margin"], fx["tolerance"])
        errs = compare_manifolds(fx, *got)
        assert not errs, "\n".join(errs)


@pytest.mark.parametrize("seed", [17, 18])
def test_gpu_capsules_match_oracle_on_fresh_pairs(ctx, seed):
    from make_fixtures import _axes, capsule_pairs
    recs = capsule_pairs(seed, 12000)
    shapes = np.zeros((len(recs), 2, 19))
    for t in range(2):
        r = recs[:, 14 * t:14 * (t + 1)]
        shapes[:, t, 0] = r[:, 0]
        shapes[:, t, 1:4] = r[:, 1:4]
        shapes[:, t, 4:13] = np.array([_axes(q) for q in r[:, 4:8]])
        shapes[:, t, 13:16] = r[:, 8:11]
        shapes[:, t, 16:19] = r[:, 11:14]
    margin, tol = recs[:, 28].copy(), recs[:, 29].copy()
    want = binding.collide(shapes[:, 0].copy(), shapes[:, 1].copy(), margin, tol)
    got = native.collide(ctx, shapes[:, 0].copy(), shapes[:, 1].copy(), margin, tol)
    fx = {"count": want[0], "normal": want[1], "points": want[2], "separations": want[3], "ids": want[4]}
    errs = compare_manifolds(fx, *got)
    assert not errs, "\n".join(errs)
    assert (want[0] > 0).sum() > 5000



def test_gpu_hulls_match_reference(ctx):
    # Hulls built by the reference's own src/hull.js build(); manifolds and loss bit for bit.
    fx = fixtures.load(fixtures.DATA + "/collide_hull.2026.bmfx.gz")
    native.set_hulls(ctx, fx)
    sh = fx["shapes"].reshape(-1, 2, 19)
    got = native.collide(ctx, sh[:, 0].copy(), sh[:, 1].copy(), fx["margin"], fx["tolerance"], with_loss=True)
    errs = compare_manifolds(fx, *got[:5])
    assert not errs, "\n".join(errs)
    assert np.array_equal(got[5].view(np.uint64), np.asarray(fx["loss"]).view(np.uint64))


@pytest.mark.parametrize("seed", [27, 28])
def test_gpu_hulls_match_oracle_on_fresh_placements(ctx, seed):
    # The same hulls, placed afresh: every shape moved and turned at random, pairs brought
    # within reach of each other.
    fx = fixtures.load(fixtures.DATA + "/collide_hull.2026.bmfx.gz")
    native.set_hulls(ctx, fx)
    binding.hulls(fx)
    from make_fixtures import _axes
    rng = np.random.default_rng(seed)
    sh = fx["shapes"].reshape(-1, 2, 19).copy()
    n = len(sh)
    for t in range(2):
        q = rng.normal(size=(n, 4))
        q /= np.linalg.norm(q, axis=1)[:, None]
        turn = sh[:, t, 0] != 0
        sh[turn, t, 4:13] = np.array([_axes(x) for x in q[turn]])
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    sh[:, 1, 1:4] = sh[:, 0, 1:4] + d * rng.uniform(0.0, 0.9, (n, 1))
    margin = rng.uniform(0, 0.3, n)
    tol = np.asarray(fx["tolerance"])
    want = binding.collide(sh[:, 0].copy(), sh[:, 1].copy(), margin, tol)
    want_loss = binding.losses(n)
    got = native.collide(ctx, sh[:, 0].copy(), sh[:, 1].copy(), margin, tol, with_loss=True)
    ref = {"count": want[0], "normal": want[1], "points": want[2], "separations": want[3], "ids": want[4]}
    errs = compare_manifolds(ref, *got[:5])
    assert not errs, "\n".join(errs)
    assert np.array_equal(got[5].view(np.uint64), want_loss.view(np.uint64))
    assert (want[0] > 0).sum() > 300
