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
