"""The oracle's level geometry (src/mesh.js Heightfield, src/triangle.js touch and
touchLevel) against the reference, and against what the geometry fixes.

1. Every triangle's edge flags and folds of three heightfields (value noise with a flat
   plateau and a raised crest), as the reference's Heightfield.edge gives them, bit for bit.
2. touchLevel as the reference computed it for 2,400 seeded spheres, capsules, boxes and
   hulls against those fields, placed, moved and turned (tests/golden level.*): manifolds a
   shape, counts, normals, points, separations and ids, bit for bit.
3. Closed forms on a flat field: a ball resting on it, and a box lying flat.
"""
import numpy as np
import pytest

import fixtures
from levelcheck import compare_level
from oracle import binding


@pytest.fixture(scope="module")
def fx():
    f = fixtures.load(fixtures.DATA + "/level.2026.bmfx.gz")
    meta = f.meta["fields"]
    binding.fields([m["columns"] for m in meta], [m["rows"] for m in meta], [m["spacing"] for m in meta],
                   [f[f"field{i}.heights"] for i in range(len(meta))])
    binding.hulls(f)
    return f


def test_field_edges_match_reference(fx):
    for i, m in enumerate(fx.meta["fields"]):
        count = 2 * (m["columns"] - 1) * (m["rows"] - 1)
        flags, folds = binding.field_edges(i, count)
        assert np.array_equal(flags, fx[f"field{i}.flags"])
        assert np.array_equal(folds.view(np.uint64), fx[f"field{i}.folds"].view(np.uint64))


def test_touch_level_matches_reference_bit_for_bit(fx):
    got = binding.level(fx["field"], fx["position"], fx["axes"], fx["shapes"], fx["boxes"], fx["margin"],
                        fx["tolerance"], fx.meta["cap"])
    errs = compare_level(fx, *got)
    assert not errs, "\n".join(errs)
    # The cases reach every kind and many directions.
    kinds = fx["shapes"].reshape(-1, 19)[:, 0]
    for k in range(4):
        assert (fx["manifolds"][kinds == k] > 0).sum() > 100
    assert fx["manifolds"].max() >= 4


def flat_field(height=0.25, n=9):
    binding.fields([n], [n], [1.0], [np.full(n * n, height)])


IDENTITY = [1.0, 0, 0, 0, 1.0, 0, 0, 0, 1.0]


def test_ball_resting_on_flat_ground():
    flat_field()
    r, y = 0.3, 0.25 + 0.3 - 0.01
    rec = [0.0, 0.3, y, -0.2, *IDENTITY, r, 0, 0, 0, 0, 0]
    box = [0.3 - 0.4, y - 0.4, -0.2 - 0.4, 0.3 + 0.4, y + 0.4, -0.2 + 0.4]
    m, count, normal, points, seps, ids = binding.level([0], [0, 0, 0], IDENTITY, rec, box, [0.02], [2 * r / 200], 8)
    # One manifold: the ball pressed 1 cm into the ground, pushed straight up (normal from the ball to the ground).
    assert m[0] == 1 and count[0] == 1
    assert normal[:3] == pytest.approx([0.0, -1.0, 0.0], abs=1e-15)
    assert seps[0] == pytest.approx(-0.01, abs=1e-15)
    assert points[:3] == pytest.approx([0.3, 0.25 - 0.005, -0.2], abs=1e-15)


def test_box_lying_flat_keeps_its_four_corners():
    flat_field()
    h = [0.4, 0.2, 0.3]
    y = 0.25 + h[1] - 0.002
    rec = [2.0, 0.1, y, 0.05, *IDENTITY, *h, 0, 0, 0]
    g = 0.6
    box = [0.1 - g, y - g, 0.05 - g, 0.1 + g, y + g, 0.05 + g]
    m, count, normal, points, seps, ids = binding.level([0], [0, 0, 0], IDENTITY, rec, box, [0.01], [0.4 / 200], 8)
    assert m[0] == 1 and count[0] == 4
    assert normal[1] == -1.0
    assert np.allclose(seps[:4], -0.002, atol=1e-15)
    # The box's own corners (ids 32 + corner), whatever triangles lie under them.
    assert sorted(ids[:4]) == [32, 33, 34, 35]
    corners = points[:12].reshape(4, 3)
    want = sorted((0.1 + sx * 0.4, 0.05 + sz * 0.3) for sx in (-1, 1) for sz in (-1, 1))
    assert np.allclose(sorted(map(tuple, corners[:, [0, 2]])), want, atol=1e-15)
