"""The oracle's narrowphase (spheres, capsules and boxes) against the reference.

1. The reference's own hand-worked cases (test/collide.test.js:61-199), to its tolerance of 1e-9.
2. collide() as the reference computed it on 4,000 seeded sphere/box pairs (tests/golden
   collide.*) and 3,000 pairs with a capsule (collide_capsule.*): counts, normals, points,
   separations and feature ids, bit for bit.
"""
import math

import numpy as np
import pytest

import fixtures
from collidecheck import compare_manifolds
from oracle import binding

COLLIDE = fixtures.find("collide") if hasattr(fixtures, "find") else []


def test_collide_matches_reference_bit_for_bit():
    paths = sorted(__import__("glob").glob(fixtures.DATA + "/collide.*.bmfx.gz"))
    assert paths
    for path in paths:
        fx = fixtures.load(path)
        sh = fx["shapes"].reshape(-1, 2, 19)
        got = binding.collide(sh[:, 0].copy(), sh[:, 1].copy(), fx["margin"], fx["tolerance"])
        errs = compare_manifolds(fx, *got)
        assert not errs, "\n".join(errs)


def shape(kind, size=(1, 1, 1), radius=0.5, height=2.0, position=(0, 0, 0), rotation=(0, 0, 0, 1)):
    """A placed shape record, as the reference's test helper make() builds one."""
    x, y, z, w = rotation
    axes = [1 - 2 * (y * y + z * z), 2 * (x * y + z * w), 2 * (x * z - y * w),
            2 * (x * y - z * w), 1 - 2 * (x * x + z * z), 2 * (y * z + x * w),
            2 * (x * z + y * w), 2 * (y * z - x * w), 1 - 2 * (x * x + y * y)]
    if kind == "box":
        return [2.0, *position, *axes, size[0] / 2, size[1] / 2, size[2] / 2, 0, 0, 0]
    if kind == "capsule":
        return [1.0, *position, *axes, radius, (height - 2 * radius) / 2, 0, 0, 0, 0]
    return [0.0, *position, *axes, radius, 0, 0, 0, 0, 0]


def turn(axis, radians):
    l = math.sqrt(sum(a * a for a in axis))
    s = math.sin(radians / 2)
    return [a / l * s for a in axis] + [math.cos(radians / 2)]


def run(a, b, margin=0.2, tolerance=0.005):
    count, normal, points, seps, ids = binding.collide(np.array([a]), np.array([b]), [margin], [tolerance])
    return int(count[0]), normal, [(points[3 * k:3 * k + 3], seps[k]) for k in range(count[0])]


def same_points(got, expected):
    assert len(got) == len(expected)
    left = list(got)
    for p, s in expected:
        i = next(i for i, (g, _) in enumerate(left) if np.linalg.norm(g - np.array(p)) < 1e-9)
        assert abs(left[i][1] - s) <= 1e-9
        left.pop(i)


def test_hand_worked_sphere_sphere():
    a = shape("sphere", radius=1)
    n, normal, pts = run(a, shape("sphere", radius=0.5, position=(2, 0, 0)), 1)
    assert np.allclose(normal, [1, 0, 0], atol=1e-9)
    same_points(pts, [([1.25, 0, 0], 0.5)])
    assert run(a, shape("sphere", radius=0.5, position=(3, 0, 0)), 0.5)[0] == 0
    same_points(run(a, shape("sphere", radius=0.5, position=(1.2, 0, 0)))[2], [([0.85, 0, 0], -0.3)])
    n, normal, pts = run(a, shape("sphere", radius=0.5))
    assert np.allclose(normal, [0, 1, 0])
    same_points(pts, [([0, 0.25, 0], -1.5)])


def test_hand_worked_sphere_box():
    box = shape("box", size=(2, 2, 2))
    n, normal, pts = run(shape("sphere", position=(0, 1.6, 0)), box)
    assert np.allclose(normal, [0, -1, 0])
    same_points(pts, [([0, 1.05, 0], 0.1)])
    n, normal, pts = run(shape("sphere", position=(1.3, 1.3, 0)), box)
    assert np.allclose(normal, [-math.sqrt(0.5), -math.sqrt(0.5), 0], atol=1e-9)
    assert abs(pts[0][1] - (0.3 * math.sqrt(2) - 0.5)) <= 1e-9
    n, normal, pts = run(shape("sphere", position=(1.2, 1.2, 1.2)), box)
    assert abs(pts[0][1] - (0.2 * math.sqrt(3) - 0.5)) <= 1e-9
    n, normal, pts = run(shape("sphere", position=(0, 0.9, 0)), box)
    assert np.allclose(normal, [0, -1, 0])
    same_points(pts, [([0, 0.7, 0], -0.6)])
    n, normal, pts = run(box, shape("sphere", position=(0, 1.6, 0)))
    assert np.allclose(normal, [0, 1, 0])
    same_points(pts, [([0, 1.05, 0], 0.1)])
    n, normal, pts = run(shape("sphere", position=(0, 1.6, 0)), shape("box", size=(2, 2, 2), rotation=turn([0, 1, 0], 0.7)))
    same_points(pts, [([0, 1.05, 0], 0.1)])


def test_hand_worked_box_box():
    a = shape("box", size=(2, 2, 2))
    n, normal, pts = run(a, shape("box", size=(2, 2, 2), position=(0, 2.1, 0)))
    assert np.allclose(normal, [0, 1, 0])
    same_points(pts, [([-1, 1.05, -1], 0.1), ([1, 1.05, -1], 0.1), ([1, 1.05, 1], 0.1), ([-1, 1.05, 1], 0.1)])
    big, small = shape("box", size=(4, 1, 4)), shape("box", size=(1, 1, 1), position=(0, 1, 0))
    corners = [[-0.5, 0.5, -0.5], [0.5, 0.5, -0.5], [0.5, 0.5, 0.5], [-0.5, 0.5, 0.5]]
    n, normal, pts = run(big, small)
    assert np.allclose(normal, [0, 1, 0])
    same_points(pts, [(p, 0) for p in corners])
    n, normal, pts = run(small, big)
    assert np.allclose(normal, [0, -1, 0])
    same_points(pts, [(p, 0) for p in corners])
    n, normal, pts = run(a, shape("box", size=(2, 2, 2), position=(0, 2.05, 0), rotation=turn([0, 1, 0], math.pi / 4)))
    assert np.allclose(normal, [0, 1, 0]) and n == 4
    for p, s in pts:
        assert abs(s - 0.05) <= 1e-9 and abs(p[0]) <= 1 + 1e-9 and abs(p[2]) <= 1 + 1e-9
    ridge = math.sqrt(2)
    n, normal, pts = run(shape("box", size=(2, 2, 2), rotation=turn([0, 0, 1], math.pi / 4)),
                         shape("box", size=(2, 2, 2), position=(0, 2.8, 0), rotation=turn([1, 0, 0], math.pi / 4)))
    assert np.allclose(normal, [0, 1, 0], atol=1e-9)
    same_points(pts, [([0, 1.4, 0], 2.8 - 2 * ridge)])
    up = [1 / math.sqrt(3)] * 3
    corner = turn([up[2], 0, -up[0]], math.acos(up[1]) + math.pi)
    n, normal, pts = run(big, shape("box", size=(1, 1, 1), position=(0, 0.5 + math.sqrt(3) / 2 + 0.01, 0), rotation=corner), 0.05)
    assert n == 1
    assert abs(pts[0][1] - 0.01) <= 1e-12
    assert np.allclose(pts[0][0], [0, 0.505, 0], atol=1e-12)


def test_capsules_match_reference_bit_for_bit():
    paths = sorted(__import__("glob").glob(fixtures.DATA + "/collide_capsule.*.bmfx.gz"))
    assert paths
    for path in paths:
        fx = fixtures.load(path)
        sh = fx["shapes"].reshape(-1, 2, 19)
        got = binding.collide(sh[:, 0].copy(), sh[:, 1].copy(), fx["margin"], fx["tolerance"])
        errs = compare_manifolds(fx, *got)
        assert not errs, "\n".join(errs)
        # every capsule pairing, and the manifold sizes each can have, are in the vectors
        kinds = set(map(tuple, sh[:, :, 0].astype(int)))
        assert {(0, 1), (1, 0), (1, 1), (1, 2), (2, 1)} <= kinds
        assert int(fx["count"].max()) == 3


def test_hand_worked_sphere_capsule():
    # test/collide.test.js:99-112
    capsule = shape("capsule", radius=0.5, height=2)
    n, normal, pts = run(shape("sphere", position=(1.2, 0, 0)), capsule)
    assert np.allclose(normal, [-1, 0, 0], atol=1e-9)
    same_points(pts, [([0.6, 0, 0], 0.2)])
    n, normal, pts = run(shape("sphere", position=(0, 1.8, 0)), capsule, margin=0.5)
    assert np.allclose(normal, [0, -1, 0], atol=1e-9)
    same_points(pts, [([0, 1.15, 0], 0.3)])
    n, normal, pts = run(shape("sphere", position=(0, 0, 0)), capsule)
    assert np.allclose(normal, [0, 0, -1], atol=1e-9)
    assert abs(pts[0][1] + 1) <= 1e-9


def test_hand_worked_capsule_capsule():
    # test/collide.test.js:114-126
    a = shape("capsule")
    n, normal, pts = run(a, shape("capsule", position=(1.1, 0, 0)))
    assert np.allclose(normal, [1, 0, 0], atol=1e-9)
    same_points(pts, [([0.55, -0.5, 0], 0.1), ([0.55, 0.5, 0], 0.1)])
    n, normal, pts = run(a, shape("capsule", position=(1.1, 0, 0), rotation=turn([1, 0, 0], math.pi / 2)))
    same_points(pts, [([0.55, 0, 0], 0.1)])
    n, normal, pts = run(a, shape("capsule", position=(0, 2.1, 0)))
    assert np.allclose(normal, [0, 1, 0], atol=1e-9)
    same_points(pts, [([0, 1.05, 0], 0.1)])
    n, normal, pts = run(a, shape("capsule", position=(1.1, 0.6, 0)))
    same_points(pts, [([0.55, 0.1, 0], 0.1), ([0.55, 0.5, 0], 0.1)])


def test_hand_worked_capsule_box():
    # test/collide.test.js:128-144
    box = shape("box", size=(4, 1, 4))
    lying = turn([0, 0, 1], -math.pi / 2)
    n, normal, pts = run(shape("capsule", position=(0, 1.1, 0), rotation=lying), box)
    assert np.allclose(normal, [0, -1, 0], atol=1e-9)
    same_points(pts, [([-0.5, 0.55, 0], 0.1), ([0.5, 0.55, 0], 0.1)])
    n, normal, pts = run(shape("capsule", position=(0, 1.6, 0)), box)
    same_points(pts, [([0, 0.55, 0], 0.1)])
    n, normal, pts = run(shape("capsule", position=(2.3, 0.8, 0), rotation=turn([1, 0, 0], math.pi / 2)), box)
    assert np.allclose(normal, [-math.sqrt(0.5), -math.sqrt(0.5), 0], atol=1e-9)
    assert n == 1 and abs(pts[0][1] - (0.3 * math.sqrt(2) - 0.5)) <= 1e-9
    n, normal, pts = run(shape("capsule", radius=0.3, height=6, position=(0, 0.3, 0), rotation=lying), box)
    assert np.allclose(normal, [0, -1, 0], atol=1e-9)
    same_points(pts, [([-2, 0.25, 0], -0.5), ([2, 0.25, 0], -0.5)])


def test_hulls_match_reference_bit_for_bit():
    # 1,200 pairs with a hull built by the reference's own src/hull.js build(): hull on hull,
    # box on hull (the box as the reference's CUBE), sphere and capsule on hull, either order;
    # manifolds and Manifold.loss (what REUSE keeps a hull manifold by) bit for bit.
    path = fixtures.DATA + "/collide_hull.2026.bmfx.gz"
    fx = fixtures.load(path)
    binding.hulls(fx)
    sh = fx["shapes"].reshape(-1, 2, 19)
    got = binding.collide(sh[:, 0].copy(), sh[:, 1].copy(), fx["margin"], fx["tolerance"])
    errs = compare_manifolds(fx, *got)
    assert not errs, "\n".join(errs)
    loss = binding.losses(len(fx["margin"]))
    assert np.array_equal(loss.view(np.uint64), np.asarray(fx["loss"]).view(np.uint64))
    assert (np.asarray(fx["loss"]) > 0).sum() > 10 and int(fx["count"].max()) == 4
    kinds = set(map(tuple, sh[:, :, 0].astype(int)))
    assert {(3, 3), (2, 3), (3, 2), (0, 3), (3, 0), (1, 3), (3, 1)} <= kinds


def test_cube_hull_is_the_references():
    # The cube the oracle and the device place a box as against a hull is hull 0 of the
    # fixture, the reference's build() of the unit cube's corners.
    fx = fixtures.load(fixtures.DATA + "/collide_hull.2026.bmfx.gz")
    assert fx["vertices"][:24].tolist() == [-0.5, -0.5, -0.5, 0.5, -0.5, -0.5, -0.5, 0.5, -0.5, 0.5, 0.5, -0.5,
                                             -0.5, -0.5, 0.5, 0.5, -0.5, 0.5, -0.5, 0.5, 0.5, 0.5, 0.5, 0.5]
    assert fx["start"][:6].tolist() == [0, 4, 8, 12, 16, 20] and fx["hullL"][1] == 24
    assert fx["loops"][:24].tolist() == [0, 1, 5, 4, 0, 4, 6, 2, 0, 2, 3, 1, 1, 3, 7, 5, 4, 5, 7, 6, 2, 6, 7, 3]
    assert fx["planes"][:24].tolist() == [0, -1, 0, 0.5, -1, 0, 0, 0.5, 0, 0, -1, 0.5, 1, 0, 0, 0.5, 0, 0, 1, 0.5,
                                           0, 1, 0, 0.5]
