"""The oracle's time of impact, core distance and trigonometry against the reference's
own outputs on 2,000 seeded pairs (tests/golden/make_impact.mjs: src/impact.js impact,
src/gjk.js closest, src/core/math.js sin and cos), bit for bit, plus what the
mathematics fixes independently of the reference."""
import math
import os

import numpy as np
import pytest

import fixtures
from oracle import binding


@pytest.fixture(scope="module")
def fx():
    return fixtures.load(os.path.join(fixtures.DATA, "impact.2026.bmfx.gz"))


def bits(a):
    return np.asarray(a, np.float64).view(np.uint64)


def shapes(bodies):
    """The 19-number shape records inside 32-number body records."""
    return np.asarray(bodies).reshape(-1, 32)[:, :19].copy()


def test_sine_and_cosine_are_the_references(fx):
    s, c = binding.sincos(fx["x"])
    assert np.array_equal(bits(s), bits(fx["sine"]))
    assert np.array_equal(bits(c), bits(fx["cosine"]))


def test_sine_and_cosine_are_within_an_ulp_of_the_true_values():
    # Independent of the reference: fdlibm's kernels are within one ulp of the exact values,
    # and sin^2 + cos^2 = 1 to a few ulps, over the range posed() uses them in.
    x = np.linspace(-10, 10, 20001)
    s, c = binding.sincos(x)
    assert np.all(np.abs(s - np.sin(x)) <= 2 * np.spacing(np.maximum(np.abs(np.sin(x)), 1e-300)) + 1e-300)
    assert np.all(np.abs(c - np.cos(x)) <= 2 * np.spacing(np.maximum(np.abs(np.cos(x)), 1e-300)) + 1e-300)
    assert np.max(np.abs(s * s + c * c - 1)) < 4e-16
    s0, c0 = binding.sincos([0.0, -0.0, math.pi / 6])
    # (the reduction's x - n pi/2 turns -0 into +0, as the reference's does: its vectors pin that)
    assert s0[0] == 0 and s0[1] == 0 and c0[0] == 1 and c0[1] == 1
    assert abs(s0[2] - 0.5) <= 1.2e-16


def test_core_distance_is_the_references(fx):
    b = np.asarray(fx["bodies"]).reshape(-1, 2, 32)
    d, na, nb = binding.closest(shapes(b[:, 0]), shapes(b[:, 1]))
    assert np.array_equal(bits(d), bits(fx["distance"]))
    assert np.array_equal(bits(na), bits(fx["nearA"]))
    assert np.array_equal(bits(nb), bits(fx["nearB"]))


def test_core_distance_of_two_spheres_is_the_centres_distance():
    # Spheres' cores are their centres: the distance is |c_b - c_a| exactly.
    rng = np.random.default_rng(5)
    a = np.zeros((50, 19)); b = np.zeros((50, 19))
    a[:, 1:4] = rng.normal(size=(50, 3)); b[:, 1:4] = rng.normal(size=(50, 3))
    a[:, 4:13] = b[:, 4:13] = [1, 0, 0, 0, 1, 0, 0, 0, 1]
    d, na, nb = binding.closest(a, b)
    assert np.allclose(d, np.linalg.norm(b[:, 1:4] - a[:, 1:4], axis=1), rtol=1e-15, atol=0)
    assert np.array_equal(na.reshape(-1, 3), a[:, 1:4])


def test_core_distance_of_box_and_point_matches_brute_force():
    # A box's core is the box itself: the distance to a point outside it is the distance to the clamped point.
    rng = np.random.default_rng(6)
    for _ in range(200):
        h = rng.uniform(0.1, 1, 3)
        q = rng.normal(size=4); q /= np.linalg.norm(q)
        x, y, z, w = q
        R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                      [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                      [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])
        c = rng.normal(size=3)
        p = c + R @ (rng.uniform(-3, 3, 3))
        box = np.concatenate([[2], c, R.T.reshape(-1), h, [0, 0, 0]])
        pt = np.concatenate([[0], p, [1, 0, 0, 0, 1, 0, 0, 0, 1], [0, 0, 0], [0, 0, 0]])
        d, _, _ = binding.closest(box[None], pt[None])
        local = R.T @ (p - c)
        want = np.linalg.norm(local - np.clip(local, -h, h))
        assert abs(d[0] - want) <= 1e-9 * max(1, want)


def test_time_of_impact_is_the_references(fx):
    b = np.asarray(fx["bodies"]).reshape(-1, 2, 32)
    t, pa, pb = binding.impact(b[:, 0], b[:, 1], fx["dt"], fx["tolerance"])
    assert np.array_equal(bits(t), bits(fx["t"]))
    assert np.array_equal(bits(pa), bits(fx["posedA"]))
    assert np.array_equal(bits(pb), bits(fx["posedB"]))
    # The vectors cover every outcome: a moment in the step, the start, apart, unsure.
    t = fx["t"]
    assert (t > 0).sum() > 100 and (t == 0).sum() > 50 and (t == -1).sum() > 100 and (t == -2).sum() > 20


def test_time_of_impact_of_spheres_in_free_flight_is_the_closed_form():
    # Two spheres with no spin and no gravity meet when |d + v t| = ra + rb: conservative
    # advancement lands within the tolerance of that time, never past it.
    rng = np.random.default_rng(7)
    n = 200
    a = np.zeros((n, 32)); b = np.zeros((n, 32))
    for r in (a, b):
        r[:, 4:13] = [1, 0, 0, 0, 1, 0, 0, 0, 1]
        r[:, 19:23] = [0, 0, 0, 1]
        r[:, 13] = rng.uniform(0.1, 0.5, n)
        r[:, 30] = r[:, 13]
        r[:, 31] = 1
    d = rng.normal(size=(n, 3)); d /= np.linalg.norm(d, axis=1)[:, None]
    gap = rng.uniform(0.05, 0.5, n)
    b[:, 1:4] = d * (a[:, 13] + b[:, 13] + gap)[:, None]
    dt = np.full(n, 1 / 60)
    speed = gap / dt * rng.uniform(1.2, 3, n)
    a[:, 23:26] = d * speed[:, None]
    tol = np.full(n, 1e-3)
    t, _, _ = binding.impact(a, b, dt, tol)
    exact = gap / speed
    assert np.all(t >= 0) and np.all(t <= exact + 1e-15)
    assert np.all((exact - t) * speed <= tol * (1 + 1e-9))
