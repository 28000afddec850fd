"""The device's touchLevel (heightfield against sphere, capsule, box and hull) against the
reference's vectors, bit for bit, and against the oracle on fresh seeded placements."""
import numpy as np
import pytest

import fixtures
from bearing_amd import native
from levelcheck import compare_level
from oracle import binding

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def fx():
    return fixtures.load(fixtures.DATA + "/level.2026.bmfx.gz")


@pytest.fixture(scope="module")
def ctx(fx):
    c = native.Context(0)
    meta = fx.meta["fields"]
    native.set_fields(c, [m["columns"] for m in meta], [m["rows"] for m in meta], [m["spacing"] for m in meta],
                      [fx[f"field{i}.heights"] for i in range(len(meta))])
    native.set_hulls(c, fx)
    yield c
    c.close()


def test_gpu_level_matches_reference(ctx, fx):
    print("level fixture", fx.meta["cases"], "cases", flush=True)
    got = native.level(ctx, fx["field"], fx["position"], fx["axes"], fx["shapes"], fx["boxes"], fx["margin"],
                       fx["tolerance"], fx.meta["cap"])
    errs = compare_level(fx, *got)
    assert not errs, "\n".join(errs)


def test_gpu_level_reports_every_direction_past_cap(ctx, fx):
    # With room for one manifold a shape, the count still says how many directions there were.
    got = native.level(ctx, fx["field"], fx["position"], fx["axes"], fx["shapes"], fx["boxes"], fx["margin"],
                       fx["tolerance"], 1)
    assert np.array_equal(got[0], fx["manifolds"])
    first = fx["count"].reshape(-1, fx.meta["cap"])[:, 0]
    assert np.array_equal(got[1], first)


@pytest.mark.parametrize("seed", [5, 6])
def test_gpu_level_matches_oracle_on_fresh_placements(ctx, fx, seed):
    from make_fixtures import level_cases
    job = level_cases(seed, 3000)
    rng = np.random.default_rng(seed)
    n = len(job["cases"])
    shapes = np.zeros((n, 19))
    field = np.zeros(n, np.int32)
    position, axes, boxes = np.zeros((n, 3)), np.zeros((n, 9)), np.zeros((n, 6))
    margin, tolerance = np.zeros(n), np.zeros(n)
    for k, c in enumerate(job["cases"]):
        s = c["shape"]
        x, y, z, w = s["q"]
        shapes[k, 0] = s["kind"]
        shapes[k, 1:4] = s["center"]
        shapes[k, 4:13] = [1 - 2 * (y * y + z * z), 2 * (x * y + z * w), 2 * (x * z - y * w),
                           2 * (x * y - z * w), 1 - 2 * (x * x + z * z), 2 * (y * z + x * w),
                           2 * (x * z + y * w), 2 * (y * z - x * w), 1 - 2 * (x * x + y * y)]
        if s["kind"] == 3:   # one of the fixture's built hulls, placed here
            shapes[k, 13] = rng.integers(0, fx.meta["hulls"])
            tolerance[k] = 0.002
        else:
            shapes[k, 13:16] = s["half"]
            tolerance[k] = (2 * min(s["half"]) if s["kind"] == 2 else 2 * s["half"][0]) / 200
        field[k], position[k], axes[k], boxes[k], margin[k] = c["field"], c["position"], c["axes"], c["box"], c["margin"]
    # The same fields as the fixture's (level_cases draws them the same way for any seed).
    meta = fx.meta["fields"]
    binding.fields([m["columns"] for m in meta], [m["rows"] for m in meta], [m["spacing"] for m in meta],
                   [fx[f"field{i}.heights"] for i in range(len(meta))])
    binding.hulls(fx)
    cap = 32
    want = binding.level(field, position, axes, shapes, boxes, margin, tolerance, cap)
    got = native.level(ctx, field, position, axes, shapes, boxes, margin, tolerance, cap)
    assert (want[0] > 0).sum() > 1000
    for name, g, w in zip(["manifolds", "count", "normal", "points", "separations", "ids"], got, want):
        if g.dtype == np.float64:
            assert np.array_equal(g.view(np.uint64), w.view(np.uint64)), name
        else:
            assert np.array_equal(g, w), name
