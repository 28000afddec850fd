"""bm_level_noalloc on the MI355X: the level pass with its working memory in the library's
fixed device block (no allocation during the call), on the caller's device arrays, against
bm_level_host on the same inputs bit for bit: the reference's level cases, then a smaller call
right after (the block keeps the last call's contents), then the refusals the plan makes
before the pass runs."""
import numpy as np
import pytest
import torch

import fixtures
from bearing_amd import native

pytestmark = pytest.mark.gpu

NAMES = ["manifolds", "count", "normal", "points", "separations", "ids"]


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


def inputs(fx, take):
    sel = lambda a, per: np.ascontiguousarray(np.asarray(a).reshape(-1, per)[take].ravel())
    return (sel(fx["field"], 1).astype(np.int32), sel(fx["position"], 3), sel(fx["axes"], 9), sel(fx["shapes"], 19),
            sel(fx["boxes"], 6), sel(fx["margin"], 1), sel(fx["tolerance"], 1))


def on_device(ctx, n, cap, ins, fill=7):
    """Runs bm_level_noalloc on device copies of `ins`; outputs start as junk (fill) so every
    slot the call leaves must be written by it."""
    dev = [torch.from_numpy(a).to("cuda") for a in ins]
    outs = [torch.full((n,), fill, dtype=torch.int32, device="cuda"),
            torch.full((cap * n,), fill, dtype=torch.int32, device="cuda"),
            torch.full((3 * cap * n,), float(fill), dtype=torch.float64, device="cuda"),
            torch.full((12 * cap * n,), float(fill), dtype=torch.float64, device="cuda"),
            torch.full((4 * cap * n,), float(fill), dtype=torch.float64, device="cuda"),
            torch.full((4 * cap * n,), fill, dtype=torch.int32, device="cuda")]
    torch.cuda.synchronize()
    native.level_noalloc(ctx, n, cap, *[t.data_ptr() for t in dev], *[t.data_ptr() for t in outs])
    got = [t.cpu().numpy() for t in outs]
    got[5] = got[5].view(np.uint32)
    return got


def same(got, want):
    for name, g, w in zip(NAMES, got, want):
        assert g.dtype.itemsize == w.dtype.itemsize, name
        assert np.array_equal(g.view(np.uint8), w.view(np.uint8)), name


def test_gpu_level_noalloc_matches_level_host_twice(ctx, fx):
    n, cap = fx.meta["cases"], fx.meta["cap"]
    print(f"bm_level_noalloc: {n} pairs, cap {cap}", flush=True)
    ins = inputs(fx, np.arange(n))
    want = native.level(ctx, *ins, cap)
    same(on_device(ctx, n, cap, ins), want)
    # A smaller call right after, the pairs in another order and one slot a pair.
    take = np.random.default_rng(3).permutation(n)[:n // 3]
    small = inputs(fx, take)
    print(f"bm_level_noalloc: {len(take)} pairs, cap 1", flush=True)
    same(on_device(ctx, len(take), 1, small, fill=-3), native.level(ctx, *small, 1))


def test_gpu_level_noalloc_refuses_before_the_pass(ctx, fx):
    n = 64
    ins = list(inputs(fx, np.arange(n)))
    bad_field = ins[0].copy()
    bad_field[5] = 99
    with pytest.raises(native.BearingError, match="pair 5: a field's number is not in the field table"):
        on_device(ctx, n, 4, [bad_field, *ins[1:]])
    # Every box over its whole field: 14 points a cell, 64 x 14 x 23 x 19 in all, past the block.
    huge = np.tile([-1e9, -1e9, -1e9, 1e9, 1e9, 1e9], n)
    with pytest.raises(native.BearingError, match="room for 391552 found points .* more than the 262144 a call holds"):
        on_device(ctx, n, 4, [*ins[:4], huge, *ins[5:]])
    with pytest.raises(native.BearingError, match="at most 65536 pairs a call"):
        native.level_noalloc(ctx, 65537, 1, *([0] * 13))
    # After the refusals the block serves a good call as before.
    same(on_device(ctx, n, 4, ins), native.level(ctx, *ins, 4))
