"""The batched form on the MI355X: the 4,096 worlds of scenes.batched (seed 4; log-uniform
8..512 bodies, box towers and sphere piles, 276,589 bodies in all) stepped together by
bm_batch_step for 300 steps at dt 1/60, every world's final world.transforms against the
reference's own trajectory of that world bit for bit, the sampled worlds' every step, and
a few worlds stepped alone through bm_world_step against their batched selves."""
import pytest

import batchedcheck
from bearing_amd import native, scenes

pytestmark = pytest.mark.gpu


@pytest.mark.timeout(900)
def test_gpu_batched_worlds_match_reference():
    fx = batchedcheck.load()
    specs = batchedcheck.scene_list(fx)
    steps, dt = fx.meta["steps"], fx.meta["dt"]
    sample = [int(i) for i in fx["sample"]]
    print(f"batched: {len(specs)} worlds, {fx.meta['bodies']} bodies, {steps} steps", flush=True)
    batch = native.WorldBatch(0, threads=16)
    worlds = []
    for bodies in specs:
        w = batch.world()
        w.add(scenes.resolve(bodies))
        worlds.append(w)
    for s in range(1, steps + 1):
        batch.step(dt)
        for i in sample:
            why = batchedcheck.check_step(fx, i, s, worlds[i])
            assert why is None, why
        if s % 50 == 0:
            print(f"batched: step {s}", flush=True)
    bad = [why for why in (batchedcheck.check_final(fx, i, w) for i, w in enumerate(worlds)) if why]
    batch.close()
    assert not bad, f"{len(bad)} worlds differ; first: {bad[0]}"


@pytest.mark.timeout(600)
def test_gpu_batched_worlds_alone_match_reference():
    # The sampled worlds (the smallest, median and largest of each kind, and the first three),
    # each stepped alone through bm_world_step, end where the reference's do.
    fx = batchedcheck.load()
    specs = batchedcheck.scene_list(fx)
    for i in (int(i) for i in fx["sample"]):
        print(f"batched world {i} alone ({len(specs[i])} bodies)", flush=True)
        w = native.World(0)
        w.add(scenes.resolve(specs[i]))
        for s in range(1, fx.meta["steps"] + 1):
            w.step(fx.meta["dt"])
            why = batchedcheck.check_step(fx, i, s, w)
            assert why is None, why
        why = batchedcheck.check_final(fx, i, w)
        w.close()
        assert why is None, why
