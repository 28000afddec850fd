"""bm_world's host half on the CPU: the AABB trees, pair list, narrow order, reuse,
wakes, freezing and sleeping of bearing_amd/csrc/world.cpp, with the oracle answering
the device calls (tests/hostworld/stub_device.cpp), stepped along the reference's own
trajectories and compared with its per-step world.transforms hashes bit for bit."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

import fixtures
from bearing_amd import native, scenes
from worldcheck import init_columns, make_world, refuses, run

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "hostworld", "_build", "libhostworld.so")


@pytest.fixture(scope="module")
def lib():
    subprocess.run(["make", "-s", "-C", os.path.join(HERE, "hostworld")], check=True)
    return ctypes.CDLL(LIB)


def traj(name):
    return fixtures.load(os.path.join(fixtures.DATA, f"{name}.traj.bmfx.gz"))


@pytest.mark.parametrize("name", ["pyramid", "tower", "throw", "rain_small", "pile_small", "darts", "capsules", "rocks", "terrain", "rain"])
def test_host_world_matches_reference_trajectory(lib, name):
    fx = traj(name)
    w = make_world(fx, lib=lib)
    matched, why = run(w, fx)
    assert why is None, why
    assert matched == fx.meta["steps"]
    state = w.read()
    assert (state["impacts"], state["catches"]) == (fx.meta["impacts"], fx.meta["catches"])
    # Every body went to sleep, as in the reference.
    assert w.read()["awake"] == int(fx["awake"][-1])


@pytest.mark.parametrize("name", ["pyramid", "rain_small", "pile_small", "darts", "rocks", "terrain"])
def test_host_world_on_its_threads_matches_reference_trajectory(lib, name):
    # Every pass's work taken to the world's host threads (the parallel paths a large world
    # takes: decisions and slots per chunk, the apart pairs found per chunk, empty meets
    # dropped and level pairs matched on the pool, pair memory, touched and sleep timers in
    # chunks) from 16 pairs on: the same trajectory, bit for bit, with its events.
    fx = traj(name)
    w = make_world(fx, lib=lib)
    w.set_parallel(16)
    matched, why = run(w, fx)
    assert why is None, why
    assert matched == fx.meta["steps"]
    state = w.read()
    assert (state["impacts"], state["catches"]) == (fx.meta["impacts"], fx.meta["catches"])


def test_host_world_refuses_what_it_does_not_handle(lib):
    refuses(lambda: native.World(lib=lib))


def test_host_batch_steps_every_world_as_alone(lib):
    # Worlds of every kind stepped together (bm_batch_step), each matching its own reference
    # trajectory on every step; two copies of one world stay the same as each other.
    from worldcheck import fill_world, run_together
    names = ["tower", "pile_small", "darts", "capsules", "rocks", "terrain", "tower", "rain_small"]
    fxs = [traj(n) for n in names]
    batch = native.WorldBatch(threads=4, lib=lib)
    worlds = [fill_world(batch.world(), fx) for fx in fxs]
    steps = min(f.meta["steps"] for f in fxs)
    matched, why = run_together(batch, worlds, fxs, steps)
    assert why is None, why
    assert matched == steps
    # A batch's world steps only with its batch.
    with pytest.raises(native.BearingError, match="steps with its batch"):
        worlds[0].step(1 / 60)
    batch.close()


def test_host_batched_sample_matches_reference(lib):
    # The batched form's sampled worlds (the smallest, median and largest box towers and sphere
    # piles of scenes.batched(4096, seed 4), and the first three of each) stepped together for
    # 300 steps, every step the reference's own trajectory of that world.
    import batchedcheck
    fx = batchedcheck.load()
    specs = batchedcheck.scene_list(fx)
    sample = [int(i) for i in fx["sample"]]
    for i in sample:   # the fixture's sampled last step is its final record
        assert bytes(fx[f"s{i}.hashes"][-32:]) == batchedcheck.final_of(fx, i)
    batch = native.WorldBatch(threads=4, lib=lib)
    worlds = {}
    for i in sample:
        worlds[i] = batch.world()
        worlds[i].add(scenes.resolve(specs[i]))
    for s in range(1, fx.meta["steps"] + 1):
        batch.step(fx.meta["dt"])
        for i, w in worlds.items():
            why = batchedcheck.check_step(fx, i, s, w)
            assert why is None, why
    for i, w in worlds.items():
        why = batchedcheck.check_final(fx, i, w)
        assert why is None, why
    batch.close()


def test_host_world_refuses_steps_after_one_fails_part_way(lib):
    # A step that fails after it began (here the soft step, with the pairs already narrowed)
    # leaves the world half stepped: every later step returns the same failure. A dt that is
    # refused at the door changes nothing and latches nothing.
    fx = traj("tower")
    w = make_world(fx, lib=lib)
    with pytest.raises(native.BearingError, match="dt must be a positive"):
        w.step(0.0)
    w.step(fx.meta["dt"])
    lib.stub_fail_next_soft_step()
    with pytest.raises(native.BearingError, match="a device failure the test asked for"):
        w.step(fx.meta["dt"])
    with pytest.raises(native.BearingError, match="takes no more steps"):
        w.step(fx.meta["dt"])
    w.close()


def test_host_batch_refuses_steps_after_one_fails_part_way(lib):
    from worldcheck import fill_world
    fx = traj("tower")
    batch = native.WorldBatch(threads=2, lib=lib)
    worlds = [fill_world(batch.world(), fx) for _ in range(3)]
    batch.step(fx.meta["dt"])
    lib.stub_fail_next_soft_step()
    with pytest.raises(native.BearingError, match="a device failure the test asked for"):
        batch.step(fx.meta["dt"])
    with pytest.raises(native.BearingError, match="takes no more steps|its worlds take no more steps"):
        batch.step(fx.meta["dt"])
    assert "no more steps" in lib.bm_world_error(worlds[1].ptr).decode()
    batch.close()


@pytest.mark.parametrize("name", ["pile_small", "rain", "pyramid"])
def test_host_world_with_workers_matches_reference_trajectory(lib, name):
    # The oracle's soft step cut across 3 more threads (oracle_soft_step_workers), as the
    # reference's workers cut it: the same trajectory bit for bit.
    fx = traj(name)
    w = make_world(fx, lib=lib)
    lib.stub_soft_workers(3)
    try:
        matched, why = run(w, fx)
    finally:
        lib.stub_soft_workers(0)
    assert why is None, why
    assert matched == fx.meta["steps"]


@pytest.mark.parametrize("name", ["pyramid", "tower", "throw", "rain_small", "pile_small", "darts", "capsules", "rocks",
                                  "terrain"])
def test_host_world_with_device_broadphase_matches_reference_trajectory(lib, name):
    # Every step's bounds, fat-box tests, tree walks and holds through the bm_broad_* calls
    # (answered here by copies of what the world sent them, so a body the host changed without
    # sending it keeps its old state there), the host re-filing and taking the results in order.
    fx = traj(name)
    w = make_world(fx, lib=lib)
    w.set_broadphase("device")
    lib.stub_broad_rows.restype = ctypes.c_int64
    before = [lib.stub_broad_rows(k) for k in range(3)]
    matched, why = run(w, fx)
    assert why is None, why
    assert matched == fx.meta["steps"]
    state = w.read()
    assert (state["impacts"], state["catches"]) == (fx.meta["impacts"], fx.meta["catches"])
    rows = [lib.stub_broad_rows(k) - before[k] for k in range(3)]
    print(f"{name}: bounds, walks and holds of {rows} rows through bm_broad_*")
    assert all(r > 0 for r in rows), rows


def test_tapped_columns_replay_the_step(lib):
    # bm_world_tap shows the columns a step hands its contact phase: replayed through a
    # context of their own (upload, colour, soft step, download) they give the world's own
    # positions and rotations after that step, with any number of workers.
    from worldcheck import replay
    fx = traj("pile_small")
    w = make_world(fx, lib=lib)
    for _ in range(69):
        w.step(fx.meta["dt"])
    w.capture_next()
    w.step(fx.meta["dt"])
    cap = w.captured
    assert cap is not None and len(cap["points"]) > 30
    after = w.read()
    w.step(fx.meta["dt"])
    assert w.captured is None or w.captured is cap  # the tap is off after one step
    out, _ = replay(cap, lib=lib)
    assert after["catches"] == 0
    assert np.array_equal(out["position"], after["position"])
    assert np.array_equal(out["rotation"], after["rotation"])
    out3, _ = replay(cap, lib=lib, workers=3)
    for key in out:
        assert np.array_equal(out[key], out3[key], equal_nan=True), key
