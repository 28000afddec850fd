"""bm_world on the MI355X: whole worlds stepped with the device narrowphase, colouring
and soft step, compared with the reference's per-step world.transforms bit for bit."""
import os

import pytest

import fixtures
from bearing_amd import native
from worldcheck import init_columns, make_world, refuses, run

pytestmark = pytest.mark.gpu


def traj(name):
    return fixtures.load(os.path.join(fixtures.DATA, f"{name}.traj.bmfx.gz"))


@pytest.mark.parametrize("name", ["pyramid", "tower", "throw", "rain_small", "pile_small", "darts", "capsules", "rocks", "terrain", "rain"])
def test_gpu_world_matches_reference_trajectory(name):
    print("trajectory", name, flush=True)
    fx = traj(name)
    w = make_world(fx, device=0)
    matched, why = run(w, fx)
    state = w.read()
    w.close()
    assert why is None, why
    assert matched == fx.meta["steps"]
    assert (state["impacts"], state["catches"]) == (fx.meta["impacts"], fx.meta["catches"])


@pytest.mark.parametrize("name", ["pyramid", "throw", "rain_small", "darts", "rocks", "terrain", "rain"])
def test_gpu_world_with_device_broadphase_matches_reference_trajectory(name):
    # Every step's bounds and fat-box tests, tree walks and pair holds on the device
    # (bm_broad_*), the host re-filing bodies and taking what the walks met in order.
    print("trajectory", name, "device broadphase", flush=True)
    fx = traj(name)
    w = make_world(fx, device=0)
    w.set_broadphase("device")
    matched, why = run(w, fx)
    state = w.read()
    w.close()
    assert why is None, why
    assert matched == fx.meta["steps"]
    assert (state["impacts"], state["catches"]) == (fx.meta["impacts"], fx.meta["catches"])


def test_gpu_world_refuses_what_it_does_not_handle():
    refuses(lambda: native.World(0))


@pytest.mark.parametrize("graph", [True, False])
def test_gpu_batch_steps_every_world_as_alone(graph):
    # Worlds of every kind stepped together on the device (bm_batch_step: one narrowphase
    # launch a pass, one colouring and one soft-step launch for all of them, as a HIP graph or
    # launch by launch), each matching its own reference trajectory on every step.
    from worldcheck import fill_world, run_together
    names = ["tower", "pile_small", "darts", "capsules", "rocks", "terrain", "tower", "rain_small", "pyramid"]
    print("batch of", names, "graph" if graph else "launches", flush=True)
    fxs = [traj(n) for n in names]
    batch = native.WorldBatch(0, threads=8)
    batch.set_graph(graph)
    worlds = [fill_world(batch.world(), fx) for fx in fxs]
    steps = min(f.meta["steps"] for f in fxs)
    matched, why = run_together(batch, worlds, fxs, steps)
    counts = batch.graph_counts()
    batch.close()
    assert why is None, why
    assert matched == steps
    # One graph launch a step when on (every step had contacts to colour), none when off.
    print("graphs", counts, flush=True)
    assert counts["launches"] == (steps if graph else 0), counts
    assert counts["made"] + counts["updated"] == counts["launches"], counts


def test_gpu_mixed_pile_first_steps():
    # The mixed pile at full size (20,001 bodies: boxes, balls, capsules and hulls onto a 256 x 256
    # heightfield), its first 150 steps as the landing starts, every step's transforms the
    # reference's.
    # By default a world this size takes its broadphase passes on the device; the host's
    # give the same steps.
    fx = traj("mixed_pile")
    for mode in ("auto", "host"):
        print("trajectory mixed_pile, 150 steps, broadphase", mode, flush=True)
        w = make_world(fx, device=0)
        w.set_broadphase(mode)
        matched, why = run(w, fx, steps=150)
        w.close()
        assert why is None, why
        assert matched == 150
