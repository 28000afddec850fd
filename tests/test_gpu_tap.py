"""A real step's contact phase, tapped (bm_world_tap) from a world stepping on the device
and replayed through a context of its own on the device and through the oracle on the host:
every column after the soft step the same bits. The worlds are reference trajectories part
way (a pile of boxes and spheres landing, the 2,000-sphere rain piling up), so the soft step
is checked on the states the plain drop-in really hands it, not only on fixtures."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

import fixtures
from bearing_amd import native
from worldcheck import make_world, replay

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.parametrize("name,step", [("pile_small", 70), ("rain", 300)])
def test_gpu_tapped_step_replays_as_the_oracle(name, step):
    print("tap", name, step, flush=True)
    subprocess.run(["make", "-s", "-C", os.path.join(HERE, "hostworld")], check=True)
    host = ctypes.CDLL(os.path.join(HERE, "hostworld", "_build", "libhostworld.so"))
    fx = fixtures.load(os.path.join(fixtures.DATA, f"{name}.traj.bmfx.gz"))
    w = make_world(fx, device=0)
    for _ in range(step - 1):
        w.step(fx.meta["dt"])
    caught = w.read()["catches"]
    w.capture_next()
    w.step(fx.meta["dt"])
    after = w.read()
    cap = w.captured
    w.close()
    gpu, _ = replay(cap, device=0)
    cpu, _ = replay(cap, lib=host, workers=3)
    if after["catches"] == caught:  # no body caught in the step: its positions are the soft step's
        assert np.array_equal(gpu["position"], after["position"])
    bad = [k for k in cpu if not np.array_equal(gpu[k], cpu[k], equal_nan=True)]
    assert not bad, bad
