"""bm_world's host half on the CPU: the AABB trees, pair list, narrow order, reuse,
wakes, freezing and sleeping of bearing_amd/csrc/world.cpp, with the oracle answering
the device calls (tests/hostworld/stub_device.cpp), stepped along the reference's own
trajectories and compared with its per-step world.transforms hashes bit for bit."""
import ctypes
import os
import subprocess

import pytest

import fixtures
from bearing_amd import native
from worldcheck import init_columns, run

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "hostworld", "_build", "libhostworld.so")


@pytest.fixture(scope="module")
def lib():
    subprocess.run(["make", "-s", "-C", os.path.join(HERE, "hostworld")], check=True)
    return ctypes.CDLL(LIB)


def traj(name):
    return fixtures.load(os.path.join(fixtures.DATA, f"{name}.traj.bmfx.gz"))


@pytest.mark.parametrize("name", ["pyramid", "tower", "throw"])
def test_host_world_matches_reference_trajectory(lib, name):
    fx = traj(name)
    w = native.World(lib=lib)
    w.add(init_columns(fx))
    matched, why = run(w, fx)
    assert why is None, why
    assert matched == fx.meta["steps"]
    # Every body went to sleep, as in the reference.
    assert w.read()["awake"] == int(fx["awake"][-1])


def test_host_world_stops_where_it_would_sweep_a_fast_body(lib):
    # Every step until then matches; the continuous catch is not restated yet, so the step
    # that has a body fast enough to be swept fails loudly instead of going on differently.
    fx = traj("rain_small")
    w = native.World(lib=lib)
    w.add(init_columns(fx))
    with pytest.raises(native.BearingError, match="swept"):
        run(w, fx)
