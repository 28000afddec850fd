"""bm_world on the MI355X: whole worlds stepped with the device narrowphase, colouring
and soft step, compared with the reference's per-step world.transforms bit for bit."""
import os

import pytest

import fixtures
from bearing_amd import native
from worldcheck import init_columns, run

pytestmark = pytest.mark.gpu


def traj(name):
    return fixtures.load(os.path.join(fixtures.DATA, f"{name}.traj.bmfx.gz"))


@pytest.mark.parametrize("name", ["pyramid", "tower", "throw"])
def test_gpu_world_matches_reference_trajectory(name):
    print("trajectory", name, flush=True)
    fx = traj(name)
    w = native.World(0)
    w.add(init_columns(fx))
    matched, why = run(w, fx)
    w.close()
    assert why is None, why
    assert matched == fx.meta["steps"]


def test_gpu_world_stops_where_it_would_sweep_a_fast_body():
    fx = traj("rain_small")
    w = native.World(0)
    w.add(init_columns(fx))
    with pytest.raises(native.BearingError, match="swept"):
        run(w, fx)
    w.close()
