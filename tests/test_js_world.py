"""The JavaScript host path: bearing_amd/js (World, addBody's resolve) over the N-API addon
(bearing_amd/_build/bearing_napi.node) over the C ABI.

CPU: the front-end's resolved columns equal the reference's own (the trajectory fixtures'
init.* columns, written by src/world.js _write) bit for bit, and the addon loads, exports its
calls and refuses to make a world without a GPU. GPU: scenes read from the reference's scene
files step through `new World()` / `addBody` / `step(dt)` in Node and every step's
`world.transforms` hashes to the reference's (SHA-256 of the Float32Array), with the awake,
contact and frozen counts, `previous`, and `moved()` since the step before and since seven
before, each list in the reference's order.
"""
import json
import os
import shutil
import subprocess
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import fixtures  # noqa: E402
from worldcheck import init_columns  # noqa: E402

SCRIPT = os.path.join(HERE, "js", "run_scene.js")
SCENES = os.path.join(fixtures.DATA, "scenes")
ADDON = os.path.join(HERE, "..", "bearing_amd", "_build", "bearing_napi.node")

NODE = shutil.which("node")
needs_node = pytest.mark.skipif(NODE is None, reason="no node on PATH")


def node(*args, timeout=60):
    out = subprocess.run([NODE, SCRIPT, *args], capture_output=True, text=True, timeout=timeout)
    assert out.returncode == 0, out.stderr[-2000:]
    return out.stdout


@needs_node
@pytest.mark.parametrize("name", ["pyramid", "tower", "rain_small", "pile_small", "throw"])
def test_js_resolve_matches_the_reference_columns(name):
    fx = fixtures.load(os.path.join(fixtures.DATA, f"{name}.traj.bmfx.gz"))
    want = init_columns(fx)
    got = json.loads(node("resolve", os.path.join(SCENES, f"{name}.json")))
    assert got["n"] == len(want["mass"])
    for col, ref in want.items():
        if col not in got:  # a frame the reference gave none of these bodies
            assert col in ("framed", "offset", "frame") and not np.asarray(want.get("framed", [0])).any(), col
            continue
        mine = np.asarray(got[col], ref.dtype).reshape(ref.shape)
        assert mine.tobytes() == ref.tobytes(), f"{name}: column {col} differs"


@needs_node
def test_js_resolve_refuses_what_it_does_not_take():
    bad = subprocess.run([NODE, "-e", "require(process.argv[1]).resolve('addBody', [{shape: 'capsule'}])",
                          os.path.join(HERE, "..", "bearing_amd", "js", "resolve.js")], capture_output=True, text=True)
    assert bad.returncode != 0 and "addBody: a capsule is not taken" in bad.stderr


@needs_node
@pytest.mark.skipif(not os.path.exists(ADDON), reason="addon not built (no Node headers)")
def test_js_addon_loads_and_fails_loudly_without_a_gpu():
    import torch

    r = json.loads(node("load"))
    assert r["exports"] == ["create", "add", "step", "read", "moved", "times", "destroy"]
    if not torch.cuda.is_available():
        assert r["threw"].startswith("World: no HIP device 0")


@pytest.mark.gpu
@needs_node
@pytest.mark.parametrize("name", ["pyramid", "tower", "rain_small", "pile_small", "throw"])
def test_gpu_js_world_steps_the_reference_trajectory(name):
    fx = fixtures.load(os.path.join(fixtures.DATA, f"{name}.traj.bmfx.gz"))
    meta = fx.meta
    print(f"js world {name}: {meta['steps']} steps", flush=True)
    lines = node("steps", os.path.join(SCENES, f"{name}.json"), str(meta["steps"]), repr(meta["dt"]),
                 timeout=110).splitlines()
    assert len(lines) == meta["steps"] + 1
    moved = meta["moved"]
    assert json.loads(lines[0]) == [moved["added"]], f"{name}: moved(0) before the first step differs"
    for s, line in enumerate(lines[1:], 1):
        h, awake, contacts, frozen, last, older, previous = json.loads(line)
        want = (meta["hashes"][s - 1], int(fx["awake"][s - 1]), int(fx["contacts"][s - 1]), int(fx["frozen"][s - 1]))
        assert (h, awake, contacts, frozen) == want, f"{name}: step {s} differs from the reference"
        assert previous == moved["previous"][s - 1], f"{name}: step {s}: previous differs from the reference"
        assert last == moved["last"][s - 1], f"{name}: step {s}: moved({s - 1}) differs from the reference"
        assert older == moved["older"][s - 1], f"{name}: step {s}: moved({max(0, s - 7)}) differs from the reference"
