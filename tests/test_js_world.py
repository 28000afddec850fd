"""The JavaScript host path: bearing_amd/js (World, addBody's resolve) over the N-API addon
(bearing_amd/_build/bearing_napi.node) over the C ABI.

CPU: the front-end's resolved columns equal the reference's own (the trajectory fixtures'
init.* columns, written by src/world.js _write) bit for bit, and the addon loads, exports its
calls and refuses to make a world without a GPU. GPU: scenes read from the reference's scene
files (boxes, spheres, capsules, hulls given as the reference built them, heightfields; the
20,001-body mixed pile among them) step through `new World()` / `addBody` / `step(dt)` in Node and every step's
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
    if out.stderr:
        print(out.stderr[-2000:], flush=True)
    return out.stdout


def built(name):
    """The reference's built hulls of scene `name` (tests/golden/make_built.mjs), if it has hulls."""
    path = os.path.join(fixtures.DATA, f"{name}.built.bmfx.gz")
    return [path] if os.path.exists(path) else []


@needs_node
@pytest.mark.parametrize("name", ["pyramid", "tower", "rain_small", "pile_small", "throw", "capsules", "rocks",
                                  "terrain", "mixed_pile"])
def test_js_resolve_matches_the_reference_columns(name):
    # Boxes and spheres; capsules; hulls given as built (their mass, centre of mass, principal
    # axes, frame and moved hull worked out in resolve.js); static heightfields: every column
    # and the hull and heightfield tables the world hands the device, bit for bit.
    fx = fixtures.load(os.path.join(fixtures.DATA, f"{name}.traj.bmfx.gz"))
    want = init_columns(fx)
    got = json.loads(node("resolve", os.path.join(SCENES, f"{name}.json"), *built(name), timeout=300))
    assert got["n"] == len(want["mass"])
    for col, ref in want.items():
        if col not in got:
            continue
        mine = np.asarray(got[col], ref.dtype).reshape(ref.shape)
        assert mine.tobytes() == ref.tobytes(), f"{name}: column {col} differs"
    for col in ("hull", "framed", "offset", "frame", "field"):  # what the reference left out is its default
        if col not in want:
            default = {"hull": -1, "field": -1}.get(col, 0)
            mine = np.asarray(got[col])
            if col == "frame":
                assert (mine.reshape(-1, 4) == [0, 0, 0, 1]).all(), f"{name}: frame"
            else:
                assert (mine == default).all(), f"{name}: {col}"
    for table, keys in (("hulls", ("hullV", "hullF", "hullL", "vertices", "start", "loops", "planes")),
                        ("fields", ("columns", "rows", "spacing", "heights"))):
        for key in keys:
            ref = fx[f"{table}.{key}"] if f"{table}.{key}" in fx else None
            mine = got[table][key]
            if ref is None:
                assert len(mine) <= 1, f"{name}: {table}.{key}"
                continue
            assert np.asarray(mine, ref.dtype).tobytes() == ref.tobytes(), f"{name}: {table}.{key} differs"


@needs_node
@pytest.mark.parametrize("body,message", [
    ("{shape: 'mesh'}", "addBody: a mesh is not taken"),
    ("{shape: 'hull', vertices: [0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0, 1]}", "addBody: a hull is taken as built"),
    ("{shape: 'capsule', radius: 1, height: 1}", "addBody: height must be at least twice radius"),
    ("{shape: 'heightfield', type: 'dynamic', heights: [0, 0, 0, 0], columns: 2, spacing: 1}",
     "addBody: type must be 'static' or 'kinematic' for a heightfield"),
    ("{shape: 'heightfield', heights: [0, 0, 0], columns: 2, spacing: 1}", "addBody: heights must be whole rows"),
    ("{shape: 'box', mass: 1, density: 2}", "addBody: give mass or density, not both"),
    ("{shape: 'box', lockRotation: true}", "addBody: lockRotation is not taken"),
])
def test_js_resolve_refuses_what_it_does_not_take(body, message):
    bad = subprocess.run([NODE, "-e", f"require(process.argv[1]).resolve('addBody', [{body}])",
                          os.path.join(HERE, "..", "bearing_amd", "js", "resolve.js")], capture_output=True, text=True)
    assert bad.returncode != 0 and message in bad.stderr, bad.stderr[-500:]


@needs_node
@pytest.mark.skipif(not os.path.exists(ADDON), reason="addon not built (no Node headers)")
def test_js_addon_loads_and_fails_loudly_without_a_gpu():
    import torch

    r = json.loads(node("load"))
    assert r["exports"] == ["create", "add", "setHulls", "setFields", "step", "read", "moved", "contacts", "times",
                            "destroy"]
    if not torch.cuda.is_available():
        assert r["threw"].startswith("World: no HIP device 0")


@pytest.mark.gpu
@needs_node
@pytest.mark.parametrize("name", ["pyramid", "tower", "rain_small", "pile_small", "throw", "capsules", "rocks",
                                  "terrain", "mixed_pile"])
@pytest.mark.timeout(400)
def test_gpu_js_world_steps_the_reference_trajectory(name):
    # mixed_pile is the 20,001-body pile of boxes, balls, capsules and built hulls on a 256 x 256
    # heightfield, all 600 of its steps, driven from Node.
    fx = fixtures.load(os.path.join(fixtures.DATA, f"{name}.traj.bmfx.gz"))
    meta = fx.meta
    print(f"js world {name}: {meta['steps']} steps", flush=True)
    lines = node("steps", os.path.join(SCENES, f"{name}.json"), str(meta["steps"]), repr(meta["dt"]), *built(name),
                 timeout=360).splitlines()
    assert len(lines) == meta["steps"] + 1
    # The mixed pile's trajectory holds the transforms and counts only, not the moved() lists.
    moved = meta.get("moved")
    assert moved is None or json.loads(lines[0]) == [moved["added"]], f"{name}: moved(0) before the first step differs"
    events = fixtures.events_of(fx)
    for s, line in enumerate(lines[1:], 1):
        h, awake, contacts, frozen, last, older, previous, ev = json.loads(line)
        want = (meta["hashes"][s - 1], int(fx["awake"][s - 1]), int(fx["contacts"][s - 1]), int(fx["frozen"][s - 1]))
        assert (h, awake, contacts, frozen) == want, f"{name}: step {s} differs from the reference"
        assert events is None or ev == events.meta["hashes"][s - 1], f"{name}: step {s}: contact events differ"
        if moved is None:
            continue
        assert previous == moved["previous"][s - 1], f"{name}: step {s}: previous differs from the reference"
        assert last == moved["last"][s - 1], f"{name}: step {s}: moved({s - 1}) differs from the reference"
        assert older == moved["older"][s - 1], f"{name}: step {s}: moved({max(0, s - 7)}) differs from the reference"
