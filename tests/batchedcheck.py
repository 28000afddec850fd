"""The batched form's worlds (scenes.batched: 4,096 worlds, body counts log-uniform in
[8, 512], even worlds box towers of height U{4..20} with each level turned U[0, 15] degrees
about y, odd worlds sphere piles; each on its own 10x1x10 ground; seed 4) stepped 300 steps
at dt 1/60, checked against the reference's own trajectories of the same worlds
(tests/golden/data/batched.traj: every world's final world.transforms hash and counts, and
every step of a sample of worlds)."""
import hashlib
import os

import fixtures
from bearing_amd import scenes


def load():
    return fixtures.load(os.path.join(fixtures.DATA, "batched.traj.bmfx.gz"))


def scene_list(fx):
    meta = fx.meta
    return scenes.batched(meta["worlds"], seed=meta["seed"])


def digest(world):
    r = world.read()
    return hashlib.sha256(r["transforms"].tobytes()).digest(), r


def final_of(fx, i):
    return bytes(fx["final"][32 * i:32 * (i + 1)])


def check_step(fx, i, s, world):
    """World i after step s (1-based) against the reference's sampled trajectory; None or why."""
    got, r = digest(world)
    want = bytes(fx[f"s{i}.hashes"][32 * (s - 1):32 * s])
    counts = (r["awake"], r["contacts"], r["frozen"])
    ref = tuple(int(fx[f"s{i}.{k}"][s - 1]) for k in ("awake", "contacts", "frozen"))
    if got != want or counts != ref:
        return (f"world {i}, step {s}: transforms {'match' if got == want else 'differ'}, "
                f"awake/contacts/frozen {counts}, reference {ref}")
    return None


def check_final(fx, i, world):
    """World i after the last step against the reference's; None or why."""
    got, r = digest(world)
    counts = (r["awake"], r["contacts"], r["frozen"], r["impacts"], r["catches"])
    ref = tuple(int(fx[k][i]) for k in ("awake", "contacts", "frozen", "impacts", "catches"))
    if got != final_of(fx, i) or counts != ref:
        return (f"world {i} after {fx.meta['steps']} steps: transforms "
                f"{'match' if got == final_of(fx, i) else 'differ'}, awake/contacts/frozen/impacts/catches "
                f"{counts}, reference {ref}")
    return None
