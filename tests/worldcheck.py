"""Steps a bm_world through a reference trajectory and compares every step's
world.transforms with the reference's (SHA-256 of the Float32Array, per step), and its
awake, contact and frozen counts."""
import hashlib

import numpy as np

from bearing_amd import native

# init.* fixture names for bm_bodies fields that the reference spells in camel case.
CAMEL = {"inv_inertia": "invInertia", "inv_mass": "invMass", "gravity_scale": "gravityScale",
         "angular_damping": "angularDamping"}


def init_columns(fx):
    """addBody's resolved columns of a trajectory fixture, as World.add takes them."""
    return {name: np.asarray(fx["init." + CAMEL.get(name, name)]).astype(dtype)
            for name, _, dtype in native.BODY_COLUMNS}


def run(world, fx, steps=None):
    """Steps `world` along fixture `fx`; returns (steps matched, the first mismatch or None)."""
    meta = fx.meta
    n = meta["steps"] if steps is None else min(steps, meta["steps"])
    for s in range(1, n + 1):
        world.step(meta["dt"])
        r = world.read()
        got = hashlib.sha256(r["transforms"].tobytes()).hexdigest()
        counts = (r["awake"], r["contacts"], r["frozen"])
        want = (int(fx["awake"][s - 1]), int(fx["contacts"][s - 1]), int(fx["frozen"][s - 1]))
        if got != meta["hashes"][s - 1] or counts != want:
            return s - 1, f"step {s}: transforms {'match' if got == meta['hashes'][s - 1] else 'differ'}, " \
                          f"awake/contacts/frozen {counts}, reference {want}"
    return n, None
