"""Steps a bm_world through a reference trajectory and compares every step's
world.transforms with the reference's (SHA-256 of the Float32Array, per step), and its
awake, contact and frozen counts."""
import hashlib
import os

import pytest

import numpy as np

from bearing_amd import native

# init.* fixture names for bm_bodies fields that the reference spells in camel case.
CAMEL = {"inv_inertia": "invInertia", "inv_mass": "invMass", "gravity_scale": "gravityScale",
         "angular_damping": "angularDamping"}


def init_columns(fx):
    """addBody's resolved columns of a trajectory fixture, as World.add takes them; the
    hull and frame columns only where the fixture has hull bodies."""
    return {name: np.asarray(fx["init." + CAMEL.get(name, name)]).astype(dtype)
            for name, _, dtype in native.BODY_COLUMNS if "init." + CAMEL.get(name, name) in fx}


def hull_table(fx):
    """The hull table of a trajectory fixture with hull bodies, or None."""
    if "hulls.hullV" not in fx:
        return None
    return {k: fx["hulls." + k] for k in ("hullV", "hullF", "hullL", "vertices", "start", "loops", "planes")}


def make_world(fx, **kw):
    """A native.World holding a trajectory fixture's bodies (and hulls and heightfields), as
    addBody left them."""
    w = native.World(**kw)
    table = hull_table(fx)
    if table is not None:
        w.set_hulls(table)
    if "fields.columns" in fx:
        cols, rows = fx["fields.columns"], fx["fields.rows"]
        sizes = [int(c) * int(r) for c, r in zip(cols, rows)]
        heights, at = [], 0
        for n in sizes:
            heights.append(fx["fields.heights"][at:at + n])
            at += n
        w.set_fields(cols, rows, fx["fields.spacing"], heights)
    w.add(init_columns(fx))
    return w


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def moved_mismatch(world, fx, s, r):
    """After step s (0: before the first), world.previous and world.moved() against the
    reference's (the fixture's SHA-256 of each; the moved lists' order counts); None or why."""
    rec = fx.meta.get("moved")
    if rec is None:
        return None
    if s == 0:
        return None if sha(world.moved(0)) == rec["added"] else "before the first step: moved(0) differs"
    if sha(r["previous"]) != rec["previous"][s - 1]:
        return f"step {s}: previous differs"
    if sha(world.moved(s - 1)) != rec["last"][s - 1]:
        return f"step {s}: moved({s - 1}) differs: {world.moved(s - 1)[:12]}"
    if sha(world.moved(max(0, s - 7))) != rec["older"][s - 1]:
        return f"step {s}: moved({max(0, s - 7)}) differs"
    return None


class Events:
    """The reference's contact events along a trajectory (fixtures.events_of), step by step:
    each step's count and SHA-256 (over [type, slot a, slot b, point 3, normal 3, speed] an
    event), and every event where the fixture kept them."""

    def __init__(self, fx):
        import fixtures
        self.ev = fixtures.events_of(fx)
        self.at = 0

    def mismatch(self, world, s):
        """After step s: the world's events against the reference's; None or why."""
        if self.ev is None:
            return None
        got = world.events()
        n = int(self.ev["counts"][s - 1])
        want = None
        if "events" in self.ev:
            want = self.ev["events"][10 * self.at:10 * (self.at + n)].reshape(n, 10)
        self.at += n
        if sha(got) == self.ev.meta["hashes"][s - 1]:
            return None
        if len(got) != n:
            return f"step {s}: {len(got)} contact events, reference {n}"
        if want is not None:
            bad = np.nonzero(np.any(got != want, axis=1))[0]
            return f"step {s}: contact event {bad[0]} is {got[bad[0]].tolist()}, reference {want[bad[0]].tolist()}"
        return f"step {s}: contact events differ"


def run(world, fx, steps=None):
    """Steps `world` along fixture `fx`; returns (steps matched, the first mismatch or None).
    Each step's transforms, counts, previous and moved() lists, and contact events where the
    fixture has them, are compared."""
    meta = fx.meta
    n = meta["steps"] if steps is None else min(steps, meta["steps"])
    why = moved_mismatch(world, fx, 0, None)
    if why:
        return 0, why
    events = Events(fx)
    for s in range(1, n + 1):
        world.step(meta["dt"])
        r = world.read()
        got = hashlib.sha256(r["transforms"].tobytes()).hexdigest()
        counts = (r["awake"], r["contacts"], r["frozen"])
        want = (int(fx["awake"][s - 1]), int(fx["contacts"][s - 1]), int(fx["frozen"][s - 1]))
        if got != meta["hashes"][s - 1] or counts != want:
            return s - 1, f"step {s}: transforms {'match' if got == meta['hashes'][s - 1] else 'differ'}, " \
                          f"awake/contacts/frozen {counts}, reference {want}"
        why = moved_mismatch(world, fx, s, r) or events.mismatch(world, s)
        if why:
            return s - 1, why
    return n, None


def refuses(new_world):
    """addBody fails loudly on what a world does not take: a compound (shape 4), a hull or a
    heightfield naming no table entry, and a framed box (a body with parts)."""
    import fixtures
    fx = fixtures.load(os.path.join(fixtures.DATA, "tower.traj.bmfx.gz"))
    cols = init_columns(fx)
    cols["shape"] = cols["shape"].copy()
    cols["shape"][3] = 4
    w = new_world()
    with pytest.raises(native.BearingError, match="boxes, spheres, capsules, hulls and heightfields"):
        w.add(cols)
    cols["shape"][3] = 6
    cols["field"] = np.zeros(len(cols["shape"]), np.int32)
    with pytest.raises(native.BearingError, match="names no field"):
        w.add(cols)
    cols.pop("field")
    cols["shape"][3] = 3
    cols["hull"] = np.zeros(len(cols["shape"]), np.int32)
    with pytest.raises(native.BearingError, match="names no hull"):
        w.add(cols)
    cols = init_columns(fx)
    cols["framed"] = cols["framed"].copy()
    cols["framed"][3] = 1
    cols["offset"] = np.zeros(3 * len(cols["shape"]))
    cols["frame"] = np.tile([0.0, 0.0, 0.0, 1.0], len(cols["shape"]))
    with pytest.raises(native.BearingError, match="has parts"):
        w.add(cols)
    assert w.read()["top"] == 0
    w.close()


def run_together(batch, worlds, fxs, steps):
    """Steps a native.WorldBatch `steps` times; after each step every world's transforms hash
    and counts must be its own fixture's. Returns (steps matched, the first mismatch or None)."""
    dt = fxs[0].meta["dt"]
    assert all(f.meta["dt"] == dt for f in fxs)
    events = [Events(fx) for fx in fxs]
    for s in range(1, steps + 1):
        batch.step(dt)
        for i, (w, fx) in enumerate(zip(worlds, fxs)):
            r = w.read()
            got = hashlib.sha256(r["transforms"].tobytes()).hexdigest()
            counts = (r["awake"], r["contacts"], r["frozen"])
            want = (int(fx["awake"][s - 1]), int(fx["contacts"][s - 1]), int(fx["frozen"][s - 1]))
            if got != fx.meta["hashes"][s - 1] or counts != want:
                return s - 1, f"world {i}, step {s}: transforms {'match' if got == fx.meta['hashes'][s - 1] else 'differ'}, " \
                              f"awake/contacts/frozen {counts}, reference {want}"
            why = moved_mismatch(w, fx, s, r) or events[i].mismatch(w, s)
            if why:
                return s - 1, f"world {i}, {why}"
    return steps, None


def fill_world(w, fx):
    """Hulls, fields and bodies of a trajectory fixture into an empty native.World."""
    table = hull_table(fx)
    if table is not None:
        w.set_hulls(table)
    if "fields.columns" in fx:
        cols, rows = fx["fields.columns"], fx["fields.rows"]
        heights, at = [], 0
        for c, r in zip(cols, rows):
            heights.append(fx["fields.heights"][at:at + int(c) * int(r)])
            at += int(c) * int(r)
        w.set_fields(cols, rows, fx["fields.spacing"], heights)
    w.add(init_columns(fx))
    return w


def replay(captured, device=0, lib=None, workers=0):
    """One step's tapped device-phase columns (World.capture_next) through a context's
    bm_upload, bm_colour, bm_soft_step and bm_download: on `device`, or on `lib` (the host
    library, the oracle answering, with `workers` more threads); the columns after, by name,
    and the soft step's seconds (the call alone)."""
    import time
    from bearing_amd import native
    ctx = native.Context(device, lib=lib)
    batch = native.Batch([{k: v.copy() for k, v in captured.items()}])
    if lib is not None:
        lib.stub_soft_workers(workers)
    try:
        batch.upload(ctx)
        batch.colour(ctx)
        t = time.perf_counter()
        batch.step(ctx)
        seconds = time.perf_counter() - t
        batch.download(ctx)
    finally:
        if lib is not None:
            lib.stub_soft_workers(0)
        ctx.close()
    out = batch.world(0)
    out["solved"], out["starts"] = batch.arrays["solved"].copy(), batch.arrays["starts"].copy()
    return out, seconds
