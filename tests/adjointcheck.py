"""Scenes and finite-difference checks for the adjoint (bm_adjoint_*), shared by the CPU
tests (the host library, tests/hostworld) and the GPU tests (the built library)."""
import numpy as np

from bearing_amd import native, scenes

GROUND = {"type": "static", "shape": "box", "size": [20, 1, 20], "position": [0, -0.5, 0]}

# Scenes whose active set (contacts, Murty sets, friction cone, clamps, bounces) holds
# within +-h of the recorded values: every body starts touching or clear of contact with
# generic velocities and spins, so no quantity sits exactly on a branch point.
SCENES = {
    # a cube sliding to rest on a 4-point face contact: friction cone, then sticking
    "slide": ([dict(GROUND, friction=0.4),
               {"shape": "box", "size": [1, 1, 1], "position": [0, 0.499, 0], "velocity": [3.0, 0, 0.5]}], 30),
    # a ball rolling and spinning: 1-point rows, roll and spin friction
    "roll": ([dict(GROUND, friction=0.4, restitution=0.3),
              {"shape": "sphere", "radius": 0.3, "position": [0, 0.299, 0], "velocity": [2.0, 0, 0.5],
               "angularVelocity": [0, 1, 0]}], 40),
    # a lopsided box sliding and turning: the tumble and its untwist
    "tumble": ([dict(GROUND, friction=0.4),
                {"shape": "box", "size": [0.6, 0.8, 1.2], "position": [0, 0.399, 0], "velocity": [2.0, 0, 0.3],
                 "angularVelocity": [0, 0.5, 0]}], 40),
    # a spinning ball dropped onto a restitutive floor: a speculative contact, the bounce
    "bounce": ([dict(GROUND, friction=0.4, restitution=0.5),
                {"shape": "sphere", "radius": 0.3, "position": [0, 1.0, 0], "velocity": [0.5, -1.0, 0.2],
                 "angularVelocity": [0.3, 0.7, -0.4], "restitution": 0.1}], 40),
}


def record(bodies, steps, dt=1 / 60, **kw):
    """A world of `bodies` (scenes.resolve) stepped `steps` times with sleeping off, recorded."""
    w = native.World(sleeping=False, **kw)
    w.add(scenes.resolve(bodies))
    w.record(True)
    for _ in range(steps):
        w.step(dt)
    return w


def weights(adj, seed=1):
    """A fixed random weighting of every dynamic body's final state: the loss is sum(W * state)."""
    rng = np.random.default_rng(seed)
    return rng.normal(size=(adj.bodies, native.HELD_STATE))


# (what, index into it, name): the inputs whose gradients are checked, for body 1 and the ground
CHECKS = [("state", (1, 7), "v0x"), ("state", (1, 8), "v0y"), ("state", (1, 9), "v0z"),
          ("state", (1, 10), "w0x"), ("state", (1, 11), "w0y"), ("state", (1, 12), "w0z"),
          ("state", (1, 0), "p0x"), ("state", (1, 1), "p0y"),
          ("friction", (0,), "friction ground"), ("friction", (1,), "friction body"),
          ("restitution", (0,), "restitution ground"), ("mass", (1,), "mass")]


def fd_check(adj, h=1e-6, seed=1, checks=CHECKS):
    """Each checked input's adjoint against a central finite difference of the loss:
    (name, adjoint, fd) rows."""
    rec = adj.recorded()
    W = weights(adj, seed)
    W[0] = 0  # the ground does not move

    def loss(args):
        return float(np.sum(W * adj.forward(args["state"], args["mass"], args["friction"], args["restitution"])))

    loss(rec)
    g = adj.backward(W)
    rows = []
    for what, idx, name in checks:
        args = {k: np.array(rec[k]) for k in ("state", "mass", "friction", "restitution")}
        a = args[what]
        base = a[idx]
        step = h * max(1.0, abs(base))
        a[idx] = base + step
        up = loss(args)
        a[idx] = base - step
        dn = loss(args)
        rows.append((name, float(g[what][idx]), (up - dn) / (2 * step)))
    return rows


def agree(rows, rel=1e-6, floor=1e-7):
    """The rows whose adjoint and finite difference differ by more than rel (relative) and
    floor (absolute: finite differences at h = 1e-6 carry about 1e-10 of rounding a unit of
    loss, more through a contact's stiff response)."""
    return [(n, a, f) for n, a, f in rows if abs(a - f) > max(rel * abs(f), floor)]
