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


# The differentiation form's checks (scenes.throws: a thrown box, a 6x6 wall of unit boxes, a
# ground; dt 1/120, 120 steps): the box is each world's last body, the ground its first.
THROW_BODIES, THROW_BOX = 38, 37
# (what, column of the state or None, body within the world, name)
THROW_CHECKS = [("state", 7, THROW_BOX, "v0x"), ("state", 8, THROW_BOX, "v0y"), ("state", 9, THROW_BOX, "v0z"),
                ("state", 10, THROW_BOX, "w0x"), ("state", 11, THROW_BOX, "w0y"), ("state", 12, THROW_BOX, "w0z"),
                ("mass", None, THROW_BOX, "mass"), ("friction", None, THROW_BOX, "friction box"),
                ("friction", None, 0, "friction ground"), ("restitution", None, 0, "restitution ground")]


def record_throws(worlds, seed=5, steps=120, lib=None, device=0, threads=4):
    """The first `worlds` of scenes.throws(512, seed), stepped together with sleeping off and
    recorded; returns (batch, worlds)."""
    specs = scenes.throws(512, seed=seed)[:worlds]
    batch = native.WorldBatch(device, threads=threads) if lib is None else native.WorldBatch(threads=threads, lib=lib)
    ws = []
    for bodies in specs:
        w = batch.world(sleeping=False)
        w.add(scenes.resolve(bodies))
        w.record(True)
        ws.append(w)
    for _ in range(steps):
        batch.step(1 / 120)
    return batch, ws


def throw_losses(adj, args, targets):
    """Each world's loss, the squared distance of its box's final centre to its target, and
    the final state."""
    out = adj.forward(args["state"], args["mass"], args["friction"], args["restitution"])
    n = len(targets)
    box = out.reshape(n, THROW_BODIES, native.HELD_STATE)[:, THROW_BOX, :3]
    return np.sum((box - np.asarray(targets)) ** 2, axis=1), out


def throws_fd_check(adj, targets, h=1e-6):
    """The differentiation form's gradients against central finite differences, every world's
    same input stepped at once (worlds are independent). Rows (world, name, adjoint, fd, raw,
    same): raw is the central difference at h; fd extrapolates those at h and h/2
    (Richardson: (4 D(h/2) - D(h)) / 3), so the differences' own h^2 error — large here, the
    contacts' response being stiff — does not count against the adjoint; same says whether
    the forward passes at +-h and +-h/2 took the branches the pass at the recorded inputs took
    (bm_adjoint_branches): only there is the finite difference the gradient's."""
    rec = adj.recorded()
    n = len(targets)
    base = {k: np.array(rec[k]) for k in ("state", "mass", "friction", "restitution")}
    _, out = throw_losses(adj, base, targets)
    b0 = adj.branches()
    bar = np.zeros((n, THROW_BODIES, native.HELD_STATE))
    box = out.reshape(n, THROW_BODIES, native.HELD_STATE)[:, THROW_BOX, :3]
    bar[:, THROW_BOX, :3] = 2 * (box - np.asarray(targets))
    g = adj.backward(bar.reshape(-1, native.HELD_STATE))
    rows = []
    for what, col, body, name in THROW_CHECKS:
        slots = np.arange(n) * THROW_BODIES + body
        diffs, same = [], np.ones(n, bool)
        for scale in (1.0, 0.5):
            losses = []
            for sign in (1, -1):
                args = {k: np.array(v) for k, v in base.items()}
                a = args[what]
                x = a[slots] if col is None else a[slots, col]
                steps = scale * h * np.maximum(1.0, np.abs(x))
                if col is None:
                    a[slots] += sign * steps
                else:
                    a[slots, col] += sign * steps
                losses.append(throw_losses(adj, args, targets)[0])
                same &= adj.branches() == b0
            diffs.append((losses[0] - losses[1]) / (2 * steps))
        fd = (4 * diffs[1] - diffs[0]) / 3
        ga = g[what][slots] if col is None else g[what][slots, col]
        for k in range(n):
            rows.append((k, name, float(ga[k]), float(fd[k]), float(diffs[0][k]), bool(same[k])))
    return rows


def throws_agree(rows, rel=1e-6, floor=3e-7):
    """Same-branch rows whose adjoint and extrapolated difference differ by more than rel
    (relative) and floor (absolute: the loss carries ~1e-13 of rounding after 120 steps, which
    the extrapolated difference at h = 1e-6 turns into ~1e-7)."""
    return [r for r in rows if r[5] and abs(r[2] - r[3]) > max(rel * abs(r[3]), floor)]


# The real engine's inputs checked: (name, body, how to perturb the scene's body by d).
def _shift(key, k):
    def apply(b, d):
        v = list(b.get(key, [0.0, 0.0, 0.0]))
        v[k] += d
        b[key] = v
    return apply


def _set(key, default):
    def apply(b, d):
        b[key] = b.get(key, default) + d
    return apply


WORLD_CHECKS = [("v0x", 1, _shift("velocity", 0)), ("v0y", 1, _shift("velocity", 1)), ("v0z", 1, _shift("velocity", 2)),
                ("w0x", 1, _shift("angularVelocity", 0)), ("w0y", 1, _shift("angularVelocity", 1)),
                ("w0z", 1, _shift("angularVelocity", 2)), ("p0x", 1, _shift("position", 0)),
                ("friction ground", 0, _set("friction", 0.5)), ("friction body", 1, _set("friction", 0.5)),
                ("restitution ground", 0, _set("restitution", 0.0)), ("mass", 1, None)]
# Where each checked input sits among the adjoint's gradients.
WORLD_GRADIENT = {"v0x": ("state", (1, 7)), "v0y": ("state", (1, 8)), "v0z": ("state", (1, 9)),
                  "w0x": ("state", (1, 10)), "w0y": ("state", (1, 11)), "w0z": ("state", (1, 12)),
                  "p0x": ("state", (1, 0)), "friction ground": ("friction", (0,)),
                  "friction body": ("friction", (1,)), "restitution ground": ("restitution", (0,)),
                  "mass": ("mass", (1,))}


def world_fd_check(bodies, steps, h=1e-6, seed=1, **kw):
    """The adjoint of the held trajectory against central finite differences of the real
    engine: each input perturbed in the scene itself, the world built, stepped and recorded
    afresh (bm_world: broadphase, narrowphase, colouring, soft step), the loss read from its
    final state. Rows (name, adjoint, fd, same) as throws_fd_check's: fd extrapolates the
    differences at h and h/2; same says whether every perturbed recording, held, took the
    branches (contacts, feature ids, active set) the recorded one took."""
    import copy

    def run(spec):
        w = record(spec, steps, **kw)
        r = w.read()
        state = np.zeros((len(spec), native.HELD_STATE))
        state[:, 0:3] = r["position"].reshape(-1, 3)
        state[:, 3:7] = r["rotation"].reshape(-1, 4)
        state[:, 7:10] = r["velocity"].reshape(-1, 3)
        state[:, 10:13] = r["angular"].reshape(-1, 3)
        adj = native.Adjoint([w], every=10, **{k: v for k, v in kw.items() if k in ("lib", "device")})
        rec = adj.recorded()
        adj.forward(rec["state"], rec["mass"], rec["friction"], rec["restitution"])
        branch = adj.branches()[0]
        return w, adj, state, branch

    w0, adj0, state0, b0 = run(bodies)
    W = np.random.default_rng(seed).normal(size=state0.shape)
    W[0] = 0
    g = adj0.backward(W)
    mass0 = adj0.recorded()["mass"][1]
    adj0.close()
    w0.close()
    rows = []
    for name, body, apply in WORLD_CHECKS:
        diffs, same = [], True
        for scale in (1.0, 0.5):
            losses = []
            for sign in (1, -1):
                spec = copy.deepcopy(bodies)
                if name == "mass":
                    step = scale * h * max(1.0, mass0)
                    spec[body]["mass"] = mass0 + sign * step
                else:
                    base = {"friction": 0.5, "restitution": 0.0}.get(name.split()[0], None)
                    x = spec[body].get(name.split()[0], base) if base is not None else None
                    step = scale * h * max(1.0, abs(x) if x is not None else 1.0)
                    apply(spec[body], sign * step)
                w, adj, state, branch = run(spec)
                losses.append(float(np.sum(W * state)))
                same = same and branch == b0
                adj.close()
                w.close()
            diffs.append((losses[0] - losses[1]) / (2 * step))
        what, idx = WORLD_GRADIENT[name]
        rows.append((name, float(g[what][idx]), (4 * diffs[1] - diffs[0]) / 3, bool(same)))
    return rows


def torch_check(adj, seed=1):
    """The adjoint against autograd through the plain FP64 PyTorch transcription of the held step
    (tests/torchheld.py, written from the reference's JavaScript): the largest difference of the
    final states, and for each of state (dynamic bodies' rows: a static body's state is a scene
    constant, its row 0), mass, friction and restitution the largest difference of the
    gradients and the largest gradient."""
    import torch

    import torchheld
    held, rec = adj.held(), adj.recorded()
    out = adj.forward(rec["state"], rec["mass"], rec["friction"], rec["restitution"])
    ins = {k: torch.tensor(np.array(rec[k]), dtype=torch.float64, requires_grad=True)
           for k in ("state", "mass", "friction", "restitution")}
    got = torchheld.forward(held, ins["state"], ins["mass"], ins["friction"], ins["restitution"])
    W = np.random.default_rng(seed).normal(size=out.shape)
    dynamic = held["type"] == 0
    W[~dynamic] = 0
    (got * torch.from_numpy(W)).sum().backward()
    g = adj.backward(W)
    diffs = {"forward": float(np.abs(got.detach().numpy() - out).max())}
    for key, t in ins.items():
        mine = t.grad.numpy() if t.grad is not None else np.zeros_like(g[key])
        theirs = g[key]
        if key == "state":
            mine, theirs = mine[dynamic], theirs[dynamic]
        diffs[key] = (float(np.abs(mine - theirs).max()), float(np.abs(theirs).max()))
    return diffs
