"""The adjoint on the CPU: bm_world recording and bm_adjoint_* with the library's own
adjoint arithmetic (bearing_amd/csrc/adjoint.h) run on the host by tests/hostworld, checked
against central finite differences of the same held trajectory (h = 1e-6), and the held
trajectory against the recorded one."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

from adjointcheck import SCENES, agree, fd_check, record
from bearing_amd import native, scenes

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "hostworld", "_build", "libhostworld.so")


@pytest.fixture(scope="module")
def lib():
    subprocess.run(["make", "-s", "-C", os.path.join(HERE, "hostworld")], check=True)
    return ctypes.CDLL(LIB)


@pytest.mark.parametrize("name", sorted(SCENES))
def test_host_adjoint_matches_finite_differences(lib, name):
    bodies, steps = SCENES[name]
    w = record(bodies, steps, lib=lib)
    adj = native.Adjoint([w], every=10, lib=lib)
    bad = agree(fd_check(adj))
    adj.close()
    assert not bad, bad


def test_host_held_trajectory_follows_the_recorded_one(lib):
    # Where the reference keeps or finds each manifold at the same place step after step
    # (no impacts), the held trajectory is the recorded one up to the rounding of the rebuild.
    bodies, steps = SCENES["slide"]
    w = record(bodies, steps, lib=lib)
    r = w.read()
    adj = native.Adjoint([w], every=7, lib=lib)
    rec = adj.recorded()
    out = adj.forward(rec["state"], rec["mass"], rec["friction"], rec["restitution"])
    adj.close()
    np.testing.assert_allclose(out[:, :3].ravel(), r["position"], rtol=0, atol=1e-12)
    np.testing.assert_allclose(out[:, 3:7].ravel(), r["rotation"], rtol=0, atol=1e-12)
    np.testing.assert_allclose(out[:, 7:10].ravel(), r["velocity"], rtol=0, atol=1e-10)


def test_host_adjoint_of_several_worlds_is_each_alone(lib):
    # Worlds of a held batch are independent: the batch's gradients are each world's own.
    ws = [record(SCENES[n][0], 20, lib=lib) for n in ("slide", "roll")]
    both = native.Adjoint(ws, every=5, lib=lib)
    rec = both.recorded()
    out = both.forward(rec["state"], rec["mass"], rec["friction"], rec["restitution"])
    rng = np.random.default_rng(3)
    W = rng.normal(size=out.shape)
    g = both.backward(W)
    nb0 = 2
    for k, w in enumerate(ws):
        one = native.Adjoint([w], every=5, lib=lib)
        r1 = one.recorded()
        o1 = one.forward(r1["state"], r1["mass"], r1["friction"], r1["restitution"])
        sl = slice(k * nb0, (k + 1) * nb0)
        assert np.array_equal(o1, out[sl])
        g1 = one.backward(W[sl])
        for key in ("state", "mass", "friction", "restitution"):
            assert np.array_equal(g1[key], g[key][sl]), key
        one.close()
    both.close()


def test_host_checkpoint_interval_changes_nothing(lib):
    bodies, steps = SCENES["tumble"]
    w = record(bodies, steps, lib=lib)
    outs = []
    for every in (1, 7, 40):
        adj = native.Adjoint([w], every=every, lib=lib)
        rec = adj.recorded()
        o = adj.forward(rec["state"], rec["mass"], rec["friction"], rec["restitution"])
        g = adj.backward(np.ones_like(o))
        outs.append((o, g))
        adj.close()
    for o, g in outs[1:]:
        assert np.array_equal(o, outs[0][0])
        for key in g:
            assert np.array_equal(g[key], outs[0][1][key]), key


def test_host_record_refuses_what_it_cannot_hold(lib):
    w = native.World(lib=lib, sleeping=True)
    w.add(scenes.resolve(SCENES["slide"][0]))
    with pytest.raises(native.BearingError, match="sleeping off"):
        w.record(True)
    w2 = native.World(lib=lib, sleeping=False)
    w2.add(scenes.resolve(SCENES["slide"][0]))
    with pytest.raises(native.BearingError, match="nothing recorded"):
        native.Adjoint([w2], lib=lib)


def test_resolve_matches_the_reference(lib):
    # scenes.resolve gives addBody's resolved columns bit for bit (the fixtures' from the
    # reference), for scenes of boxes and spheres.
    import fixtures
    from worldcheck import init_columns
    for name in ("throw", "pyramid", "rain_small", "pile_small", "tower"):
        fx = fixtures.load(os.path.join(fixtures.DATA, f"{name}.traj.bmfx.gz"))
        want, got = init_columns(fx), scenes.resolve(scenes.SCENES[name]())
        for key, col in got.items():
            assert np.array_equal(col.view(np.uint8), np.asarray(want[key], col.dtype).view(np.uint8)), (name, key)


def test_host_torch_function_gives_the_adjoint(lib):
    # The torch.autograd.Function around the adjoint: its gradients are bm_adjoint_backward's,
    # and torch's own finite-difference gradcheck agrees on the moving body's inputs.
    import torch

    from bearing_amd.adjoint import Held
    bodies, steps = SCENES["bounce"]
    w = record(bodies, 24, lib=lib)
    held = Held(native.Adjoint([w], every=6, lib=lib))
    rec = held.recorded()
    s0 = rec["state"].clone().requires_grad_(True)
    m = rec["mass"].clone().requires_grad_(True)
    f = rec["friction"].clone().requires_grad_(True)
    e = rec["restitution"].clone().requires_grad_(True)
    W = torch.from_numpy(np.random.default_rng(5).normal(size=(2, 13)))
    W[0] = 0
    (held(s0, m, f, e) * W).sum().backward()
    g = held.adj.backward(W.numpy())
    for t, key in ((s0, "state"), (m, "mass"), (f, "friction"), (e, "restitution")):
        assert np.array_equal(t.grad.numpy(), g[key]), key
    ground = rec["state"][0:1]

    def body(state1, fr, re):
        state = torch.cat([ground, state1], 0)
        return held(state, rec["mass"], fr, re)[1]

    ok = torch.autograd.gradcheck(body, (rec["state"][1:2].clone().requires_grad_(True),
                                         rec["friction"].clone().requires_grad_(True),
                                         rec["restitution"].clone().requires_grad_(True)),
                                  eps=1e-6, atol=1e-6, rtol=1e-5)
    assert ok


def test_host_batch_worlds_record_as_alone(lib):
    # Worlds recorded while stepping in a batch (bm_batch_step) hold the same trajectory as
    # the same worlds recorded alone.
    names = ("slide", "roll", "tumble")
    alone = [record(SCENES[n][0], 15, lib=lib) for n in names]
    batch = native.WorldBatch(threads=2, lib=lib)
    together = []
    for n in names:
        w = batch.world(sleeping=False)
        w.add(scenes.resolve(SCENES[n][0]))
        w.record(True)
        together.append(w)
    for _ in range(15):
        batch.step(1 / 60)
    a, b = native.Adjoint(alone, every=4, lib=lib), native.Adjoint(together, every=4, lib=lib)
    ra, rb = a.recorded(), b.recorded()
    for key in ra:
        assert np.array_equal(ra[key], rb[key]), key
    oa = a.forward(ra["state"], ra["mass"], ra["friction"], ra["restitution"])
    ob = b.forward(rb["state"], rb["mass"], rb["friction"], rb["restitution"])
    assert np.array_equal(oa, ob)
    ga, gb = a.backward(np.ones_like(oa)), b.backward(np.ones_like(ob))
    for key in ga:
        assert np.array_equal(ga[key], gb[key]), key
    a.close(); b.close(); batch.close()


def test_host_torch_function_steps_again_after_another_forward(lib):
    # Two forward passes through the Function from different states, then the first one's
    # backward: it walks its own pass (run again), not the second's checkpoints, whether the
    # Adjoint was wrapped by Held or used bare.
    import torch

    from bearing_amd.adjoint import HeldTrajectory
    bodies, _ = SCENES["slide"]
    w = record(bodies, 20, lib=lib)
    adj = native.Adjoint([w], every=5, lib=lib)
    rec = {k: torch.from_numpy(np.array(v)) for k, v in adj.recorded().items()}
    s1 = rec["state"].clone().requires_grad_(True)
    s2 = rec["state"].clone()
    s2[1, 7] += 0.05   # a second pass from another initial velocity
    W = torch.from_numpy(np.random.default_rng(9).normal(size=(2, 13)))
    out1 = HeldTrajectory.apply(adj, s1, rec["mass"], rec["friction"], rec["restitution"])
    HeldTrajectory.apply(adj, s2, rec["mass"], rec["friction"], rec["restitution"])
    (out1 * W).sum().backward()
    fresh = native.Adjoint([w], every=5, lib=lib)
    fresh.forward(*(rec[k].numpy() for k in ("state", "mass", "friction", "restitution")))
    want = fresh.backward(W.numpy())
    assert np.array_equal(s1.grad.numpy(), want["state"])
    adj.close(); fresh.close()


def test_host_adjoint_throws_matches_finite_differences(lib):
    # The differentiation form's own scene (scenes.throws, seed 5: a thrown box landing and
    # sliding towards a 6x6 wall; dt 1/120, 120 steps, checkpoints every 10), its loss the
    # squared distance of the box's final centre to its seeded target: the gradients with
    # respect to the box's initial velocity and spin, its mass, its and the ground's friction
    # and the ground's restitution against central differences at h = 1e-6 (extrapolated with
    # h / 2), wherever the passes at +-h and +-h/2 took the recorded pass's branches.
    from adjointcheck import record_throws, throws_agree, throws_fd_check
    batch, ws = record_throws(4, lib=lib)
    adj = native.Adjoint(ws, every=10, lib=lib)
    rows = throws_fd_check(adj, scenes.throw_targets(512, seed=5)[:4])
    adj.close()
    batch.close()
    same = [r for r in rows if r[5]]
    assert len(same) >= 12, f"only {len(same)} of {len(rows)} inputs kept their branches at +-h"
    bad = throws_agree(rows)
    assert not bad, bad


def test_host_adjoint_branch_records_name_the_active_set(lib):
    # Two forward passes from the same inputs keep the same branch record; a box thrown slow
    # enough to stick rather than slide takes other branches; and a different recording's
    # contacts give a different record even at the same inputs.
    bodies, steps = SCENES["slide"]
    w = record(bodies, steps, lib=lib)
    adj = native.Adjoint([w], every=10, lib=lib)
    rec = adj.recorded()
    adj.forward(rec["state"], rec["mass"], rec["friction"], rec["restitution"])
    first = adj.branches()
    adj.forward(rec["state"], rec["mass"], rec["friction"], rec["restitution"])
    assert np.array_equal(adj.branches(), first)
    slow = np.array(rec["state"])
    slow[1, 7] = 0.01
    adj.forward(slow, rec["mass"], rec["friction"], rec["restitution"])
    assert adj.branches()[0] != first[0]
    other = dict(bodies[1], velocity=[0.01, 0.0, 0.0])
    w2 = record([bodies[0], other], steps, lib=lib)
    adj2 = native.Adjoint([w2], every=10, lib=lib)
    adj2.forward(rec["state"], rec["mass"], rec["friction"], rec["restitution"])
    assert adj2.branches()[0] != first[0]
    adj.close(); adj2.close()


def test_host_adjoint_is_the_engines_gradient_for_a_rolling_ball(lib):
    # The held trajectory's gradient against finite differences of the real engine (each input
    # perturbed in the scene, the world stepped afresh through bm_world): a ball rolling and
    # spinning on a floor, whose contact the engine keeps finding at the same place relative to
    # the ball, agrees on every input; every perturbed recording took the recorded branches.
    from adjointcheck import world_fd_check
    bodies, steps = SCENES["roll"]
    rows = world_fd_check(bodies, steps, lib=lib)
    assert all(r[3] for r in rows), [r[0] for r in rows if not r[3]]
    bad = [r for r in rows if abs(r[1] - r[2]) > max(1e-6 * abs(r[2]), 3e-7)]
    assert not bad, bad


def test_host_adjoint_and_the_engine_part_where_contacts_are_found_afresh(lib):
    # Where the engine finds a sliding face contact afresh each step, or meets a body at the
    # moment of impact, its contact ends and carry back follow the state, while the held
    # trajectory keeps each step's ends fixed in the bodies' frames and the carry back fixed:
    # translating the sliding cube's start shows it (the engine's trajectory translates with
    # it exactly, so d loss / d x0 is the loss's weight on the final x; the held one's is not).
    from adjointcheck import world_fd_check
    bodies, steps = SCENES["slide"]
    rows = {r[0]: r for r in world_fd_check(bodies, steps, lib=lib)}
    W = np.random.default_rng(1).normal(size=(2, native.HELD_STATE))
    assert abs(rows["p0x"][2] - W[1, 0]) < 1e-6 * abs(W[1, 0])      # the engine: translation
    assert abs(rows["p0x"][1] - W[1, 0]) > 1e-4 * abs(W[1, 0])       # the held model: not


@pytest.mark.parametrize("name", ["slide", "roll", "bounce"])
def test_host_adjoint_is_autograd_of_a_torch_transcription(lib, name):
    # torch.autograd through a plain FP64 PyTorch transcription of the held step, written from
    # the reference's JavaScript (tests/torchheld.py), against bm_adjoint_backward: the same
    # final state and the same gradients to rounding (a box face's block LCP with Murty
    # pivoting, a ball's rolling and spinning friction, a bounce).
    from adjointcheck import torch_check
    bodies, steps = SCENES[name]
    w = record(bodies, steps, lib=lib)
    adj = native.Adjoint([w], every=10, lib=lib)
    d = torch_check(adj)
    adj.close()
    assert d["forward"] < 1e-13, d
    for key in ("state", "friction", "restitution"):
        diff, most = d[key]
        assert diff <= 1e-12 * max(most, 1.0), (key, d)
    assert d["mass"][0] < 1e-15, d   # both ~0: a body on a static floor moves as it would at any mass
