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
