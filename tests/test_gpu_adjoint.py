"""The adjoint on the MI355X (bm_adjoint_* in the built library: the held forward and
reverse kernels, one workgroup a world): its gradients against central finite differences
of the same held trajectory (h = 1e-6), and its results against the same arithmetic run on
the host (tests/hostworld) bit for bit."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

from adjointcheck import SCENES, agree, fd_check, record
from bearing_amd import native, scenes

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.parametrize("name", sorted(SCENES))
def test_gpu_adjoint_matches_finite_differences(name):
    print("adjoint fd", name, flush=True)
    bodies, steps = SCENES[name]
    w = record(bodies, steps, device=0)
    adj = native.Adjoint([w], every=10, device=0)
    bad = agree(fd_check(adj))
    adj.close()
    w.close()
    assert not bad, bad


def test_gpu_adjoint_is_the_hosts_bit_for_bit():
    # The device runs adjoint.h's arithmetic as the host does (-ffp-contract=off): a batch of
    # held worlds gives the same final state and gradients on both, bit for bit.
    print("adjoint device vs host", flush=True)
    subprocess.run(["make", "-s", "-C", os.path.join(HERE, "hostworld")], check=True)
    host = ctypes.CDLL(os.path.join(HERE, "hostworld", "_build", "libhostworld.so"))
    names = sorted(SCENES)
    gw = [record(SCENES[n][0], 30, device=0) for n in names]
    hw = [record(SCENES[n][0], 30, lib=host) for n in names]
    g, h = native.Adjoint(gw, every=7, device=0), native.Adjoint(hw, every=7, lib=host)
    rg, rh = g.recorded(), h.recorded()
    for key in rg:
        assert np.array_equal(rg[key], rh[key]), key
    og = g.forward(rg["state"], rg["mass"], rg["friction"], rg["restitution"])
    oh = h.forward(rh["state"], rh["mass"], rh["friction"], rh["restitution"])
    assert np.array_equal(og, oh)
    assert np.array_equal(g.branches(), h.branches())
    W = np.random.default_rng(2).normal(size=og.shape)
    gg, gh = g.backward(W), h.backward(W)
    for key in gg:
        assert np.array_equal(gg[key], gh[key]), key
    g.close(); h.close()


def test_gpu_adjoint_throws_batch():
    # The differentiation form's scene (a thrown box, a 6x6 wall, dt = 1/120): 32 worlds
    # recorded together in a batch, held and differentiated on the device, one workgroup a
    # world, the worlds running apart (different segments at once): the same final state and
    # gradients as the host's, bit for bit, on every run; the held trajectory the recorded one
    # to rounding.
    print("adjoint throws batch", flush=True)
    subprocess.run(["make", "-s", "-C", os.path.join(HERE, "hostworld")], check=True)
    host = ctypes.CDLL(os.path.join(HERE, "hostworld", "_build", "libhostworld.so"))
    specs = scenes.throws(32, seed=5)

    def recorded(lib):
        batch = native.WorldBatch(0, threads=4) if lib is None else native.WorldBatch(threads=4, lib=lib)
        ws = []
        for bodies in specs:
            w = batch.world(sleeping=False)
            w.add(scenes.resolve(bodies))
            w.record(True)
            ws.append(w)
        for _ in range(120):
            batch.step(1 / 120)
        real = np.concatenate([w.read()["position"].reshape(-1, 3) for w in ws])
        return batch, ws, real

    bg, wg, real = recorded(None)
    bh, wh, _ = recorded(host)
    g = native.Adjoint(wg, every=10, device=0)
    h = native.Adjoint(wh, every=10, lib=host)
    rec = g.recorded()
    W = np.random.default_rng(4).normal(size=(g.bodies, native.HELD_STATE))
    oh = h.forward(rec["state"], rec["mass"], rec["friction"], rec["restitution"])
    gh = h.backward(W)
    for run in range(3):
        og = g.forward(rec["state"], rec["mass"], rec["friction"], rec["restitution"])
        gg = g.backward(W)
        assert np.array_equal(og, oh), run
        for key in gg:
            assert np.array_equal(gg[key], gh[key]), (run, key)
    assert np.abs(og[:, :3] - real).max() < 1e-8
    g.close(); h.close(); bg.close(); bh.close()


@pytest.mark.timeout(600)
def test_gpu_adjoint_throws_matches_finite_differences():
    # The differentiation form as configured, on the first 8 of its 512 worlds (scenes.throws,
    # seed 5; dt 1/120, 120 steps, checkpoints every 10; the loss the squared distance of the
    # box's final centre to its seeded target): the device's gradients with respect to the
    # box's initial velocity and spin, its mass, its and the ground's friction and the ground's
    # restitution against central differences at h = 1e-6 (extrapolated with h / 2) within
    # 1e-6 relative, on every input whose passes at +-h and +-h/2 took the recorded pass's
    # branches (bm_adjoint_branches: the active set is the same there).
    from adjointcheck import record_throws, throws_agree, throws_fd_check
    print("adjoint fd, 8 thrown-box worlds", flush=True)
    batch, ws = record_throws(8, device=0)
    adj = native.Adjoint(ws, every=10, device=0)
    rows = throws_fd_check(adj, scenes.throw_targets(512, seed=5)[:8])
    adj.close()
    batch.close()
    same = [r for r in rows if r[5]]
    print(f"{len(same)} of {len(rows)} inputs kept their branches", flush=True)
    assert len(same) >= 24, f"only {len(same)} of {len(rows)} inputs kept their branches at +-h"
    bad = throws_agree(rows)
    assert not bad, bad


def test_gpu_adjoint_is_the_engines_gradient_for_a_rolling_ball():
    # Against finite differences of the real engine on the device (each input perturbed in the
    # scene, the world stepped afresh through bm_world).
    from adjointcheck import world_fd_check
    print("adjoint vs engine fd, roll", flush=True)
    bodies, steps = SCENES["roll"]
    rows = world_fd_check(bodies, steps, device=0)
    assert all(r[3] for r in rows), [r[0] for r in rows if not r[3]]
    bad = [r for r in rows if abs(r[1] - r[2]) > max(1e-6 * abs(r[2]), 3e-7)]
    assert not bad, bad


@pytest.mark.parametrize("name", ["slide", "roll", "bounce"])
def test_gpu_adjoint_is_autograd_of_a_torch_transcription(name):
    # The device's adjoint against torch.autograd through the plain FP64 PyTorch transcription of
    # the held step (tests/torchheld.py, from the reference's JavaScript), run on the CPU.
    from adjointcheck import torch_check
    print("adjoint vs torch transcription", name, flush=True)
    bodies, steps = SCENES[name]
    w = record(bodies, steps, device=0)
    adj = native.Adjoint([w], every=10, device=0)
    d = torch_check(adj)
    adj.close()
    w.close()
    assert d["forward"] < 1e-13, d
    for key in ("state", "friction", "restitution"):
        diff, most = d[key]
        assert diff <= 1e-12 * max(most, 1.0), (key, d)
    assert d["mass"][0] < 1e-15, d
