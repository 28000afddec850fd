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
