"""Benchmark: Bearing's contact pipeline on one MI355X, FP64, every result the reference's bits.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--parts ...] [--cpu-seconds S]

Headline (`value`): contact-iterations per second of the plain drop-in's whole step on the
20,001-body mixed pile (8,000 boxes, 5,000 balls, 4,000 capsules and 3,000 hulls fallen from a
60 x 80 x 60 m volume onto a 256 x 256 value-noise heightfield; seed 3) at rest: the world is
stepped through bm_world_step (bounds, broadphase, narrowphase with time of impact, freezing,
colouring, soft step, catch, sleep; every copy between host and device included) along the
reference's own 600-step trajectory, and the last K steps, after W untimed warm-up steps, are
timed: rows x 2 passes x 8 substeps a step over their wall time. A "step" is one world step.
Every one of the 600 steps' transforms is checked against the reference's afterwards.

`roofline` prices the step's dominant kernel, the grid-wide soft step (inertias .. store), with
the algorithmic bytes of its sweeps (bm_world_read soft_bytes: each field once a pass, each
body's fields once a manifold end) over its HIP-event time, against HBM's 8 TB/s.
`cpu_baseline` is the oracle (the CPU restatement) on the host, on the same contact phase: the
columns the pile's last untimed step handed the device, replayed through the oracle's soft step
with one thread and with the reference's 'auto' workers (half the threads less one); `value` is
the workers' rate, `cores` their threads.

Other legs: `batched_whole_worlds`, the batched form (4,096 worlds of scenes.batched, seed 4,
300 steps, bm_batch_step) against bm_world_step looped per world, bit-exact against the
reference for every world; `world`, whole reference trajectories through bm_world;
`soft_step_microbench`, the soft step alone re-run on one fixed state of the 2,000-sphere rain
(a microbenchmark of the solve, not a physical step); `batched`, many worlds' soft steps in one
launch; `adjoint`, the differentiation form (bwd/fwd).
"""
import argparse
import json
import re
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (ROOT, os.path.join(ROOT, "tests"), os.path.join(ROOT, "tests", "golden")):
    if p not in sys.path:
        sys.path.insert(0, p)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
SUBSTEPS = 8
WORKLOAD = "rain.soft.300"
BATCH_SOURCES = ["pyramid.soft.30", "tower.soft.10", "pile_small.soft.110", "rain_small.soft.90", "throw.soft.60"]


def algorithmic_bytes(points, dynamic_a, dynamic_b):
    """Compulsory HBM bytes of one soft step's sweeps over these manifolds.

    Per manifold of n points, each field read or written once per pass, body fields once
    per endpoint (the gather and scatter every colour does):
      solve (bias, separations known):  394 fixed + 56 if n <= 2 + 8 n^2 if n >= 2 (K block)
                                        + n (152 + 8 if n == 1)
                                        + bodies: v,w read 96, written 96 (48 if one end is static),
                                          inverse masses 16, inverse inertias 96
      relax (and a step's first solve): 24 fewer fixed, 56 more per point (anchors, adjusted),
                                        + 160 body (positions, starts, rotations)
      warm start:                       201 fixed + 16 if n <= 2 + 56 n + 304 body
    A step is 8 warm starts, 7 known solves, 1 first solve and 8 relaxes.
    """
    n = points.astype(np.float64)
    both = (dynamic_a & dynamic_b)
    body = 96 + np.where(both, 96, 48) + 16 + 96
    fixed = 394 + np.where(n <= 2, 56, 0) + np.where(n >= 2, 8 * n * n, 0)
    solve = fixed + n * (152 + np.where(n == 1, 8, 0)) + body
    relax = (fixed - 24) + n * (152 + 56 + np.where(n == 1, 8, 0)) + body + 160
    warm = 201 + np.where(n <= 2, 16, 0) + 56 * n + 304
    return float(np.sum(8 * warm + 7 * solve + 1 * relax + 8 * relax))


def load_world(name):
    import fixtures
    from softcheck import world_from_columns

    fx = fixtures.load(os.path.join(fixtures.DATA, name + ".bmfx.gz"))
    return fx, fixtures.soft_columns(fx), world_from_columns


def traffic_of(kernel, workload):
    """HBM bytes per launch of `kernel` on `workload` as rocprofv3's FETCH_SIZE and WRITE_SIZE
    counters measured them (each scaled by what the counter reads for a copy of known size),
    from the records profiles/*_traffic.json keep; None if not measured."""
    import glob

    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", "*_traffic.json"))):
        rec = json.load(open(path))
        if rec.get("kernel") == kernel and rec.get("workload") == workload:
            return rec["bytes"]
    return None


def micro_run(args):
    """The soft step alone (inertias .. store, one workgroup) re-run on one fixed state of the
    2,000-sphere rain (its reference state before the step-300 solve): a microbenchmark of the
    solve. After the first run the velocities are integrated against the same contacts, so the
    later runs are not physical steps; the bits of the first are checked by the tests."""
    from bearing_amd import native
    from bearing_amd.tune import tune

    fx, cols, world_from_columns = load_world(WORKLOAD)
    meta = fx.meta
    solved = meta["solved"]
    points = fx["in.c.points"][:solved].astype(np.int64)
    rows = int(points.sum())
    inv_mass = fx["in.invMass"]
    a, b = fx["in.c.a"][:solved], fx["in.c.b"][:solved]
    step_bytes = algorithmic_bytes(points, inv_mass[a] > 0, inv_mass[b] > 0)
    plain_world = world_from_columns(fx, cols)
    if args.plain_launch:
        shape, resident = (int(x) for x in args.plain_launch.split(","))
    else:
        shape, resident, _ = tune([plain_world])
    ctx = native.Context(0)
    ctx.set_launch(shape, resident)
    batch = native.Batch([plain_world])
    batch.upload(ctx)
    for _ in range(5):
        batch.step(ctx)
    kernel_ms = []
    t0 = time.perf_counter()
    for _ in range(args.micro_steps):
        kernel_ms.append(batch.step(ctx))
    wall = time.perf_counter() - t0
    avg_kernel_ms = float(np.mean(kernel_ms))
    achieved = step_bytes / (avg_kernel_ms * 1e-3) / 1e9
    ctx.set_wide(1)
    for _ in range(2):
        batch.step(ctx)
    wide_ms = float(np.mean([batch.step(ctx) for _ in range(10)]))
    ctx.set_wide(16384)
    kernel = f"soft_step_kernel<{'true' if resident else 'false'}, {512 if shape == 2 else 256}, {1 if shape == 0 else 2}>"
    out = {"workload": f"{WORKLOAD}: the rain's reference state before its step-300 solve, re-run",
           "bodies": meta["top"], "manifolds": solved, "rows": rows,
           "contact_iters_per_s": rows * 2 * SUBSTEPS * args.micro_steps / wall,
           "ms_per_launch": wall * 1000 / args.micro_steps, "kernel_ms": avg_kernel_ms,
           "launch": {"shape": shape, "resident_lds": bool(resident)}, "grid_wide_kernel_ms": wide_ms,
           "roofline": {"kernel": kernel, "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": achieved / HBM_PEAK_GBS, "traffic": traffic_of(kernel, WORKLOAD),
                        "algorithmic_bytes_per_launch": step_bytes}}
    out["batched"] = batched_soft_run(ctx, args) if "batched" in set(args.parts.split(",")) else None
    ctx.close()
    return out


def host_threads():
    """(threads this process may use, as the machine reports them; the CPU model). The threads
    are the CPUs this process is allowed to run on, no more than OMP_NUM_THREADS where it is set
    (the share of the machine a job is given)."""
    n = len(os.sched_getaffinity(0))
    share = os.environ.get("OMP_NUM_THREADS")
    if share and share.isdigit() and int(share) > 0:
        n = min(n, int(share))
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return n, model


def cpu_baseline(captured, seconds):
    """The CPU restatement (the oracle) timed on the same contact phase the headline steps: the
    columns the pile's last untimed step handed the device (bm_world_tap), replayed through the
    host library's context (tests/hostworld: the oracle answering bm_soft_step), once with one
    thread and once with the reference's 'auto' workers, half the threads less one
    (src/world.js:173-186), each about `seconds / 2`; the soft step's call alone is timed
    (upload and colouring are not)."""
    import ctypes

    from worldcheck import replay

    host_lib = os.path.join(ROOT, "tests", "hostworld", "_build", "libhostworld.so")
    if captured is None or not os.path.exists(host_lib):
        return None
    lib = ctypes.CDLL(host_lib)
    threads, model = host_threads()
    auto = 0 if threads < 2 else max(1, threads // 2 - 1)
    legs = {}
    for name, workers in (("one_thread", 0), ("workers", auto)):
        done, elapsed, rows = 0, 0.0, 0
        while elapsed < seconds / 2 and done < 200:
            out, t = replay(captured, lib=lib, workers=workers)
            solved = int(out["solved"][0])
            rows = int(np.sum(out["points"][:solved]))
            elapsed += t
            done += 1
        legs[name] = {"value": rows * 2 * SUBSTEPS * done / elapsed, "workers": workers, "threads": workers + 1,
                      "soft_steps": done, "seconds": elapsed, "rows": rows}
    best = legs["workers"]
    return {"value": best["value"], "unit": "contact-iterations/s", "cores": best["threads"], "kind": "port",
            "sample": f"the mixed pile's contact phase at rest ({best['rows']} rows, the step before the timed "
                      f"ones, tapped): {best['soft_steps']} oracle soft steps with {best['workers']} workers "
                      f"(the reference's 'auto' for {threads} threads) in {best['seconds']:.2f} s; "
                      f"one thread {legs['one_thread']['value']:.4g} contact-iterations/s",
            "legs": legs, "nproc": os.cpu_count(), "threads_available": threads, "cpu_model": model}


def cpu_baseline_worlds():
    """The CPU baseline for the whole-world runs: the same trajectories stepped on one host
    thread with the oracle standing in for every device call (tests/hostworld), ms a step."""
    import ctypes

    host_lib = os.path.join(ROOT, "tests", "hostworld", "_build", "libhostworld.so")
    if not os.path.exists(host_lib):
        return {}
    lib = ctypes.CDLL(host_lib)
    return {name: world_run(name, lib)["ms_per_step"] for name in WORLD_RUNS if name not in GPU_ONLY_RUNS}


WORLD_RUNS = ["pyramid", "rain", "darts", "terrain"]
# Trajectories too long for the host-only port's leg (the 20,001-body pile: ~10 minutes).
GPU_ONLY_RUNS = {"mixed_pile"}
# Whole worlds stepped together: the reference's own resolved bodies of these trajectories
# (13 to 401 bodies: towers, sphere piles, a pyramid, rods and balls, capsules, hulls, a
# heightfield), repeated to the batch size.
BATCH_WORLDS_FROM = ["tower", "pile_small", "pyramid", "rain_small", "capsules", "darts", "rocks", "terrain"]


PILE = "mixed_pile"
# The soft step's kernel for each launch shape with velocities in LDS (bm_set_launch).
SOFT_KERNELS = {0: "soft_step_kernel<true, 256, 1>", 1: "soft_step_kernel<true, 256, 2>", 2: "soft_step_kernel<true, 512, 2>"}


def pile_run(timed, warmup, broadphase="auto"):
    """The headline: the mixed pile's reference trajectory (600 steps) through bm_world_step;
    the first 600 - timed - warmup steps bring it to rest, then `warmup` untimed steps, then
    `timed` steps timed one by one (wall clock around bm_world_step; bm_world_read's rows,
    soft-step bytes and kernel time after each). Every step's transforms hash is checked
    against the reference's after the run."""
    import ctypes
    import hashlib

    import fixtures
    from bearing_amd import native
    from worldcheck import fill_world

    fx = fixtures.load(os.path.join(fixtures.DATA, f"{PILE}.traj.bmfx.gz"))
    meta = fx.meta
    total = int(meta["steps"])
    timed = max(1, min(timed, total - 1))
    warmup = max(0, min(warmup, total - timed))
    setup = total - timed - warmup
    w = fill_world(native.World(0), fx)
    w.set_broadphase(broadphase)
    tr = np.zeros(7 * int(meta["bodies"]), np.float32)
    state = native.WorldState()
    state.transforms = tr.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
    hashes, rows, soft_bytes, soft_ms, walls = [], [], [], [], []
    before = None
    t_all = time.perf_counter()
    for s in range(1, total + 1):
        if s == setup + warmup + 1:
            before = w.times()
        if s == setup + warmup and s >= 1:
            w.capture_next()   # this untimed step's contact phase, for the CPU baseline
        t = time.perf_counter()
        w.step(meta["dt"])
        wall = time.perf_counter() - t
        w.lib.bm_world_read(w.ptr, ctypes.byref(state))
        hashes.append(hashlib.sha256(tr.tobytes()).hexdigest())
        rows.append(state.rows)
        soft_bytes.append(state.soft_bytes)
        soft_ms.append(state.soft_ms)
        walls.append(wall)
        if s % 100 == 0:
            print(f"bench: {PILE} step {s}", file=sys.stderr, flush=True)
    all_wall = time.perf_counter() - t_all
    after = w.times()
    captured = getattr(w, "captured", None)
    w.close()
    first = setup + warmup
    tw = float(np.sum(walls[first:]))
    trows = int(np.sum(rows[first:]))
    return {
        "steps": total, "bodies": int(meta["bodies"]), "broadphase": broadphase, "setup_steps": setup, "warmup": warmup, "timed_steps": timed,
        "bit_exact": bool(all(h == r for h, r in zip(hashes, meta["hashes"]))),
        "timed_wall_s": tw, "ms_per_step": tw * 1000 / timed,
        "rows_per_step": trows / timed, "contact_iters_per_s": trows * 2 * SUBSTEPS / tw,
        "soft_kernel_ms": float(np.mean(soft_ms[first:])),
        "soft_bytes_per_launch": float(np.mean(soft_bytes[first:])),
        "soft_only_contact_iters_per_s": trows * 2 * SUBSTEPS / (float(np.sum(soft_ms[first:])) / 1000),
        "ms_per_step_by_part": {k: (after[k] - before[k]) / timed for k in after},
        "trajectory": {"ms_per_step": all_wall * 1000 / total, "world_steps_per_s": total / all_wall,
                       "contact_iters_per_s": int(np.sum(rows)) * 2 * SUBSTEPS / all_wall,
                       "note": "all 600 steps, the landing included, wall clock around bm_world_step and bm_world_read"},
    }, captured


def c4_run(threads, loop_sample=16):
    """The batched form as configured: the 4,096 worlds of scenes.batched (seed 4: body counts
    log-uniform in [8, 512], box towers and sphere piles, each on its own ground), every world
    stepped 300 steps together by bm_batch_step; against bm_world_step looped over every 256th
    world for the same 300 steps; every world's final transforms against the reference's."""
    import hashlib

    import batchedcheck
    from bearing_amd import native, scenes

    fx = batchedcheck.load()
    specs = batchedcheck.scene_list(fx)
    steps, dt = int(fx.meta["steps"]), fx.meta["dt"]
    batch = native.WorldBatch(0, threads)
    worlds = []
    for bodies in specs:
        w = batch.world()
        w.add(scenes.resolve(bodies))
        worlds.append(w)
    print(f"bench: {len(worlds)} batched worlds built", file=sys.stderr, flush=True)
    before = batch.times()
    soft_ms, soft_bytes, rows, walls = [], [], [], []
    # The autotuner on the batch's own first steps (inside the timed run): each soft-step launch
    # shape for two steps, its second step's contact-phase time kept, then the fastest for the rest.
    shapes, tried = [0, 1, 2], {}
    best = 0
    for s in range(steps):
        if s < 2 * len(shapes):
            batch.set_launch(shapes[s // 2])
        elif s == 2 * len(shapes):
            best = min(tried, key=tried.get)
            batch.set_launch(best)
        # Every tenth step launch by launch, so the shared soft launch's own time can be read;
        # the others with the contact phase as one HIP graph.
        batch.set_graph(s % 10 != 9)
        t = time.perf_counter()
        batch.step(dt)
        walls.append(time.perf_counter() - t)
        if s < 2 * len(shapes) and s % 2 == 1:
            tried[shapes[s // 2]] = worlds[0].state().soft_ms
        if s % 10 == 9:   # the shared soft launch's time and every world's bytes, every tenth step
            st = [w.state() for w in worlds]
            soft_ms.append(st[0].soft_ms)
            soft_bytes.append(sum(x.soft_bytes for x in st))
            rows.append(sum(x.rows for x in st))
        if (s + 1) % 100 == 0:
            print(f"bench: batched step {s + 1}", file=sys.stderr, flush=True)
    wall = float(np.sum(walls))
    after = batch.times()
    graphs = batch.graph_counts()
    exact = all(hashlib.sha256(w.read()["transforms"].tobytes()).digest() == batchedcheck.final_of(fx, i)
                for i, w in enumerate(worlds))
    batch.close()
    picks = list(range(0, len(specs), max(1, len(specs) // loop_sample)))[:loop_sample]
    t_loop, loop_exact = 0.0, True
    for i in picks:
        w = native.World(0)
        w.add(scenes.resolve(specs[i]))
        t = time.perf_counter()
        for _ in range(steps):
            w.step(dt)
        t_loop += time.perf_counter() - t
        loop_exact = loop_exact and hashlib.sha256(w.read()["transforms"].tobytes()).digest() == batchedcheck.final_of(fx, i)
        w.close()
    sample_bodies = sum(len(specs[i]) for i in picks)
    rate = len(specs) * steps / wall
    loop_rate = len(picks) * steps / t_loop
    achieved = float(np.mean(soft_bytes)) / (float(np.mean(soft_ms)) / 1000) / 1e9
    return {"workload": "scenes.batched(4096, seed 4): log-uniform 8..512 bodies, box towers and sphere piles, "
                        f"{steps} steps at dt 1/60", "worlds": len(specs), "bodies": int(fx.meta["bodies"]),
            "steps": steps, "host_threads": threads, "ms_per_step": wall * 1000 / steps, "world_steps_per_s": rate,
            "looped_plain_world_steps_per_s": loop_rate, "speedup_vs_looped": rate / loop_rate,
            "looped_sample": f"every {len(specs) // loop_sample}th world ({len(picks)} worlds, {sample_bodies} bodies) "
                             f"stepped alone through bm_world_step for {steps} steps",
            "bit_exact": bool(exact), "looped_bit_exact": bool(loop_exact), "contact_phase_graphs": graphs,
            "launch": {"shape": best, "contact_phase_ms_by_shape": tried,
                       "tuned": "each shape's second of two steps at the run's start (inside the timed steps)"},
            "ms_per_step_by_part": {k: (after[k] - before[k]) / steps for k in after},
            "soft_roofline": {"kernel": f"{SOFT_KERNELS[best]} (all worlds, one launch)", "bound": "hbm",
                              "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                              "frac": achieved / HBM_PEAK_GBS, "kernel_ms": float(np.mean(soft_ms)),
                              "algorithmic_bytes_per_launch": float(np.mean(soft_bytes)),
                              "rows_per_step": float(np.mean(rows)),
                              "traffic": traffic_of(SOFT_KERNELS[best], "scenes.batched"),
                              "sampled": "every tenth step"}}


def batched_world_run(count, warm, steps, threads):
    """`count` whole worlds stepped together (bm_batch_step) against the same worlds stepped one
    at a time through bm_world_step (a sample of them): world-steps per second both ways, and
    whether every batched world's transforms after warm + steps steps are its reference's bits."""
    import hashlib

    import fixtures
    from bearing_amd import native
    from worldcheck import fill_world

    fxs = [fixtures.load(os.path.join(fixtures.DATA, f"{n}.traj.bmfx.gz")) for n in BATCH_WORLDS_FROM]
    dt = fxs[0].meta["dt"]
    batch = native.WorldBatch(0, threads)
    worlds = [fill_world(batch.world(), fxs[i % len(fxs)]) for i in range(count)]
    bodies = sum(int(fxs[i % len(fxs)].meta["bodies"]) for i in range(count))
    for _ in range(warm):
        batch.step(dt)
    before = batch.times()
    t0 = time.perf_counter()
    for _ in range(steps):
        batch.step(dt)
    wall = time.perf_counter() - t0
    after = batch.times()
    parts = {k: (after[k] - before[k]) / steps for k in after}
    at = warm + steps
    exact = all(hashlib.sha256(w.read()["transforms"].tobytes()).hexdigest() == fxs[i % len(fxs)].meta["hashes"][at - 1]
                for i, w in enumerate(worlds))
    contacts = sum(w.read()["contacts"] for w in worlds)
    batch.close()
    sample = min(64, count)
    singles = [fill_world(native.World(0), fxs[i % len(fxs)]) for i in range(sample)]
    t_loop = 0.0
    for w in singles:
        for _ in range(warm):
            w.step(dt)
        t = time.perf_counter()
        for _ in range(steps):
            w.step(dt)
        t_loop += time.perf_counter() - t
        w.close()
    loop_rate = sample * steps / t_loop
    rate = count * steps / wall
    return {"worlds": count, "bodies": bodies, "contacts_at_end": int(contacts), "warm_steps": warm, "steps": steps,
            "host_threads": threads, "ms_per_step": wall * 1000 / steps, "world_steps_per_s": rate,
            "looped_plain_world_steps_per_s": loop_rate, "speedup_vs_looped": rate / loop_rate,
            "bit_exact": bool(exact), "ms_per_step_by_part": parts, "worlds_from": BATCH_WORLDS_FROM}


def adjoint_run(worlds=512, steps=120, every=10, threads=16, repeats=3):
    """The differentiation form: `worlds` thrown-box scenes (scenes.throws, seed 5; dt 1/120)
    recorded while stepping together, then held and differentiated on the device: the held
    forward pass and the backward pass (checkpoints every `every` steps) timed by their kernels'
    HIP events, and their ratio. The loss is a fixed weighting of every final state."""
    from bearing_amd import native, scenes

    dt = 1 / 120
    specs = scenes.throws(worlds, seed=5)
    batch = native.WorldBatch(0, threads)
    ws = []
    for bodies in specs:
        w = batch.world(sleeping=False)
        w.add(scenes.resolve(bodies))
        w.record(True)
        ws.append(w)
    t0 = time.perf_counter()
    for _ in range(steps):
        batch.step(dt)
    record_s = time.perf_counter() - t0
    # A world whose recorded steps cannot be held (a body caught in flight) is left out.
    keep, dropped = list(ws), 0
    while True:
        try:
            adj = native.Adjoint(keep, every=every, device=0)
            break
        except native.BearingError as e:
            m = re.search(r"world (\d+):", str(e))
            if not m or len(keep) <= 1:
                raise
            keep.pop(int(m.group(1)))
            dropped += 1
    rec = adj.recorded()
    out = adj.forward(rec["state"], rec["mass"], rec["friction"], rec["restitution"])
    real = np.concatenate([w.read()["position"].reshape(-1, 3) for w in keep])
    held_dev = float(np.abs(out[:, :3] - real).max())
    W = np.random.default_rng(5).normal(size=out.shape)
    adj.backward(W)
    fwd, bwd = [], []
    for _ in range(repeats):
        adj.forward(rec["state"], rec["mass"], rec["friction"], rec["restitution"])
        g = adj.backward(W)
        f, b = adj.times()
        fwd.append(f)
        bwd.append(b)
    finite = bool(all(np.isfinite(v).all() for v in g.values()))
    adj.close()
    batch.close()
    f, b = float(np.median(fwd)), float(np.median(bwd))
    return {"worlds": len(keep), "dropped": dropped, "bodies": int(adj.bodies), "steps": steps, "every": every,
            "dt": dt, "record_s": record_s, "forward_ms": f, "backward_ms": b, "bwd_fwd": b / f,
            "world_steps_per_s_forward": len(keep) * steps / (f / 1000),
            "held_vs_recorded_max_position_m": held_dev, "gradients_finite": finite}


def world_run(name, lib=None, steps=None):
    """A whole reference trajectory stepped through bm_world (broadphase, narrow driver,
    device meet / colour / soft step, catch, sleep): wall time per step and whether every
    step's transforms are the reference's bits (checked after the timed run)."""
    import hashlib

    import fixtures
    from bearing_amd import native
    from worldcheck import fill_world

    fx = fixtures.load(os.path.join(fixtures.DATA, f"{name}.traj.bmfx.gz"))
    meta = fx.meta
    w = fill_world(native.World(0, lib=lib), fx)
    before = w.times() if lib is None else None
    out = {"transforms": np.zeros(7 * int(meta["bodies"]), np.float32)}
    from bearing_amd.native import WorldState
    import ctypes
    kept = []
    state = WorldState()
    state.transforms = out["transforms"].ctypes.data_as(ctypes.POINTER(ctypes.c_float))
    n = meta["steps"] if steps is None else min(steps, meta["steps"])
    t0 = time.perf_counter()
    for _ in range(n):
        w.step(meta["dt"])
        w.lib.bm_world_read(w.ptr, ctypes.byref(state))
        kept.append(out["transforms"].tobytes())
    wall = time.perf_counter() - t0
    parts = None
    if before is not None:
        after = w.times()
        parts = {k: (after[k] - before[k]) / n for k in after}
    w.close()
    exact = all(hashlib.sha256(t).hexdigest() == h for t, h in zip(kept, meta["hashes"]))
    awake = int(np.count_nonzero(fx["awake"][:n]))
    return {"steps": n, "bodies": meta["bodies"], "awake_steps": awake, "ms_per_step": wall * 1000 / n,
            "ms_per_awake_step": wall * 1000 / max(awake, 1), "world_steps_per_s": n / wall,
            "bit_exact": bool(exact), "ms_per_step_by_part": parts}


def batched_soft_run(ctx, args):
    """Worlds of five scenes' soft steps, repeated to --batch-worlds, in one launch, against the
    same kernel looped one world a launch on a sample of them."""
    from bearing_amd import native
    from bearing_amd.tune import tune

    # Batched: worlds of five scenes, repeated to --batch-worlds, in one launch.
    loaded = [load_world(n) for n in BATCH_SOURCES]
    world_from_columns = loaded[0][2]
    sources = [(fx, cols) for fx, cols, _ in loaded]
    worlds = [world_from_columns(sources[i % len(sources)][0], {k: v.copy() for k, v in sources[i % len(sources)][1].items()})
              for i in range(args.batch_worlds)]
    if args.batched_launch:
        bshape, bresident = (int(x) for x in args.batched_launch.split(","))
    else:
        bshape, bresident, _ = tune(worlds)
    ctx.set_launch(bshape, bresident)
    big = native.Batch(worlds)
    big.upload(ctx)
    big.step(ctx)
    t0 = time.perf_counter()
    for _ in range(args.batch_steps):
        big.step(ctx)
    batched_wall = time.perf_counter() - t0
    batched_rate = args.batch_worlds * args.batch_steps / batched_wall
    # The same kernel looped once per world (one world a launch), on a sample of the same worlds.
    sample = min(64, args.batch_worlds)
    singles = [native.Batch([w]) for w in worlds[:sample]]
    t_loop = 0.0
    for s in singles:
        s.upload(ctx)
        s.step(ctx)
        t = time.perf_counter()
        s.step(ctx)
        t_loop += time.perf_counter() - t
    loop_rate = sample / t_loop
    return {"worlds": args.batch_worlds, "steps": args.batch_steps, "world_steps_per_s": batched_rate,
            "looped_plain_world_steps_per_s": loop_rate, "speedup_vs_looped": batched_rate / loop_rate,
            "worlds_from": BATCH_SOURCES, "launch": {"shape": bshape, "resident_lds": bool(bresident)}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50, help="timed world steps of the pile at rest (the last of its 600)")
    ap.add_argument("--warmup", type=int, default=5, help="untimed world steps before them")
    ap.add_argument("--micro-steps", type=int, default=50, help="soft steps of the rain's fixed state (microbenchmark)")
    ap.add_argument("--broadphase", default="auto", choices=["auto", "host", "device"],
                    help="where the pile's broadphase passes run (bm_world_set_broadphase)")
    ap.add_argument("--batch-worlds", type=int, default=4096)
    ap.add_argument("--batch-steps", type=int, default=10)
    ap.add_argument("--batch-threads", type=int, default=16, help="host threads for the whole-world batch")
    ap.add_argument("--whole-batch", type=int, default=None,
                    help="only step this many stand-in whole worlds together (bm_batch_step) and print the timing")
    ap.add_argument("--c4", action="store_true", help="only the batched form's 4,096 worlds and print the timing")
    ap.add_argument("--cpu-seconds", type=float, default=20.0, help="CPU baseline time, both legs together")
    ap.add_argument("--adjoint-worlds", type=int, default=512, help="worlds of the differentiation form")
    ap.add_argument("--adjoint", type=int, default=None, help="only the differentiation form, with this many worlds")
    ap.add_argument("--plain-launch", default=None,
                    help="shape,resident for the microbenchmark's world (default: autotuned, bearing_amd/tune.py)")
    ap.add_argument("--batched-launch", default=None, help="shape,resident for the batch (default: autotuned)")
    ap.add_argument("--world", default=None,
                    help="only step this reference trajectory whole through bm_world and print its timing")
    ap.add_argument("--world-steps", type=int, default=None, help="with --world: only its first steps")
    ap.add_argument("--parts", default="pile,micro,batched,c4,worlds,adjoint,cpu",
                    help="which legs of the full line to run (the rest are null): pile,micro,batched,c4,worlds,adjoint,cpu")
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="only copy this many MiB device to device 5 times (torch), to calibrate HBM byte counters")
    args = ap.parse_args()
    if args.copy_probe:
        import torch
        x = torch.ones(args.copy_probe * 1024 * 1024 // 8, dtype=torch.float64, device="cuda")
        y = torch.empty_like(x)
        for _ in range(5):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_mib": args.copy_probe, "bytes_read": x.numel() * 8, "bytes_written": y.numel() * 8}))
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(worlds would shard across GPUs with no exchange between them)", file=sys.stderr)
        sys.exit(2)

    if args.world:
        print(json.dumps({"world": args.world, **world_run(args.world, steps=args.world_steps)}))
        return
    if args.adjoint:
        print(json.dumps({"adjoint": adjoint_run(args.adjoint)}))
        return
    if args.whole_batch:
        print(json.dumps({"whole_batch": batched_world_run(args.whole_batch, 60, 20, args.batch_threads)}))
        return
    if args.c4:
        print(json.dumps({"c4": c4_run(args.batch_threads)}))
        return

    parts = set(args.parts.split(","))
    pile, captured = pile_run(args.steps, args.warmup, args.broadphase) if "pile" in parts else (None, None)
    micro = micro_run(args) if "micro" in parts else None
    c4 = c4_run(args.batch_threads) if "c4" in parts else None
    worlds_run = {name: world_run(name) for name in WORLD_RUNS} if "worlds" in parts else None
    adjoint = adjoint_run(args.adjoint_worlds) if "adjoint" in parts else None
    cpu = None
    if "cpu" in parts:
        cpu = cpu_baseline(captured, args.cpu_seconds)
        if worlds_run:
            for name, ms in cpu_baseline_worlds().items():
                worlds_run[name]["cpu_port_ms_per_step"] = ms
    if pile is not None:
        achieved = pile["soft_bytes_per_launch"] / (pile["soft_kernel_ms"] * 1e-3) / 1e9
        roofline = {"bound": "hbm", "kernel": "soft_wide_kernel<256> (the grid-wide soft step)", "achieved": achieved,
                    "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                    "traffic": traffic_of("soft_wide_kernel<256>", PILE),
                    "algorithmic_bytes_per_launch": pile["soft_bytes_per_launch"], "kernel_ms": pile["soft_kernel_ms"]}
        value, ms_per_step, steps, warmup = pile["contact_iters_per_s"], pile["ms_per_step"], pile["timed_steps"], pile["warmup"]
    else:
        roofline, value, ms_per_step, steps, warmup = None, None, None, args.steps, args.warmup
    line = {
        "metric": "contact-iters/s", "value": value, "unit": "contact-iterations/s", "n_gpus": 1,
        "steps": steps, "warmup": warmup, "ms_per_step": ms_per_step, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic: the seeded 20,001-body mixed pile (seed 3) stepped from its start along the reference's "
                "own trajectory; every step's transforms checked against the reference's",
        "config": {"workload": "mixed pile (20,001 bodies on a 256x256 heightfield), plain drop-in world.step "
                               "(bm_world_step), whole step at rest",
                   "timed": f"steps {pile['setup_steps'] + pile['warmup'] + 1}..{pile['steps']} of {pile['steps']}"
                   if pile else None,
                   "rows_per_step": pile["rows_per_step"] if pile else None},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "pile": pile,
        "batched_whole_worlds": c4,
        "world_steps_per_s_batched": c4["world_steps_per_s"] if c4 else None,
        "batched_speedup_vs_looped": c4["speedup_vs_looped"] if c4 else None,
        "world": worlds_run,
        "soft_step_microbench": micro,
        "adjoint": adjoint,
        "bwd_fwd": adjoint["bwd_fwd"] if adjoint else None,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
