"""Benchmark: the soft step (the contact solve's hot loop) on one MI355X, FP64.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--batch-worlds B] [--cpu-seconds S]

Workload (plain drop-in): the 2,000-sphere rain (2,000 spheres r U[0.1, 0.3], restitution
U[0, 0.7], friction 0.5, in a walled 20x20 m floor; seed 2) as the reference's own step 300
leaves it before the solve: its manifolds coloured and laid out, its bodies' state
(tests/golden/data/rain.soft.300). One step is one soft step of that world on the device:
inertias, prepare, 8 substeps of integrate / warm start / solve / move / relax, restitution,
store, one workgroup per world. Steps follow each other on the device-resident state.

Printed: one JSON line. `value` is contact-iterations per second (rows x 2 passes x 8
substeps per step, rows = manifold points solved). `roofline` prices the kernel against HBM
with the algorithmic bytes per manifold of the solve / relax / warm-start passes (see
`algorithmic_bytes`); `cpu_baseline` is the oracle's same soft step (a single-threaded C++
restatement) on the host on a bounded sample of the same workload; `batched` is world-steps
per second for a batch of worlds in one launch against stepping the same worlds one launch
each.
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


def cpu_baseline(fx, seconds):
    """The oracle's soft step on the host, one thread, repeated on the same input for about `seconds`."""
    from oracle import binding

    base = binding.columns(fx)
    rows = int(np.sum(fx["in.c.points"][: fx.meta["solved"]]))
    done, elapsed = 0, 0.0
    while elapsed < seconds and done < 1000:
        cols = {k: v.copy() for k, v in base.items()}  # a fresh input each time, not timed
        t = time.perf_counter()
        binding.soft_step(fx, cols)
        elapsed += time.perf_counter() - t
        done += 1
    return {"value": rows * 2 * SUBSTEPS * done / elapsed, "unit": "contact-iterations/s", "cores": 1, "kind": "port",
            "sample": f"{done} soft steps of {WORKLOAD} ({rows} rows) in {elapsed:.2f} s on one host thread"}


def cpu_baseline_worlds():
    """The CPU baseline for the whole-world runs: the same trajectories stepped on one host
    thread with the oracle standing in for every device call (tests/hostworld), ms a step."""
    import ctypes

    host_lib = os.path.join(ROOT, "tests", "hostworld", "_build", "libhostworld.so")
    if not os.path.exists(host_lib):
        return {}
    lib = ctypes.CDLL(host_lib)
    return {name: world_run(name, lib)["ms_per_step"] for name in WORLD_RUNS if name not in GPU_ONLY_RUNS}


WORLD_RUNS = ["pyramid", "rain", "darts", "terrain", "mixed_pile"]
# Trajectories too long for the host-only port's leg (the 20,001-body pile: ~10 minutes).
GPU_ONLY_RUNS = {"mixed_pile"}
# Whole worlds stepped together: the reference's own resolved bodies of these trajectories
# (13 to 401 bodies: towers, sphere piles, a pyramid, rods and balls, capsules, hulls, a
# heightfield), repeated to the batch size.
BATCH_WORLDS_FROM = ["tower", "pile_small", "pyramid", "rain_small", "capsules", "darts", "rocks", "terrain"]


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
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--batch-worlds", type=int, default=4096)
    ap.add_argument("--batch-steps", type=int, default=10)
    ap.add_argument("--batch-threads", type=int, default=16, help="host threads for the whole-world batch")
    ap.add_argument("--whole-batch", type=int, default=None,
                    help="only step this many whole worlds together (bm_batch_step) and print the timing")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--adjoint-worlds", type=int, default=512, help="worlds of the differentiation form")
    ap.add_argument("--adjoint", type=int, default=None, help="only the differentiation form, with this many worlds")
    ap.add_argument("--plain-launch", default=None,
                    help="shape,resident for the plain world (default: autotuned, bearing_amd/tune.py)")
    ap.add_argument("--batched-launch", default=None, help="shape,resident for the batch (default: autotuned)")
    ap.add_argument("--world", default=None,
                    help="only step this reference trajectory whole through bm_world and print its timing")
    ap.add_argument("--world-steps", type=int, default=None, help="with --world: only its first steps")
    ap.add_argument("--parts", default="plain,batched,whole,worlds,adjoint,cpu",
                    help="which legs of the full line to run (the rest are null): plain,batched,whole,worlds,adjoint,cpu")
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
    for _ in range(args.warmup):
        batch.step(ctx)
    kernel_ms = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        kernel_ms.append(batch.step(ctx))
    wall = time.perf_counter() - t0
    ms_per_step = wall * 1000 / args.steps
    value = rows * 2 * SUBSTEPS * args.steps / wall
    avg_kernel_ms = float(np.mean(kernel_ms))
    achieved = step_bytes / (avg_kernel_ms * 1e-3) / 1e9
    # The same world stepped by the whole grid (bm_set_wide) for comparison: same bits.
    ctx.set_wide(1)
    for _ in range(2):
        batch.step(ctx)
    wide_ms = float(np.mean([batch.step(ctx) for _ in range(10)]))
    ctx.set_wide(16384)

    parts = set(args.parts.split(","))
    batched = None
    if "batched" in parts:
        batched = batched_soft_run(ctx, args)
    ctx.close()
    whole = batched_world_run(args.batch_worlds, 60, 20, args.batch_threads) if "whole" in parts else None
    worlds_run = {name: world_run(name) for name in WORLD_RUNS} if "worlds" in parts else None
    adjoint = adjoint_run(args.adjoint_worlds) if "adjoint" in parts else None
    cpu = None
    if "cpu" in parts:
        cpu = cpu_baseline(fx, args.cpu_seconds)
        if worlds_run:
            for name, ms in cpu_baseline_worlds().items():
                worlds_run[name]["cpu_port_ms_per_step"] = ms
    # HBM bytes per launch of the plain kernel, from the rocprofv3 FETCH_SIZE / WRITE_SIZE runs
    # recorded in profiles/ (scaled by a copy kernel of known size); null if not measured
    # for this launch shape.
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "soft_step_rain2000_traffic.json")
    if os.path.exists(tpath):
        rec = json.load(open(tpath))
        if rec.get("kernel") == f"soft_step_kernel<{'true' if resident else 'false'}, {512 if shape == 2 else 256}, {1 if shape == 0 else 2}>":
            traffic = rec["bytes"]
    line = {
        "metric": "contact-iters/s", "value": value, "unit": "contact-iterations/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic: the reference's own state of the seeded 2,000-sphere rain before its step-300 solve",
        "config": {"workload": "rain-2000 soft step (inertias..store, 8 substeps), plain drop-in, one world",
                   "bodies": meta["top"], "manifolds": solved, "rows": rows, "colours": int(np.count_nonzero(np.diff(fx["in.starts"][:33]))),
                   "kernel_ms": avg_kernel_ms, "launch": {"shape": shape, "resident_lds": bool(resident)},
                   "grid_wide_kernel_ms": wide_ms},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "algorithmic_bytes_per_launch": step_bytes},
        "cpu_baseline": cpu,
        "batched": batched,
        "batched_whole_worlds": whole,
        "world": worlds_run,
        "adjoint": adjoint,
        "bwd_fwd": adjoint["bwd_fwd"] if adjoint else None,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
