"""Autotuner for the soft-step launch: which workgroup shape, and whether each world's
body velocities stay in LDS, run a given batch fastest.

Every setting computes the same bits (the tests check each one), so the choice is
only about time. The tuner runs each setting on a copy of the batch uploaded to its
own context, so the caller's state is not advanced, and keeps the fastest median
kernel time (HIP events on the launch stream).
"""
import itertools
import statistics

from . import native

SHAPES = (0, 1, 2)
RESIDENT = (True, False)


def tune(worlds, device=0, repeats=5):
    """(shape, resident, {setting: median ms}) for the batch of `worlds` (Batch world dicts)."""
    copies = [{k: (v.copy() if hasattr(v, "copy") else v) for k, v in w.items()} for w in worlds]
    batch = native.Batch(copies)
    ctx = native.Context(device)
    timings = {}
    try:
        batch.upload(ctx)
        for shape, resident in itertools.product(SHAPES, RESIDENT):
            ctx.set_launch(shape, resident)
            batch.step(ctx)
            timings[(shape, resident)] = statistics.median(batch.step(ctx) for _ in range(repeats))
    finally:
        ctx.close()
    best = min(timings, key=timings.get)
    return best[0], best[1], timings
