"""Tile-shape autotuner for the fused push kernel.

Each candidate shape (block size x particles per thread) is run for a few
real steps of the workload and timed with the library's HIP events around
the push kernel; the fastest shape is kept on the context.  Steps taken
here advance the simulation like warm-up steps.
"""
from __future__ import annotations

import ctypes as C
from typing import Callable

from . import _lib

# shape index -> (threads per block, particles per thread), as compiled in bz_push.hip
PUSH_SHAPES = {0: (256, 4), 1: (512, 2), 2: (256, 2), 3: (1024, 1)}


def push_time_ms(lib, ctx) -> tuple[float, int]:
    ms = (C.c_double * 3)()
    nl = (C.c_int64 * 3)()
    _lib.check(lib.bz_timing_get(ctx, ms, nl))
    return ms[0], int(nl[0])


def tune_shapes(lib, ctx, step: Callable[[], None], steps: int = 10) -> dict:
    """Time `steps` calls of `step` per push shape; keep the fastest."""
    results = {}
    used = 0
    for shape in PUSH_SHAPES:
        _lib.check(lib.bz_set_push_shape(ctx, shape))
        step()  # first call of a shape may set kernel attributes
        used += 1
        _lib.check(lib.bz_timing_enable(ctx, 1))
        for _ in range(steps):
            step()
            used += 1
        ms, n = push_time_ms(lib, ctx)
        _lib.check(lib.bz_timing_enable(ctx, 0))
        results[shape] = ms / max(n, 1)
    best = min(results, key=results.get)
    _lib.check(lib.bz_set_push_shape(ctx, best))
    return {"best": best, "shape": PUSH_SHAPES[best],
            "avg_ms": {str(PUSH_SHAPES[k]): v for k, v in results.items()},
            "steps_used": used}


def tune_push(engine, steps: int = 10, start_index: int = 0) -> dict:
    """Autotune the 1-D / 0-D engine's push on its own workload."""
    index = [start_index]

    def step():
        engine.step(index[0])
        index[0] += 1

    return tune_shapes(engine.lib, engine.ctx, step, steps)


def tune_swarms(batch, steps: int = 10) -> dict:
    """Autotune the batched-swarm push (all instances in one launch)."""
    return tune_shapes(batch.lib, batch.ctx, lambda: batch.step(accumulate=False), steps)
