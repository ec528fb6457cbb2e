"""Tile-shape autotuner for the fused push kernel.

Each candidate shape (block size x particles per thread) is run for a few
real steps of the workload and timed with the library's HIP events around
the push kernel; the fastest shape is kept on the context.  Steps taken
here advance the simulation like warm-up steps.
"""
from __future__ import annotations

import ctypes as C

from . import _lib

# shape index -> (threads per block, particles per thread), as compiled in bz_push.hip
PUSH_SHAPES = {0: (256, 4), 1: (512, 2), 2: (256, 2), 3: (1024, 1)}


def push_time_ms(engine) -> tuple[float, int]:
    ms = (C.c_double * 3)()
    nl = (C.c_int64 * 3)()
    _lib.check(engine.lib.bz_timing_get(engine.ctx, ms, nl))
    return ms[0], int(nl[0])


def tune_push(engine, steps: int = 10, start_index: int = 0) -> dict:
    results = {}
    index = start_index
    for shape in PUSH_SHAPES:
        _lib.check(engine.lib.bz_set_push_shape(engine.ctx, shape))
        engine.step(index)  # first call of a shape may set kernel attributes
        index += 1
        _lib.check(engine.lib.bz_timing_enable(engine.ctx, 1))
        for _ in range(steps):
            engine.step(index)
            index += 1
        ms, n = push_time_ms(engine)
        _lib.check(engine.lib.bz_timing_enable(engine.ctx, 0))
        results[shape] = ms / max(n, 1)
    best = min(results, key=results.get)
    _lib.check(engine.lib.bz_set_push_shape(engine.ctx, best))
    return {"best": best, "shape": PUSH_SHAPES[best],
            "avg_ms": {str(PUSH_SHAPES[k]): v for k, v in results.items()},
            "steps_used": index - start_index}
