"""Tile-shape and grid autotuner for the fused push kernel.

Each candidate shape (block size x particles per thread) is run for a few
real steps of the workload and timed with HIP events; the fastest is kept on
the context.  For the 1-D push the grid is tuned next, blocks per CU from 1
up to what fits, timing the push kernel: whole steps would mix in the
resampling, whose cost drifts from trial to trial as the run evolves (a
resampling discharge merges more as it goes), and the grid only changes the
push and its short reduction.  Steps taken here advance the simulation like
warm-up steps.
"""
from __future__ import annotations

import ctypes as C
from typing import Callable

from . import _lib

# shape index -> (threads per block, particles per thread), as compiled in bz_push.hip
PUSH_SHAPES = {0: (256, 4), 1: (512, 2), 2: (256, 2), 3: (1024, 1), 4: (1024, 2)}


def pick_within(results: dict, order, tie: float = None):
    """Among the candidates timed within `tie` (relative) of the fastest,
    the first by `order` (then by time)."""
    tie = TIE if tie is None else tie
    fastest = min(results.values())
    return min((k for k in results if results[k] <= fastest * (1.0 + tie)),
               key=lambda k: (order(k), results[k]))


def grid_order(bpc: int) -> int:
    """Preference among grid settings timed within a tie: the default (0,
    as many blocks as fit: one per CU where the push's LDS allows one), then
    the fewest blocks per CU."""
    return -1 if bpc == 0 else bpc


def push_time_ms(lib, ctx) -> tuple[float, int]:
    ms = (C.c_double * 3)()
    nl = (C.c_int64 * 3)()
    _lib.check(lib.bz_timing_get(ctx, ms, nl))
    return ms[0], int(nl[0])


GRID_BPC = (0, 1, 2, 4)  # blocks per CU of the 1-D push (0: as many as fit)
TIE = 0.005  # relative difference below which two trial timings count as equal
TIE_SHORT = 0.02  # the same where each CU walks few tiles
SHORT_WALK_TILES = 16  # tiles of 1024 particles per CU below which a walk is short


def _time(lib, ctx, step: Callable[[], None], steps: int, whole: bool = False) -> float:
    """Average ms of the push kernel (library events), or of whole steps
    (events on the current stream around the steps)."""
    step()  # the first call of a setting may set kernel attributes
    if whole:
        import torch
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(steps):
            step()
        e1.record()
        e1.synchronize()
        return e0.elapsed_time(e1) / steps
    _lib.check(lib.bz_timing_enable(ctx, 1))
    for _ in range(steps):
        step()
    ms, n = push_time_ms(lib, ctx)
    _lib.check(lib.bz_timing_enable(ctx, 0))
    return ms / max(n, 1)


def tune_shapes(lib, ctx, step: Callable[[], None], steps: int = 10,
                grids: bool = False, tie: float = None, max_tile: int = 0) -> dict:
    """Time `steps` calls of `step` per push shape (tiles of at most
    `max_tile` particles if set); keep the fastest (ties within `tie` go to
    the smaller tile).  With `grids`, then time the best shape per
    blocks-per-CU setting."""
    results = {}
    used = 0
    _lib.check(lib.bz_set_push_grid(ctx, 0))
    for shape in PUSH_SHAPES:
        if max_tile and PUSH_SHAPES[shape][0] * PUSH_SHAPES[shape][1] > max_tile:
            continue
        _lib.check(lib.bz_set_push_shape(ctx, shape))
        results[shape] = _time(lib, ctx, step, steps)
        used += steps + 1
    # the fastest, or a shape with a smaller tile within TIE of it: the trial
    # steps run at the starting count, and a smaller tile keeps the waves
    # evenly loaded when the count drifts (a discharge losing electrons)
    best = pick_within(results, lambda k: PUSH_SHAPES[k][0] * PUSH_SHAPES[k][1], tie)
    _lib.check(lib.bz_set_push_shape(ctx, best))
    out = {"best": best, "shape": PUSH_SHAPES[best],
           "avg_ms": {str(PUSH_SHAPES[k]): v for k, v in results.items()}}
    if grids:
        gres = {}
        for bpc in GRID_BPC:
            _lib.check(lib.bz_set_push_grid(ctx, bpc))
            gres[bpc] = _time(lib, ctx, step, steps)
            used += steps + 1
        gbest = pick_within(gres, grid_order, TIE_SHORT)
        _lib.check(lib.bz_set_push_grid(ctx, gbest))
        out["blocks_per_cu"] = gbest
        out["grid_avg_ms"] = {str(k): v for k, v in gres.items()}
    out["steps_used"] = used
    return out


def tune_push(engine, steps: int = 10, start_index: int = 0) -> dict:
    """Autotune the 1-D / 0-D engine's push on its own workload."""
    index = [start_index]

    def step():
        engine.step(index[0])
        index[0] += 1

    # a short walk (a few 1024-particle tiles per CU) is where the tile size
    # decides how evenly the waves end, and where a count that drifts after
    # the trial moves the balance: there the 2048-particle tile is left out
    # and timings within TIE_SHORT tie
    import torch
    cus = torch.cuda.get_device_properties(engine.dev).multi_processor_count
    short = engine.count() < SHORT_WALK_TILES * 1024 * cus
    # short walks make short launches: time more of them against the noise
    out = tune_shapes(engine.lib, engine.ctx, step, max(steps, 40) if short else steps,
                      grids=engine.cfg.dim == 1, tie=TIE_SHORT if short else TIE,
                      max_tile=1024 if short else 0)
    engine.drop_graphs()  # blocks captured before hold the old launch shape
    return out


def tune_swarms(batch, steps: int = 10) -> dict:
    """Autotune the batched-swarm push (all instances in one launch)."""
    return tune_shapes(batch.lib, batch.ctx, lambda: batch.step(accumulate=False), steps)
