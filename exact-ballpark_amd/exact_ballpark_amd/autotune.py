"""The launch-shape autotuner: for a park on the GPU, times each candidate (workgroup size, which decides how many
worlds share a tile; the lane schedule; ticks per launch) for a few ticks on the park's own state, which is saved
before and restored after each trial, and returns the fastest with the timings. Every candidate computes the same bits:
a world's result depends on no other world, on no lane assignment and on no tile shape (the tests hold each shape to
the plain one), so the choice is a matter of speed only.
"""

from __future__ import annotations

import itertools

import torch

from .vec_modes import ModesGraph, VecModesPark


def _restore(park: VecModesPark, saved: tuple) -> None:
    for b, old in zip((park._f, park._q, park._b), saved):
        b.copy_(old)


def time_shape(park: VecModesPark, block: int, schedule: int, ticks: int, reps: int = 3, warm: int = 1) -> float:
    """Milliseconds per tick of one shape on the park's current state (the state is given back afterwards)."""
    saved = park._save()
    t0, keys0 = park.current_time, park.keys
    old = (park.block, park.schedule)
    park.block = block
    park._set_tiles()
    park.schedule = schedule
    try:
        if park.tiles is None:
            raise ValueError(f"a world is wider than {block} rows")
        g = ModesGraph(park, ticks=ticks)
        for _ in range(warm):
            g.tick()
        s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize(park.dev)
        s.record()
        for _ in range(reps):
            g.tick()
        e.record()
        torch.cuda.synchronize(park.dev)
        ms = s.elapsed_time(e) / (reps * ticks)
    finally:
        _restore(park, saved)
        park.current_time, park.keys = t0, keys0
        park.block, park.schedule = old
        park._set_tiles()
    return ms


def tune(park: VecModesPark, blocks=(64, 128, 256), schedules=(0, 1), ticks=(1, 4, 16)) -> dict:
    """The fastest (block, schedule, ticks) for this park, applied to it, with every trial's ms per tick."""
    trials = {}
    for b, sc, t in itertools.product(blocks, schedules, ticks):
        try:
            trials[(b, sc, t)] = time_shape(park, b, sc, t)
        except ValueError:
            continue
    best = min(trials, key=trials.get)
    park.block, park.schedule = best[0], best[1]
    park._set_tiles()
    return {"block": best[0], "schedule": best[1], "ticks": best[2], "ms_per_tick": trials[best],
            "trials": {f"b{b}_s{s}_t{t}": ms for (b, s, t), ms in trials.items()}}
