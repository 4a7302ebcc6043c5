"""The launch-shape autotuner: for a park on the GPU, times each candidate (workgroup size, which decides how many
worlds share a tile; the lane schedule; ticks per launch; the world order: the row order, or an order a scheduler
installs, such as the learned scheduler's regrouping) for a few ticks on the park's own state, which is saved before
and restored after each trial, and returns the fastest with the timings. Every candidate computes the same bits: a
world's result depends on no other world, on no lane assignment, on no tile shape and on no world order (the tests hold
each to the plain one), so the choice is a matter of speed only, and the row order is always a candidate: it is the
fallback when a scheduler's order does not pay.
"""

from __future__ import annotations

import itertools

import torch

from .vec_modes import ModesGraph, VecModesPark


def _restore(park: VecModesPark, saved: tuple) -> None:
    for b, old in zip(park._blocks(), saved):
        b.copy_(old)


def time_shape(park: VecModesPark, block: int, schedule: int, ticks: int, reps: int = 0, warm: int = 1,
               order=None) -> float:
    """Milliseconds per tick of one shape on the park's current state (the state is given back afterwards). `order`:
    None for the row order, else a callable that installs a world order on the park (after its tiles are set). reps 0:
    enough launches for 32 ticks."""
    reps = reps or max(3, 32 // ticks)
    saved = park._save()
    path0 = park.path.clone() if park.path is not None else None
    iters0 = park.iters.clone() if park.iters is not None else None
    t0, keys0 = park.current_time, park.keys
    old = (park.block, park.schedule)
    park.block = block
    park._set_tiles()
    park.schedule = schedule
    try:
        if park.tiles is None:
            raise ValueError(f"a world is wider than {block} rows")
        if order is not None:
            order()
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
        if path0 is not None:
            park.path.copy_(path0)
        if iters0 is not None:
            park.iters.copy_(iters0)
        park.current_time, park.keys = t0, keys0
        park.block, park.schedule = old
        park._set_tiles()
    return ms


def tune(park: VecModesPark, blocks=(64, 128, 256), schedules=(0, 1), ticks=(1, 4, 16), orders=None) -> dict:
    """The fastest (block, schedule, ticks, order) for this park, applied to it, with every trial's ms per tick.
    `orders`: {name: callable installing a world order}; the row order ("rows") is always tried."""
    orders = {"rows": None, **(orders or {})}
    trials = {}
    for b, sc, t, o in itertools.product(blocks, schedules, ticks, orders):
        try:
            trials[(b, sc, t, o)] = time_shape(park, b, sc, t, order=orders[o])
        except ValueError:
            continue
    best = min(trials, key=trials.get)
    park.block, park.schedule = best[0], best[1]
    park._set_tiles()
    if orders[best[3]] is not None:
        orders[best[3]]()
    return {"block": best[0], "schedule": best[1], "ticks": best[2], "order": best[3], "ms_per_tick": trials[best],
            "trials": {f"b{b}_s{s}_t{t}_{o}": ms for (b, s, t, o), ms in trials.items()}}
