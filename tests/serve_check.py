"""Checks of the serving loop's episodes against the oracle: an episode as an oracle world (its initial balls, the
policy's commands replayed on the host by exact_ballpark_amd.serve.episode_commands), run to its last tick, and its
final state compared with the one the server kept, bit for bit (two NaNs equal)."""

from __future__ import annotations

import time
from collections import defaultdict

import numpy as np

import oracle as O
import park_driver as D
from exact_ballpark_amd import serve as SV

SCALAR_D = [18, 21, 22, 23, 24, 25, 19, 20]       # mass, radius, max_vel, agility, speed_fraction, follow_range, tf, lc


def short_seeds(n: int, max_len: int, max_ships: int = 64, start: int = 1) -> np.ndarray:
    """The first n seeds from `start` whose episodes last at most max_len ticks with at most max_ships ships."""
    seeds = np.arange(start, start + 400 * n * max(1, 2000 // max_len), dtype=np.uint64)
    ships, length = SV.episode_shape(seeds)
    out = seeds[(length <= max_len) & (ships <= max_ships)][:n]
    assert len(out) == n
    return out


def oracle_final(seed: int, timing: list | None = None) -> tuple:
    """(doubles, ints) of the episode's final state in the oracle, balls by id. timing: the seconds of the oracle's
    ticks (commands applied and evolved; the episode's making not counted) are appended to it."""
    seed = int(seed)
    ships, length = (int(x[0]) for x in SV.episode_shape([seed]))
    r = SV.episode_rows(np.array([seed], dtype=np.uint64), [ships])
    o = O.World(1.0e6, 1000, False, 0)
    for k in range(ships):
        o.add({"id": int(r["id"][k]), "mass": float(r["mass"][k]), "radius": float(r["radius"][k]),
               "max_vel": float(r["max_vel"][k]), "agility": float(r["agility"][k]), "sf": 1.0,
               "pos": [float(v) for v in r["pos"][k]], "vel": [0.0, 0.0, 0.0], "free": True, "massive": True,
               "interactive": bool(r["interactive"][k]), "cloaked": 0, "bubble": 0, "moribund": False, "mode": 2})
    by_tick = defaultdict(list)
    for t, op, k, other, x, y, z, rng in SV.episode_commands(seed, ships, length):
        rr, w = (SV.WARP_MIN_RANGE, SV.WARP_FACTOR) if op == SV.WARP_OP else (rng, 0)
        by_tick[t].append((t, op, k + 1, other + 1, x, y, z, rr, w))
    t0 = time.perf_counter()
    for t in range(1, length + 1):
        for c in by_tick.get(t, ()):
            o.command(c)
        o.evolve()
    if timing is not None:
        timing.append(time.perf_counter() - t0)
    d, q, _ = o.state()
    return d[:, D.D_COLS], q[:, D.Q_COLS]


def kept_as_oracle(kept: tuple) -> tuple:
    """A kept final state (f64 [NF, ships], i64 [NI, ships] with follow pointers relative to the episode's first row,
    u8 [NB, ships]) in the oracle's columns."""
    f, q, b = kept
    n = f.shape[1]
    d = np.concatenate([f[0:3].T, f[3:6].T, f[12:15].T, f[15:18].T, f[6:9].T, f[9:12].T, f[SCALAR_D].T], axis=1)
    ids, fp = q[0], q[5]
    tgt = np.where(fp >= 0, ids[np.clip(fp, 0, n - 1)], -1)
    qq = np.stack([ids, q[1], q[2], q[3], q[4], tgt, b[2], b[4], b[3], b[0], b[1]], axis=1).astype(np.int64)
    return d, qq


def same(a: tuple, b: tuple) -> bool:
    return bool(D._same_bits(a[0], b[0]).all() and np.array_equal(a[1], b[1]))


def same_kept(a: tuple, b: tuple) -> bool:
    return all(np.array_equal(np.ascontiguousarray(x).view(np.uint8), np.ascontiguousarray(y).view(np.uint8))
               for x, y in zip(a, b))
