"""Loads tests/golden/*.npz: the reference's own scalar worlds, every ball after every tick (written by the scripts
in tests/golden/), in the oracle's column order; each record's list is the ball's followers, or its collision list
in the collision fixture."""

from __future__ import annotations

import json
from pathlib import Path

import numpy as np

GOLDEN = Path(__file__).resolve().parent / "golden"
# the fixture's doubles (px .. last_collision) as the oracle's D_NAMES columns, its ints as the oracle's Q_NAMES
D_FROM_ORACLE = [0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 18, 19, 20, 21, 22, 23, 24, 25]
Q_FROM_ORACLE = [0, 1, 2, 3, 4, 5, 6, 7]


def load(name: str = "modes_golden.npz") -> dict:
    z = np.load(GOLDEN / name, allow_pickle=False)
    worlds = json.loads(str(z["worlds"]))
    for w in worlds:
        w["cmds"] = [tuple(c) for c in w["cmds"]]
    rec = {}
    d = z["d"].view(np.float64)
    fo, fi = (z["hit_off"], z["hit_ids"]) if "hit_off" in z else (z["fol_off"], z["fol_ids"])   # the per-ball lists
    world, step, ball, q = z["world"], z["step"], z["ball"], z["q"]    # each read of z[...] decompresses it again
    for r in range(len(world)):
        key = (int(world[r]), int(step[r]))
        rec.setdefault(key, []).append((int(ball[r]), d[r], q[r], [int(x) for x in fi[fo[r]:fo[r + 1]]]))
    out = {}
    for key, balls in rec.items():
        balls.sort(key=lambda t: t[0])
        out[key] = (np.stack([b[1] for b in balls]), np.stack([b[2] for b in balls]), [b[3] for b in balls])
    return {"worlds": worlds, "kept": [int(k) for k in z["kept"]], "records": out, "seed": int(z["seed"])}
