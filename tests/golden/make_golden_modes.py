#!/usr/bin/env python3
"""Writes tests/golden/modes_golden.npz: the reference's own scalar flight-modes world (src/exact_ballpark/modes.py,
`ModeBallpark`, over ref.py's `Ballpark`) run on the seeded worlds of tests/scenarios.py, every ball's state after
every tick, as bits.

Run only where the reference is mounted (/root/reference); the fixture it writes is committed and nothing else reads
the reference. The reference calls the C library's sincos, log and exp (modes.py:116-126, math.exp, math.log); here
those calls go to the reference's own transcriptions of glibc 2.34 aarch64 (dlibm.sincos, dlibm.log, vec_exp.exp),
which are the functions its batched paths and kernels are held to, so the vectors do not depend on this machine's C
library. Python 3.10 has no math.fma, which the transcriptions take at import: an exact rational a b + c rounded once
stands in for it. A world whose run raises in the reference (Python raises on a zero divisor where C gives an
infinity) is left out and listed in the fixture.
"""

from __future__ import annotations

import json
import math
import sys
import types
from fractions import Fraction
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent))
sys.path.insert(0, "/root/reference/src")

import scenarios as S  # noqa: E402


def _exact_fma(a: float, b: float, c: float) -> float:
    if not (math.isfinite(a) and math.isfinite(b) and math.isfinite(c)):
        return a * b + c
    r = Fraction(a) * Fraction(b) + Fraction(c)
    if r == 0:
        return (a * b) + c if (a * b == 0.0 and c == 0.0) else 0.0
    return float(r)


math.fma = _exact_fma                                       # before the transcriptions are imported
import torch  # noqa: E402
import exact_ballpark.dlibm as dlibm  # noqa: E402
import exact_ballpark.modes as modes  # noqa: E402
import exact_ballpark.ref as ref  # noqa: E402
import exact_ballpark.vec_exp as vec_exp  # noqa: E402


def _glibc_exp(x: float) -> float:
    return float(vec_exp.exp(torch.tensor([float(x)], dtype=torch.float64)).item())


_maths = types.SimpleNamespace(**{k: getattr(math, k) for k in dir(math) if not k.startswith("_")})
_maths.exp = _glibc_exp
_maths.log = dlibm.log
ref.math = _maths
modes.math = _maths
modes._sincos = dlibm.sincos

DOUBLES = ("px", "py", "pz", "vx", "vy", "vz", "gx", "gy", "gz", "ax", "ay", "az", "mass", "radius", "max_vel",
           "agility", "sf", "follow_range", "time_factor", "last_collision")
INTS = ("id", "mode", "effect_stamp", "follow_id", "owner_id", "follow_ptr", "massive", "sensor")


def _apply(park, c) -> None:
    _, op, a, b, x, y, z, rr, w = c
    if op == S.GOTO_OP:
        park.goto_point(a, x, y, z)
    elif op == S.DIR_OP:
        park.goto_direction(a, x, y, z)
    elif op == S.STOP_OP:
        park.stop(a)
    elif op == S.FOLLOW_OP:
        park.follow_ball(a, b, rr)
    elif op == S.ORBIT_OP:
        park.orbit(a, b, rr)
    elif op == S.WARP_OP:
        park.warp_to(a, x, y, z, rr, w)
    elif op == S.WARPIN_OP:
        park.entity_warp_in(a, x, y, z, w)
    else:
        raise ValueError(op)


def run_world(w: dict) -> list:
    """[(step, ball index, doubles, ints, followers)] after every tick."""
    park = modes.ModeBallpark(friction=w["friction"], tick_interval_ms=w["tick"], use_new_orbit=w["new_orbit"],
                              is_master=True)
    park.current_time = w["start"]
    for b in w["balls"]:
        park.add_ball(b["id"], b["mass"], b["radius"], b["max_vel"], b["pos"], b["vel"], b["agility"], b["sf"],
                      is_free=b["free"], is_massive=b["massive"], is_interactive=b["interactive"],
                      is_cloaked=b["cloaked"], bubble=b["bubble"], is_moribund=b["moribund"], mode=b["mode"])
    order = sorted(b["id"] for b in w["balls"])
    out = []
    for step in range(1, w["steps"] + 1):
        for c in w["cmds"]:
            if c[0] == step:
                _apply(park, c)
        park.evolve()
        for k, bid in enumerate(order):
            b = park.balls[bid]
            d = [*b.pos, *b.vel, *b.goto, *b.last_g, b.mass, b.radius, b.max_vel, b.agility, b.speed_fraction,
                 b.follow_range, b.time_factor, b.last_collision]
            q = [b.id, b.mode, b.effect_stamp, b.follow_id, b.owner_id,
                 -1 if b.follow_ptr is None else b.follow_ptr, int(b.is_massive), int(b.sensor_active)]
            out.append((step, k, d, q, sorted(b.followers)))
    return out


def main(seed: int = 1, scale: int = 1, name: str = "modes_golden.npz") -> None:
    worlds = S.cases(seed, scale)
    kept, skipped = [], []
    rec_world, rec_step, rec_ball, rec_d, rec_q, fol_off, fol_ids = [], [], [], [], [], [0], []
    for wi, w in enumerate(worlds):
        try:
            rows = run_world(w)
        except (ZeroDivisionError, ValueError, OverflowError) as e:
            skipped.append([wi, type(e).__name__])
            continue
        kept.append(wi)
        for step, k, d, q, fol in rows:
            rec_world.append(wi)
            rec_step.append(step)
            rec_ball.append(k)
            rec_d.append(d)
            rec_q.append(q)
            fol_ids.extend(fol)
            fol_off.append(len(fol_ids))
    d = np.array(rec_d, dtype=np.float64).view(np.uint64)
    np.savez_compressed(HERE / name, seed=np.int64(seed), worlds=np.array(json.dumps(worlds)),
                        kept=np.array(kept, dtype=np.int64), skipped=np.array(json.dumps(skipped)),
                        world=np.array(rec_world, dtype=np.int32), step=np.array(rec_step, dtype=np.int32),
                        ball=np.array(rec_ball, dtype=np.int32), d=d, q=np.array(rec_q, dtype=np.int64),
                        fol_off=np.array(fol_off, dtype=np.int64), fol_ids=np.array(fol_ids, dtype=np.int64),
                        doubles=np.array(DOUBLES), ints=np.array(INTS))
    print(f"{name}: {len(kept)} worlds kept, {len(skipped)} skipped, {len(rec_d)} records", file=sys.stderr)


if __name__ == "__main__":
    main(int(sys.argv[1]) if len(sys.argv) > 1 else 1, int(sys.argv[2]) if len(sys.argv) > 2 else 1,
         sys.argv[3] if len(sys.argv) > 3 else "modes_golden.npz")
