#!/usr/bin/env python3
"""Writes tests/golden/ypr_golden.npz: the reference's own scalar ballpark without dynamical orientation
(src/exact_ballpark/full.py, `FullBallpark(dynamical_orientation=False)`, a client's view, so with its partition and
gradient collisions) on the seeded worlds of tests/scenarios.py (`cases` less the guard worlds, whose loaded modes
AddBall cannot make, and `collide_cases`), every ball's state after every tick as bits, with the display orientation
Ball::CalculateYawPitchRoll keeps (yaw, pitch, roll, their old values, the roll speed and its old value, the yaw
change): snapped when a free ball is added and when its velocity is set, evolved at every tick's commit.

Run only where the reference is mounted (/root/reference); the fixture it writes is committed and nothing else reads
the reference. The setup of make_golden_collide.py is reused (the reference's glibc 2.34 aarch64 transcriptions in
place of the C library, the math.fma shim); sin, asin and atan2 go to the transcriptions too (dlibm.sin,
dlibm_inv.asin and atan2).
"""

from __future__ import annotations

import json
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE))
import make_golden_collide as C  # noqa: E402  (installs the redirections)
import exact_ballpark.dlibm as dlibm  # noqa: E402
import exact_ballpark.dlibm_inv as dlibm_inv  # noqa: E402

C._maths.sin = dlibm.sin
C._maths.asin = dlibm_inv.asin
C._maths.atan2 = dlibm_inv.atan2
S = C.S
full = C.full
YPR = ("yaw", "pitch", "roll", "old_yaw", "old_pitch", "old_roll", "roll_speed", "old_roll_speed", "yaw_delta")


def _state(b) -> tuple:
    d, q = C._state(b)
    return d + [float(b.yaw), float(b.pitch), float(b.roll_angle), float(b.old_yaw), float(b.old_pitch),
                float(b.old_roll), float(b.roll_speed), float(b.old_roll_speed), float(b.yaw_delta)], q


def run_world(w: dict) -> list:
    park = full.FullBallpark(friction=w["friction"], tick_interval_ms=w["tick"], use_new_orbit=w["new_orbit"],
                             dynamical_orientation=False)
    park.current_time = w["start"]
    for b in w["balls"]:
        park.add_ball(b["id"], b["mass"], b["radius"], b["max_vel"], is_free=b["free"], is_massive=b["massive"],
                      is_interactive=b["interactive"], is_space_junk=b.get("junk", False), x=b["pos"][0],
                      y=b["pos"][1], z=b["pos"][2], vx=b["vel"][0], vy=b["vel"][1], vz=b["vel"][2],
                      agility=b["agility"], speed_fraction=b["sf"])
        if b["cloaked"]:
            park.set_cloaked(b["id"], b["cloaked"])
    order = sorted(b["id"] for b in w["balls"])
    out = []
    for k, bid in enumerate(order):
        d, q = _state(park.balls[bid])
        out.append((0, k, d, q, list(park.balls[bid].collisions)))
    for step in range(1, w["steps"] + 1):
        for c in w["cmds"]:
            if c[0] == step:
                C._apply(park, c)
        park.evolve()
        for k, bid in enumerate(order):
            d, q = _state(park.balls[bid])
            out.append((step, k, d, q, list(park.balls[bid].collisions)))
    return out


def worlds(seed: int) -> list:
    out = [w for w in S.cases(seed) if w["kind"] != "guards"] + S.collide_cases(seed)
    for w in out:
        w["collisions"] = True
        w["ypr"] = True
        for b in w["balls"]:
            b.setdefault("junk", False)
    return out


def main(seed: int = 1, name: str = "ypr_golden.npz") -> None:
    ws = worlds(seed)
    kept, skipped = [], []
    rec_world, rec_step, rec_ball, rec_d, rec_q, hit_off, hit_ids = [], [], [], [], [], [0], []
    for wi, w in enumerate(ws):
        try:
            rows = run_world(w)
        except (ZeroDivisionError, ValueError, OverflowError, C.space.DestinyUndefined, NotImplementedError) as e:
            skipped.append([wi, type(e).__name__, str(e)[:80]])
            continue
        kept.append(wi)
        for step, k, d, q, hits in rows:
            rec_world.append(wi)
            rec_step.append(step)
            rec_ball.append(k)
            rec_d.append(d)
            rec_q.append(q)
            hit_ids.extend(hits)
            hit_off.append(len(hit_ids))
    d = np.array(rec_d, dtype=np.float64).view(np.uint64)
    np.savez_compressed(HERE / name, seed=np.int64(seed), worlds=np.array(json.dumps(ws)),
                        kept=np.array(kept, dtype=np.int64), skipped=np.array(json.dumps(skipped)),
                        world=np.array(rec_world, dtype=np.int32), step=np.array(rec_step, dtype=np.int32),
                        ball=np.array(rec_ball, dtype=np.int32), d=d, q=np.array(rec_q, dtype=np.int64),
                        hit_off=np.array(hit_off, dtype=np.int64), hit_ids=np.array(hit_ids, dtype=np.int64),
                        doubles=np.array(C.DOUBLES + YPR), ints=np.array(C.INTS))
    print(f"{name}: {len(kept)} worlds kept, {len(skipped)} skipped {skipped[:5]}, {len(rec_d)} records",
          file=sys.stderr)


if __name__ == "__main__":
    main(int(sys.argv[1]) if len(sys.argv) > 1 else 1)
