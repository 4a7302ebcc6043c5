#!/usr/bin/env python3
"""Writes tests/golden/ontick_golden.npz: the reference's own scalar ballpark (src/exact_ballpark/full.py,
`FullBallpark`, with the display yaw, pitch and roll, its partition and gradient method) driven by Ballpark::OnTick
alone on the seeded worlds of tests/scenarios.py (`ontick_cases`: delayed removals, BringOutTheDeads only through
OnTick, schedules of simulation times that are too early, one interval, several with a remainder carried, and
backlogs; a third on a master (`is_master=True`, every ball interactive) whose backlogs above five intervals run
one). Every ball's state as added and after every OnTick call, as bits, and each call's tick count
(`current_time`, the `clock` column).

Run only where the reference is mounted (/root/reference); the fixture it writes is committed and nothing else reads
the reference. The setup of make_golden_missile.py is reused: the reference's glibc 2.34 aarch64 transcriptions in
place of the C library, the math.fma shim.
"""

from __future__ import annotations

import json
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE))
import make_golden_missile as M  # noqa: E402  (installs the redirections)

Y, C, S = M.Y, M.C, M.S
full = C.full


def run_world(w: dict) -> list:
    park = full.FullBallpark(friction=w["friction"], tick_interval_ms=w["tick"], use_new_orbit=w["new_orbit"],
                             is_master=bool(w["master"]))
    park.current_time = w["start"]
    for b in w["balls"]:
        park.add_ball(b["id"], b["mass"], b["radius"], b["max_vel"], is_free=b["free"], is_massive=b["massive"],
                      is_interactive=b["interactive"], is_space_junk=b.get("junk", False), x=b["pos"][0],
                      y=b["pos"][1], z=b["pos"][2], vx=b["vel"][0], vy=b["vel"][1], vz=b["vel"][2],
                      agility=b["agility"], speed_fraction=b["sf"])
        if b["cloaked"]:
            park.set_cloaked(b["id"], b["cloaked"])
        if not w["master"]:                      # the master's own bubbles stand (its partition assigns them)
            park.balls[b["id"]].bubble = b["bubble"]
    if w.get("formations"):
        park.load_formations(w["formations"])
    out, ticks = [], [park.current_time]
    for k, bid in enumerate(sorted(park.balls)):
        d, q = Y._state(park.balls[bid])
        out.append((0, k, d, q, list(park.balls[bid].collisions)))
    for call, sim in enumerate(w["calls"], start=1):
        for c in w["cmds"]:
            if c[0] == call:
                M._apply(park, c)
        park.on_tick(sim)
        ticks.append(park.current_time)
        for k, bid in enumerate(sorted(park.balls)):
            d, q = Y._state(park.balls[bid])
            out.append((call, k, d, q, list(park.balls[bid].collisions)))
    return out, ticks


def main(seed: int = 1) -> None:
    ws = S.ontick_cases(seed)
    for w in ws:
        w["ypr"] = True
        w["collisions"] = True
    kept, skipped = [], []
    rec_world, rec_step, rec_ball, rec_d, rec_q, hit_off, hit_ids = [], [], [], [], [], [0], []
    for wi, w in enumerate(ws):
        try:
            rows, ticks = run_world(w)
        except (ZeroDivisionError, ValueError, OverflowError, C.space.DestinyUndefined, NotImplementedError) as e:
            skipped.append([wi, type(e).__name__, str(e)[:80]])
            continue
        kept.append(wi)
        w["ticks_after"] = ticks
        for step, k, d, q, hits in rows:
            rec_world.append(wi)
            rec_step.append(step)
            rec_ball.append(k)
            rec_d.append(d)
            rec_q.append(q)
            hit_ids.extend(hits)
            hit_off.append(len(hit_ids))
    d = np.array(rec_d, dtype=np.float64).view(np.uint64)
    np.savez_compressed(HERE / "ontick_golden.npz", seed=np.int64(seed), worlds=np.array(json.dumps(ws)),
                        kept=np.array(kept, dtype=np.int64), skipped=np.array(json.dumps(skipped)),
                        world=np.array(rec_world, dtype=np.int32), step=np.array(rec_step, dtype=np.int32),
                        ball=np.array(rec_ball, dtype=np.int32), d=d, q=np.array(rec_q, dtype=np.int64),
                        hit_off=np.array(hit_off, dtype=np.int64), hit_ids=np.array(hit_ids, dtype=np.int64),
                        doubles=np.array(C.DOUBLES + Y.YPR), ints=np.array(C.INTS))
    print(f"ontick_golden.npz: {len(kept)} worlds kept, {len(skipped)} skipped {skipped[:5]}, {len(rec_d)} records",
          file=sys.stderr)


if __name__ == "__main__":
    args = [a for a in sys.argv[1:] if not a.startswith("--")]
    main(int(args[0]) if args else 1)
