#!/usr/bin/env python3
"""Writes tests/golden/space_golden.npz: the reference's own scalar ballpark on the master (src/exact_ballpark/full.py,
`FullBallpark(is_master=True, seed=...)`: space.py's partition with its top boxes and bubbles, the bubble
dictionaries, the interactive and keep-alive sets, the dead bubbles' skip, HandleProximities with its sensors,
notification ranges and target tracking in libstdc++'s unordered_map order, glibc rand() for shuffled sensors) on the
seeded worlds of tests/scenarios.py (`space_cases`; with `--formations`, formation_master_golden.npz:
`formation_master_cases`, LoadFormations, ReserveFormationSlot and FreeFormationSlot by hand and the slots that
Stop(Ball*) frees on the master, each ball's reservation bits and each reservation's answer in `master`), after the load (step 0: the balls as added, then the step-0
commands) and after every tick:

- every ball's state as bits (the flight fields, the filing, the display yaw, pitch and roll), in ascending id;
- the master's structures as JSON (`master`, one record a world a step): each ball's bubble and old bubble, the top
  boxes that exist with their bubbles, the bubble dictionaries in their order, the interactive and keep-alive sets, the
  proximity map's keys in its order with its bucket count and next resize, each ball's sensor list likewise, the
  bubble counter, the sensor events since the last record (DoProximity, DoRange, DoTargetTracking, DoFinalCleanup),
  and the queries: CheckVisibility from the first ball to each other, GetBallIdsInRange (and with distances) from each
  ball and from a point, GetBubbleAtCoordinates at each ball, GetRemoteFollowers.

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
EVENTS = ("DoProximity", "DoRange", "DoTargetTracking", "DoFinalCleanup")


def _apply(park, c, reserved: list) -> None:
    _, op, a, b, x, y, z, rr, w = c
    if op == S.RESERVE_OP:
        reserved.append([int(a), int(park.reserve_formation_slot(a))])
    elif op == S.FREESLOT_OP:
        park.free_formation_slot(a, w)
    elif op == S.KEEPALIVE_OP:
        park.add_keep_alive_ball(a)
        bl = park.balls.get(a)
        if bl is not None:                      # the Python layer adds to the set only; the next AddToBubble counts it
            pass
    elif op == S.SENSOR_OP:
        park.add_proximity_sensor(a, rr, x, int(w & 1), bool(w & 2))
    elif op == S.UNSENSOR_OP:
        park.remove_proximity_sensor(a)
    elif op == S.NOTIFY_OP:
        park.set_notification_range(a, rr)
    elif op == S.TRACK_OP:
        park.set_target_tracking(a, b, rr)
    elif op == S.INTERACTIVE_OP:
        park.set_interactive(a, bool(w))
    elif op == S.INIT_BUBBLES_OP:
        park.initialize_bubbles(a)
    elif op == S.COPY_BUBBLES_OP:
        park.copy_bubbles()
    else:
        M._apply(park, c)


def master_state(park, w: dict, reserved: list) -> dict:
    ids = sorted(park.balls)
    tops = sorted([list(k) + [int(box.bubble)] for k, box in park.levels[0].items()])
    dicts = [[int(bub), [[int(i), int(a)] for i, a in d.items()]] for bub, d in park.bubbles.items()]
    inter = sorted([int(k), sorted(int(i) for i in v)] for k, v in park.bubble_interactives.items())
    keep = sorted([int(k), sorted(int(i) for i in v)] for k, v in park.bubble_keep_alives.items())
    sensors = []
    for i in ids:
        b = park.balls[i]
        m = b.sensor_balls
        sensors.append([int(i), [int(k) for k in m.keys()], int(m.bucket_count), int(m.next_resize)])
    pm = park.proximity_balls
    events = [[e[1], int(e[2])] + [int(x) for x in e[3:]] for e in park.events if e[1] in EVENTS]
    park.events.clear()
    q = w["queries"]
    R = q["range"]
    first = park.balls[ids[0]] if ids else None
    vis = [int(park.check_visibility(ids[0], j)) for j in ids[1:]] if ids else []
    rng = [[int(x) for x in (park.ball_ids_in_range(i, R) or [])] for i in ids]
    rngd = [[[float(d).hex(), int(x)] for d, x in (park.ball_ids_in_range(i, R, include_dist=True) or [])]
            for i in ids]
    if first is not None:
        pt = [first.pos[k] + q["point_offset"][k] for k in range(3)]
        point = [int(x) for x in (park.ball_ids_in_range(pt[0], pt[1], pt[2], R) or [])]
    else:
        pt, point = [0.0, 0.0, 0.0], []
    at = [park.bubble_at(*park.balls[i].pos) for i in ids]
    return {"ids": [int(i) for i in ids],
            "bubble": [int(park.balls[i].bubble) for i in ids],
            "old_bubble": [int(park.balls[i].old_bubble) for i in ids],
            "tops": tops, "dicts": dicts, "inter": inter, "keep": keep,
            "keep_balls": sorted(int(i) for i in park.bubble_keep_alive_balls),
            "prox": [int(k) for k in pm.keys()], "prox_bkt": [int(pm.bucket_count), int(pm.next_resize)],
            "sensors": sensors, "bubble_id": int(park.bubble_id), "events": events,
            "vis": vis, "range": rng, "range_dist": rngd, "point": [float(c).hex() for c in pt], "point_ids": point,
            "at": [None if a is None else int(a) for a in at], "remote": [int(x) for x in park.remote_followers()],
            "fslots": [int(park.balls[i].formation_slots) for i in ids], "reserved": reserved}


def run_world(w: dict) -> tuple:
    park = full.FullBallpark(friction=w["friction"], tick_interval_ms=w["tick"], use_new_orbit=w["new_orbit"],
                             is_master=True, seed=w["seed"])
    park.current_time = w["start"]
    if w.get("formations"):
        park.load_formations(w["formations"])
    for b in w["balls"]:
        park.add_ball(b["id"], b["mass"], b["radius"], b["max_vel"], is_free=b["free"], is_massive=b["massive"],
                      is_interactive=b["interactive"], x=b["pos"][0], y=b["pos"][1], z=b["pos"][2],
                      vx=b["vel"][0], vy=b["vel"][1], vz=b["vel"][2], agility=b["agility"], speed_fraction=b["sf"])
    out, master = [], []
    for step in range(0, w["steps"] + 1):
        reserved = []
        for c in w["cmds"]:
            if c[0] == step:
                _apply(park, c, reserved)
        if step:
            park.evolve()
        for k, bid in enumerate(sorted(park.balls)):
            d, q = Y._state(park.balls[bid])
            out.append((step, k, d, q, list(park.balls[bid].collisions)))
        master.append(master_state(park, w, reserved))
    return out, master


def main(seed: int = 1, formations: bool = False) -> None:
    ws = S.formation_master_cases(seed) if formations else S.space_cases(seed)
    name = "formation_master_golden.npz" if formations else "space_golden.npz"
    for w in ws:
        w["ypr"] = True
    kept, skipped, masters = [], [], []
    rec_world, rec_step, rec_ball, rec_d, rec_q, hit_off, hit_ids = [], [], [], [], [], [0], []
    for wi, w in enumerate(ws):
        try:
            rows, master = run_world(w)
        except (ZeroDivisionError, ValueError, OverflowError, C.space.DestinyUndefined, NotImplementedError) as e:
            skipped.append([wi, type(e).__name__, str(e)[:80]])
            continue
        kept.append(wi)
        masters.append(master)
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
                        doubles=np.array(C.DOUBLES + Y.YPR), ints=np.array(C.INTS),
                        master=np.array(json.dumps(masters)))
    n_ev = sum(len(r["events"]) for m in masters for r in m)
    print(f"{name}: {len(kept)} worlds kept, {len(skipped)} skipped {skipped[:5]}, {len(rec_d)} records, "
          f"{n_ev} sensor events", file=sys.stderr)


if __name__ == "__main__":
    args = [a for a in sys.argv[1:] if not a.startswith("--")]
    main(int(args[0]) if args else 1, formations="--formations" in sys.argv)
