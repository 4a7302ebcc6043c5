#!/usr/bin/env python3
"""Writes tests/golden/missile_golden.npz (with `--formations`, formation_golden.npz: `formation_cases`, LoadFormations,
SetBallFormation and FormationFollow; with `--cloak`, cloak_golden.npz: `cloak_cases`, CloakBall and UncloakBall;
with `--mushrooms`, mushroom_golden.npz: `mushroom_cases`, AddMushroom and SetBallFree, ball 0 recorded once added;
with `--minis`, mini_golden.npz: `mini_cases`, a fixed parent's miniballs, minicapsules and miniboxes, and force fields
by SetBallHarmonic; every ball of the park recorded, the miniballs with their local ids; with `--mini-own`,
mini_own_golden.npz: `mini_own_cases`, defender missiles through the miniballs of a parent whose id is their owner's
(Gradient's skip) or another's; with `--remove`,
remove_golden.npz: `remove_cases`, RemoveBall between ticks, a removed ball out of the records from then on; with
`--mass-remove`, mass_remove_golden.npz: `mass_remove_cases`, more delayed removals due at once than one
BringOutTheDeads takes): the reference's own scalar ballpark with missiles
(src/exact_ballpark/full.py, `FullBallpark(dynamical_orientation=False)`, a client's view: the partition and the
gradient method) on the seeded worlds of tests/scenarios.py (`missile_cases`: LaunchMissile and MissileFollow among
ships and fixed spheres), every ball's state as added and after every tick, as bits: the flight fields, mLastC and
mLastCollision, the filing, the collision list in order and the display yaw, pitch and roll. With `--do` it writes
missile_do_golden.npz: the same worlds with dynamical orientation (`do_missile_cases`), the turning state in place of
the yaw, pitch and roll.

Run only where the reference is mounted (/root/reference); the fixture it writes is committed and nothing else reads
the reference. The setup of make_golden_collide.py and make_golden_ypr.py is reused: the reference's glibc 2.34
aarch64 transcriptions in place of the C library, the math.fma shim.
"""

from __future__ import annotations

import json
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE))
import make_golden_ypr as Y  # noqa: E402  (installs the redirections)

C = Y.C
S = C.S
full = C.full
import exact_ballpark.dlibm_powf as dlibm_powf  # noqa: E402

full._powf = dlibm_powf.powf                        # the mushrooms' growth: the reference's glibc powf transcription


def _apply(park, c) -> None:
    _, op, a, b, x, y, z, rr, w = c
    if op == S.LAUNCH_OP:
        park.launch_missile(a, b, w, int(x), int(y))
    elif op == S.MFOLLOW_OP:
        park.missile_follow(a, b, w)
    elif op == S.SETFORM_OP:
        park.set_ball_formation(a, w)
    elif op == S.FFOLLOW_OP:
        park.formation_follow(a, b, w)
    elif op == S.CLOAK_OP:
        park.cloak_ball(a, w)
    elif op == S.UNCLOAK_OP:
        park.uncloak_ball(a)
    elif op == S.MUSHROOM_OP:
        park.add_mushroom(a, rr, x)
    elif op == S.FREE_OP:
        park.set_free(a, bool(w))
    elif op == S.HARMONIC_OP:
        park.set_harmonic(a, w, b, int(x), bool(y))
    elif op == S.REMOVE_OP:
        park.remove_ball(a, w)
    elif op == S.BRING_OP:
        park.bring_out_the_deads()
    else:
        C._apply(park, c)


def _state(b, do: bool) -> tuple:
    return C._state(b, True) if do else Y._state(b)


def run_world(w: dict) -> list:
    do = bool(w.get("dynamical"))
    park = full.FullBallpark(friction=w["friction"], tick_interval_ms=w["tick"], use_new_orbit=w["new_orbit"],
                             dynamical_orientation=do)
    park.current_time = w["start"]
    for b in w["balls"]:
        kw = {"max_ang_vel": b["max_ang_vel"], "rot_agility": b["rot_agility"]} if do else {}
        park.add_ball(b["id"], b["mass"], b["radius"], b["max_vel"], is_free=b["free"], is_massive=b["massive"],
                      is_interactive=b["interactive"], is_space_junk=b.get("junk", False), x=b["pos"][0],
                      y=b["pos"][1], z=b["pos"][2], vx=b["vel"][0], vy=b["vel"][1], vz=b["vel"][2],
                      agility=b["agility"], speed_fraction=b["sf"], **kw)
        if b["cloaked"]:
            park.set_cloaked(b["id"], b["cloaked"])
        park.balls[b["id"]].bubble = b["bubble"]
    for m in w.get("minis", []):                   # each taking the next local id, in this order
        if m["kind"] == "ball":
            park.add_mini_ball(m["parent"], *m["offset"], m["radius"])
        elif m["kind"] == "capsule":
            park.add_mini_capsule(m["parent"], *m["a"], *m["b"], m["radius"])
        else:
            park.add_mini_box(m["parent"], *m["corner"], *m["lx"], *m["ly"], *m["lz"])
    if w.get("formations"):
        park.load_formations(w["formations"])
    order = sorted(park.balls)                      # the world's balls, its miniballs, a client's mushroom (ball 0)
    out, events = [], []
    park.events.clear()
    for k, bid in enumerate(order):
        d, q = _state(park.balls[bid], do)
        out.append((0, k, d, q, list(park.balls[bid].collisions)))
    for step in range(1, w["steps"] + 1):
        for c in w["cmds"]:
            if c[0] == step:
                _apply(park, c)
        park.evolve()
        events.extend(e for e in park.events if e[1] == "DoCollision")
        park.events.clear()
        order = sorted(park.balls)
        for k, bid in enumerate(order):
            d, q = _state(park.balls[bid], do)
            out.append((step, k, d, q, list(park.balls[bid].collisions)))
    return out, events


def main(seed: int = 1, do: bool = False, what: str = "missile") -> None:
    name = f"{what}_do_golden.npz" if do else f"{what}_golden.npz"
    if what == "formation":
        ws = S.do_formation_cases(seed) if do else S.formation_cases(seed)
    elif what == "cloak":
        ws = S.do_cloak_cases(seed) if do else S.cloak_cases(seed)
    elif what == "mushroom":
        ws = S.do_mushroom_cases(seed) if do else S.mushroom_cases(seed)
    elif what == "remove":
        ws = S.remove_cases(seed) + S.delayed_remove_cases(seed)
    elif what == "mass_remove":
        ws = S.mass_remove_cases(seed)
    elif what == "mini":
        ws = S.do_mini_cases(seed) if do else S.mini_cases(seed)
    elif what == "mini_own":
        ws = S.mini_own_cases(seed)
    else:
        ws = S.do_missile_cases(seed) if do else S.missile_cases(seed)
    if not do:
        for w in ws:
            w["ypr"] = True
    kept, skipped = [], []
    rec_world, rec_step, rec_ball, rec_d, rec_q, hit_off, hit_ids = [], [], [], [], [], [0], []
    ev_world, ev_q, ev_c = [], [], []                # DoCollision: (time, missile id, collider id), mLastC
    for wi, w in enumerate(ws):
        try:
            rows, evs = run_world(w)
        except (ZeroDivisionError, ValueError, OverflowError, C.space.DestinyUndefined, NotImplementedError) as e:
            skipped.append([wi, type(e).__name__, str(e)[:80]])
            continue
        kept.append(wi)
        for e in evs:
            ev_world.append(wi)
            ev_q.append([e[0], e[2], e[3]])
            ev_c.append([e[4], e[5], e[6]])
        for step, k, d, q, hits in rows:
            rec_world.append(wi)
            rec_step.append(step)
            rec_ball.append(k)
            rec_d.append(d)
            rec_q.append(q)
            hit_ids.extend(hits)
            hit_off.append(len(hit_ids))
    d = np.array(rec_d, dtype=np.float64).view(np.uint64)
    names = C.DOUBLES + (C.TURNING if do else Y.YPR)
    np.savez_compressed(HERE / name, seed=np.int64(seed), worlds=np.array(json.dumps(ws)),
                        kept=np.array(kept, dtype=np.int64), skipped=np.array(json.dumps(skipped)),
                        world=np.array(rec_world, dtype=np.int32), step=np.array(rec_step, dtype=np.int32),
                        ball=np.array(rec_ball, dtype=np.int32), d=d, q=np.array(rec_q, dtype=np.int64),
                        hit_off=np.array(hit_off, dtype=np.int64), hit_ids=np.array(hit_ids, dtype=np.int64),
                        doubles=np.array(names), ints=np.array(C.INTS),
                        ev_world=np.array(ev_world, dtype=np.int32), ev_q=np.array(ev_q, dtype=np.int64).reshape(-1, 3),
                        ev_c=np.array(ev_c, dtype=np.float64).reshape(-1, 3).view(np.uint64))
    n_hits = sum(1 for i in range(1, len(hit_off)) if hit_off[i] > hit_off[i - 1])
    print(f"{name}: {len(ev_world)} DoCollision events", file=sys.stderr)
    print(f"{name}: {len(kept)} worlds kept, {len(skipped)} skipped {skipped[:5]}, {len(rec_d)} records, "
          f"{n_hits} with collisions", file=sys.stderr)


if __name__ == "__main__":
    args = [a for a in sys.argv[1:] if not a.startswith("--")]
    main(int(args[0]) if args else 1, do="--do" in sys.argv,
         what="formation" if "--formations" in sys.argv else "cloak" if "--cloak" in sys.argv
         else "mushroom" if "--mushrooms" in sys.argv else "mini_own" if "--mini-own" in sys.argv
         else "mini" if "--minis" in sys.argv
         else "remove" if "--remove" in sys.argv else "mass_remove" if "--mass-remove" in sys.argv else "missile")
