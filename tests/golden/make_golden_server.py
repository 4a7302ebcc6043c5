#!/usr/bin/env python3
"""Writes tests/golden/server_golden.npz (with `--do`, server_do_golden.npz; with `--space`, server_space_golden.npz
over worlds of several bubbles): the reference's own server layer
(src/exact_ballpark/server.py: Destiny's Actions, Ticker, BubbleUpdater and ParkUpdateBatcher on a master
DestinyPark over its FullBallpark, with src/exact_ballpark/client.py's PilotInterests and RecordingNetwork), driven as
Destiny's own test helper drives it: each tick the world's calls in order, the Ticker's tick, the park's Evolve, the
batcher's send_batch. The worlds are tests/server_scenarios.py's `server_cases` (`space_server_cases`). Recorded, as
bits: every ball's
state after every tick (the flight fields, the display yaw, pitch and roll or with `--do` the turning state, the
filing, the flags, the cloak byte, the bubble and old bubble), every update each client received at each tick (its
kind, method, state, wait flag and count, the ball streams as bytes), and each pending query's answer.

Run only where the reference is mounted (/root/reference); the fixture it writes is committed and nothing else reads
the reference. The setup of make_golden_missile.py is reused: the reference's glibc 2.34 aarch64 transcriptions in
place of the C library, the math.fma shim.
"""

from __future__ import annotations

import argparse
import json
import logging
import sys
from collections import defaultdict
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE))
sys.path.insert(0, str(HERE.parent))
import make_golden_missile as M  # noqa: E402  (installs the redirections)
import server_scenarios as SS  # noqa: E402
from exact_ballpark import client as cl  # noqa: E402
from exact_ballpark import server as srv  # noqa: E402

Y, C = M.Y, M.C
logging.getLogger("exact_ballpark.server").setLevel(logging.CRITICAL)    # Destiny's logged stamp mismatches


def _state(b, do: bool) -> tuple:
    d, q = C._state(b, True) if do else Y._state(b)
    q = q + [int(b.is_interactive), int(b.is_free), int(b.is_cloaked), int(b.bubble), int(b.old_bubble)]
    return d, q


def _per_client(sent: list) -> dict:
    out = defaultdict(list)
    for kind, to, method, state, wait, count in sent:
        for c in ([to] if kind == "single" else to):
            out[c].append([kind, method, state, wait, count])
    return {str(c): v for c, v in sorted(out.items())}


def run_world(w: dict, space: bool = False) -> tuple:
    do = bool(w["do"])
    if do:
        srv.enable_dynamical_orientation()
    else:
        srv.disable_dynamical_orientation()
    park = srv.DestinyPark(is_master=True)
    for b in w["balls"]:
        args = (b["id"], b["mass"], b["radius"], b["max_vel"], 1, 0, int(b["massive"]), int(b["interactive"]), 0,
                *b["pos"], *b["vel"], b["agility"], b["sf"])
        if do:
            args += (b["max_ang_vel"], b["rot_agility"])
        park.AddBall(*args)
    interests = cl.PilotInterests(w["pilots"])
    network = cl.RecordingNetwork()
    actions = srv.Actions(park)
    batcher = srv.ParkUpdateBatcher(park, network, interests, interests, srv.BubbleUpdater(park))
    ticker = srv.Ticker(park, interests, actions, batcher)
    lane = park.lane
    order = [b["id"] for b in w["balls"]]
    states, sends, answers = [], [], []
    for t, calls in enumerate(w["ticks"], start=1):
        ans = []
        for kind, name, args in calls:
            if kind == "act":
                getattr(actions, name)(*args)
            else:
                ans.append([name, args[0], SS.canon(getattr(actions, name)(*args))])
        answers.append(ans)
        ticker.tick()
        park.Evolve()
        batcher.send_batch()
        sends.append(SS.canon(_per_client(network.take())))
        if not space and len({b.bubble for b in lane.balls.values()}) != 1:
            raise ValueError("a ball left the world's bubble")
        for k, bid in enumerate(order):
            d, q = _state(lane.balls[bid], do)
            states.append((t, k, d, q))
    return states, sends, answers


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--do", action="store_true")
    ap.add_argument("--space", action="store_true")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--worlds", type=int, default=90)
    a = ap.parse_args()
    ws = SS.space_server_cases(a.seed, a.worlds) if a.space else SS.server_cases(a.seed, a.worlds, do=a.do)
    kept, skipped = [], []
    rw, rt, rb, rd, rq, sends, answers = [], [], [], [], [], [], []
    for wi, w in enumerate(ws):
        try:
            states, snd, ans = run_world(w, a.space)
        except (ZeroDivisionError, ValueError, OverflowError, C.space.DestinyUndefined, NotImplementedError) as e:
            skipped.append([wi, type(e).__name__, str(e)[:80]])
            continue
        kept.append(wi)
        for t, k, d, q in states:
            rw.append(wi), rt.append(t), rb.append(k), rd.append(d), rq.append(q)
        sends.append(snd)
        answers.append(ans)
    name = "server_space_golden.npz" if a.space else "server_do_golden.npz" if a.do else "server_golden.npz"
    np.savez_compressed(HERE / name, world=np.array(rw, dtype=np.int32), step=np.array(rt, dtype=np.int32),
                        ball=np.array(rb, dtype=np.int32), d=np.array(rd, dtype=np.float64).view(np.int64),
                        q=np.array(rq, dtype=np.int64), kept=np.array(kept, dtype=np.int32), seed=np.int64(a.seed),
                        worlds=json.dumps(ws), sends=json.dumps(sends), answers=json.dumps(answers),
                        skipped=json.dumps(skipped))
    print(f"{name}: {len(kept)} worlds kept, {len(skipped)} skipped {skipped[:4]}, {len(rw)} ball-states")


if __name__ == "__main__":
    main()
