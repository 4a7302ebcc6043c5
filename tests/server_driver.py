"""Runs tests/server_scenarios.py's worlds through the batched server layer (exact_ballpark_amd.vec_server over a
VecModesPark on the CPU or a GPU) as Destiny's test helper drives the scalar one: each tick every world's calls in
order (the j-th entries of the worlds together: Actions calls grouped by kind into pushes, undos and pending queries
batched), the pre-tick, the park's tick and the post-tick; and compares the result with the golden fixtures
tests/golden/make_golden_server.py wrote from the reference's own server layer.
"""

from __future__ import annotations

import json
import math
import struct
from collections import defaultdict
from pathlib import Path

import numpy as np
import torch

import server_scenarios as SS
from exact_ballpark_amd import vec_server as VS
from exact_ballpark_amd import vec_server_space as VSS
from exact_ballpark_amd.vec_modes import VecModesPark

GOLDEN = Path(__file__).resolve().parent / "golden"
F64, I64 = torch.float64, torch.int64
# the compared columns of the fixture's doubles: the flight fields through mLastCollision (not mLastC, the
# gradient's), then the display orientation or the turning state; of its ints: id .. sensor, interactive, free, cloak
D_FLIGHT = list(range(26))
Q_COLS = [0, 1, 2, 3, 4, 5, 6, 7, 15, 16, 17]
Q_BUBBLES = [18, 19]                # the bubble and old bubble (compared over the partition)


def load_golden(do: bool, space: bool = False) -> dict:
    name = "server_space_golden.npz" if space else "server_do_golden.npz" if do else "server_golden.npz"
    z = np.load(GOLDEN / name, allow_pickle=False)
    worlds = json.loads(str(z["worlds"]))
    kept = [int(k) for k in z["kept"]]
    d, q = z["d"].view(np.float64), z["q"]
    world, step, ball = z["world"], z["step"], z["ball"]
    states = defaultdict(dict)
    for r in range(len(world)):
        states[(int(world[r]), int(step[r]))][int(ball[r])] = (d[r], q[r])
    return {"worlds": worlds, "kept": kept, "states": states, "sends": json.loads(str(z["sends"])),
            "answers": json.loads(str(z["answers"]))}


def build(worlds: list, device: str, space: bool = False) -> tuple:
    """(VecServer, rows [world][ball index]) for the worlds on one VecModesPark (master, display orientation without
    dynamical orientation), each world's balls in AddBall's order; with `space` a SpaceServer over the master's
    partition."""
    do = bool(worlds[0]["do"])
    rows, index = [], []
    for wi, w in enumerate(worlds):
        index.append([])
        for b in w["balls"]:
            index[-1].append(len(rows))
            rows.append({"id": b["id"], "world": wi, "time": 0, "mass": b["mass"], "radius": b["radius"],
                         "max_vel": b["max_vel"], "pos": b["pos"], "vel": b["vel"], "agility": b["agility"],
                         "speed_fraction": b["sf"], "free": True, "massive": b["massive"],
                         "interactive": b["interactive"], "cloaked": 0, "bubble": -1 if space else 0,
                         "moribund": False,
                         "max_ang_vel": b["max_ang_vel"], "rot_agility": b["rot_agility"]})
    park = VecModesPark(len(rows), device=device, dynamical_orientation=do, display_ypr=not do, master=True)
    park.load(rows)
    pilots = [-1] * len(rows)
    for wi, w in enumerate(worlds):
        for k, b in enumerate(w["balls"]):
            p = w["pilots"].get(str(b["id"]), w["pilots"].get(b["id"]))
            if p is not None:
                pilots[index[wi][k]] = int(p)
    if space:
        return VSS.SpaceServer(VSS.SpaceModesAdapter(park), pilots), index
    return VS.VecServer(VS.ModesAdapter(park), pilots), index


def _push_args(kind: str, args: tuple) -> tuple:
    """(ball, target, f[4], n, loc) of one Actions call, as the batched layer takes them."""
    f, n, tgt, loc = [0.0] * 4, 0, 0, 0
    rest = list(args[1:])
    if kind in ("FollowBall", "Orbit"):
        tgt, f[0] = int(rest[0]), float(rest[1])
    elif kind == "WarpTo":
        f[:4] = [float(x) for x in rest[:4]]
        n = int(rest[4])
    elif kind == "EntityWarpIn":
        f[:3] = [float(x) for x in rest[:3]]
        n = int(rest[3])
    elif kind in ("SetBallMassive", "SetBallInteractive", "SetBallTroll", "SetBallFree"):
        n = int(rest[0])
    elif kind == "SetBallPosition":
        f[:3] = [float(x) for x in rest[:3]]
        loc = 1 if rest[3] is True else 0
    elif kind == "CloakBall":
        n, f[0] = int(rest[0]), float(rest[1])
    else:
        f[:len(rest)] = [float(x) for x in rest]
    return int(args[0]), tgt, f, n, loc


def _answer(server: VS.VecServer, name: str, mask: torch.Tensor, ball: torch.Tensor) -> list:
    """Each world's answer as Destiny's Actions returns it (None, an int 0, a bool, an int or a float)."""
    out = getattr(server, name)(mask, ball)
    if name in ("has_pending_cloak", "has_pending_uncloak"):
        return [bool(x) for x in out.tolist()]
    if name in ("has_pending_set_free", "has_pending_set_not_free"):
        ans, pend = out
        return [bool(a) if p else 0 for a, p in zip(ans.tolist(), pend.tolist())]
    v, known = out
    conv = {"get_pending_cloak_mode": int, "get_pending_interactive_state": bool,
            "get_pending_speed_fraction": float}[name]
    return [conv(x) if k else None for x, k in zip(v.tolist(), known.tolist())]


def apply_tick(server: VS.VecServer, worlds: list, t: int) -> list:
    """Tick t's entries of every world, position by position; returns each world's query answers."""
    W, d = len(worlds), server.dev
    answers = [[] for _ in range(W)]
    width = max((len(w["ticks"][t]) for w in worlds), default=0)
    for p in range(width):
        acts, undos, asks = defaultdict(list), [], defaultdict(list)
        for wi, w in enumerate(worlds):
            entries = w["ticks"][t]
            if p >= len(entries):
                continue
            kind, name, args = entries[p]
            if kind == "ask":
                asks[name].append((wi, int(args[0])))
            elif name == "undo_pending_cloak":
                undos.append((wi, int(args[0])))
            else:
                acts[VS.ACTION_KINDS[name]].append((wi, args))
        for kind, items in acts.items():
            mask = torch.zeros(W, dtype=torch.bool)
            ball, target, n, loc = (torch.zeros(W, dtype=I64) for _ in range(4))
            f = torch.zeros(W, 4, dtype=F64)
            for wi, args in items:
                b, tg, ff, nn, lc = _push_args(kind, args)
                mask[wi], ball[wi], target[wi], n[wi], loc[wi] = True, b, tg, nn, lc
                f[wi] = torch.tensor(ff, dtype=F64)
            server.push(kind, mask.to(d), ball.to(d), target.to(d), f.to(d), n.to(d), loc.to(d))
        if undos:
            mask = torch.zeros(W, dtype=torch.bool)
            ball = torch.zeros(W, dtype=I64)
            for wi, b in undos:
                mask[wi], ball[wi] = True, b
            server.undo_pending_cloak(mask.to(d), ball.to(d))
        for name, items in asks.items():
            mask = torch.zeros(W, dtype=torch.bool)
            ball = torch.zeros(W, dtype=I64)
            for wi, b in items:
                mask[wi], ball[wi] = True, b
            got = _answer(server, name, mask.to(d), ball.to(d))
            for wi, b in items:
                answers[wi].append([name, b, SS.canon(got[wi])])
    return answers


def sends_of(server: VS.VecServer) -> list:
    """Each world's sends of the last tick, canonical as the fixture holds them."""
    dec = VSS.decode if isinstance(server, VSS.SpaceServer) else VS.decode
    return [SS.canon({str(c): v for c, v in sorted(per.items())}) for per in dec(server)]


def park_rows(server: VS.VecServer, worlds: list, index: list) -> dict:
    """{(world, ball index): (doubles, ints)} in the fixture's compared columns."""
    p = server.a.park
    do = p.dynamical_orientation
    cols = [p.pos, p.vel, p.goto, p.last_g, p.old_pos, p.old_vel]
    scal = [p.mass, p.radius, p.max_vel, p.agility, p.speed_fraction, p.follow_range, p.time_factor, p.last_collision]
    dd = torch.cat([c.reshape(p.n, -1) for c in cols] + [s.reshape(p.n, 1) for s in scal], dim=1)
    if do:
        extra = torch.cat([p.rot, p.old_rot, p.ang_vel, p.old_ang_vel, p.torque, p.rot_roll.reshape(-1, 1),
                           p.max_ang_vel.reshape(-1, 1), p.rot_agility.reshape(-1, 1)], dim=1)
    else:
        extra = torch.cat([getattr(p, k).reshape(-1, 1) for k in ("yaw", "pitch", "roll", "old_yaw", "old_pitch",
                                                                   "old_roll", "roll_speed", "old_roll_speed",
                                                                   "yaw_delta")], dim=1)
    dd = torch.cat([dd, extra], dim=1).cpu().numpy()
    ids = p.ids.cpu()
    fp = p.follow_ptr.cpu()
    fid = torch.where(fp >= 0, ids[fp.clamp(min=0)], torch.full_like(fp, -1))
    cols_q = [ids, p.mode.cpu(), p.effect_stamp.cpu(), p.follow_id.cpu(), p.owner_id.cpu(), fid,
              p.massive.cpu().to(I64), p.sensor.cpu().to(I64), p.interactive.cpu().to(I64), p.free.cpu().to(I64),
              p.cloaked.cpu()]
    if isinstance(server, VSS.SpaceServer):
        cols_q += [server.bubbles().cpu(), server.old_bubbles().cpu()]
    qq = torch.stack(cols_q, dim=1).numpy()
    return {(wi, k): (dd[r], qq[r]) for wi, rs in enumerate(index) for k, r in enumerate(rs)}


def _bits(x: np.ndarray) -> np.ndarray:
    return np.ascontiguousarray(x, dtype=np.float64).view(np.int64)


def compare_golden(do: bool, device: str, limit: int | None = None, reverse: bool = False,
                   space: bool = False) -> dict:
    """The batched layer on the fixture's kept worlds (the first `limit`; in reverse order with `reverse`) against the
    fixture: rows, sends and answers that differ (`space`: the partition server on server_space_golden.npz)."""
    g = load_golden(do, space)
    pos = list(range(len(g["kept"])))[:limit] if limit else list(range(len(g["kept"])))
    if reverse:
        pos = pos[::-1]
    kept = [g["kept"][p] for p in pos]
    worlds = [g["worlds"][k] for k in kept]
    for w in worlds:
        w["ticks"] = [[(e[0], e[1], tuple(e[2])) for e in tick] for tick in w["ticks"]]
    server, index = build(worlds, device, space)
    bad_rows, bad_sends, bad_answers, compared = [], [], [], 0
    qcols = Q_COLS + (Q_BUBBLES if space else [])
    T = len(worlds[0]["ticks"])
    extra = list(range(29, 29 + (20 if do else 9)))
    for t in range(T):
        answers = apply_tick(server, worlds, t)
        server.pre_tick(snapshot=True)
        server.evolve()
        server.post_tick()
        sends = sends_of(server)
        got = park_rows(server, worlds, index)
        for i, k in enumerate(kept):
            if json.dumps(sends[i]) != json.dumps(g["sends"][pos[i]][t]):
                bad_sends.append((k, t + 1))
            if json.dumps(answers[i]) != json.dumps(g["answers"][pos[i]][t]):
                bad_answers.append((k, t + 1, answers[i], g["answers"][pos[i]][t]))
            for b in range(len(worlds[i]["balls"])):
                wd, wq = g["states"][(k, t + 1)][b]
                gd, gq = got[(i, b)]
                want_d = np.concatenate([wd[D_FLIGHT], wd[extra]])
                want_q = wq[qcols]
                compared += 1
                if not (np.array_equal(_bits(want_d), _bits(gd)) and np.array_equal(want_q, gq)):
                    cols = [c for c in range(len(want_d)) if _bits(want_d[c:c + 1])[0] != _bits(gd[c:c + 1])[0]]
                    qc = [c for c in range(len(want_q)) if want_q[c] != gq[c]]
                    bad_rows.append((k, t + 1, b, cols, qc))
    server.check()
    return {"rows": bad_rows, "sends": bad_sends, "answers": bad_answers, "compared": compared,
            "worlds": len(worlds), "ticks": T}


del math, struct


def run_streams(worlds: list, device: str, snapshot: bool = False, space: bool = False) -> list:
    """The worlds' streams tick by tick on one park; after each tick (the park's blocks, the calls' queued flags and
    routes, the send flags, the query answers; over the partition also each route's bubble, the narrowcast bubbles,
    the bubble update's lists and the partition's tables) on the host."""
    server, _ = build(worlds, device, space)
    out = []
    for t in range(len(worlds[0]["ticks"])):
        answers = apply_tick(server, worlds, t)
        server.pre_tick(snapshot=snapshot)
        server.evolve()
        server.post_tick()
        p = server.a.park
        b = server.batch
        out.append({"blocks": [x.cpu().clone() for x in p._blocks()],
                    "ok": [s.ok.cpu().clone() for s in b.slots], "route": b.route.cpu().clone(),
                    "has": b.has_character.cpu().clone(), "gets": b.gets_bubble.cpu().clone(),
                    "answers": json.dumps(answers)})
        if space:
            out[-1]["blocks"] += [b.rbubble.cpu().clone(), b.narrow.cpu().clone()] + \
                [v.cpu().clone() for v in b.ad.values()] + [v.cpu().clone() for v in server.space.st.values()]
    server.check()
    return out


def same_runs(a: list, b: list) -> list:
    """The ticks at which two run_streams results differ (blocks by their bits)."""
    bad = []
    for t, (x, y) in enumerate(zip(a, b)):
        for k, (u, v) in enumerate(zip(x["blocks"], y["blocks"])):
            if not torch.equal(u.view(torch.int64) if u.dtype == F64 else u, v.view(torch.int64) if v.dtype == F64
                               else v):
                bad.append((t, "block", k))
        if len(x["ok"]) != len(y["ok"]) or any(not torch.equal(u, v) for u, v in zip(x["ok"], y["ok"])):
            bad.append((t, "ok"))
        for key in ("route", "has", "gets"):
            if not torch.equal(x[key], y[key]):
                bad.append((t, key))
        if x["answers"] != y["answers"]:
            bad.append((t, "answers"))
    return bad
