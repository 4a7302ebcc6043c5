"""Runs tests/scenarios.py worlds through the drop-in park API (exact_ballpark_amd.vec_modes.VecModesPark) and through
the oracle, and compares them field by field, bit for bit (two NaNs equal, as the reference compares,
docs/exactness.md:33-41 in the reference).

Worlds of one (friction, tick, orbit law) share one park, each world's balls consecutive rows in the order listed (so
ids are not in row order); before each tick the j-th command of every world is applied together, op by op, as the
reference's batch adapter does (tests/modes_batch.py:100-140 in the reference).
"""

from __future__ import annotations

from collections import defaultdict

import numpy as np
import torch

import oracle as O
import scenarios as S
from exact_ballpark_amd.vec_modes import EMPTY_ID, FOLLOW, FORMATION, MISSILE, ORBIT, ModesGraph, VecModesPark

F64, I64 = torch.float64, torch.int64
# the oracle's state columns the park is compared on
D_COLS = list(range(26))                  # pos .. last_collision (the pad left out)
TURN_NAMES = ("rx", "ry", "rz", "rw", "orx", "ory", "orz", "orw", "wx", "wy", "wz", "owx", "owy", "owz", "tx", "ty",
              "tz", "rot_roll", "max_ang_vel", "rot_agility")
Q_COLS = list(range(11))                  # id .. moribund
YPR_NAMES = ("yaw", "pitch", "roll", "old_yaw", "old_pitch", "old_roll", "roll_speed", "old_roll_speed", "yaw_delta")


def _apply(park: VecModesPark, op: int, items: list, index: dict) -> None:
    if op == S.BRING_OP:                            # BringOutTheDeads in each listed world
        park.bring_out_the_deads(worlds=[wi for wi, _ in items])
        return
    n, dev = park.n, park.dev
    mask = torch.zeros(n, dtype=torch.bool)
    target = torch.full((n,), -1, dtype=I64)
    xyz = torch.zeros(n, 3, dtype=F64)
    rr = torch.zeros(n, dtype=F64)
    w = torch.zeros(n, dtype=I64)
    tid = torch.zeros(n, dtype=I64)                 # the target's Destiny id (the missile commands take ids)
    for wi, (_, _, a, b, x, y, z, r, ww) in items:
        row = index.get((wi, "empty" if op == S.MUSHROOM_OP else a))
        if op == S.MUSHROOM_OP:
            b = a
        if row is None:
            continue
        mask[row] = True
        target[row] = index.get((wi, b), -1)
        tid[row] = b
        xyz[row] = torch.tensor((x, y, z), dtype=F64)
        rr[row], w[row] = r, ww
    mask, target, xyz, rr, w, tid = (t.to(dev) for t in (mask, target, xyz, rr, w, tid))
    if op == S.GOTO_OP:
        park.goto_point(mask, xyz)
    elif op == S.DIR_OP:
        park.goto_direction(mask, xyz)
    elif op == S.STOP_OP:
        park.stop(mask)
    elif op == S.FOLLOW_OP:
        park.follow(mask, target, rr)
    elif op == S.ORBIT_OP:
        park.orbit(mask, target, rr)
    elif op == S.WARP_OP:
        park.warp_to(mask, xyz, rr, w)
    elif op == S.WARPIN_OP:
        park.entity_warp_in(mask, xyz, w)
    elif op == S.VEL_OP:
        park.set_velocity(mask, xyz)
    elif op == S.ROT_OP:
        park.set_rotation(mask, torch.cat((xyz, rr[:, None]), dim=1))
    elif op == S.ANGVEL_OP:
        park.set_angular_velocity(mask, xyz)
    elif op == S.MAXANG_OP:
        park.set_max_angular_speed(mask, rr)
    elif op == S.ROTAG_OP:
        park.set_rot_agility(mask, rr)
    elif op == S.YPR_OP:
        park.set_rotation_ypr(mask, xyz)
    elif op == S.LAUNCH_OP:
        park.launch_missile(mask, tid, w, xyz[:, 0].to(I64), xyz[:, 1].to(I64))
    elif op == S.MFOLLOW_OP:
        park.missile_follow(mask, tid, w)
    elif op == S.SETFORM_OP:
        park.set_ball_formation(mask, w)
    elif op == S.FFOLLOW_OP:
        park.formation_follow(mask, tid, w)
    elif op == S.MUSHROOM_OP:                       # into the world's empty row, the owner a Destiny id
        park.add_mushroom(mask, tid, rr, xyz[:, 0])
    elif op == S.FREE_OP:
        park.set_free(mask, w)
    elif op == S.REMOVE_OP:                         # a delay of w ticks (0: at once)
        park.remove(mask, w)
    elif op == S.HARMONIC_OP:                       # harmonic w, corporation b, alliance x, a field if y
        park.set_harmonic(mask, w, tid, xyz[:, 0].to(I64), xyz[:, 1])
    elif op == S.CLOAK_OP:
        park.cloak(mask, w)
    elif op == S.UNCLOAK_OP:
        park.uncloak(mask)
    else:
        raise ValueError(op)


def _snapshot(park: VecModesPark, members: list, rows_of: dict) -> dict:
    """{world index: (d, q, followers)} with balls by ascending id, in the oracle's column order."""
    n = park.n
    f = park._f[:, :n].cpu().numpy()
    q = park._q[:, :n].cpu().numpy()
    b = park._b[:, :n].cpu().numpy()
    ids = q[0]
    mode, fptr = q[1], q[5]
    fol = defaultdict(list)
    for row in range(n):
        if mode[row] in (FOLLOW, ORBIT, MISSILE, FORMATION) and fptr[row] >= 0:
            fol[int(fptr[row])].append(int(ids[row]))
    out = {}
    for wi in members:
        rows = sorted((r for r in rows_of[wi] if ids[r] != EMPTY_ID), key=lambda r: ids[r])
        d = np.stack([np.concatenate((f[0:3, r], f[3:6, r], f[12:15, r], f[15:18, r], f[6:9, r], f[9:12, r],
                                      [f[18, r], f[21, r], f[22, r], f[23, r], f[24, r], f[25, r], f[19, r],
                                       f[20, r]])) for r in rows])
        qq = np.stack([[ids[r], mode[r], q[2, r], q[3, r], q[4, r], ids[fptr[r]] if fptr[r] >= 0 else -1, b[2, r],
                        b[4, r], b[3, r], b[0, r], b[1, r]] for r in rows]).astype(np.int64)
        if park.dynamical_orientation:
            t = np.stack([np.concatenate((f[29:33, r], f[33:37, r], f[37:40, r], f[40:43, r], f[43:46, r],
                                          [f[28, r], f[26, r], f[27, r]])) for r in rows])
            d = np.concatenate((d, t), axis=1)
        lists = [sorted(fol.get(r, [])) for r in rows]
        if park.collisions:                 # mLastC; the filing; the collision lists beside the followers
            cf = park._cf[:, :n].cpu().numpy()
            cq = park._cq[:, :n].cpu().numpy()
            hits = park._hits[:, :n].cpu().numpy()
            d = np.concatenate((d, np.stack([cf[0:3, r] for r in rows])), axis=1)
            qq = np.concatenate((qq, np.stack([cq[0:7, r] for r in rows])), axis=1)
            lists = [(lists[k], [int(x) for x in hits[:cq[7, r], r]]) for k, r in enumerate(rows)]
        if park.display_ypr:                # the display orientation last
            y = park._y[:, :n].cpu().numpy()
            d = np.concatenate((d, np.stack([y[:, r] for r in rows])), axis=1)
        out[wi] = (d, qq, lists)
    return out


def run_park(worlds: list, device: str, block: int = 256, record=None, graph: bool = False,
             general: bool = False, order=None, initial: bool = False, events: dict | None = None,
             ontick: bool = False, make=None) -> dict:
    """{(world index, step): (d, q, followers)} after each recorded tick (`record`: a set of steps, None for all).
    `general`: force the general (untiled) launch; `order`: the world indices in row order (default ascending);
    `initial`: also step 0, the state as loaded. A world flagged `ypr` keeps the display orientation. `make(n,
    friction, tick, device, new_orbit, block, dynamical, collisions, ypr, master)`: the park (default VecModesPark)."""
    groups = defaultdict(list)
    for wi in (order if order is not None else range(len(worlds))):
        w = worlds[wi]
        groups[(w["friction"], w["tick"], bool(w["new_orbit"]), bool(w.get("dynamical", False)),
                bool(w.get("collisions", False)), bool(w.get("ypr", False)),
                repr(w.get("formations")), bool(w.get("master", False)))].append(wi)   # one formation table a park
    out = {}
    for (friction, tick, new_orbit, dynamical, coll, ypr, _, master), members in groups.items():
        minis = {wi: VecModesPark.mini_rows(worlds[wi]["balls"], worlds[wi]["minis"], wi, worlds[wi]["start"])
                 for wi in members if worlds[wi].get("minis")}
        balls = [(wi, b) for wi in members
                 for b in worlds[wi]["balls"] + [{"_row": m, "id": m["id"]} for m in minis.get(wi, ([], []))[0]]
                 + ([None] if any(c[1] == S.MUSHROOM_OP for c in worlds[wi]["cmds"])
                    else [])]     # the world's miniballs; its empty row for a mushroom
        if make is None:
            park = VecModesPark(len(balls), friction=friction, tick_interval_ms=tick, device=device,
                                use_new_orbit=new_orbit, block=block, dynamical_orientation=dynamical,
                                collisions=coll, display_ypr=ypr, master=master)
        else:
            park = make(len(balls), friction, tick, device, new_orbit, block, dynamical, coll, ypr, master)
        park.load([VecModesPark.empty_row(wi, worlds[wi]["start"]) if b is None else b["_row"] if "_row" in b else
                   {"id": b["id"], "world": wi, "time": worlds[wi]["start"], "mass": b["mass"], "radius": b["radius"],
                    "max_vel": b["max_vel"], "pos": b["pos"], "vel": b["vel"], "agility": b["agility"],
                    "speed_fraction": b["sf"], "free": b["free"], "massive": b["massive"],
                    "interactive": b["interactive"], "cloaked": b["cloaked"], "bubble": b["bubble"],
                    "moribund": b["moribund"], "mode": b["mode"], "max_ang_vel": b.get("max_ang_vel", 1.0),
                    "rot_agility": b.get("rot_agility", 0.9), "junk": b.get("junk", False)} for wi, b in balls])
        statics = {wi: worlds[wi].get("statics", []) + minis.get(wi, ([], []))[1] for wi in members
                   if worlds[wi].get("statics") or minis.get(wi, ([], []))[1]}
        if statics:
            park.load_statics(statics)
        if worlds[members[0]].get("formations") is not None:
            park.load_formations(worlds[members[0]]["formations"])
        if general:
            park.tiles = None
            park._state = None
        index = {(wi, b["id"] if b is not None else "empty"): i for i, (wi, b) in enumerate(balls)}
        index.update({(wi, 0): i for i, (wi, b) in enumerate(balls) if b is None})    # ball 0 once added
        rows_of = defaultdict(list)
        for i, (wi, _) in enumerate(balls):
            rows_of[wi].append(i)
        per_step = defaultdict(lambda: defaultdict(list))
        for wi in members:
            for c in worlds[wi]["cmds"]:
                per_step[c[0]][wi].append(c)
        last = max(worlds[wi]["steps"] for wi in members)
        if initial:
            for wi, v in _snapshot(park, members, rows_of).items():
                out[(wi, 0)] = v
        g = ModesGraph(park) if graph else None
        for s in range(1, last + 1):
            todo = per_step.get(s, {})
            for j in range(max((len(v) for v in todo.values()), default=0)):
                by_op = defaultdict(list)
                for wi, cmds in todo.items():
                    if j < len(cmds):
                        by_op[cmds[j][1]].append((wi, cmds[j]))
                for op, items in by_op.items():
                    _apply(park, op, items, index)
            if ontick:                              # Ballpark::OnTick at each world's s-th simulation time
                sim = [worlds[wi]["calls"][min(s, len(worlds[wi]["calls"])) - 1] for wi, _ in balls]
                park.on_tick(torch.tensor(sim, dtype=torch.int64))
            elif g is not None:
                g.tick()
            else:
                park.evolve()
            if events is not None:                  # {world index: [(time, missile, collider, c x, y, z)]}
                for e in park.collision_events():
                    if s <= worlds[e[0]]["steps"]:     # a world past its own last step ticks on unrecorded
                        events.setdefault(e[0], []).append(tuple(e[1:]))
            live = [wi for wi in members if s <= worlds[wi]["steps"] and (record is None or s in record)]
            if live:
                for wi, v in _snapshot(park, live, rows_of).items():
                    out[(wi, s)] = v
    return out


def _run_turning(w: dict):
    """O.run_world for a world with dynamical orientation: the doubles with the turning state appended."""
    o = O.World(w["friction"], w["tick"], w["new_orbit"], w["start"], True)
    for b in w["balls"]:
        o.add(b)
    by_step = defaultdict(list)
    for c in w["cmds"]:
        by_step[c[0]].append(c)
    for step in range(1, w["steps"] + 1):
        for c in by_step[step]:
            o.command(c)
        o.evolve()
        d, q, fol = o.state()
        yield step, np.concatenate((d[:, D_COLS], o.turning()), axis=1), q, fol      # 46 columns


def _run_flags(w: dict, initial: bool = False):
    """O.run_world for a world with collisions or the display orientation: the turning state (with dynamical
    orientation) and mLastC (with collisions) after the doubles, then the display orientation; the filing after the
    ints; with collisions each ball's (followers, collision list)."""
    dyn, coll, ypr = bool(w.get("dynamical")), bool(w.get("collisions")), bool(w.get("ypr"))
    o = O.World(w["friction"], w["tick"], w["new_orbit"], w["start"], dyn, coll, ypr)
    for b in w["balls"]:
        o.add(b)
    for s in w.get("statics", []):
        o.add_static(s)
    for mini in w.get("minis", []):
        o.add_mini(mini)
    if w.get("formations"):
        o.load_formations(w["formations"])
    by_step = defaultdict(list)
    for c in w["cmds"]:
        by_step[c[0]].append(c)
    for step in range(0 if initial else 1, w["steps"] + 1):
        if step:
            for c in by_step[step]:
                o.command(c)
            if "calls" in w:                        # an OnTick world: its step-th simulation time
                o.on_tick(w["calls"][step - 1], w.get("master", False))
            else:
                o.evolve()
        d, q, fol = o.state()
        parts, qs, lists = [d[:, D_COLS]], [q[:, Q_COLS]], fol
        if dyn:
            parts.append(o.turning())
        if coll:
            lc, filing, hits = o.collide()
            parts.append(lc)
            qs.append(filing)
            lists = [(list(f), list(h)) for f, h in zip(fol, hits)]
        if ypr:
            parts.append(o.ypr())
        yield step, np.concatenate(parts, axis=1), np.concatenate(qs, axis=1), lists


def run_oracle(worlds: list, indices, record=None, initial: bool = False) -> dict:
    out = {}
    for wi in indices:
        w = worlds[wi]
        if w.get("collisions") or w.get("ypr") or initial:
            for step, d, q, lists in _run_flags(w, initial):
                if record is None or step in record:
                    out[(wi, step)] = (d, q, lists)
            continue
        runs = _run_turning(w) if w.get("dynamical") else O.run_world(w)
        for step, d, q, fol in runs:
            if record is None or step in record:
                out[(wi, step)] = (d if w.get("dynamical") else d[:, D_COLS], q[:, Q_COLS], fol)
    return out


def _same_bits(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    a, b = a.copy(), b.copy()
    both = np.isnan(a) & np.isnan(b)
    a[both] = 0.0
    b[both] = 0.0
    return a.view(np.uint64) == b.view(np.uint64)


def compare(got: dict, want: dict, worlds: list, limit: int = 5) -> list:
    """The differences (at most `limit` described), as strings; every key of `want` must be in `got`."""
    bad = []
    for key in sorted(want):
        if key not in got:
            bad.append(f"{key}: missing")
            continue
        gd, gq, gf = got[key]
        wd, wq, wf = want[key]
        if gd.shape != wd.shape:
            bad.append(f"{key}: {gd.shape} balls vs {wd.shape}")
            continue
        same_d = _same_bits(gd, wd)
        same_q = gq == wq
        if same_d.all() and same_q.all() and gf == wf:
            continue
        if len(bad) < limit:
            w = worlds[key[0]]
            names_d = [O.D_NAMES[c] for c in D_COLS]
            names_d += list(TURN_NAMES) if w.get("dynamical") else []
            names_d += ["cx", "cy", "cz"] if w.get("collisions") else []
            names_d += list(YPR_NAMES) if w.get("ypr") else []
            names_q = [O.Q_NAMES[c] for c in Q_COLS] + ["level", "kx", "ky", "kz", "sx", "sy", "sz"]
            diffs = [f"{names_d[c]}[{k}] {gd[k, c]!r} vs {wd[k, c]!r}" for k, c in zip(*np.nonzero(~same_d))][:6]
            diffs += [f"{names_q[c]}[{k}] {gq[k, c]} vs {wq[k, c]}" for k, c in zip(*np.nonzero(~same_q))][:6]
            if gf != wf:
                diffs.append(f"lists {gf} vs {wf}")
            bad.append(f"world {key[0]} ({worlds[key[0]]['kind']}) step {key[1]}: " + "; ".join(diffs))
        else:
            bad.append(f"{key}")
            if len(bad) > 40:
                bad.append("...")
                break
    return bad
