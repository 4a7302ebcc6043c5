"""Runs tests/scenarios.py's space worlds (`space_cases`: the master's partition, bubbles and sensors) on the drop-in
park and its space layer (VecModesPark + VecSpace, on the CPU's host build or the GPU), and reads back what
tests/golden/make_golden_space.py records of the reference: each ball's flight state with the layer's filing and the
display orientation, and the master's structures, events and query answers, after the load and after every tick."""

from __future__ import annotations

from collections import defaultdict

import numpy as np
import torch

import park_driver as D
import scenarios as S
from exact_ballpark_amd.vec_modes import VecModesPark
from exact_ballpark_amd.vec_space import EVENTS, NO_TOP, OP, VecSpace

SPACE_OPS = {S.KEEPALIVE_OP, S.SENSOR_OP, S.UNSENSOR_OP, S.NOTIFY_OP, S.TRACK_OP, S.INTERACTIVE_OP,
             S.INIT_BUBBLES_OP, S.COPY_BUBBLES_OP}


def build(worlds: list, device: str, block: int = 256):
    """(park, space, per-world {id: slot}, per-world [row]) for the worlds, each world's balls in its add order."""
    rows_of, slot_of, load = [], [], []
    for wi, w in enumerate(worlds):
        rows, slots = [], {}
        for j, b in enumerate(w["balls"]):
            slots[b["id"]] = j
            rows.append(len(load))
            load.append({"id": b["id"], "world": wi, "time": w["start"], "mass": b["mass"], "radius": b["radius"],
                         "max_vel": b["max_vel"], "pos": b["pos"], "vel": b["vel"], "agility": b["agility"],
                         "speed_fraction": b["sf"], "free": b["free"], "massive": b["massive"],
                         "interactive": b["interactive"], "cloaked": 0, "bubble": -1, "moribund": False,
                         "mode": b["mode"]})
        rows_of.append(rows)
        slot_of.append(slots)
    w0 = worlds[0]
    park = VecModesPark(len(load), friction=w0["friction"], tick_interval_ms=w0["tick"], device=device,
                        use_new_orbit=w0["new_orbit"], block=block, display_ypr=True, master=True)
    park.load(load)
    width = max(len(r) for r in rows_of)
    rows = torch.full((len(worlds), width), -1, dtype=torch.int64)
    for wi, r in enumerate(rows_of):
        rows[wi, :len(r)] = torch.tensor(r)
    if worlds[0].get("formations") is not None:      # one formation table a park (the worlds share it)
        park.load_formations(worlds[0]["formations"])
    space = VecSpace(park, rows, seed=[w["seed"] for w in worlds])
    space.load()
    return park, space, slot_of, rows_of


def _apply_round(park, space, items: list, slot_of: list, rows_of: list, reserved=None) -> None:
    """One round of commands (at most one a world): the park's through park_driver, the layer's as one tape; the
    formation reservations on the park, one ball at a time (ReserveFormationSlot's answers into `reserved`)."""
    index = {(wi, bid): rows_of[wi][j] for wi in range(len(rows_of)) for bid, j in slot_of[wi].items()}
    by_op = defaultdict(list)
    tape = [[] for _ in range(space.W)]
    for wi, c in items:
        _, op, a, b, x, y, z, rr, w = c
        j = slot_of[wi].get(a, -1)
        if op in (S.RESERVE_OP, S.FREESLOT_OP):
            row = rows_of[wi][j]
            if op == S.RESERVE_OP:
                reserved[wi].append([int(a), park.reserve_formation_slot(row)])
            else:
                park.free_formation_slot(row, w)
        elif op == S.REMOVE_OP:
            by_op["remove"].append((wi, c))
        elif op == S.FREE_OP and not w and j >= 0 and bool(park.free[rows_of[wi][j]]):
            by_op[op].append((wi, c))                  # SetBallFree(false) of a free ball: InsertInBoxes again
            tape[wi].append(("REFILE", j))
        elif op == S.KEEPALIVE_OP:
            tape[wi].append(("KEEP_ALIVE", j))
        elif op == S.SENSOR_OP:
            tape[wi].append(("ADD_SENSOR", j, int(w) & 2, (rr, x, float(int(w) & 1), 0.0)))
        elif op == S.UNSENSOR_OP:
            tape[wi].append(("REMOVE_SENSOR", j))
        elif op == S.NOTIFY_OP:
            tape[wi].append(("NOTIFY_RANGE", j, 0, (rr,)))
        elif op == S.TRACK_OP:
            tape[wi].append(("TRACK", j, int(b), (rr,)))
        elif op == S.INTERACTIVE_OP:
            tape[wi].append(("SET_INTERACTIVE", j, int(w)))
        elif op == S.INIT_BUBBLES_OP:
            tape[wi].append(("INIT_BUBBLES", -1 if a == 0 else j))
        elif op == S.COPY_BUBBLES_OP:
            tape[wi].append(("COPY_BUBBLES", -1))
        else:
            by_op[op].append((wi, c))
    for op, its in by_op.items():
        if op == "remove":
            mask = torch.zeros(park.n, dtype=torch.bool)
            for wi, c in its:
                r = index.get((wi, c[2]))
                if r is not None:
                    mask[r] = True
            space.remove(mask.to(park.dev))
        else:
            D._apply(park, op, its, index)
    if any(tape):
        space.ops(tape)


def _flight(park, space, wi: int, rows: list) -> tuple:
    """World wi's balls in park now, by ascending id: (doubles [n, 38], ints [n, 15]) in the fixture's columns."""
    ids = park.ids[rows].cpu()
    live = [k for k in range(len(rows)) if int(space.st["sflags"][k, wi]) & 1]
    live.sort(key=lambda k: int(ids[k]))
    r = torch.tensor([rows[k] for k in live], dtype=torch.int64)
    f = park._f[:, r].cpu().numpy()
    q = park._q[:, r].cpu().numpy()
    b = park._b[:, r].cpu().numpy()
    y = park._y[:, r].cpu().numpy()
    d = np.concatenate([f[[0, 1, 2, 3, 4, 5, 12, 13, 14, 15, 16, 17, 6, 7, 8, 9, 10, 11, 18, 21, 22, 23, 24, 25, 19,
                           20]].T, np.zeros((len(live), 3)), y.T], axis=1)
    fp = q[5]
    all_ids = park.ids.cpu().numpy()
    level = space.st["level"][live, wi].cpu().numpy().astype(np.int64)
    key = space.st["key"][:, live, wi].cpu().numpy().T
    mod = space.st["mod"][:, live, wi].cpu().numpy().astype(np.int64).T
    qq = np.stack([q[0], q[1], q[2], q[3], q[4], np.where(fp >= 0, all_ids[np.maximum(fp, 0)], -1), b[2], b[4]],
                  axis=1)
    return d, np.concatenate([qq, level[:, None], key, mod], axis=1)


def structures(space, wi: int, live=None, ids_all=None, bub=None, old=None) -> dict:
    """World wi's master structures (the record's keys but the events and the queries); `live` its live slots by
    ascending id, the other arguments the layer's tables on the host (read here when not given)."""
    if ids_all is None:
        ids_all = space.st["ids"][:, wi:wi + 1].cpu().numpy()
        bub = space.st["bubble"][:, wi:wi + 1].cpu().numpy()
        old = space.st["old_bubble"][:, wi:wi + 1].cpu().numpy()
        fl = space.st["sflags"][:, wi].cpu().numpy()
        live = sorted((j for j in range(space.N) if fl[j] & 1), key=lambda j: ids_all[j, 0])
        col = 0
    else:
        col = wi
    return {
        "ids": [int(ids_all[j, col]) for j in live], "bubble": [int(bub[j, col]) for j in live],
        "old_bubble": [int(old[j, col]) for j in live],
        "tops": sorted(list(k) + [v] for k, v in space.tops(wi).items()), "dicts": space.dicts(wi),
        "inter": sorted([k, sorted(v)] for k, v in space.interactives(wi).items()),
        "keep": sorted([k, sorted(v)] for k, v in space.keep_alives(wi).items()),
        "keep_balls": space.keep_alive_balls(wi), "prox": space.proximity_order(wi),
        "prox_bkt": list(space.proximity_buckets(wi)),
        "sensors": [[int(ids_all[j, col]), space.sensor_list(wi, j), *space.sensor_buckets(wi, j)] for j in live],
        "bubble_id": int(space.st["bubble_id"][wi])}


def master_states(park, space, worlds: list, slot_of=None) -> list:
    """Each world's master record, as make_golden_space.py's master_state writes the reference's."""
    W, Nn = space.W, space.N
    fl = space.st["sflags"].cpu().numpy()
    ids_all = space.st["ids"].cpu().numpy()
    bub = space.st["bubble"].cpu().numpy()
    old = space.st["old_bubble"].cpu().numpy()
    live = [sorted((j for j in range(Nn) if fl[j, wi] & 1), key=lambda j: ids_all[j, wi]) for wi in range(W)]
    # the queries, one launch a query a slot rank
    R = torch.tensor([w["queries"]["range"] for w in worlds], dtype=torch.float64)
    width = max((len(v) for v in live), default=0)
    first = torch.tensor([v[0] if v else -1 for v in live], dtype=torch.int64)
    pos = park._f[0:3].T.cpu()
    vis, rng, rngd, at = ([[] for _ in range(W)] for _ in range(4))
    for k in range(width):
        slot = torch.tensor([v[k] if k < len(v) else -1 for v in live], dtype=torch.int64)
        if k:
            out, _ = space.query("VISIBILITY", first, slot)
            for wi in range(W):
                if k < len(live[wi]):
                    vis[wi].append(int(out[wi, 0]))
        f = torch.zeros(W, 4, dtype=torch.float64)
        f[:, 3] = R
        out, outf = space.query("RANGE_BALL", slot, None, f)
        out, outf = out.cpu(), outf.cpu()
        pts = torch.zeros(W, 4, dtype=torch.float64)
        for wi in range(W):
            if k < len(live[wi]):
                n = int(out[wi, 0])
                rng[wi].append([int(x) for x in out[wi, 1:1 + max(n, 0)]])
                rngd[wi].append([[float(outf[wi, i]).hex(), int(out[wi, 1 + i])] for i in range(max(n, 0))])
                pts[wi, :3] = pos[space.rows[wi, live[wi][k]]]
        out, _ = space.query("BUBBLE_AT", None, None, pts)
        for wi in range(W):
            if k < len(live[wi]):
                v = int(out[wi, 0])
                at[wi].append(None if v == NO_TOP else v)
    pt = torch.zeros(W, 4, dtype=torch.float64)
    pt[:, 3] = R
    for wi, w in enumerate(worlds):
        if live[wi]:
            p0 = pos[space.rows[wi, live[wi][0]]]
            pt[wi, :3] = torch.tensor([float(p0[q]) + w["queries"]["point_offset"][q] for q in range(3)],
                                      dtype=torch.float64)
    pout, _ = space.query("RANGE_POINT", None, None, pt)
    rout, _ = space.query("REMOTE_FOLLOWERS")
    pout, rout = pout.cpu(), rout.cpu()
    out = []
    for wi in range(W):
        out.append(structures(space, wi, live[wi], ids_all, bub, old))
        out[-1].update({
            "events": [list(e) for e in space.events(wi)],
            "vis": vis[wi], "range": rng[wi], "range_dist": rngd[wi],
            "point": [float(pt[wi, q]).hex() for q in range(3)] if live[wi] else [0.0.hex()] * 3,
            "point_ids": [int(x) for x in pout[wi, 1:1 + max(int(pout[wi, 0]), 0)]],
            "at": at[wi], "remote": [int(x) for x in rout[wi, 1:1 + int(rout[wi, 0])]]})
    space.clear_events()
    return out


def run_space(worlds: list, device: str, block: int = 256, graph: bool = False) -> dict:
    """{(world index, step): (doubles, ints, master record)} at step 0 (after the load and the step-0 commands) and
    after every tick; one park and layer for all the worlds (which must share friction, tick and orbit law)."""
    park, space, slot_of, rows_of = build(worlds, device, block)
    per_step = defaultdict(lambda: defaultdict(list))
    for wi, w in enumerate(worlds):
        for c in w["cmds"]:
            per_step[c[0]][wi].append(c)
    out = {}
    g = None
    last = max(w["steps"] for w in worlds)
    forms = worlds[0].get("formations") is not None
    for s in range(0, last + 1):
        todo = per_step.get(s, {})
        reserved = [[] for _ in worlds]
        for j in range(max((len(v) for v in todo.values()), default=0)):
            _apply_round(park, space, [(wi, cs[j]) for wi, cs in todo.items() if j < len(cs)], slot_of, rows_of,
                         reserved)
        if s:
            if graph:
                if g is None:
                    from exact_ballpark_amd.vec_space import SpaceGraph
                    g = SpaceGraph(space)
                g.tick()
            else:
                space.tick()
        masters = master_states(park, space, worlds, slot_of)
        if forms:                                      # each ball's reservation bits, in ascending id
            bits = park.formation_slots.cpu()
            for wi in range(len(worlds)):
                rows = rows_of[wi]
                live = sorted((k for k in range(len(rows)) if int(space.st["sflags"][k, wi]) & 1),
                              key=lambda k: int(park.ids[rows[k]]))
                masters[wi]["fslots"] = [int(bits[rows[k]]) for k in live]
                masters[wi]["reserved"] = reserved[wi]
        for wi, w in enumerate(worlds):
            if s <= w["steps"]:
                d, q = _flight(park, space, wi, rows_of[wi])
                out[(wi, s)] = (d, q, masters[wi])
    space.check()
    return out


def compare(got: dict, golden: dict, kept: list) -> list:
    """The (world, step) records that differ from the fixture's: which part first."""
    bad = []
    for wi in kept:
        w = golden["worlds"][wi]
        want = golden["master"][wi]
        for step in range(0, w["steps"] + 1):
            d, q, m = got[(wi, step)]
            gd, gq, _ = golden["records"][(wi, step)]
            ok = d.shape == gd.shape and D._same_bits(d, gd).all() and (q == gq).all()
            keys = [k for k in want[step] if m.get(k) != want[step][k]]
            if not ok or keys:
                cols = sorted(set(np.nonzero(~D._same_bits(d, gd))[1].tolist())) if d.shape == gd.shape else "shape"
                bad.append((wi, step, cols if not ok else "", keys[:4]))
                break
    return bad


def groups(worlds: list, kept: list) -> dict:
    """The kept worlds grouped by what one park shares: (friction, tick, orbit law)."""
    out = defaultdict(list)
    for wi in kept:
        w = worlds[wi]
        out[(w["friction"], w["tick"], bool(w["new_orbit"]), repr(w.get("formations")))].append(wi)
    return out


__all__ = ["run_space", "compare", "groups", "OP", "EVENTS"]
