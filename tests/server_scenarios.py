"""Seeded worlds for the server layer's parity tests: Destiny's Actions calls between ticks, the Ticker's pre-tick, the
park's Evolve and the batcher's send, world by world.

A world is one bubble on a master park: 2 to 6 ships within 12 km of a point near the origin (the master files them
all in one bubble), each with a pilot or not. Ship `anchor` is interactive with pilot 0 and never loses its
interactivity (a bubble without an interactive ball is dead, and the master does not evolve its balls); one `special`
ship is the one that can become massive (CloakBall, UncloakBall, SetBallMassive) or travel far (WarpTo,
EntityWarpIn), so no two massive ships ever meet.

`ticks[t]` is the list of calls made before tick t + 1, in order: ("act", method, args) an Actions method that queues
(or, for undo_pending_cloak, prunes) and ("ask", query, ball) one of the pending queries whose answer is recorded. The
kinds cover what the layer's rules turn on (the reference's tests exercise the same, tests/server_cases.py:98-214
there): the same call twice in a row (the second dropped), the same call with another between (both kept), a Stop
after a command, a SetBallPosition local and not, a cloak and an uncloak in either order, a cloak undone, calls at a
ball not in the park, refused values, follows and orbits of the special ship around its cloaks, warps and warp-ins
with followers, trolls and SetBallFree, and with dynamical orientation the turning setters (SetMaxAngularVelocity
among them, which the park never applies).
"""

from __future__ import annotations

import math
import random

MISSING = 777000                 # a ball id no world has

_MOVES = ("stop", "go_to_point", "go_to_direction", "follow_ball", "orbit", "warp_to", "entity_warp_in",
          "set_ball_velocity", "set_ball_position", "set_ball_massive", "set_ball_free", "set_ball_troll")


def _u(r: random.Random, lo: float, hi: float) -> float:
    return float(r.uniform(lo, hi))


def _world(r: random.Random, wid: int, do: bool) -> dict:
    n = r.randint(2, 6)
    ids = r.sample(range(11, 9000), n)
    centre = [_u(r, -2e4, 2e4) for _ in range(3)]
    balls, pilots = [], {}
    for k, bid in enumerate(ids):
        inter = k == 0 or r.random() < 0.55
        balls.append({"id": bid, "mass": 10 ** r.uniform(5.5, 9.0), "radius": r.choice([_u(r, 2.0, 40.0),
                                                                                          _u(r, 40.0, 400.0)]),
                      "max_vel": r.choice([_u(r, 40.0, 600.0), _u(r, 2.0, 120.0)]), "agility": _u(r, 0.08, 2.5),
                      "sf": r.choice([1.0, 1.0, _u(r, 0.05, 1.0), 0.0]),
                      "pos": [centre[i] + _u(r, -1.2e4, 1.2e4) for i in range(3)],
                      "vel": r.choice([[0.0, 0.0, 0.0], [_u(r, -120.0, 120.0) for _ in range(3)]]),
                      "interactive": inter, "massive": False,
                      "max_ang_vel": r.choice([1.0, _u(r, 0.05, 2.5)]),
                      "rot_agility": r.choice([0.9, _u(r, 0.05, 2.5)])})
        if inter and (k == 0 or r.random() < 0.75):
            pilots[bid] = len(pilots)
    order = list(range(1, n))
    r.shuffle(order)
    order.insert(r.randint(0, len(order)), 0)             # AddBall's order is the bubble's order
    return {"id": wid, "do": do, "balls": [balls[k] for k in order], "anchor": ids[0],
            "special": ids[1], "pilots": pilots, "centre": centre, "ticks": []}


def _xyz(r: random.Random, lo: float, hi: float) -> list:
    return [_u(r, lo, hi) for _ in range(3)]


def _call(r: random.Random, w: dict, do: bool) -> tuple:
    ids = [b["id"] for b in w["balls"]]
    sp, anchor = w["special"], w["anchor"]
    s = r.choice(ids + ([MISSING] if r.random() < 0.04 else []))
    other = r.choice([i for i in ids if i != s] or ids)
    bad = r.random() < 0.1
    c = w["centre"]
    k = r.random()
    if k < 0.10:
        return "go_to_point", (s, *[c[i] + _u(r, -2.5e4, 2.5e4) for i in range(3)])
    if k < 0.19:
        return "go_to_direction", (s, *([0.0, float("nan"), 1.0] if bad else _xyz(r, -1.0, 1.0)))
    if k < 0.26:
        return "set_speed_fraction", (s, r.choice([-0.25, 1.75, float("inf")]) if bad
                                      else r.choice([0.0, 1.0, _u(r, 0.0, 1.0)]))
    if k < 0.31:
        return "stop", (s,)
    if k < 0.38:
        tgt = r.choice([other, sp, MISSING, s])
        return "follow_ball", (s, tgt, _u(r, 30.0, 4000.0))
    if k < 0.44:
        return "orbit", (s, r.choice([other, sp]), math.inf if bad else _u(r, 200.0, 6000.0))
    if k < 0.48:
        d = r.choice([1.0, -1.0]) * _u(r, 1.2e5, 5e5)
        if r.random() < 0.2:
            d = _u(r, 2e3, 6e4)                             # under 100 km: a GotoPoint
        return "warp_to", (sp, c[0] + d, c[1] + _u(r, -1e4, 1e4), c[2] + _u(r, -1e4, 1e4),
                           r.choice([0.0, 15000.0, _u(r, 0.0, 4e4)]), r.choice([1, 4, 12]))
    if k < 0.51:
        return "entity_warp_in", (sp, c[0] + r.choice([1.0, -1.0]) * _u(r, 1.5e5, 5e5), c[1] + _u(r, -1e4, 1e4),
                                  c[2], r.choice([1, 6]))
    if k < 0.54:
        return "set_max_speed", (s, -3.0 if bad else _u(r, 5.0, 500.0))
    if k < 0.57:
        return "set_ball_mass", (s, 0.0 if bad else 10 ** r.uniform(5.5, 9.0))
    if k < 0.60:
        return "set_ball_agility", (s, -0.5 if bad else _u(r, 0.08, 2.5))
    if k < 0.63:
        return "set_ball_radius", (s, -1.0 if bad else _u(r, 2.0, 400.0))
    if k < 0.67:
        return "set_ball_velocity", (s, *_xyz(r, -150.0, 150.0))
    if k < 0.70 and s != anchor:
        return "set_ball_interactive", (s, r.choice([0, 1]))
    if k < 0.74:
        return "set_ball_position", (s, *[c[i] + _u(r, -1.2e4, 1.2e4) for i in range(3)], r.random() < 0.5)
    if k < 0.77:
        return "set_ball_massive", (sp, r.choice([0, 1]))
    if k < 0.83:
        return "cloak_ball", (sp, r.choice([1, 1, 2, 3, 0, 200]), _u(r, 50.0, 4000.0))
    if k < 0.88:
        return "uncloak_ball", (sp,)
    if k < 0.90 and s not in (anchor, sp):
        return "set_ball_troll", (s, r.choice([0, 1, 2, 4]))
    if k < 0.92 and s != anchor:
        return "set_ball_free", (s, r.choice([0, 1]))
    if do:
        j = r.random()
        if j < 0.3:
            return "set_ball_angular_velocity", (s, *_xyz(r, -1.2, 1.2))
        if j < 0.5:
            q = [0.0, 0.0, 0.0, 0.0] if bad else [r.gauss(0.0, 1.0) for _ in range(4)]
            return "set_ball_rotation", (s, *q)
        if j < 0.7:
            return "set_max_angular_speed", (s, -1.0 if bad else _u(r, 0.05, 2.5))
        if j < 0.85:
            return "set_ball_angular_agility", (s, 0.0 if bad else _u(r, 0.05, 2.5))
        return "set_max_angular_velocity", (s, *_xyz(r, 0.1, 2.0))
    return "go_to_direction", (s, *_xyz(r, -1.0, 1.0))


def _same(a: tuple) -> tuple:
    return (a[0], tuple(a[1]))


def _has_nan(args) -> bool:
    return any(isinstance(v, float) and v != v for v in args)


def _tick(r: random.Random, w: dict, do: bool) -> list:
    out = []
    sp = w["special"]
    frozen = w.get("warped", False)                   # the special ship once warped: nothing that moves it
    for _ in range(r.choice([0, 1, 1, 2, 2, 3, 4, 5])):
        name, args = _call(r, w, do)
        if args[0] == sp and name in _MOVES and frozen:
            name, args = "set_max_speed", (sp, _u(r, 40.0, 400.0))
        if args[0] == sp and name in ("warp_to", "entity_warp_in"):
            w["warped"] = True
        out.append(("act", name, args))
        x = r.random()
        if x < 0.10 and not _has_nan(args):
            out.append(("act", name, tuple(args)))              # twice in a row: the second dropped
        elif x < 0.16 and not (args[0] == sp and (frozen or w.get("warped"))):
            out.append(("act", "stop", (args[0],)))             # a Stop after a command
        elif name == "set_ball_position" and x < 0.45:
            out.append(("act", name, (*args[:4], not args[4])))    # the other locality: not a duplicate
        elif x < 0.22 and len(out) >= 3 and out[-3][0] == "act" and not _has_nan(out[-3][2]):
            again = out[-3]
            if not (again[2][0] == sp and again[1] in _MOVES and (frozen or w.get("warped"))):
                out.append(again)                               # the same call with another between: kept
        if name == "cloak_ball" and r.random() < 0.25:
            out.append(("act", "undo_pending_cloak", (sp,)))
    if r.random() < 0.5:
        s = r.choice([b["id"] for b in w["balls"]] + [MISSING])
        for q in r.sample(("has_pending_cloak", "has_pending_uncloak", "get_pending_cloak_mode",
                           "get_pending_interactive_state", "get_pending_speed_fraction", "has_pending_set_free",
                           "has_pending_set_not_free"), 3):
            out.append(("ask", q, (r.choice([sp, s]),)))
    return out


def server_cases(seed: int, n: int = 60, do: bool = False, ticks: int = 12) -> list:
    """n worlds (with dynamical orientation if `do`), each with `ticks` ticks of calls."""
    r = random.Random(seed * 7919 + (1 if do else 0))
    out = []
    for wid in range(n):
        w = _world(r, wid, do)
        w["ticks"] = [_tick(r, w, do) for _ in range(ticks)]
        w.pop("warped", None)
        out.append(w)
    return out


# ======================================================================== worlds of several bubbles (the partition)

TOP = 65536.0 * 480.0 * 0.25        # the partition's top box width, metres: a bubble never spans two tops' balls apart


def _space_world(r: random.Random, wid: int) -> dict:
    """2 to 6 ships around one to three anchors: near the origin, just either side of a top box's face, or far off
    (other top boxes: other bubbles). Ship `anchor` is interactive with pilot 0; ship `special` is the one that can
    cloak, turn massive or warp."""
    n = r.randint(2, 6)
    ids = r.sample(range(11, 9000), n)
    y, z = _u(r, -3e3, 3e3) + 0.3 * TOP, _u(r, -3e3, 3e3) - 0.2 * TOP
    pool = [(0.5 * TOP + _u(r, -2e4, 2e4), y, z), (TOP - _u(r, 300.0, 5000.0), y, z),
            (TOP + _u(r, 300.0, 5000.0), y, z), (-2.5 * TOP, 3.5 * TOP, 0.5 * TOP)]
    anchors = r.sample(pool, r.randint(1, 3))
    balls, pilots = [], {}
    for k, bid in enumerate(ids):
        a = anchors[0] if k == 0 else r.choice(anchors)
        inter = k == 0 or r.random() < 0.55
        balls.append({"id": bid, "mass": 10 ** r.uniform(5.5, 9.0), "radius": r.choice([_u(r, 2.0, 40.0),
                                                                                          _u(r, 40.0, 400.0)]),
                      "max_vel": r.choice([_u(r, 40.0, 600.0), _u(r, 2.0, 120.0)]), "agility": _u(r, 0.08, 2.5),
                      "sf": r.choice([1.0, 1.0, _u(r, 0.05, 1.0), 0.0]),
                      "pos": [a[i] + _u(r, -6e3, 6e3) for i in range(3)],
                      "vel": r.choice([[0.0, 0.0, 0.0], [_u(r, -120.0, 120.0) for _ in range(3)]]),
                      "interactive": inter, "massive": False, "max_ang_vel": 1.0, "rot_agility": 0.9})
        if inter and (k == 0 or r.random() < 0.75):
            pilots[bid] = len(pilots)
    order = list(range(1, n))
    r.shuffle(order)
    order.insert(r.randint(0, len(order)), 0)             # AddBall's order is the bubble's order
    return {"id": wid, "do": False, "balls": [balls[k] for k in order], "anchor": ids[0], "special": ids[1],
            "pilots": pilots, "anchors": anchors, "centre": list(anchors[0]), "ticks": []}


def _space_call(r: random.Random, w: dict) -> tuple:
    ids = [b["id"] for b in w["balls"]]
    sp, anchor = w["special"], w["anchor"]
    s = r.choice(ids + ([MISSING] if r.random() < 0.03 else []))
    other = r.choice([i for i in ids if i != s] or ids)
    bad = r.random() < 0.1
    a = r.choice(w["anchors"])
    near = lambda c, d: [c[i] + _u(r, -d, d) for i in range(3)]      # noqa: E731
    k = r.random()
    if k < 0.12:
        return "go_to_point", (s, *near(a, 1.5e4))
    if k < 0.19:
        return "go_to_direction", (s, *([0.0, float("nan"), 1.0] if bad else _xyz(r, -1.0, 1.0)))
    if k < 0.22:
        return "stop", (s,)
    if k < 0.26:
        return "set_speed_fraction", (s, r.choice([-0.25, 1.75]) if bad else r.choice([0.0, 1.0, _u(r, 0.0, 1.0)]))
    if k < 0.29:
        return "set_ball_velocity", (s, *_xyz(r, -150.0, 150.0))
    if k < 0.39:                                          # a jump: to an anchor, across a face, or to a far top
        j = r.random()
        if j < 0.45:
            dst = near(a, 8e3)
        elif j < 0.75:
            dst = [TOP + r.choice([-1.0, 1.0]) * _u(r, 100.0, 4000.0), a[1] + _u(r, -3e3, 3e3), a[2]]
        else:
            dst = near([(r.randint(-5, 5) + 0.5) * TOP, (r.randint(-5, 5) + 0.5) * TOP, 0.5 * TOP], 3e4)
        return "set_ball_position", (s, *dst, r.random() < 0.15)
    if k < 0.43:
        return "set_ball_radius", (s, -1.0 if bad else _u(r, 2.0, 400.0))
    if k < 0.47:
        return "set_max_speed", (s, -3.0 if bad else _u(r, 5.0, 600.0))
    if k < 0.50:
        return "set_ball_mass", (s, 0.0 if bad else 10 ** r.uniform(5.5, 9.0))
    if k < 0.52:
        return "set_ball_agility", (s, -0.5 if bad else _u(r, 0.08, 2.5))
    if k < 0.60:
        return "set_ball_interactive", (s, r.choice([0, 1]))
    if k < 0.66:
        return "cloak_ball", (sp, r.choice([1, 1, 2, 3, 0, 200]), _u(r, 50.0, 6000.0))
    if k < 0.71:
        return "uncloak_ball", (sp,)
    if k < 0.80:
        return r.choice(["follow_ball", "orbit"]), (s, r.choice([other, sp, MISSING]), _u(r, 100.0, 4000.0))
    if k < 0.83:
        d = 10 ** r.uniform(8.0, 10.0)
        u = [r.gauss(0.0, 1.0) for _ in range(3)]
        nn = math.sqrt(sum(x * x for x in u)) or 1.0
        c = next(b["pos"] for b in w["balls"] if b["id"] == sp)
        return "warp_to", (sp, *[c[i] + d * u[i] / nn for i in range(3)], r.choice([0.0, 2e4]), r.randint(1, 12))
    if k < 0.85:
        return "entity_warp_in", (sp, *near([(r.randint(-5, 5) + 0.5) * TOP, 0.5 * TOP, 0.5 * TOP], 1e4),
                                  r.choice([1, 6]))
    if k < 0.88:
        return "set_ball_massive", (sp, r.choice([0, 1]))
    if k < 0.91 and s != anchor:
        return "set_ball_free", (s, r.choice([0, 1]))
    return "go_to_direction", (s, *_xyz(r, -1.0, 1.0))


def _space_tick(r: random.Random, w: dict) -> list:
    out = []
    sp = w["special"]
    for _ in range(r.choice([0, 1, 1, 2, 2, 3, 4])):
        frozen = w.get("warped", False)                   # the special ship once warped: nothing that moves it
        name, args = _space_call(r, w)
        if args[0] == sp and name in _MOVES and frozen:
            name, args = "set_max_speed", (sp, _u(r, 40.0, 400.0))
        if name == "warp_to":                             # aligned at 90 % of its top speed: it enters warp at once
            b = next(x for x in w["balls"] if x["id"] == sp)
            mv = r.choice([0.5, b["max_vel"]])
            dv = [args[1 + i] - b["pos"][i] for i in range(3)]
            nn = math.sqrt(sum(x * x for x in dv))
            out.append(("act", "set_max_speed", (sp, mv)))
            out.append(("act", "set_ball_velocity", (sp, *[dv[i] / nn * 0.9 * mv for i in range(3)])))
        if args[0] == sp and name in ("warp_to", "entity_warp_in"):
            w["warped"] = True
        out.append(("act", name, args))
        x = r.random()
        if x < 0.10 and not _has_nan(args):
            out.append(("act", name, tuple(args)))              # twice in a row: the second dropped
        elif x < 0.16 and not (args[0] == sp and w.get("warped")):
            out.append(("act", "stop", (args[0],)))             # a Stop after a command
        elif name == "set_ball_position" and x < 0.35:
            out.append(("act", name, (*args[:4], not args[4])))    # the other locality: not a duplicate
        elif name == "set_ball_position" and x < 0.55 and args[0] != MISSING:
            home = next(b["pos"] for b in w["balls"] if b["id"] == args[0])
            out.append(("act", name, (args[0], *[c + _u(r, -2e3, 2e3) for c in home], False)))   # and back
        if name == "cloak_ball" and r.random() < 0.2:
            out.append(("act", "undo_pending_cloak", (sp,)))
        if name in ("follow_ball", "orbit") and args[1] != MISSING and r.random() < 0.35 and \
                not (args[1] == sp and w.get("warped")):         # the target jumps away: a remote follower
            far = [(r.randint(-5, 5) + 0.5) * TOP + _u(r, -3e4, 3e4), 2.5 * TOP, 0.5 * TOP]
            out.append(("act", "set_ball_position", (args[1], *far, False)))
    if r.random() < 0.3:
        s = r.choice([b["id"] for b in w["balls"]])
        for q in r.sample(("has_pending_cloak", "has_pending_uncloak", "get_pending_cloak_mode",
                           "get_pending_interactive_state", "get_pending_speed_fraction"), 2):
            out.append(("ask", q, (r.choice([sp, s]),)))
    return out


def space_server_cases(seed: int, n: int = 60, ticks: int = 14) -> list:
    """n worlds of several bubbles on the master's partition, each with `ticks` ticks of calls."""
    r = random.Random(seed * 104729 + 3)
    out = []
    for wid in range(n):
        w = _space_world(r, wid)
        lonely = wid % 8 == 5 and ticks > 3
        if lonely:                    # one interactive ship, which leaves its bubble and comes back within a tick
            for b in w["balls"]:
                b["interactive"] = b["id"] == w["anchor"]
            w["pilots"] = {w["anchor"]: 0}
        w["ticks"] = [_space_tick(r, w) for _ in range(ticks)]
        if lonely:
            home = next(b["pos"] for b in w["balls"] if b["id"] == w["anchor"])
            w["ticks"][3][:0] = [("act", "set_ball_position", (w["anchor"], 3.5 * TOP, -1.5 * TOP, 0.5 * TOP, False)),
                                 ("act", "set_ball_position", (w["anchor"], *[c + 100.0 for c in home], False))]
        w.pop("warped", None)
        out.append(w)
    return out


def canon(x):
    """A value as the server tests compare it, JSON-able: each item tagged with its type (bool, int, float by its
    bits, bytes in hex, str, None, list, tuple, dict in its order, and a marshalled payload by its value)."""
    import struct
    if isinstance(x, bool):
        return ["b", x]
    if isinstance(x, int):
        return ["i", int(x)]
    if isinstance(x, float):
        return ["f", struct.unpack("<q", struct.pack("<d", x))[0]]
    if isinstance(x, (bytes, bytearray)):
        return ["B", bytes(x).hex()]
    if isinstance(x, str):
        return ["s", x]
    if x is None:
        return ["n"]
    if isinstance(x, list):
        return ["l", [canon(v) for v in x]]
    if isinstance(x, tuple):
        return ["t", [canon(v) for v in x]]
    if isinstance(x, dict):
        return ["d", [[canon(k), canon(v)] for k, v in x.items()]]
    if type(x).__name__ == "Marshalled":
        return ["M", canon(x.value)]
    raise TypeError(f"canon: {type(x).__name__}")
