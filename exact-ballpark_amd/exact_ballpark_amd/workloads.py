"""Seeded synthetic workloads, generated with numpy in one pass so a park of 65,536 worlds is built in about a second.

`flight_park(seed, worlds, ships)`: the flight-modes park. Each world has `ships` ships with ids 1..ships; per ship,
position uniform in a +/-50 km cube, at rest, mass U(1e6, 1e8) kg, agility U(0.1, 5), maximum speed U(100, 3000) m/s,
radius U(20, 500) m, speed fraction 1, massive, interactive with probability 1/2. At tick 1 each ship gets a mode
uniform over {stop, goto, follow, orbit, warp}: GotoPoint to a point uniform in the ball of radius 1e10 m (1e7 km)
around the origin, FollowBall or Orbit of another ship of its world chosen uniformly at a range U(500, 20000) m, or
WarpTo a point as GotoPoint's with the thunker's defaults (minimum range 20 km, warp factor 20).
"""

from __future__ import annotations

import numpy as np
import torch

GOTO_OP, DIR_OP, STOP_OP, FOLLOW_OP, ORBIT_OP, WARP_OP = 1, 2, 3, 4, 5, 6
YPR_OP = 13                                        # SetBallRotation(RotationQuaternion(yaw, pitch, roll))
LAUNCH_OP = 14                                     # LaunchMissile(missile, target id, owner id, aimed, massive)


def _ball_points(rng: np.random.Generator, n: int, radius: float) -> np.ndarray:
    u = rng.normal(size=(n, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    return u * (radius * rng.uniform(0.0, 1.0, n) ** (1.0 / 3.0))[:, None]


def flight_park(seed: int, worlds: int, ships: int = 4) -> dict:
    """Arrays for worlds x ships rows (world-major) and each row's tick-1 command."""
    rng = np.random.default_rng(seed)
    n = worlds * ships
    world = np.repeat(np.arange(worlds, dtype=np.int64), ships)
    ids = np.tile(np.arange(1, ships + 1, dtype=np.int64), worlds)
    rows = {
        "id": ids, "world": world, "time": np.zeros(n, dtype=np.int64),
        "mass": rng.uniform(1e6, 1e8, n), "radius": rng.uniform(20.0, 500.0, n),
        "max_vel": rng.uniform(100.0, 3000.0, n), "agility": rng.uniform(0.1, 5.0, n),
        "speed_fraction": np.ones(n), "pos": rng.uniform(-5e4, 5e4, (n, 3)), "vel": np.zeros((n, 3)),
        "free": np.ones(n, dtype=bool), "massive": np.ones(n, dtype=bool),
        "interactive": rng.uniform(0.0, 1.0, n) < 0.5, "moribund": np.zeros(n, dtype=bool),
        "cloaked": np.zeros(n, dtype=np.int64), "bubble": np.zeros(n, dtype=np.int64),
        "mode": np.full(n, 2, dtype=np.int64),
    }
    kind = rng.integers(0, 5, n)                       # 0 stop, 1 goto, 2 follow, 3 orbit, 4 warp
    op = np.select([kind == 1, kind == 2, kind == 3, kind == 4], [GOTO_OP, FOLLOW_OP, ORBIT_OP, WARP_OP], 0)
    point = _ball_points(rng, n, 1e10)
    other = rng.integers(0, ships - 1, n) if ships > 1 else np.zeros(n, dtype=np.int64)
    own = ids - 1
    other = np.where(other >= own, other + 1, other)   # another ship of the world, uniform
    target = world * ships + other
    rng_ = rng.uniform(500.0, 20000.0, n)
    if ships == 1:
        op = np.where((op == FOLLOW_OP) | (op == ORBIT_OP), 0, op)
    cmds = {"op": op.astype(np.int64), "point": point, "target": target.astype(np.int64), "range": rng_,
            "min_range": np.full(n, 20000.0), "warp_factor": np.full(n, 20, dtype=np.int64)}
    return {"n": n, "worlds": worlds, "ships": ships, "rows": rows, "cmds": cmds}


def turning_park(seed: int, worlds: int, ships: int = 4) -> dict:
    """The turning park (dynamical orientation): each world's ships in a 2 km cluster around a centre uniform in a
    +/-1e6 m cube, at rest, mass U(1e6, 1e8) kg, agility U(0.1, 5), maximum speed U(100, 3000) m/s, radius U(20, 500)
    m, maximum angular speed U(0.2, 2) rad/s, rotational agility U(0.3, 3), not massive (no collisions here), each
    rotated by RotationQuaternion(yaw, pitch, roll) with the three angles U(-pi, pi). At tick 1 each ship goes to a
    point of its world's cluster (GotoPoint; 70%), orbits (15%) or follows (15%) another ship of its world at
    U(200, 2000) m: every ship turns toward its thrust, all the time."""
    rng = np.random.default_rng(seed)
    n = worlds * ships
    world = np.repeat(np.arange(worlds, dtype=np.int64), ships)
    ids = np.tile(np.arange(1, ships + 1, dtype=np.int64), worlds)
    centre = np.repeat(rng.uniform(-1e6, 1e6, (worlds, 3)), ships, axis=0)
    rows = {
        "id": ids, "world": world, "time": np.zeros(n, dtype=np.int64),
        "mass": rng.uniform(1e6, 1e8, n), "radius": rng.uniform(20.0, 500.0, n),
        "max_vel": rng.uniform(100.0, 3000.0, n), "agility": rng.uniform(0.1, 5.0, n),
        "speed_fraction": np.ones(n), "pos": centre + rng.uniform(-1000.0, 1000.0, (n, 3)), "vel": np.zeros((n, 3)),
        "free": np.ones(n, dtype=bool), "massive": np.zeros(n, dtype=bool),
        "interactive": rng.uniform(0.0, 1.0, n) < 0.5, "moribund": np.zeros(n, dtype=bool),
        "cloaked": np.zeros(n, dtype=np.int64), "bubble": np.zeros(n, dtype=np.int64),
        "mode": np.full(n, 2, dtype=np.int64), "max_ang_vel": rng.uniform(0.2, 2.0, n),
        "rot_agility": rng.uniform(0.3, 3.0, n),
    }
    ypr = rng.uniform(-np.pi, np.pi, (n, 3))
    u = rng.uniform(0.0, 1.0, n)
    op = np.where(u < 0.7, GOTO_OP, np.where(u < 0.85, ORBIT_OP, FOLLOW_OP))
    own = ids - 1
    other = rng.integers(0, ships - 1, n)
    other = np.where(other >= own, other + 1, other)
    cmds = {"op": op.astype(np.int64), "point": centre + rng.uniform(-1000.0, 1000.0, (n, 3)),
            "target": (world * ships + other).astype(np.int64), "range": rng.uniform(200.0, 2000.0, n),
            "min_range": np.full(n, 20000.0), "warp_factor": np.full(n, 20, dtype=np.int64)}
    return {"n": n, "worlds": worlds, "ships": ships, "rows": rows, "cmds": cmds, "ypr": ypr, "dynamical": True}


def _frames(rng: np.random.Generator, n: int) -> np.ndarray:
    """[n, 3, 3]: n random orthonormal frames (rows the axes), from the QR of Gaussian matrices."""
    q, r = np.linalg.qr(rng.normal(size=(n, 3, 3)))
    return q * np.sign(np.diagonal(r, axis1=1, axis2=2))[:, None, :]


def static_collidables(rng: np.random.Generator, centre: np.ndarray, per_world: int) -> dict:
    """`per_world` static collidables a world around each world's centre ([W, 3]), capsules and oriented boxes in
    turn: a capsule's axis of length U(200, 3000) m, radius U(30, 400) m; a box's half-extents U(30, 800) m; each in a
    random orientation, its centre within 2 km of the world's. Ids 1001.., parent ids 0 or 1 (so a box's walk orders
    them by parent, then id). As load_statics_arrays takes them."""
    W = len(centre)
    S = W * per_world
    world = np.repeat(np.arange(W, dtype=np.int64), per_world)
    j = np.tile(np.arange(per_world, dtype=np.int64), W)
    kind = (j % 2).astype(np.int64)
    mid = np.repeat(centre, per_world, axis=0) + rng.uniform(-2000.0, 2000.0, (S, 3))
    fr = _frames(rng, S)
    length = rng.uniform(200.0, 3000.0, S)
    radius = rng.uniform(30.0, 400.0, S)
    half = rng.uniform(30.0, 800.0, (S, 3))
    geom = np.zeros((S, 12))
    a = mid - 0.5 * length[:, None] * fr[:, 0]
    b = mid + 0.5 * length[:, None] * fr[:, 0]
    cap = kind == 0
    geom[cap, 0:3], geom[cap, 3:6], geom[cap, 6] = a[cap], b[cap], radius[cap]
    edges = 2.0 * half[:, :, None] * fr                           # [S, 3 edges, 3]
    corner = mid - 0.5 * edges.sum(axis=1)
    box = ~cap
    geom[box, 0:3] = corner[box]
    geom[box, 3:12] = edges[box].reshape(-1, 9)
    return {"world": world, "kind": kind, "id": 1001 + j, "parent": rng.integers(0, 2, S).astype(np.int64),
            "geom": geom}


def collide_park(seed: int, worlds: int, ships: int = 4, spheres: int = 4, dynamical: bool = False,
                 statics: int = 0) -> dict:
    """The turning-and-collisions park's bodies (the gradient method): each world's ships in a 2 km cluster around a
    centre uniform in a +/-1e6 m cube, at rest, mass U(1e6, 1e8) kg, agility U(0.1, 5), maximum speed U(100, 3000)
    m/s, radius U(20, 500) m, massive; and `spheres` fixed, massive spheres of radius U(100, 5000) m in a 4 km cluster
    around the same centre. At tick 1 each ship goes to a point of its world's cluster (GotoPoint), so ships cross,
    meet each other and hit or start inside the spheres. With `dynamical`, the turning values and rotations of
    turning_park (maximum angular speed U(0.2, 2) rad/s, rotational agility U(0.3, 3), RotationQuaternion(yaw, pitch,
    roll) of angles U(-pi, pi)). With `statics`, that many static collidables a world (static_collidables)."""
    rng = np.random.default_rng(seed)
    per = ships + spheres
    n = worlds * per
    world = np.repeat(np.arange(worlds, dtype=np.int64), per)
    ids = np.tile(np.arange(1, per + 1, dtype=np.int64), worlds)
    ship = np.tile(np.arange(per) < ships, worlds)
    centre = np.repeat(rng.uniform(-1e6, 1e6, (worlds, 3)), per, axis=0)
    rows = {
        "id": ids, "world": world, "time": np.zeros(n, dtype=np.int64),
        "mass": rng.uniform(1e6, 1e8, n), "radius": np.where(ship, rng.uniform(20.0, 500.0, n),
                                                             rng.uniform(100.0, 5000.0, n)),
        "max_vel": rng.uniform(100.0, 3000.0, n), "agility": rng.uniform(0.1, 5.0, n),
        "speed_fraction": np.ones(n),
        "pos": centre + np.where(ship[:, None], rng.uniform(-1000.0, 1000.0, (n, 3)),
                                 rng.uniform(-2000.0, 2000.0, (n, 3))),
        "vel": np.zeros((n, 3)), "free": ship.copy(), "massive": np.ones(n, dtype=bool),
        "interactive": rng.uniform(0.0, 1.0, n) < 0.5, "moribund": np.zeros(n, dtype=bool),
        "cloaked": np.zeros(n, dtype=np.int64), "bubble": np.zeros(n, dtype=np.int64),
        "mode": np.full(n, 2, dtype=np.int64), "junk": np.zeros(n, dtype=bool),
    }
    ypr = None
    if dynamical:
        rows["max_ang_vel"] = rng.uniform(0.2, 2.0, n)
        rows["rot_agility"] = rng.uniform(0.3, 3.0, n)
        ypr = rng.uniform(-np.pi, np.pi, (n, 3))
    cmds = {"op": np.where(ship, GOTO_OP, 0).astype(np.int64), "point": centre + rng.uniform(-1000.0, 1000.0, (n, 3)),
            "target": np.arange(n, dtype=np.int64), "range": np.zeros(n), "min_range": np.full(n, 20000.0),
            "warp_factor": np.full(n, 20, dtype=np.int64)}
    st = static_collidables(rng, centre[::per], statics) if statics else None
    return {"n": n, "worlds": worlds, "ships": per, "rows": rows, "cmds": cmds, "ypr": ypr, "dynamical": dynamical,
            "collisions": True, "statics": st}


def load_flight_park(park, wl: dict) -> None:
    """Loads the rows into a VecModesPark of wl["n"] rows and applies the tick-1 commands through its API (a turning
    park's rotations first)."""
    r = wl["rows"]
    dev = park.dev
    t = lambda a, dt=torch.float64: torch.as_tensor(a, dtype=dt).to(dev)
    park.load_arrays(ids=t(r["id"], torch.int64), world=t(r["world"], torch.int64), time=t(r["time"], torch.int64),
                     mass=t(r["mass"]), radius=t(r["radius"]), max_vel=t(r["max_vel"]), pos=t(r["pos"]),
                     vel=t(r["vel"]), agility=t(r["agility"]), speed_fraction=t(r["speed_fraction"]),
                     free=t(r["free"], torch.bool), massive=t(r["massive"], torch.bool),
                     interactive=t(r["interactive"], torch.bool), cloaked=t(r["cloaked"], torch.int64),
                     bubble=t(r["bubble"], torch.int64), moribund=t(r["moribund"], torch.bool),
                     mode=t(r["mode"], torch.int64),
                     max_ang_vel=t(r["max_ang_vel"]) if "max_ang_vel" in r else None,
                     rot_agility=t(r["rot_agility"]) if "rot_agility" in r else None,
                     junk=t(r["junk"], torch.bool) if "junk" in r else None)
    if wl.get("statics") is not None:
        park.load_statics_arrays(**wl["statics"])
    if wl.get("ypr") is not None:
        park.set_rotation_ypr(torch.ones(wl["n"], dtype=torch.bool, device=dev), t(wl["ypr"]))
    apply_commands(park, wl["cmds"])
    launch_missiles(park, wl)


def launch_missiles(park, wl: dict) -> None:
    """The workload's tick-1 missile launches (LaunchMissile through the park's API), if any."""
    L = wl.get("launch")
    if L is None or len(L["row"]) == 0:
        return
    dev = park.dev
    full = lambda v: torch.zeros(park.n, dtype=torch.int64).index_put_((torch.as_tensor(L["row"]),),
                                                                        torch.as_tensor(v, dtype=torch.int64)).to(dev)
    mask = torch.zeros(park.n, dtype=torch.bool)
    mask[torch.as_tensor(L["row"])] = True
    park.launch_missile(mask.to(dev), full(L["target"]), full(L["owner"]), full(L["aimed"]), full(L["massive"]))


def apply_commands(park, c: dict, mask=None) -> None:
    dev = park.dev
    op = torch.as_tensor(c["op"]).to(dev)
    if mask is not None:
        op = torch.where(mask, op, 0)
    pt = torch.as_tensor(c["point"], dtype=torch.float64).to(dev)
    tg = torch.as_tensor(c["target"], dtype=torch.int64).to(dev)
    rr = torch.as_tensor(c["range"], dtype=torch.float64).to(dev)
    park.goto_point(op == GOTO_OP, pt)
    park.follow(op == FOLLOW_OP, tg, rr)
    park.orbit(op == ORBIT_OP, tg, rr)
    park.warp_to(op == WARP_OP, pt, torch.as_tensor(c["min_range"], dtype=torch.float64).to(dev),
                 torch.as_tensor(c["warp_factor"], dtype=torch.int64).to(dev))


def world_rows(wl: dict, wi: int) -> range:
    """The rows of world wi (world-major layouts: fixed ships per world, or the ragged park's start and counts)."""
    if wl.get("start") is not None:
        return range(int(wl["start"][wi]), int(wl["start"][wi] + wl["counts"][wi]))
    s = wl["ships"]
    return range(wi * s, wi * s + s)


def world_scenario(wl: dict, wi: int, steps: int) -> dict:
    """World wi of the workload as a scenario dict (balls and tick-1 commands), for the oracle."""
    r, c = wl["rows"], wl["cmds"]
    rows = world_rows(wl, wi)
    balls, cmds = [], []
    for k in rows:
        balls.append({"id": int(r["id"][k]), "mass": float(r["mass"][k]), "radius": float(r["radius"][k]),
                      "max_vel": float(r["max_vel"][k]), "agility": float(r["agility"][k]),
                      "sf": float(r["speed_fraction"][k]), "pos": [float(v) for v in r["pos"][k]],
                      "vel": [float(v) for v in r["vel"][k]], "free": bool(r["free"][k]),
                      "massive": bool(r["massive"][k]), "interactive": bool(r["interactive"][k]),
                      "cloaked": int(r["cloaked"][k]), "bubble": int(r["bubble"][k]),
                      "moribund": bool(r["moribund"][k]), "mode": int(r["mode"][k]),
                      "junk": bool(r["junk"][k]) if "junk" in r else False})
        if "max_ang_vel" in r:
            balls[-1]["max_ang_vel"] = float(r["max_ang_vel"][k])
            balls[-1]["rot_agility"] = float(r["rot_agility"][k])
    if wl.get("ypr") is not None:
        for k in rows:
            y = wl["ypr"][k]
            cmds.append((1, YPR_OP, int(r["id"][k]), 0, float(y[0]), float(y[1]), float(y[2]), 0.0, 0))
    for o in (GOTO_OP, FOLLOW_OP, ORBIT_OP, WARP_OP):    # apply_commands' order
        for k in rows:
            if int(c["op"][k]) != o:
                continue
            p = [float(v) for v in c["point"][k]]
            tgt = int(r["id"][int(c["target"][k])])
            cmds.append((1, o, int(r["id"][k]), tgt, p[0], p[1], p[2],
                         float(c["min_range"][k]) if o == WARP_OP else float(c["range"][k]),
                         int(c["warp_factor"][k])))
    L = wl.get("launch")
    if L is not None:                                    # then the launches, in row order
        for i in np.flatnonzero((L["row"] >= rows.start) & (L["row"] < rows.stop)):
            cmds.append((1, LAUNCH_OP, int(r["id"][L["row"][i]]), int(L["target"][i]), float(L["aimed"][i]),
                         float(L["massive"][i]), 0.0, 0.0, int(L["owner"][i])))
    statics = []
    st = wl.get("statics")
    if st is not None:
        wid = int(r["world"][rows[0]])
        for k in np.flatnonzero(st["world"] == wid):
            g = [float(v) for v in st["geom"][k]]
            if int(st["kind"][k]) == 0:
                statics.append({"kind": "capsule", "id": int(st["id"][k]), "parent": int(st["parent"][k]),
                                "a": g[0:3], "b": g[3:6], "radius": g[6]})
            else:
                statics.append({"kind": "box", "id": int(st["id"][k]), "parent": int(st["parent"][k]),
                                "corner": g[0:3], "lx": g[3:6], "ly": g[6:9], "lz": g[9:12]})
    return {"kind": "flight", "friction": 1.0e6, "tick": 1000, "steps": steps, "new_orbit": False, "start": 0,
            "balls": balls, "cmds": cmds, "dynamical": bool(wl.get("dynamical", False)),
            "collisions": bool(wl.get("collisions", False)), "statics": statics}


def select_worlds(wl: dict, order) -> dict:
    """The workload with its worlds in `order` (a subset or a permutation of the world indices), rows world-major in
    that order and every follow target's row renumbered; each world keeps its world id."""
    s = wl["ships"]
    order = np.asarray(order, dtype=np.int64)
    rows = (order[:, None] * s + np.arange(s)[None, :]).reshape(-1)
    new_row = np.full(wl["n"], -1, dtype=np.int64)
    new_row[rows] = np.arange(rows.size)
    r = {k: v[rows] for k, v in wl["rows"].items()}
    c = {k: v[rows] for k, v in wl["cmds"].items()}
    c["target"] = new_row[c["target"]]
    out = {"n": int(rows.size), "worlds": int(order.size), "ships": s, "rows": r, "cmds": c}
    st = wl.get("statics")
    if st is not None:                                  # the selected worlds' static collidables (by world id)
        keep = np.isin(st["world"], np.unique(r["world"]))
        out["statics"] = {k: v[keep] for k, v in st.items()}
    return out


def truncated_geometric_p(mean: float, top: int) -> float:
    """p of the geometric law on {1..top}, P(k) proportional to (1-p)^(k-1), whose mean is `mean` (bisection)."""
    k = np.arange(1, top + 1, dtype=np.float64)
    lo, hi = 1e-9, 1.0 - 1e-12
    for _ in range(200):
        p = 0.5 * (lo + hi)
        w = (1.0 - p) ** (k - 1)
        m = float((k * w).sum() / w.sum())
        lo, hi = (p, hi) if m > mean else (lo, p)
    return 0.5 * (lo + hi)


# World activity profiles of mix="profiles": (weight, probabilities of stop, goto, follow, orbit, warp at tick 1)
PROFILES = {
    "quiet": (0.40, (0.90, 0.10, 0.00, 0.00, 0.00)),       # parked or drifting
    "travel": (0.25, (0.10, 0.50, 0.00, 0.00, 0.40)),      # going places, at sub-light and in warp
    "combat": (0.20, (0.05, 0.05, 0.45, 0.45, 0.00)),      # following and orbiting each other
    "mixed": (0.15, (0.20, 0.20, 0.20, 0.20, 0.20)),
}


def ragged_park(seed: int, worlds: int, mean_ships: float = 6.0, max_ships: int = 32, mix: str = "uniform") -> dict:
    """The ragged flight park: ships per world drawn from the geometric law on {1..max_ships} with mean `mean_ships`,
    each ship as flight_park's. mix="uniform": every ship's tick-1 mode uniform over stop, goto, follow, orbit, warp;
    mix="profiles": each world first draws an activity profile (PROFILES), its ships' modes then that profile's law, so
    worlds differ in cost from one another as a server's do. Rows world-major."""
    rng = np.random.default_rng(seed)
    p = truncated_geometric_p(mean_ships, max_ships)
    k = np.arange(1, max_ships + 1)
    w = (1.0 - p) ** (k - 1)
    counts = rng.choice(k, size=worlds, p=w / w.sum()).astype(np.int64)
    n = int(counts.sum())
    world = np.repeat(np.arange(worlds, dtype=np.int64), counts)
    start = np.concatenate(([0], np.cumsum(counts)[:-1]))
    own = np.arange(n, dtype=np.int64) - np.repeat(start, counts)
    ids = own + 1
    rows = {
        "id": ids, "world": world, "time": np.zeros(n, dtype=np.int64),
        "mass": rng.uniform(1e6, 1e8, n), "radius": rng.uniform(20.0, 500.0, n),
        "max_vel": rng.uniform(100.0, 3000.0, n), "agility": rng.uniform(0.1, 5.0, n),
        "speed_fraction": np.ones(n), "pos": rng.uniform(-5e4, 5e4, (n, 3)), "vel": np.zeros((n, 3)),
        "free": np.ones(n, dtype=bool), "massive": np.ones(n, dtype=bool),
        "interactive": rng.uniform(0.0, 1.0, n) < 0.5, "moribund": np.zeros(n, dtype=bool),
        "cloaked": np.zeros(n, dtype=np.int64), "bubble": np.zeros(n, dtype=np.int64),
        "mode": np.full(n, 2, dtype=np.int64),
    }
    size = np.repeat(counts, counts)
    if mix == "uniform":
        kind = rng.integers(0, 5, n)
    elif mix == "profiles":
        names = list(PROFILES)
        prof = rng.choice(len(names), size=worlds, p=[PROFILES[k][0] for k in names])
        cum = np.cumsum(np.array([PROFILES[k][1] for k in names]), axis=1)[np.repeat(prof, counts)]
        kind = np.minimum((rng.uniform(0.0, 1.0, n)[:, None] >= cum).sum(axis=1), 4)
    else:
        raise ValueError(f"mix: 'uniform' or 'profiles', not {mix!r}")
    kind = np.where((size == 1) & ((kind == 2) | (kind == 3)), 0, kind)   # a lone ship has nothing to follow
    op = np.select([kind == 1, kind == 2, kind == 3, kind == 4], [GOTO_OP, FOLLOW_OP, ORBIT_OP, WARP_OP], 0)
    point = _ball_points(rng, n, 1e10)
    other = np.floor(rng.uniform(0.0, 1.0, n) * np.maximum(size - 1, 1)).astype(np.int64)
    other = np.where(other >= own, other + 1, other)
    target = np.repeat(start, counts) + np.minimum(other, size - 1)
    cmds = {"op": op.astype(np.int64), "point": point, "target": target.astype(np.int64),
            "range": rng.uniform(500.0, 20000.0, n), "min_range": np.full(n, 20000.0),
            "warp_factor": np.full(n, 20, dtype=np.int64)}
    return {"n": n, "worlds": worlds, "ships": None, "counts": counts, "start": start, "rows": rows, "cmds": cmds,
            "mix": mix}


def _with_missiles(wl: dict, rng: np.random.Generator, most: int) -> dict:
    """The ragged workload with U{0..most} missile balls a world after its ships (ids after the ships'): radius
    U(1, 20) m, mass U(1e3, 1e5) kg, maximum speed U(500, 3000) m/s, agility U(0.2, 2), free, at rest, STOP; each
    launched at tick 1 (LaunchMissile, after the ships' commands) by a ship of its world at another (a lone ship's
    missile has no target and flies straight on), aimed half the time, massive, one in ten as a defender (a negative
    owner). Every follow target's row renumbered."""
    W, counts, start = wl["worlds"], wl["counts"], wl["start"]
    extra = rng.integers(0, most + 1, W)
    new_counts = counts + extra
    new_start = np.concatenate(([0], np.cumsum(new_counts)[:-1]))
    n_old, n = wl["n"], int(new_counts.sum())
    ship_dest = np.repeat(new_start, counts) + (np.arange(n_old) - np.repeat(start, counts))
    mis_world = np.repeat(np.arange(W), extra)
    mis_k = np.arange(int(extra.sum())) - np.repeat(np.concatenate(([0], np.cumsum(extra)[:-1])), extra)
    mis_dest = new_start[mis_world] + counts[mis_world] + mis_k
    m = len(mis_dest)
    rows = {}
    for key, v in wl["rows"].items():
        out = np.zeros((n,) + v.shape[1:], dtype=v.dtype)
        out[ship_dest] = v
        rows[key] = out
    rows["id"][mis_dest] = counts[mis_world] + 1 + mis_k
    rows["world"][mis_dest] = mis_world
    launcher = new_start[mis_world] + (rng.uniform(0.0, 1.0, m) * counts[mis_world]).astype(np.int64)
    other = (rng.uniform(0.0, 1.0, m) * np.maximum(counts[mis_world] - 1, 1)).astype(np.int64)
    own = launcher - new_start[mis_world]
    other = np.where(other >= own, other + 1, other)
    other = np.minimum(other, counts[mis_world] - 1)
    lid = rows["id"][launcher]
    tid = np.where(counts[mis_world] > 1, other + 1, lid)
    rows["mass"][mis_dest] = rng.uniform(1e3, 1e5, m)
    rows["radius"][mis_dest] = rng.uniform(1.0, 20.0, m)
    rows["max_vel"][mis_dest] = rng.uniform(500.0, 3000.0, m)
    rows["agility"][mis_dest] = rng.uniform(0.2, 2.0, m)
    rows["speed_fraction"][mis_dest] = 1.0
    rows["pos"][mis_dest] = rows["pos"][launcher]
    rows["free"][mis_dest] = True
    rows["mode"][mis_dest] = 2
    for key in ("max_ang_vel", "rot_agility"):
        if key in rows:
            rows[key][mis_dest] = rng.uniform(0.2, 2.0, m) if key == "max_ang_vel" else rng.uniform(0.3, 3.0, m)
    cmds = {}
    for key, v in wl["cmds"].items():
        out = np.zeros((n,) + v.shape[1:], dtype=v.dtype)
        out[ship_dest] = v
        cmds[key] = out
    cmds["target"][ship_dest] = ship_dest[wl["cmds"]["target"]]
    owner = np.where(rng.uniform(0.0, 1.0, m) < 0.1, -lid, lid)
    launch = {"row": mis_dest.astype(np.int64), "target": tid.astype(np.int64), "owner": owner.astype(np.int64),
              "aimed": (rng.uniform(0.0, 1.0, m) < 0.5).astype(np.int64), "massive": np.ones(m, dtype=np.int64)}
    out = dict(wl)
    ypr = wl.get("ypr")
    if ypr is not None:
        y = np.zeros((n, 3))
        y[ship_dest] = ypr
        y[mis_dest] = rng.uniform(-np.pi, np.pi, (m, 3))
        out["ypr"] = y
    out.update(n=n, counts=new_counts, start=new_start, rows=rows, cmds=cmds, launch=launch)
    return out


def composed_park(seed: int, worlds: int, mean_ships: float = 6.0, max_ships: int = 32, mix: str = "uniform",
                  missiles: int = 0) -> dict:
    """The ragged composed park without static collidables: ragged_park's worlds (massive ships), with dynamical
    orientation (maximum angular speed U(0.2, 2) rad/s, rotational agility U(0.3, 3), each ship rotated by
    RotationQuaternion(yaw, pitch, roll) of angles U(-pi, pi)) and the gradient collision method; with `missiles`,
    U{0..missiles} missiles a world launched at tick 1 (_with_missiles)."""
    wl = ragged_park(seed, worlds, mean_ships, max_ships, mix)
    rng = np.random.default_rng(seed + 0x5EED)
    n = wl["n"]
    wl["rows"]["max_ang_vel"] = rng.uniform(0.2, 2.0, n)
    wl["rows"]["rot_agility"] = rng.uniform(0.3, 3.0, n)
    wl["rows"]["junk"] = np.zeros(n, dtype=bool)
    wl.update(ypr=rng.uniform(-np.pi, np.pi, (n, 3)), dynamical=True, collisions=True)
    if missiles:
        wl = _with_missiles(wl, np.random.default_rng(seed + 0x315), missiles)
    return wl


def padded(wl: dict, width: int) -> dict:
    """The ragged workload with every world padded to `width` rows: the padding rows are fixed balls (not free, so no
    tick touches them) with ids above the world's own; real rows keep their order and their results."""
    counts, start = wl["counts"], wl["start"]
    if int(counts.max()) > width:
        raise ValueError("a world is wider than the padding")
    W = wl["worlds"]
    dest = (np.repeat(np.arange(W) * width, counts) + (np.arange(wl["n"]) - np.repeat(start, counts)))
    n = W * width
    rows = {}
    for key, v in wl["rows"].items():
        out = np.zeros((n,) + v.shape[1:], dtype=v.dtype)
        out[dest] = v
        rows[key] = out
    slot = np.tile(np.arange(width), W)
    real = np.zeros(n, dtype=bool)
    real[dest] = True
    rows["id"] = np.where(real, rows["id"], slot + 1 + width)
    rows["world"] = np.repeat(np.arange(W, dtype=np.int64), width)
    rows["free"] = np.where(real, rows["free"], False)
    rows["mass"] = np.where(real, rows["mass"], 1.0)
    rows["max_vel"] = np.where(real, rows["max_vel"], 1.0)
    rows["agility"] = np.where(real, rows["agility"], 1.0)
    rows["speed_fraction"] = np.where(real, rows["speed_fraction"], 1.0)
    rows["mode"] = np.where(real, rows["mode"], 2)
    cmds = {}
    for key, v in wl["cmds"].items():
        out = np.zeros((n,) + v.shape[1:], dtype=v.dtype)
        out[dest] = v
        cmds[key] = out
    new_row = np.full(wl["n"], -1, dtype=np.int64)
    new_row[np.arange(wl["n"])] = dest
    cmds["target"] = np.where(real, new_row[np.clip(cmds["target"], 0, wl["n"] - 1)], 0)
    ypr = None
    if wl.get("ypr") is not None:
        ypr = np.zeros((n, 3))
        ypr[dest] = wl["ypr"]
    launch = None
    if wl.get("launch") is not None:
        launch = dict(wl["launch"], row=dest[wl["launch"]["row"]])
    return {"n": n, "worlds": W, "ships": width, "counts": counts, "start": np.arange(W) * width, "rows": rows,
            "cmds": cmds, "real": real, "dest": dest, "ypr": ypr, "dynamical": wl.get("dynamical", False),
            "collisions": wl.get("collisions", False), "launch": launch}


def partition_park(seed: int, worlds: int, min_ships: int = 4, max_ships: int = 64) -> dict:
    """The partition park (the master's space layer over the flight modes): `worlds` worlds of U{min_ships..max_ships}
    ships, each world's ships in U{1..16} Gaussian clusters (centres uniform in a +/-1e9 m cube, offsets N(0, 5e4^2)
    per axis), so that clusters in different top boxes form different bubbles. Per ship as the flight park's (mass
    U(1e6, 1e8) kg, agility U(0.1, 5), maximum speed U(100, 3000) m/s, radius U(20, 500) m), not massive (no collision
    contact: the layer runs over a client park without the gradient), interactive with probability 1/2. A proximity
    sensor on half the ships (range U(1e3, 2.5e5) m, period U(1, 5) s, shuffled by the world's rand() half the time,
    sensing interactive balls only one time in five). At tick 1 modes uniform over stop, goto (a point within 200 km of
    the ship), follow and orbit (another ship of its cluster, range U(500, 20000) m; refused across bubbles) and warp (to
    another cluster's centre, the thunker's defaults). Rows world-major; each world's slots in row order."""
    rng = np.random.default_rng(seed)
    counts = rng.integers(min_ships, max_ships + 1, worlds)
    n = int(counts.sum())
    start = np.concatenate(([0], np.cumsum(counts)[:-1])).astype(np.int64)
    world = np.repeat(np.arange(worlds, dtype=np.int64), counts)
    local = np.arange(n) - start[world]
    nclus = rng.integers(1, 17, worlds)
    centres = rng.uniform(-1e9, 1e9, (worlds, 16, 3))
    clus = (rng.uniform(0.0, 1.0, n) * nclus[world]).astype(np.int64)
    pos = centres[world, clus] + rng.normal(0.0, 5e4, (n, 3))
    rows = {
        "id": (local + 1).astype(np.int64), "world": world, "time": np.zeros(n, dtype=np.int64),
        "mass": rng.uniform(1e6, 1e8, n), "radius": rng.uniform(20.0, 500.0, n),
        "max_vel": rng.uniform(100.0, 3000.0, n), "agility": rng.uniform(0.1, 5.0, n),
        "speed_fraction": np.ones(n), "pos": pos, "vel": np.zeros((n, 3)),
        "free": np.ones(n, dtype=bool), "massive": np.zeros(n, dtype=bool),
        "interactive": rng.uniform(0.0, 1.0, n) < 0.5, "moribund": np.zeros(n, dtype=bool),
        "cloaked": np.zeros(n, dtype=np.int64), "bubble": np.full(n, -1, dtype=np.int64),
        "mode": np.full(n, 2, dtype=np.int64),
    }
    kind = rng.integers(0, 5, n)                        # 0 stop, 1 goto, 2 follow, 3 orbit, 4 warp
    # another ship of the same cluster, else none
    order = np.lexsort((local, clus, world))            # rows grouped by (world, cluster)
    key = world[order] * 16 + clus[order]
    first = np.concatenate(([0], np.flatnonzero(key[1:] != key[:-1]) + 1))
    size = np.diff(np.concatenate((first, [n])))
    grp = np.repeat(np.arange(first.size), size)
    pick = rng.integers(0, np.maximum(size[grp] - 1, 1))
    pos_in = np.arange(n) - first[grp]
    pick = np.where(pick >= pos_in, pick + 1, pick)
    mate = np.full(n, -1, dtype=np.int64)
    ok = size[grp] > 1
    mate[order[ok]] = order[first[grp[ok]] + pick[ok]]
    other_clus = (clus + 1 + rng.integers(0, 15, n)) % np.maximum(nclus[world], 1)
    warp_to = centres[world, other_clus]
    point = np.where((kind == 4)[:, None], warp_to, pos + rng.uniform(-2e5, 2e5, (n, 3)))
    op = np.select([kind == 1, (kind == 2) & (mate >= 0), (kind == 3) & (mate >= 0), (kind == 4) & (nclus[world] > 1)],
                   [GOTO_OP, FOLLOW_OP, ORBIT_OP, WARP_OP], 0)
    cmds = {"op": op.astype(np.int64), "point": point, "target": np.maximum(mate, 0), "range":
            rng.uniform(500.0, 20000.0, n), "min_range": np.full(n, 20000.0),
            "warp_factor": np.full(n, 20, dtype=np.int64)}
    sensor = rng.uniform(0.0, 1.0, n) < 0.5
    sensors = {"row": np.flatnonzero(sensor), "range": rng.uniform(1e3, 2.5e5, n)[sensor],
               "period": rng.uniform(1.0, 5.0, n)[sensor], "shuffle": (rng.uniform(0.0, 1.0, n) < 0.5)[sensor],
               "only": (rng.uniform(0.0, 1.0, n) < 0.2)[sensor]}
    return {"n": n, "worlds": worlds, "ships": None, "counts": counts, "start": start, "rows": rows, "cmds": cmds,
            "sensors": sensors, "seeds": rng.integers(1, 2 ** 31, worlds), "partition": True}


def take_worlds(wl: dict, m: int) -> dict:
    """The first m worlds of a partition park, as partition_park lays them out (rows, commands and sensors of those
    worlds; the commands' targets and the sensors' rows index rows within a world's own, so they carry over)."""
    m = min(int(m), int(wl["worlds"]))
    n = int(wl["start"][m - 1] + wl["counts"][m - 1]) if m else 0
    rows = {k: v[:n] for k, v in wl["rows"].items()}
    cmds = {k: v[:n] for k, v in wl["cmds"].items()}
    keep = wl["sensors"]["row"] < n
    sensors = {k: v[keep] for k, v in wl["sensors"].items()}
    return dict(wl, n=n, worlds=m, counts=wl["counts"][:m], start=wl["start"][:m], rows=rows, cmds=cmds,
                sensors=sensors, seeds=wl["seeds"][:m])


def space_rows(wl: dict) -> torch.Tensor:
    """[W, N] the park row of each world's slots (N the widest world), -1 past a world's ships."""
    W = wl["worlds"]
    width = int(wl["counts"].max())
    rows = np.full((W, width), -1, dtype=np.int64)
    k = np.arange(width)[None, :]
    rows = np.where(k < wl["counts"][:, None], wl["start"][:, None] + k, -1)
    return torch.as_tensor(rows, dtype=torch.int64)


def load_partition_park(park, wl: dict, space_cls):
    """Loads a partition park: the rows, the master's layer over them (each world seeded with its own srand seed) and
    its AddBall filings, the sensors (one tape op a world at a time, in row order), then the tick-1 commands. Returns
    the layer."""
    r = wl["rows"]
    dev = park.dev
    t = lambda a, dt=torch.float64: torch.as_tensor(a, dtype=dt).to(dev)
    park.load_arrays(ids=t(r["id"], torch.int64), world=t(r["world"], torch.int64), time=t(r["time"], torch.int64),
                     mass=t(r["mass"]), radius=t(r["radius"]), max_vel=t(r["max_vel"]), pos=t(r["pos"]),
                     vel=t(r["vel"]), agility=t(r["agility"]), speed_fraction=t(r["speed_fraction"]),
                     free=t(r["free"], torch.bool), massive=t(r["massive"], torch.bool),
                     interactive=t(r["interactive"], torch.bool), cloaked=t(r["cloaked"], torch.int64),
                     bubble=t(r["bubble"], torch.int64), moribund=t(r["moribund"], torch.bool),
                     mode=t(r["mode"], torch.int64))
    space = space_cls(park, space_rows(wl), seed=torch.as_tensor(wl["seeds"], dtype=torch.int64))
    space.load()
    s = wl["sensors"]
    W = wl["worlds"]
    world = r["world"][s["row"]]
    slot = s["row"] - wl["start"][world]
    per = np.bincount(world, minlength=W)
    K = int(per.max()) if len(world) else 0
    if K:
        code = np.zeros((W, K), dtype=np.int64)
        a = np.full((W, K), -1, dtype=np.int64)
        b = np.zeros((W, K), dtype=np.int64)
        f = np.zeros((W, K, 4))
        col = np.arange(len(world)) - np.concatenate(([0], np.cumsum(per)[:-1]))[world]
        code[world, col] = 7                              # ADD_SENSOR
        a[world, col] = slot
        b[world, col] = s["only"].astype(np.int64)
        f[world, col, 0], f[world, col, 1], f[world, col, 2] = s["range"], s["period"], s["shuffle"]
        space.tape(torch.as_tensor(code), torch.as_tensor(a), torch.as_tensor(b), torch.as_tensor(f))
    apply_commands(park, wl["cmds"])
    return space


def partition_scenario(wl: dict, wi: int, steps: int) -> dict:
    """World wi of a partition park as a space scenario (tests/scenarios.py's space worlds' form), for the oracle."""
    r, c, s = wl["rows"], wl["cmds"], wl["sensors"]
    rows = world_rows(wl, wi)
    balls = [{"id": int(r["id"][k]), "mass": float(r["mass"][k]), "radius": float(r["radius"][k]),
              "max_vel": float(r["max_vel"][k]), "agility": float(r["agility"][k]),
              "sf": float(r["speed_fraction"][k]), "pos": [float(v) for v in r["pos"][k]], "vel": [0.0, 0.0, 0.0],
              "free": True, "massive": False, "interactive": bool(r["interactive"][k]), "cloaked": 0, "bubble": -1,
              "moribund": False, "mode": 2} for k in rows]
    cmds = []
    for i in np.flatnonzero((s["row"] >= rows.start) & (s["row"] < rows.stop)):
        cmds.append((0, 31, int(r["id"][s["row"][i]]), 0, float(s["period"][i]), 0.0, 0.0, float(s["range"][i]),
                     int(s["shuffle"][i]) | (2 if s["only"][i] else 0)))
    for o in (GOTO_OP, FOLLOW_OP, ORBIT_OP, WARP_OP):        # apply_commands' order
        for k in rows:
            if int(c["op"][k]) != o:
                continue
            p = [float(v) for v in c["point"][k]]
            tgt = int(r["id"][int(c["target"][k])])
            cmds.append((0, o, int(r["id"][k]), tgt, p[0], p[1], p[2],
                         float(c["min_range"][k]) if o == WARP_OP else float(c["range"][k]),
                         int(c["warp_factor"][k])))
    return {"kind": "partition", "friction": 1.0e6, "tick": 1000, "steps": steps, "new_orbit": False, "start": 0,
            "master": True, "seed": int(wl["seeds"][wi]), "balls": balls, "cmds": cmds,
            "queries": {"range": 5e4, "point_offset": [0.0, 0.0, 0.0]}}
