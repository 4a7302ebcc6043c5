"""THE ORACLE's Python driver -- test infrastructure only (tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
leg). It runs tests/scenarios.py-style worlds on the scalar C restatement of the reference's flight-modes world
(oracle_modes.c: src/exact_ballpark/modes.py's ModeBallpark, one world at a time) and returns every ball's state.
"""

from __future__ import annotations

import ctypes
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "_build" / "liboracle.so"

GOTO_OP, DIR_OP, STOP_OP, FOLLOW_OP, ORBIT_OP, WARP_OP, WARPIN_OP = 1, 2, 3, 4, 5, 6, 7
VEL_OP, ROT_OP, ANGVEL_OP, MAXANG_OP, ROTAG_OP, YPR_OP = 8, 9, 10, 11, 12, 13
LAUNCH_OP, MFOLLOW_OP = 14, 15
SETFORM_OP, FFOLLOW_OP = 16, 17
CLOAK_OP, UNCLOAK_OP = 18, 19
MUSHROOM_OP, FREE_OP = 20, 21
HARMONIC_OP = 22
REMOVE_OP = 23
BRING_OP = 24
RESERVE_OP, FREESLOT_OP = 40, 41
MASTER_OPS = {30: 1, 31: 2, 32: 3, 33: 4, 34: 5, 35: 6, 36: 7, 37: 8}   # keep alive .. CopyBubbles (oracle_master.c)
EVENT_NAMES = {1: "DoProximity", 2: "DoRange", 3: "DoTargetTracking", 4: "DoFinalCleanup"}
NT = 20                      # the turning state: rot, old_rot, ang_vel, old_ang_vel, torque, rot_roll, max_ang_vel, ...
ND, NQ = 27, 12
D_NAMES = ("px", "py", "pz", "vx", "vy", "vz", "gx", "gy", "gz", "ax", "ay", "az", "opx", "opy", "opz", "ovx", "ovy",
           "ovz", "mass", "radius", "max_vel", "agility", "sf", "follow_range", "time_factor", "last_collision", "pad")
Q_NAMES = ("id", "mode", "effect_stamp", "follow_id", "owner_id", "follow_ptr", "massive", "sensor", "interactive",
           "free", "moribund", "n_followers")
_LIB = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)


def lib() -> ctypes.CDLL:
    global _LIB
    if _LIB is None:
        if not LIB_PATH.exists():
            build()
        L = ctypes.CDLL(str(LIB_PATH))
        vp, d, i64, i = ctypes.c_void_p, ctypes.c_double, ctypes.c_int64, ctypes.c_int
        pd = ctypes.POINTER(ctypes.c_double)
        pq = ctypes.POINTER(ctypes.c_int64)
        L.or_park_new.restype = vp
        L.or_park_new.argtypes = [d, i, i, i64]
        L.or_park_free.argtypes = [vp]
        L.or_add_ball.argtypes = [vp, i64, d, d, d, pd, pd, d, d, i, i, i, i64, i64, i, i64]
        for name in ("or_goto_point", "or_goto_direction"):
            getattr(L, name).argtypes = [vp, i64, d, d, d]
        L.or_stop.argtypes = [vp, i64]
        L.or_follow.argtypes = [vp, i64, i64, d]
        L.or_orbit.argtypes = [vp, i64, i64, d]
        L.or_warp_to.argtypes = [vp, i64, d, d, d, d, i64]
        L.or_entity_warp_in.argtypes = [vp, i64, d, d, d, i64]
        L.or_evolve.argtypes = [vp]
        L.or_on_tick.argtypes = [vp, ctypes.c_int64, ctypes.c_int]
        L.or_on_tick.restype = ctypes.c_int
        L.or_count.argtypes = [vp]
        L.or_get_ball.argtypes = [vp, i, pd, pq]
        L.or_get_followers.argtypes = [vp, i, pq]
        L.or_maths.argtypes = [i, i64, pd, pd]
        L.or_evolve_all.argtypes = [ctypes.POINTER(vp), i, i]
        L.or_park_set_dynamical.argtypes = [vp, i]
        L.or_add_turning.argtypes = [vp, i64, d, d]
        L.or_set_velocity.argtypes = [vp, i64, d, d, d]
        L.or_set_angular_velocity.argtypes = [vp, i64, d, d, d]
        L.or_set_rotation.argtypes = [vp, i64, d, d, d, d]
        L.or_set_max_angular_speed.argtypes = [vp, i64, d]
        L.or_set_rot_agility.argtypes = [vp, i64, d]
        L.or_set_rotation_ypr.argtypes = [vp, i64, d, d, d]
        L.or_get_turning.argtypes = [vp, i, pd]
        L.or_park_set_collisions.argtypes = [vp, i]
        L.or_overflow.argtypes = [vp]
        L.or_set_junk.argtypes = [vp, i64, i]
        L.or_get_collide.argtypes = [vp, i, pd, pq, pq]
        L.or_park_set_ypr.argtypes = [vp, i]
        L.or_get_ypr.argtypes = [vp, i, pd]
        L.or_add_capsule.argtypes = [vp, i64, i64, pd, pd, d]
        L.or_add_oriented_box.argtypes = [vp, i64, i64, pd, pd, pd, pd]
        L.or_get_static_boxes.argtypes = [vp, i, pq]
        L.or_launch_missile.argtypes = [vp, i64, i64, i64, i, i]
        L.or_missile_follow.argtypes = [vp, i64, i64, i64]
        L.or_load_formations.argtypes = [vp, i, ctypes.POINTER(ctypes.c_int), pd]
        L.or_set_ball_formation.argtypes = [vp, i64, i]
        L.or_formation_follow.argtypes = [vp, i64, i64, i]
        L.or_reserve_formation_slot.argtypes = [vp, i64]
        L.or_free_formation_slot.argtypes = [vp, i64, i64]
        L.or_get_formation_slots.restype = i64
        L.or_get_formation_slots.argtypes = [vp, i]
        L.or_cloak_ball.argtypes = [vp, i64, i]
        L.or_uncloak_ball.argtypes = [vp, i64]
        L.or_add_mushroom.argtypes = [vp, i64, d, d]
        L.or_set_free.argtypes = [vp, i64, i]
        L.or_add_miniball.argtypes = [vp, i64, pd, d]
        L.or_add_mini_capsule.argtypes = [vp, i64, pd, pd, d]
        L.or_add_mini_box.argtypes = [vp, i64, pd, pd, pd, pd]
        for f in (L.or_add_miniball, L.or_add_mini_capsule, L.or_add_mini_box):
            f.restype = i64
        L.or_set_harmonic.argtypes = [vp, i64, i64, i64, i64, i]
        L.or_take_events.argtypes = [vp, pq, pd, i]
        L.or_remove_ball.argtypes = [vp, i64]
        L.or_remove_ball_delayed.argtypes = [vp, i64, i64]
        L.or_bring_out_the_deads.argtypes = [vp]
        L.or_park_set_master.argtypes = [vp, ctypes.c_uint32]
        L.or_master_op.argtypes = [vp, i, i64, i64, d, d, i64]
        for name in ("or_master_bubbles", "or_master_tops", "or_master_dicts", "or_master_keep_balls",
                     "or_master_remote"):
            getattr(L, name).argtypes = [vp, pq]
        L.or_master_pairs.argtypes = [vp, i, pq]
        L.or_master_umap.argtypes = [vp, i, pq]
        L.or_master_bubble_id.restype = i64
        L.or_master_bubble_id.argtypes = [vp]
        L.or_master_events.argtypes = [vp, pq, i]
        L.or_master_visibility.restype = i64
        L.or_master_visibility.argtypes = [vp, i64, i64]
        L.or_master_range.argtypes = [vp, i, i64, d, d, d, d, i, pq, pd]
        L.or_master_bubble_at.restype = i64
        L.or_master_bubble_at.argtypes = [vp, d, d, d, ctypes.POINTER(ctypes.c_int)]
        L.or_master_overflow.argtypes = [vp]
        _LIB = L
    return _LIB


class World:
    """One world of the oracle: add balls, command them, tick, read them back."""

    def __init__(self, friction: float, tick_ms: int, new_orbit: bool, start: int, dynamical: bool = False,
                 collisions: bool = False, ypr: bool = False):
        self.L = lib()
        self.p = self.L.or_park_new(float(friction), int(tick_ms), int(bool(new_orbit)), int(start))
        self.dynamical = bool(dynamical)
        self.collisions = bool(collisions)
        self.L.or_park_set_dynamical(self.p, int(self.dynamical))
        self.L.or_park_set_collisions(self.p, int(self.collisions))
        self.L.or_park_set_ypr(self.p, int(bool(ypr)))
        self.reserved = []                             # ReserveFormationSlot's answers since the last take_reserved

    def __del__(self):
        if getattr(self, "p", None):
            self.L.or_park_free(self.p)
            self.p = None

    def add(self, b: dict) -> None:
        pos = (ctypes.c_double * 3)(*[float(c) for c in b["pos"]])
        vel = (ctypes.c_double * 3)(*[float(c) for c in b["vel"]])
        rc = self.L.or_add_ball(self.p, int(b["id"]), float(b["mass"]), float(b["radius"]), float(b["max_vel"]),
                                pos, vel, float(b["agility"]), float(b["sf"]), int(bool(b["free"])),
                                int(bool(b["massive"])), int(bool(b["interactive"])), int(b["cloaked"]),
                                int(b["bubble"]), int(bool(b["moribund"])), int(b["mode"]))
        if rc:
            raise ValueError(f"ball {b['id']} refused")
        if b.get("junk"):
            self.L.or_set_junk(self.p, int(b["id"]), 1)
        if self.dynamical:                             # AddDynamicallyOrientedBall and the AddBall's snap
            self.L.or_add_turning(self.p, int(b["id"]), float(b.get("max_ang_vel", 1.0)),
                                  float(b.get("rot_agility", 0.9)))

    def take_events(self) -> list:
        """The DoCollision events since the last call: (time, missile id, collider id, mLastC x, y, z)."""
        q = (ctypes.c_int64 * (3 * 512))()
        c = (ctypes.c_double * (3 * 512))()
        n = self.L.or_take_events(self.p, q, c, 512)
        if n < 0:
            raise OverflowError("the oracle left DoCollision events out (more than 512 unread): read them each tick")
        return [(q[3 * i], q[3 * i + 1], q[3 * i + 2], c[3 * i], c[3 * i + 1], c[3 * i + 2]) for i in range(n)]

    def add_mini(self, m: dict) -> None:
        """A fixed parent's mini, made at once and taking the next local id: {"kind": "ball", "parent", "offset",
        "radius"}, {"kind": "capsule", "parent", "a", "b", "radius"} (offsets from the parent) or {"kind": "box",
        "parent", "corner", "lx", "ly", "lz"} (the corner an offset)."""
        v = lambda x: (ctypes.c_double * 3)(*[float(c) for c in x])
        if m["kind"] == "ball":
            self.L.or_add_miniball(self.p, int(m["parent"]), v(m["offset"]), float(m["radius"]))
        elif m["kind"] == "capsule":
            self.L.or_add_mini_capsule(self.p, int(m["parent"]), v(m["a"]), v(m["b"]), float(m["radius"]))
        else:
            self.L.or_add_mini_box(self.p, int(m["parent"]), v(m["corner"]), v(m["lx"]), v(m["ly"]), v(m["lz"]))

    def add_static(self, s: dict) -> None:
        """AddCapsule ({"kind": "capsule", "id", "parent", "a", "b", "radius"}) or AddOrientedBox ({"kind": "box",
        "id", "parent", "corner", "lx", "ly", "lz"}), filed at once."""
        v = lambda x: (ctypes.c_double * 3)(*[float(c) for c in x])
        if s["kind"] == "capsule":
            rc = self.L.or_add_capsule(self.p, int(s["id"]), int(s["parent"]), v(s["a"]), v(s["b"]),
                                       float(s["radius"]))
        elif s["kind"] == "box":
            rc = self.L.or_add_oriented_box(self.p, int(s["id"]), int(s["parent"]), v(s["corner"]), v(s["lx"]),
                                            v(s["ly"]), v(s["lz"]))
        else:
            raise ValueError(f"unknown static collidable kind {s['kind']!r}")
        if rc:
            raise ValueError(f"static collidable {s['id']} refused")

    def take_reserved(self) -> list:
        out, self.reserved = self.reserved, []
        return out

    def formation_slots(self) -> list:
        """Each ball's reserved formation slots (bits), in ascending id."""
        return [int(self.L.or_get_formation_slots(self.p, k)) for k in range(int(self.L.or_count(self.p)))]

    def load_formations(self, formations: list) -> None:
        """LoadFormations: each formation's slot offsets ([x, y, z] each)."""
        counts = (ctypes.c_int * max(1, len(formations)))(*[len(f) for f in formations])
        flat = [float(c) for f in formations for v in f for c in v]
        offs = (ctypes.c_double * max(1, len(flat)))(*flat)
        if self.L.or_load_formations(self.p, len(formations), counts, offs):
            raise ValueError("formations refused (at most 8 of at most 32 slots)")

    def static_boxes(self, k: int) -> np.ndarray:
        """[boxes, 3]: the keys static collidable k (in the order added) is filed in, sorted."""
        out = np.zeros((128, 3), dtype=np.int64)
        cnt = self.L.or_get_static_boxes(self.p, k, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)))
        return out[:cnt]

    def command(self, c) -> None:
        _, op, a, b, x, y, z, rr, w = c
        L, p = self.L, self.p
        if op == GOTO_OP:
            L.or_goto_point(p, a, x, y, z)
        elif op == DIR_OP:
            L.or_goto_direction(p, a, x, y, z)
        elif op == STOP_OP:
            L.or_stop(p, a)
        elif op == FOLLOW_OP:
            L.or_follow(p, a, b, rr)
        elif op == ORBIT_OP:
            L.or_orbit(p, a, b, rr)
        elif op == WARP_OP:
            L.or_warp_to(p, a, x, y, z, rr, w)
        elif op == WARPIN_OP:
            L.or_entity_warp_in(p, a, x, y, z, w)
        elif op == VEL_OP:
            L.or_set_velocity(p, a, x, y, z)
        elif op == ROT_OP:
            L.or_set_rotation(p, a, x, y, z, rr)
        elif op == ANGVEL_OP:
            L.or_set_angular_velocity(p, a, x, y, z)
        elif op == MAXANG_OP:
            L.or_set_max_angular_speed(p, a, rr)
        elif op == ROTAG_OP:
            L.or_set_rot_agility(p, a, rr)
        elif op == YPR_OP:
            L.or_set_rotation_ypr(p, a, x, y, z)
        elif op == LAUNCH_OP:
            L.or_launch_missile(p, a, b, w, int(x), int(y))
        elif op == MFOLLOW_OP:
            L.or_missile_follow(p, a, b, w)
        elif op == SETFORM_OP:
            L.or_set_ball_formation(p, a, w)
        elif op == FFOLLOW_OP:
            L.or_formation_follow(p, a, b, w)
        elif op == RESERVE_OP:                  # ReserveFormationSlot: its answer kept for the record
            self.reserved.append([int(a), int(L.or_reserve_formation_slot(p, a))])
        elif op == FREESLOT_OP:
            L.or_free_formation_slot(p, a, w)
        elif op == CLOAK_OP:
            L.or_cloak_ball(p, a, w)
        elif op == UNCLOAK_OP:
            L.or_uncloak_ball(p, a)
        elif op == MUSHROOM_OP:
            L.or_add_mushroom(p, a, rr, x)
        elif op == FREE_OP:
            L.or_set_free(p, a, w)
        elif op == REMOVE_OP:                   # a delay of w ticks (0: at once)
            L.or_remove_ball_delayed(p, a, w)
        elif op == BRING_OP:
            L.or_bring_out_the_deads(p)
        elif op == HARMONIC_OP:                 # harmonic w, corporation b, alliance x, a force field if y
            L.or_set_harmonic(p, a, w, b, int(x), int(y))
        elif op in MASTER_OPS:                  # the master's space layer (tests/scenarios.py's space ops)
            L.or_master_op(p, MASTER_OPS[op], a, b, rr, x, w)
        else:
            raise ValueError(op)

    def set_master(self, seed: int) -> None:
        """The server's view: the partition's tops, bubbles and sensors, srand(seed) for its shuffled sensors."""
        self.L.or_park_set_master(self.p, int(seed) & 0xFFFFFFFF)

    def master_state(self, queries: dict) -> dict:
        """The master's structures and queries, as tests/golden/make_golden_space.py records the reference's."""
        L, p = self.L, self.p
        if L.or_master_overflow(p):
            raise OverflowError("a master table past the oracle's bound, or a ball wider than a top box")
        buf = (ctypes.c_int64 * 65536)()
        dbuf = (ctypes.c_double * 1024)()
        n = L.or_count(p)
        d, q, _ = self.state()
        ids = [int(x) for x in q[:, 0]]
        L.or_master_bubbles(p, buf)
        bub = [int(buf[2 * i]) for i in range(n)]
        old = [int(buf[2 * i + 1]) for i in range(n)]
        nt = L.or_master_tops(p, buf)
        tops = sorted([int(buf[4 * i + k]) for k in range(4)] for i in range(nt))
        L.or_master_dicts(p, buf)
        dicts, k = [], 1
        for _ in range(int(buf[0])):
            b, c = int(buf[k]), int(buf[k + 1])
            k += 2
            dicts.append([b, [[int(buf[k + 2 * j]), int(buf[k + 2 * j + 1])] for j in range(c)]])
            k += 2 * c

        def pairs(which):
            m = L.or_master_pairs(p, which, buf)
            out = {}
            for i in range(m):
                out.setdefault(int(buf[2 * i]), []).append(int(buf[2 * i + 1]))
            return sorted([bb, sorted(v)] for bb, v in out.items())
        inter, keep = pairs(0), pairs(1)
        nk = L.or_master_keep_balls(p, buf)
        keep_balls = sorted(int(buf[i]) for i in range(nk))
        m = L.or_master_umap(p, -1, buf)
        prox, prox_bkt = [int(buf[i]) for i in range(m)], [int(buf[m]), int(buf[m + 1])]
        sensors = []
        for i in range(n):
            m = L.or_master_umap(p, i, buf)
            sensors.append([ids[i], [int(buf[j]) for j in range(m)], int(buf[m]), int(buf[m + 1])])
        ne = L.or_master_events(p, buf, 16384)
        events = []
        for i in range(ne):
            kind, bid, a, b = (int(buf[4 * i + j]) for j in range(4))
            name = EVENT_NAMES[kind]
            events.append([name, bid, a, b] if kind == 1 else [name, bid, a] if kind in (2, 3) else [name, bid])
        R = float(queries["range"])
        vis = [int(L.or_master_visibility(p, ids[0], j)) for j in ids[1:]] if ids else []
        rng, rngd = [], []
        for i in ids:
            c = L.or_master_range(p, 1, i, 0.0, 0.0, 0.0, R, 0, buf, dbuf)
            rng.append([int(buf[j]) for j in range(max(c, 0))])
            rngd.append([[float(dbuf[j]).hex(), int(buf[j])] for j in range(max(c, 0))])
        if ids:
            pt = [float(d[0, k] + queries["point_offset"][k]) for k in range(3)]
            c = L.or_master_range(p, 0, 0, pt[0], pt[1], pt[2], R, 0, buf, dbuf)
            point = [int(buf[j]) for j in range(max(c, 0))]
        else:
            pt, point = [0.0, 0.0, 0.0], []
        at = []
        for i in range(n):
            found = ctypes.c_int(0)
            v = L.or_master_bubble_at(p, float(d[i, 0]), float(d[i, 1]), float(d[i, 2]), ctypes.byref(found))
            at.append(int(v) if found.value else None)
        nr = L.or_master_remote(p, buf)
        return {"ids": ids, "bubble": bub, "old_bubble": old, "tops": tops, "dicts": dicts, "inter": inter,
                "keep": keep, "keep_balls": keep_balls, "prox": prox, "prox_bkt": prox_bkt, "sensors": sensors,
                "bubble_id": int(L.or_master_bubble_id(p)), "events": events, "vis": vis, "range": rng,
                "range_dist": rngd, "point": [c.hex() for c in pt], "point_ids": point, "at": at,
                "remote": [int(buf[i]) for i in range(nr)]}

    def evolve(self) -> None:
        self.L.or_evolve(self.p)

    def on_tick(self, sim_time: int, master: bool = False) -> int:
        """Ballpark::OnTick at sim_time (100 ns units): the ticks it ran."""
        return int(self.L.or_on_tick(self.p, int(sim_time), int(bool(master))))

    def turning(self) -> np.ndarray:
        """[n, 20] float64, balls by ascending id: rot, old_rot, ang_vel, old_ang_vel, torque, rot_roll, max_ang_vel,
        rot_agility."""
        n = self.L.or_count(self.p)
        t = np.zeros((n, NT), dtype=np.float64)
        for k in range(n):
            self.L.or_get_turning(self.p, k, t[k].ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        return t

    def ypr(self) -> np.ndarray:
        """[n, 9] float64, balls by ascending id: yaw, pitch, roll, old_yaw, old_pitch, old_roll, roll_speed,
        old_roll_speed, yaw_delta."""
        n = self.L.or_count(self.p)
        t = np.zeros((n, 9), dtype=np.float64)
        for k in range(n):
            self.L.or_get_ypr(self.p, k, t[k].ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        return t

    def collide(self) -> tuple:
        """(last_c [n, 3] float64, filing [n, 7] int64: level, key, side, hits: list of the collision lists), balls by
        ascending id. Raises if a collision list overflowed the oracle's bound."""
        if self.L.or_overflow(self.p):
            raise OverflowError("a collision list past the oracle's bound, or a ball wider than a top box")
        n = self.L.or_count(self.p)
        c = np.zeros((n, 3), dtype=np.float64)
        f = np.zeros((n, 8), dtype=np.int64)
        hits = []
        buf = (ctypes.c_int64 * 64)()
        for k in range(n):
            m = self.L.or_get_collide(self.p, k, c[k].ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                      f[k].ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), buf)
            hits.append(list(buf[:m]))
        return c, f[:, :7], hits

    def state(self) -> tuple:
        """(doubles [n, 27] as float64, ints [n, 12], followers: list of sorted id lists), balls by ascending id."""
        n = self.L.or_count(self.p)
        d = np.zeros((n, ND), dtype=np.float64)
        q = np.zeros((n, NQ), dtype=np.int64)
        fol = []
        buf = (ctypes.c_int64 * 64)()
        for k in range(n):
            self.L.or_get_ball(self.p, k, d[k].ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                               q[k].ctypes.data_as(ctypes.POINTER(ctypes.c_int64)))
            m = self.L.or_get_followers(self.p, k, buf)
            fol.append(list(buf[:m]))
        return d, q, fol


def run_world(w: dict, record_every_step: bool = True):
    """Yields (step, doubles, ints, followers) after each tick of world w (every step, or the last only)."""
    o = World(w["friction"], w["tick"], w["new_orbit"], w["start"], w.get("dynamical", False),
              w.get("collisions", False))
    for b in w["balls"]:
        o.add(b)
    for st in w.get("statics", []):
        o.add_static(st)
    for mini in w.get("minis", []):
        o.add_mini(mini)
    if w.get("formations"):
        o.load_formations(w["formations"])
    for step in range(1, w["steps"] + 1):
        for c in w["cmds"]:
            if c[0] == step:
                o.command(c)
        o.evolve()
        if record_every_step or step == w["steps"]:
            yield (step, *o.state())


def maths(func: str, x: np.ndarray) -> np.ndarray:
    """The oracle's glibc function (sin, cos, log, exp, sincos_sin, sincos_cos) elementwise."""
    codes = {"sin": 0, "cos": 1, "log": 2, "exp": 3, "sincos_sin": 4, "sincos_cos": 5}
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.empty_like(x)
    rc = lib().or_maths(codes[func], x.size, x.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                        out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    if rc:
        raise ValueError(func)
    return out


def time_space_worlds(worlds: list, ticks: int) -> tuple:
    """As time_worlds for space worlds (the master's view: the partition, bubbles and sensors, each world srand'ed
    with its seed): every world's balls added and its step-0 commands given, then `ticks` ticks of all of them in one
    C loop on one core; returns (seconds, the worlds' oracle objects)."""
    import time
    objs = []
    for w in worlds:
        o = World(w["friction"], w["tick"], w["new_orbit"], w["start"], False, True, True)
        o.set_master(w["seed"])
        for b in w["balls"]:
            o.add(b)
        for c in w["cmds"]:
            if c[0] == 0:
                o.command(c)
        objs.append(o)
    arr = (ctypes.c_void_p * len(objs))(*[o.p for o in objs])
    t0 = time.perf_counter()
    lib().or_evolve_all(arr, len(objs), int(ticks))
    return time.perf_counter() - t0, objs


def time_worlds(worlds: list, ticks: int) -> tuple:
    """Builds every world (balls, tick-1 commands applied), then times `ticks` ticks of all of them in one C loop on
    one core; returns (seconds, the worlds' oracle objects for checking)."""
    import time
    objs = []
    for w in worlds:
        o = World(w["friction"], w["tick"], w["new_orbit"], w["start"], w.get("dynamical", False),
                  w.get("collisions", False))
        for b in w["balls"]:
            o.add(b)
        for st in w.get("statics", []):
            o.add_static(st)
        for mini in w.get("minis", []):
            o.add_mini(mini)
        if w.get("formations"):
            o.load_formations(w["formations"])
        for c in w["cmds"]:
            if c[0] == 1:
                o.command(c)
        objs.append(o)
    arr = (ctypes.c_void_p * len(objs))(*[o.p for o in objs])
    t0 = time.perf_counter()
    lib().or_evolve_all(arr, len(objs), int(ticks))
    return time.perf_counter() - t0, objs
