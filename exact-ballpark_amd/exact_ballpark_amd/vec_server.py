"""Destiny's server layer over the batched park: every world's queued Actions calls, the Ticker's pre-tick (the calls
applied in queue order, the bubble's additions, the history's routes), the park's tick and the batcher's send flags,
for tens of thousands of one-bubble worlds at once. The drop-in for the reference's VecServer, its ModesAdapter,
ServerGraph, ClientView and decode (src/exact_ballpark/vec_server.py:193-1486), held to its scalar server layer
(server.py: Actions, Ticker, BubbleUpdater and ParkUpdateBatcher on a master FullBallpark).

The layer's algorithm is exact-ballpark_amd/csrc/xb_server.h, one lane a world: libxb_hip.so's xb_server_pre_tick on a
GPU park, the same source built for the host (libxb_cpu.so) on a CPU park. A tick's calls are pushed as the reference
pushes them (`push`: one Actions call for every masked world, the j-th calls of the worlds together) and stacked at the
pre-tick into [J, W] tables, so a wavefront's 64 lanes read one call of 64 worlds as consecutive words; the kernel
tests each call against the world's last queued one, applies it through the park's own commands and setters, routes
it and sets the send flags in one launch. Nothing here does arithmetic on the park's values: decoding the sends into
Destiny's tuples (`decode`) packs the stream records the kernel wrote into WriteBallToStream's bytes on the host.

The worlds are the park's worlds (its `world` labels, in order of first appearance), each one bubble; a ball's pilot
flies it (its client and character are the pilot's, PilotInterests' numbering). Refused here, as the reference
refuses them: a world whose bubble had no interactive ball at an evolve (its balls would not move on a master), a
world that lost a ball to OnTick's BringOutTheDeads (the bubble's deletions are not sent), a call the batched park
refuses (a mushroom as a target): `check()` raises.
"""

from __future__ import annotations

import ctypes
import struct
from dataclasses import dataclass, field

import torch

from . import _native as N
from .vec_modes import MAX_CATCH_UP_TICKS, VecModesPark

F64, I64, I32, I8, U8, BOOL = torch.float64, torch.int64, torch.int32, torch.int8, torch.uint8, torch.bool

# the queued actions by the event name the server sends (the reference's KINDS order: a kind's code is its index)
KINDS = ("", "GotoPoint", "GotoDirection", "Stop", "FollowBall", "Orbit", "WarpTo", "EntityWarpIn", "SetSpeedFraction",
         "SetMaxSpeed", "SetBallMass", "SetBallAgility", "SetBallRadius", "SetBallVelocity", "SetBallMassive",
         "SetBallInteractive", "SetBallPosition", "CloakBall", "UncloakBall", "SetMaxAngularSpeed",
         "SetBallAngularVelocity", "SetBallRotation", "SetBallAngularAgility", "SetMaxAngularVelocity",
         "LaunchMissile", "SetBallTroll", "SetBallFree")
CODE = {name: k for k, name in enumerate(KINDS) if name}
# Destiny's Actions methods (server.py:434-568) and the kinds they queue
ACTION_KINDS = {"go_to_point": "GotoPoint", "go_to_direction": "GotoDirection", "stop": "Stop",
                "follow_ball": "FollowBall", "orbit": "Orbit", "warp_to": "WarpTo", "entity_warp_in": "EntityWarpIn",
                "set_speed_fraction": "SetSpeedFraction", "set_max_speed": "SetMaxSpeed",
                "set_ball_mass": "SetBallMass", "set_ball_agility": "SetBallAgility",
                "set_ball_radius": "SetBallRadius", "set_ball_velocity": "SetBallVelocity",
                "set_ball_massive": "SetBallMassive", "set_ball_interactive": "SetBallInteractive",
                "set_ball_position": "SetBallPosition", "cloak_ball": "CloakBall", "uncloak_ball": "UncloakBall",
                "set_max_angular_speed": "SetMaxAngularSpeed", "set_ball_angular_velocity": "SetBallAngularVelocity",
                "set_ball_rotation": "SetBallRotation", "set_ball_angular_agility": "SetBallAngularAgility",
                "set_max_angular_velocity": "SetMaxAngularVelocity", "launch_missile": "LaunchMissile",
                "set_ball_troll": "SetBallTroll", "set_ball_free": "SetBallFree"}
# Actions refuses these without dynamical orientation (server.py:449-548)
DYNAMICAL_ONLY = frozenset(CODE[k] for k in ("SetBallAngularVelocity", "SetMaxAngularVelocity", "SetBallRotation",
                                             "SetBallAngularAgility", "SetMaxAngularSpeed"))
NONE, BUBBLE, CHARACTER = 0, 1, 2
CLOAK, UNCLOAK, SETPOS, INTERACTIVE = CODE["CloakBall"], CODE["UncloakBall"], CODE["SetBallPosition"], \
    CODE["SetBallInteractive"]
TROLL, FREE, SPEED, LAUNCH = CODE["SetBallTroll"], CODE["SetBallFree"], CODE["SetSpeedFraction"], \
    CODE["LaunchMissile"]
REFUSED = 1                         # a world flag: a call the batched park refuses
SREC_WORDS = 34                     # a stream record (xb_server.h's XB_SR_* words)
(SR_ID, SR_MODE, SR_RADIUS, SR_POS, SR_FLAGS, SR_MASS, SR_CLOAKED, SR_MAX_VEL, SR_VEL, SR_AGILITY, SR_SF, SR_ROT,
 SR_MAX_ANG_VEL, SR_ANG_VEL, SR_ROT_AGILITY, SR_FORMATION, SR_FOLLOW_ID, SR_FOLLOW_RANGE, SR_OWNER_ID, SR_STAMP,
 SR_GOTO, SR_LAST_COLLISION) = (0, 1, 2, 3, 6, 7, 8, 9, 10, 13, 14, 15, 19, 20, 23, 24, 25, 26, 27, 28, 29, 32)
# the event's arguments after the ball id, by kind (decode): three floats, one float, none, or the named layout
_XYZ = frozenset(CODE[k] for k in ("GotoPoint", "GotoDirection", "SetBallVelocity", "SetBallAngularVelocity",
                                   "SetMaxAngularVelocity", "SetBallPosition"))
_ONE = frozenset(CODE[k] for k in ("SetSpeedFraction", "SetMaxSpeed", "SetBallMass", "SetBallAgility",
                                   "SetBallRadius", "SetMaxAngularSpeed", "SetBallAngularAgility"))
_INT = frozenset(CODE[k] for k in ("SetBallMassive", "SetBallInteractive", "SetBallTroll", "SetBallFree"))


class ServerState(ctypes.Structure):
    """include/xb.h's xb_server_state."""
    _fields_ = [("W", ctypes.c_int64), ("N", ctypes.c_int64)] + [
        (n, ctypes.c_void_p) for n in ("rows", "pilot", "stamp_for_system", "stamp", "clock", "bubble_time",
                                       "settled", "fresh", "listed", "has_character", "gets_bubble", "history",
                                       "bubble_history", "wflags")] + [
        ("master", ctypes.c_int32), ("snapshot", ctypes.c_int32)] + [
        (n, ctypes.c_void_p) for n in ("space", "ad_kind", "ad_addn", "ad_deln", "ad_add", "ad_del", "narrow")]


class ServerQueue(ctypes.Structure):
    """include/xb.h's xb_server_queue."""
    _fields_ = [("J", ctypes.c_int64), ("J_user", ctypes.c_int64)] + [
        (n, ctypes.c_void_p) for n in ("kind", "mask", "ball", "target", "f", "n", "loc_in", "on", "ok", "row",
                                       "trow", "loc", "route", "notice", "rec_add", "rec_unc", "rbubble")]


def _lib(dev: torch.device) -> ctypes.CDLL:
    """The library for the park's device, its server entry points bound and its struct layout checked."""
    L = N.hip() if dev.type == "cuda" else N.cpu()
    if not getattr(L, "_server_bound", False):
        vp = ctypes.c_void_p
        if dev.type == "cuda":
            L.xb_server_pre_tick.argtypes = [vp, vp, vp, ctypes.c_int, vp]
            layout = L.xb_server_layout
        else:
            L.xb_cpu_server_pre_tick.argtypes = [vp, vp, vp, ctypes.c_int]
            layout = L.xb_cpu_server_layout
        out = (ctypes.c_int64 * 5)()
        layout(out)
        want = (ctypes.sizeof(ServerState), ServerState.narrow.offset, ctypes.sizeof(ServerQueue),
                ServerQueue.rbubble.offset, SREC_WORDS)
        if tuple(out) != want:
            raise N.NativeError(f"the server layer's layout {tuple(out)} is not the loader's {want}")
        L._server_bound = True
    return L


def _ptr(t):
    return None if t is None else t.data_ptr()


class Marshalled:
    """A marshalled payload (blue.marshal.Save) kept as its value, as the reference's server layer keeps it."""

    def __init__(self, value):
        self.value = value

    def __eq__(self, other):
        return isinstance(other, Marshalled) and self.value == other.value

    def __repr__(self):
        return f"Marshalled({self.value!r})"


# ============================================================================================== the park

class ModesAdapter:
    """The server layer's view of a VecModesPark (the reference's ModesAdapter, vec_server.py:264-391): its worlds in
    order of first appearance, each world's rows in the order they were added, its ids and flags, and the kinds of call
    it applies (every flight command and setter; the turning setters with dynamical orientation; not LaunchMissile)."""

    def __init__(self, park: VecModesPark, master: bool = True):
        if not isinstance(park, VecModesPark):
            raise TypeError("ModesAdapter takes a VecModesPark")
        self.park, self.R, self.dev, self.master = park, park.n, park.dev, bool(master)
        labels = park.world.cpu().tolist()
        order, groups = [], {}
        for r, w in enumerate(labels):
            if w not in groups:
                groups[w] = []
                order.append(w)
            groups[w].append(r)
        self.labels, self.W = order, len(order)
        self.N = max((len(g) for g in groups.values()), default=0)
        rows = [g + [-1] * (self.N - len(g)) for g in (groups[w] for w in order)]
        self.rows_of_world = torch.tensor(rows, dtype=I64, device=self.dev).reshape(self.W, self.N)
        index = {w: k for k, w in enumerate(order)}
        self.world = torch.tensor([index[w] for w in labels], dtype=I64, device=self.dev)
        self.rows_t = self.rows_of_world.t().contiguous().to(I32)          # [N][W] for the kernel
        do = park.dynamical_orientation
        self.kinds = frozenset(k for k in range(1, len(KINDS)) if (do or k not in DYNAMICAL_ONLY) and k != LAUNCH)

    @property
    def ids(self) -> torch.Tensor:
        return self.park.ids

    @property
    def exists(self) -> torch.Tensor:
        from .vec_modes import EMPTY_ID
        return self.park.ids != EMPTY_ID

    @property
    def ids_of_world(self) -> torch.Tensor:
        safe = self.rows_of_world.clamp(min=0)
        return torch.where(self.rows_of_world >= 0, self.park.ids[safe], 0)

    interactive = property(lambda s: s.park.interactive)
    cloaked = property(lambda s: s.park.cloaked)
    massive = property(lambda s: s.park.massive)

    def lookup(self, ball: torch.Tensor) -> torch.Tensor:
        """[W]: each world's row of the ball with Destiny id ball[w], or -1."""
        hit = (self.ids_of_world == ball.to(I64)[:, None]) & (self.rows_of_world >= 0)
        idx = hit.to(I64).argmax(dim=1)
        return torch.where(hit.any(dim=1), self.rows_of_world.gather(1, idx[:, None])[:, 0], -1)

    def any_in_world(self, flag: torch.Tensor) -> torch.Tensor:
        safe = self.rows_of_world.clamp(min=0)
        return (flag.to(BOOL)[safe] & (self.rows_of_world >= 0)).any(dim=1)

    def clock(self) -> torch.Tensor:
        """[W]: each world's mCurrentTime."""
        return self.park.clock[self.rows_of_world[:, 0].clamp(min=0)]

    def evolve(self, worlds: torch.Tensor | None = None) -> None:
        if worlds is None:
            self.park.evolve()
        else:
            self.park.evolve(active=worlds[self.world])


# ===================================================================================== the queue and the tick

@dataclass
class Slot:
    """One push: one Actions call for every masked world; `ok`, `row`, `trow` and `loc` are the pre-tick's (None
    before it)."""
    kind: int
    mask: torch.Tensor                  # [W] pushed
    ball: torch.Tensor                  # [W] Destiny id
    target: torch.Tensor                # [W] the target's Destiny id (FollowBall, Orbit)
    f: torch.Tensor                     # [W, 4] the float arguments
    n: torch.Tensor                     # [W] the int argument
    loc: torch.Tensor                   # [W] the local argument
    resolved: bool = False              # tested by an undo: `mask` is its queued flag
    frozen: bool = False                # `loc` holds CloakBall's bubble as read at the call (over the partition)
    ok: torch.Tensor | None = None
    row: torch.Tensor | None = None
    trow: torch.Tensor | None = None


@dataclass
class Batch:
    """One tick's sends for every world (the reference's Batch): the actions' stamp and the park's time, the queued
    calls and each one's route, the uncloaks' notices (slot, recipients [R], the uncloaked balls' records [W, words]),
    the bubble additions (fresh [R], listed [R], the worlds' records [N, W, words]), the post-tick's flags per row and
    the worlds this tick ran for (None: every world)."""
    stamp: torch.Tensor
    clock: torch.Tensor
    slots: list
    route: torch.Tensor                 # [W, J]
    notices: list = field(default_factory=list)
    additions: tuple | None = None
    gets_bubble: torch.Tensor | None = None
    has_character: torch.Tensor | None = None
    ticked: torch.Tensor | None = None


class VecServer:
    """The server layer over a batched park (an adapter): `push` queues, `tick` is the pre-tick, the park's tick and the
    post-tick (Destiny's test helper's order), `batch` the last tick's sends. `kernel`: the reference's choice between
    its torch path and its CUDA kernel; both name this layer's one kernel here (the GPU's on a CUDA park, the host
    build's on a CPU park)."""

    def __init__(self, adapter: ModesAdapter, pilots=None, kernel: str = "cuda"):
        if kernel not in ("torch", "cuda"):
            raise ValueError(f"kernel: 'torch' or 'cuda', not {kernel!r}")
        self.kernel = kernel
        self.a = A = adapter
        self.W, self.R, self.dev = A.W, A.R, A.dev
        d = self.dev
        self.pilot = (torch.full((self.R,), -1, dtype=I64, device=d) if pilots is None
                      else torch.as_tensor(pilots, dtype=I64).to(d).reshape(self.R))
        clock = A.clock()
        # Actions' first stamp: 0 at time 0 (server.py:408-411); otherwise the next tick's
        self.stamp_for_system = torch.where(clock == 0, 0, clock + 1).contiguous()
        self.bubble_time = torch.full((self.W,), -1, dtype=I64, device=d)
        self.settled = torch.zeros(self.R, dtype=U8, device=d)
        self.fresh = torch.zeros(self.R, dtype=U8, device=d)
        self.listed = torch.zeros(self.R, dtype=U8, device=d)
        self.has_character = torch.zeros(self.R, dtype=U8, device=d)
        self.gets_bubble = torch.zeros(self.R, dtype=U8, device=d)
        self.history = torch.zeros(self.R, dtype=U8, device=d)
        self.bubble_history = torch.zeros(self.W, dtype=U8, device=d)
        self.stamp = torch.zeros(self.W, dtype=I64, device=d)
        self.clock_out = torch.zeros(self.W, dtype=I64, device=d)
        self.wflags = torch.zeros(self.W, dtype=I64, device=d)
        self.dead = torch.zeros(self.W, dtype=BOOL, device=d)
        self.removed = None
        self.slots: list = []
        self.batch: Batch | None = None
        self._q = None
        self._seen = torch.zeros(self.W, dtype=BOOL, device=d)    # the worlds that have had a pre-tick
        self._lib = _lib(d)

    # ---- the kernel's arguments ----

    def _state(self, snapshot: bool) -> ServerState:
        s = ServerState()
        s.W, s.N = self.W, self.a.N
        s.rows, s.pilot = self.a.rows_t.data_ptr(), self.pilot.data_ptr()
        for name, t in (("stamp_for_system", self.stamp_for_system), ("stamp", self.stamp),
                        ("clock", self.clock_out), ("bubble_time", self.bubble_time), ("settled", self.settled),
                        ("fresh", self.fresh), ("listed", self.listed), ("has_character", self.has_character),
                        ("gets_bubble", self.gets_bubble), ("wflags", self.wflags), ("history", self.history),
                        ("bubble_history", self.bubble_history)):
            setattr(s, name, t.data_ptr())
        s.master, s.snapshot = int(self.a.master), int(bool(snapshot))
        return s

    def _queue(self, slots: list, on: torch.Tensor | None, snapshot: bool) -> dict:
        """The pre-tick's tables: the slots stacked [J, W], the kernel's outputs allocated."""
        J, W, R, d = len(slots), self.W, self.R, self.dev
        q = {"J": J}
        if J:
            q["kind"] = torch.tensor([s.kind for s in slots], dtype=I8, device=d)
            q["mask"] = torch.stack([s.mask.to(U8) | (2 if s.resolved else 0) for s in slots]).contiguous()
            for name in ("ball", "target", "n"):
                q[name] = torch.stack([getattr(s, name) for s in slots]).contiguous()
            q["loc_in"] = torch.stack([s.loc for s in slots]).contiguous()
            q["f"] = torch.stack([s.f for s in slots]).contiguous()
        else:
            for name, dt in (("kind", I8), ("mask", U8), ("ball", I64), ("target", I64), ("n", I64),
                             ("loc_in", I64)):
                q[name] = torch.zeros(0, W, dtype=dt, device=d)
            q["f"] = torch.zeros(0, W, 4, dtype=F64, device=d)
        q["on"] = None if on is None else on.to(device=d, dtype=U8).contiguous()
        q["ok"] = torch.zeros(J, W, dtype=U8, device=d)
        q["row"] = torch.zeros(J, W, dtype=I64, device=d)
        q["trow"] = torch.zeros(J, W, dtype=I64, device=d)
        q["loc"] = torch.zeros(J, W, dtype=I64, device=d)
        q["route"] = torch.zeros(J, W, dtype=I8, device=d)
        q["notice"] = torch.zeros(J, R, dtype=U8, device=d)
        q["rec_add"] = torch.zeros(self.a.N, W, SREC_WORDS, dtype=I64, device=d) if snapshot else None
        q["rec_unc"] = torch.zeros(J, W, SREC_WORDS, dtype=I64, device=d) if snapshot and J else None
        return q

    @staticmethod
    def _queue_struct(q: dict) -> ServerQueue:
        s = ServerQueue()
        s.J = q["J"]
        s.J_user = q.get("J_user", q["J"])
        for name in ("kind", "mask", "ball", "target", "f", "n", "loc_in", "on", "ok", "row", "trow", "loc", "route",
                     "notice", "rec_add", "rec_unc", "rbubble"):
            setattr(s, name, _ptr(q.get(name)))
        return s

    def _launch(self, q: dict, phase: int, snapshot: bool = False) -> None:
        park = self.a.park.state()
        srv = self._state(snapshot)
        qs = self._queue_struct(q)
        if self.dev.type == "cuda":
            with torch.cuda.device(self.dev):
                rc = self._lib.xb_server_pre_tick(ctypes.byref(park), ctypes.byref(srv), ctypes.byref(qs), phase,
                                                  torch.cuda.current_stream(self.dev).cuda_stream)
            N.check(rc, "xb_server_pre_tick")
        else:
            rc = self._lib.xb_cpu_server_pre_tick(ctypes.byref(park), ctypes.byref(srv), ctypes.byref(qs), phase)
            if rc:
                raise RuntimeError(f"xb_cpu_server_pre_tick failed ({rc})")

    # ---- Actions (TIME-709, TIME-710) ----

    def push(self, kind, mask: torch.Tensor, ball: torch.Tensor, target: torch.Tensor | None = None,
             f: torch.Tensor | None = None, n: torch.Tensor | None = None, loc: torch.Tensor | None = None) -> Slot:
        """An Actions call (kind: a code or an event name) for every world in mask [W]: ball [W] Destiny ids, target
        [W] ids (FollowBall, Orbit), f [W, k<=4] floats, n [W] the int argument (warp factor, cloak mode, flag, delay),
        loc [W] SetBallPosition's is_local_ball (1 for True). Unused arguments are zero, so the duplicate test compares
        them alike. The pre-tick tests it against the world's last queued call and looks the ball up."""
        k = CODE[kind] if isinstance(kind, str) else int(kind)
        if k not in self.a.kinds:
            raise ValueError(f"{KINDS[k] if 0 < k < len(KINDS) else kind} is not applied by the park")
        W, d = self.W, self.dev
        z = torch.zeros(W, dtype=I64, device=d)
        ff = torch.zeros(W, 4, dtype=F64, device=d)
        if f is not None:
            f = f.to(device=d, dtype=F64).reshape(W, -1)
            ff[:, :f.shape[1]] = f
        slot = Slot(k, mask.to(device=d, dtype=BOOL).reshape(W).clone(), ball.to(device=d, dtype=I64).reshape(W),
                    z if target is None else target.to(device=d, dtype=I64).reshape(W), ff,
                    z if n is None else n.to(device=d, dtype=I64).reshape(W),
                    z if loc is None else loc.to(device=d, dtype=I64).reshape(W))
        self.slots.append(slot)
        return slot

    # ---- the pending queries and the undo (Destiny's Actions, server.py:576-642) ----

    def _tested(self) -> dict:
        """The current queue's calls tested as the pre-tick will test them (each call's queued flag, rows and local
        argument), none applied: the kernel's test alone (phase 2)."""
        q = self._queue(self.slots, None, False)
        if q["J"]:
            self._launch(q, 2)
        return q

    def _last_pending(self, kind: int, mask: torch.Tensor, ball: torch.Tensor) -> tuple:
        """(found [W], the last such queued call's slot index [W], -1 where none) for calls of `kind` at ball[w]."""
        mask = mask.to(device=self.dev, dtype=BOOL).reshape(self.W)
        ball = ball.to(device=self.dev, dtype=I64).reshape(self.W)
        q = self._tested()
        found = torch.zeros(self.W, dtype=BOOL, device=self.dev)
        at = torch.full((self.W,), -1, dtype=I64, device=self.dev)
        for j, s in enumerate(self.slots):
            if s.kind != kind:
                continue
            hit = mask & (q["ok"][j] != 0) & (s.ball == ball)
            found = found | hit
            at = torch.where(hit, j, at)
        return found, at

    def _arg(self, at: torch.Tensor, attr: str, col: int | None = None) -> torch.Tensor:
        if not self.slots:
            return torch.zeros(self.W, dtype=F64 if col is not None else I64, device=self.dev)
        vals = torch.stack([getattr(s, attr) if col is None else getattr(s, attr)[:, col] for s in self.slots])
        return vals.gather(0, at.clamp(min=0)[None, :])[0]

    def _pending_or_current(self, kind: int, mask, ball, attr: str, col, current: torch.Tensor) -> tuple:
        found, at = self._last_pending(kind, mask, ball)
        row = self.a.lookup(ball.to(device=self.dev, dtype=I64).reshape(self.W))
        in_park = row >= 0
        pend = self._arg(at, attr, col)
        cur = current[row.clamp(min=0)].to(pend.dtype)
        known = mask.to(device=self.dev, dtype=BOOL).reshape(self.W) & (found | in_park)
        return torch.where(found, pend, torch.where(in_park, cur, torch.zeros_like(pend))), known

    def has_pending_cloak(self, mask: torch.Tensor, ball: torch.Tensor) -> torch.Tensor:
        """Actions.has_pending_cloak for every world in mask: [W] bool."""
        return self._last_pending(CLOAK, mask, ball)[0]

    def has_pending_uncloak(self, mask: torch.Tensor, ball: torch.Tensor) -> torch.Tensor:
        return self._last_pending(UNCLOAK, mask, ball)[0]

    def get_pending_cloak_mode(self, mask: torch.Tensor, ball: torch.Tensor) -> tuple:
        """The last queued CloakBall's mode as pushed, else the ball's cloak byte, else None: (value [W], known [W])."""
        return self._pending_or_current(CLOAK, mask, ball, "n", None, self.a.cloaked)

    def get_pending_interactive_state(self, mask: torch.Tensor, ball: torch.Tensor) -> tuple:
        v, known = self._pending_or_current(INTERACTIVE, mask, ball, "n", None, self.a.interactive.to(I64))
        return v != 0, known

    def get_pending_speed_fraction(self, mask: torch.Tensor, ball: torch.Tensor) -> tuple:
        return self._pending_or_current(SPEED, mask, ball, "f", 0, self.a.park.speed_fraction)

    def _pending_free(self, mask, ball, value: int) -> tuple:
        found, at = self._last_pending(FREE, mask, ball)
        return found & (self._arg(at, "n") == value), found

    def has_pending_set_free(self, mask: torch.Tensor, ball: torch.Tensor) -> tuple:
        """(answer [W], pending [W]): Destiny answers the int 0 where nothing is pending (`pending` False)."""
        return self._pending_free(mask, ball, 1)

    def has_pending_set_not_free(self, mask: torch.Tensor, ball: torch.Tensor) -> tuple:
        return self._pending_free(mask, ball, 0)

    def undo_pending_cloak(self, mask: torch.Tensor, ball: torch.Tensor) -> None:
        """Actions.undo_pending_cloak for every world in mask: the ball's queued CloakBalls pruned. The queue's calls
        are tested now and kept as tested (`resolved`), so the calls pushed later are tested against the queue's last
        call as it then stands."""
        mask = mask.to(device=self.dev, dtype=BOOL).reshape(self.W)
        ball = ball.to(device=self.dev, dtype=I64).reshape(self.W)
        q = self._tested()
        for j, s in enumerate(self.slots):
            ok = q["ok"][j] != 0
            if s.kind == CLOAK:
                ok = ok & ~(mask & (s.ball == ball))
            s.mask, s.resolved = ok, True

    # ---- the tick ----

    def pre_tick(self, snapshot: bool | None = None, mask: torch.Tensor | None = None) -> None:
        """DoPreTick's Ticker.tick for every world (or those in `mask` [W], OnTick's): the stamps, the queue applied in
        order (CloakBall after the bubble update), the bubble additions, the routes, the send flags. `snapshot`: write
        the stream records `decode` needs (by default when a stream can be sent: a world's first pre-tick, an
        UncloakBall, SetBallInteractive or SetBallTroll queued). A world outside `mask` keeps its queued calls, its
        stamp and its bubble update, and sends nothing."""
        W, d = self.W, self.dev
        on = None if mask is None else mask.to(device=d, dtype=BOOL).reshape(W)
        slots = self.slots
        if on is None:
            self.slots = []
            first = not bool(self._seen.all())
            self._seen.fill_(True)
        else:
            kept = [Slot(s.kind, s.mask & ~on, s.ball, s.target, s.f, s.n, s.loc, s.resolved) for s in slots]
            hold = torch.stack([s.mask for s in kept]).any(dim=1).tolist() if kept else []
            self.slots = [s for s, h in zip(kept, hold) if h]
            first = bool((on & ~self._seen).any())
            self._seen |= on
        kinds = {s.kind for s in slots}
        if snapshot is None:
            snapshot = first or bool(kinds & {UNCLOAK, INTERACTIVE, TROLL})
        q = self._queue(slots, on, snapshot)
        self._launch(q, 0, snapshot)
        self._q = q
        A = self.a
        for j, s in enumerate(slots):
            s.ok, s.row, s.trow = q["ok"][j] != 0, q["row"][j], q["trow"][j]
            s.loc = q["loc"][j]
        notices = [(j, q["notice"][j] != 0, q["rec_unc"][j] if q["rec_unc"] is not None else None)
                   for j, s in enumerate(slots) if s.kind == UNCLOAK]
        fresh = self.fresh != 0
        if on is not None:
            fresh = fresh & on[A.world]
        self.batch = Batch(stamp=self.stamp.clone(), clock=self.clock_out.clone(), slots=slots, route=q["route"].t(),
                           notices=notices, additions=(fresh, self.listed != 0, q["rec_add"]),
                           gets_bubble=self.gets_bubble != 0, has_character=self.has_character != 0, ticked=on)

    def evolve(self, mask: torch.Tensor | None = None) -> None:
        """The park's tick (Ballpark::Evolve); a world whose bubble has no interactive ball is flagged (`check`). `mask`
        [W] (OnTick's): those worlds alone, every other world's balls and clock kept bit for bit."""
        A = self.a
        live = A.any_in_world(A.interactive)
        if mask is None:
            self.dead |= ~live
            A.evolve()
            return
        on = mask.to(device=self.dev, dtype=BOOL).reshape(self.W)
        self.dead |= on & ~live
        A.evolve(on)

    def post_tick(self, mask: torch.Tensor | None = None) -> None:
        """DoPostTick's send_batch: the send flags computed again after the park's tick (a petrified troll leaves the
        interactive set), in place in the batch; the worlds outside `mask` (by default the pre-tick's) send nothing."""
        if self._q is None:
            return
        on = mask if mask is not None else (self.batch.ticked if self.batch is not None else None)
        q = dict(self._q)
        q["on"] = None if on is None else on.to(device=self.dev, dtype=U8).reshape(self.W).contiguous()
        self._launch(q, 1)
        if self.batch is not None:
            self.batch.gets_bubble = self.gets_bubble != 0
            self.batch.has_character = self.has_character != 0

    def tick(self) -> None:
        self.pre_tick()
        self.evolve()
        self.post_tick()

    def check(self) -> None:
        """Raise on a world whose bubble was dead at an evolve, one that lost a ball to OnTick's BringOutTheDeads, or
        one with a call the batched park refuses."""
        if bool(self.dead.any()):
            bad = torch.nonzero(self.dead).flatten().tolist()
            raise RuntimeError(f"worlds {bad[:8]} had no interactive ball at an evolve: on the master their balls would "
                               f"not be evolved (TIME-714), which this layer does not model")
        if self.removed is not None and bool(self.removed.any()):
            bad = torch.nonzero(self.removed).flatten().tolist()
            raise RuntimeError(f"worlds {bad[:8]} lost a ball to OnTick's BringOutTheDeads: the bubble's deletions "
                               f"(RemoveBalls) are not sent by the one-bubble layer")
        if bool((self.wflags & REFUSED).any()):
            bad = torch.nonzero(self.wflags & REFUSED).flatten().tolist()
            raise RuntimeError(f"worlds {bad[:8]} queued a call the batched park refuses (a mushroom as a target)")

    # ---- OnTick (Ballpark.cpp:219-330, TIME-702 to TIME-705) ----

    def on_tick(self, sim_time) -> list:
        """Ballpark::OnTick on the master for every world at its own simulation time (sim_time [W], 100 ns units), with
        the Ticker's tick at DoPreTick and send_batch at DoPostTick (full.py:2590-2624): BringOutTheDeads first (a
        world that loses a ball is flagged, `check`); less than an interval since the world's time: nothing; a master's
        backlog of more than MAX_CATCH_UP_TICKS intervals cut to one; a world's first call evolves once with no
        pre-tick (its queue kept); every other world its whole intervals, those with one left ticking together under a
        mask, the remainder carried in its time. Returns the batches of the ticks run, in order (a world's sends are in
        those whose `ticked` holds it); `tick_counts` holds the call's counts."""
        A, W, d = self.a, self.W, self.dev
        if getattr(self, "time", None) is None:
            self.time = torch.zeros(W, dtype=I64, device=d)
            self.first_time = torch.ones(W, dtype=BOOL, device=d)
        sim = torch.as_tensor(sim_time, dtype=I64).to(d).reshape(W)
        step = int(A.park.tick_interval) * 10000
        park = A.park
        before = park.ids.clone()
        park.bring_out_the_deads()
        gone = A.any_in_world(before != park.ids)
        self.removed = gone if self.removed is None else self.removed | gone
        behind = sim - self.time
        due = behind >= step
        whole = torch.where(due, torch.div(behind, step, rounding_mode="floor"), 0)
        resync = due & (whole > MAX_CATCH_UP_TICKS) & bool(A.master)
        behind = torch.where(resync, step, behind)
        whole = torch.where(resync, 1, whole)
        first = due & self.first_time
        rest = due & ~first
        runs = torch.where(rest, whole, 0)
        carry = torch.where(rest, behind - whole * step, 0)
        counts = torch.stack(((~due).sum(), first.sum(), first.all().to(I64), resync.sum(), runs.sum(),
                              (rest & (carry > 0)).sum(), runs.max() if W else runs.sum())).tolist()
        self.tick_counts = {"worlds too early": counts[0], "first calls": counts[1], "resynchronisations": counts[3],
                            "intervals run": counts[4], "remainders carried": counts[5]}
        if counts[1]:
            self.evolve(mask=None if counts[2] else first)
        self.first_time = self.first_time & ~first
        batches = []
        for k in range(counts[6]):
            m = rest & (runs > k)
            mm = None if bool(m.all()) else m
            self.pre_tick(mask=mm)
            self.evolve(mask=mm)
            self.post_tick(mask=mm)
            self.batch.ticked = m
            batches.append(self.batch)
        self.time = torch.where(due, sim - carry, self.time)
        return batches


# ===================================================================================== a tick as a HIP graph

class ServerGraph:
    """A VecServer's tick with a fixed schedule of pushes, captured once as a HIP graph: the pre-tick, the park's tick
    and the post-tick (the reference's ServerGraph, vec_server.py:1200-1290). `schedule` is the pushes' kinds in order;
    before each replay the caller writes each push's arguments into `inputs[j]` (mask, ball, target, f, n, loc: views
    of the queue's own tables) and `tick()` replays and returns the graph's Batch (no stream records: `decode` needs
    an eager tick). The queue's tables, the server's state and the park's blocks are fixed buffers the kernels read in
    place; the park's tick must not replace a block between replays (`VecModesPark` commands and setters write in
    place)."""

    def __init__(self, server: VecServer, schedule: list, warm: int = 1):
        if server.dev.type != "cuda":
            raise ValueError("a HIP graph needs a GPU park")
        if server.slots:
            raise ValueError("calls are queued: tick them eagerly first")
        self.server = S = server
        self.schedule = [CODE[k] if isinstance(k, str) else int(k) for k in schedule]
        for k in self.schedule:
            if k not in S.a.kinds:
                raise ValueError(f"{KINDS[k]} is not applied by the park")
        W, d = S.W, S.dev
        J = len(self.schedule)
        slots = [Slot(k, torch.zeros(W, dtype=BOOL, device=d), torch.zeros(W, dtype=I64, device=d),
                      torch.zeros(W, dtype=I64, device=d), torch.zeros(W, 4, dtype=F64, device=d),
                      torch.zeros(W, dtype=I64, device=d), torch.zeros(W, dtype=I64, device=d))
                 for k in self.schedule]
        self.q = S._queue(slots, None, False)
        self.inputs = [{"mask": self.q["mask"][j], "ball": self.q["ball"][j], "target": self.q["target"][j],
                        "f": self.q["f"][j], "n": self.q["n"][j], "loc": self.q["loc_in"][j]} for j in range(J)]
        self.slots = slots
        park = S.a.park
        self._ptrs = self._pointers()
        saved = park._save()
        srv_saved = self._server_tensors()
        srv_copy = [t.clone() for t in srv_saved]
        time0 = park.current_time
        torch.cuda.synchronize(d)
        for _ in range(max(0, int(warm))):          # the modules loaded outside the capture
            self._body()
        torch.cuda.synchronize(d)
        park._put_back(saved, torch.zeros(S.R, dtype=BOOL, device=d))
        for cur, old in zip(srv_saved, srv_copy):
            cur.copy_(old)
        park.current_time = time0
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            self._body()
        park.current_time = time0
        torch.cuda.synchronize(d)

    def _server_tensors(self) -> tuple:
        S = self.server
        return (S.stamp_for_system, S.bubble_time, S.settled, S.fresh, S.listed, S.has_character, S.gets_bubble,
                S.stamp, S.clock_out, S.wflags, S.history, S.bubble_history)

    def _pointers(self) -> tuple:
        s = self.server.a.park.state()
        return (s.f64, s.i64, s.u8, s.tiles, s.rowmap, s.coll_f64, s.coll_i64, s.hits, s.ypr, s.form, s.fld)

    def _body(self) -> None:
        S = self.server
        S._launch(self.q, 0)
        S.a.park._tick()
        S._launch(self.q, 1)

    def tick(self) -> Batch:
        S = self.server
        if S.slots:
            raise RuntimeError("eager pushes before a graph tick: the graph's tick pushes its own schedule")
        if self._pointers() != self._ptrs:
            raise RuntimeError("the park replaced a buffer the graph's launches point at: capture again")
        self.graph.replay()
        S.a.park.current_time += 1
        q = self.q                               # the kernel writes 0 or 1: the bytes viewed as bools, no launch
        for j, s in enumerate(self.slots):
            s.mask = q["mask"][j].view(BOOL)
            s.ok, s.row, s.trow, s.loc = q["ok"][j].view(BOOL), q["row"][j], q["trow"][j], q["loc"][j]
        S.batch = self._batch()
        return S.batch

    def _batch(self) -> Batch:
        S, q = self.server, self.q
        return Batch(stamp=S.stamp, clock=S.clock_out, slots=self.slots, route=q["route"].t(),
                     notices=[(j, q["notice"][j].view(BOOL), None) for j, s in enumerate(self.slots)
                              if s.kind == UNCLOAK],
                     additions=(S.fresh.view(BOOL), S.listed.view(BOOL), None),
                     gets_bubble=S.gets_bubble.view(BOOL), has_character=S.has_character.view(BOOL))


# ===================================================================================== what each client holds

class ClientView:
    """What each pilot's client holds after each tick, `down` ticks behind the server (the reference's ClientView,
    vec_server.py:1293-1350): the last `down` + 1 recorded states of the observed fields; another pilot's cloaked ship
    is absent from a client's park, its own is there. `record(time)` after each server tick; `view()` the state."""

    def __init__(self, server: VecServer, down: int, fields: tuple = ("pos", "vel", "mode")):
        if down < 0:
            raise ValueError("down: whole ticks, at least 0")
        self.server, self.down, self.fields = server, int(down), tuple(fields)
        self.depth = self.down + 1
        self.ring: dict = {}
        self.times = torch.full((self.depth,), -1, dtype=I64)
        self.k = 0

    def record(self, time: int) -> None:
        slot = self.k % self.depth
        park = self.server.a.park
        for name in self.fields + ("cloaked",):
            t = getattr(park, name)
            if name not in self.ring:
                self.ring[name] = torch.zeros((self.depth, *t.shape), dtype=t.dtype, device=t.device)
            self.ring[name][slot].copy_(t)
        self.times[slot] = int(time)
        self.k += 1

    def view(self) -> tuple:
        """(time, {field: [R, ...]}, present [R, R]: row x pilot's ship row, the ball is in that client's park); time
        -1 before the state `down` ticks back exists."""
        if self.k <= self.down:
            return -1, {}, None
        slot = (self.k - 1 - self.down) % self.depth
        out = {name: self.ring[name][slot] for name in self.fields}
        cl = self.ring["cloaked"][slot]
        a = self.server.a
        same = a.world[:, None] == a.world[None, :]
        present = same & ((cl == 0)[:, None] | torch.eye(a.R, dtype=BOOL, device=cl.device))
        return int(self.times[slot]), out, present


# ================================================================================ decoding (the host, the tests)

def _f(bits: int) -> float:
    return struct.unpack("<d", struct.pack("<q", int(bits)))[0]


def _i8(x: int) -> int:
    return ((int(x) + 128) & 0xFF) - 128


def _i32(x: int) -> int:
    x = int(x) & 0xFFFFFFFF
    return x - (1 << 32) if x >= (1 << 31) else x


def ball_bytes(rec, dynamical: bool) -> bytes:
    """Ballpark::WriteBallToStream of one stream record (Thunkers.cpp:3180-3377; the reference's transcription,
    server.py:1892-1956): little-endian, unpadded; a moribund ball writes nothing."""
    rec = [int(v) for v in rec]
    fl = rec[SR_FLAGS]
    if fl & 16:
        return b""
    mode = rec[SR_MODE] & 0xFF
    flags = (0x01 if fl & 1 else 0) | (0x04 if fl & 2 else 0) | (0x08 if fl & 4 else 0) | (0x10 if fl & 8 else 0)
    out = [struct.pack("<qBf3dB", rec[SR_ID], mode, _f(rec[SR_RADIUS]), *(_f(rec[SR_POS + k]) for k in range(3)),
                       flags)]
    if mode != 11:                                      # RIGID: no mass block
        out.append(struct.pack("<dbqii", _f(rec[SR_MASS]), _i8(rec[SR_CLOAKED]), -1, -1, -1))
    if flags & 0x01:
        out.append(struct.pack("<f3dff", _f(rec[SR_MAX_VEL]), *(_f(rec[SR_VEL + k]) for k in range(3)),
                               _f(rec[SR_AGILITY]), _f(rec[SR_SF])))
        if dynamical:
            out.append(struct.pack("<4ff3ff", *(_f(rec[SR_ROT + k]) for k in range(4)), _f(rec[SR_MAX_ANG_VEL]),
                                   *(_f(rec[SR_ANG_VEL + k]) for k in range(3)), _f(rec[SR_ROT_AGILITY])))
    out.append(struct.pack("<b", _i8(rec[SR_FORMATION])))
    stamp = _i32(rec[SR_STAMP])
    goto = [_f(rec[SR_GOTO + k]) for k in range(3)]
    if mode in (1, 4):                                  # FOLLOW, ORBIT
        out.append(struct.pack("<qf", rec[SR_FOLLOW_ID], _f(rec[SR_FOLLOW_RANGE])))
    elif mode == 12:                                    # FORMATION
        out.append(struct.pack("<qfi", rec[SR_FOLLOW_ID], _f(rec[SR_FOLLOW_RANGE]), stamp))
    elif mode == 5:                                     # MISSILE
        out.append(struct.pack("<qfqi3d", rec[SR_FOLLOW_ID], _f(rec[SR_FOLLOW_RANGE]), rec[SR_OWNER_ID], stamp, *goto))
    elif mode == 0:                                     # GOTO
        out.append(struct.pack("<3d", *goto))
    elif mode == 3:                                     # WARP
        out.append(struct.pack("<3didqq", *goto, stamp, _f(rec[SR_LAST_COLLISION]), rec[SR_FOLLOW_ID],
                               rec[SR_OWNER_ID]))
    elif mode == 6:                                     # MUSHROOM
        out.append(struct.pack("<fdiq", _f(rec[SR_FOLLOW_RANGE]), goto[0], stamp, rec[SR_OWNER_ID]))
    elif mode == 8:                                     # TROLL
        out.append(struct.pack("<i", stamp))
    return b"".join(out)


def ball_stream(recs: list, time: int, dynamical: bool) -> bytes:
    """WriteBallsToStream (Thunkers.cpp:2202-2257): DESTINY_BALLS, the time as an int32, then each ball."""
    return struct.pack("<Bi", 1, _i32(time)) + b"".join(ball_bytes(r, dynamical) for r in recs)


def event_of(slot: Slot, w: int) -> tuple:
    """The event (name, args) Destiny queued for world w's call in this slot, with the caller's argument types."""
    k = slot.kind
    name = KINDS[k]
    b = int(slot.ball[w])
    f = [float(x) for x in slot.f[w].tolist()]
    n = int(slot.n[w])
    if k in _XYZ:
        return name, (b, f[0], f[1], f[2])
    if k in _ONE:
        return name, (b, f[0])
    if k in _INT:
        return name, (b, n)
    if name in ("Stop", "UncloakBall"):
        return name, (b,)
    if name in ("FollowBall", "Orbit"):
        return name, (b, int(slot.target[w]), f[0])
    if name == "WarpTo":
        return name, (b, f[0], f[1], f[2], f[3], n)
    if name == "EntityWarpIn":
        return name, (b, f[0], f[1], f[2], n)
    if name == "CloakBall":
        return name, (b, n, f[0])
    if name == "SetBallRotation":
        return name, (b, f[0], f[1], f[2], f[3])
    raise ValueError(name)


def decode(server: VecServer, batch: Batch | None = None, client_base: int = 10000) -> list:
    """The batch as Destiny's network sends it, per world: {client id: [[kind, method, state, wait_for_bubble, count]]}
    in the order the client receives them (its singlecast, then the bubble's narrowcast); a state is [(stamp, (name,
    args))], the additions a PackagedAction of a Marshalled list. Host only."""
    A, b = server.a, batch or server.batch
    do = bool(A.park.dynamical_orientation)
    rows_of = [[r for r in rw if r >= 0] for rw in A.rows_of_world.tolist()]
    pilot = server.pilot.tolist()
    route = b.route.t().cpu().tolist()                  # [J][W]
    stamp, clock = b.stamp.tolist(), b.clock.tolist()
    oks = [s.ok.tolist() for s in b.slots]
    srows = [s.row.tolist() for s in b.slots]
    has, gets = b.has_character.tolist(), b.gets_bubble.tolist()
    ticked = None if b.ticked is None else b.ticked.tolist()
    notices = [(j, rec.tolist(), None if unc is None else unc.cpu()) for j, rec, unc in b.notices]
    fresh, listed, rec_add = b.additions
    fresh, listed = fresh.tolist(), listed.tolist()
    rec_add = None if rec_add is None else rec_add.cpu()
    out = []
    for w in range(A.W):
        if ticked is not None and not ticked[w]:
            out.append({})
            continue
        bubble, chars = [], {}
        for j, s in enumerate(b.slots):
            if not oks[j][w]:
                continue
            if route[j][w] == BUBBLE:
                bubble.append((stamp[w], event_of(s, w)))
            elif route[j][w] == CHARACTER:
                chars.setdefault(pilot[srows[j][w]], []).append((stamp[w], event_of(s, w)))
        front = {}
        for j, rec, unc in notices:                 # each uncloak's AddBallsToPark, put in front as processed
            if not oks[j][w] or srows[j][w] < 0 or not any(rec[r] for r in rows_of[w]):
                continue
            if unc is None:
                raise ValueError("this tick ran without stream records: its uncloaks' streams were not kept")
            payload = ("AddBallsToPark", (ball_stream([unc[w].tolist()], clock[w], do),))
            for r in rows_of[w]:
                if rec[r]:
                    front.setdefault(pilot[r], []).insert(0, (stamp[w], payload))
        adds = {}
        listing = [k for k, r in enumerate(rows_of[w]) if listed[r]]
        for r in rows_of[w]:
            if fresh[r] and pilot[r] >= 0 and listing:
                if rec_add is None:
                    raise ValueError("this tick ran without stream records: its additions' streams were not kept")
                blob = Marshalled([(clock[w], ("AddBallsToPark",
                                               (ball_stream([rec_add[k, w].tolist() for k in listing], clock[w],
                                                            do),)))])
                adds[pilot[r]] = [(clock[w], ("PackagedAction", blob))]
        per = {}
        for r in rows_of[w]:
            p = pilot[r]
            if p < 0:
                continue
            state = front.get(p, []) + adds.get(p, []) + chars.get(p, [])
            got = []
            if has[r]:
                got.append(["single", "DoDestinyUpdate", state, bool(gets[r]), 2 if gets[r] else 1])
            if gets[r]:
                got.append(["narrow", "DoDestinyUpdate", bubble, False, 2 if has[r] else 1])
            if got:
                per[client_base + p] = got
        out.append(per)
    return out
