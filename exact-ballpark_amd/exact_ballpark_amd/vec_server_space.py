"""The server layer over the master's partition: what each client is sent when a world holds several bubbles. The
drop-in for the reference's SpaceServer and CudaSpaceServer (src/exact_ballpark/vec_server_space.py:225-661,
vec_server_space_cuda.py:331-480): VecServer's queue, stamps, duplicate test, pending queries and OnTick, over the
partition of exact_ballpark_amd.vec_space.VecSpace, held to the reference's scalar server layer (server.py: Destiny's
Actions, Ticker, BubbleUpdater and ParkUpdateBatcher on a master DestinyPark over FullBallpark).

The pre-tick is one kernel launch (include/xb.h's xb_server_pre_tick with `space` set; one lane a world, the same
source on the host build): per world, in Destiny's order,

- the follower clean-up (Actions.clean_up_followers, MOV-731): a Stop queued behind the pushed calls for every free
  ball following a ball missing or in another bubble (GetRemoteFollowers: ascending id), under the duplicate test;
- the calls in queue order, each applied to the park, then its part in the partition: SetBallPosition re-files its
  ball, SetBallRadius and SetMaxSpeed too when they take the value, SetBallFree(false) of a free ball too;
  SetBallInteractive moves its ball in or out of its bubble's interactive set (the partition's op in place of the
  park's flag); UncloakBall drops its normal cloak's sensor and leaves the sensor lists of its bubble first;
- BubbleUpdater once per park time: AdditionsAndDeletions (per user ball its bubble's lists when it stayed, its own
  lists across its old and new bubbles when it moved) and CopyBubbles;
- the history: CloakBall to the bubble it was queued in (its old bubble while the two differ; read at the call),
  then its 2 s sensor (a normal cloak) and the park's cloak; UncloakBall to its own character and its AddBallsToPark
  notice to the other piloted balls of its bubble's interactive set; every other call to its ball's bubble (a
  bubble with a dictionary and an interactive ball; a local SetBallPosition nowhere), or to its ball's character
  while the ball is cloaked.

The park's tick is VecSpace.tick (the dead bubbles' balls left out, the commits re-filed, HandleProximities); the
post-tick (the same launch, phase 1) gives each pilot its ship's bubble's narrowcast when its ship is in that
bubble's interactive set and a call went to that bubble, and a singlecast when its character has history (its
bubble update's lists, a call routed to it, an uncloak's notice). `decode` turns a batch into Destiny's sends.

Not built, and refused: global balls (make_ball_global, make_ball_local and their EVERY route), missiles,
SetBallTroll (its PetrifyTroll's re-filing is the park's tick's, not the partition's: the reference's SpaceModesAdapter
leaves it out too), and parks other than VecModesPark.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass, field

import torch

from . import vec_server as VS
from . import vec_space as SP

I64, I32, I8, U8, BOOL, F64 = torch.int64, torch.int32, torch.int8, torch.uint8, torch.bool, torch.float64
STOP = VS.CODE["Stop"]
AD = {"NONE": 0, "SAME_BUBBLE": 1, "SAME_OWN": 2, "CHANGED": 3}     # the bubble update's lists per user ball
FROZEN = 4                          # a queued call's mask bit: its local argument (CloakBall's bubble) held in loc_in


class SpaceModesAdapter(VS.ModesAdapter):
    """ModesAdapter over the master's partition: the proximity sensors are the partition's; SetBallTroll is not
    applied (the reference's SpaceModesAdapter, vec_server_space.py:103-118)."""

    def __init__(self, park, master: bool = True):
        if not master:
            raise ValueError("the partition is the master's")
        super().__init__(park, master=True)
        self.kinds = frozenset(self.kinds - {VS.TROLL})


def space_for(adapter: VS.ModesAdapter, global_balls=None) -> SP.VecSpace:
    """The partition over an adapter's park: its worlds and rows, each world's slots in the order its balls were added
    (the reference's space_for, vec_server_space.py:186-205). Global balls are not built."""
    if global_balls is not None:
        raise NotImplementedError("global balls (make_ball_global, make_ball_local) are not built")
    if not isinstance(adapter, VS.ModesAdapter):
        raise ValueError(f"{type(adapter).__name__}: the partition server runs over a VecModesPark's ModesAdapter")
    return SP.VecSpace(SP.ModesView(adapter.park), adapter.rows_of_world, server=True)


@dataclass
class SpaceBatch(VS.Batch):
    """VecServer's Batch with the bubbles: each BUBBLE route's bubble, the bubble update's lists per row (kind, the
    additions and removals as world slots, their counts), the worlds' stream records at the bubble update, and each
    row's narrowcast bubble (-1: none)."""
    rbubble: torch.Tensor | None = None     # [W, J]
    ad: dict = field(default_factory=dict)  # kind [R] int8, addn [R], deln [R], add [N, R] slots, dele [N, R] slots
    ad_records: torch.Tensor | None = None  # [N, W, words]: each world's balls as the bubble update left them
    narrow: torch.Tensor | None = None      # [R]


class SpaceServer(VS.VecServer):
    """VecServer over the master's partition (several bubbles a world). `space`: the VecSpace over the adapter's park
    (by default `space_for(adapter)`), loaded here unless its caller has loaded it. `kernel`: the reference's choice of path; both are this layer's
    one kernel (the GPU's for a GPU park, the host build's for a CPU park)."""

    def __init__(self, adapter: VS.ModesAdapter, pilots=None, space: SP.VecSpace | None = None,
                 kernel: str = "cuda"):
        if not adapter.master:
            raise ValueError("the partition is the master's")
        super().__init__(adapter, pilots, kernel)
        sp = space if space is not None else space_for(adapter)
        if sp.park is not adapter.park or sp.W != self.W or sp.N != adapter.N or \
                not torch.equal(sp.rows.cpu(), adapter.rows_of_world.cpu()):
            raise ValueError("the partition must cover the adapter's park, worlds and rows (space_for)")
        self.space = sp
        if not sp.loaded:                     # (a layer loaded with its sensors by its caller is taken as it is)
            sp.load()
        R, N, d = self.R, adapter.N, self.dev
        self.ad_kind = torch.zeros(R, dtype=I8, device=d)
        self.ad_addn = torch.zeros(R, dtype=I32, device=d)
        self.ad_deln = torch.zeros(R, dtype=I32, device=d)
        self.ad_add = torch.zeros(N, R, dtype=I8, device=d)
        self.ad_del = torch.zeros(N, R, dtype=I8, device=d)
        self.narrow = torch.full((R,), -1, dtype=I64, device=d)
        self._stops = torch.zeros(self.W, dtype=I64, device=d)

    @property
    def cleanup_stops(self) -> int:
        """The follower clean-up's Stops queued so far (before their duplicate test), over every world."""
        return int(self._stops.sum())

    # ---- the kernel's arguments ----

    def _state(self, snapshot: bool) -> VS.ServerState:
        s = super()._state(snapshot)
        s.space = ctypes.addressof(self.space.state())
        for name in ("ad_kind", "ad_addn", "ad_deln", "ad_add", "ad_del", "narrow"):
            setattr(s, name, getattr(self, name).data_ptr())
        return s

    def _queue(self, slots: list, on: torch.Tensor | None, snapshot: bool) -> dict:
        """VecServer's tables with the clean-up's room behind the pushed calls: N Stops a world, which the kernel
        fills (ball, mask) at phase 0."""
        W, N, d = self.W, self.a.N, self.dev
        z = torch.zeros(W, dtype=I64, device=d)
        room = [VS.Slot(STOP, torch.zeros(W, dtype=BOOL, device=d), z, z, torch.zeros(W, 4, dtype=F64, device=d), z, z)
                for _ in range(N)]
        q = super()._queue(list(slots) + room, on, snapshot)
        q["J_user"] = len(slots)
        if slots:
            frozen = torch.tensor([FROZEN if s.frozen else 0 for s in slots], dtype=U8, device=d)
            q["mask"][:len(slots)] |= frozen[:, None]
        q["rbubble"] = torch.full((q["J"], W), -1, dtype=I64, device=d)
        return q

    # ---- the tick ----

    def pre_tick(self, snapshot: bool | None = None, mask: torch.Tensor | None = None) -> None:
        """DoPreTick's Ticker.tick over the partition for every world (or those in `mask` [W], OnTick's): the
        clean-up's Stops, the stamps, the calls with their partition parts, the bubble update, the history's routes.
        `snapshot` (default True): write the stream records `decode` needs. A world outside `mask` keeps its queue, its
        stamp and its lists, and sends nothing."""
        W, d = self.W, self.dev
        on = None if mask is None else mask.to(device=d, dtype=BOOL).reshape(W)
        slots = self.slots
        if on is None:
            self.slots = []
        else:
            kept = [VS.Slot(s.kind, s.mask & ~on, s.ball, s.target, s.f, s.n, s.loc, s.resolved, s.frozen)
                    for s in slots]
            hold = torch.stack([s.mask for s in kept]).any(dim=1).tolist() if kept else []
            self.slots = [s for s, h in zip(kept, hold) if h]
        if on is None:
            self._seen.fill_(True)
        else:
            self._seen |= on
        if snapshot is None:
            snapshot = True
        q = self._queue(slots, on, snapshot)
        self._launch(q, 0, snapshot)
        self._q = q
        Ju = q["J_user"]
        self._stops += (q["mask"][Ju:] & 1).sum(dim=0).to(I64)
        every = list(slots)
        for j in range(Ju, q["J"]):                      # the clean-up's Stops as the kernel queued them
            every.append(VS.Slot(STOP, q["mask"][j] != 0, q["ball"][j], q["target"][j], q["f"][j], q["n"][j],
                                 q["loc_in"][j]))
        for j, s in enumerate(every):
            s.ok, s.row, s.trow, s.loc = q["ok"][j] != 0, q["row"][j], q["trow"][j], q["loc"][j]
        notices = [(j, q["notice"][j] != 0, q["rec_unc"][j] if q["rec_unc"] is not None else None)
                   for j, s in enumerate(every) if s.kind == VS.UNCLOAK]
        ad = {"kind": self.ad_kind.clone(), "addn": self.ad_addn.clone(), "deln": self.ad_deln.clone(),
              "add": self.ad_add.clone(), "dele": self.ad_del.clone()}
        self.batch = SpaceBatch(stamp=self.stamp.clone(), clock=self.clock_out.clone(), slots=every,
                                route=q["route"].t(), notices=notices, additions=None,
                                gets_bubble=self.gets_bubble != 0, has_character=self.has_character != 0, ticked=on,
                                rbubble=q["rbubble"].t(), ad=ad, ad_records=q["rec_add"], narrow=self.narrow.clone())

    def _freeze_cloaks(self) -> None:
        """CloakBall's bubble is read at the call: before a park tick with calls still queued (a world's first OnTick
        evolves without a pre-tick), the queued CloakBalls' bubbles are taken now and held."""
        if not any(s.kind == VS.CLOAK and not s.frozen for s in self.slots):
            return
        q = self._tested()
        for j, s in enumerate(self.slots):
            if s.kind == VS.CLOAK and not s.frozen:
                s.loc, s.frozen = q["loc"][j].clone(), True

    def evolve(self, mask: torch.Tensor | None = None) -> None:
        """The park's tick with the partition (VecSpace.tick: the dead bubbles' balls left out, the commits re-filed,
        HandleProximities); `mask` [W] (OnTick's): those worlds alone."""
        self._freeze_cloaks()
        if mask is None:
            self.space.tick(self.a.park)
        else:
            self.space.tick(self.a.park, worlds=mask.to(device=self.dev, dtype=BOOL).reshape(self.W))

    def post_tick(self, mask: torch.Tensor | None = None) -> None:
        super().post_tick(mask)
        if self.batch is not None:
            self.batch.narrow = self.narrow.clone()

    def check(self) -> None:
        """Raise on a world outside the partition's domain (its tables' rooms, a ball in two interactive sets) or with
        a call the batched park refuses."""
        self.space.check()
        if bool((self.wflags & VS.REFUSED).any()):
            bad = torch.nonzero(self.wflags & VS.REFUSED).flatten().tolist()
            raise RuntimeError(f"worlds {bad[:8]} queued a call the batched park refuses (a mushroom as a target)")

    # ---- readers ----

    def bubbles(self) -> torch.Tensor:
        """[R] each row's bubble (newBubbleId)."""
        out = torch.full((self.R + 1,), -1, dtype=I64, device=self.dev)
        out.scatter_(0, self.space._flat, self.space.st["bubble"].t().reshape(-1))
        return out[:self.R]

    def old_bubbles(self) -> torch.Tensor:
        """[R] each row's old bubble (oldBubbleId)."""
        out = torch.full((self.R + 1,), -1, dtype=I64, device=self.dev)
        out.scatter_(0, self.space._flat, self.space.st["old_bubble"].t().reshape(-1))
        return out[:self.R]


# ===================================================================================== a tick as a HIP graph

class SpaceServerGraph(VS.ServerGraph):
    """A SpaceServer's tick with a fixed schedule of pushes captured once as a HIP graph: the pre-tick (the follower
    clean-up included), the partition's tick (dead bubbles, the park's kernel, the re-filings and HandleProximities)
    and the post-tick (the reference's SpaceServerGraph, vec_server_space_cuda.py:483-541). As ServerGraph: before
    each replay the caller writes each push's arguments into `inputs[j]`; `tick()` replays and returns the SpaceBatch
    (views of the fixed buffers; no stream records). The partition's tables are fixed buffers too."""

    def __init__(self, server: SpaceServer, schedule: list, warm: int = 1):
        if not isinstance(server, SpaceServer):
            raise TypeError("SpaceServerGraph takes a SpaceServer")
        super().__init__(server, schedule, warm)
        q, J = self.q, len(self.schedule)
        self.cleanup = [VS.Slot(STOP, q["mask"][j].view(BOOL), q["ball"][j], q["target"][j], q["f"][j], q["n"][j],
                                q["loc_in"][j]) for j in range(J, q["J"])]

    def _server_tensors(self) -> tuple:
        S = self.server
        return super()._server_tensors() + (S.ad_kind, S.ad_addn, S.ad_deln, S.ad_add, S.ad_del, S.narrow,
                                            S._stops) + tuple(S.space.st.values())

    def _body(self) -> None:
        S, q = self.server, self.q
        S._launch(q, 0)
        S._stops += (q["mask"][q["J_user"]:] & 1).sum(dim=0).to(I64)
        S.space.tick(S.a.park)
        S._launch(q, 1)

    def _batch(self) -> SpaceBatch:
        S, q = self.server, self.q
        J = len(self.schedule)
        for k, s in enumerate(self.cleanup):
            j = J + k
            s.ok, s.row, s.trow, s.loc = q["ok"][j].view(BOOL), q["row"][j], q["trow"][j], q["loc"][j]
        slots = self.slots + self.cleanup
        return SpaceBatch(stamp=S.stamp, clock=S.clock_out, slots=slots, route=q["route"].t(),
                          notices=[(j, q["notice"][j].view(BOOL), None) for j, s in enumerate(slots)
                                   if s.kind == VS.UNCLOAK],
                          additions=None, gets_bubble=S.gets_bubble.view(BOOL),
                          has_character=S.has_character.view(BOOL), rbubble=q["rbubble"].t(),
                          ad={"kind": S.ad_kind, "addn": S.ad_addn, "deln": S.ad_deln, "add": S.ad_add,
                              "dele": S.ad_del}, ad_records=None, narrow=S.narrow)


# ================================================================================ decoding (the host, the tests)

def decode(server: SpaceServer, batch: SpaceBatch | None = None, client_base: int = 10000) -> list:
    """The batch as Destiny's network sends it, per world: {client id: [[kind, method, state, wait_for_bubble,
    count]]} in the order each client receives them (its singlecast, then its ship's bubble's narrowcast); a state is
    [(stamp, (name, args))], PackagedAction payloads VS.Marshalled (ParkUpdateBatcher.send_batch and its character
    updates, server.py:947-1001 and _parkupdatebatcher.py:225-316). Host only."""
    A, b = server.a, batch or server.batch
    do = bool(A.park.dynamical_orientation)
    rows_of = [[r for r in rw if r >= 0] for rw in A.rows_of_world.tolist()]
    pilot = server.pilot.tolist()
    ids = A.ids.tolist()
    route = b.route.t().cpu().tolist()                   # [J][W]
    rbub = b.rbubble.t().cpu().tolist()
    stamp, clock = b.stamp.tolist(), b.clock.tolist()
    oks = [s.ok.tolist() for s in b.slots]
    srows = [s.row.tolist() for s in b.slots]
    has, gets, narrow = b.has_character.tolist(), b.gets_bubble.tolist(), b.narrow.tolist()
    ticked = None if b.ticked is None else b.ticked.tolist()
    notices = [(j, rec.tolist(), None if unc is None else unc.cpu()) for j, rec, unc in b.notices]
    kind, addn, deln = b.ad["kind"].tolist(), b.ad["addn"].tolist(), b.ad["deln"].tolist()
    adds, dels = b.ad["add"].t().tolist(), b.ad["dele"].t().tolist()     # [R][N] slots
    recs = None if b.ad_records is None else b.ad_records.cpu()
    out = []
    for w in range(A.W):
        if ticked is not None and not ticked[w]:
            out.append({})
            continue
        rows_w = A.rows_of_world[w].tolist()
        bubbles, chars = {}, {}
        for j, s in enumerate(b.slots):
            if not oks[j][w]:
                continue
            if route[j][w] == VS.BUBBLE:
                bubbles.setdefault(rbub[j][w], []).append((stamp[w], VS.event_of(s, w)))
            elif route[j][w] == VS.CHARACTER:
                chars.setdefault(pilot[srows[j][w]], []).append((stamp[w], VS.event_of(s, w)))
        front = {}
        for j, rec, unc in notices:                      # each uncloak's AddBallsToPark, put in front as processed
            if not oks[j][w] or srows[j][w] < 0 or not any(rec[r] for r in rows_of[w]):
                continue
            if unc is None:
                raise ValueError("this tick ran without stream records: its uncloaks' streams were not kept")
            payload = ("AddBallsToPark", (VS.ball_stream([unc[w].tolist()], clock[w], do),))
            for r in rows_of[w]:
                if rec[r]:
                    front.setdefault(pilot[r], []).insert(0, (stamp[w], payload))
        upd = {}
        for r in rows_of[w]:
            p, k = pilot[r], kind[r]
            if p < 0 or k == AD["NONE"]:
                continue
            t = clock[w]
            gone = [ids[rows_w[x]] for x in dels[r][:deln[r]]]
            new = adds[r][:addn[r]]
            entries = []
            if gone:
                rm = (t, ("RemoveBalls", (gone,)))
                entries.append((t, ("PackagedAction", VS.Marshalled((rm,)))) if k == AD["SAME_BUBBLE"] else rm)
            if new:
                if recs is None:
                    raise ValueError("this tick ran without stream records: its additions' streams were not kept")
                stream = VS.ball_stream([recs[x, w].tolist() for x in new], t, do)
                entries.append((t, ("PackagedAction", VS.Marshalled([(t, ("AddBallsToPark", (stream,)))]))))
            upd[p] = entries
        per = {}
        for r in rows_of[w]:
            p = pilot[r]
            if p < 0:
                continue
            state = front.get(p, []) + upd.get(p, []) + chars.get(p, [])
            got = []
            if has[r]:
                got.append(["single", "DoDestinyUpdate", state, bool(gets[r]), 2 if gets[r] else 1])
            if gets[r]:
                got.append(["narrow", "DoDestinyUpdate", bubbles.get(narrow[r], []), False, 2 if has[r] else 1])
            if got:
                per[client_base + p] = got
        out.append(per)
    return out
