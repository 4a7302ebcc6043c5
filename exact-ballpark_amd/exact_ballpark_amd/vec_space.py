"""The master's space layer over a batched park: the partition's top boxes and bubbles, the bubble dictionaries, the
interactive and keep-alive sets, the dead bubbles' skip (TIME-714), the proximity sensors, notification ranges and
target tracking (HandleProximities, SEN-701 to SEN-704) in libstdc++'s unordered_map order, glibc rand() for shuffled
sensors, CheckVisibility, GetBallIdsInRange, GetBubbleAtCoordinates and GetRemoteFollowers, for W worlds of up to N
slots each over a VecModesPark's rows. The drop-in for the reference's VecSpace and SpaceGraph
(src/exact_ballpark/vec_space.py:777-1260): the same constructor, tape of per-world operations, broadcast, query,
load, dead_rows, after_evolve, tick, check and readers, held to FullBallpark(is_master=True).

The layer's algorithm is exact-ballpark_amd/csrc/xb_space.h: one sequential algorithm a world, run by one gfx950 lane
a world (libxb_hip.so's xb_space_tape and xb_space_query) or, for a park on the CPU, by the same source built for the
host (libxb_cpu.so). Its tables are torch tensors laid out world-interleaved ([rows, W]: a wavefront's 64 worlds read
one table entry as consecutive words); the park's own rows carry each ball's bubble, interactive flag and sensor flag,
which the layer keeps current.

A tick with the layer (`tick`): DEAD flags the live rows of dead bubbles moribund for the park's tick (the park's
kernel skips a moribund row in every phase, as Evolve skips a ball in a dead bubble) and notes the rows in warp proper;
the park ticks; AFTER_EVOLVE undoes the flags, re-files the warps' drop-outs at rest and, in ascending id, every
committed ball that moved at least sqrt(0.1) m (the master's InsertInBoxes with its bubble rules), then runs
HandleProximities. No host synchronisation: a tick is capturable (`SpaceGraph`).
"""

from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _native as N
from .vec_modes import EMPTY_ID, ModesGraph, VecModesPark

I64, F64, I32, I8, U8 = torch.int64, torch.float64, torch.int32, torch.int8, torch.uint8
MAX_SLOTS = 64
OPS = ("NONE", "ADD", "REFILE", "INIT_BUBBLES", "COPY_BUBBLES", "KEEP_ALIVE", "SET_INTERACTIVE", "ADD_SENSOR",
       "REMOVE_SENSOR", "NOTIFY_RANGE", "TRACK", "REMOVE_BALL", "DEAD", "AFTER_EVOLVE", "CLEAR_EVENTS",
       "CLOAK_SENSOR", "UNCLOAK")
OP = {name: k for k, name in enumerate(OPS)}
QUERIES = ("NONE", "VISIBILITY", "RANGE_BALL", "RANGE_POINT", "BUBBLE_AT", "REMOTE_FOLLOWERS")
Q = {name: k for k, name in enumerate(QUERIES)}
EVENTS = ("", "DoProximity", "DoRange", "DoTargetTracking", "DoFinalCleanup")
FLAGS = ("EVENTS", "TOPS", "ENTRIES", "DICTS", "SETS", "PROXIMITY", "HUGE", "BAD_OP")
FLAG = {name: 1 << k for k, name in enumerate(FLAGS)}
NO_TOP = -(1 << 63)                  # GetBubbleAtCoordinates' None
IN, DEAD, KEEP = 1, 2, 128           # a slot's flag bits (xb_space.h's XS_IN, XS_DEAD, XS_KEEP)
RAND_WORDS = 33


class SpaceState(ctypes.Structure):
    """include/xb.h's xb_space_state."""
    _fields_ = [(n, ctypes.c_int64) for n in ("W", "N", "T", "E", "D", "I", "M")] + [
        (n, ctypes.c_void_p) for n in (
            "rows", "by_id", "ids", "level", "key", "mod", "bubble", "old_bubble", "seq", "sflags", "s_range",
            "s_period", "s_elapsed", "trk", "trk_range", "ntf_range", "trk_pos", "s_keys", "s_len", "s_bkt",
            "s_next", "tn", "top_key", "top_bub", "top_ref", "en", "e_bub", "e_slot", "e_act", "dn", "d_bub", "in_", "i_bub",
            "i_slot", "kn", "k_bub", "k_slot", "p_keys", "p_len", "p_bkt", "p_next", "bubble_id", "seq_next", "evn",
            "ev", "rnd", "wflags")]


def _lib(dev: torch.device) -> ctypes.CDLL:
    L = N.hip() if dev.type == "cuda" else N.cpu()
    if not getattr(L, "_space_bound", False):
        vp, i64 = ctypes.c_void_p, ctypes.c_int64
        if dev.type == "cuda":
            L.xb_space_tape.argtypes = [vp, vp, i64, vp, vp, vp, vp, vp]
            L.xb_space_query.argtypes = [vp, vp, i64, vp, vp, vp, vp, vp, vp]
            L.xb_space_layout.argtypes = [ctypes.POINTER(ctypes.c_int64)]
            out = (ctypes.c_int64 * 3)()
            L.xb_space_layout(out)
            if out[0] != ctypes.sizeof(SpaceState):
                raise N.NativeError(f"libxb_hip.so: xb_space_state is {out[0]} bytes, the loader's "
                                    f"{ctypes.sizeof(SpaceState)}")
        else:
            L.xb_cpu_space_tape.argtypes = [vp, vp, i64, vp, vp, vp, vp]
            L.xb_cpu_space_query.argtypes = [vp, vp, i64, vp, vp, vp, vp, vp]
        L._space_bound = True
    return L


class ModesView:
    """The reference's ParkView over a VecModesPark (vec_space.py's ModesView): here the park itself is the view."""

    def __init__(self, park: VecModesPark):
        self.park = park
        self.dev = park.dev
        self.R = park.n


class VecSpace:
    """The master's partition, bubbles and sensors for W worlds of N slots over a park's rows. `rows` [W, N]: the
    park's row of each slot (-1 for none), each world's slots in the order its balls are added (the ball dictionary's
    order). `kernel`: "host" or "cuda" (the park's device's by default). `events`: the room for each world's sensor
    events between clear_events. `seed`: each world's srand seed (an int for every world, or [W])."""

    def __init__(self, view, rows, kernel: str | None = None, events: int = 256, hits: bool = False,
                 block: int = 64, lib=None, server: bool = False, umap_cap: int | None = None, seed=1):
        park = view.park if isinstance(view, ModesView) else view
        if not isinstance(park, VecModesPark):
            raise ValueError("VecSpace runs over a VecModesPark (or its ModesView)")
        if lib is not None:
            raise ValueError("lib: the layer binds the park's own library")
        if hits:
            raise ValueError("the branch counters are not kept by this layer")
        self.server = bool(server)          # the server's bubble update's lists are the server layer's own tables
        self.view = view if isinstance(view, ModesView) else ModesView(park)
        self.park = park
        self.dev = park.dev
        self.kernel = kernel or ("cuda" if self.dev.type == "cuda" else "host")
        if (self.kernel == "cuda") != (self.dev.type == "cuda"):
            raise ValueError(f"kernel {self.kernel!r} on {self.dev}: 'host' on the CPU, 'cuda' on a GPU")
        rows = torch.as_tensor(rows, dtype=I64)
        if rows.dim() != 2:
            raise ValueError("rows: [W, N]")
        W, Nn = int(rows.shape[0]), int(rows.shape[1])
        if not 1 <= Nn <= MAX_SLOTS:
            raise ValueError(f"{Nn} slots a world: from 1 to {MAX_SLOTS}")
        if umap_cap is not None and int(umap_cap) < Nn:
            raise ValueError("umap_cap: this layer keeps every map's room at N (no smaller cap)")
        self.W, self.N = W, Nn
        self.T, self.E, self.D, self.I, self.M = 8 * Nn + 8, 4 * Nn + 8, 8 * Nn + 16, 2 * Nn + 8, int(events)
        d = self.dev
        ids = park.ids.cpu()
        rows_c = rows.cpu()
        safe = rows_c.clamp(min=0)
        idw = torch.where(rows_c >= 0, ids[safe], torch.full_like(rows_c, 1 << 62))
        order = torch.sort(idw, dim=1, stable=True).indices
        by_id = torch.where(torch.gather(rows_c, 1, order) >= 0, order, -1)
        self.rows = rows_c.to(d)
        t = {}

        def z(name, shape, dtype, fill=0):
            t[name] = torch.full(shape, fill, dtype=dtype, device=d)
        t["rows"] = rows_c.T.contiguous().to(I32).to(d)
        t["by_id"] = by_id.T.contiguous().to(I32).to(d)
        z("ids", (Nn, W), I64)
        z("level", (Nn, W), I8, -1)
        z("key", (3, Nn, W), I64)
        z("mod", (3, Nn, W), I8)
        z("bubble", (Nn, W), I64, -1)
        z("old_bubble", (Nn, W), I64, -1)
        z("seq", (Nn, W), I64)
        z("sflags", (Nn, W), U8)
        z("s_range", (Nn, W), torch.float32)
        z("s_period", (Nn, W), F64, 2.0)
        z("s_elapsed", (Nn, W), F64)
        z("trk", (Nn, W), I32, -1)
        z("trk_range", (Nn, W), F64)
        z("ntf_range", (Nn, W), F64)
        z("trk_pos", (3, Nn, W), F64)
        z("s_keys", (Nn, Nn, W), I8)
        z("s_len", (Nn, W), I32)
        z("s_bkt", (Nn, W), I32, 1)
        z("s_next", (Nn, W), I32)
        z("tn", (W,), I32)
        z("top_key", (3, self.T, W), I32)
        z("top_bub", (self.T, W), I64, -1)
        z("top_ref", (self.T, W), I32)
        z("en", (W,), I32)
        z("e_bub", (self.E, W), I64)
        z("e_slot", (self.E, W), I8)
        z("e_act", (self.E, W), I8)
        z("dn", (W,), I32)
        z("d_bub", (self.D, W), I64)
        z("in_", (W,), I32)
        z("i_bub", (self.I, W), I64)
        z("i_slot", (self.I, W), I8)
        z("kn", (W,), I32)
        z("k_bub", (self.I, W), I64)
        z("k_slot", (self.I, W), I8)
        z("p_keys", (Nn, W), I8)
        z("p_len", (W,), I32)
        z("p_bkt", (W,), I32, 1)
        z("p_next", (W,), I32)
        z("bubble_id", (W,), I64)
        z("seq_next", (W,), I64)
        z("evn", (W,), I32)
        z("ev", (4, self.M, W), I64)
        z("wflags", (W,), I64)
        seeds = torch.as_tensor([int(seed)] * W if isinstance(seed, int) else seed, dtype=I64).reshape(-1)
        if seeds.numel() != W:
            raise ValueError(f"seed: an int or one a world ({W}), not {seeds.numel()}")
        t["rnd"] = self._srand(seeds)
        self.st = t
        self._state = None
        self.loaded = False                 # load() has filed the balls
        self._flat = torch.where(self.rows >= 0, self.rows, park.n).flatten()   # slot -> row (past R: none)

    def _srand(self, seeds: torch.Tensor) -> torch.Tensor:
        """Each world's glibc srand(seed) state, [33, W] int32, by the library's own xb_rand_seed."""
        s32 = (seeds & 0xFFFFFFFF).to(torch.int64).numpy().astype(np.uint32)
        out = np.zeros((self.W, RAND_WORDS), dtype=np.int32)
        L = N.cpu()
        L.xb_cpu_rand_seed(s32.ctypes.data, self.W, out.ctypes.data)
        return torch.from_numpy(out.T.copy()).to(self.dev)

    # ---- plumbing ----

    def state(self) -> SpaceState:
        if self._state is None:
            s = SpaceState()
            s.W, s.N, s.T, s.E, s.D, s.I, s.M = self.W, self.N, self.T, self.E, self.D, self.I, self.M
            for name, _ in SpaceState._fields_[7:]:
                setattr(s, name, self.st[name].data_ptr())
            self._state = s
        return self._state

    def _stream(self):
        return torch.cuda.current_stream(self.dev).cuda_stream if self.kernel == "cuda" else None

    def tape(self, code, a=None, b=None, f=None) -> None:
        """Per world, K operations in order: code [W, K] (OP), a [W, K] (a slot), b [W, K] (an int), f [W, K, 4]."""
        W, d = self.W, self.dev
        code = torch.as_tensor(code, dtype=I64).to(d).reshape(W, -1)
        K = code.shape[1]
        a = torch.full_like(code, -1) if a is None else torch.as_tensor(a, dtype=I64).to(d).reshape(W, K)
        b = torch.zeros_like(code) if b is None else torch.as_tensor(b, dtype=I64).to(d).reshape(W, K)
        f = (torch.zeros(W, K, 4, dtype=F64, device=d) if f is None else
             torch.as_tensor(f, dtype=F64).to(d).reshape(W, K, 4))
        code_t, a_t, b_t = (x.T.contiguous() for x in (code, a, b))          # [K, W]
        f_t = f.permute(2, 1, 0).contiguous()                                 # [4, K, W]
        park = self.park.state()
        if self.kernel == "cuda":
            with torch.cuda.device(d):
                rc = _lib(d).xb_space_tape(ctypes.byref(park), ctypes.byref(self.state()), K, code_t.data_ptr(),
                                           a_t.data_ptr(), b_t.data_ptr(), f_t.data_ptr(), self._stream())
        else:
            rc = _lib(d).xb_cpu_space_tape(ctypes.byref(park), ctypes.byref(self.state()), K, code_t.data_ptr(),
                                           a_t.data_ptr(), b_t.data_ptr(), f_t.data_ptr())
        N.check(rc, "xb_space_tape")

    def ops(self, per_world: list) -> None:
        """A tape from per-world lists of (op name, slot, int, floats)."""
        K = max((len(x) for x in per_world), default=0)
        if K == 0:
            return
        code = torch.zeros(self.W, K, dtype=I64)
        a = torch.full((self.W, K), -1, dtype=I64)
        b = torch.zeros(self.W, K, dtype=I64)
        f = torch.zeros(self.W, K, 4, dtype=F64)
        for w, lst in enumerate(per_world):
            for k, op in enumerate(lst):
                code[w, k] = OP[op[0]]
                a[w, k] = int(op[1]) if len(op) > 1 else -1
                b[w, k] = int(op[2]) if len(op) > 2 else 0
                for i, x in enumerate(op[3] if len(op) > 3 else ()):
                    f[w, k, i] = float(x)
        self.tape(code, a, b, f)

    def broadcast(self, name: str, mask: torch.Tensor | None = None) -> None:
        """One operation without arguments for every world (or the worlds in mask [W]), on the device."""
        code = torch.full((self.W, 1), OP[name], dtype=I64, device=self.dev)
        if mask is not None:
            code = torch.where(mask.to(self.dev).reshape(self.W, 1), code, torch.zeros_like(code))
        self.tape(code)

    def query(self, kind: str, a=None, b=None, f=None) -> tuple:
        """A query for every world: (out [W, 1 + N], outf [W, N])."""
        W, Nn, d = self.W, self.N, self.dev
        a = torch.full((W,), -1, dtype=I64, device=d) if a is None else torch.as_tensor(a, dtype=I64).to(d).reshape(W)
        b = torch.zeros(W, dtype=I64, device=d) if b is None else torch.as_tensor(b, dtype=I64).to(d).reshape(W)
        ff = torch.zeros(5, W, dtype=F64, device=d)
        if f is not None:
            f = torch.as_tensor(f, dtype=F64).to(d).reshape(W, -1)
            ff[:f.shape[1]] = f.T
        out = torch.zeros(1 + Nn, W, dtype=I64, device=d)
        outf = torch.zeros(Nn, W, dtype=F64, device=d)
        a, b = a.contiguous(), b.contiguous()
        park = self.park.state()
        if self.kernel == "cuda":
            with torch.cuda.device(d):
                rc = _lib(d).xb_space_query(ctypes.byref(park), ctypes.byref(self.state()), Q[kind], a.data_ptr(),
                                            b.data_ptr(), ff.data_ptr(), out.data_ptr(), outf.data_ptr(),
                                            self._stream())
        else:
            rc = _lib(d).xb_cpu_space_query(ctypes.byref(park), ctypes.byref(self.state()), Q[kind], a.data_ptr(),
                                            b.data_ptr(), ff.data_ptr(), out.data_ptr(), outf.data_ptr())
        N.check(rc, "xb_space_query")
        return out.T, outf.T

    # ---- the load and the tick ----

    def load(self) -> None:
        """AddBall of every ball, each world's slots in order (the master files each: InsertBallInBoxes)."""
        code = torch.where(self.rows >= 0, OP["ADD"], OP["NONE"]).to(I64)
        slots = torch.arange(self.N, dtype=I64, device=self.dev).expand(self.W, self.N).contiguous()
        self.tape(code, slots)
        self.loaded = True

    def dead_rows(self, worlds: torch.Tensor | None = None) -> torch.Tensor:
        """Phase 1's dead bubbles (TIME-714): their live rows flagged moribund for the park's tick, and [R] of them
        (no host synchronisation); `worlds` [W] limits it to those worlds."""
        self.broadcast("DEAD", worlds)
        dead = (self.st["sflags"] & DEAD).T.reshape(-1) != 0
        out = torch.zeros(self.park.n + 1, dtype=torch.bool, device=self.dev)
        out.scatter_(0, self._flat, dead)
        return out[:self.park.n]

    def after_evolve(self, dropped_rows=None, worlds: torch.Tensor | None = None) -> None:
        """After the park's tick: the flags undone, the warps' drop-outs and the commits re-filed, HandleProximities
        (the drop-outs are the rows DEAD noted in warp proper that left WARP; `dropped_rows` is not needed)."""
        self.broadcast("AFTER_EVOLVE", worlds)

    def row_mask(self, worlds: torch.Tensor) -> torch.Tensor:
        """[R] the park's rows of the worlds in `worlds` [W]."""
        out = torch.zeros(self.park.n + 1, dtype=torch.bool, device=self.dev)
        out.scatter_(0, self._flat, worlds.to(torch.bool).to(self.dev).reshape(self.W, 1).expand(self.W, self.N)
                     .flatten())
        return out[:self.park.n]

    def tick(self, park: VecModesPark | None = None, worlds: torch.Tensor | None = None) -> None:
        """One tick of the park with this layer, as the master's Evolve runs it (no host synchronisation):
        the dead bubbles' balls out of the park's tick, the tick, the re-filings and HandleProximities. `worlds`
        [W] (OnTick's mask) ticks those worlds alone."""
        park = park or self.park
        if park is not self.park:
            raise ValueError("the layer ticks its own park")
        self.last_dead = self.dead_rows(worlds)
        if worlds is None:
            park.evolve()
        else:
            park.evolve(active=self.row_mask(worlds))
        self.after_evolve(None, worlds)

    def clear_events(self) -> None:
        self.st["evn"].zero_()

    def check(self) -> None:
        """Raises for a world flagged outside this layer's domain (a table past its room, a huge ball)."""
        bad = torch.nonzero(self.st["wflags"]).flatten().tolist()
        if bad:
            fl = int(self.st["wflags"][bad[0]])
            raise RuntimeError(f"worlds {bad[:8]} left this layer's domain (world {bad[0]}: "
                               f"{[n for n, bit in FLAG.items() if fl & bit]})")

    # ---- commands that touch the layer ----

    def remove(self, mask: torch.Tensor) -> None:
        """RemoveBall at once on the master for the park rows in mask: the park's removal (stopped, followers
        released, the row emptied), then the layer's (DoFinalCleanup; out of its boxes, its bubble and the proximity
        map)."""
        mask = mask.to(self.dev).reshape(self.park.n)
        slots = (mask[self.rows.clamp(min=0)] & (self.rows >= 0))
        self.park.remove(mask)
        K = int(slots.sum(dim=1).max()) if self.W else 0
        if K == 0:
            return
        code = torch.zeros(self.W, K, dtype=I64, device=self.dev)
        a = torch.full((self.W, K), -1, dtype=I64, device=self.dev)
        order = torch.argsort((~slots).to(torch.int8), dim=1, stable=True)[:, :K]
        take = torch.gather(slots, 1, order)
        code[take] = OP["REMOVE_BALL"]
        a[take] = order[take]
        self.tape(code, a)

    # ---- readers (host) ----

    def _w(self, name: str, w: int):
        return self.st[name][..., w].cpu()

    def slot_ids(self, w: int) -> dict:
        ids = self._w("ids", w)
        fl = self._w("sflags", w)
        return {j: int(ids[j]) for j in range(self.N) if int(self.rows[w, j]) >= 0 and (int(fl[j]) & IN or ids[j])}

    def dicts(self, w: int) -> list:
        """The bubble dictionaries of world w in their order: [[bubble, [[ball id, action], ...]], ...]."""
        ids = self._w("ids", w)
        dn, en = int(self.st["dn"][w]), int(self.st["en"][w])
        eb, es, ea = self._w("e_bub", w)[:en], self._w("e_slot", w)[:en], self._w("e_act", w)[:en]
        out = []
        for bub in self._w("d_bub", w)[:dn].tolist():
            out.append([int(bub), [[int(ids[int(es[i])]), int(ea[i])] for i in range(en) if int(eb[i]) == bub]])
        return out

    def tops(self, w: int) -> dict:
        n = int(self.st["tn"][w])
        keys = self.st["top_key"][:, :n, w].T.cpu().tolist()
        return {tuple(k): int(b) for k, b in zip(keys, self._w("top_bub", w)[:n].tolist())}

    def _pairs(self, w: int, b: str, s: str, n: str) -> dict:
        ids = self._w("ids", w)
        k = int(self.st[n][w])
        out = {}
        for bub, slot in zip(self._w(b, w)[:k].tolist(), self._w(s, w)[:k].tolist()):
            out.setdefault(int(bub), set()).add(int(ids[int(slot)]))
        return out

    def interactives(self, w: int) -> dict:
        return self._pairs(w, "i_bub", "i_slot", "in_")

    def keep_alives(self, w: int) -> dict:
        return self._pairs(w, "k_bub", "k_slot", "kn")

    def keep_alive_balls(self, w: int) -> list:
        ids, fl = self._w("ids", w), self._w("sflags", w)
        return sorted(int(ids[j]) for j in range(self.N) if int(fl[j]) & KEEP)

    def proximity_order(self, w: int) -> list:
        ids = self._w("ids", w)
        return [int(ids[int(j)]) for j in self._w("p_keys", w)[:int(self.st["p_len"][w])].tolist()]

    def proximity_buckets(self, w: int) -> tuple:
        return int(self.st["p_bkt"][w]), int(self.st["p_next"][w])

    def sensor_list(self, w: int, slot: int) -> list:
        ids = self._w("ids", w)
        n = int(self.st["s_len"][slot, w])
        return [int(ids[int(j)]) for j in self.st["s_keys"][:n, slot, w].cpu().tolist()]

    def sensor_buckets(self, w: int, slot: int) -> tuple:
        return int(self.st["s_bkt"][slot, w]), int(self.st["s_next"][slot, w])

    def events(self, w: int) -> list:
        """The world's events since the last clear: (name, ball id, a[, b])."""
        n = int(self.st["evn"][w])
        out = []
        for kind, ball, a, b in self.st["ev"][:, :n, w].T.cpu().tolist():
            name = EVENTS[kind]
            out.append((name, ball, a, b) if name == "DoProximity" else (name, ball, a) if name != "DoFinalCleanup"
                       else (name, ball))
        return out


class SpaceGraph:
    """A tick of a VecModesPark with its VecSpace layer (`VecSpace.tick`) captured once as a HIP graph and replayed:
    the park's kernel and the layer's state are updated in place, so the state carries from replay to replay (the
    reference's SpaceGraph, src/exact_ballpark/vec_space.py:1182-1260). A call that replaces one of the park's buffers
    makes tick() raise (capture again)."""

    def __init__(self, space: VecSpace, park: VecModesPark | None = None, warm: int = 0):
        park = park or space.park
        if space.dev.type != "cuda" or park is not space.park:
            raise ValueError("SpaceGraph: the layer's own park on a GPU")
        self.space, self.park = space, park
        for _ in range(warm):
            space.tick()
        saved = [t.clone() for t in space.st.values()]          # the warm-up launch runs the modules' loading
        blocks = [b.clone() for b in park._blocks()]
        space.tick()
        for t, old in zip(space.st.values(), saved):
            t.copy_(old)
        for b, old in zip(park._blocks(), blocks):
            b.copy_(old)
        park.current_time -= 1
        torch.cuda.synchronize(space.dev)
        self._ptrs = ModesGraph._pointers(self)
        self.graph = torch.cuda.CUDAGraph()
        t0 = park.current_time
        with torch.cuda.graph(self.graph):
            space.tick()
        park.current_time = t0

    def stale(self) -> bool:
        return ModesGraph._pointers(self) != self._ptrs

    def tick(self) -> None:
        if self.stale():
            raise RuntimeError("a buffer of the park was replaced since the capture: capture again")
        self.graph.replay()
        self.park.current_time += 1


__all__ = ["VecSpace", "SpaceGraph", "ModesView", "OPS", "QUERIES", "EMPTY_ID"]
