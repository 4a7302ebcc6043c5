"""Rollout serving with continuous batching: episodes arrive, wait in a queue, and are admitted into ONE resident park
on the GPU as rows and HBM budget free up; every serving step runs the seeded policy's commands and one tick of every
live episode in a single tiled launch, whatever step each was admitted at; finished episodes hand back their final
state (or its digest) and free their rows for the next arrivals. Nothing of an episode depends on the mix: its rows,
its admission step and its neighbours change only where its worlds are placed, and a world's bits depend on no other
world, so an episode's result is the same alone and in the mix (the tests hold it so, and to the oracle).

Episodes: each is a seed. Its ship count (2..64), its length (50..2,000 ticks), its ships' initial state and the
policy's commands all come from the counter-based draws of exact-ballpark_amd/csrc/xb_hash.h, restated here with numpy's
wrapping uint64 arithmetic for the host side (the initial rows, and `episode_commands`, the policy replayed for one
episode, which the GPU computes with xb_serve_policy).

The park: `capacity` rows of a VecModesPark, the rows handed out in power-of-two runs (2..64) by `RowAllocator`; the
tiled launch's row map lists the live episodes' rows in admission order (or in a scheduler's order) and the tiles pack
whole episodes, so free rows cost HBM only, never lanes. Admission is first come, first served, while the allocated
rows' bytes stay within the HBM budget.
"""

from __future__ import annotations

import ctypes
import time
from collections import deque

import numpy as np
import torch

from . import _native as N
from .vec_modes import VecModesPark

GOTO_OP, STOP_OP, FOLLOW_OP, ORBIT_OP, WARP_OP = 1, 3, 4, 5, 6
POLICY_OPS = (STOP_OP, GOTO_OP, FOLLOW_OP, ORBIT_OP, WARP_OP)     # by draw, as xb_policy_ops
APPLY_ORDER = (STOP_OP, GOTO_OP, FOLLOW_OP, ORBIT_OP, WARP_OP)    # the commands of one tick, op by op
(F_SHIPS, F_LENGTH, F_MASS, F_RADIUS, F_MAX_VEL, F_AGILITY, F_POS, _, _, F_INTERACTIVE, F_GAP, F_OP, F_POINT, _, _,
 F_TARGET, F_RANGE) = range(17)
POINT_HALF = 5.0e9
WARP_MIN_RANGE, WARP_FACTOR = 20000.0, 20
MIN_SHIPS, MAX_SHIPS, MIN_LENGTH, MAX_LENGTH = 2, 64, 50, 2000
_M64 = (1 << 64) - 1


# ---- the draws (xb_hash.h) ----

def _u64(x) -> np.ndarray:
    return np.asarray(x, dtype=np.uint64)


def mix64(z: np.ndarray) -> np.ndarray:
    z = _u64(z)
    z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def key(seed, field, ship, tick) -> np.ndarray:
    g = np.uint64(0x9E3779B97F4A7C15)
    with np.errstate(over="ignore"):
        h = mix64(_u64(seed) + g * (_u64(field) + np.uint64(1)))
        h = mix64(h + g * (_u64(ship) + np.uint64(1)))
        return mix64(h + g * (_u64(tick) + np.uint64(1)))


def u01(h: np.ndarray) -> np.ndarray:
    return (_u64(h) >> np.uint64(11)).astype(np.float64) * 2.0 ** -53


# ---- episodes ----

def episode_shape(seeds) -> tuple:
    """(ships, length) of each episode seed."""
    s = _u64(seeds)
    ships = MIN_SHIPS + (key(s, F_SHIPS, 0, 0) % np.uint64(MAX_SHIPS - MIN_SHIPS + 1)).astype(np.int64)
    length = MIN_LENGTH + (key(s, F_LENGTH, 0, 0) % np.uint64(MAX_LENGTH - MIN_LENGTH + 1)).astype(np.int64)
    return ships, length


def episode_rows(seeds, ships) -> dict:
    """The initial rows of episodes (their ships consecutive, ids 1..ships): mass U(1e6, 1e8) kg, radius U(20, 500) m,
    maximum speed U(100, 3000) m/s, agility U(0.1, 5), position uniform in a +/-50 km cube, at rest, free, massive,
    interactive with probability 1/2, in STOP."""
    ships = np.asarray(ships, dtype=np.int64)
    seed = np.repeat(_u64(seeds), ships)
    start = np.concatenate(([0], np.cumsum(ships)[:-1]))
    k = np.arange(int(ships.sum()), dtype=np.int64) - np.repeat(start, ships)
    u = lambda f: u01(key(seed, f, k, 0))
    n = k.size
    pos = np.stack([(2.0 * u(F_POS + a) - 1.0) * 5.0e4 for a in range(3)], axis=1)
    return {"ship": k, "id": k + 1, "mass": 1.0e6 + u(F_MASS) * 99.0e6, "radius": 20.0 + u(F_RADIUS) * 480.0,
            "max_vel": 100.0 + u(F_MAX_VEL) * 2900.0, "agility": 0.1 + u(F_AGILITY) * 4.9, "pos": pos,
            "vel": np.zeros((n, 3)), "interactive": u(F_INTERACTIVE) < 0.5}


def episode_commands(seed: int, ships: int, length: int) -> list:
    """The policy's commands of one episode, replayed on the host: (tick, op, ship, target ship, x, y, z, range)
    for ticks 1..length, each tick's commands in APPLY_ORDER and by ship. Ship k commands at tick 1 and then every
    1 + (draw mod 10) ticks."""
    k_all = np.arange(ships, dtype=np.int64)
    t = np.ones(ships, dtype=np.int64)
    ks, ts = [], []
    while True:
        on = t <= length
        if not on.any():
            break
        ks.append(k_all[on])
        ts.append(t[on])
        t = np.where(on, t + 1 + (key(seed, F_GAP, k_all, t) % np.uint64(10)).astype(np.int64), t)
    k = np.concatenate(ks)
    t = np.concatenate(ts)
    op = np.asarray(POLICY_OPS, dtype=np.int64)[(key(seed, F_OP, k, t) % np.uint64(5)).astype(np.int64)]
    p = np.stack([(2.0 * u01(key(seed, F_POINT + a, k, t)) - 1.0) * POINT_HALF for a in range(3)], axis=1)
    other = (key(seed, F_TARGET, k, t) % np.uint64(max(ships - 1, 1))).astype(np.int64)
    other = np.minimum(np.where(other >= k, other + 1, other), ships - 1)
    rng = 500.0 + u01(key(seed, F_RANGE, k, t)) * 19500.0
    rank = np.asarray([APPLY_ORDER.index(o) if o in APPLY_ORDER else 9 for o in range(8)])[op]
    order = np.lexsort((k, rank, t))
    return [(int(t[i]), int(op[i]), int(k[i]), int(other[i]), float(p[i, 0]), float(p[i, 1]), float(p[i, 2]),
             float(rng[i])) for i in order]


# ---- rows ----

class RowAllocator:
    """Runs of rows in power-of-two size classes (2..64) over `capacity` rows, with the HBM budget as the limit on the
    allocated rows' bytes. Freed runs are reused by their class."""

    def __init__(self, capacity: int, bytes_per_row: int, budget_bytes: float):
        self.capacity, self.bytes_per_row, self.budget = int(capacity), int(bytes_per_row), float(budget_bytes)
        self.top = 0
        self.free_runs = {}
        self.allocated = 0

    @staticmethod
    def run_size(ships: int) -> int:
        return 1 << max(1, (int(ships) - 1).bit_length())

    def alloc(self, ships: int):
        size = self.run_size(ships)
        if (self.allocated + size) * self.bytes_per_row > self.budget:
            return None
        runs = self.free_runs.get(size)
        if runs:
            base = runs.pop()
        elif self.top + size <= self.capacity:
            base = self.top
            self.top += size
        else:
            return None
        self.allocated += size
        return base

    def free(self, base: int, ships: int) -> None:
        size = self.run_size(ships)
        self.free_runs.setdefault(size, []).append(int(base))
        self.allocated -= size


# ---- the server ----

# per-field weights of an episode's state digest (odd 64-bit constants: each field's bits count differently)
_DIGEST_W = mix64(np.arange(1, 72, dtype=np.uint64)) | np.uint64(1)


class RolloutServer:
    """A resident park of `capacity` rows serving episodes (`submit(seed)`), one tick of every live episode a `step()`.

    keep: seeds whose whole final state is kept (the others keep a 64-bit digest of it), for checking."""

    def __init__(self, capacity: int, budget_bytes: float = 200e9, block: int = 256, device="cuda",
                 friction: float = 1.0e6, tick_interval_ms: int = 1000, keep=()):
        self.dev = torch.device(device)
        self.park = VecModesPark(int(capacity), friction=friction, tick_interval_ms=tick_interval_ms, device=device,
                                 block=block)
        cap, d = int(capacity), self.dev
        self.e_seed = torch.zeros(cap, dtype=torch.int64, device=d)         # uint64 bits
        self.e_ship = torch.zeros(cap, dtype=torch.int32, device=d)
        self.e_size = torch.zeros(cap, dtype=torch.int32, device=d)
        self.e_base = torch.zeros(cap, dtype=torch.int64, device=d)
        self.e_admit = torch.zeros(cap, dtype=torch.int64, device=d)
        self.e_next = torch.zeros(cap, dtype=torch.int64, device=d)
        self.op = torch.zeros(cap, dtype=torch.uint8, device=d)
        self.point = torch.zeros(cap, 3, dtype=torch.float64, device=d)
        self.target = torch.zeros(cap, dtype=torch.int64, device=d)
        self.range = torch.zeros(cap, dtype=torch.float64, device=d)
        self.min_range = torch.full((cap,), WARP_MIN_RANGE, dtype=torch.float64, device=d)
        self.warp_factor = torch.full((cap,), WARP_FACTOR, dtype=torch.int64, device=d)
        pk = self.park
        per_row = (pk._f.shape[0] * 8 + pk._q.shape[0] * 8 + pk._b.shape[0] + pk._sf.shape[0] * 8 + 1 + 4 + sum(
            t.element_size() * (t.numel() // cap) for t in (
                self.e_seed, self.e_ship, self.e_size, self.e_base, self.e_admit, self.e_next, self.op, self.point,
                self.target, self.range, self.min_range, self.warp_factor)))   # the blocks, scratch, row map, these
        self.bytes_per_row = per_row
        self.rows = RowAllocator(cap, per_row, budget_bytes)
        self.queue = deque()
        self.live = {}                      # seed -> [base, ships, length, admit step, arrival step, slot]
        self.l_slot = np.zeros(0, dtype=np.int64)      # the live episodes in launch (admission) order: slot numbers,
        self.l_base = np.zeros(0, dtype=np.int64)      # first rows
        self.l_ships = np.zeros(0, dtype=np.int64)     # and ship counts
        self.slots = 0
        self.ending = {}                    # step -> seeds whose last tick it is
        self.step_no = 0
        self.keep = set(int(s) for s in keep)
        self.results = {}                   # seed -> (digest, arrival, admit, finish)
        self.kept = {}                      # seed -> (f64 [NF, ships], i64 [NI, ships], u8 [NB, ships])
        self._dirty = True
        self.world_ticks = 0
        self._e = N.hip() if self.dev.type == "cuda" else None
        if self._e is not None:
            self._e.xb_serve_policy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64,
                                                ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                                ctypes.c_void_p]

    # -- queue and admission --

    def submit(self, seed: int, arrival=None, shape=None) -> None:
        """Queues an episode; shape: its (ships, length) if the caller has them (episode_shape)."""
        seed = int(seed) & _M64
        if shape is None:
            shape = tuple(int(x[0]) for x in episode_shape([seed]))
        self.queue.append((seed, self.step_no if arrival is None else int(arrival), int(shape[0]), int(shape[1])))

    def _admit(self) -> list:
        batch = []
        while self.queue:
            seed, arrival, ships, length = self.queue[0]
            if seed in self.live:
                raise ValueError(f"episode {seed} is already live")
            base = self.rows.alloc(ships)
            if base is None:
                break                                   # first come, first served: the head waits for room
            self.queue.popleft()
            self.live[seed] = [base, ships, length, self.step_no, arrival, self.slots]
            self.slots += 1
            self.ending.setdefault(self.step_no + length - 1, []).append(seed)
            batch.append(seed)
        if batch:
            info = np.array([self.live[x] for x in batch], dtype=np.int64)
            self.l_slot = np.concatenate((self.l_slot, info[:, 5]))
            self.l_base = np.concatenate((self.l_base, info[:, 0]))
            self.l_ships = np.concatenate((self.l_ships, info[:, 1]))
            self._load(batch)
            self._dirty = True
        return batch

    def _load(self, seeds: list) -> None:
        """The admitted episodes' initial rows, made by a staging park's AddBall (its clamps and decay) and copied
        into the episodes' runs of the resident park; their policy state set."""
        p, d = self.park, self.dev
        ships = np.array([self.live[s][1] for s in seeds], dtype=np.int64)
        bases = np.array([self.live[s][0] for s in seeds], dtype=np.int64)
        r = episode_rows(np.array(seeds, dtype=np.uint64), ships)
        m = int(ships.sum())
        rows = torch.from_numpy(np.repeat(bases, ships) + r["ship"]).to(d)
        stage = VecModesPark(m, friction=p.friction, tick_interval_ms=p.tick_interval, device=d, block=p.block)
        t = lambda a, dt=torch.float64: torch.from_numpy(np.ascontiguousarray(a)).to(device=d, dtype=dt)
        world = np.repeat(np.arange(len(seeds), dtype=np.int64), ships)
        stage.load_arrays(ids=t(r["id"], torch.int64), world=t(world, torch.int64),
                          time=torch.zeros(m, dtype=torch.int64, device=d), mass=t(r["mass"]),
                          radius=t(r["radius"]), max_vel=t(r["max_vel"]), pos=t(r["pos"]), vel=t(r["vel"]),
                          agility=t(r["agility"]), speed_fraction=torch.ones(m, dtype=torch.float64, device=d),
                          free=torch.ones(m, dtype=torch.bool, device=d),
                          massive=torch.ones(m, dtype=torch.bool, device=d),
                          interactive=t(r["interactive"], torch.bool),
                          cloaked=torch.zeros(m, dtype=torch.int64, device=d),
                          bubble=torch.zeros(m, dtype=torch.int64, device=d),
                          moribund=torch.zeros(m, dtype=torch.bool, device=d),
                          mode=torch.full((m,), 2, dtype=torch.int64, device=d))
        stage.world = t(np.repeat(np.array(seeds, dtype=np.uint64).view(np.int64), ships), torch.int64)
        p._f[:, rows] = stage._f[:, :m]
        p._q[:, rows] = stage._q[:, :m]
        p._b[:, rows] = stage._b[:, :m]
        self.e_seed[rows] = t(np.repeat(np.array(seeds, dtype=np.uint64).view(np.int64), ships), torch.int64)
        self.e_ship[rows] = t(r["ship"], torch.int32)
        self.e_size[rows] = t(np.repeat(ships, ships), torch.int32)
        self.e_base[rows] = t(np.repeat(bases, ships), torch.int64)
        self.e_admit[rows] = self.step_no
        self.e_next[rows] = 1

    def _relaunch(self) -> None:
        """The row map (live episodes' rows in launch order) and the tiles packing whole episodes."""
        p = self.park
        if not self.l_slot.size:
            p.tiles, p.rowmap, p._state = None, None, None
            self.live_rows = None
            self._dirty = False
            return
        bases = self.l_base
        ships = self.l_ships.astype(np.int32)
        total = int(ships.sum())
        offs = np.concatenate(([0], np.cumsum(ships, dtype=np.int64)[:-1]))
        tiles = np.zeros(len(ships) + 1, dtype=np.int32)
        nt = ctypes.c_int64()
        rc = N.cpu().xb_cpu_pack_tiles(ships.ctypes.data, len(ships), p.block, tiles.ctypes.data, ctypes.byref(nt))
        if rc:
            raise ValueError(f"an episode is wider than the block ({p.block} rows)")
        shift = torch.from_numpy(bases - offs).to(self.dev)
        rowmap = (torch.repeat_interleave(shift, torch.from_numpy(ships.astype(np.int64)).to(self.dev),
                                          output_size=total) + torch.arange(total, device=self.dev)).to(torch.int32)
        p.rowmap = rowmap
        p.tiles = torch.from_numpy(tiles[:nt.value + 1].copy()).to(self.dev)
        p._state = None
        self.live_rows = rowmap
        self._dirty = False

    # -- one step --

    def _policy_host(self) -> None:
        """xb_serve_policy for a park in host memory (the same draws, numpy)."""
        r = self.live_rows.to(torch.int64).numpy()
        seed = self.e_seed.numpy()[r].view(np.uint64)
        ship = self.e_ship.numpy()[r].astype(np.int64)
        t = self.step_no - self.e_admit.numpy()[r] + 1
        nxt = self.e_next.numpy()
        go = t == nxt[r]
        self.op.zero_()
        if not go.any():
            return
        r, seed, ship, t = r[go], seed[go], ship[go], t[go]
        size = self.e_size.numpy()[r].astype(np.int64)
        nxt[r] = t + 1 + (key(seed, F_GAP, ship, t) % np.uint64(10)).astype(np.int64)
        self.op.numpy()[r] = np.asarray(POLICY_OPS, dtype=np.uint8)[(key(seed, F_OP, ship, t) % np.uint64(5)).astype(
            np.int64)]
        pt = self.point.numpy()
        for a in range(3):
            pt[r, a] = (2.0 * u01(key(seed, F_POINT + a, ship, t)) - 1.0) * POINT_HALF
        other = (key(seed, F_TARGET, ship, t) % np.maximum(size - 1, 1).astype(np.uint64)).astype(np.int64)
        other = np.where(other >= ship, other + 1, other)
        self.target.numpy()[r] = self.e_base.numpy()[r] + np.minimum(other, size - 1)
        self.range.numpy()[r] = 500.0 + u01(key(seed, F_RANGE, ship, t)) * 19500.0

    def _policy(self) -> None:
        if self._e is None:
            self._policy_host()
            self._commands()
            return
        e = N.EpisodeRows(self.e_seed.data_ptr(), self.e_ship.data_ptr(), self.e_size.data_ptr(),
                          self.e_base.data_ptr(), self.e_admit.data_ptr(), self.e_next.data_ptr())
        self.op.zero_()
        n_live = self.live_rows.numel()
        N.check(self._e.xb_serve_policy(ctypes.byref(e), self.live_rows.data_ptr(), n_live, self.step_no,
                                        self.op.data_ptr(), self.point.data_ptr(), self.target.data_ptr(),
                                        self.range.data_ptr(), torch.cuda.current_stream(self.dev).cuda_stream),
                "xb_serve_policy")
        self._commands()

    def _commands(self) -> None:
        """The step's commands, one launch over the live rows (each row has at most one)."""
        self.park.commands(self.op, self.point, self.target, self.range, self.min_range, self.warp_factor,
                           rows=self.live_rows)

    def _retire(self) -> list:
        done = self.ending.pop(self.step_no, [])
        if not done:
            return done
        p = self.park
        ships = np.array([self.live[s][1] for s in done], dtype=np.int64)
        bases = np.array([self.live[s][0] for s in done], dtype=np.int64)
        k = np.arange(int(ships.sum())) - np.repeat(np.concatenate(([0], np.cumsum(ships)[:-1])), ships)
        rows = torch.from_numpy(np.repeat(bases, ships) + k).to(self.dev)
        base_of = torch.from_numpy(np.repeat(bases, ships)).to(self.dev)
        f = p._f[:, rows]
        q = p._q[:, rows].clone()
        b = p._b[:N.NB, rows]
        q[5] = torch.where(q[5] >= 0, q[5] - base_of, q[5])                  # follow pointers relative to the run
        bits = torch.cat((f.view(torch.int64), q, b.to(torch.int64)), dim=0)   # [NF + NI + NB, rows]
        w = torch.from_numpy(_DIGEST_W[:bits.shape[0]].view(np.int64)).to(self.dev)
        shipw = torch.from_numpy(mix64(k.astype(np.uint64) + np.uint64(977)).view(np.int64)).to(self.dev)
        per_row = (bits * w[:, None]).sum(0) * shipw
        which = torch.from_numpy(np.repeat(np.arange(len(done)), ships)).to(self.dev)
        dig = torch.zeros(len(done), dtype=torch.int64, device=self.dev).index_add_(0, which, per_row).cpu().numpy()
        offs = np.concatenate(([0], np.cumsum(ships)))
        keep_idx = [i for i, s in enumerate(done) if s in self.keep]
        if keep_idx:
            fc, qc, bc = f.cpu(), q.cpu(), b.cpu()
        slots = []
        for i, s in enumerate(done):
            base, n, length, admit, arrival, slot = self.live.pop(s)
            slots.append(slot)
            self.results[s] = (int(dig[i]) & _M64, arrival, admit, self.step_no)
            if i in keep_idx:
                sl = slice(int(offs[i]), int(offs[i + 1]))
                self.kept[s] = (fc[:, sl].numpy(), qc[:, sl].numpy(), bc[:, sl].numpy())
            self.rows.free(base, n)
        keep = np.ones(self.l_slot.size, dtype=bool)
        keep[np.searchsorted(self.l_slot, np.array(slots, dtype=np.int64))] = False
        self.l_slot, self.l_base, self.l_ships = self.l_slot[keep], self.l_base[keep], self.l_ships[keep]
        self._dirty = True
        return done

    def step(self) -> dict:
        """Admit what fits, apply this tick's policy commands, tick every live episode once (one launch), retire the
        episodes that have run their length. Returns the step's counts."""
        admitted = self._admit()
        if self._dirty:
            self._relaunch()
        n_live = int(self.l_slot.size)
        if n_live:
            self._policy()
            self.park._tick()
            self.world_ticks += n_live
        retired = self._retire()
        self.step_no += 1
        return {"admitted": len(admitted), "live": n_live, "retired": len(retired), "queued": len(self.queue)}

    def idle(self) -> bool:
        return not self.queue and not self.live


def poisson_arrivals(seed: int, episodes: int, rate_per_step: float) -> np.ndarray:
    """The serving step each episode arrives at: a Poisson stream of `rate_per_step` episodes a step (exponential gaps,
    seeded), in order."""
    rng = np.random.default_rng(seed)
    return np.floor(np.cumsum(rng.exponential(1.0 / rate_per_step, episodes))).astype(np.int64)


def serve(server: RolloutServer, seeds, arrivals, max_steps: int | None = None, on_step=None) -> dict:
    """Runs the stream to the end: each step submits the episodes arriving at it, then steps the server; per-step
    latency (host wall time of the step, its device work synchronised) recorded."""
    seeds = np.asarray(seeds, dtype=np.uint64)
    arrivals = np.asarray(arrivals, dtype=np.int64)
    ships, length = episode_shape(seeds)
    i, lat = 0, []
    t_start = time.perf_counter()
    while i < len(seeds) or not server.idle():
        while i < len(seeds) and arrivals[i] <= server.step_no:
            server.submit(int(seeds[i]), int(arrivals[i]), (int(ships[i]), int(length[i])))
            i += 1
        t0 = time.perf_counter()
        info = server.step()
        if server.dev.type == "cuda":
            torch.cuda.synchronize(server.dev)
        lat.append(time.perf_counter() - t0)
        if on_step is not None:
            on_step(info)
        if max_steps is not None and server.step_no >= max_steps:
            break
    wall = time.perf_counter() - t_start
    lat_ms = np.array(lat) * 1000.0
    return {"steps": server.step_no, "wall_s": wall, "world_ticks": server.world_ticks,
            "world_ticks_per_s": server.world_ticks / wall if wall > 0 else 0.0,
            "step_ms": {"p50": float(np.percentile(lat_ms, 50)), "p99": float(np.percentile(lat_ms, 99)),
                        "max": float(lat_ms.max())},
            "episodes_done": len(server.results)}
