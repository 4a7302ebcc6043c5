"""The reference's collision park on the row park: VecCollidePark (src/exact_ballpark/vec_collide.py:431-1051), W
worlds of N ball slots, each world's balls in its first slots by ascending id (GetNearbyBalls' order), its state as
[W, N, ...] tensors, GOTO, GotoDirection, STOP and the setters, the gradient collision method (Destiny's partition
filing and Potential) and static collidables.

The park is a VecModesPark of W x N rows with the gradient method on, slot s of world w at row w N + s, an empty slot an
empty row. Every [W, N] attribute is a view of the row park's blocks (a world's N rows consecutive, so a field's first
W N entries split into [W, N] without a copy): reading it costs nothing and writing into it writes the park. Commands
reshape their [W, N] masks and arguments to rows and call the row park's own.

Not built, and refused at construction: the iterative collision method (`iterative=True`), free balls restricted to
the first `rows` slots, minis (local ids) and global balls.
"""

from __future__ import annotations

import torch

from .vec_modes import COLL_FIELDS, EMPTY_ID, FIELD_NAMES, YPR_FIELDS, VecModesPark

I64, BOOL = torch.int64, torch.bool
_NAMES = {"id": "ids", "mod": "side"}                   # the reference's names for the row park's
_VIEWS = frozenset(FIELD_NAMES) | frozenset(f[0] for f in COLL_FIELDS) | frozenset(f[0] for f in YPR_FIELDS) | \
    frozenset(_NAMES)


class VecCollidePark:
    """W worlds of N ball slots (a slot holds a ball or nothing); in each world the balls fill the first slots in
    ascending id (the reference's constructor and API)."""

    def __init__(self, worlds: int, slots: int, friction: float = 1.0e6, tick_interval_ms: int = 1000,
                 iterative: bool = False, collisions: bool = True, contact_slots: int | None = None,
                 locations: int = 48, device="cpu", rows: int | None = None, impact_slots: int | None = None,
                 kernel: str = "torch", statics: int = 0, minis: bool = False, fields: bool = False,
                 global_balls: bool = False):
        if kernel not in ("torch", "cuda"):
            raise ValueError(f"kernel: 'torch' or 'cuda', not {kernel!r}")
        if iterative:
            raise NotImplementedError("the iterative collision method is not built")
        if rows is not None and int(rows) != int(slots):
            raise NotImplementedError("free balls restricted to the first `rows` slots are not built")
        if minis or global_balls:
            raise NotImplementedError("minis and global balls in the [W, N] park are not built")
        if not 1 <= int(slots) <= 64:
            raise ValueError(f"slots={slots}: from 1 to 64 (a world's contacts are one wavefront's mask)")
        W, N = int(worlds), int(slots)
        object.__setattr__(self, "W", W)
        object.__setattr__(self, "N", N)
        object.__setattr__(self, "S", int(statics))
        object.__setattr__(self, "fields", bool(fields))
        object.__setattr__(self, "kernel", kernel)
        object.__setattr__(self, "contact_slots", contact_slots)     # every contact is taken here: no limit
        object.__setattr__(self, "park", VecModesPark(W * N, friction, tick_interval_ms, device,
                                                      collisions=bool(collisions), block=256))
        object.__setattr__(self, "dev", self.park.dev)

    # ---- the [W, N] views ----

    def __getattr__(self, name):
        if name in _VIEWS:
            t = getattr(self.park, _NAMES.get(name, name))
            if not torch.is_tensor(t):
                return t
            return t.reshape(self.W, self.N, *t.shape[1:])
        if name == "hits":
            h = self.park.hits
            return h.reshape(self.W, self.N, h.shape[1])
        raise AttributeError(name)

    def __setattr__(self, name, value):
        if name in _VIEWS:
            v = torch.as_tensor(value)
            if v.dim() >= 2:
                v = v.reshape(self.W * self.N, *v.shape[2:])
            setattr(self.park, _NAMES.get(name, name), v)
            return
        object.__setattr__(self, name, value)

    @property
    def exists(self) -> torch.Tensor:
        return (self.park.ids != EMPTY_ID).reshape(self.W, self.N)

    @property
    def current_time(self) -> int:
        return self.park.current_time

    @current_time.setter
    def current_time(self, t: int) -> None:
        self.park.current_time = int(t)
        self.park.clock = torch.full((self.park.n,), int(t), dtype=I64, device=self.dev)

    @property
    def dt(self) -> float:
        return self.park.dt

    def use_kernel(self, kernel: str) -> None:
        """The reference's choice of its torch path or its CUDA kernels: both are this park's one tick here."""
        if kernel not in ("torch", "cuda"):
            raise ValueError(f"kernel: 'torch' or 'cuda', not {kernel!r}")
        object.__setattr__(self, "kernel", kernel)

    # ---- loading ----

    def load(self, worlds: list, statics: list | None = None) -> None:
        """Each world a list of balls (AddBall's arguments: id, mass, radius, max_vel, agility, speed_fraction, pos,
        vel, free, massive, and optionally junk and cloaked), at most N a world, into its first slots by ascending id;
        `statics` each world's static collidables (load_statics)."""
        if len(worlds) != self.W:
            raise ValueError(f"{len(worlds)} worlds for a park of {self.W}")
        rows = []
        t = self.park.current_time
        for w, balls in enumerate(worlds):
            if len(balls) > self.N:
                raise ValueError(f"world {w}: {len(balls)} balls for {self.N} slots")
            for b in sorted(balls, key=lambda b: int(b["id"])):
                rows.append({"id": int(b["id"]), "world": w, "time": t, "mass": b["mass"], "radius": b["radius"],
                             "max_vel": b["max_vel"], "pos": b["pos"], "vel": b["vel"], "agility": b["agility"],
                             "speed_fraction": b["speed_fraction"], "free": bool(b.get("free", True)),
                             "massive": bool(b.get("massive", True)), "interactive": False,
                             "cloaked": int(bool(b.get("cloaked", False))), "bubble": -1, "moribund": False,
                             "junk": bool(b.get("junk", False))})
            rows += [VecModesPark.empty_row(w, t) for _ in range(self.N - len(balls))]
        self.park.load(rows)
        if statics is not None:
            self.load_statics(statics)

    def load_statics(self, statics: list) -> None:
        """Each world's static collidables (at most `statics` a world, as the park was made)."""
        if any(len(s) > self.S for s in statics):
            raise ValueError(f"more static collidables in a world than the park's {self.S} slots")
        self.park.load_statics({w: s for w, s in enumerate(statics) if s})

    # ---- commands ([W, N] masks and arguments) ----

    def _m(self, mask: torch.Tensor) -> torch.Tensor:
        return mask.to(device=self.dev, dtype=BOOL).reshape(self.W * self.N)

    def _x(self, x: torch.Tensor, k: int = 0) -> torch.Tensor:
        x = torch.as_tensor(x, device=self.dev)
        return x.reshape(self.W * self.N, k) if k else x.reshape(self.W * self.N)

    def goto_point(self, mask: torch.Tensor, point: torch.Tensor) -> None:
        self.park.goto_point(self._m(mask), self._x(point, 3))

    def goto_direction(self, mask: torch.Tensor, direction: torch.Tensor) -> None:
        self.park.goto_direction(self._m(mask), self._x(direction, 3))

    def stop(self, mask: torch.Tensor) -> None:
        self.park.stop(self._m(mask))

    def set_speed_fraction(self, mask: torch.Tensor, s: torch.Tensor) -> None:
        self.park.set_speed_fraction(self._m(mask), self._x(s))

    def set_mass(self, mask: torch.Tensor, mass: torch.Tensor) -> None:
        self.park.set_mass(self._m(mask), self._x(mass))

    def set_agility(self, mask: torch.Tensor, agility: torch.Tensor) -> None:
        self.park.set_agility(self._m(mask), self._x(agility))

    def set_radius(self, mask: torch.Tensor, r: torch.Tensor) -> None:
        self.park.set_radius(self._m(mask), self._x(r))

    def set_max_velocity(self, mask: torch.Tensor, v: torch.Tensor) -> None:
        self.park.set_max_velocity(self._m(mask), self._x(v))

    def set_velocity(self, mask: torch.Tensor, v: torch.Tensor) -> None:
        self.park.set_velocity(self._m(mask), self._x(v, 3))

    def set_harmonic(self, mask: torch.Tensor, harmonic: torch.Tensor, corporation: torch.Tensor,
                     alliance: torch.Tensor, is_field: torch.Tensor) -> None:
        if not self.fields:
            raise ValueError("set_harmonic: a park made with fields=True")
        self.park.set_harmonic(self._m(mask), self._x(harmonic), self._x(corporation), self._x(alliance),
                               self._x(is_field))

    # ---- the tick ----

    def evolve(self, active: torch.Tensor | None = None, skip: torch.Tensor | None = None) -> None:
        """One tick in every world, or in the worlds where `active` [W] (the others kept bit for bit). `skip` (the
        master's dead bubbles) is the partition layer's: refused here."""
        if skip is not None:
            raise NotImplementedError("skip: the master's dead bubbles are the space layer's (VecSpace)")
        if active is None:
            self.park.evolve()
        else:
            act = active.to(device=self.dev, dtype=BOOL).reshape(self.W)
            self.park.evolve(active=act.repeat_interleave(self.N))

    def collision_lists(self) -> list:
        """[W][N] each slot's collision list of its last tick."""
        flat = self.park.collision_lists()
        return [flat[w * self.N:(w + 1) * self.N] for w in range(self.W)]
