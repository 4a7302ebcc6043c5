"""The reference's combat and composed parks on the row park: VecCombatPark (missiles, formations, cloaking, mushrooms,
removals and balls added mid-run, its DoCollision events and its domain flags; src/exact_ballpark/vec_combat.py:
147-1120) and VecComposedPark (the same with the gradient method's collisions between every massive ball;
vec_composed.py:183-194), with their constructors, `load` rows (an empty row: "exists": False), `lookup(rows, ball)`,
`add_balls`, `events`, `check`, `watch` and `world_flags`.

Both are a VecModesPark with the gradient method on (the tiled launch: each world's rows consecutive, at most 64): a
missile's one candidate is its target, as the combat park's contact test has it, and the collisions between other
massive balls are FullBallpark's gradient, which the composed park runs. The combat park's domain has no such
collisions: `watch()` flags a world holding two massive balls outside MISSILE mode, as the reference's does, and
`check()` raises on any flag.
"""

from __future__ import annotations

import torch

from .vec_modes import EMPTY_ID, MISSILE, VecModesPark

I64, BOOL = torch.int64, torch.bool
# the domain's flags (the reference's FLAG_* bits and names, vec_combat.py:96-127)
FLAG_MASSIVE, FLAG_BUBBLES, FLAG_DEAD, FLAG_POOL, FLAG_UNDEFINED = 1, 2, 4, 8, 16
FLAG_NAMES = {FLAG_MASSIVE: "two massive balls outside MISSILE mode in a world (their contact is not computed)",
              FLAG_BUBBLES: "balls of one world in more than one bubble (master)",
              FLAG_DEAD: "a world without an interactive ball (master: a dead bubble)",
              FLAG_POOL: "add_balls into a row that is not empty, or an id already in its world",
              FLAG_UNDEFINED: "Destiny's undefined behaviour"}


def _block(b: int) -> int:
    """The reference's workgroup size as this park's (a multiple of 64, at most 256)."""
    return max(64, min(256, (int(b) + 63) // 64 * 64))


class VecCombatPark(VecModesPark):
    """VecModesPark with MISSILE and FORMATION, the missiles' contacts and DoCollision events, cloaking at the park,
    AddBall and RemoveBall mid-run (the reference's VecCombatPark). `disable_do_for_missiles` (FullBallpark's missiles
    snapped to their velocity instead of turning) is not built: refused."""

    _DOMAIN_COLLISIONS = False              # the combat park's domain: no contact but a missile's with its target

    def __init__(self, n: int, friction: float = 1.0e6, tick_interval_ms: int = 1000, device="cuda",
                 dynamical_orientation: bool = False, use_new_orbit: bool = False, block: int = 128,
                 master: bool = False, disable_do_for_missiles: bool = False, ypr: str = "host"):
        if disable_do_for_missiles:
            raise NotImplementedError("disable_do_for_missiles: missiles snapped to their velocity are not built")
        super().__init__(n, friction, tick_interval_ms, device, dynamical_orientation=dynamical_orientation,
                         use_new_orbit=use_new_orbit, block=_block(block), ypr=ypr, collisions=True,
                         display_ypr=not dynamical_orientation, master=master)
        self.tick_ms = int(tick_interval_ms)
        self.disable_do_for_missiles = False
        self.flags = torch.zeros(self.n, dtype=I64, device=self.dev)

    # ---- loading ----

    def load(self, rows: list) -> None:
        """VecModesPark.load, every row a ball or, with "exists": False, an empty row of its world (inert until
        add_balls or add_mushroom fills it); then the world index (rows_of_world, world_index, labels)."""
        out = []
        for r in rows:
            if r.get("exists", True):
                out.append({k: v for k, v in r.items() if k != "exists"})
            else:
                out.append(VecModesPark.empty_row(int(r.get("world", 0)), int(r.get("time", 0))))
        super().load(out)
        self.flags = torch.zeros(self.n, dtype=I64, device=self.dev)
        self._index_worlds()

    def _index_worlds(self) -> None:
        labels = self.world.cpu().tolist()
        order, groups = [], {}
        for r, w in enumerate(labels):
            if w not in groups:
                groups[w] = []
                order.append(w)
            groups[w].append(r)
        self.labels = order
        width = max((len(g) for g in groups.values()), default=0)
        self.rows_of_world = torch.tensor([g + [-1] * (width - len(g)) for g in (groups[w] for w in order)],
                                          dtype=I64, device=self.dev).reshape(len(order), width)
        index = {w: k for k, w in enumerate(order)}
        self.world_index = torch.tensor([index[w] for w in labels], dtype=I64, device=self.dev)

    @property
    def exists(self) -> torch.Tensor:
        """[n] the row holds a ball (not an empty row)."""
        return self.ids != EMPTY_ID

    @property
    def mod(self) -> torch.Tensor:
        """The filing's side on each axis (the reference's name for `side`)."""
        return self.side

    def lookup(self, rows: torch.Tensor, ball: torch.Tensor | None = None) -> torch.Tensor:
        """For each listed row (rows [R]), the row of its own world holding the ball with Destiny id ball[i], or -1
        (the reference's lookup(rows, ball)). With one argument, the row park's: ball[row] for every row."""
        if ball is None:
            return super().lookup(rows)
        rows = rows.to(device=self.dev, dtype=I64).reshape(-1)
        probe = torch.full((self.n,), EMPTY_ID, dtype=I64, device=self.dev)
        out = torch.full((rows.numel(),), -1, dtype=I64, device=self.dev)
        ball = ball.to(device=self.dev, dtype=I64).reshape(-1)
        for k in range(rows.numel()):                     # rows may repeat: one probe each
            probe.fill_(EMPTY_ID)
            probe[rows[k]] = ball[k]
            out[k] = super().lookup(probe)[rows[k]]
        return out

    def add_balls(self, mask: torch.Tensor, rows: list) -> None:
        """AddBall into the empty rows in `mask` (VecModesPark.add_balls); a refusal also flags the rows (FLAG_POOL)."""
        try:
            super().add_balls(mask, rows)
        except ValueError:
            self.flags = self.flags | torch.where(mask.to(self.dev).reshape(self.n), FLAG_POOL, 0)
            raise
        self._index_worlds()

    # ---- the tick's events and the domain ----

    def events(self) -> list:
        """The last tick's DoCollision events (COL-719) as the reference lists them: per world (its label, in
        first-appearance order) the missiles by ascending id, each list in its order: (world label, time, missile id,
        collider id)."""
        order = {w: k for k, w in enumerate(self.labels)}
        ev = self.collision_events()
        ev.sort(key=lambda e: order.get(e[0], 0))           # stable: the balls' order within a world kept
        return [(e[0], e[1], e[2], e[3]) for e in ev]

    def watch(self) -> None:
        """The combat park's domain condition at a tick's start and end: a world holding two massive balls outside
        MISSILE mode is flagged on all its rows (their contact is the composed park's)."""
        if self._DOMAIN_COLLISIONS:
            return
        m = self.exists & ~self.moribund & self.massive & (self.mode != MISSILE)
        rw = self.rows_of_world
        safe = rw.clamp(min=0)
        two = (m[safe] & (rw >= 0)).sum(dim=1) > 1
        self.flags = self.flags | torch.where(two[self.world_index], FLAG_MASSIVE, 0)

    def world_flags(self) -> list:
        """Each world's flags (the OR of its rows'), by `labels` order."""
        rw = self.rows_of_world
        safe = rw.clamp(min=0)
        f = torch.where(rw >= 0, self.flags[safe], 0).cpu().tolist()
        out = []
        for row in f:
            v = 0
            for x in row:
                v |= int(x)
            out.append(v)
        return out

    def check(self) -> None:
        """Raise if a world left the domain: the flags, and on the master a world with balls in more than one bubble
        or without an interactive ball."""
        flags = 0
        for v in set(self.flags.cpu().tolist()):
            flags |= int(v)
        if self.master:
            rw = self.rows_of_world
            safe = rw.clamp(min=0)
            ok = (rw >= 0) & self.exists[safe]
            b = self.bubble[safe]
            first = torch.where(ok, b, torch.iinfo(torch.int64).max).min(dim=1).values
            if bool((ok & (b != first[:, None])).any()):
                flags |= FLAG_BUBBLES
            if bool((ok.any(dim=1) & ~(ok & self.interactive[safe]).any(dim=1)).any()):
                flags |= FLAG_DEAD
        if flags:
            names = [v for k, v in FLAG_NAMES.items() if flags & k]
            raise RuntimeError(f"outside the batched domain: {', '.join(names)}")


class VecComposedPark(VecCombatPark):
    """VecCombatPark with the gradient method's collisions between every massive ball, missiles included (the
    reference's VecComposedPark). `collisions=False` is the combat park's domain (no contact but a missile's with its
    target, flagged by `watch`)."""

    def __init__(self, n: int, friction: float = 1.0e6, tick_interval_ms: int = 1000, device="cpu",
                 dynamical_orientation: bool = False, use_new_orbit: bool = False, block: int = 128,
                 master: bool = False, disable_do_for_missiles: bool = False, collisions: bool = True):
        super().__init__(n, friction, tick_interval_ms, device, dynamical_orientation, use_new_orbit, block, master,
                         disable_do_for_missiles)
        self._DOMAIN_COLLISIONS = bool(collisions)
