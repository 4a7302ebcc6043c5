"""OnTick's masked tick in the kernels (`evolve(active=...)`: the rows outside the mask skipped, not ticked and put
back), on the host build: a row outside the mask keeps every field bit for bit, a row inside it equals the unmasked
tick's, on the composed park (orientation, missiles, collisions), the flight park with the display orientation, and
the general (untiled) launch."""

import torch

from exact_ballpark_amd import workloads as WL
from exact_ballpark_amd.vec_modes import VecModesPark


def _park(wl, **kw):
    p = VecModesPark(wl["n"], device="cpu", dynamical_orientation=bool(wl.get("dynamical")),
                     collisions=bool(wl.get("collisions")), **kw)
    WL.load_flight_park(p, wl)
    return p


def _bits(p):
    return [b[:, :p.n].clone().view(torch.uint8) if b.dtype != torch.uint8 else b[:, :p.n].clone()
            for b in p._blocks()]


def _compare(before, after, full, sel):
    for a, b, f in zip(before, after, full):
        cols = sel.repeat_interleave(a.shape[1] // sel.numel())
        assert torch.equal(b[:, ~cols], a[:, ~cols]), "a row outside the mask changed"
        assert torch.equal(b[:, cols], f[:, cols]), "a row inside the mask differs from the unmasked tick"


def _check(wl, ticks=5, **kw):
    full, masked = _park(wl, **kw), _park(wl, **kw)
    for t in range(ticks):
        active = (masked.world % 3 != t % 3)
        before = _bits(masked)
        masked.evolve(active=active)
        full_before = [b.clone() for b in full._blocks()]
        full.evolve()
        _compare(before, _bits(masked), _bits(full), active)
        for blk, old, cur in zip(full._blocks(), full_before, masked._blocks()):   # the park it continues from
            blk.copy_(cur)


def test_masked_composed_park():
    _check(WL.composed_park(4, 300, missiles=8))


def test_masked_flight_park_display_orientation():
    _check(WL.flight_park(4, 300, 4), display_ypr=True)


def test_masked_general_launch():
    wl = WL.flight_park(5, 200, 4)
    p = _park(wl)
    p.tiles = None                           # the general launch: any row order, one tick
    q = _park(wl)
    q.tiles = None
    p._state = q._state = None
    active = p.world % 2 == 0
    before = _bits(p)
    p.evolve(active=active)
    q.evolve()
    _compare(before, _bits(p), _bits(q), active)
