"""OnTick's masked tick in the gfx950 kernels (the rows outside the mask skipped by the tiled, collision and general
launches, and by a masked ModesGraph's replays): a row outside the mask keeps every field bit for bit, a row inside it
equals the unmasked tick's, on the composed park at 16,384 worlds, the flight park with the display orientation and
the general launch."""

import pytest
import torch

from exact_ballpark_amd import workloads as WL
from exact_ballpark_amd.vec_modes import ModesGraph, VecModesPark

pytestmark = pytest.mark.gpu
DEV = "cuda"


def _park(wl, **kw):
    p = VecModesPark(wl["n"], device=DEV, dynamical_orientation=bool(wl.get("dynamical")),
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


@pytest.mark.parametrize("kind", ["composed", "flight_ypr"])
def test_masked_tick_eager_and_graph(kind):
    if kind == "composed":
        wl, kw = WL.composed_park(6, 16384, missiles=8), {"block": 64}
    else:
        wl, kw = WL.flight_park(6, 16384, 4), {"display_ypr": True}
    print(f"masked tick, {kind}: {wl['n']} rows", flush=True)
    full, masked, graphed = _park(wl, **kw), _park(wl, **kw), _park(wl, **kw)
    g = ModesGraph(graphed, masked=True)
    for t in range(4):
        active = (masked.world % 3 != t % 3)
        before = _bits(masked)
        masked.evolve(active=active)
        g.tick(active)
        full.evolve()
        torch.cuda.synchronize()
        _compare(before, _bits(masked), _bits(full), active)
        for x, y in zip(_bits(masked), _bits(graphed)):
            assert torch.equal(x, y), "the masked graph differs from the masked eager tick"
        for blk, cur in zip(full._blocks(), masked._blocks()):
            blk.copy_(cur)


def test_masked_general_launch():
    wl = WL.flight_park(7, 8192, 4)
    p, q = _park(wl), _park(wl)
    p.tiles = q.tiles = None
    p._state = q._state = None
    active = p.world % 2 == 0
    before = _bits(p)
    p.evolve(active=active)
    q.evolve()
    torch.cuda.synchronize()
    _compare(before, _bits(p), _bits(q), active)
