"""FullBallpark's static collidables on gfx950 (the collision tick kernels with a static table), bit for bit (two NaNs
equal):

- against the reference's own vectors (tests/golden/static_golden.npz with the display yaw, pitch and roll,
  static_do_golden.npz with dynamical orientation): every ball as added and after every tick, with mLastC, the filing
  and the collision lists;
- the turning-and-collisions park with 8 static collidables a world (exact_ballpark_amd.workloads.collide_park, 4
  ships, 4 fixed spheres, capsules and oriented boxes in turn): 2,048 worlds against the host build after 30 ticks,
  sampled worlds against the oracle; at full size (65,536 worlds) a run repeated, several ticks a launch and the HIP
  graph's replays equal to eager ticks; a world alone equal to the same world inside the batch.
"""

import pytest
import torch

import golden_fixture as G
import park_driver as D
from exact_ballpark_amd import workloads as WL
from exact_ballpark_amd.vec_modes import ModesGraph, VecModesPark
from test_statics import compare_golden

pytestmark = pytest.mark.gpu
DEV = "cuda"


@pytest.mark.parametrize("block", [256, 64])
def test_park_matches_reference_vectors(block):
    golden = G.load("static_golden.npz")
    kept = golden["kept"]
    got = D.run_park(golden["worlds"], DEV, order=kept, initial=True, block=block)
    bad = compare_golden(got, golden, kept)
    assert not bad, f"{len(bad)} (world, step) records differ, first {bad[:6]}"


def test_park_with_orientation_matches_reference_vectors():
    golden = G.load("static_do_golden.npz")
    kept = golden["kept"]
    got = D.run_park(golden["worlds"], DEV, order=kept, initial=True, block=128)
    bad = compare_golden(got, golden, kept, do=True)
    assert not bad, f"{len(bad)} (world, step) records differ, first {bad[:6]}"


def _blocks(park):
    return [b.detach().cpu().clone() for b in park._blocks()]


def _same(a, b):
    return all(torch.equal(x.view(torch.uint8), y.view(torch.uint8)) for x, y in zip(a, b))


def _load(wl, device):
    park = VecModesPark(wl["n"], device=device, collisions=True, dynamical_orientation=wl.get("dynamical", False))
    WL.load_flight_park(park, wl)
    return park


def test_statics_park_matches_host_and_oracle():
    wl = WL.collide_park(7, 2048, dynamical=True, statics=8)
    g, h = _load(wl, DEV), _load(wl, "cpu")
    for _ in range(30):
        g.evolve()
        h.evolve()
    assert _same(_blocks(g), _blocks(h))
    hits = h._hits[:, :h.n]
    static_hits = int(((hits >= 1001) & (hits <= 1008) & (torch.arange(hits.shape[0])[:, None] < h.hits_n)).sum())
    assert static_hits > 20
    for wi in range(0, 2048, 131):
        sc = WL.world_scenario(wl, wi, 30)
        w = D.run_oracle([sc], [0], record={30})[(0, 30)]
        got = D._snapshot(g, [wi], {wi: list(WL.world_rows(wl, wi))})[wi]
        assert D._same_bits(got[0], w[0]).all() and (got[1] == w[1]).all() and got[2] == w[2], wi


def test_full_size_repeat_resident_and_graph():
    wl = WL.collide_park(11, 65536, dynamical=True, statics=8)
    a, b, c = _load(wl, DEV), _load(wl, DEV), _load(wl, DEV)
    for _ in range(8):
        a.evolve()
    b.evolve_ticks(4)
    b.evolve_ticks(4)
    g = ModesGraph(c, ticks=2, launches=2)
    g.tick()
    g.tick()
    torch.cuda.synchronize()
    sa = _blocks(a)
    assert _same(sa, _blocks(b)) and _same(sa, _blocks(c))
    a2 = _load(wl, DEV)
    for _ in range(8):
        a2.evolve()
    assert _same(sa, _blocks(a2))


def test_world_alone_equals_batch():
    wl = WL.collide_park(13, 4096, statics=8)
    full = _load(wl, DEV)
    for _ in range(20):
        full.evolve()
    for wi in (0, 1234, 4095):
        one = WL.select_worlds(wl, [wi])
        one.update(ypr=None, dynamical=False, collisions=True)
        p = _load(one, DEV)
        for _ in range(20):
            p.evolve()
        rows = list(WL.world_rows(wl, wi))
        for x, y in zip(_blocks(full), _blocks(p)):
            assert torch.equal(x[..., rows].view(torch.uint8), y[..., :len(rows)].view(torch.uint8))
