"""The display orientation without dynamical orientation on gfx950 (xb_collide_tiled_flight_ypr, xb_modes_tiled_flight_ypr,
the general launch's step C, xb_ypr_snap_kernel), bit for bit (two NaNs equal):

- against the reference's own FullBallpark vectors (tests/golden/ypr_golden.npz: flight and collision worlds, every
  ball as added and after every tick, with the collision lists);
- the flight worlds without collisions against the oracle, tiled and general launch;
- at size: 8,192 collision worlds (4 ships, 4 fixed spheres) and 65,536 flight worlds against the host build, and at
  full size a run repeated, several ticks a launch and the HIP graph's replays equal to eager ticks.
"""

import pytest
import torch

import golden_fixture as G
import park_driver as D
import scenarios as S
from exact_ballpark_amd import workloads as WL
from exact_ballpark_amd.vec_modes import ModesGraph, VecModesPark
from test_ypr import compare_golden

pytestmark = pytest.mark.gpu
DEV = "cuda"


def test_park_matches_reference_vectors():
    golden = G.load("ypr_golden.npz")
    kept = golden["kept"]
    got = D.run_park(golden["worlds"], DEV, order=kept, initial=True)
    bad = compare_golden(got, golden, kept)
    assert not bad, f"{len(bad)} (world, step) records differ, first {bad[:6]}"


@pytest.mark.parametrize("general", [False, True])
def test_flight_worlds_without_collisions_match_oracle(general):
    ws = [w for w in S.cases(1) if w["kind"] != "guards"]
    for w in ws:
        w["ypr"] = True
    idx = list(range(0, len(ws), 2))
    got = D.run_park(ws, DEV, order=idx, general=general, initial=True)
    bad = D.compare(got, D.run_oracle(ws, idx, initial=True), ws)
    assert not bad, f"{len(bad)} differ: {bad}"


def _blocks(park):
    return [b.detach().cpu().clone() for b in park._blocks()]


def _same(a, b):
    return all(torch.equal(x.view(torch.uint8), y.view(torch.uint8)) for x, y in zip(a, b))


def _load(wl, device, collisions):
    park = VecModesPark(wl["n"], device=device, collisions=collisions, display_ypr=True)
    WL.load_flight_park(park, wl)
    return park


@pytest.mark.parametrize("collisions", [True, False])
def test_at_size_matches_host(collisions):
    wl = WL.collide_park(17, 8192) if collisions else WL.flight_park(17, 65536)
    g, h = _load(wl, DEV, collisions), _load(wl, "cpu", collisions)
    for t in range(24):
        g.evolve()
        h.evolve()
        if t == 11:                     # SetBallVelocity's snap on a third of the rows, mid-run
            mask = torch.arange(wl["n"]) % 3 == 0
            v = torch.full((wl["n"], 3), 40.0, dtype=torch.float64)
            g.set_velocity(mask.to(DEV), v.to(DEV))
            h.set_velocity(mask, v)
    assert _same(_blocks(g), _blocks(h))
    assert float((h.roll != 0.0).double().mean()) > 0.3


def test_full_size_repeat_resident_and_graph():
    wl = WL.collide_park(19, 65536)
    a, b, c, d = (_load(wl, DEV, True) for _ in range(4))
    for _ in range(8):
        a.evolve()
        d.evolve()
    b.evolve_ticks(4)
    b.evolve_ticks(4)
    g = ModesGraph(c, ticks=2, launches=2)
    g.tick()
    g.tick()
    torch.cuda.synchronize()
    sa = _blocks(a)
    assert _same(sa, _blocks(b)) and _same(sa, _blocks(c)) and _same(sa, _blocks(d))
