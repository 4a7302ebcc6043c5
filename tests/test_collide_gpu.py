"""The gradient collision method on gfx950 (xb_collide_tiled_kernel) against the oracle and the host build, bit for
bit (two NaNs equal):

- every ball of every collision world (tests/scenarios.py collide_cases) after every tick: the flight fields, mLastC,
  the filing and the collision list, in workgroups of 256 and 64 rows;
- the turning-and-collisions park's bodies (exact_ballpark_amd.workloads.collide_park, 4 ships and 4 fixed spheres a
  world): 2,048 worlds against the host build after 30 ticks, sampled worlds against the oracle; at full size (65,536
  worlds) a run repeated is identical, several ticks a launch equal one a launch, and the HIP graph's replays equal
  eager ticks;
- a world alone equal to the same world inside the full batch.
"""

import numpy as np
import pytest
import torch

import park_driver as D
import scenarios as S
from exact_ballpark_amd import workloads as WL
from exact_ballpark_amd.vec_modes import ModesGraph, VecModesPark

pytestmark = pytest.mark.gpu
DEV = "cuda"


@pytest.fixture(scope="module")
def worlds():
    return S.collide_cases(1)


@pytest.fixture(scope="module")
def want(worlds):
    return D.run_oracle(worlds, range(len(worlds)))


@pytest.mark.parametrize("block", [256, 64])
def test_collision_worlds_match_oracle(worlds, want, block):
    got = D.run_park(worlds, DEV, block=block)
    bad = D.compare(got, want, worlds)
    assert not bad, f"{len(bad)} differ: {bad}"


def _blocks(park):
    return [b.detach().cpu().clone() for b in park._blocks()]


def _same(a, b):
    return all(torch.equal(x.view(torch.uint8), y.view(torch.uint8)) for x, y in zip(a, b))


def _load(wl, device, **kw):
    park = VecModesPark(wl["n"], device=device, collisions=True, dynamical_orientation=wl["dynamical"], **kw)
    WL.load_flight_park(park, wl)
    return park


def test_collide_park_matches_host_and_oracle():
    wl = WL.collide_park(7, 2048)
    g, h = _load(wl, DEV), _load(wl, "cpu")
    for _ in range(30):
        g.evolve()
        h.evolve()
    assert _same(_blocks(g), _blocks(h))
    assert int((h.hits_n > 0).sum()) > 200
    for wi in range(0, 2048, 97):
        sc = WL.world_scenario(wl, wi, 30)
        w = D.run_oracle([sc], [0], record={30})[(0, 30)]
        got = D._snapshot(g, [wi], {wi: list(WL.world_rows(wl, wi))})[wi]
        assert D._same_bits(got[0], w[0]).all() and (got[1] == w[1]).all() and got[2] == w[2], wi


def test_full_size_repeat_resident_and_graph():
    wl = WL.collide_park(11, 65536)
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
    assert int((a.hits_n > 0).sum()) > 5000


def test_world_alone_equals_batch():
    wl = WL.collide_park(13, 4096)
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


def test_collision_worlds_with_orientation_match_oracle():
    w = S.do_collide_cases(1)
    got = D.run_park(w, DEV, block=128)
    bad = D.compare(got, D.run_oracle(w, range(len(w))), w)
    assert not bad, f"{len(bad)} differ: {bad}"


def test_composed_park_host_oracle_and_padded():
    """The ragged composed park (ragged worlds of 1..32 massive ships and 0..8 missiles launched at tick 1, dynamical
    orientation, collisions): the device against the host build after 20 ticks, sampled worlds against the oracle, and
    the padded layout's real rows equal to the packed layout's (follow pointers mapped through the padding)."""
    wl = WL.composed_park(5, 4096, missiles=8)
    g, h = _load(wl, DEV), _load(wl, "cpu")
    for _ in range(20):
        g.evolve()
        h.evolve()
    assert _same(_blocks(g), _blocks(h))
    for wi in range(0, 4096, 211):
        sc = WL.world_scenario(wl, wi, 20)
        w = D.run_oracle([sc], [0], record={20})[(0, 20)]
        got = D._snapshot(g, [wi], {wi: list(WL.world_rows(wl, wi))})[wi]
        assert D._same_bits(got[0], w[0]).all() and (got[1] == w[1]).all() and got[2] == w[2], wi
    pad = WL.padded(wl, max(32, int(wl["counts"].max())))
    p = _load(pad, DEV)
    for _ in range(20):
        p.evolve()
    rows = torch.as_tensor(pad["dest"])
    back = torch.full((pad["n"],), -1, dtype=torch.int64)
    back[rows] = torch.arange(wl["n"])
    for k, (x, y) in enumerate(zip(_blocks(g)[:3], _blocks(p)[:3])):
        x, y = x[..., :wl["n"]], y[..., rows]
        if k == 1:                                  # follow pointers: rows of each layout
            fp = y[5]
            y = y.clone()
            y[5] = torch.where(fp >= 0, back[fp.clamp(min=0)], fp)
        assert torch.equal(x.view(torch.uint8), y.view(torch.uint8)), k
