"""The gradient collision method through the drop-in park (VecModesPark(collisions=True)), host build of the kernel
source, against the oracle (itself pinned to the reference's SpaceBallpark vectors, tests/test_oracle_collide_golden.py):
every ball of every collision world (tests/scenarios.py collide_cases) after every tick, bit for bit: the flight
fields, mLastC, the filing (level, key, sides) and the collision list in order. Also: workgroups of 64 and 128 rows,
the lanes sorted by branch class, several ticks a launch, OnTick's masked tick, and the refusals (a world past 64
rows, the general launch)."""

import numpy as np
import pytest
import torch

import park_driver as D
import scenarios as S
from exact_ballpark_amd.vec_modes import VecModesPark


@pytest.fixture(scope="module")
def worlds():
    return S.collide_cases(1)


@pytest.fixture(scope="module")
def want(worlds):
    return D.run_oracle(worlds, range(len(worlds)))


@pytest.mark.parametrize("block", [256, 64])
def test_collision_worlds_match_oracle(worlds, want, block):
    got = D.run_park(worlds, "cpu", block=block)
    bad = D.compare(got, want, worlds)
    assert not bad, f"{len(bad)} differ: {bad}"


def test_collision_worlds_with_orientation_match_oracle():
    """Dynamical orientation with collisions (the oracle pinned to the reference's FullBallpark on these worlds)."""
    w = S.do_collide_cases(1)
    got = D.run_park(w, "cpu", block=128)
    bad = D.compare(got, D.run_oracle(w, range(len(w))), w)
    assert not bad, f"{len(bad)} differ: {bad}"


def test_collision_worlds_in_reverse_order(worlds, want):
    got = D.run_park(worlds, "cpu", order=list(reversed(range(len(worlds)))))
    assert not D.compare(got, want, worlds)


def _park(worlds, idx, **kw):
    balls = [(wi, b) for wi in idx for b in worlds[wi]["balls"]]
    park = VecModesPark(len(balls), device="cpu", collisions=True, **kw)
    park.load([{"id": b["id"], "world": wi, "time": 0, "mass": b["mass"], "radius": b["radius"],
                "max_vel": b["max_vel"], "pos": b["pos"], "vel": b["vel"], "agility": b["agility"],
                "speed_fraction": b["sf"], "free": b["free"], "massive": b["massive"], "junk": b["junk"],
                "cloaked": b["cloaked"]} for wi, b in balls])
    goto = torch.zeros(park.n, dtype=torch.bool)
    pts = torch.zeros(park.n, 3, dtype=torch.float64)
    for k, (wi, b) in enumerate(balls):            # every free ball toward the centre of its world
        c = np.mean([x["pos"] for x in worlds[wi]["balls"]], axis=0)
        goto[k] = b["free"]
        pts[k] = torch.tensor(2 * c - np.array(b["pos"]))
    park.goto_point(goto, pts)
    return park


def _blocks(park):
    return [b.clone() for b in park._blocks()]


def test_resident_ticks_and_lane_order_equal_plain(worlds):
    idx = [wi for wi, w in enumerate(worlds) if w["friction"] == 1e6 and w["tick"] == 1000][:60]
    a, b, c = (_park(worlds, idx) for _ in range(3))
    for _ in range(12):
        a.evolve()
    b.evolve_ticks(12)
    c.schedule = 1
    for _ in range(12):
        c.evolve()
    for x, y in zip(_blocks(a), _blocks(b)):
        assert torch.equal(x.view(torch.uint8) if x.dtype == torch.float64 else x,
                           y.view(torch.uint8) if y.dtype == torch.float64 else y)
    for x, y in zip(_blocks(a), _blocks(c)):
        assert torch.equal(x, y) or torch.equal(x.view(torch.int64), y.view(torch.int64))
    assert int(a.hits_n.sum()) > 0


def test_masked_tick_keeps_other_worlds(worlds):
    idx = [wi for wi, w in enumerate(worlds) if w["friction"] == 1e6 and w["tick"] == 1000][:40]
    park = _park(worlds, idx)
    before = _blocks(park)
    active = (park.world % 2 == 0)
    park.evolve(active=active)
    after = _blocks(park)
    keep = ~active
    for x, y in zip(before, after):
        xs, ys = x[..., :park.n], y[..., :park.n]
        assert torch.equal(xs[..., keep].view(torch.uint8), ys[..., keep].view(torch.uint8))


def test_refusals(worlds):
    big = [{"id": k + 1, "world": 0, "mass": 1e6, "radius": 10.0, "max_vel": 100.0, "pos": [k * 50.0, 0.0, 0.0],
            "vel": [0.0, 0.0, 0.0], "agility": 1.0, "speed_fraction": 1.0} for k in range(65)]
    park = VecModesPark(65, device="cpu", collisions=True)
    with pytest.raises(ValueError):
        park.load(big)
    park = _park(worlds, [0, 1])
    park.tiles = None
    park._state = None
    with pytest.raises(RuntimeError):
        park.evolve()


def test_collision_lists_and_fixture_coverage(worlds, want):
    """The worlds meet every kind of contact the reference's Potential has: impacts on fixed and free balls, overlaps
    pushed apart through the recursion, several contacts of one ball in one tick."""
    lists = [h for (d, q, ls) in want.values() for (_, h) in ls]
    assert sum(len(h) > 2 for h in lists) > 10 and sum(len(h) == 1 for h in lists) > 100
    d = np.concatenate([v[0] for v in want.values()])
    assert (d[:, 26:29] != 0.0).any(axis=1).sum() > 300
