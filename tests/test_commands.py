"""The one-launch command batch (VecModesPark.commands, xb_modes_commands) against the park's masked commands applied op
by op (stop, goto_point, follow, orbit, warp_to), bit for bit over every state field, on random commands with the
cases each command refuses or turns: NaN and infinite points and ranges, targets in another world, the row itself, out
of the park, moribund or cloaked targets, short warps (under 100 km: a GotoPoint), a warp to the point already being
warped to, negative and NaN minimum ranges, warp factors below 1. On the host build, and (marked gpu) on the device."""

import numpy as np
import pytest
import torch

from exact_ballpark_amd import workloads as WL
from exact_ballpark_amd.vec_modes import VecModesPark

STOP, GOTO, FOLLOW, ORBIT, WARP = 3, 1, 4, 5, 6


def _park(dev, seed):
    wl = WL.flight_park(seed, 600, 5)
    park = VecModesPark(wl["n"], device=dev, block=256)
    WL.load_flight_park(park, wl)
    for _ in range(3):
        park.evolve()
    n = park.n
    rng = np.random.default_rng(seed)
    park.moribund = torch.from_numpy(rng.uniform(size=n) < 0.05).to(dev)
    park.cloaked = torch.from_numpy((rng.uniform(size=n) < 0.05).astype(np.int64)).to(dev)
    return park


def _batch(park, seed):
    n = park.n
    rng = np.random.default_rng(seed + 100)
    op = rng.choice([0, STOP, GOTO, FOLLOW, ORBIT, WARP], n).astype(np.uint8)
    pos = park.pos.cpu().numpy()
    point = rng.uniform(-1e10, 1e10, (n, 3))
    near = rng.uniform(size=n) < 0.15
    point[near] = pos[near] + rng.uniform(-5e4, 5e4, (near.sum(), 3))
    same = rng.uniform(size=n) < 0.1
    point[same] = park.goto.cpu().numpy()[same]
    bad = rng.uniform(size=n) < 0.05
    point[bad, rng.integers(0, 3, bad.sum())] = rng.choice([np.nan, np.inf, -np.inf], bad.sum())
    world = park.world.cpu().numpy()
    target = (np.arange(n) // 5) * 5 + rng.integers(0, 5, n)
    odd = rng.uniform(size=n)
    target = np.where(odd < 0.05, -1, np.where(odd < 0.1, n + 3, np.where(odd < 0.15, rng.integers(0, n, n), target)))
    rr = rng.uniform(500.0, 20000.0, n)
    rr[rng.uniform(size=n) < 0.05] = np.nan
    rr[rng.uniform(size=n) < 0.03] = 1e300                     # rounds to an infinite float: refused
    mr = rng.choice([20000.0, 0.0, -5.0, np.nan, 1e6], n)
    wf = rng.choice([20, 0, -3, 1, 7], n).astype(np.int64)
    assert (world >= 0).all()
    t = lambda a, dt: torch.from_numpy(np.ascontiguousarray(a)).to(device=park.dev, dtype=dt)
    return (t(op, torch.uint8), t(point, torch.float64), t(target, torch.int64), t(rr, torch.float64),
            t(mr, torch.float64), t(wf, torch.int64))


def _op_by_op(park, op, point, target, rr, mr, wf):
    park.stop(op == STOP)
    park.goto_point(op == GOTO, point)
    park.follow(op == FOLLOW, target, rr)
    park.orbit(op == ORBIT, target, rr)
    park.warp_to(op == WARP, point, mr, wf)


def _bits(park):
    return [park._f.view(torch.int64).clone(), park._q.clone(), park._b.clone()]


def _check(dev, seed, rows_subset):
    a, b = _park(dev, seed), _park(dev, seed)
    batch = _batch(a, seed)
    op = batch[0]
    rows = None
    if rows_subset:
        keep = torch.from_numpy(np.random.default_rng(seed).uniform(size=a.n) < 0.5).to(dev)
        rows = torch.nonzero(keep).reshape(-1).to(torch.int32)
        op = torch.where(keep, op, torch.zeros_like(op))
    _op_by_op(a, op, *batch[1:])
    b.commands(batch[0] if rows_subset else op, *batch[1:], rows=rows)
    for x, y in zip(_bits(a), _bits(b)):
        assert torch.equal(x, y)
    changed = (a.mode != _park(dev, seed).mode).sum().item()
    assert changed > a.n // 4


@pytest.mark.parametrize("seed,subset", [(1, False), (2, True)])
def test_commands_equal_op_by_op_host(seed, subset):
    _check("cpu", seed, subset)


@pytest.mark.gpu
@pytest.mark.parametrize("seed,subset", [(3, False), (4, True)])
def test_commands_equal_op_by_op_device(seed, subset):
    _check("cuda", seed, subset)
