"""The learned scheduler's host side and its oracle, on the CPU:

- the network oracle's gradient against central finite differences of its own loss (float64), every parameter block;
- the oracle's Adam step against its closed form at step 1 (the update is lr * g / (|g| + eps));
- the oracle's features on a hand-computed world;
- a regrouped world order (rowmap and tiles) ticks every world to the same bits as the row order, on the host build of
  the tick, for a random order and for the ascending-cost order, with one and several ticks a launch;
- install_order's refusals, and the cost model's preference for grouped worlds.
"""

import numpy as np
import pytest
import torch

import mlp_oracle as M
from exact_ballpark_amd import scheduler as S
from exact_ballpark_amd import workloads as WL
from exact_ballpark_amd.vec_modes import VecModesPark


def _net(seed, n=24):
    rng = np.random.default_rng(seed)
    p = rng.normal(0.0, 0.4, M.NPARAM)
    x = rng.normal(0.0, 1.0, (n, M.NI))
    y = rng.integers(0, M.NO, n)
    return p, x, y


def test_oracle_gradient_matches_finite_differences():
    p, x, y = _net(1)
    g = M.grad(p, x, y)
    rng = np.random.default_rng(2)
    offs = np.cumsum([0] + [int(np.prod(s)) for _, s in M.SIZES])
    picks = np.concatenate([rng.choice(np.arange(offs[k], offs[k + 1]), 8, replace=False) for k in range(6)])
    h = 1e-6
    for i in picks:
        e = np.zeros_like(p)
        e[i] = h
        fd = (M.loss(p + e, x, y) - M.loss(p - e, x, y)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-7 + 1e-5 * abs(g[i]), (i, fd, g[i])


def test_oracle_gradient_of_a_dead_relu_is_zero():
    """A first-layer unit whose input is negative for every sample passes no gradient to its weights."""
    p, x, y = _net(3)
    q = M.unpack(p)
    q["b1"][5] = -1e6
    g = M.unpack(M.grad(p, x, y))
    assert np.all(g["W1"][:, 5] == 0.0) and g["b1"][5] == 0.0 and np.all(g["W2"][5, :] == 0.0)


def test_oracle_adam_first_step_closed_form():
    p, x, y = _net(4)
    g = M.grad(p, x, y)
    lr = 1e-3
    p1, m, v = M.adam(p, g, np.zeros_like(p), np.zeros_like(p), 1, lr)
    assert np.allclose(p1, p - lr * g / (np.abs(g) + 1e-8), rtol=0, atol=1e-12)
    assert np.allclose(m, 0.1 * g) and np.allclose(v, 0.001 * g * g)


def test_oracle_features_hand_computed():
    pos = np.array([[0.0, 0.0, 0.0], [2.0, 0.0, 0.0], [5.0, 5.0, 5.0]])
    mode = np.array([0, 3, 2])                        # goto, warp (aligning), stop
    stamp = np.array([0, -1, 0])
    free = np.array([1, 1, 0], dtype=np.uint8)
    mori = np.zeros(3, dtype=np.uint8)
    path = np.array([2, 2, 0], dtype=np.uint8)
    f = M.features(pos, mode, stamp, free, mori, path, np.array([0, 2]), np.array([2, 1]))
    assert np.allclose(f[0], [0, 0.5, 0, 0, 0.5, 2 / 32, 0.5, 0, np.log10(2) / 10, np.log10(2) / 10, 2 / 38,
                              4 / 2 / 9])
    assert np.allclose(f[1], [1, 0, 0, 0, 0, 1 / 32, 0, 1, 0, 0, 0, 0])


def _cpu_park(seed, worlds, mix, block=128):
    wl = WL.ragged_park(seed, worlds, mix=mix)
    park = VecModesPark(wl["n"], device="cpu", block=block)
    WL.load_flight_park(park, wl)
    return park


def _bits(park):
    return [b.view(torch.int64) if b.dtype == torch.float64 else b for b in (park._f.clone(), park._q.clone(),
                                                                               park._b.clone())]


@pytest.mark.parametrize("mix,ticks", [("uniform", 1), ("profiles", 4)])
def test_regrouped_order_ticks_to_the_same_bits(mix, ticks):
    T = 12
    plain = _cpu_park(5, 1500, mix)
    for _ in range(T // ticks):
        plain.evolve_ticks(ticks)
    ref = _bits(plain)
    for how in ("random", "cost"):
        park = _cpu_park(5, 1500, mix)
        park.record_paths(True)
        park.schedule = 1
        start, count = S.world_ranges(park)
        park.evolve()
        if how == "random":
            order = torch.randperm(start.numel(), generator=torch.Generator().manual_seed(9))
        else:
            order = torch.sort(S.world_cost(park, start, count), stable=True).indices
        info = S.install_order(park, start, count, order)
        assert info["tiles"] == park.tiles.numel() - 1 and park.rowmap is not None
        assert not torch.equal(park.rowmap, torch.arange(park.n, dtype=torch.int32))
        left = T - 1
        while left:
            k = min(ticks, left)
            park.evolve_ticks(k)
            left -= k
        for a, b in zip(ref, _bits(park)):
            assert torch.equal(a, b), how


def test_install_order_refuses_a_bad_order_and_restores():
    park = _cpu_park(6, 200, "uniform")
    start, count = S.world_ranges(park)
    W = start.numel()
    with pytest.raises(ValueError):
        S.install_order(park, start, count, torch.cat((torch.arange(W - 1), torch.tensor([0]))))
    S.install_order(park, start, count, torch.arange(W).flip(0))
    assert park.rowmap is not None
    S.install_order(park, start, count, None)
    assert park.rowmap is None and park.tiles is not None


def test_cost_model_prefers_grouped_worlds():
    """On worlds of differing activity, ordering by divergence cost lowers the modeled launch cost."""
    park = _cpu_park(7, 4000, "profiles", block=256)
    park.record_paths(True)
    park.schedule = 1
    for _ in range(3):
        park.evolve()
    start, count = S.world_ranges(park)
    base = S.modeled_cost(park)
    S.install_order(park, start, count, torch.sort(S.world_cost(park, start, count), stable=True).indices)
    assert S.modeled_cost(park) < 0.7 * base
