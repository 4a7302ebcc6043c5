"""The learned scheduler's gfx950 kernels against the oracle (oracle/mlp_oracle.py), and what the trained network buys:

- each world's features (xb_world_features) against the oracle's, within 1e-6;
- the forward pass (logits) and the backward pass (mean loss, gradient of every parameter) against the oracle's float64
  restatement, within stated float32 tolerances; the backward pass repeated is bit-identical (no atomics);
- one Adam step against the oracle's, from the same gradient;
- training on the park's own recorded ticks lowers the loss and beats the majority bucket;
- the learned regrouping ticks every world to the same bits as the unsorted order (the unsorted order is the reference);
- the trained network's regrouping lowers the modeled launch cost and the measured time per tick against the unsorted
  order on the same worlds (the modeled cost and the measured median time asserted).
"""

import numpy as np
import pytest
import torch

import mlp_oracle as M
from exact_ballpark_amd import scheduler as S
from exact_ballpark_amd import workloads as WL
from exact_ballpark_amd.vec_modes import ModesGraph, VecModesPark

pytestmark = pytest.mark.gpu
DEV = "cuda"


def _park(seed, worlds, mix="profiles", block=256):
    wl = WL.ragged_park(seed, worlds, mix=mix)
    park = VecModesPark(wl["n"], device=DEV, block=block)
    WL.load_flight_park(park, wl)
    return park


def _batch(seed, n):
    rng = np.random.default_rng(seed)
    net = S.Network(DEV, seed)
    x = rng.normal(0.0, 1.0, (n, M.NI)).astype(np.float32)
    y = rng.integers(0, M.NO, n).astype(np.int32)
    return net, x, y


def test_features_match_oracle():
    park = _park(11, 3000)
    park.record_paths(True)
    for _ in range(3):
        park.evolve()
    start, count = S.world_ranges(park)
    got = S.features(park, start, count).cpu().numpy()
    torch.cuda.synchronize()
    want = M.features(park.pos.cpu().numpy(), park.mode.cpu().numpy(), park.effect_stamp.cpu().numpy(),
                      park.free.cpu().numpy(), park.moribund.cpu().numpy(), park.path.cpu().numpy(),
                      start.cpu().numpy(), count.cpu().numpy())
    assert np.abs(got - want).max() <= 1e-6


@pytest.mark.parametrize("n", [1, 300, 8192])
def test_forward_and_gradient_match_oracle(n):
    net, x, y = _batch(n, n)
    p = net.params.cpu().numpy()
    xd, yd = torch.from_numpy(x).to(DEV), torch.from_numpy(y).to(DEV)
    logits = net.forward(xd)[0].cpu().numpy()
    _, _, z = M.forward(p, x)
    assert np.abs(logits - z).max() <= 1e-5 * max(1.0, np.abs(z).max())
    loss, g = net.gradient(xd, yd)
    g = g.cpu().numpy()
    gw = M.grad(p, x, y)
    assert abs(float(loss) - M.loss(p, x, y)) <= 1e-5 * max(1.0, M.loss(p, x, y))
    # float32 sums of n terms against float64, block by block (a wrong index or operand errs by the gradient's size)
    k = 0
    for name, shape in M.SIZES:
        m = int(np.prod(shape))
        err = np.abs(g[k:k + m] - gw[k:k + m]).max()
        assert err <= 1e-3 * np.abs(gw[k:k + m]).max() + 1e-9, (name, err)
        k += m
    loss2, g2 = net.gradient(xd, yd)
    assert torch.equal(torch.from_numpy(g), g2.cpu()) and float(loss2) == float(loss)


def test_adam_step_matches_oracle():
    net, x, y = _batch(5, 1024)
    xd, yd = torch.from_numpy(x).to(DEV), torch.from_numpy(y).to(DEV)
    p = net.params.cpu().numpy().astype(np.float64)
    m = np.zeros_like(p)
    v = np.zeros_like(p)
    for step in (1, 2, 3):
        _, g = net.gradient(xd, yd)
        gh = g.cpu().numpy().astype(np.float64)
        net.adam(g, 1e-3)
        p, m, v = M.adam(p, gh, m, v, step, 1e-3)
        assert np.abs(net.params.cpu().numpy() - p).max() <= 1e-6 * max(1.0, np.abs(p).max())
        p = net.params.cpu().numpy().astype(np.float64)   # carry the device's rounding forward


def test_labels_are_checked():
    net, x, y = _batch(6, 16)
    y[3] = 8
    with pytest.raises(ValueError):
        net.gradient(torch.from_numpy(x).to(DEV), torch.from_numpy(y).to(DEV))


def test_training_learns_the_park():
    park = _park(12, 8192)
    park.schedule = 1
    L = S.LearnedScheduler(park, seed=0)
    x, c = L.collect(6)
    st = L.fit(x, c, epochs=30, batch=4096, lr=3e-3)
    y = L.bucket(c)
    majority = float(torch.bincount(y).max()) / y.numel()
    print("training:", st, "majority", round(majority, 3))
    assert st["loss_last"] < 0.6 * st["loss_first"]
    assert st["accuracy"] > majority + 0.1


def _bits(park):
    return [b.view(torch.int64) if b.dtype == torch.float64 else b.clone() for b in (park._f.clone(), park._q.clone(),
                                                                                       park._b.clone())]


def _ms_per_tick(park, ticks, reps=10, windows=5):
    """The median over `windows` timed windows of `reps` graph replays, ms a tick."""
    g = ModesGraph(park, ticks=ticks)
    g.tick()
    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    got = []
    for _ in range(windows):
        torch.cuda.synchronize()
        s.record()
        for _ in range(reps):
            g.tick()
        e.record()
        torch.cuda.synchronize()
        got.append(s.elapsed_time(e) / (reps * ticks))
    return sorted(got)[len(got) // 2]


def test_learned_regroup_is_exact_and_gains():
    """Train on the park's first ticks; then, from one saved state, the unsorted order and the learned order run the
    same ticks: every field bit-identical, the learned order cheaper in the cost model and in measured time (the median
    of 5 timed windows of 10 replays of 16 ticks each)."""
    park = _park(13, 32768)
    park.schedule = 1
    L = S.LearnedScheduler(park, seed=0)
    L.train(ticks=6, epochs=20, batch=8192)
    saved, t0 = park._save(), park.current_time
    out = {}
    for how in ("unsorted", "learned"):
        for b, old in zip((park._f, park._q, park._b), saved):
            b.copy_(old)
        park.current_time = t0
        if how == "unsorted":
            S.install_order(park, L.start, L.count, None)
        else:
            L.regroup()
        park.evolve()                                    # one tick: paths recorded under this order
        model = S.modeled_cost(park)
        ms = _ms_per_tick(park, 16)                      # timed on a copy of the state, then given back
        for b, old in zip((park._f, park._q, park._b), saved):
            b.copy_(old)
        park.current_time = t0
        for _ in range(3):
            park.evolve_ticks(16)
        torch.cuda.synchronize()
        out[how] = (_bits(park), model, ms)
    for a, b in zip(out["unsorted"][0], out["learned"][0]):
        assert torch.equal(a, b)
    print("modeled cost unsorted", out["unsorted"][1], "learned", out["learned"][1],
          "| ms per tick unsorted", round(out["unsorted"][2], 5), "learned", round(out["learned"][2], 5))
    assert out["learned"][1] < 0.8 * out["unsorted"][1]
    assert out["learned"][2] < out["unsorted"][2], "the learned order is not faster (median of 5 windows)"
