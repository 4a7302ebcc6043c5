"""The rollout-serving loop on the CPU (the host build of the tick, the host restatement of the policy):

- the draws: SplitMix64's published outputs; the episode law's ranges;
- the row allocator: power-of-two runs, reuse, the HBM budget's refusal;
- a mix of episodes arriving at different steps, each finished episode's final state bit-exact against the oracle;
- each episode alone gives the same final state and digest as in the mix;
- a budget that holds two episodes at a time queues the rest first come, first served, with the same results.
"""

import numpy as np
import pytest

import serve_check as C
from exact_ballpark_amd import serve as SV


def test_splitmix64_known_outputs():
    """splitmix64's first outputs from state 0 (Vigna's splitmix64.c: next() adds the golden gamma, then mixes)."""
    g = 0x9E3779B97F4A7C15
    got = [int(SV.mix64(np.array([(g * k) & ((1 << 64) - 1)], dtype=np.uint64))[0]) for k in (1, 2, 3)]
    assert got == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


def test_episode_law():
    seeds = np.arange(1, 20001, dtype=np.uint64)
    ships, length = SV.episode_shape(seeds)
    assert ships.min() == 2 and ships.max() == 64 and length.min() == 50 and length.max() <= 2000
    assert abs(length.mean() - 1025) < 20 and abs(ships.mean() - 33) < 1
    r = SV.episode_rows(seeds[:50], ships[:50])
    assert (r["mass"] >= 1e6).all() and (r["mass"] < 1e8).all() and (np.abs(r["pos"]) <= 5e4).all()
    cmds = SV.episode_commands(int(seeds[0]), int(ships[0]), 200)
    ticks = [c[0] for c in cmds if c[2] == 0]
    assert ticks[0] == 1 and all(1 <= b - a <= 10 for a, b in zip(ticks, ticks[1:]))
    assert all(c[3] != c[2] and 0 <= c[3] < ships[0] for c in cmds)


def test_row_allocator():
    a = SV.RowAllocator(capacity=256, bytes_per_row=100, budget_bytes=100 * 128)
    assert a.run_size(2) == 2 and a.run_size(3) == 4 and a.run_size(33) == 64 and a.run_size(64) == 64
    b0 = a.alloc(33)
    b1 = a.alloc(40)
    assert (b0, b1) == (0, 64) and a.allocated == 128
    assert a.alloc(2) is None                       # the budget is spent
    a.free(b0, 33)
    assert a.alloc(60) == 0                         # the run reused


def _run(seeds, arrivals, capacity=512, budget=200e9, keep=True):
    srv = SV.RolloutServer(capacity, budget_bytes=budget, device="cpu", block=64,
                           keep=[int(s) for s in seeds] if keep else ())
    stats = SV.serve(srv, seeds, arrivals)
    return srv, stats


@pytest.fixture(scope="module")
def mix():
    seeds = C.short_seeds(10, 90, max_ships=12)
    srv, stats = _run(seeds, np.arange(len(seeds)) * 4)
    return seeds, srv, stats


def test_mix_matches_oracle(mix):
    seeds, srv, stats = mix
    assert stats["episodes_done"] == len(seeds)
    assert max(r[2] for r in srv.results.values()) > 0              # episodes admitted at different steps
    for s in seeds:
        assert C.same(C.kept_as_oracle(srv.kept[int(s)]), C.oracle_final(int(s))), int(s)


def test_alone_equals_mix(mix):
    seeds, srv, _ = mix
    for s in seeds[:3]:
        alone, _ = _run([s], [0])
        assert C.same_kept(alone.kept[int(s)], srv.kept[int(s)])
        assert alone.results[int(s)][0] == srv.results[int(s)][0]


def test_budget_queues_first_come_first_served(mix):
    seeds, srv, _ = mix
    small, _ = _run(seeds, np.zeros(len(seeds), dtype=np.int64), capacity=64,
                    budget=2 * 16 * srv.bytes_per_row, keep=False)
    assert len(small.results) == len(seeds)
    admits = [small.results[int(s)][2] for s in seeds]
    assert admits == sorted(admits) and admits[-1] > 0
    for s in seeds:
        assert small.results[int(s)][0] == srv.results[int(s)][0]
