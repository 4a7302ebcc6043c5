"""The rollout-serving loop on the GPU (xb_serve_policy and the tiled tick over the live episodes' row map):

- a Poisson stream of 300 episodes (2..64 ships, at most 200 ticks each) under continuous batching: every finished
  episode's final state bit-exact against the oracle;
- the same stream on the host build (the numpy restatement of the policy, the host tick): every episode's digest equal,
  so the policy kernel and the host policy issue the same commands;
- episodes run alone give the same final state as in the mix;
- a tight HBM budget queues episodes and changes no result.
"""

import numpy as np
import pytest

import serve_check as C
from exact_ballpark_amd import serve as SV

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def stream():
    seeds = C.short_seeds(300, 200, start=7)
    arrivals = SV.poisson_arrivals(3, len(seeds), 12.0)
    srv = SV.RolloutServer(1 << 16, device="cuda", keep=[int(s) for s in seeds])
    stats = SV.serve(srv, seeds, arrivals)
    print("serving:", stats, "peak rows", srv.rows.top)
    return seeds, arrivals, srv


def test_every_episode_matches_oracle(stream):
    seeds, _, srv = stream
    assert len(srv.results) == len(seeds)
    bad = [int(s) for s in seeds if not C.same(C.kept_as_oracle(srv.kept[int(s)]), C.oracle_final(int(s)))]
    assert not bad, bad[:5]


def test_host_build_gives_the_same_digests(stream):
    seeds, arrivals, srv = stream
    host = SV.RolloutServer(1 << 16, device="cpu")
    SV.serve(host, seeds[:120], arrivals[:120])
    for s in seeds[:120]:
        assert host.results[int(s)][0] == srv.results[int(s)][0], int(s)


def test_alone_equals_mix(stream):
    seeds, _, srv = stream
    for s in seeds[::100]:
        alone = SV.RolloutServer(256, device="cuda", keep=[int(s)])
        SV.serve(alone, [s], [0])
        assert C.same_kept(alone.kept[int(s)], srv.kept[int(s)])
        assert alone.results[int(s)][0] == srv.results[int(s)][0]


def test_tight_budget_queues(stream):
    seeds, arrivals, srv = stream
    small = SV.RolloutServer(1 << 12, budget_bytes=512 * srv.bytes_per_row, device="cuda")
    SV.serve(small, seeds[:60], np.zeros(60, dtype=np.int64))
    assert len(small.results) == 60 and max(r[2] for r in small.results.values()) > 0
    for s in seeds[:60]:
        assert small.results[int(s)][0] == srv.results[int(s)][0]
