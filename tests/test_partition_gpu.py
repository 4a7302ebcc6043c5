"""The partition park on gfx950: a small park against the oracle world by world after every tick (eagerly and as
SpaceGraph replays), and a park of 2,048 worlds (about 70,000 ships) as SpaceGraph replays against the host build,
every table of the park and the layer bit for bit, every tick's events included, over 16 ticks."""

import pytest
import torch

import test_partition as TP
from exact_ballpark_amd import workloads as WL
from exact_ballpark_amd.vec_space import SpaceGraph

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("graph", [False, True], ids=["eager", "graph"])
def test_partition_park_matches_oracle_on_gpu(graph):
    wl = WL.partition_park(3, 48, 4, 40)
    print(f"partition park on the GPU: 48 worlds, {wl['n']} ships, graph={graph}", flush=True)
    got, ws = TP.run_product(wl, "cuda", 24, graph=graph)
    bad = TP.compare_oracle(got, ws, range(48), 24)
    assert not bad, f"{len(bad)} worlds differ, first {bad[:6]}"


def _bits(t: torch.Tensor) -> torch.Tensor:
    t = t.cpu()
    return t.view(torch.int64) if t.dtype == torch.float64 else t.view(torch.int32) if t.dtype == torch.float32 else t


def _diff(host, gpu) -> list:
    out = [f"park block {i}" for i, (a, b) in enumerate(zip(host[0]._blocks(), gpu[0]._blocks()))
           if not torch.equal(_bits(a), _bits(b))]
    return out + [k for k in host[1].st if not torch.equal(_bits(host[1].st[k]), _bits(gpu[1].st[k]))]


def test_partition_park_gpu_matches_host_at_scale():
    wl = WL.partition_park(11, 2048)
    print(f"partition park: 2048 worlds, {wl['n']} ships, 16 ticks as SpaceGraph replays", flush=True)
    host, gpu = TP.build(wl, "cpu", 1024), TP.build(wl, "cuda", 1024)
    assert not _diff(host, gpu), "after the load"
    g = SpaceGraph(gpu[1])
    events = 0
    for s in range(1, 17):
        host[1].tick()
        g.tick()
        d = _diff(host, gpu)
        assert not d, f"tick {s}: {d[:6]}"
        events += int(host[1].st["evn"].sum())
        host[1].clear_events()
        gpu[1].clear_events()
    host[1].check()
    gpu[1].check()
    assert events > 1000
