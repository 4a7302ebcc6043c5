"""The partition server on gfx950 (libxb_hip.so's xb_server_pre_tick with the partition, one lane a world, and the
partition's own kernels) through the C ABI: bit for bit against the reference's own server layer on worlds of several
bubbles (tests/golden/server_space_golden.npz: ball states and bubbles, every client's updates with their ball
streams, the pending queries' answers), at 1,024 worlds against the host build of the same source (the park's
blocks, each call's queued flag, route and bubble, the send flags and narrowcast bubbles, the bubble update's lists,
every partition table, every tick), and the HIP graph of a tick with a fixed schedule of pushes against eager ticks,
replayed with eager work between."""

import pytest
import torch

import server_driver as SD
import server_scenarios as SS
from exact_ballpark_amd import vec_server_space as VSS

pytestmark = pytest.mark.gpu


def test_space_server_gpu_matches_reference_vectors():
    print("partition server golden on cuda:0", flush=True)
    r = SD.compare_golden(False, "cuda:0", space=True)
    assert r["compared"] > 8000
    assert not r["rows"], f"{len(r['rows'])} ball-states differ, first {r['rows'][:4]}"
    assert not r["sends"], f"{len(r['sends'])} world-ticks' sends differ, first {r['sends'][:4]}"
    assert not r["answers"], f"{len(r['answers'])} answers differ, first {r['answers'][:2]}"


def test_space_server_gpu_equals_host_build_1024_worlds():
    worlds = SS.space_server_cases(11, 1024, ticks=8)
    for w in worlds:
        w["ticks"] = [[(e[0], e[1], tuple(e[2])) for e in tick] for tick in w["ticks"]]
    print("partition server 1024 worlds x 8 ticks: cuda:0 against the host build", flush=True)
    gpu = SD.run_streams(worlds, "cuda:0", space=True)
    cpu = SD.run_streams(worlds, "cpu", space=True)
    bad = SD.same_runs(gpu, cpu)
    assert not bad, bad[:8]


def test_space_server_graph_replays_equal_eager_ticks():
    worlds = SS.space_server_cases(21, 512, ticks=1)
    for w in worlds:
        w["ticks"] = [[]]
    eager, _ = SD.build(worlds, "cuda:0", space=True)
    graphed, _ = SD.build(worlds, "cuda:0", space=True)
    W, d = eager.W, eager.dev
    schedule = ["GotoPoint", "SetSpeedFraction", "FollowBall", "SetBallPosition", "Stop", "SetBallVelocity"]
    ids = eager.a.ids_of_world.to(d)
    print("partition server graph: 512 worlds, 6 replays of 6 pushes against eager ticks", flush=True)
    g = VSS.SpaceServerGraph(graphed, schedule)
    gen = torch.Generator(device="cpu").manual_seed(5)
    for tick in range(6):
        args = []
        for k, kind in enumerate(schedule):
            mask = torch.rand(W, generator=gen) < 0.5
            pick = torch.randint(0, 2, (W,), generator=gen)
            ball = ids[torch.arange(W, device=d), pick.to(d)]
            target = ids[torch.arange(W, device=d), (1 - pick).to(d)]
            f = torch.randn(W, 4, generator=gen, dtype=torch.float64) * 5000.0
            if kind == "SetSpeedFraction":
                f = torch.rand(W, 4, generator=gen, dtype=torch.float64)
            if kind == "SetBallPosition":         # now and then to another top box: another bubble
                far = (torch.rand(W, generator=gen) < 0.3).to(torch.float64) * SS.TOP * 2.0
                f = f + far[:, None]
            args.append((kind, mask.to(d), ball, target, f.to(d)))
        for kind, mask, ball, target, f in args:
            eager.push(kind, mask, ball, target, f)
        eager.pre_tick(snapshot=False)
        eager.evolve()
        eager.post_tick()
        for j, (kind, mask, ball, target, f) in enumerate(args):
            inp = g.inputs[j]
            inp["mask"].copy_(mask.to(torch.uint8))
            inp["ball"].copy_(ball)
            inp["target"].copy_(target)
            inp["f"].copy_(f)
        b = g.tick()
        torch.cuda.synchronize()
        for u, v in zip(eager.a.park._blocks(), graphed.a.park._blocks()):
            assert torch.equal(u, v), f"tick {tick}: a park block differs"
        for name, u in eager.space.st.items():
            assert torch.equal(u, graphed.space.st[name]), f"tick {tick}: the partition's {name} differs"
        assert torch.equal(eager.batch.route, b.route), tick
        assert torch.equal(eager.batch.rbubble, b.rbubble), tick
        assert torch.equal(eager.batch.has_character, b.has_character), tick
        assert torch.equal(eager.batch.gets_bubble, b.gets_bubble), tick
        assert torch.equal(eager.batch.narrow, b.narrow), tick
        assert [bool(s.ok.any()) for s in eager.batch.slots] == [bool(s.ok.any()) for s in b.slots], tick
        eager.a.park.pos.sum().item()               # eager work between replays
    assert eager.cleanup_stops == graphed.cleanup_stops
    graphed.check()
