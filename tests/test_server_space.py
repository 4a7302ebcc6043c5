"""The server layer over the master's partition (SpaceServer over a VecModesPark and its VecSpace,
exact-ballpark_amd/csrc/xb_server.h with the partition) on the host build, bit for bit against the reference's own
server layer on worlds of several bubbles (tests/golden/server_space_golden.npz, written by
tests/golden/make_golden_server.py --space from src/exact_ballpark/server.py on a master FullBallpark): every ball's
state, bubble and old bubble after every tick, every update every client received (the follower clean-up's Stops, the
bubble update's RemoveBalls and PackagedActions across bubbles, the narrowcasts per bubble, the cloaks sent to the
bubble they were queued in, the uncloaks' notices, the ball streams as bytes, the wait flags and counts) and every
pending query's answer; the worlds in reverse order; what the fixture reaches; the refusals."""

import json

import pytest
import torch

import server_driver as SD
from exact_ballpark_amd import vec_server as VS
from exact_ballpark_amd import vec_server_space as VSS
from exact_ballpark_amd.vec_modes import VecModesPark


def test_space_server_matches_reference_vectors():
    r = SD.compare_golden(False, "cpu", space=True)
    assert r["compared"] > 8000 and r["worlds"] >= 150
    assert not r["rows"], f"{len(r['rows'])} ball-states differ, first {r['rows'][:4]}"
    assert not r["sends"], f"{len(r['sends'])} world-ticks' sends differ, first {r['sends'][:4]}"
    assert not r["answers"], f"{len(r['answers'])} query answers differ, first {r['answers'][:2]}"


def test_space_server_reversed_worlds():
    r = SD.compare_golden(False, "cpu", limit=60, reverse=True, space=True)
    assert not r["rows"] and not r["sends"] and not r["answers"], (r["rows"][:3], r["sends"][:3])


def _flatten(x, out):
    if isinstance(x, list):
        if len(x) == 2 and x[0] == "s":
            out.append(x[1])
        else:
            for v in x:
                _flatten(v, out)
    return out


def test_space_fixture_reaches_the_rules():
    """The golden streams reach what the partition's rules turn on: several bubbles in a world, several bubbles
    narrowcasting in one tick, clients with both updates, the bubble update's own-list removals and the bubble's
    packaged lists, cloaks sent to a bubble, uncloak notices; and the batched run queues the follower clean-up's
    Stops and lists the bubble update's kinds."""
    g = SD.load_golden(False, space=True)
    several = sum(1 for (k, t), st in g["states"].items() if len({int(q[18]) for _, q in st.values()}) > 1)
    assert several > 500
    own_removals = packaged_removals = two_bubbles = two_updates = cloaks = notices = 0
    for w_sends in g["sends"]:
        for snd in w_sends:
            per = snd[1]
            narrow = set()
            for client in per:
                ups = client[1][1]
                if len(ups) == 2:
                    two_updates += 1
                for up in ups:
                    kind = up[1][0][1]
                    names = _flatten(up[1][2], [])
                    if kind == "narrow":
                        narrow.add(json.dumps(up[1][2]))
                        cloaks += names.count("CloakBall")
                    else:
                        for k, name in enumerate(names):
                            if name == "RemoveBalls":
                                packed = k > 0 and names[k - 1] == "PackagedAction"
                                packaged_removals += packed
                                own_removals += not packed
                        notices += sum(1 for a in names if a == "AddBallsToPark")
            two_bubbles += len(narrow) > 1
    assert own_removals > 20 and packaged_removals > 20 and cloaks > 50 and notices > 100
    assert two_bubbles > 20 and two_updates > 100
    worlds = [g["worlds"][k] for k in g["kept"]]
    for w in worlds:
        w["ticks"] = [[(e[0], e[1], tuple(e[2])) for e in tick] for tick in w["ticks"]]
    server, _ = SD.build(worlds, "cpu", space=True)
    kinds = set()
    for t in range(len(worlds[0]["ticks"])):
        SD.apply_tick(server, worlds, t)
        server.tick()
        kinds |= set(server.batch.ad["kind"].tolist())
    assert server.cleanup_stops > 20
    # SAME_OWN (a user ball listed in the interactive set of a bubble other than its own, which stayed) is not
    # reached by these streams: its lists are PopulateAddDelForBubble's, as SAME_BUBBLE's, sent unpackaged
    assert {VSS.AD["SAME_BUBBLE"], VSS.AD["CHANGED"]} <= kinds


def _park(n_worlds=2, ships=3):
    rows = [{"id": 10 + k, "world": w, "time": 0, "mass": 1e7, "radius": 50.0, "max_vel": 200.0,
             "pos": [100.0 * k, 0.0, 0.0], "vel": [0.0, 0.0, 0.0], "agility": 1.0, "speed_fraction": 1.0,
             "interactive": k == 0, "bubble": -1} for w in range(n_worlds) for k in range(ships)]
    park = VecModesPark(len(rows), device="cpu", display_ypr=True, master=True)
    park.load(rows)
    return park


def test_space_server_refusals():
    park = _park()
    srv = VSS.SpaceServer(VSS.SpaceModesAdapter(park), pilots=[0, -1, -1] * 2)
    m, b = torch.ones(2, dtype=torch.bool), torch.full((2,), 10, dtype=torch.int64)
    with pytest.raises(ValueError):
        srv.push("SetBallTroll", m, b, n=torch.ones(2, dtype=torch.int64))   # PetrifyTroll's re-filing: not here
    with pytest.raises(NotImplementedError):
        VSS.space_for(VSS.SpaceModesAdapter(_park()), global_balls=[False] * 6)
    with pytest.raises(ValueError):
        VSS.SpaceModesAdapter(_park(), master=False)
    other = VSS.SpaceModesAdapter(_park())
    with pytest.raises(ValueError, match="partition must cover"):
        VSS.SpaceServer(VSS.SpaceModesAdapter(_park()), space=VSS.space_for(other))


def test_space_server_tick_and_on_tick_agree():
    """OnTick with every world one whole interval behind (after the first call's lone evolve) runs the same ticks as
    tick(): the same park, bubbles and sends."""
    worlds = SD.SS.space_server_cases(5, 24, ticks=6)
    for w in worlds:
        w["ticks"] = [[(e[0], e[1], tuple(e[2])) for e in tick] for tick in w["ticks"]]
    a, _ = SD.build(worlds, "cpu", space=True)
    b, _ = SD.build(worlds, "cpu", space=True)
    step = int(a.a.park.tick_interval) * 10000
    W = a.W
    a.evolve()
    b.on_tick(torch.full((W,), step, dtype=torch.int64))          # the first call: one evolve, no pre-tick
    for t in range(len(worlds[0]["ticks"])):
        SD.apply_tick(a, worlds, t)
        SD.apply_tick(b, worlds, t)
        a.tick()
        got = b.on_tick(torch.full((W,), (t + 2) * step, dtype=torch.int64))
        assert len(got) == 1
        assert SD.sends_of(a) == SD.sends_of(b)
        for x, y in zip(a.a.park._blocks(), b.a.park._blocks()):
            assert torch.equal(x, y)
        assert torch.equal(a.bubbles(), b.bubbles())
