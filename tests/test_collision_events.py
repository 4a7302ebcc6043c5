"""DoCollision events (Ball::DispatchCollisions: a missile that moved at least sqrt(0.1) m in a tick dispatches one
event per entry of its collision list, with its mLastC): the oracle's and the drop-in park's (`collision_events`, host
build) against the events the reference's own FullBallpark emitted on the missile worlds (tests/golden/
missile_golden.npz and missile_do_golden.npz, written by tests/golden/make_golden_missile.py), bit for bit."""

from collections import defaultdict

import numpy as np
import pytest

import golden_fixture as G
import oracle as O
import park_driver as D


def golden_events(name: str) -> dict:
    z = np.load(G.GOLDEN / name, allow_pickle=False)
    c = z["ev_c"].view(np.float64)
    out = defaultdict(list)
    for i, wi in enumerate(z["ev_world"].tolist()):
        t, a, b = z["ev_q"][i].tolist()
        out[wi].append((t, a, b, *c[i].tolist()))
    return out


def oracle_events(w: dict, do: bool) -> list:
    o = O.World(w["friction"], w["tick"], w["new_orbit"], w["start"], do, True, not do)
    for b in w["balls"]:
        o.add(b)
    by_step = defaultdict(list)
    for cmd in w["cmds"]:
        by_step[cmd[0]].append(cmd)
    out = []
    for step in range(1, w["steps"] + 1):
        for cmd in by_step[step]:
            o.command(cmd)
        o.evolve()
        out.extend(o.take_events())
    return out


def same(a: list, b: list) -> bool:
    return len(a) == len(b) and all(x[:3] == y[:3] and np.array(x[3:]).view(np.uint64).tolist() ==
                                    np.array(y[3:]).view(np.uint64).tolist() for x, y in zip(a, b))


@pytest.mark.parametrize("name,do", [("missile_golden.npz", False), ("missile_do_golden.npz", True)])
def test_oracle_events_match_reference(name, do):
    golden, want = G.load(name), golden_events(name)
    assert sum(len(v) for v in want.values()) > (1000 if not do else 500)
    bad = [wi for wi in golden["kept"] if not same(oracle_events(golden["worlds"][wi], do), want.get(wi, []))]
    assert not bad, f"{len(bad)} worlds differ, first {bad[:6]}"


@pytest.mark.parametrize("name,do", [("missile_golden.npz", False), ("missile_do_golden.npz", True)])
def test_park_events_match_reference(name, do):
    golden, want = G.load(name), golden_events(name)
    got: dict = {}
    D.run_park(golden["worlds"], "cpu", order=golden["kept"], block=128, events=got)
    bad = [wi for wi in golden["kept"] if not same(got.get(wi, []), want.get(wi, []))]
    assert not bad, f"{len(bad)} worlds differ, first {bad[:6]}"


def test_oracle_flags_events_it_left_out():
    """A reader of the oracle's events gets every one or an error: a world whose run piles up more events than a
    read takes (here a read of one at a time) makes take_events raise instead of truncating."""
    import ctypes
    golden, want = G.load("missile_golden.npz"), golden_events("missile_golden.npz")
    wi = next(w for w in golden["kept"] if len(want.get(w, [])) >= 2)
    w = golden["worlds"][wi]
    o = O.World(w["friction"], w["tick"], w["new_orbit"], w["start"], False, True, True)
    for b in w["balls"]:
        o.add(b)
    by_step = defaultdict(list)
    for cmd in w["cmds"]:
        by_step[cmd[0]].append(cmd)
    for step in range(1, w["steps"] + 1):
        for cmd in by_step[step]:
            o.command(cmd)
        o.evolve()
    q, c = (ctypes.c_int64 * 3)(), (ctypes.c_double * 3)()
    assert o.L.or_take_events(o.p, q, c, 1) == -1          # more unread than the read takes
    assert o.take_events() == []                           # the buffer emptied, the flag cleared
