"""The server layer (VecServer over a VecModesPark, exact-ballpark_amd/csrc/xb_server.h) on the host build, bit for bit
against the reference's own server layer (tests/golden/server_golden.npz and server_do_golden.npz, written by
tests/golden/make_golden_server.py from src/exact_ballpark/server.py on a master FullBallpark): every ball's state
after every tick (flight fields, display orientation or turning state, the flags, the cloak byte), every update every
client received (singlecasts and narrowcasts in order, with their wait flags and counts, the events with their
stamps, the bubble additions' and the uncloaks' ball streams as bytes) and every pending query's answer; the worlds
in reverse order as well (a world's result does not depend on its neighbours); the refusals."""

import ctypes
import json

import pytest
import torch

import server_driver as SD
from exact_ballpark_amd import _native as N
from exact_ballpark_amd import vec_server as VS
from exact_ballpark_amd.vec_modes import VecModesPark


@pytest.mark.parametrize("do", [False, True], ids=["display", "dynamical"])
def test_server_layer_matches_reference_vectors(do):
    r = SD.compare_golden(do, "cpu")
    assert r["compared"] > 3000 and r["worlds"] > 80
    assert not r["rows"], f"{len(r['rows'])} ball-states differ, first {r['rows'][:4]}"
    assert not r["sends"], f"{len(r['sends'])} world-ticks' sends differ, first {r['sends'][:4]}"
    assert not r["answers"], f"{len(r['answers'])} query answers differ, first {r['answers'][:2]}"


def test_server_layer_reversed_worlds():
    r = SD.compare_golden(False, "cpu", reverse=True)
    assert not r["rows"] and not r["sends"] and not r["answers"], (r["rows"][:3], r["sends"][:3])


def test_fixture_covers_the_rules():
    """The golden streams reach what the layer's rules turn on."""
    g = SD.load_golden(False)
    names = [e[1] for w in g["worlds"] for tick in w["ticks"] for e in tick]
    for name in ("cloak_ball", "uncloak_ball", "undo_pending_cloak", "set_ball_position", "set_ball_troll",
                 "set_ball_free", "set_ball_interactive", "warp_to", "entity_warp_in", "follow_ball", "orbit",
                 "has_pending_cloak", "get_pending_speed_fraction", "has_pending_set_free"):
        assert name in names, name
    flat = json.dumps(g["sends"])
    assert "AddBallsToPark" in flat and "PackagedAction" in flat and '"narrow"' in flat and '"single"' in flat


def _park(n_worlds=3, ships=3, device="cpu"):
    rows = [{"id": 10 + k, "world": w, "time": 0, "mass": 1e7, "radius": 50.0, "max_vel": 200.0,
             "pos": [100.0 * k, 0.0, 0.0], "vel": [0.0, 0.0, 0.0], "agility": 1.0, "speed_fraction": 1.0,
             "interactive": k == 0} for w in range(n_worlds) for k in range(ships)]
    park = VecModesPark(len(rows), device=device, display_ypr=True, master=True)
    park.load(rows)
    return park


def test_push_refuses_kinds_the_park_does_not_apply():
    srv = VS.VecServer(VS.ModesAdapter(_park()))
    m, b = torch.ones(3, dtype=torch.bool), torch.full((3,), 10, dtype=torch.int64)
    with pytest.raises(ValueError):
        srv.push("LaunchMissile", m, b)
    with pytest.raises(ValueError):
        srv.push("SetBallRotation", m, b)        # dynamical orientation only


def test_check_refuses_a_dead_bubble():
    srv = VS.VecServer(VS.ModesAdapter(_park()), pilots=[0, -1, -1] * 3)
    m = torch.tensor([True, False, False])
    srv.push("SetBallInteractive", m, torch.full((3,), 10, dtype=torch.int64), n=torch.zeros(3, dtype=torch.int64))
    srv.tick()
    with pytest.raises(RuntimeError, match="no interactive ball"):
        srv.check()


def test_abi_refusals():
    park = _park()
    srv = VS.VecServer(VS.ModesAdapter(park))
    s = park.state()
    st = srv._state(False)
    q = VS.VecServer._queue_struct(srv._queue([], None, False))
    L = N.cpu()
    L.xb_cpu_server_pre_tick.argtypes = [ctypes.c_void_p] * 3 + [ctypes.c_int]
    assert L.xb_cpu_server_pre_tick(ctypes.byref(s), ctypes.byref(st), ctypes.byref(q), 0) == 0
    assert L.xb_cpu_server_pre_tick(ctypes.byref(s), ctypes.byref(st), ctypes.byref(q), 3) == -1
    bad = VS.ServerQueue.from_buffer_copy(q)
    bad.J = 5000
    assert L.xb_cpu_server_pre_tick(ctypes.byref(s), ctypes.byref(st), ctypes.byref(bad), 0) == -1
    bad = VS.ServerQueue()                     # a call with no tables
    bad.J = 1
    assert L.xb_cpu_server_pre_tick(ctypes.byref(s), ctypes.byref(st), ctypes.byref(bad), 0) == -1
    st2 = VS.ServerState.from_buffer_copy(st)
    st2.rows = None
    assert L.xb_cpu_server_pre_tick(ctypes.byref(s), ctypes.byref(st2), ctypes.byref(q), 0) == -1
