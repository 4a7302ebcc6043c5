"""Mushrooms through the drop-in park (VecModesPark(collisions=True): add_mushroom into an empty row, set_free, the
mushroom's turn in the tick), host build of the kernel source: against the reference's own vectors
(tests/golden/mushroom_golden.npz with the display yaw, pitch and roll, mushroom_do_golden.npz with dynamical orientation:
every ball, ball 0 from its AddMushroom on, after every tick, bit for bit with mLastC, the filing and the collision lists),
and against the oracle in workgroups of 64 rows in reverse world order."""

import pytest
import torch

import golden_fixture as G
import park_driver as D
import scenarios as S
from exact_ballpark_amd.vec_modes import VecModesPark
from test_statics import compare_golden


@pytest.mark.parametrize("block", [256, 64])
def test_park_matches_reference_vectors(block):
    golden = G.load("mushroom_golden.npz")
    kept = golden["kept"]
    got = D.run_park(golden["worlds"], "cpu", order=kept, initial=True, block=block)
    bad = compare_golden(got, golden, kept)
    assert not bad, f"{len(bad)} (world, step) records differ, first {bad[:6]}"


def test_park_with_orientation_matches_reference_vectors():
    golden = G.load("mushroom_do_golden.npz")
    kept = golden["kept"]
    got = D.run_park(golden["worlds"], "cpu", order=kept, initial=True, block=128)
    bad = compare_golden(got, golden, kept, do=True)
    assert not bad, f"{len(bad)} (world, step) records differ, first {bad[:6]}"


def test_worlds_match_oracle_reversed():
    ws = S.mushroom_cases(11)[:48]
    for w in ws:
        w["ypr"] = True
    idx = list(reversed(range(len(ws))))
    got = D.run_park(ws, "cpu", order=idx, block=64)
    bad = D.compare(got, D.run_oracle(ws, idx), ws)
    assert not bad, f"{len(bad)} differ: {bad}"


def test_mushroom_rows_are_checked():
    """add_mushroom takes empty rows only; a mushroom is no follow target; a second ball 0 in a world is an error."""
    rows = [{"id": 5 + i, "world": 0, "time": 0, "mass": 1e6, "radius": 50.0, "max_vel": 300.0,
             "pos": [1000.0 * i, 0.0, 0.0], "vel": [0.0, 0.0, 0.0], "agility": 1.0, "speed_fraction": 1.0,
             "free": True, "massive": True, "interactive": False, "cloaked": 0, "bubble": -1, "moribund": False,
             "mode": 2} for i in range(2)]
    p = VecModesPark(4, device="cpu", collisions=True)
    p.load(rows + [VecModesPark.empty_row(0), VecModesPark.empty_row(0)])
    t = lambda *x: torch.tensor(x)
    with pytest.raises(ValueError):
        p.add_mushroom(t(True, False, False, False), t(5, 0, 0, 0), t(100.0, 0, 0, 0), t(3.0, 0, 0, 0))
    p.add_mushroom(t(False, False, True, False), t(0, 0, 5, 0), t(100.0, 0, 100.0, 0), t(0, 0, 3.0, 0))
    assert int(p.ids[2]) == 0 and int(p.mode[2]) == 6 and float(p.goto[2, 0]) == 3.0
    with pytest.raises(ValueError):
        p.add_mushroom(t(False, False, False, True), t(0, 0, 0, 5), t(0, 0, 0, 100.0), t(0, 0, 0, 3.0))
    with pytest.raises(ValueError):
        p.follow(t(True, False, False, False), t(2, 0, 0, 0), t(500.0, 0, 0, 0))
    p.add_mushroom(t(False, False, False, True), t(0, 0, 0, 5), t(0, 0, 0, -1.0), t(0, 0, 0, 3.0))   # refused
    assert int(p.ids[3]) == -(2 ** 62)
