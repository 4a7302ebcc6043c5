"""Cloaking through the drop-in park (VecModesPark(collisions=True): cloak and uncloak between ticks, the ticks around
them), host build of the kernel source: against the reference's own vectors (tests/golden/cloak_golden.npz with the
display yaw, pitch and roll, cloak_do_golden.npz with dynamical orientation: every ball of the cloaking worlds as added
and after every tick, bit for bit with mLastC, the filing and the collision lists), and against the oracle in
workgroups of 64 rows in reverse world order."""

import pytest
import torch

import golden_fixture as G
import park_driver as D
import scenarios as S
from exact_ballpark_amd.vec_modes import VecModesPark
from test_statics import compare_golden


@pytest.mark.parametrize("block", [256, 64])
def test_park_matches_reference_vectors(block):
    golden = G.load("cloak_golden.npz")
    kept = golden["kept"]
    got = D.run_park(golden["worlds"], "cpu", order=kept, initial=True, block=block)
    bad = compare_golden(got, golden, kept)
    assert not bad, f"{len(bad)} (world, step) records differ, first {bad[:6]}"


def test_park_with_orientation_matches_reference_vectors():
    golden = G.load("cloak_do_golden.npz")
    kept = golden["kept"]
    got = D.run_park(golden["worlds"], "cpu", order=kept, initial=True, block=128)
    bad = compare_golden(got, golden, kept, do=True)
    assert not bad, f"{len(bad)} (world, step) records differ, first {bad[:6]}"


def test_worlds_match_oracle_reversed():
    ws = S.cloak_cases(11)[:48]
    for w in ws:
        w["ypr"] = True
    idx = list(reversed(range(len(ws))))
    got = D.run_park(ws, "cpu", order=idx, block=64)
    bad = D.compare(got, D.run_oracle(ws, idx), ws)
    assert not bad, f"{len(bad)} differ: {bad}"


def test_the_cloak_byte():
    """isCloaked is the mode's byte (a signed char above 0); a refused mode leaves it."""
    p = VecModesPark(3, device="cpu", collisions=True)
    p.load([{"id": 5 + i, "world": 0, "time": 0, "mass": 1e6, "radius": 50.0, "max_vel": 300.0,
             "pos": [1000.0 * i, 0.0, 0.0], "vel": [0.0, 0.0, 0.0], "agility": 1.0, "speed_fraction": 1.0,
             "free": True, "massive": True, "interactive": False, "cloaked": 0, "bubble": 1, "moribund": False,
             "mode": 2} for i in range(3)])
    p.cloak(torch.tensor([True, True, True]), torch.tensor([3, 200, 127]))
    assert p.cloaked.tolist() == [3, 0, 127] and p.massive.tolist() == [False, True, False]
    p.uncloak(torch.tensor([True, False, False]))
    assert p.cloaked.tolist() == [0, 0, 127] and p.massive.tolist() == [True, True, False]
