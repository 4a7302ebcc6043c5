"""Missiles through the drop-in park (VecModesPark(collisions=True): launch_missile, missile_follow, the MISSILE tick),
host build of the kernel source: against the reference's own vectors (tests/golden/missile_golden.npz: every ball of
the missile worlds as added and after every tick, bit for bit with mLastC, the filing, the collision lists and the
display yaw, pitch and roll; missile_do_golden.npz with dynamical orientation), and against the oracle in workgroups
of 64 rows in reverse world order."""

import pytest

import golden_fixture as G
import park_driver as D
import scenarios as S
from test_statics import compare_golden


@pytest.mark.parametrize("block", [256, 64])
def test_park_matches_reference_vectors(block):
    golden = G.load("missile_golden.npz")
    kept = golden["kept"]
    got = D.run_park(golden["worlds"], "cpu", order=kept, initial=True, block=block)
    bad = compare_golden(got, golden, kept)
    assert not bad, f"{len(bad)} (world, step) records differ, first {bad[:6]}"


def test_park_with_orientation_matches_reference_vectors():
    golden = G.load("missile_do_golden.npz")
    kept = golden["kept"]
    got = D.run_park(golden["worlds"], "cpu", order=kept, initial=True, block=128)
    bad = compare_golden(got, golden, kept, do=True)
    assert not bad, f"{len(bad)} (world, step) records differ, first {bad[:6]}"


def test_worlds_match_oracle_reversed():
    ws = S.missile_cases(5)
    for w in ws:
        w["ypr"] = True
    idx = list(reversed(range(0, len(ws), 3)))
    got = D.run_park(ws, "cpu", order=idx, block=64)
    bad = D.compare(got, D.run_oracle(ws, idx), ws)
    assert not bad, f"{len(bad)} differ: {bad}"
