"""Minis and force fields through the drop-in park (VecModesPark(collisions=True): mini_rows' miniballs and static
collidables, set_harmonic, the ticks), host build of the kernel source: against the reference's own vectors
(tests/golden/mini_golden.npz with the display yaw, pitch and roll, mini_do_golden.npz with dynamical orientation: every
ball, the miniballs with their local ids, after every tick, bit for bit with mLastC, the filing and the collision lists),
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
    golden = G.load("mini_golden.npz")
    kept = golden["kept"]
    got = D.run_park(golden["worlds"], "cpu", order=kept, initial=True, block=block)
    bad = compare_golden(got, golden, kept)
    assert not bad, f"{len(bad)} (world, step) records differ, first {bad[:6]}"


def test_park_with_orientation_matches_reference_vectors():
    golden = G.load("mini_do_golden.npz")
    kept = golden["kept"]
    got = D.run_park(golden["worlds"], "cpu", order=kept, initial=True, block=128)
    bad = compare_golden(got, golden, kept, do=True)
    assert not bad, f"{len(bad)} (world, step) records differ, first {bad[:6]}"


def test_worlds_match_oracle_reversed():
    ws = S.mini_cases(11)[:48]
    for w in ws:
        w["ypr"] = True
    idx = list(reversed(range(len(ws))))
    got = D.run_park(ws, "cpu", order=idx, block=64)
    bad = D.compare(got, D.run_oracle(ws, idx), ws)
    assert not bad, f"{len(bad)} differ: {bad}"


@pytest.mark.parametrize("block", [256, 64])
def test_defenders_through_minis_match_reference_vectors(block):
    """Defender missiles through the miniballs of a parent whose id is their owner's (skipped) or another's (met)."""
    golden = G.load("mini_own_golden.npz")
    kept = golden["kept"]
    got = D.run_park(golden["worlds"], "cpu", order=kept, initial=True, block=block)
    bad = compare_golden(got, golden, kept)
    assert not bad, f"{len(bad)} (world, step) records differ, first {bad[:6]}"
