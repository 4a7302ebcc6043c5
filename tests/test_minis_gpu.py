"""Minis and force fields on gfx950 (the collision tick kernels with miniballs in BallSortComparer's order, a parent's
minicapsules and miniboxes, the force field filter), bit for bit (two NaNs equal): against the reference's own vectors
(tests/golden/mini_golden.npz with the display yaw, pitch and roll, mini_do_golden.npz with dynamical
orientation): every ball as added and after every tick, with mLastC, the filing and the collision lists, in workgroups
of 256 and 64 rows."""

import pytest

import golden_fixture as G
import park_driver as D
from test_statics import compare_golden

pytestmark = pytest.mark.gpu
DEV = "cuda"


@pytest.mark.parametrize("block", [256, 64])
def test_park_matches_reference_vectors(block):
    golden = G.load("mini_golden.npz")
    kept = golden["kept"]
    got = D.run_park(golden["worlds"], DEV, order=kept, initial=True, block=block)
    bad = compare_golden(got, golden, kept)
    assert not bad, f"{len(bad)} (world, step) records differ, first {bad[:6]}"


def test_park_with_orientation_matches_reference_vectors():
    golden = G.load("mini_do_golden.npz")
    kept = golden["kept"]
    got = D.run_park(golden["worlds"], DEV, order=kept, initial=True, block=128)
    bad = compare_golden(got, golden, kept, do=True)
    assert not bad, f"{len(bad)} (world, step) records differ, first {bad[:6]}"


@pytest.mark.parametrize("block", [256, 64])
def test_defenders_through_minis_match_reference_vectors(block):
    """Defender missiles through the miniballs of a parent whose id is their owner's (skipped) or another's (met)."""
    golden = G.load("mini_own_golden.npz")
    kept = golden["kept"]
    got = D.run_park(golden["worlds"], DEV, order=kept, initial=True, block=block)
    bad = compare_golden(got, golden, kept)
    assert not bad, f"{len(bad)} (world, step) records differ, first {bad[:6]}"
