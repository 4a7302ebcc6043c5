"""Crowded delayed removals (tests/scenarios.py `mass_remove_cases`: 8 or more balls of a world removed with delays
of 1-3 ticks in one step, more than the 7 that one BringOutTheDeads call takes within two ticks of their stamp, the
rest left to a later call): the oracle and the drop-in park's host build against the reference's own FullBallpark
(tests/golden/mass_remove_golden.npz), every ball after every tick, bit for bit; the GPU in test_mass_remove_gpu.py."""

import pytest

import golden_fixture as G
import park_driver as D
from test_oracle_missile_golden import check
from test_statics import compare_golden

FIXTURE = "mass_remove_golden.npz"


@pytest.fixture(scope="module")
def golden():
    return G.load(FIXTURE)


def test_oracle_matches_reference(golden):
    bad = check(golden, False)
    assert not bad, f"{len(bad)} worlds differ, first {bad[:6]}"


def test_fixture_reaches_the_cap(golden):
    """Some BringOutTheDeads call removes 7 balls while more are due (the fixture pins the cap)."""
    capped = 0
    for wi in golden["kept"]:
        w = golden["worlds"][wi]
        n = [golden["records"][(wi, s)][0].shape[0] if (wi, s) in golden["records"] else 0
             for s in range(w["steps"] + 1)]
        drops = [a - b for a, b in zip(n, n[1:])]
        capped += max(drops) == 7 and sum(drops) > 7
    assert capped >= 10


@pytest.mark.parametrize("block", [256, 64])
def test_park_matches_reference_vectors(golden, block):
    kept = golden["kept"]
    got = D.run_park(golden["worlds"], "cpu", order=kept, initial=True, block=block)
    bad = compare_golden(got, golden, kept)
    assert not bad, f"{len(bad)} (world, step) records differ, first {bad[:6]}"
