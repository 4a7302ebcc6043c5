"""RemoveBall between ticks without a delay (a client's view): the oracle's and the drop-in park's (`remove`, host
build; GPU in test_remove_gpu.py) against the reference's own FullBallpark (tests/golden/remove_golden.npz: the
removal worlds, every ball as added and after every tick, a removed ball gone from the records from then on, bit for
bit with mLastC, the filing, the collision lists and the display yaw, pitch and roll)."""

import pytest

import golden_fixture as G
import park_driver as D
from test_oracle_missile_golden import check
from test_statics import compare_golden


@pytest.fixture(scope="module")
def golden():
    return G.load("remove_golden.npz")


def test_oracle_removal_matches_reference(golden):
    bad = check(golden, False)
    assert not bad, f"{len(bad)} worlds differ, first {bad[:6]}"


def test_fixture_removes_balls(golden):
    gone = 0
    for wi in golden["kept"]:
        w = golden["worlds"][wi]
        first, last = golden["records"][(wi, 0)], golden["records"][(wi, w["steps"])]
        gone += int(len(last[1]) < len(first[1]))
    assert gone >= 60, gone


@pytest.mark.parametrize("block", [256, 64])
def test_park_matches_reference_vectors(golden, block):
    kept = golden["kept"]
    got = D.run_park(golden["worlds"], "cpu", order=kept, initial=True, block=block)
    bad = compare_golden(got, golden, kept)
    assert not bad, f"{len(bad)} (world, step) records differ, first {bad[:6]}"
