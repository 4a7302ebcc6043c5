"""The oracle's mushrooms (AddMushroom: a client's ball 0, fixed, massive, radius 0 at its owner; SetBallFree; a free
mushroom's turn: range x powf(fraction, 0.25f) and re-filed, or, its lifetime passed, RemoveBall inside Evolve; no swap;
Gradient's skip of a ball's own mushroom) pinned to the reference: every ball of the mushroom worlds (tests/scenarios.py
`mushroom_cases`), ball 0 from its AddMushroom on, as added and after every tick, bit for bit with mLastC, the filing,
the collision lists and the display yaw, pitch and roll, against the vectors the reference's own FullBallpark wrote
(tests/golden/make_golden_missile.py --mushrooms); and with dynamical orientation (mushroom_do_golden.npz)."""

import pytest

import golden_fixture as G
from test_oracle_missile_golden import check

RADIUS, MODE, MASSIVE, STOP, MUSHROOM = 19, 1, 6, 2, 6


@pytest.fixture(scope="module")
def golden():
    return G.load("mushroom_golden.npz")


def test_oracle_mushrooms_match_reference(golden):
    bad = check(golden, False)
    assert not bad, f"{len(bad)} worlds differ, first {bad[:6]}"


def test_oracle_mushrooms_with_orientation_match_reference():
    bad = check(G.load("mushroom_do_golden.npz"), True)
    assert not bad, f"{len(bad)} worlds differ, first {bad[:6]}"


def test_fixture_covers_the_mushrooms(golden):
    """Mushrooms that grow over several ticks, that are removed (STOP, non-massive), balls that meet a mushroom (ball 0
    in their collision lists) and owners that do not."""
    grew = removed = met = 0
    for wi in golden["kept"]:
        w = golden["worlds"][wi]
        radii, gone = [], False
        for step in range(w["steps"] + 1):
            d, q, hits = golden["records"][(wi, step)]
            if q[0, 0] != 0:
                continue
            radii.append(d[0, RADIUS])
            gone |= bool(q[0, MODE] == STOP and q[0, MASSIVE] == 0)
            met += sum(1 for h in hits[1:] if 0 in h)
        grew += int(len(set(radii)) >= 3)
        removed += int(gone)
    assert grew >= 40 and removed >= 30 and met >= 30, (grew, removed, met)
