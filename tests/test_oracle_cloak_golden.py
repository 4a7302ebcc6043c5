"""The oracle's cloaking (CloakBall: refused unless the signed char mode is positive, the followers released, the ball
cloaked and non-massive; UncloakBall: massive again unless in warp proper) pinned to the reference: every ball of the
cloaking worlds (tests/scenarios.py `cloak_cases`) as added and after every tick, bit for bit with mLastC, the filing,
the collision lists and the display yaw, pitch and roll, against the vectors the reference's own FullBallpark wrote
(tests/golden/make_golden_missile.py --cloak); and with dynamical orientation (cloak_do_golden.npz)."""

import pytest

import golden_fixture as G
import scenarios as S

FOLLOW = 1                      # the FOLLOW mode
from test_oracle_missile_golden import check


@pytest.fixture(scope="module")
def golden():
    return G.load("cloak_golden.npz")


def test_oracle_cloaking_matches_reference(golden):
    bad = check(golden, False)
    assert not bad, f"{len(bad)} worlds differ, first {bad[:6]}"


def test_oracle_cloaking_with_orientation_matches_reference():
    bad = check(G.load("cloak_do_golden.npz"), True)
    assert not bad, f"{len(bad)} worlds differ, first {bad[:6]}"


def test_fixture_covers_the_cloaking(golden):
    """The cloaking ball turns non-massive when its cloak applies and massive when it uncloaks (also one that was not
    massive before); its followers leave their modes at the cloak; a refused cloak changes nothing."""
    dropped = regained = released = refused_kept = 0
    for wi in golden["kept"]:
        w = golden["worlds"][wi]
        cl = w["balls"][0]["id"]
        ids = sorted(b["id"] for b in w["balls"])
        k = ids.index(cl)
        for c in w["cmds"]:
            if c[1] not in (S.CLOAK_OP, S.UNCLOAK_OP) or c[0] > w["steps"]:
                continue
            before = golden["records"][(wi, c[0] - 1)][1]
            after = golden["records"][(wi, c[0])][1]
            if c[1] == S.CLOAK_OP and 0 < c[8] < 128:
                dropped += int(before[k, 6] == 1 and after[k, 6] == 0)
                released += int(((before[:, 1] == FOLLOW) & (before[:, 5] == cl)
                                 & (after[:, 5] != cl)).sum())
            elif c[1] == S.CLOAK_OP:
                refused_kept += int(before[k, 6] == after[k, 6])
            else:
                regained += int(before[k, 6] == 0 and after[k, 6] == 1)
    assert dropped >= 30 and regained >= 20 and refused_kept >= 15 and released >= 10, \
        (dropped, regained, refused_kept, released)
