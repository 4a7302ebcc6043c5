"""The oracle's formations (LoadFormations, SetBallFormation with its leave and shrink rules, FormationFollow,
EvolveFormation: the slot offset scaled to both radii plus its length and turned by the leader's rotation, from its
display yaw, pitch and roll without dynamical orientation) pinned to the reference: every ball of the formation worlds
(tests/scenarios.py `formation_cases`) as added and after every tick, bit for bit, against the vectors the reference's
own FullBallpark wrote (tests/golden/make_golden_missile.py --formations); and with dynamical orientation."""

import numpy as np
import pytest

import golden_fixture as G
from test_oracle_missile_golden import check


@pytest.fixture(scope="module")
def golden():
    return G.load("formation_golden.npz")


def test_oracle_formations_match_reference(golden):
    bad = check(golden, False)
    assert not bad, f"{len(bad)} worlds differ, first {bad[:6]}"


def test_oracle_formations_with_orientation_match_reference():
    bad = check(G.load("formation_do_golden.npz"), True)
    assert not bad, f"{len(bad)} worlds differ, first {bad[:6]}"


def test_fixture_covers_the_formations(golden):
    in_slot = stopped = 0
    for (wi, step), (d, q, lists) in golden["records"].items():
        in_slot += int((q[:, 1] == 12).sum())
    kinds = {golden["worlds"][wi]["kind"] for wi in golden["kept"]}
    assert in_slot > 1500, in_slot
    assert {"formation", "formation_change", "formation_leave", "formation_past_table"} <= kinds
