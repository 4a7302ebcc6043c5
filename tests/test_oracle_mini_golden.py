"""The oracle's minis and force fields (COL-713, MOV-737): a fixed parent's miniballs (local ids, fixed and massive
with the parent's mass, MINIBALL, owned by the parent), minicapsules and miniboxes (static collidables with the
parent); GetNearbyBalls' BallSortComparer order within a box (two local ids by their owners first); a missile skipping
its owner's miniballs; SetBallHarmonic and the candidates' force field filter (harmonic -2 passes, or the field's
harmonic, corporation or alliance the ball's). Pinned to the reference: every ball of the mini and field worlds
(tests/scenarios.py `mini_cases`) as added and after every tick, bit for bit with mLastC, the filing, the collision
lists and the display yaw, pitch and roll, against the vectors the reference's own FullBallpark wrote
(tests/golden/make_golden_missile.py --minis); and with dynamical orientation (mini_do_golden.npz). The skip of a
missile's owner's miniballs on the defender worlds of mini_own_golden.npz (`mini_own_cases`, --mini-own): a defender
whose owner is its parent's negative id passes through the miniballs, another parent's it meets."""

import numpy as np

import pytest

import golden_fixture as G
from test_oracle_missile_golden import check

MODE, FIELD, MINIBALL = 1, 10, 9


@pytest.fixture(scope="module")
def golden():
    return G.load("mini_golden.npz")


def test_oracle_minis_and_fields_match_reference(golden):
    bad = check(golden, False)
    assert not bad, f"{len(bad)} worlds differ, first {bad[:6]}"


def test_oracle_minis_and_fields_with_orientation_match_reference():
    bad = check(G.load("mini_do_golden.npz"), True)
    assert not bad, f"{len(bad)} worlds differ, first {bad[:6]}"


def test_fixture_covers_minis_and_fields(golden):
    """Miniballs met by ships, two local balls in one list, fields passed (a ball inside a field's radius that does
    not list it) and met."""
    met_mini = two_local = field_met = 0
    for (wi, step), (d, q, hits) in golden["records"].items():
        ids = q[:, 0]
        modes = dict(zip(ids.tolist(), q[:, MODE].tolist()))
        for h in hits:
            locals_ = [x for x in h if x < 0]
            met_mini += int(bool(locals_))
            two_local += int(len(locals_) >= 2)
            field_met += int(any(modes.get(x) == FIELD for x in h))
    kinds = {golden["worlds"][wi]["kind"] for wi in golden["kept"]}
    assert met_mini >= 100 and two_local >= 20 and field_met >= 50, (met_mini, two_local, field_met)
    assert {"minis", "minis_missile", "field", "field_cleared"} <= kinds


def test_oracle_defenders_through_minis_match_reference():
    bad = check(G.load("mini_own_golden.npz"), False)
    assert not bad, f"{len(bad)} worlds differ, first {bad[:6]}"


def test_fixture_reaches_the_owner_miniball_skip():
    """In every world whose parent is the defender's owner, the defender overlaps one of the parent's miniballs at
    some tick and never lists one (nor the parent); where the parent is another's, it lists them."""
    g = G.load("mini_own_golden.npz")
    own = other = 0
    for wi in g["kept"]:
        w = g["worlds"][wi]
        missile, parent = w["balls"][3], w["balls"][1]
        overlapped, listed = False, 0
        for step in range(w["steps"] + 1):
            d, q, hits = g["records"][(wi, step)]
            ids = q[:, 0].tolist()
            k = ids.index(missile["id"])
            listed += sum(1 for x in hits[k] if x < -1_000_000 or x == parent["id"])
            minis = [m for m in w["minis"] if m["kind"] == "ball"]
            local = sorted((i for i in ids if i < -1_000_000), reverse=True)
            for m, mid in zip(minis, local):
                j = ids.index(mid)
                if np.linalg.norm(d[k, 0:3] - d[j, 0:3]) < m["radius"] + missile["radius"]:
                    overlapped = True
        if w["kind"] == "defender_own_minis":
            assert listed == 0, wi
            own += overlapped
        else:
            other += listed > 0
    assert own >= 15 and other >= 15, (own, other)
