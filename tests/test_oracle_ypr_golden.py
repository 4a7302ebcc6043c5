"""The oracle's display orientation without dynamical orientation (Ball::CalculateYawPitchRoll: yaw, pitch, roll, their
old values, the roll speed, the yaw change) pinned to the reference: every ball of the flight worlds (tests/scenarios.py
`cases`, less the guard worlds) and the collision worlds (`collide_cases`) as added and after every tick, bit for bit
(two NaNs equal), with the rest of the state and the collision lists, against the vectors the reference's own
FullBallpark(dynamical_orientation=False) wrote (tests/golden/make_golden_ypr.py)."""

from collections import defaultdict

import numpy as np
import pytest

import golden_fixture as G
import oracle as O
from park_driver import _same_bits


@pytest.fixture(scope="module")
def golden():
    return G.load("ypr_golden.npz")


def run(w):
    o = O.World(w["friction"], w["tick"], w["new_orbit"], w["start"], False, True, True)
    for b in w["balls"]:
        o.add(b)
    by_step = defaultdict(list)
    for c in w["cmds"]:
        by_step[c[0]].append(c)
    out = []
    for step in range(0, w["steps"] + 1):
        if step:
            for c in by_step[step]:
                o.command(c)
            o.evolve()
        d, q, _ = o.state()
        lc, filing, hits = o.collide()
        out.append((step, np.concatenate([d[:, :26], lc, o.ypr()], axis=1),
                    np.concatenate([q[:, G.Q_FROM_ORACLE], filing], axis=1), hits))
    return out


def test_fixture_covers_the_orientation(golden):
    d = np.concatenate([r[0] for r in golden["records"].values()])
    assert (d[:, 29] != 0.0).mean() > 0.5 and (d[:, 31] != 0.0).mean() > 0.2 and (d[:, 37] != 0.0).mean() > 0.2


def test_oracle_yaw_pitch_roll_match_reference(golden):
    bad = []
    for wi in golden["kept"]:
        for step, d, q, hits in run(golden["worlds"][wi]):
            gd, gq, gh = golden["records"][(wi, step)]
            same_d = _same_bits(d, gd)
            if not (same_d.all() and (q == gq).all() and hits == gh):
                cols = sorted(set(np.nonzero(~same_d)[1].tolist()))
                bad.append((wi, golden["worlds"][wi]["kind"], step, cols[:6], hits != gh))
    assert not bad, f"{len(bad)} (world, step) records differ, first {bad[:6]}"
