"""The oracle's dynamical orientation pinned to the reference: every ball of every turning world (tests/scenarios.py
do_cases) as added and after every tick, bit for bit (two NaNs equal), position .. last_collision, old position and
velocity, rotation, old rotation, angular velocity, old angular velocity, torque, roll torque and the turning values,
against the vectors the reference's own FullBallpark with dynamical orientation wrote
(tests/golden/make_golden_orient.py)."""

from collections import defaultdict

import numpy as np
import pytest

import golden_fixture as G
import oracle as O
from park_driver import _same_bits


@pytest.fixture(scope="module")
def golden():
    return G.load("orient_golden.npz")


def run(w):
    """[(step, doubles [n, 46], ints [n, 8], followers)] at step 0 and after every tick."""
    o = O.World(w["friction"], w["tick"], w["new_orbit"], w["start"], True)
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
        d, q, fol = o.state()
        out.append((step, np.concatenate([d[:, :26], o.turning()], axis=1), q[:, G.Q_FROM_ORACLE], fol))
    return out


def test_fixture_covers_the_turning(golden):
    kinds = {golden["worlds"][wi]["kind"] for wi in golden["kept"]}
    assert kinds >= {"do_goto", "do_orbit_old", "do_orbit_new", "do_warp", "do_warpin", "do_setters"}
    d = np.concatenate([r[0] for r in golden["records"].values()])
    assert (d[:, 29] != 1.0).mean() > 0.5 and (d[:, 40] != 0.0).mean() > 0.3      # turned; torque applied


def test_oracle_orientation_matches_reference(golden):
    bad = []
    for wi in golden["kept"]:
        for step, d, q, fol in run(golden["worlds"][wi]):
            gd, gq, gf = golden["records"][(wi, step)]
            same_d = _same_bits(d, gd)
            if not (same_d.all() and (q == gq).all() and fol == gf):
                cols = sorted(set(np.nonzero(~same_d)[1].tolist()))
                bad.append((wi, golden["worlds"][wi]["kind"], step, cols[:6]))
    assert not bad, f"{len(bad)} (world, step) records differ, first {bad[:6]}"
