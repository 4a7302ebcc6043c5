"""The oracle's missiles (LaunchMissile, MissileFollow, EvolveMissile, a missile's collision candidates: a normal
missile's target only, a defender's every nearby ball with the missiles kept; Gradient's missile skips; COL-711's
exception) pinned to the reference: every ball of the missile worlds (tests/scenarios.py `missile_cases`) as added and
after every tick, bit for bit (two NaNs equal) with mLastC, the filing, the collision lists and the display yaw, pitch
and roll, against the vectors the reference's own FullBallpark wrote (tests/golden/make_golden_missile.py); and with
dynamical orientation (missile_do_golden.npz)."""

from collections import defaultdict

import numpy as np
import pytest

import golden_fixture as G
import oracle as O
from park_driver import _same_bits


def run(w, do=False):
    o = O.World(w["friction"], w["tick"], w["new_orbit"], w["start"], do, True, not do)
    for b in w["balls"]:
        o.add(b)
    for mini in w.get("minis", []):
        o.add_mini(mini)
    if w.get("formations"):
        o.load_formations(w["formations"])
    by_step = defaultdict(list)
    for c in w["cmds"]:
        by_step[c[0]].append(c)
    out = []
    for step in range(0, w["steps"] + 1):
        if step:
            for c in by_step[step]:
                o.command(c)
            if "calls" in w:                        # an OnTick world: its step-th simulation time
                o.on_tick(w["calls"][step - 1], w.get("master", False))
            else:
                o.evolve()
        d, q, _ = o.state()
        lc, filing, hits = o.collide()
        tail = o.turning() if do else o.ypr()
        out.append((step, np.concatenate([d[:, :26], lc, tail], axis=1),
                    np.concatenate([q[:, G.Q_FROM_ORACLE], filing], axis=1), hits))
    return out


def check(golden, do):
    bad = []
    for wi in golden["kept"]:
        for step, d, q, hits in run(golden["worlds"][wi], do):
            gd, gq, gh = golden["records"][(wi, step)]
            same_d = _same_bits(d, gd)
            if not (same_d.all() and (q == gq).all() and hits == gh):
                diff = sorted(set(np.nonzero(~same_d)[1].tolist()))
                bad.append((wi, golden["worlds"][wi]["kind"], step, diff[:6], hits != gh))
                break
    return bad


@pytest.fixture(scope="module")
def golden():
    return G.load("missile_golden.npz")


def test_oracle_missiles_match_reference(golden):
    bad = check(golden, False)
    assert not bad, f"{len(bad)} worlds differ, first {bad[:6]}"


def test_oracle_missiles_with_orientation_match_reference():
    bad = check(G.load("missile_do_golden.npz"), True)
    assert not bad, f"{len(bad)} worlds differ, first {bad[:6]}"


def test_fixture_covers_the_missiles(golden):
    pursuing = hit_target = 0
    for (wi, step), (d, q, lists) in golden["records"].items():
        mode = q[:, 1]
        pursuing += int((mode == 5).sum())
        for k in np.flatnonzero(mode == 5):
            if q[k, 3] in lists[k]:                  # its target (follow id) in its collision list
                hit_target += 1
    assert pursuing > 2000 and hit_target > 50, (pursuing, hit_target)
