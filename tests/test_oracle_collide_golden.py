"""The oracle's partition and gradient collisions pinned to the reference: every ball of every collision world
(tests/scenarios.py collide_cases) as added and after every tick, bit for bit (two NaNs equal): position .. last
collision, old position and velocity, the collision acceleration (mLastC), the filing (level, key, sides) and the
collision list in order, against the vectors the reference's own SpaceBallpark (a client's view, the gradient
method) wrote (tests/golden/make_golden_collide.py)."""

from collections import defaultdict

import numpy as np
import pytest

import golden_fixture as G
import oracle as O
from park_driver import _same_bits


@pytest.fixture(scope="module")
def golden():
    return G.load("collide_golden.npz")


def run(w):
    """[(step, doubles [n, 29], ints [n, 15], collision lists)] at step 0 and after every tick."""
    o = O.World(w["friction"], w["tick"], w["new_orbit"], w["start"], False, True)
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
        out.append((step, np.concatenate([d[:, :26], lc], axis=1),
                    np.concatenate([q[:, G.Q_FROM_ORACLE], filing], axis=1), hits))
    return out


def test_fixture_covers_the_collisions(golden):
    kinds = {golden["worlds"][wi]["kind"] for wi in golden["kept"]}
    assert kinds >= {"head_on", "overlap", "spheres", "levels", "symmetric", "follow", "orbit", "warp", "dropout"}
    d = np.concatenate([r[0] for r in golden["records"].values()])
    q = np.concatenate([r[1] for r in golden["records"].values()])
    lists = [h for r in golden["records"].values() for h in r[2]]
    assert (d[:, 26:29] != 0.0).any(axis=1).sum() > 500                   # mLastC set
    assert sum(len(h) > 1 for h in lists) > 80                             # several contacts in a tick
    assert len(set(q[:, 8].tolist())) >= 4 and (q[:, 12:15] != 0).any(axis=1).mean() > 0.2   # levels; sides


def test_oracle_collisions_match_reference(golden):
    bad = []
    for wi in golden["kept"]:
        for step, d, q, hits in run(golden["worlds"][wi]):
            gd, gq, gh = golden["records"][(wi, step)]
            same_d = _same_bits(d, gd)
            if not (same_d.all() and (q == gq).all() and hits == gh):
                cols = sorted(set(np.nonzero(~same_d)[1].tolist()))
                qcols = sorted(set(np.nonzero(q != gq)[1].tolist()))
                bad.append((wi, golden["worlds"][wi]["kind"], step, cols[:6], qcols[:6], hits != gh))
    assert not bad, f"{len(bad)} (world, step) records differ, first {bad[:6]}"


@pytest.fixture(scope="module")
def golden_do():
    return G.load("collide_do_golden.npz")


def run_do(w):
    """As run, with dynamical orientation: the turning state after mLastC."""
    o = O.World(w["friction"], w["tick"], w["new_orbit"], w["start"], True, True)
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
        out.append((step, np.concatenate([d[:, :26], lc, o.turning()], axis=1),
                    np.concatenate([q[:, G.Q_FROM_ORACLE], filing], axis=1), hits))
    return out


def test_oracle_collisions_with_orientation_match_reference(golden_do):
    """The same worlds with dynamical orientation against the reference's FullBallpark(dynamical_orientation=True):
    the heading thrust's mLastG feeds the gradient, the rotation a tick on commits beside the collision response."""
    bad = []
    turned = 0
    for wi in golden_do["kept"]:
        for step, d, q, hits in run_do(golden_do["worlds"][wi]):
            gd, gq, gh = golden_do["records"][(wi, step)]
            same_d = _same_bits(d, gd)
            turned += int((gd[:, 29] != 1.0).sum())
            if not (same_d.all() and (q == gq).all() and hits == gh):
                cols = sorted(set(np.nonzero(~same_d)[1].tolist()))
                bad.append((wi, golden_do["worlds"][wi]["kind"], step, cols[:6], hits != gh))
    assert not bad, f"{len(bad)} (world, step) records differ, first {bad[:6]}"
    assert turned > 1000
