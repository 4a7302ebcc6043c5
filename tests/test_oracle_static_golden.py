"""The oracle's static collidables (capsules and oriented boxes: their filing, their place in the gradient's walk, their
responses) pinned to the reference: every ball of the static collidable worlds (tests/scenarios.py `static_cases`) as
added and after every tick, bit for bit (two NaNs equal) with mLastC, the filing, the collision lists and the display
yaw, pitch and roll, and each static collidable's boxes, against the vectors the reference's own FullBallpark wrote
(tests/golden/make_golden_static.py); and with dynamical orientation (static_do_golden.npz)."""

from collections import defaultdict

import numpy as np
import pytest

import golden_fixture as G
import oracle as O
from park_driver import _same_bits


def _static_boxes(name):
    z = np.load(G.GOLDEN / name, allow_pickle=False)
    sw, si, so, sk = z["st_world"], z["st_index"], z["st_off"], z["st_keys"]
    return {(int(sw[k]), int(si[k])): sk[so[k]:so[k + 1]] for k in range(len(sw))}


def run(w, do=False, boxes=None):
    o = O.World(w["friction"], w["tick"], w["new_orbit"], w["start"], do, True, not do)
    for b in w["balls"]:
        o.add(b)
    for s in w["statics"]:
        o.add_static(s)
    if boxes is not None:
        boxes.extend(o.static_boxes(k) for k in range(len(w["statics"])))
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
        tail = o.turning() if do else o.ypr()
        parts = [d[:, :26], lc, tail] if not do else [d[:, :26], lc, tail]
        out.append((step, np.concatenate(parts, axis=1), np.concatenate([q[:, G.Q_FROM_ORACLE], filing], axis=1),
                    hits))
    return out


def _check(golden, do):
    bad = []
    for wi in golden["kept"]:
        for step, d, q, hits in run(golden["worlds"][wi], do):
            gd, gq, gh = golden["records"][(wi, step)]
            if do:                                  # the fixture's turning columns follow mLastC
                gd = gd
            same_d = _same_bits(d, gd)
            if not (same_d.all() and (q == gq).all() and hits == gh):
                cols = sorted(set(np.nonzero(~same_d)[1].tolist()))
                bad.append((wi, golden["worlds"][wi]["kind"], step, cols[:6], hits != gh))
    return bad


@pytest.fixture(scope="module")
def golden():
    return G.load("static_golden.npz")


def test_oracle_static_filing_matches_reference(golden):
    want = _static_boxes("static_golden.npz")
    bad = []
    for wi in golden["kept"]:
        boxes = []
        run(dict(golden["worlds"][wi], steps=0), boxes=boxes)
        for j, got in enumerate(boxes):
            if not np.array_equal(got, want[(wi, j)]):
                bad.append((wi, j, len(got), len(want[(wi, j)])))
    assert not bad, f"{len(bad)} filings differ, first {bad[:6]}"
    assert max(len(v) for v in want.values()) > 20          # long capsules span many boxes


def test_oracle_statics_match_reference(golden):
    bad = _check(golden, False)
    assert not bad, f"{len(bad)} (world, step) records differ, first {bad[:6]}"


def test_oracle_statics_with_orientation_match_reference():
    bad = _check(G.load("static_do_golden.npz"), True)
    assert not bad, f"{len(bad)} (world, step) records differ, first {bad[:6]}"


def test_fixture_covers_the_statics(golden):
    sids = {(wi, s["id"]) for wi in golden["kept"] for s in golden["worlds"][wi]["statics"]}
    hit_kinds = set()
    n = 0
    for (wi, _), (_, _, lists) in golden["records"].items():
        if any((wi, x) in sids for lst in lists for x in lst):
            n += 1
            hit_kinds.add(golden["worlds"][wi]["kind"])
    assert n > 300
    assert {"capsule_through", "capsule_cap", "capsule_inside", "capsule_rest", "capsule_axis", "box_face",
            "box_corner", "box_inside", "box_centre", "box_huge", "two_statics", "faces", "field"} <= hit_kinds
