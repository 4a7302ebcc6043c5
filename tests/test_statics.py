"""FullBallpark's static collidables (capsules and oriented boxes) through the drop-in park
(VecModesPark(collisions=True).load_statics), host build of the kernel source: against the reference's own vectors
(tests/golden/static_golden.npz: every ball of the static collidable worlds as added and after every tick, bit for bit
with mLastC, the filing, the collision lists and the display yaw, pitch and roll; static_do_golden.npz with dynamical
orientation), each static collidable's filing against the reference's, the same worlds against the oracle in
workgroups of 64 rows, and the refusals."""

import numpy as np
import pytest
import torch

import golden_fixture as G
import park_driver as D
import scenarios as S
from exact_ballpark_amd.vec_modes import VecModesPark
from test_oracle_static_golden import _static_boxes

Q_GOLDEN = list(range(8)) + list(range(11, 18))     # the park's ints as the fixture's: id .. interactive, the filing


def compare_golden(got, golden, kept, do=False):
    """The park's records against the fixture's; with dynamical orientation the park's columns (flight, turning,
    mLastC) in the fixture's order (flight, mLastC, turning)."""
    cols = list(range(26)) + [46, 47, 48] + list(range(26, 46)) if do else None
    bad = []
    for wi in kept:
        w = golden["worlds"][wi]
        for step in range(0, w["steps"] + 1):
            gd, gq, gl = got[(wi, step)]
            wd, wq, wl = golden["records"][(wi, step)]
            if cols is not None:
                gd = gd[:, cols]
            same_d = D._same_bits(gd, wd)
            same_q = gq[:, Q_GOLDEN] == wq
            hits = [h for _, h in gl]
            if not (same_d.all() and same_q.all() and hits == wl):
                diff = sorted(set(np.nonzero(~same_d)[1].tolist()))
                qdiff = sorted(set(np.nonzero(~same_q)[1].tolist()))
                bad.append((wi, w["kind"], step, diff[:6], qdiff[:4], hits if hits != wl else None, wl))
    return bad


@pytest.fixture(scope="module")
def golden():
    return G.load("static_golden.npz")


@pytest.mark.parametrize("block", [256, 64])
def test_park_matches_reference_vectors(golden, block):
    kept = golden["kept"]
    got = D.run_park(golden["worlds"], "cpu", order=kept, initial=True, block=block)
    bad = compare_golden(got, golden, kept)
    assert not bad, f"{len(bad)} (world, step) records differ, first {bad[:6]}"


def test_park_with_orientation_matches_reference_vectors():
    golden = G.load("static_do_golden.npz")
    kept = golden["kept"]
    got = D.run_park(golden["worlds"], "cpu", order=kept, initial=True, block=128)
    bad = compare_golden(got, golden, kept, do=True)
    assert not bad, f"{len(bad)} (world, step) records differ, first {bad[:6]}"


def test_filing_matches_reference(golden):
    want = _static_boxes("static_golden.npz")
    kept = golden["kept"]
    balls = [(wi, b) for wi in kept for b in golden["worlds"][wi]["balls"]]
    park = VecModesPark(len(balls), device="cpu", collisions=True)
    park.load([{"id": b["id"], "world": wi, "time": 0, "mass": b["mass"], "radius": b["radius"],
                "max_vel": b["max_vel"], "pos": b["pos"], "vel": b["vel"], "agility": b["agility"],
                "speed_fraction": b["sf"], "free": b["free"], "massive": b["massive"]} for wi, b in balls])
    park.load_statics({wi: golden["worlds"][wi]["statics"] for wi in kept})
    for wi in kept:
        for j in range(len(golden["worlds"][wi]["statics"])):
            assert np.array_equal(park.static_boxes(wi, j), want[(wi, j)]), (wi, j)


def test_worlds_match_oracle_reversed():
    ws = S.static_cases(3)
    idx = list(reversed(range(0, len(ws), 2)))
    got = D.run_park(ws, "cpu", order=idx, block=64)
    bad = D.compare(got, D.run_oracle(ws, idx), ws)
    assert not bad, f"{len(bad)} differ: {bad}"


def _one(**kw):
    p = VecModesPark(2, device="cpu", collisions=True, **kw)
    p.load([{"id": k + 1, "world": 0, "time": 0, "mass": 1e6, "radius": 10.0, "max_vel": 100.0,
             "pos": [0.0, 0.0, 50.0 * k], "vel": [0.0, 0.0, 0.0], "agility": 1.0, "speed_fraction": 1.0,
             "free": True, "massive": True} for k in range(2)])
    return p


def test_refusals():
    cap = {"kind": "capsule", "id": 5, "parent": 0, "a": [0.0, 0.0, 0.0], "b": [100.0, 0.0, 0.0], "radius": 10.0}
    with pytest.raises(ValueError):
        VecModesPark(2, device="cpu").load_statics([[cap]])
    p = _one()
    with pytest.raises(ValueError):
        p.load_statics([[cap, dict(cap)]])                    # ids not distinct
    with pytest.raises(ValueError):
        p.load_statics([[dict(cap, kind="cone")]])
    with pytest.raises(ValueError):
        p.load_statics([[dict(cap, id=k) for k in range(65)]])
    with pytest.raises(ValueError):                          # a walk that never ends: 2r below half an ulp of |AB|
        p.load_statics([[dict(cap, a=[1e12, 0.0, 0.0], b=[-1e12, 0.0, 0.0], radius=1e-3)]])
    p.load_statics([[cap]])
    assert p._hits.shape[0] == 3 * (2 - 1) + 1 + 1
    p.load_statics([])
    assert p._st is None and p._hits.shape[0] == 4
