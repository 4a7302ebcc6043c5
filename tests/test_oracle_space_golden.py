"""The oracle's master (oracle/oracle_master.c: the top boxes and their bubbles, the bubble dictionaries, the
interactive and keep-alive sets, InitializeBubbles, UpdateBallBubble, CopyBubbles, the dead bubbles' skip,
HandleProximities with its sensors, notification ranges and target tracking in libstdc++'s unordered_map order,
glibc rand() for shuffled sensors, RemoveBall's partition side, and the queries) pinned to the reference: every ball
of the space worlds (tests/scenarios.py `space_cases`) after the load and after every tick, bit for bit with the
filing and the display yaw, pitch and roll, and every master structure and query answer, against the vectors the
reference's own FullBallpark(is_master=True) wrote (tests/golden/make_golden_space.py)."""

import json
from collections import defaultdict

import numpy as np
import pytest

import golden_fixture as G
import oracle as O
from park_driver import _same_bits


def load_space():
    g = G.load("space_golden.npz")
    z = np.load(G.GOLDEN / "space_golden.npz", allow_pickle=False)
    g["master"] = dict(zip(g["kept"], json.loads(str(z["master"]))))
    return g


def run_oracle_space(w: dict):
    """[(step, doubles, ints, master state)] for a space world on the oracle: the balls added, the step-0 commands,
    then every tick after its commands."""
    o = O.World(w["friction"], w["tick"], w["new_orbit"], w["start"], False, True, True)
    o.set_master(w["seed"])
    for b in w["balls"]:
        o.add(b)
    by_step = defaultdict(list)
    for c in w["cmds"]:
        by_step[c[0]].append(c)
    out = []
    for step in range(0, w["steps"] + 1):
        for c in by_step[step]:
            o.command(c)
        if step:
            o.evolve()
        d, q, _ = o.state()
        lc, filing, hits = o.collide()
        out.append((step, np.concatenate([d[:, :26], lc, o.ypr()], axis=1),
                    np.concatenate([q[:, G.Q_FROM_ORACLE], filing], axis=1), o.master_state(w["queries"])))
    return out


@pytest.fixture(scope="module")
def golden():
    return load_space()


def test_oracle_master_matches_reference(golden):
    bad = []
    for wi in golden["kept"]:
        w = golden["worlds"][wi]
        want = golden["master"][wi]
        for step, d, q, m in run_oracle_space(w):
            gd, gq, _ = golden["records"][(wi, step)]
            ok_d = d.shape == gd.shape and _same_bits(d, gd).all() and (q == gq).all()
            keys = [k for k in want[step] if m.get(k) != want[step][k]]
            if not ok_d or keys:
                bad.append((wi, step, "balls" if not ok_d else "", keys[:4]))
                break
    assert not bad, f"{len(bad)} worlds differ, first {bad[:6]}"


def test_fixture_covers_the_master(golden):
    recs = [r for m in golden["master"].values() for r in m]
    kinds = {e[0] for r in recs for e in r["events"]}
    assert kinds == {"DoProximity", "DoRange", "DoTargetTracking", "DoFinalCleanup"}
    assert any(r["remote"] for r in recs)
    assert max(r["bubble_id"] for r in recs) >= 3
    assert sum(1 for m in golden["master"].values() if len({len(r["tops"]) for r in m}) > 1) >= 30
    dead = sum(1 for r in recs for b in set(r["bubble"]) if b not in {x for x, _ in r["inter"]} | {x for x, _ in
                                                                                                    r["keep"]})
    assert dead >= 100
    assert any(a == -1 for r in recs for ds in r["dicts"] for _, a in ds[1])     # left, not yet copied
    assert any(s[2] > 13 for r in recs for s in r["sensors"]) or any(r["prox_bkt"][0] > 13 for r in recs)
