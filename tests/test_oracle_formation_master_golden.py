"""The oracle's formation reservations on the master (oracle/oracle_modes.c: ReserveFormationSlot, FreeFormationSlot,
Stop(Ball*) freeing a FORMATION follower's slot in its leader, SetBallFormation(-1) clearing every reservation) pinned
to the reference: every ball of the master's formation worlds (tests/scenarios.py `formation_master_cases`) after the
load and after every tick, bit for bit with the filing and the display yaw, pitch and roll, each ball's reservation
bits and each ReserveFormationSlot's answer, with every master structure, against the vectors the reference's own
FullBallpark(is_master=True) wrote (tests/golden/make_golden_space.py --formations)."""

import json
from collections import defaultdict

import numpy as np
import pytest

import golden_fixture as G
import oracle as O
from park_driver import _same_bits

FIXTURE = "formation_master_golden.npz"


def load():
    g = G.load(FIXTURE)
    z = np.load(G.GOLDEN / FIXTURE, allow_pickle=False)
    g["master"] = dict(zip(g["kept"], json.loads(str(z["master"]))))
    return g


def run_oracle(w: dict):
    """[(step, doubles, ints, master state with the reservation bits and answers)] for a master formation world."""
    o = O.World(w["friction"], w["tick"], w["new_orbit"], w["start"], False, True, True)
    o.set_master(w["seed"])
    o.load_formations(w["formations"])
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
        m = o.master_state(w["queries"])
        m["fslots"] = o.formation_slots()
        m["reserved"] = o.take_reserved()
        out.append((step, np.concatenate([d[:, :26], lc, o.ypr()], axis=1),
                    np.concatenate([q[:, G.Q_FROM_ORACLE], filing], axis=1), m))
    return out


@pytest.fixture(scope="module")
def golden():
    return load()


def test_oracle_formation_reservations_match_reference(golden):
    bad = []
    for wi in golden["kept"]:
        w = golden["worlds"][wi]
        want = golden["master"][wi]
        for step, d, q, m in run_oracle(w):
            gd, gq, _ = golden["records"][(wi, step)]
            ok_d = d.shape == gd.shape and _same_bits(d, gd).all() and (q == gq).all()
            keys = [k for k in want[step] if m.get(k) != want[step][k]]
            if not ok_d or keys:
                bad.append((wi, w["kind"], step, "balls" if not ok_d else "", keys[:4]))
                break
    assert not bad, f"{len(bad)} worlds differ, first {bad[:6]}"


def test_fixture_covers_the_reservations(golden):
    """The vectors reach what they pin: reservations granted and refused, bits that Stop(Ball*) cleared (no
    FreeFormationSlot that step) in worlds where a follower stopped or went elsewhere, where the leader warped out
    (StopAllFollowers inside the tick) and where it left its formation."""
    granted = refused = 0
    cleared = defaultdict(int)
    for wi, m in golden["master"].items():
        w = golden["worlds"][wi]
        for step, rec in enumerate(m):
            granted += sum(1 for _, s in rec["reserved"] if s >= 0)
            refused += sum(1 for _, s in rec["reserved"] if s < 0)
            if step == 0:
                continue
            prev = dict(zip(m[step - 1]["ids"], m[step - 1]["fslots"]))
            by_hand = any(c[0] == step and c[1] == O.FREESLOT_OP for c in w["cmds"])
            for i, v in zip(rec["ids"], rec["fslots"]):
                if i in prev and prev[i] & ~v and not by_hand:
                    cleared[w["kind"]] += 1
    assert granted >= 100 and refused >= 5
    for kind in ("formation_master_stop", "formation_master_warp", "formation_master_leave",
                 "formation_master_goto", "formation_master_remove", "formation_master_fixed"):
        assert cleared[kind] >= 1, (kind, dict(cleared))
