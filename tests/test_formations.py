"""Formations through the drop-in park (VecModesPark(collisions=True): load_formations, set_ball_formation,
formation_follow, the FORMATION tick), host build of the kernel source: against the reference's own vectors
(tests/golden/formation_golden.npz with the display yaw, pitch and roll, formation_do_golden.npz with dynamical
orientation: every ball of the formation worlds as added and after every tick, bit for bit with mLastC, the filing and
the collision lists), and against the oracle in workgroups of 64 rows in reverse world order."""

import pytest
import torch

import golden_fixture as G
import park_driver as D
import scenarios as S
from exact_ballpark_amd.vec_modes import VecModesPark
from test_statics import compare_golden


@pytest.mark.parametrize("block", [256, 64])
def test_park_matches_reference_vectors(block):
    golden = G.load("formation_golden.npz")
    kept = golden["kept"]
    got = D.run_park(golden["worlds"], "cpu", order=kept, initial=True, block=block)
    bad = compare_golden(got, golden, kept)
    assert not bad, f"{len(bad)} (world, step) records differ, first {bad[:6]}"


def test_park_with_orientation_matches_reference_vectors():
    golden = G.load("formation_do_golden.npz")
    kept = golden["kept"]
    got = D.run_park(golden["worlds"], "cpu", order=kept, initial=True, block=128)
    bad = compare_golden(got, golden, kept, do=True)
    assert not bad, f"{len(bad)} (world, step) records differ, first {bad[:6]}"


def test_worlds_match_oracle_reversed():
    ws = [w for w in S.formation_cases(7)[:48]          # a negative slot is Destiny's undefined behaviour
          if not any(c[1] == S.FFOLLOW_OP and c[8] < 0 for c in w["cmds"])]
    for w in ws:
        w["ypr"] = True
    idx = list(reversed(range(len(ws))))
    got = D.run_park(ws, "cpu", order=idx, block=64)
    bad = D.compare(got, D.run_oracle(ws, idx), ws)
    assert not bad, f"{len(bad)} differ: {bad}"


def _park(**kw):
    p = VecModesPark(2, device="cpu", collisions=True, **kw)
    p.load([{"id": 10 + i, "world": 0, "time": 0, "mass": 1e6, "radius": 50.0, "max_vel": 300.0,
             "pos": [1000.0 * i, 0.0, 0.0], "vel": [0.0, 0.0, 0.0], "agility": 1.0, "speed_fraction": 1.0,
             "free": True, "massive": True, "interactive": False, "cloaked": 0, "bubble": 1, "moribund": False,
             "mode": 2} for i in range(2)])
    return p


def test_undefined_and_unsupported_calls_are_refused():
    p = _park(display_ypr=True)
    one = torch.tensor([True, False])
    with pytest.raises(ValueError):
        p.set_ball_formation(one, torch.tensor([0, 0]))              # no table loaded
    p.load_formations([[[0.0, 100.0, 0.0]], [[1.0, 0.0, 0.0], [0.0, 0.0, 1.0]]])
    p.set_ball_formation(one, torch.tensor([1, 0]))
    with pytest.raises(ValueError):
        p.set_ball_formation(one, torch.tensor([-2, 0]))             # a negative formation read
    with pytest.raises(ValueError):
        p.formation_follow(torch.tensor([False, True]), torch.tensor([0, 10]), torch.tensor([0, -1]))
    p.formation_follow(torch.tensor([False, True]), torch.tensor([0, 10]), torch.tensor([0, 1]))
    assert int(p.mode[1]) == 12 and int(p.follow_ptr[1]) == 0 and int(p.effect_stamp[1]) == 1
    p.set_ball_formation(one, torch.tensor([0, 0]))                  # shrink: slot 1 is gone
    assert int(p.mode[1]) == 2
    q = _park()
    q.load_formations([[[0.0, 100.0, 0.0]]])
    with pytest.raises(ValueError):
        q.formation_follow(torch.tensor([False, True]), torch.tensor([0, 10]), torch.tensor([0, 0]))
