"""The master's formation reservations through the drop-in park (VecModesPark(master=True) with its space layer:
reserve_formation_slot, free_formation_slot, the slot that Stop(Ball*) of a FORMATION follower frees in its leader, by
a command or by StopAllFollowers inside the tick, SetBallFormation(-1) clearing every reservation, RemoveBall), host
build of the kernel source: bit for bit against the reference's own vectors (tests/golden/formation_master_golden.npz,
FullBallpark(is_master=True) on tests/scenarios.py's `formation_master_cases`): every ball's flight state with the
filing and the display orientation, each ball's reservation bits and each ReserveFormationSlot's answer, every master
structure, after the load and after every tick. One park a world (a park holds one formation table)."""

import pytest
import torch

import space_driver as SD
from test_oracle_formation_master_golden import load
from exact_ballpark_amd.vec_modes import VecModesPark


@pytest.fixture(scope="module")
def golden():
    return load()


def run(golden, device, members=None):
    bad = []
    for group in SD.groups(golden["worlds"], golden["kept"]).values():
        for wi in group if members is None else [m for m in group if m in members]:
            got = SD.run_space([golden["worlds"][wi]], device)
            bad += SD.compare({(wi, s): v for (_, s), v in got.items()}, golden, [wi])
    return bad


def test_master_reservations_match_reference_vectors(golden):
    bad = run(golden, "cpu")
    assert not bad, f"{len(bad)} worlds differ, first {bad[:6]}"


def test_reservation_refusals():
    """Destiny's undefined cases are refused with an error: a slot past 15 freed in a longer formation, a follower
    put in such a slot on the master (its Stop would free it); a ball without a formation reserves nothing."""
    p = VecModesPark(3, device="cpu", display_ypr=True, master=True)
    p.load([{"id": i + 1, "world": 0, "time": 0, "mass": 1e6, "radius": 50.0, "max_vel": 100.0, "pos": [0.0, 0.0, 0.0],
             "vel": [0.0, 0.0, 0.0], "agility": 1.0, "speed_fraction": 1.0, "free": True, "massive": False,
             "interactive": True, "cloaked": 0, "bubble": 0, "moribund": False, "mode": 2} for i in range(3)])
    assert p.reserve_formation_slot(0) == -1                            # no formations loaded
    p.load_formations([[[float(k), 0.0, 0.0] for k in range(20)], [[0.0, 0.0, 0.0]]])
    assert p.reserve_formation_slot(0) == -1                            # no formation
    one = torch.tensor([True, False, False])
    p.set_ball_formation(one, torch.tensor([1, 0, 0]))
    assert [p.reserve_formation_slot(0), p.reserve_formation_slot(0)] == [0, -1]
    p.free_formation_slot(0, 5)                                         # past the formation: nothing
    p.free_formation_slot(0, 0)
    assert int(p.formation_slots[0]) == 0
    p.set_ball_formation(one, torch.tensor([0, 0, 0]))
    assert p.reserve_formation_slot(0) == -1                            # more than 16 slots
    with pytest.raises(ValueError):
        p.free_formation_slot(0, 17)
    with pytest.raises(ValueError):
        p.formation_follow(torch.tensor([False, True, False]), torch.tensor([0, 1, 0]), torch.tensor([0, 16, 0]))
