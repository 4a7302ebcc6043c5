"""The reference's other park classes on gfx950, through the C ABI: VecComposedPark on the missile worlds (with and
without dynamical orientation) and VecCollidePark with its [W, N] layout on the collision worlds, against the
reference's own vectors; balls added mid-run on the GPU equal to the host build's."""

import pytest
import torch

import golden_fixture as G
import park_driver as D
import scenarios as S
from exact_ballpark_amd.vec_modes import VecModesPark
from test_parks import _collide_worlds, _composed, _run_collide_park
from test_statics import compare_golden

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name,do", [("missile_golden.npz", False), ("missile_do_golden.npz", True)])
def test_composed_park_gpu_matches_reference_vectors(name, do):
    print(f"VecComposedPark on cuda:0, {name}", flush=True)
    golden = G.load(name)
    kept = golden["kept"]
    got = D.run_park(golden["worlds"], "cuda:0", order=kept, initial=True, block=128, make=_composed)
    bad = compare_golden(got, golden, kept, do=do)
    assert not bad, f"{len(bad)} (world, step) records differ, first {bad[:4]}"


def test_collide_park_gpu_matches_reference_vectors():
    print("VecCollidePark on cuda:0, collide_golden.npz", flush=True)
    golden = G.load("collide_golden.npz")
    members = _collide_worlds(golden)
    got = _run_collide_park(golden, members, device="cuda:0")
    bad = compare_golden(got, golden, members)
    assert not bad, f"{len(bad)} (world, step) records differ, first {bad[:4]}"


@pytest.mark.parametrize("do", [False, True])
def test_add_balls_gpu_equals_host(do):
    worlds = S.flight_park(9, 256, 3)
    parks = {}
    for dev in ("cuda:0", "cpu"):
        rows, index = [], []
        for wi, w in enumerate(worlds):
            for b in w["balls"]:
                rows.append({"id": b["id"], "world": wi, "time": w["start"], "mass": b["mass"], "radius": b["radius"],
                             "max_vel": b["max_vel"], "pos": b["pos"], "vel": b["vel"], "agility": b["agility"],
                             "speed_fraction": b["sf"], "free": True, "massive": False, "interactive": False,
                             "cloaked": 0, "bubble": -1, "moribund": False, "mode": 2})
            index.append(len(rows))
            rows.append(VecModesPark.empty_row(wi, w["start"]))
        park = VecModesPark(len(rows), device=dev, dynamical_orientation=do, display_ypr=not do, collisions=True)
        park.load(rows)
        for _ in range(3):
            park.evolve()
        mask = torch.zeros(len(rows), dtype=torch.bool)
        mask[index] = True
        park.add_balls(mask, [{"id": 5000 + wi, "mass": 1e7 + wi, "radius": 40.0 + wi, "max_vel": 300.0,
                               "pos": [1000.0 * wi, -5.0 * wi, 2.5], "vel": [float(wi % 7), 3.0, -1.0],
                               "agility": 0.7, "speed_fraction": 1.0, "free": True, "massive": True}
                              for wi in range(len(worlds))])
        for _ in range(6):
            park.evolve()
        parks[dev] = park
    for a, b in zip(parks["cuda:0"]._blocks(), parks["cpu"]._blocks()):
        assert torch.equal(a.cpu(), b), "a block differs"
