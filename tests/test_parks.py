"""The reference's other park classes at the drop-in boundary, constructed as the reference constructs them, on the
host build, against the reference's own vectors: VecComposedPark(n, friction, tick_interval_ms, device,
dynamical_orientation, use_new_orbit, block, master, disable_do_for_missiles, collisions) on the missile worlds
(missile_golden.npz, missile_do_golden.npz); VecCombatPark on the formation worlds (formation_golden.npz), its domain
flags (`watch`, `check`) raised where a world holds two massive balls; VecCollidePark(worlds, slots, ...) with its
[W, N] layout, `load` of per-world ball lists and [W, N] commands on the collision worlds whose commands it takes
(collide_golden.npz: GotoPoint, GotoDirection, Stop, SetBallVelocity); and balls added mid-run (add_balls) against the
oracle's AddBall at the same clock."""

import numpy as np
import pytest
import torch

import golden_fixture as G
import oracle as O
import park_driver as D
import scenarios as S
from exact_ballpark_amd.vec_collide import VecCollidePark
from exact_ballpark_amd.vec_combat import FLAG_MASSIVE, VecCombatPark, VecComposedPark
from exact_ballpark_amd.vec_modes import EMPTY_ID, VecModesPark
from test_statics import compare_golden


def _composed(n, friction, tick, device, new_orbit, block, dynamical, coll, ypr, master):
    return VecComposedPark(n, friction=friction, tick_interval_ms=tick, device=device,
                           dynamical_orientation=dynamical, use_new_orbit=new_orbit, block=block, master=master,
                           disable_do_for_missiles=False, collisions=True)


def _combat(n, friction, tick, device, new_orbit, block, dynamical, coll, ypr, master):
    return VecCombatPark(n, friction, tick, device, dynamical, new_orbit, block, master)


@pytest.mark.parametrize("name,do", [("missile_golden.npz", False), ("missile_do_golden.npz", True)])
def test_composed_park_matches_reference_vectors(name, do):
    golden = G.load(name)
    kept = golden["kept"]
    got = D.run_park(golden["worlds"], "cpu", order=kept, initial=True, block=128, make=_composed)
    bad = compare_golden(got, golden, kept, do=do)
    assert not bad, f"{len(bad)} (world, step) records differ, first {bad[:4]}"


def test_combat_park_matches_reference_vectors_and_flags_its_domain():
    golden = G.load("formation_golden.npz")
    kept = golden["kept"]
    got = D.run_park(golden["worlds"], "cpu", order=kept, initial=True, block=128, make=_combat)
    assert not compare_golden(got, golden, kept)
    # the combat park's domain: a world with two massive balls outside MISSILE mode is flagged, and check() raises
    park = VecCombatPark(4, device="cpu")
    park.load([{"id": k + 1, "world": 0, "mass": 1e7, "radius": 50.0, "max_vel": 100.0, "pos": [100.0 * k, 0, 0],
                "vel": [0.0, 0.0, 0.0], "agility": 1.0, "speed_fraction": 1.0, "massive": k < 2}
               for k in range(3)] + [{"world": 0, "exists": False}])
    assert park.exists.tolist() == [True, True, True, False]
    park.watch()
    assert park.world_flags() == [FLAG_MASSIVE]
    with pytest.raises(RuntimeError, match="two massive balls"):
        park.check()


def test_combat_park_lookup_takes_rows_and_ids():
    park = VecCombatPark(5, device="cpu")
    park.load([{"id": i, "world": w, "mass": 1e7, "radius": 10.0, "max_vel": 100.0, "pos": [0.0, 0.0, float(i)],
                "vel": [0.0, 0.0, 0.0], "agility": 1.0, "speed_fraction": 1.0}
               for i, w in ((7, 0), (9, 0), (7, 1), (3, 1))] + [{"world": 1, "exists": False}])
    got = park.lookup(torch.tensor([0, 1, 2, 3, 4, 2]), torch.tensor([9, 7, 7, 7, 3, 9]))
    assert got.tolist() == [1, 0, 2, 2, 3, -1]


def _collide_worlds(golden):
    ok = {S.GOTO_OP, S.DIR_OP, S.STOP_OP, S.VEL_OP}
    return [wi for wi in golden["kept"] if {c[1] for c in golden["worlds"][wi]["cmds"]} <= ok]


def _run_collide_park(golden, members, device="cpu"):
    """The worlds through VecCollidePark's own API, one park per (friction, tick, orbit law, start)."""
    groups = {}
    for wi in members:
        w = golden["worlds"][wi]
        groups.setdefault((w["friction"], w["tick"], w["new_orbit"], w["start"]), []).append(wi)
    out = {}
    for (friction, tick, new_orbit, start), ws in groups.items():
        N = max(len(golden["worlds"][wi]["balls"]) for wi in ws)
        vp = VecCollidePark(len(ws), N, friction=friction, tick_interval_ms=tick, device=device, kernel="cuda"
                            if device != "cpu" else "torch")
        vp.park.use_new_orbit = bool(new_orbit)
        vp.current_time = start
        vp.load([[{"id": b["id"], "mass": b["mass"], "radius": b["radius"], "max_vel": b["max_vel"],
                   "agility": b["agility"], "speed_fraction": b["sf"], "pos": b["pos"], "vel": b["vel"],
                   "free": b["free"], "massive": b["massive"], "junk": b.get("junk", False),
                   "cloaked": bool(b["cloaked"])} for b in golden["worlds"][wi]["balls"]] for wi in ws])
        slot = {}
        for k, wi in enumerate(ws):
            for s, b in enumerate(sorted(golden["worlds"][wi]["balls"], key=lambda b: b["id"])):
                slot[(wi, b["id"])] = (k, s)
        rows_of = {wi: [k * N + s for s in range(len(golden["worlds"][wi]["balls"]))] for k, wi in enumerate(ws)}
        for wi, v in D._snapshot(vp.park, ws, rows_of).items():
            out[(wi, 0)] = v
        steps = max(golden["worlds"][wi]["steps"] for wi in ws)
        for step in range(1, steps + 1):
            cmds = [(wi, c) for wi in ws for c in golden["worlds"][wi]["cmds"] if c[0] == step]
            for j in range(max([sum(1 for x in cmds if x[0] == wi) for wi in ws] + [0])):
                by_op = {}
                for wi in ws:
                    mine = [c for x, c in cmds if x == wi]
                    if j < len(mine):
                        by_op.setdefault(mine[j][1], []).append((wi, mine[j]))
                for op, items in by_op.items():
                    mask = torch.zeros(len(ws), N, dtype=torch.bool)
                    xyz = torch.zeros(len(ws), N, 3, dtype=torch.float64)
                    for wi, c in items:
                        k, s = slot[(wi, c[2])]
                        mask[k, s] = True
                        xyz[k, s] = torch.tensor(c[4:7], dtype=torch.float64)
                    if op == S.GOTO_OP:
                        vp.goto_point(mask, xyz)
                    elif op == S.DIR_OP:
                        vp.goto_direction(mask, xyz)
                    elif op == S.STOP_OP:
                        vp.stop(mask)
                    else:
                        vp.set_velocity(mask, xyz)
            vp.evolve()
            live = [wi for wi in ws if step <= golden["worlds"][wi]["steps"]]
            for wi, v in D._snapshot(vp.park, live, rows_of).items():
                out[(wi, step)] = v
        assert vp.pos.shape == (len(ws), N, 3) and vp.mod.shape == (len(ws), N, 3)
        assert bool((vp.id[vp.exists] == vp.park.ids[vp.park.ids != EMPTY_ID]).all())
    return out


def test_collide_park_matches_reference_vectors():
    golden = G.load("collide_golden.npz")
    members = _collide_worlds(golden)
    assert len(members) > 100
    got = _run_collide_park(golden, members)
    bad = compare_golden(got, golden, members)
    assert not bad, f"{len(bad)} (world, step) records differ, first {bad[:4]}"


def test_collide_park_views_write_through():
    vp = VecCollidePark(2, 3, device="cpu")
    vp.load([[{"id": 5, "mass": 1e7, "radius": 10.0, "max_vel": 100.0, "agility": 1.0, "speed_fraction": 1.0,
               "pos": [0.0, 0.0, 0.0], "vel": [1.0, 0.0, 0.0]}], []])
    assert vp.exists.tolist() == [[True, False, False], [False, False, False]]
    vp.speed_fraction = torch.full((2, 3), 0.5, dtype=torch.float64)
    assert float(vp.park.speed_fraction[0]) == 0.5
    with pytest.raises(NotImplementedError):
        VecCollidePark(2, 3, iterative=True)


def test_add_balls_mid_run_matches_oracle_addball():
    """Balls added after some ticks (the empty rows filled by add_balls) against the oracle's AddBall at that clock."""
    worlds = S.flight_park(3, 24, 3)
    added = {}
    rng = np.random.default_rng(4)
    for wi, w in enumerate(worlds):
        w["steps"] = 9
        added[wi] = {"id": 900 + wi, "mass": float(10 ** rng.uniform(6, 8)), "radius": float(rng.uniform(10, 300)),
                     "max_vel": float(rng.uniform(50, 900)), "pos": [float(x) for x in rng.uniform(-4e4, 4e4, 3)],
                     "vel": [float(x) for x in rng.uniform(-50, 50, 3)] if wi % 2 else [0.0, 0.0, 0.0],
                     "agility": float(rng.uniform(0.2, 3)), "sf": 1.0, "free": True, "massive": False,
                     "interactive": False, "cloaked": 0, "bubble": -1, "moribund": False, "mode": 2}
    for do in (False, True):
        rows, index = [], []
        for wi, w in enumerate(worlds):
            for b in w["balls"]:
                rows.append({"id": b["id"], "world": wi, "time": w["start"], "mass": b["mass"], "radius": b["radius"],
                             "max_vel": b["max_vel"], "pos": b["pos"], "vel": b["vel"], "agility": b["agility"],
                             "speed_fraction": b["sf"], "free": b["free"], "massive": b["massive"],
                             "interactive": b["interactive"], "cloaked": 0, "bubble": b["bubble"], "moribund": False,
                             "mode": 2})
            index.append(len(rows))
            rows.append(VecModesPark.empty_row(wi, w["start"]))
        park = VecModesPark(len(rows), device="cpu", dynamical_orientation=do, display_ypr=not do)
        park.load(rows)
        for _ in range(4):
            park.evolve()
        mask = torch.zeros(len(rows), dtype=torch.bool)
        mask[index] = True
        park.add_balls(mask, [{"id": added[wi]["id"], "mass": added[wi]["mass"], "radius": added[wi]["radius"],
                               "max_vel": added[wi]["max_vel"], "pos": added[wi]["pos"], "vel": added[wi]["vel"],
                               "agility": added[wi]["agility"], "speed_fraction": 1.0, "free": True}
                              for wi in range(len(worlds))])
        with pytest.raises(ValueError):
            park.add_balls(mask, [{"id": 1}] * len(worlds))        # no longer empty
        for _ in range(5):
            park.evolve()
        for wi, w in enumerate(worlds):
            o = O.World(w["friction"], w["tick"], w["new_orbit"], w["start"], dynamical=do, ypr=not do)
            for b in w["balls"]:
                o.add(b)
            for _ in range(4):
                o.evolve()
            o.add(added[wi])
            for _ in range(5):
                o.evolve()
            d, q, _ = o.state()
            rows_w = list(range(index[wi] - len(w["balls"]), index[wi] + 1))
            snap = D._snapshot(park, [wi], {wi: rows_w})[wi]
            gd = snap[0][:, :26]
            assert np.array_equal(gd.view(np.int64), np.ascontiguousarray(d[:, :26]).view(np.int64)), (wi, do)
            if do:
                assert np.array_equal(snap[0][:, 26:46].view(np.int64),
                                      np.ascontiguousarray(o.turning()).view(np.int64)), (wi, "turning")
            else:
                assert np.array_equal(snap[0][:, 26:35].view(np.int64),
                                      np.ascontiguousarray(o.ypr()).view(np.int64)), (wi, "ypr")
