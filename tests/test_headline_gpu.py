"""The benchmark's own shapes on gfx950, bit for bit:

- the ragged composed park at its full size (65,536 worlds, ships geometric on 1..32, 0..8 missiles launched at tick
  1, dynamical orientation, gradient collisions) in the headline's launch shape (workgroups of 64 rows, lanes sorted
  by branch key, 16 ticks a launch, worlds regrouped by the learned scheduler trained on the park's own ticks) against
  the plain launch (workgroups of 256 rows in row order, one tick a launch) over the whole park, and sampled worlds
  against the oracle;
- the composed park with the lanes placed by each row's last orientation search steps (schedule 3) against the plain
  launch;
- the flight-modes park of 1,024 worlds x 4 ships with the new orbit law (use_new_orbit) for 100 ticks against the
  oracle, the old law's run beside it in tests/test_modes_gpu.py.
"""

import pytest
import torch

import park_driver as D
import scenarios as S
from exact_ballpark_amd import scheduler as SC
from exact_ballpark_amd import workloads as WL
from exact_ballpark_amd.vec_modes import VecModesPark

pytestmark = pytest.mark.gpu
DEV = "cuda"


def _load(wl, **kw):
    park = VecModesPark(wl["n"], device=DEV, collisions=True, dynamical_orientation=wl["dynamical"], **kw)
    WL.load_flight_park(park, wl)
    return park


def _blocks(park):
    return [b[:, :park.n].clone() for b in (park._f, park._q, park._b, park._cf, park._cq)]


def test_composed_full_size_headline_shape_equals_plain_and_oracle():
    wl = WL.composed_park(1, 65536, missiles=8)
    plain, shaped = _load(wl, block=256), _load(wl, block=64)
    shaped.schedule = 1
    shaped._set_tiles()
    shaped.record_paths(True)
    for _ in range(4):
        plain.evolve()
        shaped.evolve()
    print(f"headline shape: {wl['n']} rows, training the scheduler", flush=True)
    L = SC.LearnedScheduler(shaped, seed=0)
    L.train(ticks=8, epochs=4)                      # 8 ticks of the park, recorded
    for _ in range(8):
        plain.evolve()
    L.regroup()
    assert shaped.rowmap is not None
    for _ in range(2):
        shaped.evolve_ticks(16)
    for _ in range(32):
        plain.evolve()
    torch.cuda.synchronize()
    for k, (x, y) in enumerate(zip(_blocks(plain), _blocks(shaped))):
        assert torch.equal(x.view(torch.uint8), y.view(torch.uint8)), f"block {k} differs"
    ticks = 4 + 8 + 32
    for wi in range(0, 65536, 1311):
        sc = WL.world_scenario(wl, wi, ticks)
        want = D.run_oracle([sc], [0], record={ticks})[(0, ticks)]
        got = D._snapshot(shaped, [wi], {wi: list(WL.world_rows(wl, wi))})[wi]
        assert D._same_bits(got[0], want[0]).all() and (got[1] == want[1]).all() and got[2] == want[2], wi


def test_composed_search_steps_schedule_equals_plain():
    """Lanes placed by each row's last orientation search steps (schedule 3, the engine's own recorded counters), 16
    ticks a launch, against the plain launch: the same bits."""
    wl = WL.composed_park(2, 16384, missiles=8)
    plain, shaped = _load(wl, block=256), _load(wl, block=128)
    shaped.record_paths(True)
    for _ in range(3):
        plain.evolve()
        shaped.evolve()
    assert int(shaped.iters.max()) > 0
    shaped.schedule = 3
    print(f"composed park, {wl['n']} rows: lanes by search steps, 2 x 16 ticks a launch", flush=True)
    for _ in range(2):
        shaped.evolve_ticks(16)
    for _ in range(32):
        plain.evolve()
    torch.cuda.synchronize()
    for k, (x, y) in enumerate(zip(_blocks(plain), _blocks(shaped))):
        assert torch.equal(x.view(torch.uint8), y.view(torch.uint8)), f"block {k} differs"


def test_flight_park_new_orbit_matches_oracle():
    ws = S.flight_park(2, 1024)
    for w in ws:
        w["steps"] = 100
        w["new_orbit"] = True
    record = {1, 2, 10, 50, 100}
    want = D.run_oracle(ws, range(len(ws)), record)
    got = D.run_park(ws, DEV, record=record)
    bad = D.compare(got, want, ws)
    assert not bad, f"{len(bad)} differences: " + " | ".join(bad[:5])
    assert any(w["balls"][k]["mode"] == 4 or any(c[1] == S.ORBIT_OP for c in w["cmds"])
               for w in ws for k in range(len(w["balls"])))
