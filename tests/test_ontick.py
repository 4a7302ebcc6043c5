"""Ballpark::OnTick through the drop-in park (VecModesPark.on_tick, host build of the kernel source) against the
reference's own vectors (tests/golden/ontick_golden.npz: FullBallpark driven by OnTick alone on delayed-removal
worlds, every ball as added and after every call, bit for bit with the clock, mLastC, the filing and the collision
lists): BringOutTheDeads before the interval check (delayed removals brought out by OnTick itself, more than seven
balls due within two ticks at once), too-early calls, a world's first call, whole intervals with the remainder carried,
and a master's backlog above five intervals run as one; the oracle's OnTick against the same vectors."""

import numpy as np
import pytest

import golden_fixture as G
import park_driver as D
from test_statics import compare_golden


@pytest.fixture(scope="module")
def golden():
    return G.load("ontick_golden.npz")


@pytest.mark.parametrize("block", [256, 64])
def test_park_on_tick_matches_reference_vectors(golden, block):
    kept = golden["kept"]
    got = D.run_park(golden["worlds"], "cpu", order=kept, initial=True, block=block, ontick=True)
    bad = compare_golden(got, golden, kept)
    assert not bad, f"{len(bad)} (world, call) records differ, first {bad[:6]}"


def test_fixture_covers_the_cases(golden):
    ws = [golden["worlds"][k] for k in golden["kept"]]
    interval = {id(w): w["tick"] * 10000 for w in ws}
    gaps = [(b - a) / interval[id(w)] for w in ws for a, b in zip([0] + w["calls"][:-1], w["calls"])]
    assert min(gaps) < 1.0 and max(gaps) > 6.0                  # too early, and backlogs past five intervals
    assert any(w["master"] for w in ws) and not all(w["master"] for w in ws)
    ticks = [np.diff(w["ticks_after"]) for w in ws]
    assert any((t == 0).any() for t in ticks) and any((t > 1).any() for t in ticks)
    master_backlog = [t for w, t in zip(ws, ticks) if w["master"]]
    assert all((t <= 5).all() for t in master_backlog)          # a master never runs more than five in a call
    assert any(w["kind"].endswith("remove_buffer_ontick") for w in ws)


def test_oracle_on_tick_matches_reference_vectors(golden):
    from test_oracle_missile_golden import check
    bad = check(golden, False)
    assert not bad, f"{len(bad)} worlds differ, first {bad[:6]}"
