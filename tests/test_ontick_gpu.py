"""Ballpark::OnTick through the drop-in park on gfx950 (VecModesPark.on_tick: BringOutTheDeads first, per-world
interval counts, masked rounds, a master's resynchronisation), bit for bit against the reference's own vectors
(tests/golden/ontick_golden.npz), in workgroups of 256 and 64 rows."""

import pytest

import golden_fixture as G
import park_driver as D
from test_statics import compare_golden

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("block", [256, 64])
def test_park_on_tick_matches_reference_vectors(block):
    golden = G.load("ontick_golden.npz")
    kept = golden["kept"]
    got = D.run_park(golden["worlds"], "cuda", order=kept, initial=True, block=block, ontick=True)
    bad = compare_golden(got, golden, kept)
    assert not bad, f"{len(bad)} (world, call) records differ, first {bad[:6]}"
