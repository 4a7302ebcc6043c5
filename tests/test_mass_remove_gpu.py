"""Crowded delayed removals on gfx950 (more balls due than one BringOutTheDeads call takes), bit for bit against the
reference's own vectors (tests/golden/mass_remove_golden.npz), in workgroups of 256 and 64 rows."""

import pytest

import golden_fixture as G
import park_driver as D
from test_statics import compare_golden

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("block", [256, 64])
def test_park_matches_reference_vectors_gpu(block):
    golden = G.load("mass_remove_golden.npz")
    kept = golden["kept"]
    got = D.run_park(golden["worlds"], "cuda", order=kept, initial=True, block=block)
    bad = compare_golden(got, golden, kept)
    assert not bad, f"{len(bad)} (world, step) records differ, first {bad[:6]}"
