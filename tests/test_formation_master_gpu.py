"""The master's formation reservations on gfx950 (the command kernels' Stop(Ball*) and the tick's StopAllFollowers
clearing a FORMATION follower's slot with an atomic AND in its leader's word), bit for bit against the reference's
own vectors (tests/golden/formation_master_golden.npz), one park a world."""

import pytest

from test_formation_master import run
from test_oracle_formation_master_golden import load

pytestmark = pytest.mark.gpu


def test_master_reservations_match_reference_vectors_gpu():
    golden = load()
    print(f"formation reservations on the GPU: {len(golden['kept'])} worlds", flush=True)
    bad = run(golden, "cuda")
    assert not bad, f"{len(bad)} worlds differ, first {bad[:6]}"
