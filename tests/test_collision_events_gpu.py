"""DoCollision events on gfx950 (`collision_events` after each one-tick launch of the collision tick kernels) against
the events the reference's own FullBallpark emitted on the missile worlds (tests/golden/missile_golden.npz and
missile_do_golden.npz), bit for bit."""

import pytest

import golden_fixture as G
import park_driver as D
from test_collision_events import golden_events, same

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["missile_golden.npz", "missile_do_golden.npz"])
def test_park_events_match_reference(name):
    golden, want = G.load(name), golden_events(name)
    got: dict = {}
    D.run_park(golden["worlds"], "cuda", order=golden["kept"], block=64, events=got)
    bad = [wi for wi in golden["kept"] if not same(got.get(wi, []), want.get(wi, []))]
    assert not bad, f"{len(bad)} worlds differ, first {bad[:6]}"
