"""The master's space layer on gfx950 (xb_space_tape and xb_space_query: one lane a world over world-interleaved
tables), bit for bit against the reference's own vectors (tests/golden/space_golden.npz), eagerly and as SpaceGraph
replays (the tick with its layer captured as a HIP graph)."""

import pytest

import space_driver as SD
from test_oracle_space_golden import load_space

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def golden():
    return load_space()


@pytest.mark.parametrize("graph", [False, True], ids=["eager", "graph"])
def test_space_layer_matches_reference_vectors(golden, graph):
    bad = []
    for members in SD.groups(golden["worlds"], golden["kept"]).values():
        print(f"space layer on the GPU: {len(members)} worlds, graph={graph}", flush=True)
        got = SD.run_space([golden["worlds"][i] for i in members], "cuda", graph=graph)
        bad += SD.compare({(members[k], s): v for (k, s), v in got.items()}, golden, members)
    assert not bad, f"{len(bad)} worlds differ, first {bad[:6]}"
