"""The oracle pinned to the reference: every ball of every golden world after every tick, bit for bit (two NaNs equal),
against the vectors the reference's own ModeBallpark wrote (tests/golden/make_golden_modes.py)."""

import numpy as np
import pytest

import golden_fixture as G
import oracle as O
from park_driver import _same_bits


@pytest.fixture(scope="module")
def golden():
    return G.load()


def test_fixture_covers_every_kind(golden):
    kinds = {golden["worlds"][wi]["kind"] for wi in golden["kept"]}
    assert kinds >= {"follow", "orbit_old", "orbit_new", "warp", "warpin", "guards"}
    assert len(golden["records"]) > 1000


def test_oracle_matches_reference_bit_for_bit(golden):
    bad = []
    for wi in golden["kept"]:
        for step, d, q, fol in O.run_world(golden["worlds"][wi]):
            gd, gq, gf = golden["records"][(wi, step)]
            ok = _same_bits(d[:, G.D_FROM_ORACLE], gd).all() and (q[:, G.Q_FROM_ORACLE] == gq).all() and fol == gf
            if not ok:
                bad.append((wi, step))
    assert not bad, f"{len(bad)} (world, step) records differ from the reference, first {bad[:5]}"


def test_planted_mistake_is_caught(golden, monkeypatch):
    """The comparison is not vacuous: the oracle's state with one last bit flipped in one position differs."""
    wi = golden["kept"][0]
    step, d, q, fol = next(iter(O.run_world(golden["worlds"][wi])))
    gd = golden["records"][(wi, step)][0]
    d2 = d[:, G.D_FROM_ORACLE].copy()
    d2.view(np.uint64)[0, 0] ^= np.uint64(1)
    assert not _same_bits(d2, gd).all()
