"""The drop-in park on the CPU (the kernel's tick source built for the host, libxb_cpu.so) against the oracle, bit for
bit, on the golden worlds: the tiled launch shape (whole worlds per group) and the general one."""

import pytest

import golden_fixture as G
import park_driver as D


@pytest.fixture(scope="module")
def golden():
    return G.load()


@pytest.fixture(scope="module")
def oracle_states(golden):
    return D.run_oracle(golden["worlds"], golden["kept"])


@pytest.mark.parametrize("general", [False, True], ids=["tiled", "general"])
def test_cpu_park_matches_oracle(golden, oracle_states, general):
    worlds = golden["worlds"]
    got = D.run_park(worlds, "cpu", general=general)
    bad = D.compare(got, oracle_states, worlds)
    assert not bad, f"{len(bad)} differences: " + " | ".join(bad[:5])
