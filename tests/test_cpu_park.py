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


def test_cpu_resident_ticks_equal_single_ticks():
    import torch
    from exact_ballpark_amd import workloads as WL
    from exact_ballpark_amd.vec_modes import VecModesPark
    wl = WL.flight_park(2, 512)
    a = VecModesPark(wl["n"], device="cpu")
    b = VecModesPark(wl["n"], device="cpu")
    WL.load_flight_park(a, wl)
    WL.load_flight_park(b, wl)
    for _ in range(9):
        a.evolve()
    b.evolve_ticks(4)
    b.evolve_ticks(5)
    for x, y in zip((a._f, a._q, a._b), (b._f, b._q, b._b)):
        assert torch.equal(x.view(torch.int64) if x.dtype == torch.float64 else x,
                           y.view(torch.int64) if y.dtype == torch.float64 else y)


@pytest.fixture(scope="module")
def turning():
    g = G.load("orient_golden.npz")
    return g, D.run_oracle(g["worlds"], g["kept"])


@pytest.mark.parametrize("general", [False, True], ids=["tiled", "general"])
def test_cpu_park_dynamical_orientation_matches_oracle(turning, general):
    """The turning worlds (dynamical orientation, its setters and snaps) through the park on the host build, every
    field including the rotation, angular velocity, torque and roll, against the oracle (itself held to the reference's
    FullBallpark vectors)."""
    g, want = turning
    got = D.run_park(g["worlds"], "cpu", general=general)
    bad = D.compare(got, want, g["worlds"])
    assert not bad, f"{len(bad)} differences: " + " | ".join(bad[:5])


def test_cpu_turning_park_matches_oracle():
    """The turning park workload (dynamical orientation, rotations from yaw, pitch and roll, ships turning toward
    points of their cluster, orbiting and following) on the host build against the oracle."""
    from exact_ballpark_amd import workloads as WL
    from exact_ballpark_amd.vec_modes import VecModesPark
    wl = WL.turning_park(5, 96)
    park = VecModesPark(wl["n"], device="cpu", dynamical_orientation=True)
    WL.load_flight_park(park, wl)
    T = 30
    for _ in range(T):
        park.evolve()
    worlds = [WL.world_scenario(wl, wi, T) for wi in range(wl["worlds"])]
    want = D.run_oracle(worlds, range(len(worlds)), {T})
    rows_of = {wi: list(range(wi * 4, wi * 4 + 4)) for wi in range(wl["worlds"])}
    got = {(wi, T): v for wi, v in D._snapshot(park, list(range(wl["worlds"])), rows_of).items()}
    bad = D.compare(got, want, worlds)
    assert not bad, f"{len(bad)} differences: " + " | ".join(bad[:5])
    rw = park.rot[:, 3]
    assert (rw != 1.0).all() and (park.torque != 0.0).any()
