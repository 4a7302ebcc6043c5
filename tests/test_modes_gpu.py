"""The gfx950 flight-modes kernel against the oracle, bit for bit (two NaNs equal), through the drop-in park API:

- the golden worlds (every kind of scenario), every ball after every tick, in both launch shapes;
- the flight-modes park workload (1,024 worlds x 4 ships, 100 ticks, modes uniform over stop, goto, follow, orbit and
  warp), at recorded ticks, and against the host build of the same source;
- the tick's maths on the device against the host (sin, cos, log, exp, sincos; sqrt and division IEEE), and the
  float functions, pow, powf, asin and atan2 of the device against the host build of the same source;
- properties at full size (65,536 worlds x 4 ships): tiled and general launches identical, a run repeated identical, a
  world alone identical to the same world in the batch and in reversed order, block sizes 64, 128 and 256 identical;
- the HIP graph replay identical to eager ticks, with eager commands between replays; OnTick's masked tick.
"""

import math

import numpy as np
import pytest
import torch

import golden_fixture as G
import park_driver as D
import scenarios as S
from exact_ballpark_amd import _native as N
from exact_ballpark_amd import workloads as WL
from exact_ballpark_amd.vec_modes import ModesGraph, VecModesPark

pytestmark = pytest.mark.gpu
DEV = "cuda"


def _bits(a):
    return np.asarray(a, dtype=np.float64).view(np.uint64)


@pytest.fixture(scope="module")
def golden():
    return G.load()


@pytest.fixture(scope="module")
def golden_oracle(golden):
    return D.run_oracle(golden["worlds"], golden["kept"])


@pytest.mark.parametrize("name", ["sin", "cos", "log", "exp", "sincos_sin", "sincos_cos"])
def test_device_maths_equals_host(name):
    import oracle as O
    rng = np.random.default_rng(21)
    x = np.concatenate([rng.uniform(-10, 10, 100000), rng.uniform(-1e4, 1e4, 100000),
                        np.exp(rng.uniform(-700, 700, 100000)) * rng.choice([-1, 1], 100000),
                        rng.uniform(-1e8, 1e8, 50000), rng.uniform(1e8, 1e300, 20000),
                        [0.0, -0.0, math.inf, -math.inf, math.nan, 5e-324, float.fromhex("0x1.ad1fb54442d18p+0")]])
    if name == "log":
        x = np.abs(x)
    got = N.maths(name, torch.from_numpy(x).to(DEV)).cpu().numpy()
    want = O.maths(name, x)
    same = (_bits(got) == _bits(want)) | (np.isnan(got) & np.isnan(want))
    assert same.all(), f"{int((~same).sum())} differ, first at x = {x[~same][:3]}"


def test_device_sqrt_and_division_are_ieee():
    rng = np.random.default_rng(22)
    k = np.arange(1, 200001, dtype=np.float64)
    x = np.concatenate([np.exp(rng.uniform(-744, 709, 400000)), k * k, k * k + 1, k * k - 1,
                        rng.uniform(0, 1e-310, 10000), [0.0, -0.0, np.inf, 4.0]])
    got = N.maths("sqrt", torch.from_numpy(x).to(DEV)).cpu().numpy()
    assert (_bits(got) == _bits(np.sqrt(x))).all()
    y = np.exp(rng.uniform(-300, 300, x.size)) * rng.choice([-1, 1], x.size)
    q = N.maths("div", torch.from_numpy(x).to(DEV), torch.from_numpy(y).to(DEV)).cpu().numpy()
    with np.errstate(all="ignore"):
        assert (_bits(q) == _bits(x / y)).all()


UNARY_F = ("expf", "logf", "sinf", "cosf", "acosf", "atanf", "asin")
BINARY_F = ("atan2f", "pow", "atan2", "powf")


@pytest.mark.parametrize("name", UNARY_F + BINARY_F)
def test_device_float_and_inverse_maths_equal_host(name):
    """Seeded arguments over the functions' domains and their special values: random bit patterns (as floats for the
    float functions), the unit interval, wide magnitudes; the device result equal to libxb_cpu.so's bit for bit."""
    rng = np.random.default_rng(23)
    m = 200000
    raw = rng.integers(0, 2 ** 64, m, dtype=np.uint64).view(np.float64)
    if name in ("expf", "logf", "sinf", "cosf", "acosf", "atanf", "atan2f", "powf"):
        with np.errstate(invalid="ignore"):                 # signalling NaNs among the patterns
            raw = rng.integers(0, 2 ** 32, m, dtype=np.uint64).astype(np.uint32).view(np.float32).astype(np.float64)
    x = np.concatenate([raw, rng.uniform(-1, 1, m), rng.uniform(-100, 100, m),
                        np.exp(rng.uniform(-300, 300, m)) * rng.choice([-1, 1], m),
                        [0.0, -0.0, 1.0, -1.0, math.inf, -math.inf, math.nan, 5e-324, 0.25, 0.5]])
    y = None
    if name in BINARY_F:
        y = np.roll(x, 12345).copy()
        if name == "pow":
            x = np.abs(x)
            with np.errstate(invalid="ignore"):             # the infinities and NaNs stay NaN
                y = np.round(y) % 200 - 100
        if name == "powf":
            y[m:2 * m] = 0.25
    args = [torch.from_numpy(x)] + ([torch.from_numpy(y)] if y is not None else [])
    want = N.maths(name, *args).numpy()
    got = N.maths(name, *[a.to(DEV) for a in args]).cpu().numpy()
    same = (_bits(got) == _bits(want)) | (np.isnan(got) & np.isnan(want))
    assert same.all(), f"{int((~same).sum())} differ, first at x = {x[~same][:3]}"


@pytest.mark.parametrize("general", [False, True], ids=["tiled", "general"])
def test_golden_worlds_match_oracle(golden, golden_oracle, general):
    got = D.run_park(golden["worlds"], DEV, general=general)
    bad = D.compare(got, golden_oracle, golden["worlds"])
    assert not bad, f"{len(bad)} differences: " + " | ".join(bad[:5])


def test_golden_worlds_block_64(golden, golden_oracle):
    got = D.run_park(golden["worlds"], DEV, block=64)
    bad = D.compare(got, golden_oracle, golden["worlds"])
    assert not bad, f"{len(bad)} differences: " + " | ".join(bad[:5])


def _flight(worlds: int, steps: int, seed: int = 1) -> list:
    ws = S.flight_park(seed, worlds)
    for w in ws:
        w["steps"] = steps
    return ws


def test_flight_park_matches_oracle():
    ws = _flight(1024, 100)
    record = {1, 2, 10, 50, 100}
    want = D.run_oracle(ws, range(len(ws)), record)
    got = D.run_park(ws, DEV, record=record)
    bad = D.compare(got, want, ws)
    assert not bad, f"{len(bad)} differences: " + " | ".join(bad[:5])
    cpu = D.run_park(ws, "cpu", record={100})
    assert not D.compare(cpu, {k: v for k, v in got.items() if k[1] == 100}, ws)


def _big_park(worlds: int, block: int = 256, seed: int = 3, order=None) -> VecModesPark:
    wl = WL.flight_park(seed, worlds)
    if order is not None:
        wl = WL.select_worlds(wl, order)
    park = VecModesPark(wl["n"], device=DEV, block=block)
    WL.load_flight_park(park, wl)
    return park


def _blocks(park: VecModesPark) -> list:
    return [b[:, :park.n].clone() for b in (park._f, park._q, park._b)]


def _same(a: list, b: list) -> bool:
    return all(torch.equal(x.view(torch.int64) if x.dtype == torch.float64 else x,
                           y.view(torch.int64) if y.dtype == torch.float64 else y) for x, y in zip(a, b))


def test_full_size_tiled_equals_general_and_repeats():
    W, T = 65536, 20
    runs = []
    for general in (False, True, False):
        park = _big_park(W)
        assert park.tiles is not None
        if general:
            park.tiles, park._state = None, None
        for _ in range(T):
            park.evolve()
        torch.cuda.synchronize()
        runs.append(_blocks(park))
    assert _same(runs[0], runs[1]), "tiled and general launches differ"
    assert _same(runs[0], runs[2]), "a repeated run differs"


def test_batch_independence_and_block_sizes():
    W, T = 4096, 30
    base = _big_park(W)
    for _ in range(T):
        base.evolve()
    ref = _blocks(base)
    rev = list(range(W))[::-1]
    park = _big_park(W, order=rev)
    for _ in range(T):
        park.evolve()
    got = _blocks(park)
    # row 4k + j of the reversed park holds ship j of world W-1-k: compare through that permutation
    perm = torch.tensor([4 * wi + j for wi in rev for j in range(4)], device=DEV)
    keep_q = [0, 1, 2, 3, 4, 6, 7, 8, 9]                  # follow_ptr holds row numbers
    for a, b in zip(ref, got):
        x = a[:, perm]
        if x.dtype == torch.float64:
            x, b = x.view(torch.int64), b.view(torch.int64)
        elif x.dtype == torch.int64:
            x, b = x[keep_q], b[keep_q]
        assert torch.equal(x, b), "a world's result depends on where it sits in the batch"
    for block in (64, 128):
        p = _big_park(W, block=block)
        for _ in range(T):
            p.evolve()
        assert _same(_blocks(p), ref), f"block {block} differs"
    alone = _big_park(W, order=[5])
    for _ in range(T):
        alone.evolve()
    one = _blocks(alone)
    x = [b[:, 20:24] for b in ref]
    x[1] = x[1][keep_q]
    one[1] = one[1][keep_q]
    assert _same(x, one), "a world alone differs from the same world in the batch"


def test_graph_replay_equals_eager():
    W, T = 2048, 25
    eager = _big_park(W)
    graphed = _big_park(W)
    g = ModesGraph(graphed)
    assert _same(_blocks(eager), _blocks(graphed)), "building the graph changed the park"
    for t in range(T):
        if t == 10:                                  # eager commands between replays
            for p in (eager, graphed):
                m = torch.zeros(p.n, dtype=torch.bool, device=DEV)
                m[::7] = True
                p.stop(m)
                p.set_speed_fraction(m, torch.full((p.n,), 0.5, dtype=torch.float64, device=DEV))
        eager.evolve()
        g.tick()
        torch.cuda.synchronize()
        assert _same(_blocks(eager), _blocks(graphed)), f"replay {t} differs from the eager tick"


def test_masked_tick_and_graph():
    W = 1024
    full = _big_park(W)
    masked = _big_park(W)
    gm_park = _big_park(W)
    g = ModesGraph(gm_park, masked=True)
    active = (masked.world % 3 != 0)
    before = _blocks(masked)
    masked.evolve(active=active)
    g.tick(active)
    full.evolve()
    torch.cuda.synchronize()
    sel = active.cpu()
    for a, b, f, m in zip(before, _blocks(masked), _blocks(full), _blocks(gm_park)):
        a, b, f, m = (x.cpu() for x in (a, b, f, m))
        if a.dtype == torch.float64:
            a, b, f, m = (x.view(torch.int64) for x in (a, b, f, m))
        assert torch.equal(b[:, ~sel], a[:, ~sel]), "a row outside the mask changed"
        assert torch.equal(b[:, sel], f[:, sel]), "a row inside the mask differs from the unmasked tick"
        assert torch.equal(m, b), "the masked graph differs from the masked eager tick"


@pytest.mark.parametrize("schedule,ticks", [(1, 1), (2, 1), (0, 5), (1, 7)])
def test_schedules_and_resident_ticks_equal_plain(schedule, ticks):
    """Sorting a tile's lanes by branch key (built in, or random keys) and running several ticks per launch (rows held
    in registers) give the plain launch's bits, at full size."""
    W, T = 65536, 14
    plain = _big_park(W)
    for _ in range(T):
        plain.evolve()
    park = _big_park(W)
    park.schedule = schedule
    if schedule == 2:
        g = torch.Generator().manual_seed(5)
        park.keys = torch.randint(0, 16, (park.n,), generator=g, dtype=torch.uint8).to(DEV)
    done = 0
    while done < T:
        k = min(ticks, T - done)
        park.evolve_ticks(k)
        done += k
    torch.cuda.synchronize()
    assert _same(_blocks(plain), _blocks(park))
    assert park.current_time == plain.current_time == T


def test_ragged_park_tiled_padded_and_alone_agree():
    """The ragged park (ships per world geometric on 1..32, mean 6): packed tiles, the padded-to-32 layout, the lane
    schedule and a world taken alone give the same bits for every real row."""
    wl = WL.ragged_park(4, 8192)
    T = 12
    packed = VecModesPark(wl["n"], device=DEV, block=256)
    WL.load_flight_park(packed, wl)
    pad = WL.padded(wl, 32)
    padded = VecModesPark(pad["n"], device=DEV, block=256)
    WL.load_flight_park(padded, pad)
    sched = VecModesPark(wl["n"], device=DEV, block=128)
    WL.load_flight_park(sched, wl)
    sched.schedule = 1
    for p in (packed, padded, sched):
        assert p.tiles is not None
        for _ in range(T):
            p.evolve()
    torch.cuda.synchronize()
    a, b, c = _blocks(packed), _blocks(padded), _blocks(sched)
    dest = torch.as_tensor(pad["dest"], device=DEV)
    keep_q = [0, 1, 2, 3, 4, 6, 7, 8, 9]
    for x, y, z in zip(a, b, c):
        y = y[:, dest]
        if x.dtype == torch.float64:
            x, y, z = (t.view(torch.int64) for t in (x, y, z))
        elif x.dtype == torch.int64:
            x, y, z = x[keep_q], y[keep_q], z[keep_q]
        assert torch.equal(x, y), "the padded layout differs from the packed one"
        assert torch.equal(x, z), "the scheduled lanes differ from the plain ones"


# ---- dynamical orientation ----

@pytest.fixture(scope="module")
def turning():
    g = G.load("orient_golden.npz")
    return g, D.run_oracle(g["worlds"], g["kept"])


@pytest.mark.parametrize("general", [False, True], ids=["tiled", "general"])
def test_turning_worlds_match_oracle(turning, general):
    """The turning worlds (dynamical orientation: the torque and roll searches, ApplyTorque, the heading thrust, the
    snaps and the setters) on the device, every field against the oracle (held to the reference's FullBallpark)."""
    g, want = turning
    got = D.run_park(g["worlds"], DEV, general=general)
    bad = D.compare(got, want, g["worlds"])
    assert not bad, f"{len(bad)} differences: " + " | ".join(bad[:5])


def test_turning_park_full_size():
    """The turning park at 65,536 worlds x 4: the tiled and general launches and 8 ticks a launch identical; 256
    worlds evenly spaced against the oracle after 40 ticks."""
    W, T = 65536, 40
    wl = WL.turning_park(3, W)
    runs = []
    for shape in ("tiled", "general", "resident"):
        park = VecModesPark(wl["n"], device=DEV, dynamical_orientation=True)
        WL.load_flight_park(park, wl)
        if shape == "general":
            park.tiles, park._state = None, None
        if shape == "resident":
            for _ in range(T // 8):
                park.evolve_ticks(8)
        else:
            for _ in range(T):
                park.evolve()
        torch.cuda.synchronize()
        runs.append((park, _blocks(park)))
    assert _same(runs[0][1], runs[1][1]) and _same(runs[0][1], runs[2][1])
    park = runs[0][0]
    sample = np.linspace(0, W - 1, 256).astype(int).tolist()
    worlds = [WL.world_scenario(wl, wi, T) for wi in sample]
    want = D.run_oracle(worlds, range(len(worlds)), {T})
    rows_of = {k: list(range(wi * 4, wi * 4 + 4)) for k, wi in enumerate(sample)}
    snap = D._snapshot(park, list(range(len(sample))), rows_of)
    got = {(k, T): v for k, v in snap.items()}
    bad = D.compare(got, want, worlds)
    assert not bad, f"{len(bad)} differences: " + " | ".join(bad[:5])
