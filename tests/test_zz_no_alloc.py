"""The allocation-free tick (xb_modes_tick_noalloc: the general launch with the library's own static scratch) against
xb_modes_tick with the caller's scratch, bit for bit, through the C ABI: two calls in a row on one device, the second
on a smaller park (the static array keeps the first call's contents, which the second must not read); its refusals
(a tiled state, a stride above XB_NOALLOC_MAX_ROWS) before any launch."""

import ctypes

import pytest
import torch

from exact_ballpark_amd import _native as N
from exact_ballpark_amd import workloads as WL
from exact_ballpark_amd.vec_modes import VecModesPark

pytestmark = pytest.mark.gpu
DEV = "cuda"


def _park(worlds: int, seed: int, dynamical: bool = False) -> VecModesPark:
    wl = WL.turning_park(seed, worlds, 4) if dynamical else WL.flight_park(seed, worlds, 4)
    park = VecModesPark(wl["n"], device=DEV, block=256, dynamical_orientation=dynamical)
    WL.load_flight_park(park, wl)
    park.tiles, park._state = None, None               # the general shape (any row order)
    return park


def _blocks(park: VecModesPark) -> list:
    return [b[:, :park.n].clone() for b in (park._f, park._q, park._b)]


def _same(a: list, b: list) -> bool:
    return all(torch.equal(x.view(torch.int64) if x.dtype == torch.float64 else x,
                           y.view(torch.int64) if y.dtype == torch.float64 else y) for x, y in zip(a, b))


def _tick(park: VecModesPark, noalloc: bool) -> None:
    s = park.state()
    stream = torch.cuda.current_stream().cuda_stream
    fn = N.hip().xb_modes_tick_noalloc if noalloc else N.hip().xb_modes_tick
    N.check(fn(ctypes.byref(s), stream, 256), "xb_modes_tick_noalloc" if noalloc else "xb_modes_tick")


@pytest.mark.parametrize("dynamical", [False, True], ids=["flight", "turning"])
def test_noalloc_equals_caller_scratch_twice(dynamical):
    # two calls in a row, the second on a smaller park: 16,384 worlds x 4 ships, then 1,000 worlds x 4
    for worlds, ticks in ((16384, 6), (1000, 9)):
        print(f"noalloc: {worlds} worlds x 4 ships, {ticks} ticks, dynamical={dynamical}", flush=True)
        want, got = _park(worlds, 3, dynamical), _park(worlds, 3, dynamical)
        assert _same(_blocks(want), _blocks(got))
        for _ in range(ticks):
            _tick(want, False)
            _tick(got, True)
        torch.cuda.synchronize()
        assert _same(_blocks(want), _blocks(got)), f"{worlds} worlds: the allocation-free tick differs"


def test_noalloc_refusals():
    park = _park(64, 5)
    s = park.state()
    stream = torch.cuda.current_stream().cuda_stream
    tiled = VecModesPark(park.n, device=DEV, block=256)
    WL.load_flight_park(tiled, WL.flight_park(5, 64, 4))
    assert tiled.tiles is not None
    assert N.hip().xb_modes_tick_noalloc(ctypes.byref(tiled.state()), stream, 256) == -1
    big = N.ModesState.from_buffer_copy(s)
    big.stride = N.NOALLOC_MAX_ROWS + 64                # refused on the stride alone: nothing is launched
    rc = N.hip().xb_modes_tick_noalloc(ctypes.byref(big), stream, 256)
    assert rc == N.ERR_TOO_MANY_ROWS
    assert b"XB_NOALLOC_MAX_ROWS" in N.hip().xb_error_string(rc)
    before = _blocks(park)
    torch.cuda.synchronize()
    assert _same(before, _blocks(park))
