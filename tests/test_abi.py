"""The C ABI on the CPU: both libraries load and export every function include/xb.h declares (their own sets); the
loader's struct and the library's layout agree; refusals come back before any launch. No kernel is launched."""

import ctypes
import re
from pathlib import Path

import pytest
import torch  # noqa: F401

from exact_ballpark_amd import _native as N

HEADER = Path(__file__).resolve().parent.parent / "include" / "xb.h"


def declared() -> list:
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    return re.findall(r"^\s*(?:int|const char\*|void)\s+(xb_\w+)\s*\(", text, flags=re.M)


def test_header_declares_the_entry_points():
    names = declared()
    assert {"xb_abi", "xb_error_string", "xb_layout", "xb_modes_tick", "xb_maths_eval", "xb_modes_info"} <= set(names)


def test_hip_library_exports_every_declared_symbol():
    lib = N.hip()
    for name in declared():
        if name.startswith("xb_cpu_"):
            continue
        assert hasattr(lib, name), name
    assert lib.xb_abi() == N.ABI_VERSION


def test_cpu_library_exports_its_symbols():
    lib = N.cpu()
    names = [n for n in declared() if n.startswith("xb_cpu_")]
    assert {"xb_cpu_abi", "xb_cpu_modes_tick", "xb_cpu_maths_eval", "xb_cpu_pack_tiles"} <= set(names)
    for name in names:
        assert hasattr(lib, name), name


def test_pack_tiles():
    import numpy as np
    sizes = np.array([3, 5, 2, 6, 1, 4], dtype=np.int32)
    tiles = np.zeros(len(sizes) + 1, dtype=np.int32)
    nt = ctypes.c_int64()
    rc = N.cpu().xb_cpu_pack_tiles(sizes.ctypes.data, len(sizes), 8, tiles.ctypes.data, ctypes.byref(nt))
    assert rc == 0 and nt.value == 3
    assert tiles[:4].tolist() == [0, 8, 16, 21]              # [3 5] [2 6] [1 4]
    assert N.cpu().xb_cpu_pack_tiles(sizes.ctypes.data, len(sizes), 4, tiles.ctypes.data, ctypes.byref(nt)) == -1


def test_layout_matches_loader():
    out = (ctypes.c_int64 * 5)()
    N.hip().xb_layout(out)
    assert tuple(out) == (N.NF, N.NI, N.NB, N.NS, ctypes.sizeof(N.ModesState))


def test_refusals_before_launch():
    lib = N.hip()
    s = N.ModesState()
    s.n, s.stride = 4, 64
    assert lib.xb_modes_tick(ctypes.byref(s), None, 100) == -1          # block not a multiple of 64
    assert lib.xb_modes_tick(ctypes.byref(s), None, 512) == -1          # block above 256
    assert lib.xb_modes_tick(ctypes.byref(s), None, 128) == -1          # no state blocks
    s.stride = 2
    assert lib.xb_modes_tick(ctypes.byref(s), None, 128) == -1          # stride below n
    assert lib.xb_maths_eval(99, 1, None, None, None, 128) == -1
    assert lib.xb_error_string(-1).decode().startswith("refused")
