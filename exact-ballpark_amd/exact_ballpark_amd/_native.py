"""Loading the native libraries: libxb_hip.so (the gfx950 kernels behind include/xb.h) and libxb_cpu.so (the same tick
source built for the host). The ctypes side of the C ABI: `ModesState` mirrors `xb_modes_state` field for field, and
the library's own layout (xb_layout) is checked against it at load, as the reference checks its ABI
(src/exact_ballpark/vec_modes_cuda.py:248-269).

The HIP library must run on the HIP runtime torch has loaded: torch is imported first, so the loader resolves the
library's libamdhip64.so.7 to torch's copy (both carry that soname) and launches go on torch's streams. A missing
library is an error, never a fallback.
"""

from __future__ import annotations

import ctypes
from pathlib import Path

import torch  # noqa: F401  (first: its HIP runtime is the one the library binds)

LIB_DIR = Path(__file__).resolve().parent.parent / "lib"
ABI_VERSION = 9
NF, NI, NB, NS = 46, 10, 7, 7
NCF, NCI = 3, 8                 # the collision blocks: mLastC; level, key (3), side (3), the list's length
NY = 9                          # the display orientation block: yaw, pitch, roll, their old values, roll speed (2), yaw delta
NSF, NSI = 93, 5                # the static collidables' table: geometry; kind, id, parent, level, box count
NOALLOC_MAX_ROWS = 1 << 18      # the most rows (stride) xb_modes_tick_noalloc takes
ERR_TOO_MANY_ROWS = -2          # its refusal of a larger park


class EpisodeRows(ctypes.Structure):
    """include/xb.h's xb_episode_rows: the serving loop's per-row episode arrays."""
    _fields_ = [("seed", ctypes.c_void_p), ("ship", ctypes.c_void_p), ("size", ctypes.c_void_p),
                ("base", ctypes.c_void_p), ("admit", ctypes.c_void_p), ("next_cmd", ctypes.c_void_p)]


class ModesState(ctypes.Structure):
    """include/xb.h's xb_modes_state."""
    _fields_ = [("f64", ctypes.c_void_p), ("i64", ctypes.c_void_p), ("u8", ctypes.c_void_p),
                ("scratch_f64", ctypes.c_void_p), ("scratch_u8", ctypes.c_void_p), ("tiles", ctypes.c_void_p),
                ("rowmap", ctypes.c_void_p), ("n", ctypes.c_int64), ("stride", ctypes.c_int64),
                ("n_tiles", ctypes.c_int32), ("new_orbit", ctypes.c_int32), ("friction", ctypes.c_double),
                ("dt", ctypes.c_double), ("ffriction", ctypes.c_float), ("fdt", ctypes.c_float),
                ("dynamical", ctypes.c_int32), ("ticks", ctypes.c_int32),
                ("schedule", ctypes.c_int32), ("n_keys", ctypes.c_int32), ("keys", ctypes.c_void_p),
                ("path", ctypes.c_void_p), ("iters", ctypes.c_void_p), ("active", ctypes.c_void_p),
                ("coll_f64", ctypes.c_void_p), ("coll_i64", ctypes.c_void_p),
                ("hits", ctypes.c_void_p), ("hits_cap", ctypes.c_int32), ("collisions", ctypes.c_int32),
                ("ypr", ctypes.c_void_p), ("st_f64", ctypes.c_void_p), ("st_i64", ctypes.c_void_p),
                ("st_keys", ctypes.c_void_p), ("st_range", ctypes.c_void_p), ("st_stride", ctypes.c_int64),
                ("st_boxes", ctypes.c_int32), ("statics", ctypes.c_int32), ("form", ctypes.c_void_p),
                ("form_len", ctypes.c_void_p), ("n_form", ctypes.c_int32), ("form_stride", ctypes.c_int32),
                ("fld", ctypes.c_void_p), ("form_slots", ctypes.c_void_p)]


_LIBS: dict = {}


class NativeError(RuntimeError):
    pass


def _load(name: str) -> ctypes.CDLL:
    path = LIB_DIR / name
    if not path.exists():
        raise NativeError(f"{path} is not built: run __graft_entry__.build() (or exact_ballpark_amd.build) first")
    return ctypes.CDLL(str(path))


def hip() -> ctypes.CDLL:
    """libxb_hip.so, loaded once and checked (ABI version, state layout)."""
    if "hip" not in _LIBS:
        L = _load("libxb_hip.so")
        L.xb_abi.restype = ctypes.c_int
        L.xb_error_string.restype = ctypes.c_char_p
        L.xb_error_string.argtypes = [ctypes.c_int]
        L.xb_layout.argtypes = [ctypes.POINTER(ctypes.c_int64)]
        L.xb_modes_tick.argtypes = [ctypes.POINTER(ModesState), ctypes.c_void_p, ctypes.c_int]
        L.xb_modes_tick_noalloc.argtypes = [ctypes.POINTER(ModesState), ctypes.c_void_p, ctypes.c_int]
        L.xb_maths_eval.argtypes = [ctypes.c_int, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                    ctypes.c_int]
        L.xb_modes_info.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_int64)]
        L.xb_rand_seed.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p]
        L.xb_rand_draw.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32, ctypes.c_void_p, ctypes.c_void_p]
        L.xb_collide_file.argtypes = [ctypes.POINTER(ModesState), ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p]
        L.xb_modes_ypr_snap.argtypes = [ctypes.POINTER(ModesState), ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p]
        L.xb_modes_orient_setup.argtypes = [ctypes.POINTER(ModesState), ctypes.c_void_p, ctypes.c_int64, ctypes.c_int,
                                            ctypes.c_void_p]
        L.xb_modes_commands.argtypes = [ctypes.POINTER(ModesState), ctypes.c_void_p, ctypes.c_int64] + \
            [ctypes.c_void_p] * 7
        L.xb_modes_command_check.argtypes = [ctypes.POINTER(ModesState), ctypes.c_void_p, ctypes.c_int64,
                                             ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        L.xb_modes_command_apply.argtypes = [ctypes.POINTER(ModesState), ctypes.c_void_p, ctypes.c_int64,
                                             ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        _check_layout(L, "libxb_hip.so")
        _LIBS["hip"] = L
    return _LIBS["hip"]


def cpu() -> ctypes.CDLL:
    """libxb_cpu.so, the host build of the tick."""
    if "cpu" not in _LIBS:
        L = _load("libxb_cpu.so")
        if L.xb_cpu_abi() != ABI_VERSION:
            raise NativeError("libxb_cpu.so: ABI version mismatch")
        L.xb_cpu_modes_tick.argtypes = [ctypes.POINTER(ModesState), ctypes.c_int]
        L.xb_cpu_maths_eval.argtypes = [ctypes.c_int, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p]
        L.xb_cpu_rand_seed.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p]
        L.xb_cpu_rand_draw.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32, ctypes.c_void_p]
        L.xb_cpu_collide_file.argtypes = [ctypes.POINTER(ModesState), ctypes.c_void_p, ctypes.c_int64]
        L.xb_cpu_modes_ypr_snap.argtypes = [ctypes.POINTER(ModesState), ctypes.c_void_p, ctypes.c_int64]
        L.xb_cpu_statics_setup.argtypes = [ctypes.c_int64] + [ctypes.c_void_p] * 7 + [ctypes.c_int64, ctypes.c_int32]
        L.xb_cpu_modes_orient_setup.argtypes = [ctypes.POINTER(ModesState), ctypes.c_void_p, ctypes.c_int64,
                                                ctypes.c_int]
        L.xb_cpu_modes_commands.argtypes = [ctypes.POINTER(ModesState), ctypes.c_void_p, ctypes.c_int64] + \
            [ctypes.c_void_p] * 6
        L.xb_cpu_modes_command_check.argtypes = [ctypes.POINTER(ModesState), ctypes.c_void_p, ctypes.c_int64,
                                                 ctypes.c_void_p, ctypes.c_void_p]
        L.xb_cpu_modes_command_apply.argtypes = [ctypes.POINTER(ModesState), ctypes.c_void_p, ctypes.c_int64,
                                                 ctypes.c_void_p, ctypes.c_void_p]
        L.xb_cpu_pack_tiles.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32, ctypes.c_void_p,
                                        ctypes.POINTER(ctypes.c_int64)]
        _LIBS["cpu"] = L
    return _LIBS["cpu"]


def _check_layout(L: ctypes.CDLL, name: str) -> None:
    if L.xb_abi() != ABI_VERSION:
        raise NativeError(f"{name}: ABI version {L.xb_abi()}, expected {ABI_VERSION}")
    out = (ctypes.c_int64 * 5)()
    L.xb_layout(out)
    want = (NF, NI, NB, NS, ctypes.sizeof(ModesState))
    if tuple(out) != want:
        raise NativeError(f"{name}: layout {tuple(out)} differs from the loader's {want}")


def check(rc: int, fn: str) -> None:
    if rc:
        msg = hip().xb_error_string(rc).decode() if "hip" in _LIBS else str(rc)
        raise RuntimeError(f"{fn} failed ({rc}: {msg})")


MATHS = {"sin": 0, "cos": 1, "log": 2, "exp": 3, "sincos_sin": 4, "sincos_cos": 5, "sqrt": 6, "div": 7, "expf": 8,
         "logf": 9, "sinf": 10, "cosf": 11, "acosf": 12, "atanf": 13, "atan2f": 14, "pow": 15, "asin": 16,
         "atan2": 17, "powf": 18}
BINARY = ("div", "atan2f", "pow", "atan2", "powf")


def maths(name: str, x: torch.Tensor, y: torch.Tensor | None = None, block: int = 256) -> torch.Tensor:
    """The tick's own maths function `name` elementwise (on the device of x: the GPU kernel or the host build): the
    double sin, cos, log, exp, sincos's two parts, IEEE sqrt and division x / y; the float expf, logf, sinf, cosf,
    acosf, atanf, atan2f(x, y) (x the ordinate) and powf(x, y) on floats held in doubles; the double pow(x, y), asin
    and atan2(x, y) (x the ordinate)."""
    x = x.to(torch.float64).contiguous()
    n = x.numel()
    inp = torch.cat((x.reshape(-1), y.to(torch.float64).reshape(-1))) if name in BINARY else x.reshape(-1)
    out = torch.empty(n, dtype=torch.float64, device=x.device)
    if x.device.type == "cuda":
        with torch.cuda.device(x.device):
            rc = hip().xb_maths_eval(MATHS[name], n, inp.data_ptr(), out.data_ptr(),
                                     torch.cuda.current_stream(x.device).cuda_stream, block)
        check(rc, "xb_maths_eval")
    else:
        rc = cpu().xb_cpu_maths_eval(MATHS[name], n, inp.data_ptr(), out.data_ptr())
        if rc:
            raise RuntimeError(f"xb_cpu_maths_eval failed ({rc})")
    return out.reshape(x.shape)


def kernel_info(which: int = 0) -> dict:
    out = (ctypes.c_int64 * 4)()
    check(hip().xb_modes_info(which, out), "xb_modes_info")
    return {"registers": out[0], "private_bytes": out[1], "lds_bytes": out[2], "max_threads": out[3]}
