"""Pins of the glibc maths the tick calls (exact-ballpark_amd/csrc/xb_maths_double.h, used by the kernels, their host
build and the oracle), none of them a restatement of the formula:

- the exp table: the SHA-256 of its 256 words that the reference records (src/exact_ballpark/vec_exp.py:52);
- sin, cos, log and exp against this machine's C library, bit for bit: glibc 2.35's x86-64 FMA builds of the same
  IBM and Arm routines (the reference's target is glibc 2.34 aarch64's; on these samples the two agree);
- sincos: the reference's known answer at X0 = 0x1.ad1fb54442d18p+0, where glibc 2.34's sincos and sin differ in the
  last bit (docs/numerics.md:121-127 in the reference), and sincos equal to (sin, cos) on every other sample;
- special values (zeros, infinities, NaN, subnormals, the branch boundaries);
- the host build of the kernel source (libxb_cpu.so) equal to the oracle's build, and its sqrt and division IEEE;
- the float functions, powf and the double asin and atan2 swept against the C library (tests/maths_float_sweep.cpp):
  powf(x, 0.25f), the mushroom growth's call, agrees on all 2^32 bit patterns of x.
"""

import ctypes
import ctypes.util
import hashlib
import math
import re
import struct
from pathlib import Path

import numpy as np
import pytest
import torch

import oracle as O
from exact_ballpark_amd import _native as N

TABLES = Path(__file__).resolve().parent.parent / "exact-ballpark_amd" / "csrc" / "xb_glibc_tables.h"
X0 = float.fromhex("0x1.ad1fb54442d18p+0")


def _bits(a):
    return np.asarray(a, dtype=np.float64).view(np.uint64)


def _samples(seed: int, n: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    return np.concatenate([rng.uniform(-10, 10, n), rng.uniform(-1e3, 1e3, n), rng.uniform(0.85, 2.5, n),
                           np.exp(rng.uniform(-700, 700, n)) * rng.choice([-1, 1], n), rng.uniform(-1e8, 1e8, n),
                           rng.uniform(1e8, 1e300, n // 4)])


def test_exp_table_sha256():
    text = TABLES.read_text()
    body = re.search(r"XB_EXP_TAB\[256\] = \{([^}]*)\}", text).group(1)
    words = [int(w.strip().rstrip("ULL"), 16) for w in body.split(",")]
    assert len(words) == 256
    digest = hashlib.sha256(struct.pack("<256Q", *words)).hexdigest()
    assert digest == "a4014caf214c2bab90ea0fdf3f57941e0235ae04dc2282a6c2f72e7a480c87e4"


@pytest.mark.parametrize("name,ref", [("sin", math.sin), ("cos", math.cos)])
def test_sin_cos_match_c_library(name, ref):
    x = _samples(11, 20000)
    got = O.maths(name, x)
    want = np.array([ref(v) for v in x])
    assert (_bits(got) == _bits(want)).all()


def test_log_matches_c_library():
    x = np.abs(_samples(12, 20000))
    x = np.concatenate([x[x > 0], [5e-324, 2.2e-308, 1e-310, 0.9375, 1.0625, 1.0 + 2 ** -52, 1.0 - 2 ** -53]])
    got = O.maths("log", x)
    want = np.array([math.log(v) for v in x])
    assert (_bits(got) == _bits(want)).all()


def test_exp_matches_c_library():
    rng = np.random.default_rng(13)
    x = np.concatenate([rng.uniform(-745, 709, 50000), rng.uniform(-20, 0, 50000), rng.uniform(-1e-3, 1e-3, 20000),
                        [-745.1, -708.5, 709.7, 512.0, -512.0, 2 ** -55, -(2 ** -55)]])
    got = O.maths("exp", x)
    want = np.array([math.exp(v) for v in x])
    assert (_bits(got) == _bits(want)).all()


def test_sincos_known_answer_and_agreement():
    s = O.maths("sincos_sin", np.array([X0, -X0]))
    c = O.maths("sincos_cos", np.array([X0, -X0]))
    assert s[0].hex() == "0x1.fd27acee550dbp-1" and s[1].hex() == "-0x1.fd27acee550dbp-1"
    assert O.maths("sin", np.array([X0]))[0].hex() == "0x1.fd27acee550dcp-1"
    assert (_bits(c) == _bits(O.maths("cos", np.array([X0, -X0])))).all()
    x = _samples(14, 20000)
    assert (_bits(O.maths("sincos_sin", x)) == _bits(O.maths("sin", x))).all()
    assert (_bits(O.maths("sincos_cos", x)) == _bits(O.maths("cos", x))).all()


def test_special_values():
    inf, nan = math.inf, math.nan
    assert O.maths("exp", np.array([0.0, -inf, inf]))[:3].tolist() == [1.0, 0.0, inf]
    assert math.isnan(O.maths("exp", np.array([nan]))[0])
    lg = O.maths("log", np.array([1.0, 0.0, -0.0, inf, -1.0, nan]))
    assert lg[0] == 0.0 and lg[1] == -inf and lg[2] == -inf and lg[3] == inf and math.isnan(lg[4]) and math.isnan(lg[5])
    sn = O.maths("sin", np.array([0.0, -0.0, inf, nan, 1e-300]))
    assert _bits(sn[:2]).tolist() == _bits([0.0, -0.0]).tolist() and math.isnan(sn[2]) and math.isnan(sn[3])
    assert sn[4] == 1e-300
    assert O.maths("cos", np.array([0.0]))[0] == 1.0


@pytest.mark.parametrize("name", ["sin", "cos", "log", "exp", "sincos_sin", "sincos_cos"])
def test_host_build_equals_oracle_build(name):
    x = _samples(15, 5000)
    if name == "log":
        x = np.abs(x)
    if name == "exp":
        x = np.clip(x, -800, 800)
    got = N.maths(name, torch.from_numpy(x)).numpy()
    assert (_bits(got) == _bits(O.maths(name, x))).all()


def test_host_build_sqrt_and_division_are_ieee():
    rng = np.random.default_rng(16)
    x = np.concatenate([np.exp(rng.uniform(-700, 700, 20000)), [4.0, 2.0, 1e-320, 0.0, np.inf]])
    got = N.maths("sqrt", torch.from_numpy(x)).numpy()
    assert (_bits(got) == _bits(np.sqrt(x))).all()
    y = np.exp(rng.uniform(-300, 300, x.size))
    q = N.maths("div", torch.from_numpy(x), torch.from_numpy(y)).numpy()
    assert (_bits(q) == _bits(x / y)).all()


def _sweep(tmp_path, unfused: bool, stride: int) -> dict:
    import json
    import subprocess
    src = Path(__file__).resolve().parent / "maths_float_sweep.cpp"
    exe = tmp_path / ("sweep_unfused" if unfused else "sweep")
    cmd = ["g++", "-O2", "-fopenmp", "-ffp-contract=off", "-march=x86-64-v3", str(src), "-o", str(exe), "-lm"]
    if unfused:
        cmd.insert(1, "-DXB_FDLIBM_FMAF(a,b,c)=((a)*(b)+(c))")
    subprocess.run(cmd, check=True)
    out = subprocess.run([str(exe), str(stride)], check=True, capture_output=True, text=True).stdout
    return json.loads(out)


def test_float_maths_sweep_against_c_library(tmp_path):
    """Every 16th float bit pattern. Exhaustive, with stride 1: logf, sinf, cosf and powf(x, 0.25f) 0 differences;
    expf 2, at x = 0x1.04845ep+5 and -0x1.f8cbb2p+5, where this x86-64 library's last bit is one off the correctly
    rounded value (float(exp(double x))), which the aarch64 algorithm gives."""
    r = _sweep(tmp_path, unfused=False, stride=16)
    assert r["expf"] == 0 and r["logf"] == 0 and r["sinf"] == 0 and r["cosf"] == 0 and r["pow"] == 0
    assert r["powf_quarter"] == 0 and r["powf"] == 0 and r["asin"] == 0 and r["atan2"] == 0
    u = _sweep(tmp_path, unfused=True, stride=16)
    assert u["acosf"] == 0 and u["atanf"] == 0 and u["atan2f"] == 0
    assert r["acosf"] > 0 and r["atanf"] > 0      # the fused steps do change bits: the two builds really differ
