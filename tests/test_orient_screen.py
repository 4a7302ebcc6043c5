"""The orientation searches' screen (exact-ballpark_amd/csrc/xb_orient.h, XO_SCREEN) changes no result: the searches
with it and without it, bit for bit, on a million seeded states over and beyond the tick's ranges
(tests/orient_screen_check.cpp, built with the product's flags for the host)."""

import json
import subprocess
from pathlib import Path

HERE = Path(__file__).resolve().parent
FLAGS = ["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-fno-fast-math", "-march=x86-64-v3"]


def test_screened_searches_equal_plain(tmp_path):
    src = HERE / "orient_screen_check.cpp"
    a, b, exe = tmp_path / "a.o", tmp_path / "b.o", tmp_path / "check"
    subprocess.run(FLAGS + ["-c", "-DSIDE=screened", "-DMAIN", str(src), "-o", str(a)], check=True)
    subprocess.run(FLAGS + ["-c", "-DSIDE=plain", "-DXO_SCREEN=0", str(src), "-o", str(b)], check=True)
    subprocess.run(FLAGS + [str(a), str(b), "-o", str(exe)], check=True)
    r = json.loads(subprocess.run([str(exe), "1000000"], check=True, capture_output=True, text=True).stdout)
    assert r["differ"] == 0, r
    assert r["torque_and_roll_nonzero"] > 100000, r
