"""Builds the native libraries in-tree: libxb_hip.so (the gfx950 kernels and the C ABI, hipcc) and libxb_cpu.so (the
same tick source for the host, g++), into exact-ballpark_amd/lib/, and checks the exactness of the GPU build.

The checks are the build's part of the exactness contract (the reference's own are its source scan and PTX check,
src/exact_ballpark/vec_orient_cuda.py:189-227): no OCML call or device math builtin in the sources; in our own LLVM IR
no floating-point instruction carries a fast-math flag (contract, afn, reassoc, nnan, ninf, nsz, arcp) and every
llvm.fma comes from an explicit call; in the gfx950 code object denormals kept for float and double and IEEE mode on,
and no hardware transcendental (v_exp, v_log, v_sin, v_cos, v_rcp_f64 outside the division and square-root sequences is
not checked: the compiler's IEEE division and sqrt expansions use v_rcp/v_rsq with correction steps).
"""

from __future__ import annotations

import os
import re
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = ROOT / "csrc"
LIB = ROOT / "lib"
ARCH = "gfx950"

HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-fno-gpu-flush-denormals-to-zero",
             "-fhip-fp32-correctly-rounded-divide-sqrt", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wall", "-Wno-unused-function", "-Wno-unused-variable"]
CPU_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-march=x86-64-v3", "-fopenmp",
             "-Wall", "-Wno-unused-function"]
HIP_SOURCES = ["xb_modes.hip", "xb_sched.hip", "xb_serve.hip", "xb_rand.hip", "xb_space.hip", "xb_server.hip"]
BANNED_SOURCE = re.compile(r"__ocml_|__builtin_amdgcn_(rcp|rsq|sqrt|exp|log|sin|cos)|\b__f?(exp|log|sin|cos)f?\(")
FAST_FLAGS = re.compile(r"\b(fadd|fsub|fmul|fdiv|frem|fneg|fcmp|call)\b[^\n]*\b(fast|contract|afn|reassoc|nnan|ninf|nsz|arcp)\b")


def _run(cmd: list, **kw) -> subprocess.CompletedProcess:
    p = subprocess.run(cmd, capture_output=True, text=True, **kw)
    if p.returncode != 0:
        raise RuntimeError(f"build step failed: {' '.join(map(str, cmd))}\n{p.stdout}\n{p.stderr}")
    return p


def scan_sources() -> None:
    for f in sorted(CSRC.glob("*.h")) + sorted(CSRC.glob("*.hip")) + sorted(CSRC.glob("*.cpp")):
        text = f.read_text()
        m = BANNED_SOURCE.search(text)
        if m:
            raise RuntimeError(f"{f.name}: calls {m.group(0)!r}, outside the exactness contract")


def check_ir(src: Path) -> dict:
    """Our device code's LLVM IR: no fast-math flags on any floating-point operation; the count of llvm.fma calls."""
    out = LIB / (src.stem + ".gfx950.ll")
    _run(["hipcc", f"--offload-arch={ARCH}", *HIP_FLAGS, "--cuda-device-only", "-S", "-emit-llvm", str(src), "-o",
          str(out)], cwd=CSRC)
    ir = out.read_text()
    bad = FAST_FLAGS.search(ir)
    if bad:
        raise RuntimeError(f"{src.name}: the device IR carries a fast-math flag: {bad.group(0)[:160]}")
    if "@__ocml_" in ir:
        raise RuntimeError(f"{src.name}: the device IR calls OCML")
    return {"llvm_fma": len(re.findall(r"@llvm\.fma\.f64", ir))}


def check_isa(src: Path) -> dict:
    out = LIB / (src.stem + ".gfx950.s")
    _run(["hipcc", f"--offload-arch={ARCH}", *HIP_FLAGS, "--cuda-device-only", "-S", str(src), "-o", str(out)],
         cwd=CSRC)
    asm = out.read_text()
    for key in (".amdhsa_float_denorm_mode_32 3", ".amdhsa_float_denorm_mode_16_64 3", ".amdhsa_ieee_mode 1"):
        n_kernels = asm.count(".amdhsa_kernel ")
        if asm.count(key) != n_kernels:
            raise RuntimeError(f"{src.name}: not every kernel has {key!r}")
    hw = re.findall(r"^\s*(v_(exp|log|sin|cos)_f(16|32|64))\b", asm, flags=re.M)
    if hw:
        raise RuntimeError(f"{src.name}: hardware transcendental instructions {sorted(set(h[0] for h in hw))}")
    return {"kernels": asm.count(".amdhsa_kernel ")}


def _jobs() -> int:
    return max(1, min(8, int(os.environ.get("MAX_JOBS", os.cpu_count() or 1))))


def build_hip(check: bool = True) -> Path:
    """Each source compiled to its own object (host and gfx950 code) and the checks' IR and ISA, all in parallel, then
    one link."""
    from concurrent.futures import ThreadPoolExecutor
    LIB.mkdir(exist_ok=True)
    out = LIB / "libxb_hip.so"
    objs = [LIB / (Path(s).stem + ".o") for s in HIP_SOURCES]
    steps = [lambda s=s, o=o: _run(["hipcc", f"--offload-arch={ARCH}", *HIP_FLAGS, "-c", str(CSRC / s), "-o", str(o)],
                                   cwd=CSRC) for s, o in zip(HIP_SOURCES, objs)]
    if check:
        steps += [lambda s=s: check_ir(CSRC / s) for s in HIP_SOURCES]
        steps += [lambda s=s: check_isa(CSRC / s) for s in HIP_SOURCES]
    with ThreadPoolExecutor(_jobs()) as ex:
        for f in [ex.submit(st) for st in steps]:
            f.result()
    _run(["hipcc", f"--offload-arch={ARCH}", "-shared", *map(str, objs), "-o", str(out)], cwd=CSRC)
    for o in objs:
        o.unlink()
    return out


def build_cpu() -> Path:
    LIB.mkdir(exist_ok=True)
    out = LIB / "libxb_cpu.so"
    _run(["g++", *CPU_FLAGS, "-shared", str(CSRC / "xb_cpu.cpp"), "-o", str(out)], cwd=CSRC)
    return out


def build(check: bool = True) -> None:
    scan_sources()
    build_hip(check)
    build_cpu()


if __name__ == "__main__":
    build(check="--no-check" not in sys.argv)
    print("built", *(p.name for p in sorted(LIB.glob("*.so"))), file=sys.stderr)
