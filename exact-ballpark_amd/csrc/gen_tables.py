#!/usr/bin/env python3
# SPDX-License-Identifier: LGPL-2.1-or-later
"""Writes `xb_glibc_tables.h`: the constants and tables of the GNU C Library 2.34 (aarch64) maths that the destiny
tick calls, as C literals (hex floats, exact), for the HIP kernels and for the host builds of the same maths.

The values are data, not code: the reference keeps them as hex literals in its maths transcriptions
(src/exact_ballpark/dlibm.py:47-266, vec_exp.py:46-57, vec_libm.py:45-280, dlibm_powf.py:48-101,
dlibm_inv.py:49-1590). This script reads those files as text with `ast` (nothing is imported or executed) and
evaluates only literal expressions: numbers, strings, `_H(...)`/`float.fromhex(...)` of a literal, tuples, unary minus
and arithmetic between values already read. The exp table (`__exp_data.tab`) is not read but generated from its
definition, scale_i = 2^(i/128) and tail_i = 2^(i/128)/scale_i - 1, each rounded once from 80 digits, and checked
against the SHA-256 of its 256 little-endian words that the reference records (vec_exp.py:52).

Run in a checkout that has the reference at /root/reference; the header it writes is committed, so builds never
need the reference.
"""

from __future__ import annotations

import ast
import hashlib
import struct
import sys
from decimal import Context
from fractions import Fraction
from pathlib import Path

REF = Path("/root/reference/src/exact_ballpark")
OUT = Path(__file__).resolve().with_name("xb_glibc_tables.h")
MODULES = (("DL", "dlibm.py"), ("EX", "vec_exp.py"), ("LF", "vec_libm.py"), ("PF", "dlibm_powf.py"),
           ("IV", "dlibm_inv.py"))
EXP_TABLE_SHA256 = "a4014caf214c2bab90ea0fdf3f57941e0235ae04dc2282a6c2f72e7a480c87e4"


class _Skip(Exception):
    pass


def _eval(node: ast.AST, env: dict):
    """The value of a literal expression, or _Skip."""
    if isinstance(node, ast.Constant) and isinstance(node.value, (int, float, str)) and not isinstance(node.value, bool):
        return node.value
    if isinstance(node, ast.Tuple):
        return tuple(_eval(e, env) for e in node.elts)
    if isinstance(node, ast.Name):
        if node.id in env:
            return env[node.id]
        raise _Skip
    if isinstance(node, ast.UnaryOp) and isinstance(node.op, ast.USub):
        return -_eval(node.operand, env)
    if isinstance(node, ast.BinOp):
        a, b = _eval(node.left, env), _eval(node.right, env)
        ops = {ast.Add: lambda: a + b, ast.Sub: lambda: a - b, ast.Mult: lambda: a * b, ast.Div: lambda: a / b,
               ast.Pow: lambda: a ** b, ast.LShift: lambda: a << b, ast.BitOr: lambda: a | b,
               ast.BitAnd: lambda: a & b, ast.FloorDiv: lambda: a // b}
        f = ops.get(type(node.op))
        if f is None:
            raise _Skip
        return f()
    if isinstance(node, ast.Call) and len(node.args) == 1 and not node.keywords:
        fn = node.func
        is_h = isinstance(fn, ast.Name) and fn.id == "_H"
        is_fromhex = (isinstance(fn, ast.Attribute) and fn.attr == "fromhex" and isinstance(fn.value, ast.Name)
                      and fn.value.id == "float")
        if is_h or is_fromhex:
            s = _eval(node.args[0], env)
            if not isinstance(s, str):
                raise _Skip
            return float.fromhex(s)
        if isinstance(fn, ast.Name) and fn.id == "_fh":       # a float's bit pattern as a Python float
            u = _eval(node.args[0], env)
            return struct.unpack("<f", struct.pack("<I", u))[0]
    raise _Skip


def read_constants(path: Path) -> dict:
    tree = ast.parse(path.read_text(), filename=str(path))
    env: dict = {}
    for stmt in tree.body:
        if not isinstance(stmt, ast.Assign):
            continue
        targets = stmt.targets[0]
        try:
            if isinstance(targets, ast.Name):
                env[targets.id] = _eval(stmt.value, env)
            elif isinstance(targets, ast.Tuple) and isinstance(stmt.value, ast.Tuple):
                vals = [_eval(v, env) for v in stmt.value.elts]
                for t, v in zip(targets.elts, vals):
                    if isinstance(t, ast.Name):
                        env[t.id] = v
        except (_Skip, ZeroDivisionError, OverflowError, TypeError):
            continue
    return env


def exp_table() -> list:
    """`__exp_data.tab` from its definition: (bits of tail_i, bits of scale_i - i 2^45) for i < 128."""
    ctx = Context(prec=80)
    ln2 = ctx.ln(2)
    words = []
    for i in range(128):
        v = Fraction(ctx.exp(ctx.divide(ctx.multiply(ln2, i), 128)))
        scale = float(v)                                  # rounded to nearest, ties to even
        tail = float(v / Fraction(scale) - 1)
        u_tail = struct.unpack("<Q", struct.pack("<d", tail))[0]
        u_scale = struct.unpack("<Q", struct.pack("<d", scale))[0]
        words += [u_tail, (u_scale - (i << 45)) & ((1 << 64) - 1)]
    if hashlib.sha256(struct.pack("<256Q", *words)).hexdigest() != EXP_TABLE_SHA256:
        raise SystemExit("the generated exp table does not match the recorded SHA-256")
    return words


def _flat(v):
    if isinstance(v, tuple):
        out = []
        for e in v:
            out.extend(_flat(e))
        return out
    return [v]


def _dlit(x: float) -> str:
    if x != x or x in (float("inf"), float("-inf")):
        u = struct.unpack("<Q", struct.pack("<d", x))[0]
        return f"XB_BITS_DOUBLE(0x{u:016x}ULL)"
    return x.hex()


def _ilit(i: int) -> str:
    if i < 0:
        return f"((int64_t)0x{i & ((1 << 64) - 1):016x}ULL)"
    return f"0x{i:x}ULL" if i >= 1 << 31 else str(i)


def main() -> None:
    lines = ["/* SPDX-License-Identifier: LGPL-2.1-or-later */",
             "/* The GNU C Library 2.34 (aarch64) maths constants and tables used by the destiny tick: data, written by",
             "   gen_tables.py from the literals the reference keeps (src/exact_ballpark/dlibm.py, vec_exp.py, vec_libm.py,",
             "   dlibm_powf.py, dlibm_inv.py); the exp table generated from its definition and checked by SHA-256. Do not",
             "   edit. Scalars are macros; tables take XB_TABLE (the including file defines it: __constant__ on the GPU). */",
             "#ifndef XB_GLIBC_TABLES_H", "#define XB_GLIBC_TABLES_H", "#include <stdint.h>", "#ifndef XB_TABLE",
             "#define XB_TABLE", "#endif", ""]
    for prefix, name in MODULES:
        env = read_constants(REF / name)
        lines.append(f"/* ---- from {name} ---- */")
        for key, v in env.items():
            if key.startswith("_") or isinstance(v, str) or key in ("LIBM_SHA256", "M64", "TABLE_SHA256"):
                continue
            cname = f"XB_{prefix}_{key}"
            if isinstance(v, tuple):
                flat = _flat(v)
                if all(isinstance(e, float) for e in flat):
                    body = ", ".join(_dlit(e) for e in flat)
                    lines.append(f"XB_TABLE static const double {cname}[{len(flat)}] = {{{body}}};")
                elif all(isinstance(e, int) for e in flat):
                    body = ", ".join(f"0x{e & ((1 << 64) - 1):x}ULL" for e in flat)
                    lines.append(f"XB_TABLE static const uint64_t {cname}[{len(flat)}] = {{{body}}};")
                else:
                    continue
                if isinstance(v[0], tuple):
                    lines.append(f"#define {cname}_COLS {len(v[0])}")
            elif isinstance(v, float):
                lines.append(f"#define {cname} {_dlit(v)}")
            elif isinstance(v, int):
                lines.append(f"#define {cname} {_ilit(v)}")
        lines.append("")
    words = exp_table()
    body = ", ".join(f"0x{w:016x}ULL" for w in words)
    lines += ["/* ---- __exp_data.tab, generated: (tail_i bits, scale_i bits - i 2^45), i < 128 ---- */",
              f"XB_TABLE static const uint64_t XB_EXP_TAB[256] = {{{body}}};", "", "#endif", ""]
    OUT.write_text("\n".join(lines))
    print(f"wrote {OUT} ({len(lines)} lines)", file=sys.stderr)


if __name__ == "__main__":
    main()
