"""THE ORACLE for glibc's srand() and rand() per world -- test infrastructure only (tests/). A pure-Python restatement
of glibc's TYPE_3 generator (stdlib/random_r.c), as the reference restates it (src/exact_ballpark/glibc_rand.py:14-45):
the 31-word table seeded by 16807 x mod 2^31 - 1 (Schrage), 310 draws discarded, each draw the sum of the front and
rear words modulo 2^32 stored at the front, shifted right by one.

What pins it: tests/test_rand.py holds it to this machine's own C library (srand and rand through ctypes, the same
integer algorithm in every glibc), draw for draw, on edge and random seeds.
"""

from __future__ import annotations


def _cdiv(a: int, b: int) -> int:
    q = abs(a) // abs(b)
    return q if (a >= 0) == (b >= 0) else -q


class Rand:
    def __init__(self, seed: int):
        seed &= 0xFFFFFFFF
        if seed == 0:
            seed = 1
        word = seed - (1 << 32) if seed >= (1 << 31) else seed
        self.table = [word]
        for _ in range(30):
            hi = _cdiv(word, 127773)
            lo = word - 127773 * hi
            word = 16807 * lo - 2836 * hi
            if word < 0:
                word += 2147483647
            self.table.append(word)
        self.front, self.rear = 3, 0
        for _ in range(310):
            self.next()

    def next(self) -> int:
        val = (self.table[self.front] + self.table[self.rear]) & 0xFFFFFFFF
        self.table[self.front] = val - (1 << 32) if val >= (1 << 31) else val
        self.front += 1
        if self.front >= 31:
            self.front = 0
            self.rear += 1
        else:
            self.rear += 1
            if self.rear >= 31:
                self.rear = 0
        return val >> 1

    def state(self) -> list:
        return list(self.table) + [self.front, self.rear]
