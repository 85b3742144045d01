"""Deterministic sequence generators used by the reference's own tests.

body(n)      crates/map/src/fragment.rs:141-151 (xorshift64, seed 0xe911da37, ACGT)
Rng          crates/map/src/extend.rs:2481-2498 (xorshift64 <<13 >>7 <<17 over ACGTN)
rand_window  crates/map/src/extend.rs:2509-2525 (read embedded with ~1/8 subs, 1/16 drops)
revcomp      crates/map/src/align.rs:2518-2533 (N stays N; IUPAC -> N)
"""
from __future__ import annotations

M64 = (1 << 64) - 1


def xorshift(x: int) -> int:
    x ^= (x << 13) & M64
    x ^= x >> 7
    x ^= (x << 17) & M64
    return x


def body(n: int) -> bytes:
    s = 0xE911DA37
    out = bytearray()
    for _ in range(n):
        s = xorshift(s)
        out.append(b"ACGT"[s & 3])
    return bytes(out)


class Rng:
    def __init__(self, seed: int):
        self.s = seed & M64

    def next(self) -> int:
        self.s = xorshift(self.s)
        return self.s

    def below(self, n: int) -> int:
        return self.next() % n

    def base(self) -> int:
        return b"ACGTN"[self.below(5)]


def rand_seq(rng: Rng, n: int) -> bytes:
    return bytes(rng.base() for _ in range(n))


def rand_window(rng: Rng, read: bytes, n: int) -> bytes:
    w = bytearray(rand_seq(rng, n))
    if n > len(read):
        off = rng.below(n - len(read) + 1)
        k = 0
        while k < len(read) and off + k < n:
            b = read[k]
            w[off + k] = rng.base() if rng.below(8) == 0 else b
            if rng.below(16) == 0:
                k += 2
            else:
                k += 1
    return bytes(w)


_COMP = {ord("A"): ord("T"), ord("a"): ord("T"), ord("C"): ord("G"), ord("c"): ord("G"),
         ord("G"): ord("C"), ord("g"): ord("C"), ord("T"): ord("A"), ord("t"): ord("A")}


def revcomp(seq: bytes) -> bytes:
    return bytes(_COMP.get(b, ord("N")) for b in reversed(seq))
