"""Pins the oracle's restatement of the engine's EA-Myers
(crates/engine/src/myers.rs: infix / long::infix and global / long::global),
for which the reference holds no test of its own:
- against the O(m n) edit-distance recurrence it computes (EA rules of
  myers.rs:1-7: identity and pattern G / text A, pattern C / text T cost 0,
  N matches nothing, case folded, indels cost 1; infix frees the text
  overhang on both sides, global forces both ends), at pattern lengths on
  both sides of the 128-bit block boundary (1..256);
- the infix distance equals the map's EA-Myers distance (myers_ea.rs, same
  recurrence, edlib-form 64-bit words), an independent implementation;
- patterns of length 0 or above 256 are refused (the reference asserts).
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle


@pytest.fixture(scope="module", autouse=True)
def _lib():
    oracle.lib()


def _code(b: int) -> int:
    return {65: 0, 97: 0, 67: 1, 99: 1, 71: 2, 103: 2, 84: 3, 116: 3}.get(b, 4)


def _cost(p: int, t: int) -> int:
    cp, ct = _code(p), _code(t)
    if ct > 3 or cp > 3:
        return 1
    if cp == ct or (ct == 0 and cp == 2) or (ct == 3 and cp == 1):
        return 0
    return 1


def ea_dp(pat: bytes, text: bytes, global_: bool) -> int:
    m, n = len(pat), len(text)
    prev = np.arange(n + 1) if global_ else np.zeros(n + 1, dtype=np.int64)
    prev = prev.astype(np.int64)
    tc = np.array([_code(t) for t in text], dtype=np.int64)
    for i in range(1, m + 1):
        cp = _code(pat[i - 1])
        if cp > 3:
            sub = np.ones(n, dtype=np.int64)
        else:
            sub = np.where((tc == cp) | ((tc == 0) & (cp == 2)) | ((tc == 3) & (cp == 1)), 0, 1)
        cur = np.empty(n + 1, dtype=np.int64)
        cur[0] = i
        diag = prev[:-1] + sub
        up = prev[1:] + 1
        best = np.minimum(diag, up)
        for j in range(1, n + 1):  # left dependency
            v = best[j - 1]
            left = cur[j - 1] + 1
            cur[j] = v if v < left else left
        prev = cur
    return int(prev[n]) if global_ else int(prev.min())


def _mutate(rng, s: bytearray, rate: float) -> bytearray:
    out = bytearray()
    for b in s:
        r = rng.random()
        if r < rate / 3:
            continue
        if r < 2 * rate / 3:
            out.append(int(rng.choice(list(b"ACGT"))))
        if r < rate:
            out.append(int(rng.choice(list(b"ACGTN"))))
            continue
        out.append(b)
    return out


@pytest.mark.parametrize("global_", [False, True])
def test_engine_myers_against_dp(global_):
    rng = np.random.default_rng(31 + int(global_))
    alphabet = list(b"ACGTACGTacgtNRY")
    for m in (1, 2, 7, 63, 64, 65, 127, 128, 129, 200, 255, 256):
        for _ in range(3):
            pat = bytes(rng.choice(alphabet, m).tolist())
            core = _mutate(rng, bytearray(pat), 0.1)
            # A>G / T>C seen from the text: text A under pattern G is free
            core = bytearray({71: 65, 67: 84}.get(b, b) if rng.random() < 0.1 else b for b in core)
            flank = int(rng.integers(0, 20)) if not global_ else int(rng.integers(0, 3))
            text = bytes(rng.choice(list(b"ACGT"), flank).tolist()) + bytes(core) + \
                bytes(rng.choice(list(b"ACGT"), flank).tolist())
            assert oracle.engine_myers(pat, text, global_) == ea_dp(pat, text, global_), (m, pat, text)
    # empty text
    assert oracle.engine_myers(b"ACGT", b"", global_) == 4


def test_engine_infix_equals_map_infix():
    rng = np.random.default_rng(41)
    for _ in range(300):
        m = int(rng.integers(1, 257))
        pat = bytes(rng.choice(list(b"ACGTacgtN"), m).tolist())
        text = bytes(rng.choice(list(b"ACGTN"), int(rng.integers(0, 400))).tolist())
        assert oracle.engine_myers(pat, text) == oracle.infix(pat, text)


def test_engine_myers_refuses_lengths():
    assert oracle.engine_myers(b"", b"ACGT") == -1
    assert oracle.engine_myers(b"A" * 257, b"ACGT") == -1
    assert oracle.engine_myers(b"A" * 256, b"") == 256
