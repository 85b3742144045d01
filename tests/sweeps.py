"""The reference's seeded equivalence sweeps, regenerated as case lists.

Each case: (read, window, edit_dir, offset, half) with half = None for the
full matrix.  Seeds, shapes and offsets follow crates/map/src/extend.rs:
  full_sweep        avx2_full_matrix_matches_scalar  :2578-2627  seed 0x9E3779B97F4A7C15
  full_blocks       avx2_full_matrix_block_boundaries :2632-2658 seed 0x0F1E2D3C4B5A6978
  full_dispatch     avx2_dispatch_path_matches_scalar :2665-2691 seed 0xDEADBEEFCAFE1234
  band_sweep        avx2_band_matches_scalar         :2723-2822  seed 0xA5A55A5AC3C33C3C
  band_blocks       avx2_band_block_boundaries       :2828-2862  seed 0x13579BDF2468ACE0
  band_dispatch     avx2_band_dispatch_path_matches_scalar :2869-2913 seed 0x0BADC0DE12345678
The long n = 20,000 windows are returned separately (long_full, long_band).

band_interior is not one of the reference's sweeps: it places pairs inside
the interior-band envelope of every specialised half (offset >= half and
offset + m + half <= n) so that each one exercises the band kernel.
"""
from __future__ import annotations

from esperanto_amd.records import EditDir
from seqgen import Rng, rand_seq, rand_window

DIRS = [EditDir.Off, EditDir.Ag, EditDir.Tc]


def full_sweep(include_long=False):
    rng = Rng(0x9E3779B97F4A7C15)
    lens_m = [1, 2, 3, 7, 8, 15, 16, 17, 31, 32, 33, 63, 64, 100, 150, 151, 200]
    lens_n = [1, 2, 3, 7, 8, 15, 16, 17, 31, 48, 63, 64, 100, 127, 255, 257, 1000, 3000]
    cases = []
    for m in lens_m:
        for n in lens_n:
            for rep in range(3):
                read = rand_seq(rng, m)
                win = rand_seq(rng, n) if (rep == 2 or n < m) else rand_window(rng, read, n)
                d = DIRS[rng.below(3)]
                cases.append((read, win, d, 0, None))
    longs = []
    for rep in range(3):
        read = rand_seq(rng, 150)
        win = rand_seq(rng, 20000) if rep == 2 else rand_window(rng, read, 20000)
        longs.append((read, win, EditDir.Ag, 0, None))
    cases.append((b"", b"ACGT", EditDir.Off, 0, None))
    cases.append((b"ACGT", b"", EditDir.Off, 0, None))
    return (cases, longs) if include_long else cases


def full_blocks():
    rng = Rng(0x0F1E2D3C4B5A6978)
    cases = []
    for m in range(14, 21):
        for n in range(1, 21):
            for rep in range(4):
                read = rand_seq(rng, m)
                win = rand_window(rng, read, n) if (rep & 1 == 0 and n >= m) else rand_seq(rng, n)
                cases.append((read, win, DIRS[rep % 3], 0, None))
    for n in (30, 31, 32, 33, 47, 63, 64, 65):
        read = rand_seq(rng, 17)
        win = rand_window(rng, read, n)
        cases.append((read, win, EditDir.Tc, 0, None))
    return cases


def full_dispatch():
    rng = Rng(0xDEADBEEFCAFE1234)
    cases = []
    for m, n in ((6, 6), (8, 8), (9, 40), (50, 300), (120, 500)):
        for rep in range(4):
            read = rand_seq(rng, m)
            win = rand_seq(rng, n) if rep == 3 else rand_window(rng, read, n)
            cases.append((read, win, DIRS[rep % 3], 0, None))
    return cases


def band_sweep(include_long=False):
    rng = Rng(0xA5A55A5AC3C33C3C)
    shapes = [(8, 9), (8, 40), (15, 16), (16, 16), (17, 33), (20, 200), (31, 31), (32, 64),
              (60, 61), (60, 300), (100, 300), (150, 151), (150, 1000), (40, 20), (17, 9)]
    cases = []
    for m, n in shapes:
        mn = max(m, n)
        halfs = [0, 1, 2, 30, mn + 5] if n <= 64 else [0, 1, 2, 30, 300]
        for rep in range(2):
            read = rand_seq(rng, m)
            win = rand_seq(rng, n) if rep == 1 else rand_window(rng, read, n)
            for half in halfs:
                hh = min(half, mn)
                for off in (-hh - 5, -1, 0, 5, n - m, n, n + hh + 5, 4 * n):
                    d = DIRS[rng.below(3)]
                    cases.append((read, win, d, off, half))
    longs = []
    for rep in range(3):
        read = rand_seq(rng, 150)
        win = rand_seq(rng, 20000) if rep == 2 else rand_window(rng, read, 20000)
        for off, half in ((-40, 30), (0, 30), (5000, 30), (19850, 30), (-20000, 20000),
                          (20000, 20000)):
            longs.append((read, win, EditDir.Ag, off, half))
    return (cases, longs) if include_long else cases


def band_blocks():
    rng = Rng(0x13579BDF2468ACE0)
    halfs = [0, 1, 2, 7, 8, 15, 16, 17, 30, 31, 32, 33, 48]
    cases = []
    for m in range(8, 25):
        for dn in (0, 1, 15, 16, 17, 31, 32, 33):
            n = m + dn
            for rep in range(2):
                read = rand_seq(rng, m)
                win = rand_seq(rng, n) if rep == 1 else rand_window(rng, read, n)
                for half in halfs:
                    off = rng.below(120) - 60 + (n - m)
                    cases.append((read, win, DIRS[rep % 3], off, half))
    return cases


def band_dispatch():
    rng = Rng(0x0BADC0DE12345678)
    cases = []
    for m, n in ((8, 9), (10, 300), (60, 400), (120, 500)):
        for rep in range(3):
            read = rand_seq(rng, m)
            win = rand_seq(rng, n) if rep == 2 else rand_window(rng, read, n)
            off = rng.below(n) - 30
            half = [0, 1, 2, 30][rep % 4]
            cases.append((read, win, DIRS[rep % 3], off, half))
    return cases


def hint_cases(seed=0x5EED_4711, count=400):
    """extend_hint inputs: realistic windows with hints near the true diagonal,
    so both the kept-band and the escalation paths occur."""
    rng = Rng(seed)
    cases = []
    for _ in range(count):
        m = 20 + rng.below(180)
        n = max(1, m + rng.below(200) - 20)
        read = rand_seq(rng, m)
        win = rand_window(rng, read, n) if rng.below(4) else rand_seq(rng, n)
        off = rng.below(n + 40) - 20
        cases.append((read, win, DIRS[rng.below(3)], off, 30))
    for n in (1, 10, 29, 30, 31):
        read = rand_seq(rng, 25)
        cases.append((read, rand_window(rng, read, n), EditDir.Off, 0, 30))
    return cases


def _edit_near_edges(rng: Rng, read: bytes, half: int) -> bytes:
    """The read with a gap of about half bases (inserted or deleted) in its
    middle, so the optimal path runs along a band edge."""
    mid = len(read) // 2
    g = max(1, half - rng.below(3))
    if rng.below(2) == 0:
        return read[:mid] + rand_seq(rng, g) + read[mid:]
    return read[:mid] + read[mid + g:] if len(read) > mid + g else read


def band_interior(seed=0x1A7E_B0D5):
    """(read, window, dir, offset, half) for halves 16/30/32/48/64 with the
    band strictly inside the window: similar and random pairs, gaps near the
    band edge, reads of 1..300 bases, offsets at both interior limits."""
    rng = Rng(seed)
    cases = []
    for half in (16, 30, 32, 48, 64):
        for m in (1, 2, 7, 31, 64, 65, 100, 129, 150, 199, 250, 300):
            for rep in range(6):
                read = rand_seq(rng, m)
                extra = (0, 1, 17, 40)[rep % 4]
                n = m + 2 * half + extra
                lo, hi = half, n - m - half
                off = (lo, hi, lo + rng.below(hi - lo + 1))[rep % 3]
                if rep == 5:
                    win = rand_seq(rng, n)  # unrelated: ties and zero rows
                else:
                    body = read if rep < 3 else _edit_near_edges(rng, read, half)
                    w = bytearray(rand_seq(rng, n))
                    for k, b in enumerate(body):
                        j = off + k
                        if j < n:
                            w[j] = rng.base() if rng.below(25) == 0 else b
                    win = bytes(w)
                if rep == 4:
                    read = read.lower()
                cases.append((read, win, DIRS[rep % 3], off, half))
    return cases


def i16_wrap_cases():
    """Pure pairs whose i16 route (extend.rs:473-501) can pass 32767: scaled
    scores on 40-nt reads, a 3000-nt read at a 5-bit match score, positive
    mismatches.  Cases are (read, window, ExtendParams); the callers add
    lowercase (impure, i32-route) copies where they need them."""
    from esperanto_amd.records import ExtendParams
    rng = Rng(0x16_0001)
    out = []
    for match, mm, go, ge, m, n in [(1024, -4, 4, 2, 40, 50), (900, -4, 4, 2, 40, 45),
                                    (1024, -2048, 4000, 50, 40, 50), (1000, -300, 100, 10, 64, 90),
                                    (15, -8, 3, 1, 3000, 3060), (1024, 1024, 0, 0, 36, 40),
                                    (800, 300, 16, 8, 50, 70)]:
        for rep in range(3):
            read = rand_seq(rng, m)
            win = rand_window(rng, read, n) if rep != 2 else rand_seq(rng, n)
            if rep == 1:
                b = bytearray(read)
                b[m // 2] = ord("N")
                read = bytes(b)
            out.append((read, win, ExtendParams(match, mm, go, ge, DIRS[rep % 3])))
    return out
