"""CPU model of the pruned full-matrix rerun (the claim k_gotoh_full's pruned
mode rests on), checked against the unpruned local affine DP.

The model sweeps the matrix column by column like the kernel: head groups of
8 rows plus the groups meeting rows [j - d - h, j - d + h], with margins
(H = 0, E = F = NEG) for every uncomputed neighbour.  Where its exit check
passes (every computed cell with an uncomputed down, right or diagonal
neighbour has H < L - S * (m - i)), every live value (V + S * (m - i) >= L),
every comparison the traceback reads at a live cell (D vs H, E vs H, open vs
extension of E and of F) and the first row-major maximum must equal the
unpruned DP's; the test also plants a skipped live cell to show the check
catches it.  L is the band attempt's best from the oracle (extend.rs:606
run_band with half 30, as extend_hint runs it).
"""
from __future__ import annotations

import numpy as np

from esperanto_amd.records import ExtendParams
from oracle import oracle

NEG = -10 ** 9
MATCH, MIS, OE, EXT = 2, -4, -6, -2
S = MATCH


def _codes(b: bytes) -> np.ndarray:
    t = np.full(256, 4, np.int8)
    for i, ch in enumerate(b"ACGT"):
        t[ch] = i
        t[ch + 32] = i
    return t[np.frombuffer(b, np.uint8)]


def _sweep(read: bytes, win: bytes, computed=None):
    """Column sweep of the local affine DP (H floor 0).  computed(i, j) -> bool
    selects the cells evaluated; the others read as margins.  Returns H, E, F,
    D and the open-vs-extension comparisons of E and F, as (m+1) x (n+1)."""
    r, w = _codes(read), _codes(win)
    m, n = len(r), len(w)
    H = np.zeros((m + 1, n + 1), np.int64)
    E = np.full((m + 1, n + 1), NEG, np.int64)
    F = np.full((m + 1, n + 1), NEG, np.int64)
    D = np.full((m + 1, n + 1), NEG, np.int64)
    Eo = np.zeros((m + 1, n + 1), bool)  # E from open (H above + oe >= E above + ext)
    Fo = np.zeros((m + 1, n + 1), bool)
    live = np.ones((m + 1, n + 1), bool)
    for j in range(1, n + 1):
        for i in range(1, m + 1):
            if computed is not None and not computed(i, j):
                live[i, j] = False
                continue  # margin: H = 0, E = F = NEG
            sc = MATCH if (r[i - 1] == w[j - 1] and w[j - 1] < 4) else MIS
            D[i, j] = H[i - 1, j - 1] + sc
            uo, ue = H[i - 1, j] + OE, E[i - 1, j] + EXT
            E[i, j] = max(uo, ue)
            Eo[i, j] = uo >= ue
            fo, fe = H[i, j - 1] + OE, F[i, j - 1] + EXT
            F[i, j] = max(fo, fe)
            Fo[i, j] = fo >= fe
            H[i, j] = max(0, D[i, j], E[i, j], F[i, j])
    return H, E, F, D, Eo, Fo, live


def _first_max(H):
    best = H.max()
    i, j = np.argwhere(H == best)[0]
    return int(best), int(i), int(j)


def _check(H, comp, m, n, L):
    """Exit check of the pruned pass: computed cells with an uncomputed down,
    right or diagonal neighbour must hold H below L - S * (m - i)."""
    for i in range(1, m + 1):
        for j in range(1, n + 1):
            if not comp[i, j]:
                continue
            exits = (i < m and not comp[i + 1, j]) or (j < n and not comp[i, j + 1]) or \
                    (i < m and j < n and not comp[i + 1, j + 1])
            if exits and H[i, j] >= L - S * (m - i):
                return False
    return True


def _case(rng, kind):
    m = int(rng.integers(24, 48))
    n = int(rng.integers(m + 40, m + 140))
    win = bytearray(np.frombuffer(b"ACGT", np.uint8)[rng.integers(0, 4, n)].tobytes())
    p = int(rng.integers(0, n - m))
    read = bytearray(win[p:p + m])
    if kind == 1:  # substitutions
        for q in rng.choice(m, int(rng.integers(1, 4)), replace=False):
            read[q] = b"ACGT"[(b"ACGT".index(bytes([read[q]])) + 1) % 4]
    elif kind == 2:  # an indel in the middle
        k = int(rng.integers(1, 5))
        read = read[:m // 2] + win[p + m // 2 + k:p + m + k] if p + m + k <= n else read
    elif kind == 3:  # a second copy of the locus: ties elsewhere
        q = int(rng.integers(0, n - m))
        win[q:q + m] = read
    return bytes(read[:m]), bytes(win), p


def test_pruned_model_keeps_every_live_value_and_the_maximum():
    rng = np.random.default_rng(20)
    proved = 0
    for k in range(120):
        read, win, p = _case(rng, k % 4)
        m, n = len(read), len(win)
        ext, _ = oracle.band(read, win, ExtendParams(), (p, 30) if p >= 30 and p + m + 30 <= n else None)
        L = ext.score
        if L <= 0:
            continue
        Hf, Ef, Ff, Df, Eof, Fof, _ = _sweep(read, win)
        best = _first_max(Hf)
        assert best[0] >= L  # the band attempt's best bounds the full matrix's from below
        d = p  # the diagonal the kernel takes from the band maximum (j - i)
        nh = (max(0, m - (L + S - 1) // S) + 8 + 7) // 8
        comp = np.zeros((m + 2, n + 2), bool)
        for j in range(1, n + 1):
            comp[1:min(m, 8 * nh) + 1, j] = True
            lo, hi = j - d - 8, j - d + 8
            if hi >= 1 and lo <= m:
                g0, g1 = max(lo - 1, 0) // 8, (min(hi, m) - 1) // 8
                comp[8 * g0 + 1:min(m, 8 * g1 + 8) + 1, j] = True
        Hp, Ep, Fp, Dp, Eop, Fop, _ = _sweep(read, win, lambda i, j: comp[i, j])
        if not _check(Hp, comp, m, n, L):
            continue  # the kernel hands this pair to the unpruned pass
        proved += 1
        i = np.arange(m + 1)[:, None]
        thr = L - S * (m - i)
        liveH = Hf >= thr
        assert not (liveH & ~comp[:m + 1, :n + 1])[1:, 1:].any(), "a skipped cell was live"
        assert np.array_equal(Hp[liveH], Hf[liveH])
        for Vf, Vp in ((Ef, Ep), (Ff, Fp), (Df, Dp)):
            lv = (Vf >= thr) & comp[:m + 1, :n + 1]
            assert np.array_equal(Vp[lv], Vf[lv])
        lc = liveH & comp[:m + 1, :n + 1]
        assert np.array_equal((Dp == Hp)[lc], (Df == Hf)[lc])
        assert np.array_equal((Ep == Hp)[lc], (Ef == Hf)[lc])
        le = (Ef >= thr) & comp[:m + 1, :n + 1]
        assert np.array_equal(Eop[le], Eof[le])
        lf = (Ff >= thr) & comp[:m + 1, :n + 1]
        assert np.array_equal(Fop[lf], Fof[lf])
        assert _first_max(Hp) == best
    assert proved >= 40


def test_pruned_model_check_catches_a_skipped_live_cell():
    """A computed region that leaves out part of the alignment path must fail
    the exit check."""
    rng = np.random.default_rng(3)
    read, win, p = _case(rng, 0)
    m, n = len(read), len(win)
    L = 2 * m
    comp = np.zeros((m + 2, n + 2), bool)
    comp[1:9, 1:n + 1] = True  # the head only: the path below row 8 is skipped
    Hp = _sweep(read, win, lambda i, j: comp[i, j])[0]
    assert not _check(Hp, comp, m, n, L)
