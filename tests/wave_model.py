"""CPU model of the one-pair-per-wavefront kernel (esp_wave.hip), step for
step: 64 lanes in lockstep, lane l on row i0 + l and column J0 + t - l at
step t, H and E moved one lane down each step, the first lane reading the row
above from the copy the previous block's last lane wrote, margins (H = 0,
E = F = NEG) for every cell outside its row window, per-block reductions of
the row maxima and edge values, the row-window trace layout and the
single-lane traceback and finish.  The tests run it against the oracle
before the kernel ever runs on a device, and the kernel follows it line for
line.
"""
from __future__ import annotations

import numpy as np

NEG = -1_000_000_000
LANES = 64
HD_DIAG, HD_E, HD_F, HD_START = 0, 1, 2, 3
OP_M, OP_I, OP_D, OP_S = 0, 1, 2, 4


def _code_table(other: int) -> np.ndarray:
    t = np.full(256, other, np.int64)
    for i, ch in enumerate(b"ACGT"):
        t[ch] = i
        t[ch + 32] = i
    return t


RCODE, WCODE = _code_table(4), _code_table(5)


def _partner(rc: int, edit: int) -> int:
    return 0 if (edit == 1 and rc == 2) else 3 if (edit == 2 and rc == 1) else 6


def _row_window(off, half, n, i):
    c = min(max(off + i, 1), n)
    return max(c - half, 1), min(c + half, n)


def _clamp_off(off, m, n):
    return max(-m - 1, min(off, n))


def _clamp_half(half, n):
    return max(0, min(half, n))


def _lut(P, q, r):
    if q == r:
        return P["match"]
    if r == _partner(q, P["edit"]):
        return 0
    return P["mismatch"]


def _neutral(q, r):
    return q < 4 and r < 4 and (q == r or (r == 0 and q == 2) or (r == 3 and q == 1))


def fill(read: bytes, win: bytes, P: dict, off: int, half: int, wrap16: bool):
    """The kernel's sweep.  Returns (best, bi, bj, touch, trace) with trace a
    dict (i, k) -> nibble of the cell (i, lo(i) + k)."""
    m, n = len(read), len(win)
    rcodes = RCODE[np.frombuffer(read, np.uint8)]
    wcodes = WCODE[np.frombuffer(win, np.uint8)]
    lane = np.arange(LANES)
    bh = np.zeros(n + 1, np.int64)
    be = np.full(n + 1, NEG, np.int64)
    trace = {}
    best, bi, bj = 0, 0, 0
    touch = (0, 0, 0, 0)
    oe, ext = P["oe"], P["ext"]
    for i0 in range(1, m + 1, LANES):
        i1 = min(i0 + 63, m)
        i = i0 + lane
        rowok = i <= m
        win_l = [_row_window(off, half, n, int(min(x, m))) for x in i]
        lo = np.array([w[0] for w in win_l])
        hi = np.array([w[1] for w in win_l])
        lo0, _ = _row_window(off, half, n, i0)
        _, hi1 = _row_window(off, half, n, i1)
        plo, phi = _row_window(off, half, n, i0 - 1) if i0 > 1 else (1, 0)
        J0 = lo0
        T = hi1 - lo0 + (i1 - i0) + 1
        rc = np.where(rowok, rcodes[np.minimum(i, m) - 1], 4)
        partner = np.array([_partner(int(c), P["edit"]) for c in rc])
        oH = np.zeros(LANES, np.int64)
        oE = np.full(LANES, NEG, np.int64)
        oF = np.full(LANES, NEG, np.int64)
        dH = np.zeros(LANES, np.int64)
        if i0 > 1 and plo <= J0 - 1 <= phi:
            dH[0] = bh[J0 - 1]
        rmax = np.full(LANES, -1, np.int64)
        rarg = np.zeros(LANES, np.int64)
        hl = np.full(LANES, -1, np.int64)
        hr = np.full(LANES, -1, np.int64)
        for t in range(T):
            j = J0 + t - lane
            uH = np.concatenate(([0], oH[:-1]))
            uE = np.concatenate(([0], oE[:-1]))
            j0 = int(j[0])
            inb = i0 > 1 and plo <= j0 <= phi
            uH[0] = bh[j0] if inb else 0
            uE[0] = be[j0] if inb else NEG
            valid = rowok & (j >= lo) & (j <= hi)
            jin = (j >= 1) & (j <= n)
            fc = np.where(jin, wcodes[np.clip(j - 1, 0, max(n - 1, 0))], 7)
            eo, ee = uH + oe, uE + ext
            E = np.maximum(eo, ee)
            edir = (ee > eo).astype(np.int64)
            fo, fe = oH + oe, oF + ext
            F = np.maximum(fo, fe)
            fdir = (fe > fo).astype(np.int64)
            sc = np.where(rc == fc, P["match"], P["mismatch"])
            if P["edit"]:
                sc = np.where(fc == partner, 0, sc)
            D0 = dH + sc
            D = np.where(wrap16 & (D0 > 32767), NEG, D0)
            H = np.maximum(np.maximum(D, E), np.maximum(F, 0))
            hd = np.where(H == 0, HD_START, np.where(D == H, HD_DIAG, np.where(E == H, HD_E, HD_F)))
            dH = uH
            oH = np.where(valid, H, 0)
            oE = np.where(valid, E, NEG)
            oF = np.where(valid, F, NEG)
            for l in np.nonzero(valid)[0]:
                trace[(int(i[l]), int(j[l] - lo[l]))] = int(hd[l] | (edir[l] << 2) | (fdir[l] << 3))
            up = valid & (H > rmax)
            rmax = np.where(up, H, rmax)
            rarg = np.where(up, j, rarg)
            hl = np.where(valid & (j == lo), H, hl)
            hr = np.where(valid & (j == hi), H, hr)
            if valid[63]:
                bh[j[63]] = H[63]
                be[j[63]] = E[63]
        vmax = int(np.max(np.where(rowok, rmax, -1)))
        if vmax > best:
            f = int(np.nonzero(rowok & (rmax == vmax))[0][0])
            best, bi, bj = vmax, i0 + f, int(rarg[f])
        if not touch[0]:
            L = rowok & (hl == rmax)
            R = rowok & (hi < n) & (hr == rmax)
            C = np.nonzero(L | R)[0]
            if C.size:
                f = int(C[0])
                flo, _ = _row_window(off, half, n, i0 + f)
                left = int(L[f])
                touch = (1, i0 + f, left, int(flo <= 1) if left else 0)
    return best, bi, bj, touch, trace


def traceback(read, win, off, half, trace, bi, bj):
    n = len(win)
    i, j, state = bi, bj, 0
    ops = []
    while i > 0 and j > 0:
        lo, hi = _row_window(off, half, n, i)
        if j < lo or j > hi:
            break
        nib = trace[(i, j - lo)]
        if state == 0:
            hd = nib & 3
            if hd == HD_DIAG:
                ops.append(OP_M); i -= 1; j -= 1
            elif hd == HD_E:
                ops.append(OP_I); state = 1 if (nib >> 2) & 1 else 0; i -= 1
            elif hd == HD_F:
                ops.append(OP_D); state = 2 if (nib >> 3) & 1 else 0; j -= 1
            else:
                break
        elif state == 1:
            ops.append(OP_I)
            if not (nib >> 2) & 1:
                state = 0
            i -= 1
        else:
            ops.append(OP_D)
            if not (nib >> 3) & 1:
                state = 0
            j -= 1
    return max(i, 0), max(j, 0), ops[::-1]


def finish(read, win, P, best, rs, fs, ops):
    m, n = len(read), len(win)
    nm, ni, nd = ops.count(OP_M), ops.count(OP_I), ops.count(OP_D)
    read_end = rs + nm + ni
    nl = nr = 0
    if best > 0:
        delta = 0
        for k in range(1, min(rs, fs) + 1):
            q, r = int(RCODE[read[rs - k]]), int(WCODE[win[fs - k]])
            if not _neutral(q, r):
                break
            delta += _lut(P, q, r)
            if delta == 0:
                nl = k
        ref_end = fs + nm + nd
        delta = 0
        for k in range(min(m - read_end, n - ref_end)):
            q, r = int(RCODE[read[read_end + k]]), int(WCODE[win[ref_end + k]])
            if not _neutral(q, r):
                break
            delta += _lut(P, q, r)
            if delta == 0:
                nr = k + 1
        rs -= nl
        fs -= nl
        read_end += nr
    cig = []

    def push(kind, ln):
        if cig and cig[-1][0] == kind:
            cig[-1][1] += ln
        else:
            cig.append([kind, ln])

    if rs > 0:
        push(OP_S, rs)
    if nl:
        push(OP_M, nl)
    for op in ops:
        push(op, 1)
    if nr:
        push(OP_M, nr)
    if read_end < m:
        push(OP_S, m - read_end)
    return rs, read_end, fs, best, [(k, l) for k, l in cig]


def _pure(s: bytes) -> bool:
    return all(b in b"ACGTN" for b in s)


def wave_pair(read: bytes, win: bytes, params, mode: str, offset: int = 0, half: int = 0):
    """One call as the kernel runs it: mode 'band', 'full' or 'hint'.
    Returns (read_start, read_end, ref_start, score, cigar [(op, len)],
    touch (touched, row, left, clamped), path)."""
    m, n = len(read), len(win)
    P = {"match": params.match_score, "mismatch": params.mismatch, "oe": -(params.gap_open + params.gap_ext),
         "ext": -params.gap_ext, "edit": int(params.edit_dir)}
    if m == 0 or n == 0:
        return 0, 0, 0, 0, ([(OP_S, m)] if m else []), (0, 0, 0, 0), 2
    path = 0
    if mode == "full":
        band = False
    elif mode == "hint":
        band = 30 < n
        path = 0 if band else 3
    else:
        band = True
    off = _clamp_off(offset, m, n) if band else 0
    hf = _clamp_half(30 if mode == "hint" else half, n) if band else n
    smax = max(P["match"], P["mismatch"], 0)
    wrap16 = m * smax > 32767 and _pure(read) and _pure(win)
    best, bi, bj, touch, tr = fill(read, win, P, off, hf, wrap16)
    if mode == "hint" and band:
        lhs = np.int32(np.uint32((best * 20) & 0xFFFFFFFF).astype(np.int32))
        rhs = np.int32(np.uint32((m * P["match"] * 17) & 0xFFFFFFFF).astype(np.int32))
        if touch[0] or lhs < rhs:
            off, hf = 0, n
            best, bi, bj, touch, tr = fill(read, win, P, off, hf, wrap16)
            path = 1
    rs, fs, ops = traceback(read, win, off, hf, tr, bi, bj)
    a, b, c, d, cig = finish(read, win, P, best, rs, fs, ops)
    return a, b, c, d, cig, touch, path
