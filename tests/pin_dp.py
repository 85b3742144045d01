"""An independent statement of the banded local Gotoh DP and of the traceback
choice, used to pin the oracle where the reference's literal tests do not.

Nothing here reads direction bits.  The DP is held as whole matrices of
values (H, E, F and the diagonal candidate D).  The set of co-optimal paths
is every backward walk from the best cell that only takes steps whose value
reproduces the cell's value; the reference's choice among them is stated as
a preference order over those steps (crates/map/src/extend.rs):
  H state: diagonal, then E, then F (the cascade F -> E -> diag with the
           later source winning ties, :763-782); a cell of value 0 ends the
           walk (:783-787, :918);
  E / F state: close the gap (go back to H) unless extending is strictly
           better (`better(e_ext, e_open)` is strict `>`, :731, :746).
The best cell is the first maximum in row-major order (:789-792, :823-827).
finish() restates end trimming, the neutral-flank scan and the soft clips of
finish_extension (:932-1059) on a step list.
"""
from __future__ import annotations

from typing import List, Optional, Sequence, Tuple

NINF = -(1 << 40)
_CODE = {ord("A"): 0, ord("a"): 0, ord("C"): 1, ord("c"): 1, ord("G"): 2, ord("g"): 2,
         ord("T"): 3, ord("t"): 3}


def code(b: int) -> int:
    return _CODE.get(b, 4)


def subst(p, q: int, r: int) -> int:
    """SubstLut score of read code q against reference code r (extend.rs:161-186)."""
    if q < 4 and q == r:
        return p.match_score
    d = int(p.edit_dir)
    if d == 1 and q == 2 and r == 0:
        return 0
    if d == 2 and q == 1 and r == 3:
        return 0
    return p.mismatch


def windows(m: int, n: int, band: Optional[Tuple[int, int]]):
    off, half = (0, n) if band is None else (band[0], max(band[1], 0))
    w = [None]
    for i in range(1, m + 1):
        c = min(max(off + i, 1), n)
        w.append((max(c - half, 1), min(c + half, n)))
    return w


class Dp:
    def __init__(self, read: bytes, win: bytes, p, band=None, diag_cap: Optional[int] = None):
        self.read, self.win, self.p = read, win, p
        m, n = len(read), len(win)
        self.m, self.n = m, n
        self.W = windows(m, n, band)
        go, ge = p.gap_open, p.gap_ext
        self.go, self.ge = go, ge
        H = [[0] * (n + 1) for _ in range(m + 1)]
        E = [[NINF] * (n + 1) for _ in range(m + 1)]
        F = [[NINF] * (n + 1) for _ in range(m + 1)]
        D = [[NINF] * (n + 1) for _ in range(m + 1)]
        self.qc = [code(b) for b in read]
        self.rc = [code(b) for b in win]
        for i in range(1, m + 1):
            lo, hi = self.W[i]
            for j in range(lo, hi + 1):
                e = H[i - 1][j] - go - ge
                if E[i - 1][j] > NINF:
                    e = max(e, E[i - 1][j] - ge)
                f = H[i][j - 1] - go - ge
                if F[i][j - 1] > NINF:
                    f = max(f, F[i][j - 1] - ge)
                d = H[i - 1][j - 1] + subst(p, self.qc[i - 1], self.rc[j - 1])
                if diag_cap is not None and d > diag_cap:
                    d = NINF
                H[i][j] = max(0, d, e, f)
                E[i][j], F[i][j], D[i][j] = e, f, d
        self.H, self.E, self.F, self.D = H, E, F, D

    def inside(self, i: int, j: int) -> bool:
        if i <= 0 or j <= 0 or i > self.m:
            return False
        lo, hi = self.W[i]
        return lo <= j <= hi

    def best_cell(self):
        best, bi, bj = 0, 0, 0
        for i in range(1, self.m + 1):
            lo, hi = self.W[i]
            for j in range(lo, hi + 1):
                if self.H[i][j] > best:
                    best, bi, bj = self.H[i][j], i, j
        return best, bi, bj

    # A step is (op, delta) with op in "MID", listed end-first (walk order).
    def _h_moves(self, i, j):
        """Co-optimal moves out of H state at (i, j), most preferred first."""
        if not self.inside(i, j) or self.H[i][j] == 0:
            return []
        h = self.H[i][j]
        mv = []
        if self.D[i][j] == h:
            s = subst(self.p, self.qc[i - 1], self.rc[j - 1])
            mv.append((("M", s), (i - 1, j - 1, "H")))
        if self.E[i][j] == h:
            mv.extend(self._gap_moves(i, j, "E", first=True))
        if self.F[i][j] == h:
            mv.extend(self._gap_moves(i, j, "F", first=True))
        return mv

    def _gap_moves(self, i, j, st, first=False):
        """One gap step at (i, j) in state st: close (back to H) before extend."""
        oe, ext = -(self.go + self.ge), -self.ge
        delta = oe if first else ext
        if st == "E":
            v, prev_h, prev_g, nxt, op = self.E[i][j], self.H[i - 1][j], self.E[i - 1][j], (i - 1, j), "I"
        else:
            v, prev_h, prev_g, nxt, op = self.F[i][j], self.H[i][j - 1], self.F[i][j - 1], (i, j - 1), "D"
        mv = []
        if v == prev_h + oe:
            mv.append(((op, delta), (nxt[0], nxt[1], "H")))
        if prev_g > NINF and v == prev_g + ext:
            mv.append(((op, delta), (nxt[0], nxt[1], st)))
        return mv

    def _moves(self, i, j, st):
        if st == "H":
            return self._h_moves(i, j)
        if not self.inside(i, j):
            return []
        return self._gap_moves(i, j, st)

    def co_optimal_paths(self, limit: int = 5000):
        """Every co-optimal walk from the best cell as (steps end-first,
        start (i, j)), in preference order; None when more than `limit`."""
        _, bi, bj = self.best_cell()
        out = []

        def walk(i, j, st, acc):
            if len(out) > limit:
                return
            mv = self._moves(i, j, st)
            if not mv:
                out.append((list(acc), (i, j)))
                return
            for step, (ni, nj, ns) in mv:
                acc.append(step)
                walk(ni, nj, ns, acc)
                acc.pop()

        walk(bi, bj, "H", [])
        return None if len(out) > limit else out

    def touch(self):
        """TouchRec from the values: the first row whose left edge, else
        whose right edge below n, holds the row maximum (extend.rs:803-821)."""
        for i in range(1, self.m + 1):
            lo, hi = self.W[i]
            if lo > hi:
                continue
            row = self.H[i][lo:hi + 1]
            mx = max(row)
            if self.H[i][lo] == mx:
                return (True, i, True, lo <= 1)
            if hi < self.n and self.H[i][hi] == mx:
                return (True, i, False, hi >= self.n)
        return (False, 0, False, False)

    def finish(self, steps_end_first: Sequence[Tuple[str, int]], start: Tuple[int, int]):
        """finish_extension's tail (extend.rs:932-1059) on a walked path.
        Returns (read_start, read_end, ref_start, score, [(op, len)])."""
        steps = list(reversed(steps_end_first))
        ns = len(steps)
        lead = trail = 0
        acc = 0
        for k, (_, d) in enumerate(steps):
            acc += d
            if acc <= 0:
                lead = k + 1
        acc = 0
        for k in range(ns - 1, -1, -1):
            acc += steps[k][1]
            if acc <= 0:
                trail = ns - k
        if ns > 0 and lead + trail >= ns:
            lead = ns // 2
            trail = ns - lead - 1
        i0, j0 = start
        read_start = max(i0, 0) + sum(1 for op, _ in steps[:lead] if op in "MI")
        ref_start = max(j0, 0) + sum(1 for op, _ in steps[:lead] if op in "MD")
        kept = steps[lead:ns - trail]
        score = sum(d for _, d in kept)
        ops: List[list] = []
        for op, _ in kept:
            if ops and ops[-1][0] == op:
                ops[-1][1] += 1
            else:
                ops.append([op, 1])
        read_end = read_start + sum(1 for op, _ in kept if op in "MI")
        if score > 0:
            def neutral(q, r):
                return q < 4 and r < 4 and (q == r or (r == 0 and q == 2) or (r == 3 and q == 1))
            nl = nr = 0
            delta = 0
            for k in range(1, min(read_start, ref_start) + 1):
                q, r = self.qc[read_start - k], self.rc[ref_start - k]
                if not neutral(q, r):
                    break
                delta += subst(self.p, q, r)
                if delta == 0:
                    nl = k
            ref_end = ref_start + sum(1 for op, _ in kept if op in "MD")
            delta = 0
            for k in range(min(self.m - read_end, self.n - ref_end)):
                q, r = self.qc[read_end + k], self.rc[ref_end + k]
                if not neutral(q, r):
                    break
                delta += subst(self.p, q, r)
                if delta == 0:
                    nr = k + 1
            if nl:
                ops = [["M", nl]] + ops
            if nr:
                ops = ops + [["M", nr]]
            merged: List[list] = []
            for op, ln in ops:
                if merged and merged[-1][0] == op:
                    merged[-1][1] += ln
                else:
                    merged.append([op, ln])
            ops = merged
            read_start -= nl
            ref_start -= nl
            read_end += nr
        cig = []
        if read_start > 0:
            cig.append(["S", read_start])
        for op, ln in ops:
            if cig and cig[-1][0] == op:
                cig[-1][1] += ln
            else:
                cig.append([op, ln])
        if read_end < self.m:
            cig.append(["S", self.m - read_end])
        return read_start, read_end, ref_start, score, [tuple(x) for x in cig]


def brute_force_co_optimal(read: bytes, win: bytes, p):
    """Full-matrix only, for tiny inputs: every local alignment (start cell and
    op string) ending at the best cell, scored by its ops alone (a maximal run
    of k gap ops costs gap_open + k * gap_ext), kept when its score is the
    best cell's value and every prefix that ends before an M or before a new
    gap run is positive (a path starts where its running score last touched
    0).  Returns (best, {(start, ops end-first)})."""
    dp = Dp(read, win, p, None)
    best, bi, bj = dp.best_cell()
    if best == 0:
        return 0, set()
    found = set()

    def score(start, ops_end_first):
        i, j = start
        s, prev, pref = 0, None, []
        for op in reversed(ops_end_first):
            if op == "M" or op != prev:
                pref.append(s)
            if op == "M":
                s += subst(p, code(read[i]), code(win[j]))
                i, j = i + 1, j + 1
            else:
                s -= p.gap_ext + (p.gap_open if prev != op else 0)
                i, j = (i + 1, j) if op == "I" else (i, j + 1)
            prev = op
        return s, pref

    def rec(i, j, ops):
        if ops:
            s, pref = score((i, j), ops)
            if s == best and all(x > 0 for x in pref[1:]):
                found.add(((i, j), "".join(ops)))
        if i > 0 and j > 0:
            ops.append("M")
            rec(i - 1, j - 1, ops)
            ops.pop()
        if i > 0:
            ops.append("I")
            rec(i - 1, j, ops)
            ops.pop()
        if j > 0:
            ops.append("D")
            rec(i, j - 1, ops)
            ops.pop()

    rec(bi, bj, [])
    return best, found
