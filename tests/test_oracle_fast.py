"""Pins the oracle's restatement of try_fast_lane (align.rs:2274-2401).

The reference has no test of its own for this lane, so the restatement is
pinned by what the mathematics fixes:
- against an O(m^2) enumeration of every subarray of the diagonal's column
  scores (the lane's result is the maximum sum, ending at the first column
  where it is reached, starting at the latest start with that sum, and it is
  refused when that subarray holds a scored mismatch or the sum is <= 0);
- against the reference's own stated equivalence: when the lane accepts, the
  banded affine DP (extend_hint over continuous_extension_window with the
  chain-diagonal hint, align.rs:1886-1919) reaches the same optimum score,
  and the same read span, genomic start and CIGAR unless the DP's tie rules
  pick another alignment of that score (a run extended past a mismatch by
  matches that win it back, which the lane's first-maximum rule stops
  before).  The equivalence holds where the diagonal holds
  only upper-case ACGT on both sides and lies inside the contig: the lane
  scores N as 0 and compares bytes as they are, while the DP's SubstLut
  (extend.rs:161-186) scores N as a mismatch and folds lower case, so the
  check is made on those diagonals (the lane itself is reproduced as written
  everywhere, and the enumeration pins it there);
- each acceptance condition on hand-made chains (no anchors, an anchor off
  the diagonal, an anchor gap above k + w + 5, a read end uncovered by more
  than 2 bases, a mismatch inside the best run).
"""
from __future__ import annotations

import numpy as np
import pytest

from esperanto_amd.records import DiagHint, ExtendParams, Match, SoftClip
from esperanto_amd.reference import Reference
from esperanto_amd.unspliced import continuous_extension_window
from fastgen import fast_reference, random_chain
from oracle import oracle


@pytest.fixture(scope="module", autouse=True)
def _lib():
    oracle.lib()


def _column_scores(seq_upper: bytes, q: bytes, diag: int, p: ExtendParams, clen: int):
    """(score, is_mismatch) per read column on the diagonal, from the N-masked contig."""
    out = []
    for i, qb in enumerate(q):
        r = i + diag
        if r < 0 or r >= clen:
            out.append((0, False))
            continue
        rb = seq_upper[r]
        rb = rb if rb in b"ACGT" else ord("N")
        if qb == ord("N") or rb == ord("N"):
            out.append((0, False))
        elif qb == rb:
            out.append((p.match_score, False))
        elif (p.edit_dir == 1 and (rb, qb) == (65, 71)) or (p.edit_dir == 2 and (rb, qb) == (84, 67)):
            out.append((0, False))
        else:
            out.append((p.mismatch, True))
    return out


def _brute(cols):
    """Best subarray by enumeration: (sum, start, end, mismatches) with the
    lane's tie rules, or None when no subarray sums above 0."""
    best = None
    m = len(cols)
    pre = [0]
    for s, _ in cols:
        pre.append(pre[-1] + s)
    for e in range(1, m + 1):
        for s in range(e - 1, -1, -1):
            v = pre[e] - pre[s]
            if v > 0 and (best is None or v > best[0]):
                best = (v, s, e)
            elif best is not None and v == best[0] and e == best[2] and s > best[1]:
                best = (v, s, e)
    if best is None:
        return None
    v, s, e = best
    return v, s, e, sum(1 for c in cols[s:e] if c[1])


def _accepts(q, anchors, m, k=15, w=5):
    if not anchors:
        return False
    d = anchors[0][1] - anchors[0][0]
    if any(r - a != d for a, r in anchors):
        return False
    if any(((b[0] - a[0]) & 0xFFFFFFFF) > k + w + 5 for a, b in zip(anchors, anchors[1:])):
        return False
    qs, qe = anchors[0][0], anchors[-1][0] + k
    return qs <= 2 and ((m - qe) & 0xFFFFFFFF) <= 2


@pytest.mark.parametrize("edit_dir", [0, 1, 2])
def test_fast_lane_against_enumeration(edit_dir):
    ref, seqs = fast_reference(3)
    o = oracle.OracleRef(ref)
    p = ExtendParams(edit_dir=edit_dir)
    rng = np.random.default_rng(17 + edit_dir)
    accepted = refused = 0
    for _ in range(400):
        q, c, anchors = random_chain(rng, seqs)
        got = oracle.fast_lane(o, q, c, anchors, params=p)
        if not _accepts(q, anchors, len(q)):
            assert got is None
            continue
        diag = anchors[0][1] - anchors[0][0]
        b = _brute(_column_scores(seqs[c].upper(), q, diag, p, len(seqs[c])))
        if b is None or b[3] > 0:
            assert got is None, (q, c, anchors)
            refused += 1
            continue
        ext, wstart = got
        v, s, e, _ = b
        flank = len(q) // 2 + 30
        assert wstart == max(anchors[0][1] - flank, 0)
        assert (ext.score, ext.read_start, ext.read_end) == (v, s, e)
        assert wstart + ext.ref_start == s + diag
        want = ([SoftClip(s)] if s else []) + [Match(e - s)] + ([SoftClip(len(q) - e)] if e < len(q) else [])
        assert ext.cigar == want
        accepted += 1
    assert accepted > 50 and refused > 10


def test_fast_lane_equals_banded_dp_when_it_accepts():
    """align.rs:2268-2273: zero mismatches in the best run make the gap-free
    diagonal the DP optimum; the DP fallback (align.rs:1886-1919) agrees."""
    ref, seqs = fast_reference(4)
    o = oracle.OracleRef(ref)
    rng = np.random.default_rng(5)
    checked = ties = 0
    for _ in range(1500):
        q, c, anchors = random_chain(rng, seqs)
        got = oracle.fast_lane(o, q, c, anchors)
        if got is None:
            continue
        ext, wstart = got
        k = 15
        diag = anchors[0][1] - anchors[0][0]
        span = seqs[c][diag:diag + len(q)] if diag >= 0 else b""
        if len(span) != len(q) or any(b not in b"ACGT" for b in q + span):
            continue
        rs, re_ = anchors[0][1], anchors[-1][1] + k
        qs, qe = anchors[0][0], anchors[-1][0] + k
        clen = int(ref.lens[c])
        ws, we = continuous_extension_window((rs, re_), (qs, qe), len(q), clen)
        window = oracle.decode(o, c, ws, we)
        dp = oracle.extend_hint(q, window, ExtendParams(), DiagHint((rs - ws) - qs))
        assert dp.score == ext.score, (q, c, anchors)
        same = ((dp.read_start, dp.read_end, ws + dp.ref_start, dp.cigar) ==
                (ext.read_start, ext.read_end, wstart + ext.ref_start, ext.cigar))
        if not same:
            # a tie: the DP's alignment reaches the same score over a longer
            # span that contains the lane's run
            assert dp.read_start <= ext.read_start and dp.read_end >= ext.read_end
            ties += 1
        checked += 1
    assert checked > 60 and ties < checked // 3


def test_fast_lane_conditions():
    seq = ("ACGTTGCAAGGCATTACGACCAAGTTACAGGATCCAAGTTGACGTTACAGGCATTAC" * 4).encode()
    ref = Reference.from_sequences([seq])
    o = oracle.OracleRef(ref)
    q = seq[40:120]
    good = [(0, 40), (20, 60), (40, 80), (65, 105)]
    ext, wstart = oracle.fast_lane(o, q, 0, good)
    assert (ext.score, ext.read_start, ext.read_end, wstart) == (160, 0, 80, 0)
    assert wstart + ext.ref_start == 40 and ext.cigar == [Match(80)]
    assert oracle.fast_lane(o, q, 0, []) is None
    assert oracle.fast_lane(o, q, 0, [(0, 40), (20, 61), (40, 80), (65, 105)]) is None
    assert oracle.fast_lane(o, q, 0, [(0, 40), (26, 66), (45, 85), (65, 105)]) is None  # gap 26 > 25
    assert oracle.fast_lane(o, q, 0, [(0, 40), (25, 65), (45, 85), (65, 105)]) is not None  # gap 25
    assert oracle.fast_lane(o, q, 0, [(3, 43), (20, 60), (40, 80), (65, 105)]) is None  # qstart 3
    assert oracle.fast_lane(o, q, 0, [(0, 40), (20, 60), (40, 80), (62, 102)]) is None  # 3 uncovered
    assert oracle.fast_lane(o, q, 0, [(0, 40), (20, 60), (40, 80), (63, 103)]) is not None
    # a mismatch inside the best run refuses; one at the read end is clipped
    mid = bytearray(q)
    mid[40] = ord("A") if mid[40] != ord("A") else ord("C")
    assert oracle.fast_lane(o, bytes(mid), 0, good) is None
    end = bytearray(q)
    end[79] = ord("A") if end[79] != ord("A") else ord("C")
    ext, _ = oracle.fast_lane(o, bytes(end), 0, good)
    assert (ext.score, ext.read_end, ext.cigar) == (158, 79, [Match(79), SoftClip(1)])
    assert oracle.fast_lane(o, b"N" * 80, 0, good) is None  # best sum 0
