"""Pins the placement-walk restatements of the oracle (rescore_with,
recount_mm_ea_in, exact_kmer_support, intron_length_penalty) before they
check the device walk kernel.

- the reference's own known answers: rescore_strict_manual_cigars
  (align.rs:4216-4262), rescore_strict_splice_signal (align.rs:4265-4303),
  exact_kmer_support_counts (align.rs:4306-4317);
- intron_length_penalty against its closed form at the decade points and
  against the product's threshold-table evaluation over every length below
  2^21 and at every threshold of the full u32 range;
- rescore invariants: the StrictScore fields reassemble the score, recount
  equals rescore's (mm, ea_ag, ea_tc), the EditDir exemption moves exactly
  the exempted substitutions' penalty;
- exact_kmer_support against an independent run-length count (a k-mer is
  counted each time k consecutive good columns accumulate in one M block).
"""
from __future__ import annotations

import math

import numpy as np
import pytest

from esperanto_amd import _native
from esperanto_amd.records import Ins, Match, RefSkip
from esperanto_amd.reference import Reference
from oracle import oracle
from seqgen import revcomp
from walkgen import random_placement, random_walk_reference


@pytest.fixture(scope="module", autouse=True)
def _lib():
    oracle.lib()


def _w(ops):
    return [o.word for o in ops]


def _on(seq: str):
    ref = Reference.from_sequences([seq.encode()])
    return ref, oracle.OracleRef(ref)


READ50 = b"ACGT" * 12 + b"AC"


def test_rescore_strict_manual_cigars():
    """align.rs:4216-4262"""
    _, o = _on("ACGT" * 50)
    s = oracle.rescore(o, READ50, 0, 0, _w([Match(50)]))
    assert (s["score"], s["matches"], s["mm"], s["aligned"]) == (100, 50, 0, 50)
    r2 = bytearray(READ50)
    r2[25] = ord("C") if r2[25] == ord("A") else ord("A")
    s = oracle.rescore(o, bytes(r2), 0, 0, _w([Match(50)]))
    assert (s["score"], s["mm"]) == (49 * 2 - 4, 1)
    # ref C at 25, so the test's branch plants read G over ref A at 24
    r3 = bytearray(READ50)
    r3[24] = ord("G")
    s = oracle.rescore(o, bytes(r3), 0, 0, _w([Match(50)]))
    assert (s["score"], s["mm"]) == (49 * 2 - 4, 0)
    assert s["ea_ag"] + s["ea_tc"] == 1
    ins_read = READ50[:24] + b"TT" + READ50[24:48]
    s = oracle.rescore(o, ins_read, 0, 0, _w([Match(24), Ins(2), Match(24)]))
    assert (s["score"], s["ins"]) == (48 * 2 - (4 + 2 * 2), 1)


EXON1 = "ACGTTGCAAGGCATTACGACCAAGTTACAGGATCCAAGTTGACGTTACAGGCATTAC"
INTRON = "GT" + "N" * 56 + "AG"
EXON2 = "CATGGACTTGAACCTTGACGTACATGGACTTGAACCTTGACGTA"


def _splice_case():
    seq = EXON1 + INTRON + EXON2
    start = len(EXON1) - 20
    e2 = len(EXON1) + len(INTRON)
    read = (seq[start:len(EXON1)] + seq[e2:e2 + 20]).encode()
    return seq, start, read, len(INTRON)


def test_rescore_strict_splice_signal():
    """align.rs:4265-4303: 40 x 2 + GT-AG bonus - penalty(60), on either read strand."""
    seq, start, read, n = _splice_case()
    _, o = _on(seq)
    cig = _w([Match(20), RefSkip(n), Match(20)])
    s = oracle.rescore(o, read, 0, start, cig)
    expected = 40 * 2 + 4 - oracle.intron_length_penalty(n)
    assert oracle.intron_length_penalty(60) == 9
    assert (s["score"], s["n_splice"], s["aligned"]) == (expected, 1, 40)
    assert s["splice_score"] == 4 - 9
    # the minus-strand placement scores the same query (revcomp of revcomp)
    assert oracle.rescore(o, revcomp(revcomp(read)), 0, start, cig)["score"] == expected


def test_rescore_known_junction_and_signals():
    seq, start, read, n = _splice_case()
    _, o = _on(seq)
    cig = _w([Match(20), RefSkip(n), Match(20)])
    j = [(0, len(EXON1), len(EXON1) + n)]
    s = oracle.rescore(o, read, 0, start, cig, junctions=j)
    assert s["score"] == 80 + 6 and s["splice_score"] == 6
    # a library entry on another contig or with another end does not hit
    for miss in ([(1, len(EXON1), len(EXON1) + n)], [(0, len(EXON1), len(EXON1) + n + 1)]):
        assert oracle.rescore(o, read, 0, start, cig, junctions=miss)["score"] == 80 + 4 - 9
    # splice.rs:78-103 motifs on both strands; unknown bytes are non-canonical
    sig = oracle.splice_signal_score
    assert sig(b"GT", b"AG", False) == 4 and sig(b"GC", b"AG", False) == 2
    assert sig(b"AT", b"AC", False) == 1 and sig(b"GT", b"AC", False) == -8
    assert sig(b"CT", b"AC", True) == 4 and sig(b"CT", b"GC", True) == 2
    assert sig(b"GT", b"AT", True) == 1 and sig(b"GT", b"AG", True) == -8
    assert sig(b"NT", b"AG", False) == -8


def test_exact_kmer_support_counts():
    """align.rs:4306-4317"""
    _, o = _on("ACGT" * 50)
    read = (b"ACGT" * 13)[:50]
    assert oracle.kmer_support(o, read, 0, 0, _w([Match(50)])) == 3
    r2 = bytearray(read)
    r2[20] = ord("C") if r2[20] == ord("A") else ord("A")
    assert oracle.kmer_support(o, bytes(r2), 0, 0, _w([Match(50)])) < 3
    assert oracle.kmer_support(o, b"", 0, 0, _w([Match(50)])) == 0
    assert oracle.kmer_support(o, read, 0, 0, _w([Match(50)]), k=0) == 0


def _penalty_closed(n: int) -> int:
    return int(8.0 * abs(math.log10(max(n, 1)) - 3.0))


def test_intron_length_penalty_closed_form():
    """splice.rs:119-121: 8 x |log10(len) - 3|, truncated; 0 at 1 kb, 8 per decade."""
    for n, want in ((0, 24), (1, 24), (10, 16), (100, 8), (1000, 0), (10_000, 8), (100_000, 16),
                    (1_000_000, 24), (60, 9), (2**32 - 1, 53)):
        assert oracle.intron_length_penalty(n) == want == _penalty_closed(n)
    rng = np.random.default_rng(3)
    for n in rng.integers(0, 2**32, 2000).tolist():
        assert oracle.intron_length_penalty(n) == _penalty_closed(n)


def test_product_intron_penalty_table_matches_formula():
    """The walk kernel's threshold table (built on the host) against the
    oracle's f64 formula: every length below 2^21, then one either side of
    every step of the penalty up to u32::MAX."""
    L = _native.lib()
    for n in range(0, 1 << 21, 1):
        if L.esp_intron_length_penalty(n) != _penalty_closed(n):
            pytest.fail(f"len {n}: {L.esp_intron_length_penalty(n)} != {_penalty_closed(n)}")
    for p in range(0, 54):
        # lengths where the closed form reaches p above 1 kb
        lo, hi = 1000, 2**32 - 1
        while lo < hi:
            mid = (lo + hi) // 2
            if _penalty_closed(mid) >= p:
                hi = mid
            else:
                lo = mid + 1
        for n in (lo - 1, lo, lo + 1):
            if 0 <= n < 2**32:
                assert L.esp_intron_length_penalty(n) == oracle.intron_length_penalty(n), n


def test_rescore_field_invariants():
    rng = np.random.default_rng(11)
    ref, seqs = random_walk_reference(5)
    o = oracle.OracleRef(ref)
    for _ in range(300):
        q, c, pos, cig = random_placement(rng, ref, seqs)
        for d in (0, 1, 2):
            s = {k: int(v) for k, v in zip(oracle.STRICT_DTYPE.names, oracle.rescore(o, q, c, pos, cig, dir=d))}
            gaps = sum(4 + 2 * (w >> 4) for w in cig if (w & 15) in (1, 2))
            exempt = s["ea_ag"] if d == 1 else s["ea_tc"] if d == 2 else 0
            subs = s["mm"] + s["ea_ag"] + s["ea_tc"]
            assert s["score"] == 2 * s["matches"] - 4 * (subs - exempt) - gaps + s["splice_score"]
            assert s["aligned"] == s["matches"] + subs + sum((w >> 4) for w in cig if (w & 15) in (1, 2))
            assert s["clipped"] == sum((w >> 4) for w in cig if (w & 15) == 4)
            assert s["n_splice"] == sum(1 for w in cig if (w & 15) == 3)
            if len(q) >= sum((w >> 4) for w in cig if (w & 15) in (0, 1, 4)):
                assert oracle.recount(o, q, c, pos, cig) == (s["mm"], s["ea_ag"], s["ea_tc"])


def _kmer_runs(o, q, contig, pos, cig, d, k):
    """Run-length count: each M block keeps the length of its current run of
    good columns; a run reaching k counts one k-mer and starts again."""
    clen = None
    count, ro, rf = 0, 0, pos
    for w in cig:
        op, n = w & 15, w >> 4
        if op == 0:
            run = 0
            for t in range(n):
                qb = q[ro + t] if ro + t < len(q) else ord("N")
                rb = oracle.decode(o, contig, rf + t, rf + t + 1)
                rb = rb[0] if rb else ord("N")
                good = qb == rb or (d == 1 and (rb, qb) == (65, 71)) or (d == 2 and (rb, qb) == (84, 67))
                run = run + 1 if good else 0
                if run == k:
                    count += 1
                    run = 0
            ro += n
            rf += n
        elif op in (1, 4):
            ro += n
        else:
            rf += n
    return count if q and k else 0


def test_kmer_support_equals_run_length_count():
    rng = np.random.default_rng(12)
    ref, seqs = random_walk_reference(6)
    o = oracle.OracleRef(ref)
    for _ in range(150):
        q, c, pos, cig = random_placement(rng, ref, seqs)
        for d in (0, 1, 2):
            for k in (1, 5, 15):
                assert oracle.kmer_support(o, q, c, pos, cig, dir=d, k=k) == _kmer_runs(o, q, c, pos, cig, d, k)


def test_walk_past_contig_end_reads_n():
    """Columns past the contig end read N: unscored, uncounted (align.rs:1082-1092)."""
    ref, o = _on("ACGTACGTAC")
    s = oracle.rescore(o, b"ACGTACGTACGGGG", 0, 0, _w([Match(14)]))
    assert (s["score"], s["aligned"], s["matches"]) == (20, 10, 10)
    s = oracle.rescore(o, b"ACGT", 0, 50, _w([Match(4)]))
    assert (s["score"], s["aligned"]) == (0, 0)
    assert oracle.kmer_support(o, b"ACGTACGTACGGGG", 0, 0, _w([Match(14)]), k=5) == 2
