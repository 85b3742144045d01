"""Known-answer tests of the reference, restated over an implementation object.

Each ``kat_*`` function takes ``impl`` with three methods mirroring the
reference API:
  impl.extend(read, window, params) -> Extension          (extend.rs:347)
  impl.extend_hint(read, window, params, hint) -> Extension  (extend.rs:392)
  impl.band(read, window, params, band) -> (Extension, TouchRec)
        band = (offset, half) -> run_dispatch(Some((offset,1,1,half)))
        band = None           -> run_dispatch(None)       (extend.rs:293)
The expected values are the literal ones of the reference's tests
(file:line on each function).  The same functions run against the CPU
restatement (tests/test_oracle_kat.py) and the HIP path (tests/test_gpu_kat.py).
"""
from __future__ import annotations

from esperanto_amd.records import (DiagHint, Del, EditDir, ExtendParams, Ins, Match, OP_D, OP_I,
                                   SoftClip)
from seqgen import body, revcomp

DIRS = [EditDir.Off, EditDir.Ag, EditDir.Tc]


def _exempt_none(r: int, q: int) -> bool:
    return not ((r == ord("A") and q == ord("G")) or (r == ord("T") and q == ord("C")))


def kat_affine_gap_closed_form(impl):
    """extend.rs:2033-2128: score = 2*len - open - ext*n for a central indel."""
    reference = body(160)
    for insertion in (False, True):
        for n in (1, 2, 3, 7):
            query = bytearray(reference)
            if insertion:
                b = next(c for c in b"ACGT" if c != reference[79] and c != reference[80])
                query[80:80] = bytes([b]) * n
            else:
                del query[80:80 + n]
            query = bytes(query)
            for reverse in (False, True):
                q, r = (revcomp(query), revcomp(reference)) if reverse else (query, reference)
                for d in DIRS:
                    p = ExtendParams(edit_dir=d)
                    for offset in (0, 37):
                        window = b"N" * offset + r + b"N" * 11
                        for bnd in (None, (offset, 30)):
                            expected_match = len(r) if insertion else len(q)
                            expected = expected_match * p.match_score - p.gap_open - p.gap_ext * n
                            outs = [impl.band(q, window, p, bnd)[0],
                                    impl.extend_hint(q, window, p, DiagHint(offset))]
                            for ext in outs:
                                ctx = (insertion, n, reverse, d, offset, bnd, ext.cigar)
                                assert ext.score == expected, ctx
                                assert ext.read_start == 0, ctx
                                assert ext.read_end == len(q), ctx
                                assert ext.ref_start == offset, ctx
                                want = Ins(n) if insertion else Del(n)
                                assert want in ext.cigar, ctx
                                assert ext.cigar == outs[0].cigar, ctx


def kat_terminal_gap(impl):
    """extend.rs:2131-2162: terminal 1D3M costs exactly six match points."""
    b = body(120)
    reference = b + b"TAGC"
    query = b + b"AGC" + b"N" * 27
    for reverse in (False, True):
        q, r = (revcomp(query), revcomp(reference)) if reverse else (query, reference)
        ext = impl.extend(q, r, ExtendParams())
        assert ext.score == 240
        assert not any(c.op in (OP_I, OP_D) for c in ext.cigar), ext.cigar


def kat_terminal_gap_local_body(impl):
    """align.rs:2856-2875 at the extension level: the same body scores 240 as
    [M120, S30] on the forward strand."""
    b = body(120)
    ext = impl.extend(b + b"AGC" + b"N" * 27, b + b"TAGC", ExtendParams())
    assert ext.score == 240
    assert ext.ref_start == 0
    assert ext.cigar == [Match(120), SoftClip(30)], ext.cigar


def kat_rescue_penalties(impl):
    """extend.rs:2165-2180: 400/400 gap penalties keep an ungapped body."""
    reference = body(160)
    query = reference[:80] + reference[81:]
    ext = impl.extend(query, reference, ExtendParams(gap_open=400, gap_ext=400))
    assert ext.score > 0
    assert not any(c.op in (OP_I, OP_D) for c in ext.cigar), ext.cigar


CORE = b"GACGATCGGATACGCTAGCA"


def kat_neutral_outer_flanks(impl):
    """extend.rs:2183-2226"""
    for rflank, qflank in ((b"AAT", b"AAC"), (b"CCA", b"CCG")):
        for left in (True, False):
            reference, query = bytearray(b"NN"), bytearray()
            if left:
                reference += rflank
                query += b"NN" + qflank
            reference += CORE
            query += CORE
            if not left:
                reference += rflank[::-1]
                query += qflank[::-1] + b"NN"
            reference += b"NN"
            ext = impl.extend_hint(bytes(query), bytes(reference), ExtendParams(),
                                   DiagHint(0 if left else 2))
            assert ext.score == len(CORE) * 2
            assert ext.ref_start == 2
            want = [SoftClip(2), Match(23)] if left else [Match(23), SoftClip(2)]
            assert ext.cigar == want, (left, ext.cigar)


def kat_neutral_no_negative_clips(impl):
    """extend.rs:2229-2250"""
    reference = b"AATGACGATCGGATACGCTAGCA"
    for prefix in (b"NNN", b"ACC"):
        query = prefix + reference[3:]
        ext = impl.extend_hint(query, reference, ExtendParams(), DiagHint(0))
        assert ext.score == 40
        assert ext.read_start == 3
        assert ext.ref_start == 3
        assert ext.cigar == [SoftClip(3), Match(20)], ext.cigar


def kat_technical_tails(impl):
    """extend.rs:2253-2319 (uppercase and lowercase, all EditDirs)."""
    tails = [b"AGATCGGAAGAGCACACGTCTGAACTCCAGTCA", b"AAAAAAAAAAAAAAAAAAAA",
             b"CAGTCGTACGATCTAG"]
    for tail in tails:
        for left in (True, False):
            flank = bytearray(tail[-3:] if left else tail[:3])
            mm = 2 if left else 0
            probe = tail[-1] if left else tail[0]
            flank[mm] = next(r for r in b"ACGT" if r != flank[mm] and _exempt_none(r, probe))
            padding = len(tail) + 2
            reference, query = bytearray(b"N" * padding), bytearray()
            if left:
                reference += flank
                query += tail
            reference += CORE
            query += CORE
            if not left:
                reference += flank
                query += tail
            reference += b"N" * padding
            for d in DIRS:
                for lower in (False, True):
                    r, q = bytes(reference), bytes(query)
                    if lower:
                        r, q = r.lower(), q.lower()
                    ext = impl.extend(q, r, ExtendParams(edit_dir=d))
                    assert ext.score == 40
                    assert ext.ref_start == padding + (3 if left else 0)
                    want = ([SoftClip(len(tail)), Match(20)] if left
                            else [Match(20), SoftClip(len(tail))])
                    assert ext.cigar == want, (tail, left, d, lower, ext.cigar)


def kat_flank_scan_stops_at_mismatch(impl):
    """extend.rs:2322-2341"""
    ext = impl.extend(b"NNAAGAAC" + CORE, b"NNAATAAT" + CORE, ExtendParams())
    assert ext.score == 40
    assert ext.read_start == 5
    assert ext.ref_start == 5
    assert ext.cigar == [SoftClip(5), Match(23)], ext.cigar


def kat_terminal_edit_ties(impl):
    """extend.rs:2344-2390"""
    for rf, qf, direction in ((b"AAT", b"AAC", EditDir.Tc), (b"CCA", b"CCG", EditDir.Ag)):
        for left in (True, False):
            for d in DIRS:
                for lower in (False, True):
                    r, q = bytearray(b"NN"), bytearray(b"NN")
                    if left:
                        r += rf
                        q += qf
                    r += CORE
                    q += CORE
                    if not left:
                        r += rf[::-1]
                        q += qf[::-1]
                    r += b"NN"
                    q += b"NN"
                    r, q = bytes(r), bytes(q)
                    if lower:
                        r, q = r.lower(), q.lower()
                    ext = impl.extend(q, r, ExtendParams(edit_dir=d))
                    assert ext.score == 40 + (4 if d == direction else 0), (rf, left, d, lower)
                    assert ext.ref_start == 2
                    assert ext.cigar == [SoftClip(2), Match(23), SoftClip(2)], ext.cigar


def kat_identical_sequences(impl):
    """extend.rs:2393-2416"""
    reference = b"NNAATGACGATCGGATACGCTAGCA"
    same = b"NNAACGACGATCGGATACGCTAGCA"
    a = impl.extend(same, reference, ExtendParams())
    b = impl.extend(same, reference, ExtendParams())
    assert a.score == b.score and a.cigar == b.cigar
    assert a.cigar == [SoftClip(2), Match(23)], a.cigar


def kat_positive_ordinary_terminal(impl):
    """extend.rs:2419-2435"""
    ext = impl.extend(b"NNAAAA" + CORE, b"NNAAAC" + CORE, ExtendParams())
    assert ext.score == 42
    assert ext.cigar == [SoftClip(2), Match(24)], ext.cigar


def kat_full_matrix_short_reference(impl):
    """extend.rs:2542-2571: rf = (ACGT)*, read = N^(m-n) rf -> [S(m-n), M(n)], 2n."""
    for m, n in ((4, 1), (33, 7), (80, 15), (150, 55), (301, 31)):
        rf = bytes(b"ACGT"[i % 4] for i in range(n))
        read = b"N" * (m - n) + rf
        for ext in (impl.band(read, rf, ExtendParams(), None)[0], impl.extend(read, rf, ExtendParams())):
            assert ext.read_start == m - n
            assert ext.read_end == m
            assert ext.ref_start == 0
            assert ext.score == 2 * n
            assert ext.cigar == [SoftClip(m - n), Match(n)], ext.cigar


def kat_escape_hatch_inputs(impl):
    """extend.rs:2920-2959 child-process body: both dispatches still align."""
    read = b"ACGTACGTACGTACGTACGTACGTACGTACGT"
    win = b"TTTACGTACGTACGTACGTAAGGTACGTACGTACGTACGTAA"
    assert impl.extend(read, win, ExtendParams()).score > 0
    assert impl.extend_hint(read, win, ExtendParams(), DiagHint(3)).score > 0


def kat_empty_inputs(impl):
    """extend.rs:392-442: m == 0 or n == 0 -> soft-clip-only, score 0."""
    e = impl.extend_hint(b"", b"ACGT", ExtendParams(), DiagHint(0))
    assert e.score == 0 and e.cigar == [] and e.read_start == 0
    e = impl.extend_hint(b"ACGT", b"", ExtendParams(), DiagHint(0))
    assert e.score == 0 and e.cigar == [SoftClip(4)]
    e = impl.extend(b"ACGT", b"", ExtendParams())
    assert e.score == 0 and e.cigar == [SoftClip(4)]


ALL_EXTEND_KATS = [
    kat_affine_gap_closed_form, kat_terminal_gap, kat_terminal_gap_local_body,
    kat_rescue_penalties, kat_neutral_outer_flanks, kat_neutral_no_negative_clips,
    kat_technical_tails, kat_flank_scan_stops_at_mismatch, kat_terminal_edit_ties,
    kat_identical_sequences, kat_positive_ordinary_terminal, kat_full_matrix_short_reference,
    kat_escape_hatch_inputs, kat_empty_inputs,
]
