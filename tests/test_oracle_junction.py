"""The oracle's read-aware junction refinement, segment_chain, prepare_spliced
+ align_spliced and reextend_with, pinned by the reference's own tests and
by planted junctions:
  - read_boundary_competition_is_invariant_to_mate_orientation
    (crates/map/src/splice.rs:894-955), every combination of transcript
    strand, library annotation, read strand and shift;
  - try_spliced_stitches_two_segment_chain (crates/map/src/align.rs:5591-5639):
    genomic start 37, [M20, N60, M20], junction 57..117 GT-AG;
  - planted GT-AG / CT-AC introns: the refined junction is the planted one, the
    stitched CIGAR consumes the read and its RefSkip is the planted intron;
    short and non-canonical introns are refused without library support and
    accepted with it; one-segment chains are refused;
  - reextend_with: a placement whose soft clip hides editing mismatches is
    re-extended under the matching EditDir only, and an accepted extension is
    extend_hint of the stated window at the stated hint.
"""
from __future__ import annotations

import numpy as np
import pytest

from esperanto_amd.records import EditDir, ExtendParams, Match, RefSkip, SoftClip, cigar_from_words
from esperanto_amd.reference import Reference
from oracle import oracle
import splicegen


def _xorshift_genome(n=500, state=0x243F6A88):
    out = bytearray()
    M = (1 << 64) - 1
    for _ in range(n):
        state ^= (state << 13) & M
        state ^= state >> 7
        state ^= (state << 17) & M
        out.append(b"ACGT"[state & 3])
    return out


def test_refine_read_boundary_kat():
    """splice.rs:894-955."""
    for transcript_minus in (False, True):
        genome = _xorshift_genome()
        donor, acceptor = (b"CT", b"AC") if transcript_minus else (b"GT", b"AG")
        genome[120:122] = donor
        genome[298:300] = acceptor
        genome[300:302] = donor
        genome[119] = ord("T")
        query = bytearray(genome[60:120])
        query[59] = ord("C")
        query += genome[300:379]
        ref = Reference.from_sequences([bytes(genome)])
        o = oracle.OracleRef(ref)
        for annotated in (False, True):
            lib = oracle.JLib([(0, 120, 300, int(transcript_minus))] if annotated else [])
            for read_minus in (False, True):
                for shift in (1, 2):
                    got = oracle.refine_junction_read(o, bytes(query), 0, 120 + shift, 300 + shift, 60 + shift,
                                                      60 + shift, read_minus, lib)
                    assert got == (120, 300, transcript_minus, "GtAg", int(annotated)), (transcript_minus, annotated,
                                                                                         read_minus, shift, got)
            lib.close()


def spliced_chain_fixture():
    """align.rs:5591-5620."""
    exon1 = b"ACGTTGCAAGGCATTACGACCAAGTTACAGGATCCAAGTTGACGTTACAGGCATTAC"
    intron = b"GT" + b"N" * 56 + b"AG"
    exon2 = b"CATGGACTTGAACCTTGACGTACATGGACTTGAACCTTGACGTA"
    seq = exon1 + intron + exon2
    start = len(exon1) - 20
    e2 = len(exon1) + len(intron)
    read = seq[start:len(exon1)] + seq[e2:e2 + 20]
    return seq, read, [(0, start), (20, e2)]


def test_try_spliced_kat():
    """align.rs:5624-5638 through align_spliced with k = 15 (SeedParams of the
    test's aligner) and the default scoring."""
    seq, read, anchors = spliced_chain_fixture()
    assert len(seq) == 57 + 60 + 44
    o = oracle.OracleRef(Reference.from_sequences([seq]))
    r = oracle.align_spliced(o, None, read, 0, False, anchors, 15)
    assert r["status"] == 1
    assert r["genomic_start"] == 37
    assert r["stitched"] == [Match(20), RefSkip(60), Match(20)]
    assert r["junctions"] == [(57, 117, False, "GtAg", 0)]
    assert oracle.segment_chain(anchors, 15) == [(0, 15, 37, 52), (20, 35, 117, 132)]


def test_segment_chain_breaks():
    """splice.rs:364-396: a break needs ref gap > min_intron and read gap <= 10."""
    k = 15
    assert oracle.segment_chain([], k) == []
    assert oracle.segment_chain([(0, 100)], k) == [(0, 15, 100, 115)]
    # ref gap 21 > 20, read gap 0: break
    assert oracle.segment_chain([(0, 100), (15, 136)], k) == [(0, 15, 100, 115), (15, 30, 136, 151)]
    # ref gap 20: no break
    assert oracle.segment_chain([(0, 100), (15, 135)], k) == [(0, 30, 100, 150)]
    # read gap 11: no break
    assert oracle.segment_chain([(0, 100), (26, 500)], k) == [(0, 41, 100, 515)]
    # overlapping anchors saturate the gaps at 0
    assert oracle.segment_chain([(0, 100), (3, 103)], k) == [(0, 18, 100, 118)]


def _run_workload(seed, n):
    ref, chains, quads, truth = splicegen.spliced_workload(seed, n)
    o = oracle.OracleRef(ref)
    lib = oracle.JLib(quads)
    res = [oracle.align_spliced(o, lib, r, c, m, a, splicegen.K) for r, c, m, a in chains]
    return ref, chains, quads, truth, lib, res


def test_planted_junctions():
    ref, chains, quads, truth, lib, res = _run_workload(0x5E1, 240)
    hit = tot = 0
    annotated = {(q[1], q[2], q[3]) for q in quads}
    for (read, c, minus, anchors), (s, e, kind), r in zip(chains, truth, res):
        if kind == "onesegment":
            assert r["status"] == 0 and r["refusal"] == 1
            continue
        known = (s, e, int(minus)) in annotated
        if kind in ("short", "noncanon") and not known and r["n_segments"] >= 2:
            # splice support guard (splice.rs:107-115) unless the refinement
            # picked a neighbouring boundary
            if r["status"] == 1:
                assert any(j[4] > 0 for j in r["junctions"]) or all(
                    (j[1] - j[0]) >= 50 and j[3] != "NonCanonical" for j in r["junctions"])
            continue
        if r["status"] != 1:
            continue
        # the stitched CIGAR consumes the read and carries the refined introns
        q = sum(op.len for op in r["stitched"] if op.op in (0, 1, 4))
        assert q == len(read)
        skips = [op.len for op in r["stitched"] if op.op == 3]
        assert all(j[1] - j[0] in skips or not skips for j in r["junctions"])
        if kind == "canon":
            tot += 1
            hit += r["junctions"][0][:2] == (s, e)
    assert tot > 40 and hit >= 0.9 * tot, (hit, tot)
    lib.close()


def test_short_intron_needs_support():
    """A 40-base GT-AG intron is refused (guard 4) without library support and
    stitched with it."""
    rng = np.random.default_rng(3)
    g = bytearray(np.frombuffer(b"ACGT", dtype=np.uint8)[rng.integers(0, 4, 2000)].tobytes())
    s, e = 700, 740
    g[s:s + 2] = b"GT"
    g[e - 2:e] = b"AG"
    read = bytes(g[s - 60:s] + g[e:e + 60])
    anchors = [(0, s - 60), (15, s - 45), (30, s - 30), (45, s - 15), (60, e), (75, e + 15), (90, e + 30)]
    o = oracle.OracleRef(Reference.from_sequences([bytes(g)]))
    r = oracle.align_spliced(o, None, read, 0, False, anchors, 15)
    assert r["status"] == 0 and r["refusal"] == 4
    lib = oracle.JLib([(0, s, e, 0)])
    r = oracle.align_spliced(o, lib, read, 0, False, anchors, 15)
    assert r["status"] == 1 and r["junctions"][0] == (s, e, False, "GtAg", 1)
    assert r["stitched"] == [Match(60), RefSkip(40), Match(60)]
    assert r["genomic_start"] == s - 60


def test_refine_queries_are_consistent():
    """Over shifted naive breakpoints: a found junction keeps the intron length
    ref_gap - read_gap and lies within the shift window; read gaps above 20
    find nothing."""
    ref, chains, quads, truth = splicegen.spliced_workload(0x5E2, 64)
    o = oracle.OracleRef(ref)
    lib = oracle.JLib(quads)
    found = 0
    for q in splicegen.refine_queries(7, chains, truth, 400):
        c, contig, ns, ne, qa, qb, minus = q[:7]
        got = oracle.refine_junction_read(o, chains[c][0], contig, ns, ne, qa, qb, minus, lib)
        gap = qb - qa
        if contig != 0 or gap > 20:
            assert got is None
            continue
        if got is None:
            continue
        found += 1
        s, e = got[0], got[1]
        assert e - s == (ne - ns) - gap
        assert min(gap, 0) - 2 <= s - ns <= max(gap, 0) + 2
        assert got[4] == lib.support(0, s, e, int(got[2]))
    assert found > 100
    lib.close()


# ---- reextend_with --------------------------------------------------------------

def _reext_case(g, pos, lead, span, tail, edits):
    read = bytearray(g[pos - lead:pos + span + tail])
    for i in edits:
        read[i] = ord("G") if read[i] == ord("A") else read[i]
    return bytes(read), [SoftClip(lead), Match(span), SoftClip(tail)] if lead and tail else \
        ([SoftClip(lead), Match(span)] if lead else [Match(span), SoftClip(tail)])


def test_reextend_under_the_matching_edit_dir():
    rng = np.random.default_rng(11)
    g = bytearray(np.frombuffer(b"ACGT", dtype=np.uint8)[rng.integers(0, 4, 3000)].tobytes())
    pos, lead, span = 1000, 12, 100
    # plant A's in the clipped lead and turn them into G's in the read
    for i in range(pos - 12, pos, 2):
        g[i] = ord("A")
    read, cig = _reext_case(g, pos, lead, span, 0, [0, 2, 4, 6, 8, 10])
    o = oracle.OracleRef(Reference.from_sequences([bytes(g)]))
    words = [c.word for c in cig]
    p = ExtendParams()
    ag = oracle.reextend_with(o, read, 0, pos, words, p, int(EditDir.Ag))
    assert ag["status"] == 1 and ag["pos"] == pos - lead
    assert cigar_from_words(ag["cigar"]) == [Match(lead + span)]
    # without the exemption the six G-over-A columns cost more than they gain:
    # the re-extension keeps a leading clip (and is accepted, aligning as much)
    off = oracle.reextend_with(o, read, 0, pos, words, p, int(EditDir.Off))
    assert off["pos"] > pos - lead and cigar_from_words(off["cigar"])[0].op == 4
    assert off["new_aligned"] < lead + span
    # the accepted extension is extend_hint of the stated window at the stated hint
    win = oracle.decode(o, 0, ag["wstart"], ag["wstart"] + ag["wlen"])
    e = oracle.extend_hint(read, win, ExtendParams(edit_dir=EditDir.Ag), oracle.DiagHint(pos - ag["wstart"] - lead))
    assert e.score == ag["score"] and [c.word for c in e.cigar] == ag["cigar"]
    assert ag["wstart"] == pos - lead - 30 and ag["wlen"] == lead + span + 60
    # spliced placements are left alone; a contig out of range is refused
    assert oracle.reextend_with(o, read, 0, pos, [Match(50).word, RefSkip(100).word, Match(62).word], p, 1)["status"] == 2
    assert oracle.reextend_with(o, read, 3, pos, words, p, 1)["status"] == -1


# ---- try_intron_chain_placement ----------------------------------------------------

def test_intron_chain_known_answer():
    """intron_fast_path_reconstructs_asymmetric_exons_with_technical_clips
    (crates/map/src/align.rs:5067-5120): pos 1000, [S11, M60, N1940, M79], score 278,
    for both motif orientations and both anchor strands."""
    genome = _xorshift_genome(5000, 0x517CC1B7)
    for donor, acceptor in ((b"GT", b"AG"), (b"CT", b"AC")):
        genome[1060:1062] = donor
        genome[2998:3000] = acceptor
        genome[1059] = ord("T")
        genome[3000] = ord("A")
        query = b"N" * 11 + bytes(genome[1000:1060]) + bytes(genome[3000:3079])
        o = oracle.OracleRef(Reference.from_sequences([bytes(genome)]))
        for strand in (0, 1):
            anchors = [(q + off, r + off, 0, strand) for q, r, span in ((11, 1000, 60), (71, 3000, 79))
                       for off in range(3, span - 14, 4)]
            h = oracle.intron_chain(o, None, query, anchors)
            assert h["status"] == 1, h
            assert h["genomic_start"] == 1000
            assert h["stitched"] == [SoftClip(11), Match(60), RefSkip(1940), Match(79)]
            assert h["score"] == 278
            assert (h["jstart"], h["jend"]) == (1060, 3000)


def test_intron_chain_places_planted_junctions():
    """Balanced two-exon reads land on their planted intron with a read-conserving
    [.., N(intron), ..] CIGAR; reads with one dominant fragment go to the
    unspliced paths (reason 3)."""
    ref, reads, pools, quads = splicegen.intron_chain_workload(5, 240)
    o = oracle.OracleRef(ref)
    jl = oracle.JLib(quads)
    placed = hits = 0
    for c, (r, p) in enumerate(zip(reads, pools)):
        h = oracle.intron_chain(o, jl, r, p)
        kind = c % 6
        if kind == 5:
            assert h["status"] == 0 and h["reason"] == 3
        if h["status"] != 1:
            continue
        placed += 1
        assert sum(op.len for op in h["stitched"] if op.op in (0, 1, 4)) == len(r)
        assert any(op.op == 3 and op.len == h["jend"] - h["jstart"] for op in h["stitched"])
        # a substitution near the seam can stop an exon's MEM early; when the
        # read gap then exceeds 20 the pair's own naive junction is kept
        hits += h["jstart"] == 2000 + c * 9000 + 700
    assert placed > 120 and hits >= 0.85 * placed, (placed, hits)
    jl.close()
