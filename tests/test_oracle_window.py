"""Pins the window restatements of the oracle (reference decode, mate-window
rescue, spliced window DP) before they check the device pipelines.

- decode: the 2-bit layout of fasta.rs:1-6 round-trips every ACGT byte and
  turns every other byte into N (Base::from_ascii, fasta.rs:30-39); the
  reverse-complement route equals an independent dictionary revcomp;
- pack_contig reproduces the byte layout on a hand-worked example;
- mate_window_placement: the reference's own known answers
  (align.rs:4694-4751) and planted reads whose true locus and strand are
  known, including the saturating window geometry at both contig ends;
- spliced window DP: the reference's own known answer (align.rs:5591-5639:
  pos 37, 20M60N20M) and planted junction reads, whose stitched CIGAR and
  genomic start are fixed by construction.
"""
from __future__ import annotations

import numpy as np
import pytest

from esperanto_amd.records import Match, RefSkip, SoftClip, ExtendParams
from esperanto_amd.reference import Reference, pack_contig
from esperanto_amd.rescue import MateRequests
from esperanto_amd.splice import SpliceRequests
from esperanto_amd.synth_window import splice_rescue_workload
from oracle import oracle
from seqgen import revcomp


@pytest.fixture(scope="module", autouse=True)
def _lib():
    oracle.lib()


def _upper_n(seq: bytes) -> bytes:
    return bytes(b if b in b"ACGT" else ord("N") for b in seq.upper())


def test_pack_layout_example():
    packed, iv = pack_contig(b"ACGTNacgT")
    # A=0 C=1 G=2 T=3 MSB-first; N codes 0 in the packed bytes; tail byte padded
    assert packed.tolist() == [0b00011011, 0b00000110, 0b11000000]
    assert iv.tolist() == [[4, 5]]
    packed, iv = pack_contig(b"NNAC" + b"RYN")
    assert iv.tolist() == [[0, 2], [4, 7]]
    assert packed.shape[0] == 2


def test_decode_round_trip_and_revcomp():
    rng = np.random.default_rng(7)
    alphabet = np.frombuffer(b"ACGTacgtNRYKM", dtype=np.uint8)
    seqs = []
    for L in (1, 3, 4, 5, 97, 1000, 4099):
        s = alphabet[rng.integers(0, 4, L)].copy()
        for _ in range(5):  # runs of other bytes
            a = int(rng.integers(0, L))
            s[a:a + int(rng.integers(1, 40))] = alphabet[rng.integers(4, 13)]
        seqs.append(s.tobytes())
    ref = Reference.from_sequences(seqs)
    o = oracle.OracleRef(ref)
    for k, s in enumerate(seqs):
        exp = _upper_n(s)
        assert oracle.decode(o, k, 0, len(s)) == exp
        for _ in range(40):
            a = int(rng.integers(0, len(s) + 1))
            b = int(rng.integers(a, len(s) + 1))
            assert oracle.decode(o, k, a, b) == exp[a:b]
            assert oracle.decode(o, k, a, b, rc=True) == revcomp(exp[a:b])
        assert oracle.decode(o, k, 0, len(s) + 1) is None
        assert oracle.decode(o, k, 2, 1) is None
    assert oracle.decode(o, len(seqs), 0, 0) is None


def test_revcomp_matches_dictionary_revcomp():
    for s in (b"", b"ACGT", b"acgtNRYkm", b"AAAACCCGGT" * 7):
        assert oracle.revcomp(s) == revcomp(s)


def test_mate_window_reference_known_answers():
    """align.rs:4694-4751 on a 1000 x A reference, mate at 100."""
    o = oracle.OracleRef(Reference.from_sequences([b"A" * 1000]))
    rec, cig = oracle.mate_window(o, b"A" * 98 + b"G" * 52, 0, 100)
    assert rec["placed"] == 1 and rec["dist"] == 0  # EA: read G matches ref A
    rec, cig = oracle.mate_window(o, b"A" * 150, 0, 100)
    assert rec["placed"] == 1 and cig == [Match(150)]
    rec, cig = oracle.mate_window(o, b"", 0, 100)
    assert rec["placed"] == 0 and cig == []


def _planted(rng, contig: bytes, m: int):
    P = int(rng.integers(0, len(contig) - m))
    seg = contig[P:P + m]
    minus = bool(rng.integers(0, 2))
    return P, minus, (revcomp(seg) if minus else seg)


def test_mate_window_planted_loci():
    rng = np.random.default_rng(11)
    contig = np.frombuffer(b"ACGT", dtype=np.uint8)[rng.integers(0, 4, 30000)].tobytes()
    o = oracle.OracleRef(Reference.from_sequences([contig]))
    for _ in range(60):
        m = int(rng.integers(30, 257))
        P, minus, read = _planted(rng, contig, m)
        mate = P + int(rng.integers(-700, 700))
        mate = max(0, mate)
        rec, cig = oracle.mate_window(o, read, 0, mate)
        assert rec["placed"] == 1
        assert bool(rec["strand"]) == minus
        assert rec["pos"] == P and rec["dist"] == 0
        assert cig == [Match(m)]
        span = 750 + m
        assert rec["wstart"] == max(0, mate - span)
        assert rec["wstart"] + rec["wlen"] == min(mate + span + m, len(contig))


def test_mate_window_guards():
    rng = np.random.default_rng(12)
    contig = np.frombuffer(b"ACGT", dtype=np.uint8)[rng.integers(0, 4, 5000)].tobytes()
    o = oracle.OracleRef(Reference.from_sequences([contig]))
    assert oracle.mate_window(o, contig[100:357], 0, 100)[0]["placed"] == 0   # m = 257 > 256
    assert oracle.mate_window(o, contig[100:256], 0, 100)[0]["placed"] == 1   # m = 156
    assert oracle.mate_window(o, contig[100:250], 1, 100)[0]["placed"] == -1  # no contig 1
    rec, _ = oracle.mate_window(o, contig[100:250], 0, 9000)  # window starts past the end
    assert rec["placed"] == 0 and rec["wlen"] == 0
    rec, _ = oracle.mate_window(o, b"N" * 150, 0, 100)  # N matches nothing: d = m > m/3
    assert rec["placed"] == 0 and rec["wlen"] > 0


def test_spliced_reference_known_answer():
    """align.rs:5587-5639: exon1 57 | GT..AG intron 60 | exon2 44, read 20+20."""
    exon1 = b"ACGTTGCAAGGCATTACGACCAAGTTACAGGATCCAAGTTGACGTTACAGGCATTAC"
    intron = b"GT" + b"N" * 56 + b"AG"
    exon2 = b"CATGGACTTGAACCTTGACGTACATGGACTTGAACCTTGACGTA"
    seq = exon1 + intron + exon2
    read = seq[37:57] + seq[117:137]
    o = oracle.OracleRef(Reference.from_sequences([seq]))
    # prepare_spliced: first exon from 37 - ALIGN_SPLICED_FLANK, junction 57..117,
    # last exon to the contig end (the 30-base flank is clipped at 161)
    rec, ext, st = oracle.spliced(o, read, 0, [(7, 57), (117, 161)], 0)
    assert rec["status"] == 1 and rec["genomic_start"] == 37
    assert st == [Match(20), RefSkip(60), Match(20)]


def test_spliced_planted_junctions():
    ref, sp, _, truth = splice_rescue_workload(64, ref_len=400_000, sub_rate=0.0)
    o = oracle.OracleRef(ref)
    rec, ext, cig, off, cig_n, st, soff = oracle.spliced_batch(o, sp)
    for k in range(sp.npairs):
        s, e, u = int(truth["s"][k]), int(truth["e"][k]), int(truth["u"][k])
        assert rec[k]["status"] == 1
        assert rec[k]["genomic_start"] == s - u
        words = st[int(soff[k]):int(soff[k]) + int(rec[k]["n_stitched"])].tolist()
        assert words == [Match(u).word, RefSkip(e - s).word, Match(150 - u).word]
        assert ext[k, 3] == 300


def test_spliced_rejections():
    rng = np.random.default_rng(5)
    contig = np.frombuffer(b"ACGT", dtype=np.uint8)[rng.integers(0, 4, 3000)].tobytes()
    o = oracle.OracleRef(Reference.from_sequences([contig]))
    read = contig[500:600]
    assert oracle.spliced(o, read, 0, [(400, 600), (700, 700)], 100)[0]["status"] == 0
    assert oracle.spliced(o, read, 0, [(400, 600), (2900, 3001)], 100)[0]["status"] == -1
    assert oracle.spliced(o, read, 0, [], 0)[0]["status"] == -1
    assert oracle.spliced(o, read, 1, [(400, 600)], 100)[0]["status"] == -1
    rec, ext, st = oracle.spliced(o, read, 0, [(400, 600)], 100)
    assert rec["status"] == 1 and rec["genomic_start"] == 500 and st == [Match(100)]


def test_workload_truth_consistency():
    ref, sp, mt, truth = splice_rescue_workload(32, ref_len=300_000, sub_rate=0.0)
    o = oracle.OracleRef(ref)
    for k in range(8):
        p2 = int(truth["p2"][k])
        r2 = mt.reads[int(mt.read_off[k]):int(mt.read_off[k]) + 150].tobytes()
        assert r2 == oracle.decode(o, 0, p2, p2 + 150, rc=True)
        rec, cig = oracle.mate_window(o, r2, 0, int(mt.mate_pos[k]))
        assert rec["placed"] == 1 and rec["pos"] == p2 and rec["strand"] == 1


def _body(n: int) -> bytes:
    """align.rs:2776-2786 sparse_window_body: xorshift64 from 0x75bcd15 over ACGT."""
    s, out, M = 0x75BCD15, bytearray(), (1 << 64) - 1
    for _ in range(n):
        s ^= (s << 13) & M
        s ^= s >> 7
        s ^= (s << 17) & M
        out.append(b"ACGT"[s & 3])
    return bytes(out)


def _sparse_chain(n: int, near_head: bool, pos: int, k: int = 15):
    """align.rs:2877-2897 sparse_window_chain as (ref_span, read_span) with k = 15."""
    qs = [2, 12, 27] if near_head else [n - 42, n - 32, n - 17]
    return (pos + qs[0], pos + qs[-1] + k), (qs[0], qs[-1] + k)


def test_continuous_extension_window_reference_properties():
    """align.rs:3116-3173 sparse_window_geometry_preserves_old_windows_and_bounds."""
    from esperanto_amd.unspliced import continuous_extension_window
    for n in (60, 150, 300, 1024):
        flank = n // 2 + 30
        for qs in range(0, n - 15 + 1, 7):
            for qe in range(qs + 15, n + 1, 11):
                for pos in (0, 1, 1000):
                    for extra in (0, 1000):
                        rsp, qsp = (pos + qs, pos + qe - 15 + 15), (qs, qe)
                        length = pos + n + extra
                        start, end = oracle.ext_window(rsp, qsp, n, length)
                        assert (start, end) == continuous_extension_window(rsp, qsp, n, length)
                        assert start <= pos and end >= pos + n
                        assert start <= end <= length
                        assert end - start <= n + 2 * flank
                        if qs <= flank and n - qe <= flank:
                            assert (start, end) == (max(rsp[0] - flank, 0),
                                                    min(rsp[1] + flank, length, rsp[0] + n + flank))
                        disp = (rsp[0], rsp[1] + 50000)  # displaced last anchor
                        a, b = oracle.ext_window(disp, qsp, n, length + 50000)
                        assert a <= b <= length + 50000 and b - a <= n + 2 * flank
    rsp, qsp = _sparse_chain(150, True, 0xFFFFFFFF - 100)
    rsp = (rsp[0], min(rsp[1], 0xFFFFFFFF))
    a, b = oracle.ext_window(rsp, qsp, 150, 0xFFFFFFFF)
    assert a <= b and b - a <= 360


def test_unspliced_reference_known_answers():
    """align.rs:2899-2972: the sparse-window chain keeps the whole ordinary body
    (EA lane for n <= 256, banded DP above), respects contig edges, and the EA
    lane places without touching the DP budget."""
    for n in (150, 300):
        body = _body(n)
        for left, right in ((1000, 1000), (0, 1000), (1000, 0)):
            o = oracle.OracleRef(Reference.from_sequences([b"N" * left + body + b"N" * right]))
            for near_head in (True, False):
                for minus in (False, True):
                    seq = revcomp(body) if minus else body
                    rsp, qsp = _sparse_chain(n, near_head, left)
                    rec, cig = oracle.unspliced(o, seq, 0, minus, rsp, qsp, True)
                    assert rec["status"] == (1 if n <= 256 else 2)
                    assert rec["pos"] == left and cig == [Match(n)]
                    rec, cig = oracle.unspliced(o, seq, 0, minus, rsp, qsp, False)
                    if n <= 256:
                        assert rec["status"] == 1 and cig == [Match(n)]
                    else:
                        assert rec["status"] == 0 and cig == []


def test_unspliced_ea_lane_soft_clips_and_fallback():
    """Planted reads: ends mismatching on purpose are soft-clipped by Kadane; an
    indel makes the recount differ from the distance and sends the read to the
    banded DP (or to nothing without budget); an N on the diagonal does too."""
    rng = np.random.default_rng(4)
    contig = np.frombuffer(b"ACGT", dtype=np.uint8)[rng.integers(0, 4, 20000)].tobytes()
    o = oracle.OracleRef(Reference.from_sequences([contig]))
    comp = {ord("A"): "C", ord("C"): "A", ord("G"): "T", ord("T"): "G"}
    for _ in range(40):
        m = int(rng.integers(60, 200))
        P = int(rng.integers(500, 19000 - m))
        read = bytearray(contig[P:P + m])
        for q in (2, 3):  # 2 + 2 - 4 - 4 < 0: Kadane clips the first four bases
            read[q] = ord(comp[read[q]])
        rsp, qsp = (P + 20, P + 60), (20, 60)
        rec, cig = oracle.unspliced(o, bytes(read), 0, False, rsp, qsp, True)
        ws, we = oracle.ext_window(rsp, qsp, m, len(contig))
        d, start = oracle.infix_best_start(bytes(read), contig[ws:we])
        if ws + start == P:  # the substitution reading is the rightmost optimal start
            assert rec["status"] == 1 and rec["dist"] == 2 == d
            assert rec["pos"] == P + 4 and cig == [SoftClip(4), Match(m - 4)]
        else:  # another optimal start: the recount disagrees and the DP decides
            assert rec["status"] == 2
        ind = bytes(read[:m // 2]) + b"T" + bytes(read[m // 2:])  # an insertion
        rec, _ = oracle.unspliced(o, ind, 0, False, rsp, qsp, True)
        assert rec["status"] in (0, 2)
        rec, _ = oracle.unspliced(o, ind, 0, False, rsp, qsp, False)
        assert rec["status"] in (0, 1)


def _run_plan_on_oracle(o, reads, jobs, mode):
    out = []
    for j in jobs:
        read = reads[int(j["read_off"])]
        pat = read[int(j["qstart"]):int(j["qend"])]
        text = oracle.decode(o, int(j["contig"]), int(j["tstart"]), int(j["tend"]))
        out.append(oracle.infix_best_end(pat, text) if mode == 0 else
                   oracle.infix_best_start(pat, text) if mode == 1 else (oracle.infix(pat, text), 0))
    return out


def test_intron_chain_refinements_on_planted_junctions():
    """intron_chain.rs:285-373 geometry with the oracle's Myers: on a read
    spanning a planted junction the donor search ends at the donor and the
    acceptor search starts at the acceptor (unique random sequence), and the
    split-read verification finds both halves at distance 0."""
    from esperanto_amd.intron_chain import acceptor_plan, donor_plan, verify_plan
    rng = np.random.default_rng(6)
    contig = np.frombuffer(b"ACGT", dtype=np.uint8)[rng.integers(0, 4, 50000)].tobytes()
    ref = Reference.from_sequences([contig])
    o = oracle.OracleRef(ref)
    reads, s_, e_, u_ = [], [], [], []
    for _ in range(50):
        m = int(rng.integers(80, 200))
        u = int(rng.integers(30, m - 30))
        s = int(rng.integers(1000, 20000))
        e = s + int(rng.integers(60, 20000))
        reads.append(contig[s - u:s] + contig[e:e + m - u])
        s_.append(s)
        e_.append(e)
        u_.append(u)
    n = len(reads)
    out, jobs, who, base = donor_plan(ref.lens, reads, [0] * n, u_, s_, [s - u for s, u in zip(s_, u_)])
    assert len(who) == n
    res = _run_plan_on_oracle(o, reads, jobs, 0)
    assert [base[j] + res[j][1] for j in range(n)] == s_
    out, jobs, who, base = acceptor_plan(ref.lens, reads, [0] * n, u_, e_, [e - u for e, u in zip(e_, u_)])
    res = _run_plan_on_oracle(o, reads, jobs, 1)
    assert [base[j] + res[j][1] for j in range(n)] == e_
    jobs, who = verify_plan(ref.lens, reads, [0] * n, [s - u for s, u in zip(s_, u_)],
                            [e + len(r) - u for r, e, u in zip(reads, e_, u_)], [x - 1 for x in s_], e_)
    res = _run_plan_on_oracle(o, reads, jobs, 2)
    for j in range(len(who)):
        k = who[j]
        mid = len(reads[k]) // 2
        # a half that crosses the junction is split by it; the other matches exactly
        if mid <= u_[k]:
            assert res[2 * j][0] == 0
        if mid >= u_[k]:
            assert res[2 * j + 1][0] == 0
