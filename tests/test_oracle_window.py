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
