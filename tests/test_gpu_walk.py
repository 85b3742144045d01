"""Parity of the device placement walks (esp_walk_placements) with the
oracle's rescore_with / recount_mm_ea_in / exact_kmer_support, through the
C ABI: every StrictScore field and count on random placements mixing
M/I/D/N/S over references with N runs and soft-masked runs, queries with
substitutions, N and lower-case bytes, positions past the contig end,
queries shorter than their CIGAR, all three EditDirs, known junctions from
an uploaded library (shuffled, with duplicates), custom scoring, k = 0 / 2 / 5 / 15 (k < 4 takes the
per-column k-mer path), empty queries and CIGARs, and an out-of-range
contig; then the reference's own known answers (align.rs:4216-4317) through
the device entry point."""
from __future__ import annotations

import sys

import numpy as np
import pytest

from esperanto_amd.records import ExtendParams, Ins, Match, RefSkip
from esperanto_amd.reference import Reference
from oracle import oracle
from walkgen import random_placement, random_walk_reference, words

pytestmark = pytest.mark.gpu

FIELDS = ("score", "aligned", "matches", "mm", "ea_ag", "ea_tc", "ins", "del", "clipped", "n_splice",
          "splice_score")


def say(msg):
    print(f"[gpu-test] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def buf():
    from esperanto_amd.extend import ExtendBuffer
    say("creating device context")
    b = ExtendBuffer(0)
    yield b
    b.close()


def _cases(seed: int, n: int):
    rng = np.random.default_rng(seed)
    ref, seqs = random_walk_reference(seed)
    qs, cs, ps, cigs, dirs = [], [], [], [], []
    for _ in range(n):
        q, c, p, cig = random_placement(rng, ref, seqs)
        qs.append(q), cs.append(c), ps.append(p), cigs.append(cig), dirs.append(int(rng.integers(0, 3)))
    # edge placements: empty query, empty CIGAR, past the contig end, a bad contig
    qs += [b"", b"ACGT", b"ACGTACGTAC", b"ACGT"]
    cs += [0, 1, 2, 7]
    ps += [5, 3, int(ref.lens[2]) - 4, 0]
    cigs += [words([Match(20)]), [], words([Match(10)]), words([Match(4)])]
    dirs += [0, 1, 2, 0]
    return ref, qs, cs, ps, cigs, dirs, rng


def _junctions(rng, cs, ps, cigs, frac=0.5):
    """(contig, start, end) of about frac of the RefSkips the placements cross."""
    out = set()
    for c, p, cig in zip(cs, ps, cigs):
        rf = p
        for w in cig:
            op, n = w & 15, w >> 4
            if op == 3 and rng.random() < frac:
                out.add((c, rf, rf + n))
            if op in (0, 2, 3):
                rf += n
    return sorted(out)


@pytest.mark.parametrize("seed,scoring", [(1, "default"), (2, "custom")])
def test_walk_matches_oracle(buf, seed, scoring):
    from esperanto_amd.reference import DeviceReference
    from esperanto_amd.walk import DeviceJunctions, Placements, WalkParams, walk_batch
    ref, qs, cs, ps, cigs, dirs, rng = _cases(seed, 3000)
    if scoring == "custom":
        ep = ExtendParams(match_score=1, mismatch=-3, gap_open=5, gap_ext=1)
        sp = (9, 5, 3, 2, 7)
    else:
        ep, sp = ExtendParams(), oracle.SPLICE_DEFAULTS
    o = oracle.OracleRef(ref)
    dref = DeviceReference(ref, buf)
    jl_triples = _junctions(rng, cs, ps, cigs)
    # the device set takes the triples in any order, duplicates included
    shuffled = [jl_triples[i] for i in rng.permutation(len(jl_triples))] + jl_triples[: len(jl_triples) // 3]
    jl = DeviceJunctions(shuffled, buf)
    pl = Placements.from_lists(qs, cs, ps, cigs, dirs)
    try:
        for k in (15, 5, 2, 0):
            wp = WalkParams(*sp, seed_k=k)
            say(f"walk seed {seed} {scoring} k {k}: {pl.n} placements, {len(jl_triples)} junctions")
            strict, counts, status = walk_batch(dref, pl, ep, wp, jl)
            for j in range(pl.n):
                if cs[j] >= ref.ncontig:
                    assert status[j] != 0 and strict[j]["score"] == 0 and counts[j]["kmers"] == 0
                    continue
                assert status[j] == 0, j
                if k == 15:
                    e = oracle.rescore(o, qs[j], cs[j], ps[j], cigs[j], dirs[j], ep, sp, jl_triples)
                    got = tuple(int(strict[j][f]) for f in FIELDS)
                    want = tuple(int(e[f]) for f in FIELDS)
                    assert got == want, (j, qs[j], cs[j], ps[j], cigs[j], dirs[j], got, want)
                    mmc = (int(counts[j]["mm"]), int(counts[j]["ea_ag"]), int(counts[j]["ea_tc"]))
                    assert mmc == (want[3], want[4], want[5]), j
                    if len(qs[j]) >= sum(w >> 4 for w in cigs[j] if (w & 15) in (0, 1, 4)):
                        assert mmc == oracle.recount(o, qs[j], cs[j], ps[j], cigs[j]), j
                kw = oracle.kmer_support(o, qs[j], cs[j], ps[j], cigs[j], dirs[j], k)
                assert int(counts[j]["kmers"]) == kw, (j, k, qs[j], cs[j], ps[j], cigs[j], dirs[j])
        # without a library every RefSkip is scored by its signal and length
        say("walk without junction library")
        strict, _, _ = walk_batch(dref, pl, ep, WalkParams(*sp))
        for j in range(0, pl.n, 7):
            if cs[j] < ref.ncontig:
                e = oracle.rescore(o, qs[j], cs[j], ps[j], cigs[j], dirs[j], ep, sp, None)
                assert int(strict[j]["score"]) == int(e["score"]), j
    finally:
        jl.close()
        dref.close()


def test_walk_reference_known_answers(buf):
    """align.rs:4216-4317 through the device walk."""
    from esperanto_amd.reference import DeviceReference
    from esperanto_amd.walk import exact_kmer_support, rescore_strict
    read = b"ACGT" * 12 + b"AC"
    dref = DeviceReference(Reference.from_sequences([b"ACGT" * 50]), buf)
    try:
        s = rescore_strict(dref, read, 0, 0, words([Match(50)]))
        assert (int(s["score"]), int(s["matches"]), int(s["mm"]), int(s["aligned"])) == (100, 50, 0, 50)
        r3 = bytearray(read)
        r3[24] = ord("G")
        s = rescore_strict(dref, bytes(r3), 0, 0, words([Match(50)]))
        assert (int(s["score"]), int(s["mm"]), int(s["ea_ag"] + s["ea_tc"])) == (94, 0, 1)
        ins_read = read[:24] + b"TT" + read[24:48]
        s = rescore_strict(dref, ins_read, 0, 0, words([Match(24), Ins(2), Match(24)]))
        assert (int(s["score"]), int(s["ins"])) == (88, 1)
        r = (b"ACGT" * 13)[:50]
        assert exact_kmer_support(dref, r, 0, 0, words([Match(50)])) == 3
        r2 = bytearray(r)
        r2[20] = ord("C") if r2[20] == ord("A") else ord("A")
        assert exact_kmer_support(dref, bytes(r2), 0, 0, words([Match(50)])) < 3
    finally:
        dref.close()
    exon1 = "ACGTTGCAAGGCATTACGACCAAGTTACAGGATCCAAGTTGACGTTACAGGCATTAC"
    intron = "GT" + "N" * 56 + "AG"
    exon2 = "CATGGACTTGAACCTTGACGTACATGGACTTGAACCTTGACGTA"
    seq = exon1 + intron + exon2
    start = len(exon1) - 20
    e2 = len(exon1) + len(intron)
    read = (seq[start:len(exon1)] + seq[e2:e2 + 20]).encode()
    dref = DeviceReference(Reference.from_sequences([seq.encode()]), buf)
    try:
        s = rescore_strict(dref, read, 0, start, words([Match(20), RefSkip(60), Match(20)]))
        assert (int(s["score"]), int(s["n_splice"]), int(s["aligned"])) == (40 * 2 + 4 - 9, 1, 40)
    finally:
        dref.close()


def test_walk_large_batch_properties(buf):
    """100k placements in one launch: the sampled records equal the oracle's,
    and the counts repeat the StrictScore's substitution counts."""
    from esperanto_amd.reference import DeviceReference
    from esperanto_amd.walk import Placements, walk_batch
    ref, qs, cs, ps, cigs, dirs, rng = _cases(9, 100_000)
    o = oracle.OracleRef(ref)
    dref = DeviceReference(ref, buf)
    pl = Placements.from_lists(qs, cs, ps, cigs, dirs)
    try:
        say(f"walk {pl.n} placements")
        strict, counts, status = walk_batch(dref, pl)
        ok = np.asarray(cs) < ref.ncontig
        assert (status[ok] == 0).all() and (status[~ok] != 0).all()
        for f in ("mm", "ea_ag", "ea_tc"):
            assert (counts[f] == strict[f]).all()
        for j in rng.choice(pl.n, 2000, replace=False).tolist():
            if ok[j]:
                e = oracle.rescore(o, qs[j], cs[j], ps[j], cigs[j], dirs[j])
                assert tuple(int(strict[j][f]) for f in FIELDS) == tuple(int(e[f]) for f in FIELDS), j
                assert int(counts[j]["kmers"]) == oracle.kmer_support(o, qs[j], cs[j], ps[j], cigs[j], dirs[j])
    finally:
        dref.close()
