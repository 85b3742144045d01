"""Parity of the device fast lane (esp_fast_lane_batch) with the oracle's
try_fast_lane (align.rs:2274-2401), through the C ABI: status, Extension and
CIGAR of every chain on random chains over references with N runs and
soft-masked runs -- reads on their diagonal with substitutions, N, lower-case
bytes and A>G / T>C edits, reads overhanging either contig end, chains with an
anchor off the diagonal, an anchor gap above k + w + 5 or an uncovered end --
under all three EditDirs and custom scoring, plus empty anchor lists, empty
reads and an out-of-range contig; then 200,000 chains of 150-nt reads in one
launch against a sample."""
from __future__ import annotations

import sys

import numpy as np
import pytest

from esperanto_amd.records import ExtendParams
from fastgen import fast_reference, random_chain
from oracle import oracle

pytestmark = pytest.mark.gpu


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


def _check(o, recs, cigs, qs, cs, ans, params, k=15, w=5):
    for j in range(len(qs)):
        got = recs[j]
        if cs[j] >= 2:
            assert got["status"] == -1, j
            continue
        exp = oracle.fast_lane(o, qs[j], cs[j], ans[j], k=k, w=w, params=params)
        if exp is None:
            assert got["status"] == 0, (j, qs[j], cs[j], ans[j])
            continue
        ext, wstart = exp
        assert got["status"] == 1, (j, qs[j], cs[j], ans[j])
        assert (int(got["score"]), int(got["read_start"]), int(got["read_end"]), int(got["ref_start"]),
                int(got["wstart"])) == (ext.score, ext.read_start, ext.read_end, ext.ref_start, wstart), j
        assert int(got["pos"]) == wstart + ext.ref_start
        assert cigs[j][: int(got["n_ops"])].tolist() == [op.word for op in ext.cigar], j


@pytest.mark.parametrize("seed,params", [(1, ExtendParams()), (2, ExtendParams(edit_dir=1)),
                                         (3, ExtendParams(edit_dir=2)),
                                         (4, ExtendParams(match_score=1, mismatch=-3, gap_open=5, gap_ext=1))])
def test_fast_lane_matches_oracle(buf, seed, params):
    from esperanto_amd.fast_lane import fast_lane_batch
    from esperanto_amd.reference import DeviceReference
    ref, seqs = fast_reference(seed)
    o = oracle.OracleRef(ref)
    dref = DeviceReference(ref, buf)
    rng = np.random.default_rng(seed)
    qs, cs, ans = [], [], []
    for _ in range(3000):
        q, c, a = random_chain(rng, seqs)
        qs.append(q), cs.append(c), ans.append(a)
    qs += [b"ACGT" * 10, b"", b"ACGTACGT", b"ACGT" * 10]
    cs += [0, 1, 0, 7]
    ans += [[], [], [(0, 5)], [(0, 0), (25, 25)]]
    try:
        say(f"fast lane seed {seed}: {len(qs)} chains")
        recs, cigs = fast_lane_batch(dref, qs, cs, ans, params)
        _check(o, recs, cigs, qs, cs, ans, params)
        assert (recs["status"] == 1).sum() > 300 and (recs["status"] == 0).sum() > 300
        # other seed parameters move the gap and span conditions
        say("fast lane k 11 w 3")
        recs, cigs = fast_lane_batch(dref, qs, cs, ans, params, k=11, w=3)
        _check(o, recs, cigs, qs, cs, ans, params, k=11, w=3)
    finally:
        dref.close()


def test_fast_lane_reference_diagonal_cases(buf):
    from esperanto_amd.fast_lane import try_fast_lane
    from esperanto_amd.reference import DeviceReference, Reference
    seq = ("ACGTTGCAAGGCATTACGACCAAGTTACAGGATCCAAGTTGACGTTACAGGCATTAC" * 4).encode()
    dref = DeviceReference(Reference.from_sequences([seq]), buf)
    q = seq[40:120]
    good = [(0, 40), (20, 60), (40, 80), (65, 105)]
    try:
        ext, wstart = try_fast_lane(dref, q, 0, good)
        assert (ext.score, ext.read_start, ext.read_end, wstart, wstart + ext.ref_start) == (160, 0, 80, 0, 40)
        assert try_fast_lane(dref, q, 0, [(0, 40), (26, 66), (45, 85), (65, 105)]) is None
        mid = bytearray(q)
        mid[40] = ord("A") if mid[40] != ord("A") else ord("C")
        assert try_fast_lane(dref, bytes(mid), 0, good) is None
    finally:
        dref.close()


def test_fast_lane_large_batch(buf):
    """200,000 chains of 150-nt reads on their true diagonals (1% substitutions)
    in one launch; a sample of 3000 against the oracle, and the accepted ones
    place the read where it came from."""
    from esperanto_amd.fast_lane import fast_lane_batch
    from esperanto_amd.reference import DeviceReference, Reference
    rng = np.random.default_rng(11)
    L = 2_000_000
    genome = np.frombuffer(b"ACGT", dtype=np.uint8)[rng.integers(0, 4, L)]
    ref = Reference.from_sequences([genome.tobytes()])
    n, m = 200_000, 150
    loci = rng.integers(0, L - m, n)
    reads = genome[loci[:, None] + np.arange(m)[None, :]].copy()
    hit = rng.random(reads.shape) < 0.01
    reads[hit] = np.frombuffer(b"ACGT", dtype=np.uint8)[rng.integers(0, 4, int(hit.sum()))]
    qs = [reads[j].tobytes() for j in range(n)]
    ans = [[(qp, int(loci[j]) + qp) for qp in range(0, m - 15 + 1, 15)] for j in range(n)]
    cs = [0] * n
    dref = DeviceReference(ref, buf)
    try:
        say(f"fast lane {n} chains")
        recs, cigs = fast_lane_batch(dref, qs, cs, ans)
        acc = recs["status"] == 1
        assert acc.mean() > 0.3
        assert (recs["pos"][acc].astype(np.int64) - recs["read_start"][acc] == loci[acc]).all()
        o = oracle.OracleRef(ref)
        pick = rng.choice(n, 3000, replace=False)
        _check(o, recs[pick], cigs[pick], [qs[j] for j in pick], [0] * 3000, [ans[j] for j in pick],
               ExtendParams())
    finally:
        dref.close()
