"""Parity of the device window pipelines with the oracle, through the C ABI.

- window gather (esp_gather_batch) against Contig::decode_into / revcomp_into
  over multi-contig references with N runs, including empty windows, windows
  ending at a contig end and out-of-range jobs;
- mate-window rescue (esp_mate_rescue_batch) against mate_window_placement:
  every record field and CIGAR word, on planted Plus/Minus reads, reads that
  place nowhere, the m = 0 / m > 256 guards, mate positions at and past both
  contig ends, lowercase and IUPAC read bytes, an out-of-range contig index;
- spliced window DP (esp_spliced_batch) against the oracle's pseudo-reference
  extend_hint + stitch_cigar: status, extension, both CIGARs, genomic start,
  on the reference's known answer, planted junction reads with substitutions,
  three-exon requests, random hints and rejected / invalid exon lists.
"""
from __future__ import annotations

import sys

import numpy as np
import pytest

from esperanto_amd.records import ExtendParams, Match, RefSkip
from esperanto_amd.reference import Reference
from esperanto_amd.rescue import MateRequests
from esperanto_amd.splice import SpliceRequests
from esperanto_amd.synth_window import splice_rescue_workload
from oracle import oracle
from seqgen import revcomp

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


def _ref_with_runs(seed: int, lens, runs=6):
    return Reference.synthetic(seed, lens, n_runs=runs, n_run_max=150)


def test_gather_matches_decode(buf):
    from esperanto_amd.reference import DeviceReference
    ref = _ref_with_runs(21, [5000, 1, 7, 20000, 333])
    o = oracle.OracleRef(ref)
    dref = DeviceReference(ref, buf)
    rng = np.random.default_rng(3)
    jobs = []
    for k, L in enumerate(ref.lens.tolist()):
        jobs += [(k, 0, L, False), (k, 0, L, True), (k, L, L, False), (k, max(0, L - 5), L, True)]
        for _ in range(60):
            a = int(rng.integers(0, L + 1))
            b = int(rng.integers(a, min(L, a + 3000) + 1))
            jobs.append((k, a, b, bool(rng.integers(0, 2))))
    bad = [(5, 0, 10, False), (0, 10, 5, False), (0, 4990, 5001, True)]
    jobs += bad
    say(f"gather {len(jobs)} jobs")
    out, status = dref.gather(jobs)
    for j, (c, a, b, rc) in enumerate(jobs):
        exp = oracle.decode(o, c, a, b, rc)
        if exp is None:
            assert status[j] == 1, j
        else:
            assert status[j] == 0 and out[j] == exp, (j, c, a, b, rc)
    assert dref.device_bytes() > 0
    dref.close()


def _mate_cases(rng, ref: Reference, contig_seqs, n: int):
    reads, contigs, poss = [], [], []
    nc = len(contig_seqs)
    for _ in range(n):
        kind = int(rng.integers(0, 10))
        c = int(rng.integers(0, nc))
        seq = contig_seqs[c]
        L = len(seq)
        m = int(rng.integers(20, 257))
        if kind <= 5 and L > m + 1:  # planted, with substitutions and an indel now and then
            P = int(rng.integers(0, L - m))
            r = bytearray(seq[P:P + m])
            for _ in range(int(rng.integers(0, 6))):
                r[int(rng.integers(0, m))] = b"ACGT"[int(rng.integers(0, 4))]
            if rng.integers(0, 4) == 0:
                q = int(rng.integers(1, m - 1))
                del r[q]
            r = bytes(r)
            if rng.integers(0, 2):
                r = revcomp(r)
            if rng.integers(0, 8) == 0:
                r = r.lower()
            pos = max(0, P + int(rng.integers(-900, 900)))
        elif kind == 6:  # random read: mostly no placement
            r = np.frombuffer(b"ACGTNRY", dtype=np.uint8)[rng.integers(0, 7, m)].tobytes()
            pos = int(rng.integers(0, L + 10))
        elif kind == 7:  # guards
            r = [b"", b"A" * 257, b"ACGT" * 80][int(rng.integers(0, 3))]
            pos = int(rng.integers(0, L + 1))
        elif kind == 8:  # at / past the contig ends
            r = seq[max(0, L - m):L] if L >= m else seq
            pos = [0, L, L + 5000, 0xFFFFFFF0][int(rng.integers(0, 4))]
        else:
            P = int(rng.integers(0, max(1, L - m)))
            r = seq[P:P + m]
            pos = P
        reads.append(r)
        contigs.append(c)
        poss.append(pos)
    return reads, contigs, poss


def _compare_mate(req, g, c):
    grec, gcig, goff, gn = g
    crec, ccig, coff, cn = c
    assert np.array_equal(goff, coff)
    for k in range(req.npairs):
        assert grec[k].tolist() == crec[k].tolist(), (k, grec[k], crec[k])
        assert gn[k] == cn[k], k
        a, b = int(goff[k]), int(goff[k]) + int(gn[k])
        assert np.array_equal(gcig[a:b], ccig[a:b]), k


def test_mate_rescue_matches_oracle(buf):
    from esperanto_amd.reference import DeviceReference
    from esperanto_amd.rescue import DeviceMateBatch
    ref = _ref_with_runs(31, [40000, 3000, 900])
    seqs = [oracle.decode(oracle.OracleRef(ref), k, 0, int(L)) for k, L in enumerate(ref.lens)]
    rng = np.random.default_rng(8)
    reads, contigs, poss = _mate_cases(rng, ref, seqs, 3000)
    req = MateRequests.from_lists(reads, contigs, poss)
    o = oracle.OracleRef(ref)
    for params in (ExtendParams(), ExtendParams(2, -4, 400, 400)):
        dref = DeviceReference(ref, buf)
        db = DeviceMateBatch(req, dref, buf)
        say(f"mate rescue {req.npairs} requests params={params}")
        db.launch(params)
        db.sync()
        g = db.raw_results()
        c = oracle.mate_batch(o, req, params, nthreads=8)
        _compare_mate(req, g, c)
        placed = int((g[0]["placed"] == 1).sum())
        say(f"placed {placed}")
        assert placed > req.npairs // 3
        db.free()
        dref.close()


def test_mate_rescue_invalid_contig_and_known_answers(buf):
    from esperanto_amd.reference import DeviceReference
    from esperanto_amd.rescue import mate_window_placement, DeviceMateBatch
    ref = Reference.from_sequences([b"A" * 1000])
    dref = DeviceReference(ref, buf)
    p = mate_window_placement(dref, b"A" * 98 + b"G" * 52, 0, 100)
    assert p is not None and p.dist == 0
    p = mate_window_placement(dref, b"A" * 150, 0, 100)
    assert p is not None and p.cigar == [Match(150)]
    assert mate_window_placement(dref, b"", 0, 100) is None
    req = MateRequests.from_lists([b"A" * 150], [3], [100])
    db = DeviceMateBatch(req, dref, buf)
    db.launch()
    db.sync()
    assert db.raw_results()[0][0]["placed"] == -1
    with pytest.raises(ValueError):
        db.placements()
    db.free()
    dref.close()


def _splice_cases(rng, seqs, n: int):
    reads, contigs, exl, hints = [], [], [], []
    for _ in range(n):
        kind = int(rng.integers(0, 8))
        c = int(rng.integers(0, len(seqs)))
        seq = seqs[c]
        L = len(seq)
        if kind <= 4 and L > 3000:
            nex = 2 if kind <= 3 else 3
            cuts = np.sort(rng.choice(np.arange(400, L - 400), size=2 * nex - 2, replace=False))
            starts = [int(rng.integers(100, 400))] if False else []
            exons, segs = [], []
            lo = int(cuts[0]) - int(rng.integers(40, 300))
            pts = [lo] + cuts.tolist() + [min(L, int(cuts[-1]) + int(rng.integers(40, 300)))]
            for q in range(nex):
                exons.append((int(pts[2 * q]), int(pts[2 * q + 1])))
            pseudo = b"".join(seq[a:b] for a, b in exons)
            m = int(rng.integers(30, min(250, len(pseudo)) + 1))
            off = int(rng.integers(0, len(pseudo) - m + 1))
            r = bytearray(pseudo[off:off + m])
            for _ in range(int(rng.integers(0, 5))):
                r[int(rng.integers(0, m))] = b"ACGT"[int(rng.integers(0, 4))]
            reads.append(bytes(r))
            hints.append(off + int(rng.integers(-3, 4)) if rng.integers(0, 3) else int(rng.integers(-50, len(pseudo))))
            exl.append(exons)
        elif kind == 5:  # rejected: an exon with re <= rs
            reads.append(seq[10:80])
            exl.append([(0, 100), (500, 500)])
            hints.append(10)
        elif kind == 6:  # invalid: past the contig end / no exons
            reads.append(seq[10:80])
            exl.append([(0, 100), (L - 10, L + 1)] if rng.integers(0, 2) else [])
            hints.append(0)
        else:  # one exon, plain window
            a = int(rng.integers(0, max(1, L - 400)))
            b = min(L, a + 400)
            m = int(rng.integers(1, min(200, b - a) + 1))
            reads.append(seq[a + 50:a + 50 + m] if b - a > 50 + m else seq[a:a + m])
            exl.append([(a, b)])
            hints.append(50)
        contigs.append(c)
    return reads, contigs, exl, hints


def _compare_splice(req, g, c):
    grec, gext, gcig, goff, gn, gst, gsoff = g
    crec, cext, ccig, coff, cn, cst, csoff = c
    for k in range(req.npairs):
        assert grec[k]["status"] == crec[k]["status"], (k, grec[k], crec[k])
        if crec[k]["status"] != 1:
            continue
        assert grec[k].tolist() == crec[k].tolist(), (k, grec[k], crec[k])
        assert gext[k].tolist() == cext[k].tolist(), k
        assert gn[k] == cn[k]
        a = int(goff[k])
        assert np.array_equal(gcig[a:a + int(gn[k])], ccig[a:a + int(cn[k])]), k
        a = int(gsoff[k])
        ns = int(crec[k]["n_stitched"])
        assert np.array_equal(gst[a:a + ns], cst[a:a + ns]), k


def test_spliced_matches_oracle(buf):
    from esperanto_amd.reference import DeviceReference
    from esperanto_amd.splice import DeviceSpliceBatch
    ref = _ref_with_runs(41, [60000, 5000, 700])
    o = oracle.OracleRef(ref)
    seqs = [oracle.decode(o, k, 0, int(L)) for k, L in enumerate(ref.lens)]
    rng = np.random.default_rng(9)
    reads, contigs, exl, hints = _splice_cases(rng, seqs, 2000)
    req = SpliceRequests.from_lists(reads, contigs, exl, hints)
    dref = DeviceReference(ref, buf)
    for params in (ExtendParams(), ExtendParams(edit_dir=1)):
        db = DeviceSpliceBatch(req, dref, buf)
        say(f"spliced {req.npairs} requests params={params}")
        db.launch(params)
        db.sync()
        g = db.raw_results()
        c = oracle.spliced_batch(o, req, params, nthreads=8)
        _compare_splice(req, g, c)
        db.free()
    dref.close()


def test_spliced_known_answer_and_workload(buf):
    from esperanto_amd.reference import DeviceReference
    from esperanto_amd.splice import spliced_batch, DeviceSpliceBatch
    exon1 = b"ACGTTGCAAGGCATTACGACCAAGTTACAGGATCCAAGTTGACGTTACAGGCATTAC"
    seq = exon1 + b"GT" + b"N" * 56 + b"AG" + b"CATGGACTTGAACCTTGACGTACATGGACTTGAACCTTGACGTA"
    dref = DeviceReference(Reference.from_sequences([seq]), buf)
    (aln,) = spliced_batch(dref, [seq[37:57] + seq[117:137]], [0], [[(7, 57), (117, 161)]], [0])
    assert aln.genomic_start == 37 and aln.cigar == [Match(20), RefSkip(60), Match(20)]
    dref.close()
    ref, sp, mt, truth = splice_rescue_workload(4000, ref_len=2_000_000)
    o = oracle.OracleRef(ref)
    dref = DeviceReference(ref, buf)
    expect = oracle.spliced_batch(o, sp, nthreads=8)
    from esperanto_amd import _native
    for kernel in (0, 1):  # auto (wave per pair at this size), then lane per pair
        buf.set_option(_native.ESP_OPT_FULL_KERNEL, kernel)
        db = DeviceSpliceBatch(sp, dref, buf)
        say(f"splice/rescue workload 4000 pairs: spliced, full kernel {kernel}")
        db.launch()
        db.sync()
        _compare_splice(sp, db.raw_results(), expect)
        db.free()
    buf.set_option(_native.ESP_OPT_FULL_KERNEL, 0)
    from esperanto_amd.rescue import DeviceMateBatch
    mb = DeviceMateBatch(mt, dref, buf)
    say("splice/rescue workload 4000 pairs: mate rescue")
    mb.launch()
    mb.sync()
    _compare_mate(mt, mb.raw_results(), oracle.mate_batch(o, mt, nthreads=8))
    mb.free()
    dref.close()


def _chain_cases(rng, seqs, n: int):
    reads, contigs, minus, rsp, qsp, dp = [], [], [], [], [], []
    for _ in range(n):
        c = int(rng.integers(0, len(seqs)))
        seq = seqs[c]
        L = len(seq)
        m = int(rng.choice([int(rng.integers(1, 257)), int(rng.integers(257, 400)), 0]
                           if rng.integers(0, 10) == 0 else [int(rng.integers(20, 257))]))
        m = min(m, max(L - 1, 0))
        P = int(rng.integers(0, max(1, L - m)))
        r = bytearray(seq[P:P + m])
        for _ in range(int(rng.integers(0, 5))):
            if m:
                r[int(rng.integers(0, m))] = b"ACGTN"[int(rng.integers(0, 5))]
        if m > 10 and rng.integers(0, 5) == 0:
            q = int(rng.integers(1, m - 1))
            r = r[:q] + (b"T" * int(rng.integers(1, 4)) if rng.integers(0, 2) else b"") + r[q + 2:]
        r = bytes(r)
        mi = bool(rng.integers(0, 2))
        if mi:
            r = revcomp(r)
        qs = int(rng.integers(0, max(1, len(r) - 15))) if len(r) > 15 else 0
        qe = min(len(r), qs + int(rng.integers(15, 80)))
        rs = max(0, P + qs + int(rng.integers(-3, 4)))
        reads.append(r)
        contigs.append(c)
        minus.append(int(mi))
        rsp.append((rs, rs + (qe - qs)))
        qsp.append((qs, qe))
        dp.append(int(rng.integers(0, 4) != 0))
    return reads, contigs, minus, rsp, qsp, dp


def test_unspliced_matches_oracle(buf):
    from esperanto_amd.reference import DeviceReference
    from esperanto_amd.unspliced import ChainRequests, DeviceChainBatch
    ref = _ref_with_runs(51, [30000, 2000, 600])
    o = oracle.OracleRef(ref)
    seqs = [oracle.decode(o, k, 0, int(L)) for k, L in enumerate(ref.lens)]
    rng = np.random.default_rng(10)
    reads, contigs, minus, rsp, qsp, dp = _chain_cases(rng, seqs, 2500)
    contigs[7] = 9  # an invalid contig index
    req = ChainRequests.from_lists(reads, contigs, minus, rsp, qsp, dp)
    dref = DeviceReference(ref, buf)
    for params in (ExtendParams(), ExtendParams(edit_dir=2)):
        db = DeviceChainBatch(req, dref, buf)
        say(f"unspliced {req.npairs} chains params={params}")
        db.launch(params)
        db.sync()
        rec, cig, off, cig_n = db.raw_results()
        crec, ccig = oracle.unspliced_batch(o, req, params)
        lanes = [0, 0, 0]
        for k in range(req.npairs):
            assert rec[k].tolist() == crec[k].tolist(), (k, rec[k], crec[k])
            words = cig[int(off[k]):int(off[k]) + int(cig_n[k])].tolist()
            assert words == [c.word for c in ccig[k]], k
            lanes[max(int(rec[k]["status"]), 0)] += 1
        say(f"none/EA/DP = {lanes}")
        assert lanes[1] > 200 and lanes[2] > 50
        db.free()
    dref.close()


def test_unspliced_reference_known_answers(buf):
    """align.rs:2899-2972 on the device (see tests/test_oracle_window.py)."""
    from esperanto_amd.reference import DeviceReference
    from esperanto_amd.unspliced import ChainRequests, unspliced_batch
    from test_oracle_window import _body, _sparse_chain
    for n in (150, 300):
        body = _body(n)
        for left, right in ((1000, 1000), (0, 1000), (1000, 0)):
            dref = DeviceReference(Reference.from_sequences([b"N" * left + body + b"N" * right]), buf)
            reads, minus, rsp, qsp, dp = [], [], [], [], []
            for near_head in (True, False):
                for mi in (False, True):
                    for allow in (1, 0):
                        a, b = _sparse_chain(n, near_head, left)
                        reads.append(revcomp(body) if mi else body)
                        minus.append(int(mi))
                        rsp.append(a)
                        qsp.append(b)
                        dp.append(allow)
            req = ChainRequests.from_lists(reads, [0] * len(reads), minus, rsp, qsp, dp)
            out = unspliced_batch(dref, req)
            for p, allow in zip(out, dp):
                if n > 256 and not allow:
                    assert p is None
                else:
                    assert p is not None and p.pos == left and p.cigar == [Match(n)]
                    assert p.lane == (1 if n <= 256 else 2)
            dref.close()


def test_window_myers_and_intron_refinements(buf):
    from esperanto_amd.intron_chain import (WINDOW_JOB_DTYPE, donor_plan, acceptor_plan,
                                            refine_acceptor_start, refine_donor_end,
                                            verify_split_read, window_myers)
    from esperanto_amd.reference import DeviceReference
    from test_oracle_window import _run_plan_on_oracle
    ref = _ref_with_runs(61, [40000, 800])
    o = oracle.OracleRef(ref)
    seqs = [oracle.decode(o, k, 0, int(L)) for k, L in enumerate(ref.lens)]
    dref = DeviceReference(ref, buf)
    rng = np.random.default_rng(12)
    reads = []
    for _ in range(300):
        c = 0
        m = int(rng.integers(20, 300))
        P = int(rng.integers(0, len(seqs[c]) - m))
        r = bytearray(seqs[c][P:P + m])
        for _ in range(int(rng.integers(0, 6))):
            r[int(rng.integers(0, m))] = b"ACGTNacgt"[int(rng.integers(0, 9))]
        reads.append(bytes(r))
    jobs = []
    for k, r in enumerate(reads):
        for _ in range(3):
            qs = int(rng.integers(0, len(r)))
            qe = int(rng.integers(qs, min(len(r), qs + 260) + 1))
            c = int(rng.integers(0, 2))
            L = len(seqs[c])
            ts = int(rng.integers(0, L))
            te = min(L, ts + int(rng.integers(0, 400)))
            jobs.append((k, qs, qe, c, ts, te, 0))
    jobs.append((0, 0, 10, 5, 0, 10, 0))       # no such contig
    jobs.append((0, 0, 10, 1, 700, 900, 0))    # past the contig end
    J = np.array(jobs, dtype=WINDOW_JOB_DTYPE)
    for mode in (0, 1, 2):
        say(f"window myers {len(jobs)} jobs mode {mode}")
        dist, pos, st = window_myers(dref, reads, J, mode)
        exp = _run_plan_on_oracle(o, reads, J[:-2], mode)
        for j in range(len(jobs) - 2):
            assert st[j] == 0
            assert dist[j] == exp[j][0], (mode, j)
            if mode != 2:
                assert int(pos[j]) == exp[j][1], (mode, j)
        assert st[-1] == 1 and st[-2] == 1
    # the refinement functions: searches on the device, same answers as the plan on the oracle
    n = len(reads)
    ctg = [0] * n
    joff = [int(rng.integers(0, len(r) + 1)) for r in reads]
    hi = [int(rng.integers(0, 40000)) for _ in reads]
    imp = [h - int(rng.integers(0, 300)) for h in hi]
    got = refine_donor_end(dref, reads, ctg, joff, hi, imp)
    out, jobs2, who, base = donor_plan(ref.lens, reads, ctg, joff, hi, imp)
    res = _run_plan_on_oracle(o, reads, jobs2, 0)
    for j, k in enumerate(who):
        out[k] = base[j] + res[j][1]
    assert got == out
    got = refine_acceptor_start(dref, reads, ctg, joff, hi, imp)
    out, jobs2, who, base = acceptor_plan(ref.lens, reads, ctg, joff, hi, imp)
    res = _run_plan_on_oracle(o, reads, jobs2, 1)
    for j, k in enumerate(who):
        out[k] = base[j] + res[j][1]
    assert got == out
    got = verify_split_read(dref, reads, ctg, imp, [h + 200 for h in hi], hi, [h + 100 for h in hi])
    assert len(got) == n and all(isinstance(x, tuple) for x in got)
    dref.close()
