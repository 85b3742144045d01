"""Device junction refinement, align_spliced over chains and reextend_with
against the oracle, field by field: refined junctions (coordinates, strand,
signal, support), refusal codes, segment counts, exon bounds, pseudo-reference
extension records and CIGARs, stitched CIGARs, genomic starts, and the
re-extension records and CIGARs.  The reference's own known answers
(splice.rs:894-955, align.rs:5591-5639) run through the device entry points
too."""
from __future__ import annotations

import sys

import numpy as np
import pytest

from esperanto_amd.records import EditDir, ExtendParams, Match, RefSkip, SoftClip, cigar_from_words
from esperanto_amd.reference import Reference
from oracle import oracle
import splicegen

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-test] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def buf():
    from esperanto_amd.extend import ExtendBuffer
    b = ExtendBuffer(0)
    yield b
    b.close()


def _fallback(o, jl, contig, start, end, minus):
    """junction_signal (splice.rs:864-868) and support of the fallback junction."""
    def dinuc(p, donor):
        if not donor and p < 2:
            return b"NN"
        a, b = (p, p + 2) if donor else (p - 2, p)
        L = int(o.arrs[0][contig])
        if a >= L:
            return b"NN"
        s = oracle.decode(o, contig, a, min(b, L))
        return (s + b"NN")[:2]
    d, a = dinuc(start, True), dinuc(end, False)
    tab = {False: {(b"GT", b"AG"): "GtAg", (b"GC", b"AG"): "GcAg", (b"AT", b"AC"): "AtAc"},
           True: {(b"CT", b"AC"): "GtAg", (b"CT", b"GC"): "GcAg", (b"GT", b"AT"): "AtAc"}}
    sig = tab[bool(minus)].get((d.upper(), a.upper()), "NonCanonical")
    return (start, end, bool(minus), sig, jl.support(contig, start, end, int(minus)))


def test_refine_junction_batch_matches_oracle(buf):
    from esperanto_amd.junction import JunctionLib, refine_junction_read_batch
    from esperanto_amd.reference import DeviceReference
    ref, chains, quads, truth = splicegen.spliced_workload(0x5E3, 96)
    o = oracle.OracleRef(ref)
    jl = oracle.JLib(quads)
    dref = DeviceReference(ref, buf)
    dlib = JunctionLib(dref, quads)
    try:
        qs = splicegen.refine_queries(9, chains, truth, 1500)
        reads = [c[0] for c in chains]
        say(f"refine_junction_read: {len(qs)} queries")
        got = refine_junction_read_batch(dref, dlib, reads, qs)
        found = 0
        for q, g in zip(qs, got):
            c, contig, ns, ne, qa, qb, minus = q[:7]
            want = oracle.refine_junction_read(o, reads[c], contig, ns, ne, qa, qb, minus, jl)
            if want is None and len(q) > 7 and contig < ref.ncontig:
                want = _fallback(o, jl, contig, q[7][0], q[7][1], minus)
            assert (None if g is None else (g.start, g.end, g.minus, g.signal, g.support)) == want, q
            found += want is not None
        assert found > 500
        # the reference's boundary-competition test (splice.rs:894-955)
        for transcript_minus in (False, True):
            from test_oracle_junction import _xorshift_genome
            genome = _xorshift_genome()
            donor, acceptor = (b"CT", b"AC") if transcript_minus else (b"GT", b"AG")
            genome[120:122] = donor
            genome[298:300] = acceptor
            genome[300:302] = donor
            genome[119] = ord("T")
            query = bytearray(genome[60:120])
            query[59] = ord("C")
            query += genome[300:379]
            r2 = DeviceReference(Reference.from_sequences([bytes(genome)]), buf)
            for annotated in (False, True):
                lib2 = JunctionLib(r2, [(0, 120, 300, int(transcript_minus))] if annotated else [])
                qq = [(0, 0, 120 + s, 300 + s, 60 + s, 60 + s, rm) for rm in (False, True) for s in (1, 2)]
                for g in refine_junction_read_batch(r2, lib2, [bytes(query)], qq):
                    assert (g.start, g.end, g.minus, g.signal, g.support) == (120, 300, transcript_minus, "GtAg",
                                                                              int(annotated))
                lib2.close()
            r2.close()
    finally:
        dlib.close()
        dref.close()
        jl.close()


def _check_chains(buf, ref, chains, quads, label, params=None):
    from esperanto_amd.junction import JunctionLib, align_spliced_batch
    from esperanto_amd.reference import DeviceReference
    o = oracle.OracleRef(ref)
    jl = oracle.JLib(quads)
    dref = DeviceReference(ref, buf)
    dlib = JunctionLib(dref, quads)
    try:
        say(f"align_spliced {label}: {len(chains)} chains")
        got = align_spliced_batch(dref, dlib, [c[0] for c in chains], [c[1] for c in chains],
                                  [c[2] for c in chains], [c[3] for c in chains], splicegen.K,
                                  extend_params=params)
        status = {}
        for k, ((read, contig, minus, anchors), g) in enumerate(zip(chains, got)):
            w = oracle.align_spliced(o, jl, read, contig, minus, anchors, splicegen.K, params=params)
            ctx = (label, k, w["status"], w["refusal"])
            st = w["status"] if contig < ref.ncontig else -1
            assert g.status == st and g.n_segments == w["n_segments"], (ctx, g.status, g.n_segments)
            status[g.status] = status.get(g.status, 0) + 1
            if st == 0:
                assert g.refusal == w["refusal"], ctx
                continue
            if st != 1:
                continue
            assert g.extension == w["ext"], ctx
            assert g.cigar == w["stitched"], ctx
            assert g.genomic_start == w["genomic_start"], ctx
            assert [(j.start, j.end, j.minus, j.signal, j.support) for j in g.junctions] == w["junctions"], ctx
            assert g.exon_bounds == w["bounds"], ctx
        return status
    finally:
        dlib.close()
        dref.close()
        jl.close()


def test_align_spliced_batch_matches_oracle(buf):
    ref, chains, quads, truth = splicegen.spliced_workload(0x5E4, 400)
    two = next(c for c in chains if len(oracle.segment_chain(c[3], splicegen.K)) >= 2)
    chains = chains + [(two[0], 3, False, two[3])]  # contig out of range
    st = _check_chains(buf, ref, chains, quads, "planted")
    assert st.get(1, 0) > 150 and st.get(0, 0) > 50 and st.get(-1, 0) == 1, st
    st = _check_chains(buf, ref, chains[:120], quads[:5], "custom scoring", ExtendParams(1, -3, 5, 2, EditDir.Ag))
    assert st.get(1, 0) > 30, st


def test_align_spliced_known_answer(buf):
    """align.rs:5591-5639: genomic start 37, [M20, N60, M20], junction 57..117."""
    from esperanto_amd.junction import align_spliced_batch
    from esperanto_amd.reference import DeviceReference
    from test_oracle_junction import spliced_chain_fixture
    seq, read, anchors = spliced_chain_fixture()
    dref = DeviceReference(Reference.from_sequences([seq]), buf)
    try:
        g = align_spliced_batch(dref, None, [read], [0], [False], [anchors], 15)[0]
        assert g.status == 1 and g.genomic_start == 37
        assert g.cigar == [Match(20), RefSkip(60), Match(20)]
        assert [(j.start, j.end, j.signal) for j in g.junctions] == [(57, 117, "GtAg")]
    finally:
        dref.close()


def test_unstranded_library_is_refused(buf):
    import ctypes
    from esperanto_amd import _native
    from esperanto_amd.junction import refine_junction_read_batch
    from esperanto_amd.reference import DeviceReference
    dref = DeviceReference(Reference.from_sequences([b"ACGT" * 100]), buf)
    t = np.array([[0, 10, 100]], dtype=np.uint32)
    h = ctypes.c_void_p()
    _native.check(_native.lib().esp_junctions_upload(buf.handle, 1, ctypes.c_void_p(t.ctypes.data), ctypes.byref(h)))

    class Fake:
        handle = h
    try:
        with pytest.raises(_native.EspError):
            refine_junction_read_batch(dref, Fake(), [b"ACGT" * 10], [(0, 0, 10, 100, 5, 5, False)])
    finally:
        _native.lib().esp_junctions_free(buf.handle, h)
        dref.close()


def _reext_placements(seed, n, genome_len=200_000):
    rng = np.random.default_rng(seed)
    g = bytearray(np.frombuffer(b"ACGT", dtype=np.uint8)[rng.integers(0, 4, genome_len)].tobytes())
    out = []
    for i in range(n):
        m = int(rng.integers(40, 250))
        lead = int(rng.integers(0, 25))
        tail = int(rng.integers(0, 25))
        lead, tail = min(lead, m // 3), min(tail, m // 3)
        span = m - lead - tail
        pos = int(rng.integers(lead + 1, genome_len - span - tail - 1)) if i % 50 else int(rng.integers(0, 20))
        read = bytearray(g[max(pos - lead, 0):pos + span + tail])
        read = read + bytearray(b"A" * (m - len(read)))
        for j in range(m):  # editing-direction substitutions, more inside the clips
            rate = 0.2 if (j < lead or j >= lead + span) else 0.01
            if rng.random() < rate:
                read[j] = {65: 71, 84: 67}.get(read[j], read[j]) if i % 2 else {71: 65}.get(read[j], 84)
        cig = ([SoftClip(lead)] if lead else []) + [Match(span)] + ([SoftClip(tail)] if tail else [])
        contig = 0
        if i % 97 == 5:
            cig = [Match(span // 2), RefSkip(100), Match(span - span // 2)] + ([SoftClip(tail)] if tail else [])
            if lead:
                cig = [SoftClip(lead)] + cig
        if i % 89 == 7:
            contig = 2
        out.append((bytes(read), contig, pos, cig))
    return Reference.from_sequences([bytes(g)]), out


@pytest.mark.parametrize("edit_dir", [0, 1, 2])
def test_reextend_batch_matches_oracle(buf, edit_dir):
    from esperanto_amd.junction import reextend_batch
    from esperanto_amd.reference import DeviceReference
    ref, pl = _reext_placements(31 + edit_dir, 1500)
    o = oracle.OracleRef(ref)
    dref = DeviceReference(ref, buf)
    try:
        say(f"reextend_with dir {edit_dir}: {len(pl)} placements")
        got = reextend_batch(dref, [p[0] for p in pl], [p[1] for p in pl], [p[2] for p in pl],
                             [p[3] for p in pl], ExtendParams(), edit_dir)
        counts = {}
        for k, (p, g) in enumerate(zip(pl, got)):
            w = oracle.reextend_with(o, p[0], p[1], p[2], [c.word for c in p[3]], ExtendParams(), edit_dir)
            assert (g.status, g.pos, g.score, g.wstart, g.wlen, g.aligned, g.new_aligned) == (
                w["status"], w["pos"], w["score"], w["wstart"], w["wlen"], w["aligned"], w["new_aligned"]), (k, w)
            assert [c.word for c in g.cigar] == w["cigar"], k
            counts[g.status] = counts.get(g.status, 0) + 1
        assert counts.get(1, 0) > 500 and counts.get(2, 0) > 5 and counts.get(-1, 0) > 5, counts
        moved = sum(1 for p, g in zip(pl, got) if g.status == 1 and g.pos != p[2])
        assert moved > 50, moved
    finally:
        dref.close()


IC_FIELDS = ("status", "reason", "genomic_start", "score", "left_start", "jstart", "jend", "right_end", "support",
             "minus", "n_stitched", "donor_end", "acceptor_start", "vd", "va")
SIG = ("GtAg", "GcAg", "AtAc", "NonCanonical")


@pytest.mark.parametrize("seed,n_runs", [(3, False), (4, False), (5, True)])
def test_intron_chain_batch_matches_oracle(buf, seed, n_runs):
    """Every record field, extension and stitched CIGAR against the oracle;
    with n_runs, N runs in the genome beside and inside the exons and
    lower-case read stretches (the seed extensions stop at N and fold case)."""
    from esperanto_amd.junction import JunctionLib, intron_chain_batch
    from esperanto_amd.reference import DeviceReference
    ref, reads, pools, quads = splicegen.intron_chain_workload(seed, 600, n_runs=n_runs)
    o = oracle.OracleRef(ref)
    jl = oracle.JLib(quads)
    dref = DeviceReference(ref, buf)
    dlib = JunctionLib(dref, quads)
    try:
        say(f"intron chain: {len(reads)} reads")
        rec, exts, stitched = intron_chain_batch(dref, dlib, reads, pools)
        placed = 0
        for k, (r, p) in enumerate(zip(reads, pools)):
            w = oracle.intron_chain(o, jl, r, p)
            got = {f: int(rec[k][f]) for f in IC_FIELDS}
            want = {f: int(w[f]) for f in IC_FIELDS}
            assert got == want, (k, got, want)
            assert SIG[int(rec[k]["signal"])] == w["signal"], k
            assert stitched[k] == w["stitched"], k
            if w["status"] == 1:
                assert exts[k] == w["ext"], k
                placed += 1
        assert placed > (250 if n_runs else 300), placed
    finally:
        dlib.close()
        dref.close()
        jl.close()


def test_intron_chain_known_answer(buf):
    """align.rs:5067-5120: pos 1000, [S11, M60, N1940, M79], score 278."""
    from esperanto_amd.junction import intron_chain_batch
    from esperanto_amd.reference import DeviceReference
    from test_oracle_junction import _xorshift_genome
    genome = _xorshift_genome(5000, 0x517CC1B7)
    for donor, acceptor in ((b"GT", b"AG"), (b"CT", b"AC")):
        genome[1060:1062] = donor
        genome[2998:3000] = acceptor
        genome[1059] = ord("T")
        genome[3000] = ord("A")
        query = b"N" * 11 + bytes(genome[1000:1060]) + bytes(genome[3000:3079])
        dref = DeviceReference(Reference.from_sequences([bytes(genome)]), buf)
        try:
            pools = [[(q + off, r + off, 0, strand) for q, r, span in ((11, 1000, 60), (71, 3000, 79))
                      for off in range(3, span - 14, 4)] for strand in (0, 1)]
            rec, exts, stitched = intron_chain_batch(dref, None, [query, query], pools)
            for k in range(2):
                assert int(rec[k]["status"]) == 1 and int(rec[k]["genomic_start"]) == 1000
                assert stitched[k] == [SoftClip(11), Match(60), RefSkip(1940), Match(79)]
                assert int(rec[k]["score"]) == 278
        finally:
            dref.close()
