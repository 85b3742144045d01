"""The batched paths at the workload sizes they are measured on, through
properties that hold at any size, plus the oracle on samples:

- the 2,000,000-pair uneven band workload: every reported path fits its pair
  (soft clips, read span, window) and rescores, from its CIGAR over the read
  and window, to the reported score (espo_path_scores, pinned against the
  oracle's own paths in test_oracle.py); the graph replay equals the eager
  call on every pair; 10,000 random pairs equal the oracle field by field;
- the 1,000,000-pair Myers workload (150 nt against 214 nt): every pair's
  distance, end and start equal the oracle's.
"""
from __future__ import annotations

import sys

import numpy as np
import pytest

from esperanto_amd.records import ExtendParams
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


def test_uneven_band_2m_pairs(buf):
    from esperanto_amd.batch import DeviceBatch, workload_c3
    P = ExtendParams()
    say("generating 2,000,000 uneven pairs")
    hb = workload_c3(2_000_000)
    db = DeviceBatch(hb, buf)
    try:
        say("eager batch")
        db.launch_gotoh(P, 0)
        db.sync()
        ext, touch, cig, off, n, aux = db.gotoh_results()
        say("rescoring every path")
        score, bad = oracle.path_scores(hb, ext, cig, off, n, P, nthreads=16)
        assert not bad.any(), np.nonzero(bad)[0][:10]
        assert np.array_equal(score, ext[:, 3]), np.nonzero(score != ext[:, 3])[0][:10]
        assert (n > 0).all() and (ext[:, 3] > 0).all()
        say("graph replay")
        g = db.capture_gotoh(P, 0)
        try:
            g.launch()
            g.check()
            ext2, touch2, cig2, _, n2, _ = db.gotoh_results()
        finally:
            g.destroy()
        assert np.array_equal(ext, ext2) and np.array_equal(touch, touch2) and np.array_equal(n, n2)
        assert np.array_equal(cig, cig2)
        say("oracle on 10,000 pairs")
        rng = np.random.default_rng(1)
        pick = np.sort(rng.choice(hb.npairs, 10_000, replace=False))
        sub = hb.subset(pick)
        o = oracle.band_batch(sub, P, 0, nthreads=16)
        assert np.array_equal(ext[pick], o[0]) and np.array_equal(touch[pick], o[1])
        assert np.array_equal(n[pick], o[4])
        for j, k in enumerate(pick.tolist()):
            a, b = int(off[k]), int(o[3][j])
            assert np.array_equal(cig[a:a + n[k]], o[2][b:b + o[4][j]]), k
    finally:
        db.free()


def test_myers_1m_pairs(buf):
    from esperanto_amd.batch import DeviceBatch, workload_c2
    say("generating 1,000,000 Myers pairs")
    hb = workload_c2(1_000_000)
    db = DeviceBatch(hb, buf)
    try:
        say("device Myers (end and start)")
        db.launch_myers(True)
        db.sync()
        dist, end, start = db.myers_results()
        say("oracle Myers")
        od, oe, os_ = oracle.myers_batch(hb, True, nthreads=16)
        assert np.array_equal(dist, od)
        assert np.array_equal(end, oe)
        assert np.array_equal(start, os_)
    finally:
        db.free()


def _query_len(words) -> int:
    return int(sum(int(w) >> 4 for w in words if (int(w) & 15) in (0, 1, 4)))


def _ref_len(words) -> int:
    return int(sum(int(w) >> 4 for w in words if (int(w) & 15) in (0, 2, 3)))


def test_splice_rescue_500k_pairs(buf):
    """The 500,000-pair splice / mate-rescue workload on a 64 Mb reference:
    every placed spliced read's stitched CIGAR consumes its 150 bases and
    stays on the contig, planted junctions are found where the read was cut
    (genomic start s - u with one RefSkip of e - s, or a read clipped at the
    cut); every placed mate's CIGAR consumes the read inside its window and
    lands at the planted position on the minus strand; the first 5,000 pairs
    of each equal the oracle field by field."""
    from esperanto_amd.reference import DeviceReference
    from esperanto_amd.rescue import DeviceMateBatch, MateRequests
    from esperanto_amd.splice import DeviceSpliceBatch, SpliceRequests
    from esperanto_amd.synth_window import splice_rescue_workload
    from test_gpu_window import _compare_mate, _compare_splice
    say("generating the 500,000-pair workload")
    ref, sp, mt, truth = splice_rescue_workload(500_000)
    dref = DeviceReference(ref, buf)
    L = int(ref.lens[0])
    try:
        say("spliced batch")
        db = DeviceSpliceBatch(sp, dref, buf)
        db.launch()
        db.sync()
        g = db.raw_results()
        db.free()
        rec, _, _, _, _, st, soff = g
        placed = rec["status"] == 1
        assert placed.mean() > 0.99
        s, e, u = truth["s"], truth["e"], truth["u"]
        at_cut = 0
        for k in np.nonzero(placed)[0].tolist():
            a = int(soff[k])
            words = st[a:a + int(rec[k]["n_stitched"])]
            assert _query_len(words) == 150, k
            assert int(rec[k]["genomic_start"]) + _ref_len(words) <= L, k
            skips = [int(w) >> 4 for w in words if (int(w) & 15) == 3]
            if int(rec[k]["genomic_start"]) == int(s[k] - u[k]) and skips == [int(e[k] - s[k])]:
                at_cut += 1
        assert at_cut > 0.95 * placed.sum()
        say("mate rescue batch")
        mb = DeviceMateBatch(mt, dref, buf)
        mb.launch()
        mb.sync()
        gm = mb.raw_results()
        mb.free()
        mrec, mcig, moff, mn = gm
        mplaced = mrec["placed"] == 1
        assert mplaced.mean() > 0.95
        p2 = truth["p2"]
        home = 0
        for k in np.nonzero(mplaced)[0].tolist():
            a = int(moff[k])
            words = mcig[a:a + int(mn[k])]
            assert _query_len(words) == 150, k
            r = mrec[k]
            assert r["wstart"] <= r["pos"] and r["pos"] + _ref_len(words) <= r["wstart"] + r["wlen"], k
            home += int(r["strand"] == 1 and int(r["pos"]) == int(p2[k]) + int(r["read_start"]))
        assert home > 0.95 * mplaced.sum()
        say("oracle on the first 5,000 pairs")
        o = oracle.OracleRef(ref)
        k = 5000
        r1 = int(sp.read_off[k - 1] + sp.read_len[k - 1])
        sub = SpliceRequests(sp.reads[:r1], sp.read_off[:k], sp.read_len[:k], sp.contig[:k], sp.hint[:k],
                             sp.exons[:int(sp.exon_off[k])], sp.exon_off[:k + 1])
        exp = oracle.spliced_batch(o, sub, nthreads=16)
        _compare_splice(sub, tuple(x[:k] if i in (0, 1, 4) else x for i, x in enumerate(g)), exp)
        r1 = int(mt.read_off[k - 1] + mt.read_len[k - 1])
        msub = MateRequests(mt.reads[:r1], mt.read_off[:k], mt.read_len[:k], mt.contig[:k], mt.mate_pos[:k])
        mexp = oracle.mate_batch(o, msub, nthreads=16)
        _compare_mate(msub, (mrec[:k], mcig, moff[:k + 1], mn[:k]), mexp)
    finally:
        dref.close()
