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
