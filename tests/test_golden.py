"""The C restatement against the committed golden fixtures (CPU only).

extend_cases.json holds inputs from the reference's seeded sweeps
(extend.rs:2578-2913), the 100-nt/164-nt band-32 workload, edge cases and
the rescue penalties; myers_cases.json holds EA-Myers cases whose expected
values come from a textbook edit-distance DP.  The tests here check that
the oracle still reproduces every fixture, that every fixture's CIGAR
re-walks to its score, and that every modes-0/1 score is the maximum of an
independent local affine DP over the same rows.
"""
from __future__ import annotations

import numpy as np

from oracle import oracle
import golden_cases as G


def test_fixture_inventory():
    ext = G.extend_cases()
    modes = {m: sum(1 for c in ext if c["mode"] == m) for m in (0, 1, 2)}
    assert all(v >= 100 for v in modes.values()), modes
    assert any(len(c["read"]) == 0 for c in ext) and any(len(c["window"]) == 0 for c in ext)
    assert any(c["params"][2] == 400 for c in ext)
    assert any(c["path"] == 1 for c in ext if c["mode"] == 2)
    mye = G.myers_cases()
    assert any(len(c["read"]) > 256 for c in mye) and any(len(c["read"]) == 0 for c in mye)


def test_oracle_reproduces_extend_fixtures():
    for (mode, _, _), cases in G.groups(G.extend_cases()).items():
        hb = G.batch_of(cases)
        ext, touch, cig, off, n, aux = oracle.band_batch(hb, cases[0]["prm"], mode, nthreads=4)
        G.check_batch(cases, ext, touch, cig, off, n, aux, "oracle")


def test_fixture_cigars_rescore_and_scores_are_textbook_maxima():
    for (mode, _, _), cases in G.groups(G.extend_cases()).items():
        hb = G.batch_of(cases)
        ext = np.array([c["ext"] for c in cases], dtype=np.int64).reshape(len(cases), 4)
        words = [c["cigar"] for c in cases]
        cig_n = np.array([len(w) for w in words], dtype=np.uint32)
        cig_off = np.zeros(len(cases) + 1, dtype=np.uint64)
        cig_off[1:] = np.cumsum(cig_n, dtype=np.uint64)
        cig = np.array([x for w in words for x in w] or [0], dtype=np.uint32)
        score, bad = oracle.path_scores(hb, ext, cig, cig_off, cig_n, cases[0]["prm"])
        # the empty extension [S(m)] with score 0 (extend.rs:430-442) is the
        # one result the path checker does not model: its clip starts at 0
        empty = np.array([c["ext"] == [0, 0, 0, 0] and c["cigar"] == ([len(c["read"]) << 4 | 4]
                                                                      if c["read"] else [])
                          for c in cases])
        assert ((bad == 0) | empty).all() and (score == ext[:, 3]).all()
        if mode != 2:
            for c in cases:
                band = (c["offset"], c["half"]) if mode == 0 else None
                assert oracle.bf_local_affine(c["read"], c["window"], c["prm"], band) == c["ext"][3]


def test_oracle_reproduces_myers_fixtures():
    cases = G.myers_cases()
    d, e, s = oracle.myers_batch(G.myers_batch_of(cases), True, nthreads=4)
    wd, we, ws = G.myers_expected(cases)
    assert (d == wd).all() and (e == we).all() and (s == ws).all()
    for c in cases[:: 7]:
        assert oracle.infix_best_end(c["read"], c["text"]) == (c["dist"], c["end"])
        assert oracle.infix_best_start(c["read"], c["text"]) == (c["dist"], c["start"])
