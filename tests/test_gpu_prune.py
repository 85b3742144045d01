"""Exactness of the pruned full-matrix pass of extend_hint's escalations
(ESP_OPT_FULL_PRUNE), through the C ABI.

extend_hint (extend.rs:392-428) reruns an escalated band attempt on the full
matrix.  The pruned pass computes only the cells that can still reach the
band attempt's score and proves per pair that no skipped cell could; a pair
it cannot prove reruns one wavefront per pair over the whole matrix.  Every
record, CIGAR and path must equal the oracle's extend_hint and the unpruned
run's, on batches built
so that the pass both succeeds (mate-rescue-like 1,950-nt windows, spliced-
like pseudo-references whose read sits off the hint diagonal) and must fall
back (indels that walk the path off its stripe, tandem copies that tie at
other diagonals, heavy substitution that leaves a low band score, N and
lower-case bytes), at read lengths of every row tile (64 .. 192 pruned; a
batch of 200-250-nt reads takes the 264-row tile, which has no pruned pass,
so its escalations go to the one-wave-per-pair rerun directly),
with the default and a custom scoring (EditDir Ag included).
"""
from __future__ import annotations

import sys

import numpy as np
import pytest

from esperanto_amd.records import EditDir, ExtendParams
from oracle import oracle

pytestmark = pytest.mark.gpu

ACGT = np.frombuffer(b"ACGT", dtype=np.uint8)


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


def _rand(rng, n):
    return ACGT[rng.integers(0, 4, n)].tobytes()


def _subst(rng, seq, rate):
    s = bytearray(seq)
    for q in np.nonzero(rng.random(len(s)) < rate)[0]:
        s[q] = b"ACGT"[(b"ACGT".index(bytes([s[q]])) + int(rng.integers(1, 4))) % 4]
    return bytes(s)


def _mate_like(rng, n, starts=None):
    """150-nt reads with 1% substitutions in 1,950-nt windows, hint = the
    planted start (drawn from `starts` when given: many pairs a diagonal, as
    in a large batch)."""
    out = []
    for _ in range(n):
        w = _rand(rng, 1950)
        p = int(rng.integers(0, 1950 - 150)) if starts is None else int(rng.choice(starts))
        out.append((_subst(rng, w[p:p + 150], 0.01), w, p, 0))
    return out


def _spliced_like(rng, n):
    """Reads of 80-190 nt placed 70 nt into 600-nt windows, hint offset 40 (the
    band attempt runs 30 diagonals off the path and touches its edge)."""
    out = []
    for _ in range(n):
        w = _rand(rng, 600)
        m = int(rng.integers(80, 191))
        out.append((_subst(rng, w[70:70 + m], 0.01), w, 40, 0))
    return out


def _adversarial(rng, n):
    out = []
    for k in range(n):
        kind = k % 6
        m = int(rng.integers(50, 191))
        wlen = int(rng.integers(m + 60, 1400))
        w = bytearray(_rand(rng, wlen))
        p = int(rng.integers(0, wlen - m - 20))
        if kind == 0:  # an indel of 1-12 bases in the middle of the read
            d = int(rng.integers(1, 13))
            r = bytes(w[p:p + m // 2]) + (bytes(w[p + m // 2 + d:p + m + d]) if rng.random() < 0.5
                                          else _rand(rng, d) + bytes(w[p + m // 2:p + m - d]))
        elif kind == 1:  # a tandem copy of the locus elsewhere in the window: ties
            r = bytes(w[p:p + m])
            q = int(rng.integers(0, wlen - m))
            w[q:q + m] = r
        elif kind == 2:  # heavy substitution: a low band score, a deep head
            r = _subst(rng, bytes(w[p:p + m]), float(rng.choice([0.08, 0.15, 0.3])))
        elif kind == 3:  # N runs and lower-case bytes in both sequences
            r = bytearray(_subst(rng, bytes(w[p:p + m]), 0.01))
            for q in rng.choice(m, 4, replace=False):
                r[q] = ord("N")
            for q in rng.choice(wlen, 20, replace=False):
                w[q] = w[q] | 0x20
            w[p + 5:p + 9] = b"NNNN"
            r = bytes(r)
        elif kind == 4:  # a hint off by a few diagonals
            r = _subst(rng, bytes(w[p:p + m]), 0.02)
            p = max(0, p + int(rng.integers(-20, 21)))
        else:  # a random read: nothing to find
            r = _rand(rng, m)
        out.append((r, bytes(w), p, 0))
    return out


def _short_reads(rng, n, lo, hi):
    """Reads of lo..hi nt with 1% substitutions in 900-nt windows, hint = the
    planted start: the 64- and 128-row tiles of the pruned passes."""
    out = []
    for _ in range(n):
        m = int(rng.integers(lo, hi + 1))
        w = _rand(rng, 900)
        p = int(rng.integers(0, 900 - m))
        out.append((_subst(rng, w[p:p + m], 0.01), w, p, 0))
    return out


def _long_reads(rng, n):
    """200-250-nt reads: the 264-row tile, which has no pruned pass."""
    out = []
    for _ in range(n):
        m = int(rng.integers(200, 251))
        w = _rand(rng, 1200)
        p = int(rng.integers(0, 1200 - m))
        out.append((_subst(rng, w[p:p + m], 0.01), w, p, 0))
    return out


def _run(buf, part, params, prune):
    """prune: 0 off, 1 the register-window pass for small heads beside the
    whole-tile pass for the rest, 2 the whole-tile pass for all."""
    from esperanto_amd import _native
    from esperanto_amd.batch import DeviceBatch
    buf.set_option(_native.ESP_OPT_FULL_PRUNE, int(prune))
    try:
        db = DeviceBatch(part, buf)
        db.launch_gotoh(params, 2)
        db.sync()
        g = db.gotoh_results()
        st = db.last_plan_stats()
        kinds = [t[0] for t in db.last_timings()]
        db.free()
    finally:
        buf.set_option(_native.ESP_OPT_FULL_PRUNE, 1)
    return g, st, kinds


def _same(g, o, label):
    assert np.array_equal(g[0], o[0]), label
    assert np.array_equal(g[4], o[4]), label
    assert np.array_equal(g[5], o[5]), label
    off = g[3]
    for k in range(len(g[4])):
        a = int(off[k])
        assert np.array_equal(g[2][a:a + g[4][k]], o[2][a:a + o[4][k]]), (label, k)


@pytest.mark.parametrize("scoring", ["default", "custom"])
def test_pruned_full_matrix_equals_oracle_and_unpruned(buf, scoring):
    from esperanto_amd.batch import PairBatch
    params = ExtendParams() if scoring == "default" else ExtendParams(
        match_score=1, mismatch=-3, gap_open=5, gap_ext=1, edit_dir=EditDir.Ag)
    rng = np.random.default_rng(91 if scoring == "default" else 92)
    starts = [int(x) for x in rng.integers(0, 1950 - 150, 12)]
    for label, pairs in (("mate-like", _mate_like(rng, 1500)),
                         ("mate-like, shared diagonals", _mate_like(rng, 1500, starts)),
                         ("spliced-like", _spliced_like(rng, 1500)),
                         ("short reads, 64-row tile", _short_reads(rng, 1200, 40, 64)),
                         ("short reads, 128-row tile", _short_reads(rng, 1200, 70, 128)),
                         ("adversarial", _adversarial(rng, 1800)),
                         ("long reads", _long_reads(rng, 400))):
        part = PairBatch.from_pairs(pairs)
        say(f"{scoring} {label}: {part.npairs} pairs, pruning on")
        gp, sp, kp = _run(buf, part, params, 1)
        say(f"{scoring} {label}: routing {sp}")
        say(f"{scoring} {label}: whole-tile pruned pass only")
        gw, sw, kw = _run(buf, part, params, 2)
        say(f"{scoring} {label}: routing {sw}")
        say(f"{scoring} {label}: pruning off")
        gu, su, ku = _run(buf, part, params, 0)
        _same(gp, gu, f"{label} pruned vs unpruned")
        _same(gw, gu, f"{label} whole-tile pruned vs unpruned")
        o = oracle.band_batch(part, params, 2, nthreads=8)
        _same(gp, o, f"{label} pruned vs oracle")
        assert sp["escalated"] == su["escalated"] == sw["escalated"] > 0
        assert sp["full_pruned"] + sp["full_unpruned"] == sp["escalated"]
        assert su["full_pruned"] == 0 and su["full_unpruned"] == su["escalated"]
        assert 12 not in ku
        if label == "long reads":
            # no pruned pass for the 264-row tile: every escalation goes to
            # the one-wave-per-pair rerun first
            assert 12 not in kp and 13 in kp and sp["full_second"] == sp["escalated"], sp
            assert sp["full_pruned"] == sp["escalated"] - sp["full_unpruned"], sp
            continue
        assert 12 in kp and 13 in kp and 11 in kp
        if label != "adversarial":
            # the passes prove nearly every escalation of a clean read, or
            # hand it to the one-wave-per-pair rerun (the register-window
            # pass narrows the stripe of a wave whose diagonals spread far)
            assert sp["full_pruned"] >= 0.9 * sp["escalated"], sp
            assert sw["full_pruned"] >= 0.9 * sw["escalated"], sw
            # the whole-tile pass itself proves nearly all of them
            assert sw["full_second"] <= 0.1 * sw["escalated"], sw
        else:
            assert sp["full_second"] > 0, sp  # the one-wave-per-pair rerun was exercised


def test_full_matrix_census(buf):
    """ESP_OPT_FULL_CENSUS counts the cells the full-matrix passes compute
    without changing any result.  Unpruned, the computed cells are exactly the
    matrices of the pairs the full-matrix kernel ran; on mate-like pairs whose
    hint band lies inside the window (so that no band attempt runs in the
    general kernel, which escalates by itself) those are the m x n of every
    pair whose path says it escalated, summed on the host.  Pruned, the
    computed cells are a fraction of them and at most the lane-cells the waves
    issued."""
    from esperanto_amd import _native
    from esperanto_amd.batch import PairBatch
    params = ExtendParams()
    rng = np.random.default_rng(93)
    mate = [q for q in _mate_like(rng, 900) if 31 <= q[2] <= 1950 - 150 - 31][:700]
    for label, pairs in (("mate-like", mate), ("adversarial", _adversarial(rng, 600))):
        part = PairBatch.from_pairs(pairs)
        say(f"census {label}: {part.npairs} pairs")
        ref_p, _, _ = _run(buf, part, params, 1)
        buf.set_option(_native.ESP_OPT_FULL_CENSUS, 1)
        try:
            # the whole-tile pass's count (in a small batch the register-
            # window pass's waves span diagonals far apart, its narrowed
            # stripes leave more pairs to the one-wave rerun, and the census
            # counts the rerun's full matrices too)
            gp, sp, _ = _run(buf, part, params, 2)
            g1, s1, _ = _run(buf, part, params, 1)
            gu, su, _ = _run(buf, part, params, 0)
        finally:
            buf.set_option(_native.ESP_OPT_FULL_CENSUS, 0)
        _same(gp, ref_p, f"{label} census on vs off")
        _same(g1, ref_p, f"{label} census on vs off, register-window pass")
        _same(gu, ref_p, f"{label} census unpruned vs pruned")
        assert 0 < s1["full_cells_computed"] <= s1["full_lane_cells_issued"]
        assert s1["full_cells_matrix"] == su["full_cells_matrix"]
        esc = np.nonzero(gp[5] == 1)[0]
        mn = int(sum(int(part.read_len[k]) * int(part.win_len[k]) for k in esc))
        say(f"census {label}: pruned {sp['full_cells_computed']} of {su['full_cells_matrix']} cells, "
            f"issued {sp['full_lane_cells_issued']}; unpruned {su['full_cells_computed']}; "
            f"{len(esc)} escalated, {sp['escalated']} through the full-matrix kernel")
        assert sp["escalated"] > 0
        assert su["full_cells_computed"] == su["full_cells_matrix"] == sp["full_cells_matrix"]
        assert su["full_lane_cells_issued"] >= su["full_cells_computed"]
        assert 0 < sp["full_cells_computed"] <= sp["full_lane_cells_issued"]
        if label == "mate-like":
            assert len(esc) == sp["escalated"]
            assert su["full_cells_matrix"] == mn
            assert sp["full_cells_computed"] < mn
        else:
            assert su["full_cells_matrix"] <= mn


def test_prune_autotune_changes_shapes_not_results(buf):
    """esp_autotune_prune times every stripe/head shape of its list on the
    probe (extend_hint mode) and keeps one; the records, CIGARs and extend_hint
    paths after tuning equal those before it, and the oracle's."""
    from esperanto_amd import _native
    from esperanto_amd.batch import DeviceBatch, PairBatch
    params = ExtendParams()
    rng = np.random.default_rng(95)
    part = PairBatch.from_pairs(_mate_like(rng, 1200))
    before, st, _ = _run(buf, part, params, 1)
    db = DeviceBatch(part, buf)
    try:
        rep = db.autotune_prune(params, reps=1)
        say(f"prune autotune: {rep}")
        assert len(rep["candidates"]) == 8 and rep["candidates"][0][:2] == (32, 16)
        assert rep["chosen"] in [c[:2] for c in rep["candidates"]]
        assert all(c[2] > 0 for c in rep["candidates"]) and rep["ms_best"] <= rep["ms_default"]
        db.launch_gotoh(params, 2)
        db.sync()
        after = db.gotoh_results()
    finally:
        db.free()
        buf.set_option(_native.ESP_OPT_FULL_PRUNE_SHAPE, 0)
    _same(after, before, "tuned vs default shape")
    _same(after, oracle.band_batch(part, params, 2, nthreads=8), "tuned vs oracle")
