"""Batched calls captured as HIP graphs (esp_gotoh_graph_create,
esp_mate_rescue_graph_create, esp_spliced_graph_create / esp_graph_launch)
against the eager calls on the same batch: an uneven band batch with a tenth
of its pairs outside their band class (the strip kernel), full mode, and
extend_hint mode with its escalations (the sort, the pruned passes and the
unpruned kernel all inside the graph); the mate-rescue and spliced window
pipelines.  Every replay is checked field by field (extension, TouchRec,
CIGAR words and counts, aux; pipeline records and stitched CIGARs) after the
outputs are cleared, with other eager work on the context between replays;
then new read contents of the same lengths (new mate positions for the
rescue pipeline) are written into the device arrays and replayed; then a
graph captured for shorter pairs than the device holds leaves the
over-length pairs alone, reports them through esp_graph_check, and still gets
every other pair right."""
from __future__ import annotations

import ctypes
import sys

import numpy as np
import pytest

from esperanto_amd import _native
from esperanto_amd.records import ExtendParams

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


def _clear_outputs(db):
    n = db.npairs
    for arr, nbytes in ((db.ext, 16 * n), (db.touch, 24 * n), (db.cig_n, 4 * n), (db.aux, 4 * n),
                        (db.cigar, 4 * int(db.cig_off_host[-1]))):
        z = np.full(max(nbytes, 1), 0xA5, dtype=np.uint8)
        _native.check(_native.lib().esp_copy_to_device(db.buf.handle, ctypes.c_void_p(arr.ptr),
                                                       ctypes.c_void_p(z.ctypes.data), nbytes), db.buf.handle)


def _same(a, b, rows=None):
    ext_a, touch_a, cig_a, off, n_a, aux_a = a
    ext_b, touch_b, cig_b, _, n_b, aux_b = b
    idx = range(len(n_a)) if rows is None else rows
    for k in idx:
        assert np.array_equal(ext_a[k], ext_b[k]), k
        assert np.array_equal(touch_a[k], touch_b[k]), k
        assert n_a[k] == n_b[k] and aux_a[k] == aux_b[k], k
        o = int(off[k])
        assert np.array_equal(cig_a[o:o + n_a[k]], cig_b[o:o + n_b[k]]), k


def _mixed_batch(seed):
    from esperanto_amd.batch import workload_c3
    hb = workload_c3(20_000)
    rng = np.random.default_rng(seed)
    pick = rng.choice(hb.npairs, hb.npairs // 10, replace=False)
    hb.offset[pick] = rng.integers(0, 8, pick.shape[0])  # clamped bands: strip class
    return hb


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_graph_replays_equal_eager(buf, mode):
    from esperanto_amd.batch import DeviceBatch, workload_c1
    hb = _mixed_batch(5)
    if mode == 1:
        hb = workload_c1(3000)
    P = ExtendParams()
    db = DeviceBatch(hb, buf)
    other = DeviceBatch(workload_c1(2000), buf)
    say(f"graph mode {mode}: eager")
    _clear_outputs(db)  # fields a mode leaves unwritten then match across runs
    db.launch_gotoh(P, mode)
    db.sync()
    eager = db.gotoh_results()
    g = db.capture_gotoh(P, mode)
    try:
        for rep in range(3):
            _clear_outputs(db)
            db.sync()
            say(f"graph mode {mode}: replay {rep}")
            g.launch()
            g.check()
            _same(db.gotoh_results(), eager)
            other.launch_gotoh(P, 0)
            other.sync()
        # new contents, same lengths: substitutions in every read
        rng = np.random.default_rng(9)
        reads = hb.reads.copy()
        hit = rng.random(reads.shape) < 0.02
        reads[hit] = np.frombuffer(b"ACGT", dtype=np.uint8)[rng.integers(0, 4, int(hit.sum()))]
        _native.check(_native.lib().esp_copy_to_device(buf.handle, ctypes.c_void_p(db.d["reads"].ptr),
                                                       ctypes.c_void_p(reads.ctypes.data), reads.nbytes),
                      buf.handle)
        hb2 = type(hb)(**{**hb.__dict__, "reads": reads})
        ref = DeviceBatch(hb2, buf)
        _clear_outputs(ref)
        _clear_outputs(db)
        ref.launch_gotoh(P, mode)
        ref.sync()
        say(f"graph mode {mode}: replay on new reads")
        g.launch()
        g.check()
        _same(db.gotoh_results(), ref.gotoh_results())
        ref.free()
    finally:
        g.destroy()
        db.free()
        other.free()


def test_graph_reports_pairs_longer_than_at_capture(buf):
    from esperanto_amd.batch import DeviceBatch
    hb = _mixed_batch(6)
    P = ExtendParams()
    db = DeviceBatch(hb, buf)
    _clear_outputs(db)
    db.launch_gotoh(P, 0)
    db.sync()
    eager = db.gotoh_results()
    cap = int(np.percentile(hb.read_len, 60))
    rl = np.minimum(hb.read_len, cap).astype(np.uint32)
    wl = hb.win_len.astype(np.uint32).copy()
    meta = _native.CHostMeta(rl.ctypes.data, wl.ctypes.data, hb.offset.ctypes.data, hb.half.ctypes.data)
    pc = P.to_c()
    h = ctypes.c_void_p()
    _native.check(_native.lib().esp_gotoh_graph_create(buf.handle, ctypes.byref(db._cb), ctypes.byref(meta),
                                                       ctypes.byref(pc), 0, ctypes.byref(db._out),
                                                       ctypes.byref(h)), buf.handle)
    try:
        say(f"graph captured for reads <= {cap}")
        _clear_outputs(db)
        db.sync()
        _native.check(_native.lib().esp_graph_launch(buf.handle, h), buf.handle)
        with pytest.raises(_native.EspError):
            _native.check(_native.lib().esp_graph_check(buf.handle, h), buf.handle)
        rows = np.nonzero(hb.read_len <= cap)[0]
        _same(db.gotoh_results(), eager, rows)
    finally:
        _native.lib().esp_graph_destroy(buf.handle, h)
        db.free()


def test_graph_reports_a_long_read_in_a_long_window(buf):
    """Captured with a long read in a short window and a short read in a long
    one; a replay whose pair has both the long read and the long window passes
    the separate length limits but not the strip scratch: it is reported, and
    the other pairs are still right."""
    from esperanto_amd.batch import DeviceBatch, PairBatch
    from seqgen import Rng, rand_seq, rand_window
    rng = Rng(0x6A_0001)
    pairs = []
    for m, n in [(200, 300), (200, 3000), (60, 90), (150, 200)]:
        r = rand_seq(rng, m)
        pairs.append((r, rand_window(rng, r, n), 0, 0))
    hb = PairBatch.from_pairs(pairs)
    P = ExtendParams()
    db = DeviceBatch(hb, buf)
    _clear_outputs(db)
    db.launch_gotoh(P, 1)
    db.sync()
    eager = db.gotoh_results()
    rl = hb.read_len.astype(np.uint32).copy()
    rl[1] = 20  # at capture: a short read in the long window
    wl = hb.win_len.astype(np.uint32).copy()
    meta = _native.CHostMeta(rl.ctypes.data, wl.ctypes.data, hb.offset.ctypes.data, hb.half.ctypes.data)
    pc = P.to_c()
    h = ctypes.c_void_p()
    _native.check(_native.lib().esp_gotoh_graph_create(buf.handle, ctypes.byref(db._cb), ctypes.byref(meta),
                                                       ctypes.byref(pc), 1, ctypes.byref(db._out),
                                                       ctypes.byref(h)), buf.handle)
    try:
        say("graph captured with (200, 300) and (20, 3000); replayed with (200, 3000)")
        _clear_outputs(db)
        db.sync()
        _native.check(_native.lib().esp_graph_launch(buf.handle, h), buf.handle)
        with pytest.raises(_native.EspError):
            _native.check(_native.lib().esp_graph_check(buf.handle, h), buf.handle)
        _same(db.gotoh_results(), eager, [0, 2, 3])
    finally:
        _native.lib().esp_graph_destroy(buf.handle, h)
        db.free()


def test_autotune_changes_shapes_not_results(buf):
    """esp_autotune_band picks each half's grid from 1 .. its occupancy limit
    by timing a probe; batches after it give the same records as before."""
    from esperanto_amd.batch import DeviceBatch, workload_c3
    from esperanto_amd.extend import ExtendBuffer
    P = ExtendParams()
    hb = _mixed_batch(7)
    b2 = ExtendBuffer(0)  # a context of its own, so other tests keep the default grids
    try:
        db = DeviceBatch(hb, b2)
        _clear_outputs(db)
        db.launch_gotoh(P, 0)
        db.sync()
        before = db.gotoh_results()
        probe = DeviceBatch(workload_c3(30_000), b2)
        say("autotune on a 30,000-pair probe")
        rep = probe.autotune_band(P, reps=1)
        probe.free()
        for half, (chosen, occ, ms_def, ms_best) in rep.items():
            assert half in (16, 30, 32, 48, 64) and 1 <= chosen <= occ
            assert ms_best <= ms_def + 1e-9
        _clear_outputs(db)
        db.launch_gotoh(P, 0)
        db.sync()
        _same(db.gotoh_results(), before)
        db.free()
    finally:
        b2.close()


def _clear(buf, arr, nbytes):
    z = np.full(max(nbytes, 1), 0xA5, dtype=np.uint8)
    _native.check(_native.lib().esp_copy_to_device(buf.handle, ctypes.c_void_p(arr.ptr),
                                                   ctypes.c_void_p(z.ctypes.data), nbytes), buf.handle)


def _mate_state(mb):
    rec, cig, off, cig_n = mb.raw_results()
    words = [cig[int(off[k]):int(off[k]) + int(cig_n[k])].tobytes() for k in range(mb.npairs)]
    return rec.tobytes(), cig_n.tobytes(), words


def _splice_state(sb):
    rec, ext, cig, off, cig_n, st, soff = sb.raw_results()
    words = [cig[int(off[k]):int(off[k]) + int(cig_n[k])].tobytes() for k in range(sb.npairs)]
    stw = [st[int(soff[k]):int(soff[k]) + int(rec[k]["n_stitched"])].tobytes() for k in range(sb.npairs)]
    return rec.tobytes(), ext.tobytes(), cig_n.tobytes(), words, stw


def test_window_pipeline_graphs_equal_eager(buf):
    """Mate rescue (prep, EA-Myers on both strands, select, extend_hint with
    escalations, finish) and the spliced pipeline (pseudo-reference decode,
    extend_hint, stitch) each recorded as one graph: three replays equal the
    eager call; then the mates are moved and a replay equals a fresh eager
    call on the moved batch."""
    from esperanto_amd.reference import DeviceReference
    from esperanto_amd.rescue import MATE_REC_DTYPE, DeviceMateBatch, MateRequests
    from esperanto_amd.splice import SPLICE_REC_DTYPE, DeviceSpliceBatch
    from esperanto_amd.synth_window import splice_rescue_workload
    from esperanto_amd.batch import DeviceBatch, workload_c1
    P = ExtendParams()
    ref, sp, mt, _ = splice_rescue_workload(3000, ref_len=2_000_000)
    dref = DeviceReference(ref, buf)
    other = DeviceBatch(workload_c1(1000), buf)
    mb = DeviceMateBatch(mt, dref, buf)
    sb = DeviceSpliceBatch(sp, dref, buf)
    gm = gs = None
    try:
        n = mb.npairs
        mb.launch(P)
        mb.sync()
        eager_m = _mate_state(mb)
        sb.launch(P)
        sb.sync()
        eager_s = _splice_state(sb)
        say("mate rescue graph: recording")
        gm = mb.capture(P)
        say("spliced graph: recording")
        gs = sb.capture(P)
        for rep in range(3):
            _clear(buf, mb.rec, MATE_REC_DTYPE.itemsize * n)
            _clear(buf, mb.cig_n, 4 * n)
            _clear(buf, mb.cigar, 4 * int(mb.cig_off_host[-1]))
            _clear(buf, sb.rec, SPLICE_REC_DTYPE.itemsize * n)
            _clear(buf, sb.ext, 16 * n)
            _clear(buf, sb.cig_n, 4 * n)
            _clear(buf, sb.cigar, 4 * int(sb.cig_off_host[-1]))
            _clear(buf, sb.stitched, 4 * int(sb.stitch_off_host[-1]))
            buf_sync(buf)
            say(f"window graphs: replay {rep}")
            gm.launch()
            gs.launch()
            gm.check()
            gs.check()
            assert _mate_state(mb) == eager_m
            assert _splice_state(sb) == eager_s
            other.launch_gotoh(P, 0)
            other.sync()
        # move every mate by a few bases: a replay reads the new positions
        rng = np.random.default_rng(12)
        pos = (mt.mate_pos.astype(np.int64) + rng.integers(-40, 41, n)).clip(0).astype(np.uint32)
        _native.check(_native.lib().esp_copy_to_device(buf.handle, ctypes.c_void_p(mb.d["mate_pos"].ptr),
                                                       ctypes.c_void_p(pos.ctypes.data), pos.nbytes),
                      buf.handle)
        moved = MateRequests(mt.reads, mt.read_off, mt.read_len, mt.contig, pos)
        fresh = DeviceMateBatch(moved, dref, buf)
        fresh.launch(P)
        fresh.sync()
        want = _mate_state(fresh)
        fresh.free()
        _clear(buf, mb.rec, MATE_REC_DTYPE.itemsize * n)
        _clear(buf, mb.cig_n, 4 * n)
        _clear(buf, mb.cigar, 4 * int(mb.cig_off_host[-1]))
        buf_sync(buf)
        say("mate rescue graph: replay on moved mates")
        gm.launch()
        gm.check()
        assert _mate_state(mb) == want
    finally:
        for g in (gm, gs):
            if g is not None:
                g.destroy()
        mb.free()
        sb.free()
        other.free()
        dref.close()


def buf_sync(buf):
    _native.check(_native.lib().esp_sync(buf.handle), buf.handle)
