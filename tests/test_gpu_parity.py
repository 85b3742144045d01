"""Parity of the HIP path with the CPU restatement, through the C ABI.

Every comparison is bit-exact: score, read_start, read_end, ref_start,
CIGAR words, TouchRec fields, extend_hint path, Myers (distance, end, start).
Inputs are the reference's own seeded sweeps (tests/sweeps.py), its literal
known-answer tests (tests/kats.py), the synthetic workloads at reduced size,
and edge cases (empty inputs, N-rich and lowercase bytes, rescue penalties,
band clamps at both window edges, n = 20,000 windows).
"""
from __future__ import annotations

import sys

import numpy as np
import pytest

from esperanto_amd.records import DiagHint, EditDir, ExtendParams
from oracle import oracle
from kats import ALL_EXTEND_KATS
import sweeps

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


class GpuPlain:
    def __init__(self, buf):
        from esperanto_amd import extend as X
        self.X = X
        self.buf = buf

    def extend(self, r, w, p):
        return self.X.extend(r, w, p, self.buf)

    def extend_hint(self, r, w, p, hint):
        return self.X.extend_hint(r, w, p, self.buf, hint)

    def band(self, r, w, p, b):
        return self.X.band(r, w, p, b, self.buf)


@pytest.mark.parametrize("kat", ALL_EXTEND_KATS, ids=lambda f: f.__name__)
def test_reference_kats_plain(buf, kat):
    say(f"kat {kat.__name__}")
    kat(GpuPlain(buf))


def _group_by_dir(cases):
    groups = {}
    for c in cases:
        groups.setdefault(int(c[2]), []).append(c)
    return groups


def compare_batch(buf, cases, mode, base=ExtendParams(), label=""):
    from esperanto_amd.batch import DeviceBatch, PairBatch
    for d, group in _group_by_dir(cases).items():
        p = ExtendParams(base.match_score, base.mismatch, base.gap_open, base.gap_ext, EditDir(d))
        hb = PairBatch.from_pairs([(c[0], c[1], c[3], c[4] if c[4] is not None else 0)
                                   for c in group])
        say(f"{label} mode={mode} dir={d} pairs={hb.npairs}")
        db = DeviceBatch(hb, buf)
        db.launch_gotoh(p, mode)
        db.sync()
        g_ext, g_touch, g_cig, off, g_n, g_aux = db.gotoh_results()
        o_ext, o_touch, o_cig, _, o_n, o_aux = oracle.band_batch(hb, p, mode, nthreads=8)
        for k in range(hb.npairs):
            ctx = (label, mode, d, k, group[k][3], group[k][4], len(group[k][0]), len(group[k][1]))
            assert tuple(g_ext[k]) == tuple(o_ext[k]), ctx
            a, b = int(off[k]), int(off[k])
            assert g_n[k] == o_n[k], ctx
            assert np.array_equal(g_cig[a:a + g_n[k]], o_cig[b:b + o_n[k]]), ctx
            if mode != 2:
                assert tuple(g_touch[k]) == tuple(o_touch[k]), ctx
            else:
                assert g_aux[k] == o_aux[k], ctx
        db.free()


def test_full_matrix_sweeps(buf):
    compare_batch(buf, sweeps.full_sweep(), 1, label="full_sweep")
    compare_batch(buf, sweeps.full_blocks(), 1, label="full_blocks")
    compare_batch(buf, sweeps.full_dispatch(), 1, label="full_dispatch")


def test_band_sweeps(buf):
    compare_batch(buf, sweeps.band_sweep(), 0, label="band_sweep")
    compare_batch(buf, sweeps.band_blocks(), 0, label="band_blocks")
    compare_batch(buf, sweeps.band_dispatch(), 0, label="band_dispatch")


def test_band_kernel_interior_sweep(buf):
    """Pairs inside the interior envelope of every specialised half: all of
    them take the band kernel (planner routing checked), bit-exact."""
    from esperanto_amd.batch import DeviceBatch, PairBatch
    cases = sweeps.band_interior()
    for d, group in _group_by_dir(cases).items():
        hb = PairBatch.from_pairs([(c[0], c[1], c[3], c[4]) for c in group])
        db = DeviceBatch(hb, buf)
        db.launch_gotoh(ExtendParams(edit_dir=EditDir(d)), 0)
        db.sync()
        st = db.last_plan_stats()
        # the band kernel's LDS holds m + 2 * half + 1 window rows: 320 for
        # halves up to 32, 352 for 48, 384 for 64; longer reads take the strip
        span = {16: 320, 30: 320, 32: 320, 48: 352, 64: 384}
        want = {h: sum(1 for c in group if c[4] == h and 0 < len(c[0]) <= span[h] - 2 * h - 1)
                for h in (16, 30, 32, 48, 64)}
        say(f"interior dir={d} routing {st}")
        assert st["band_pairs"] == want, (st, want)
        db.free()
    compare_batch(buf, cases, 0, label="band_interior")
    # gap_open + gap_ext = 43 does not fold into the 5-bit LUT: the packed
    # halves take their unfolded kernels
    compare_batch(buf, cases, 0, ExtendParams(gap_open=40, gap_ext=3), label="band_interior_unfolded")
    compare_batch(buf, cases, 0, ExtendParams(match_score=1, mismatch=-3, gap_open=5, gap_ext=2),
                  label="band_interior_custom")
    # scores outside the band kernel's int16 envelope (5-bit LUT fields, gap
    # penalties over 8000): every pair takes the int32 strip kernel instead
    wide = ExtendParams(match_score=20, mismatch=-30, gap_open=9000, gap_ext=100)
    group = [c for c in cases if int(c[2]) == 0]
    hb = PairBatch.from_pairs([(c[0], c[1], c[3], c[4]) for c in group])
    db = DeviceBatch(hb, buf)
    db.launch_gotoh(wide, 0)
    db.sync()
    st = db.last_plan_stats()
    assert sum(st["band_pairs"].values()) == 0 and st["strip_pairs"] == len(group), st
    db.free()
    compare_batch(buf, group, 0, wide, label="band_interior_int32")


def _left_clamped_cases():
    """Bands with 0 <= offset < half (clamped on the left only): reads placed
    at the window's first columns, on the band edge and inside it, with
    substitutions and indels, random reads (zero rows), lower case, N."""
    from seqgen import Rng, rand_seq
    rng = Rng(0x1EF7)
    out = []
    for half in (16, 30, 32, 48, 64):
        for off in sorted({0, 1, half // 2, half - 1}):
            for m in (1, 20, 64, 65, 150, 200):
                for place in (0, 3, off, off + half // 2):
                    n = off + m + half + int(rng.below(40))
                    w = bytearray(rand_seq(rng, n))
                    read = bytearray(w[place:place + m]) if place + m <= n else bytearray(rand_seq(rng, m))
                    for _ in range(int(rng.below(4))):
                        q = int(rng.below(max(m, 1)))
                        read[q] = b"ACGT"[int(rng.below(4))]
                    if m > 30 and rng.below(3) == 0:  # an indel of 1-3 bases
                        q, d = int(rng.below(m - 10)) + 5, int(rng.below(3)) + 1
                        read = read[:q] + read[q + d:] + bytearray(rand_seq(rng, d))
                    if rng.below(6) == 0:
                        read = bytearray(rand_seq(rng, m))
                    if rng.below(8) == 0:
                        read[int(rng.below(m))] = ord("N")
                    if rng.below(8) == 0:
                        read = bytearray(bytes(read).lower())
                    out.append((bytes(read), bytes(w), int(rng.below(3)), off, half))
    return out


def test_left_clamped_bands_take_the_band_kernel(buf):
    """Left-clamped bands (0 <= offset < half) run in the band kernel: the
    columns left of the window are 'N' columns holding H = 0 and the left
    touch is the first maximum at or before the first real cell.  Routing
    checked, records, CIGARs and TouchRec bit-exact against the oracle; in
    extend_hint mode (hint = offset, half 30) the kept/escalated path too;
    with a positive mismatch score they keep the strip kernel."""
    from esperanto_amd.batch import DeviceBatch, PairBatch
    cases = _left_clamped_cases()
    span = {16: 320, 30: 320, 32: 320, 48: 352, 64: 384}
    for d, group in _group_by_dir(cases).items():
        hb = PairBatch.from_pairs([(c[0], c[1], c[3], c[4]) for c in group])
        db = DeviceBatch(hb, buf)
        db.launch_gotoh(ExtendParams(edit_dir=EditDir(d)), 0)
        db.sync()
        st = db.last_plan_stats()
        want = {h: sum(1 for c in group if c[4] == h and 0 < len(c[0]) <= span[h] - 2 * h - 1
                       and c[3] + len(c[0]) + h <= len(c[1])) for h in (16, 30, 32, 48, 64)}
        say(f"left-clamped dir={d} routing {st}")
        assert st["band_pairs"] == want, (st, want)
        db.free()
    compare_batch(buf, cases, 0, label="left_clamped")
    compare_batch(buf, cases, 0, ExtendParams(match_score=1, mismatch=-3, gap_open=5, gap_ext=2),
                  label="left_clamped_custom")
    compare_batch(buf, cases, 0, ExtendParams(gap_open=40, gap_ext=3), label="left_clamped_unfolded")
    hint = [(r, w, d, off, 30) for r, w, d, off, half in cases if half == 30 or off < 30]
    compare_batch(buf, hint, 2, label="left_clamped_hint")
    pos = ExtendParams(match_score=2, mismatch=1, gap_open=4, gap_ext=2)
    group = [c for c in cases if c[2] == 0]
    hb = PairBatch.from_pairs([(c[0], c[1], c[3], c[4]) for c in group])
    db = DeviceBatch(hb, buf)
    db.launch_gotoh(pos, 0)
    db.sync()
    st = db.last_plan_stats()
    assert sum(st["band_pairs"].values()) == 0, st
    db.free()
    compare_batch(buf, group, 0, pos, label="left_clamped_positive_mismatch")


def test_long_windows(buf):
    _, lf = sweeps.full_sweep(include_long=True)
    _, lb = sweeps.band_sweep(include_long=True)
    compare_batch(buf, lf, 1, label="long_full")
    compare_batch(buf, lb, 0, label="long_band")


def test_extend_hint_batched(buf):
    compare_batch(buf, sweeps.hint_cases(), 2, label="hint")


def test_rescue_and_custom_scoring(buf):
    cases = sweeps.band_dispatch() + sweeps.full_dispatch()
    compare_batch(buf, [c for c in cases if c[4] is not None], 0,
                  ExtendParams(gap_open=400, gap_ext=400), label="rescue_band")
    compare_batch(buf, [c for c in cases if c[4] is None], 1,
                  ExtendParams(match_score=1, mismatch=-3, gap_open=5, gap_ext=2), label="custom_full")


def test_lowercase_and_ambiguity_bytes(buf):
    from seqgen import Rng
    rng = Rng(0x10CA5E)
    alphabet = b"ACGTNacgtnRYKM"
    cases = []
    for _ in range(200):
        m, n = 5 + rng.below(60), 5 + rng.below(120)
        r = bytes(alphabet[rng.below(len(alphabet))] for _ in range(m))
        w = bytes(alphabet[rng.below(len(alphabet))] for _ in range(n))
        cases.append((r, w, EditDir(rng.below(3)), rng.below(n + 10) - 5, rng.below(20)))
    compare_batch(buf, cases, 0, label="bytes_band")
    compare_batch(buf, [(c[0], c[1], c[2], 0, None) for c in cases], 1, label="bytes_full")


def test_plain_equals_batched_c1_sample(buf):
    from esperanto_amd import extend as X
    from esperanto_amd.batch import workload_c1
    hb = workload_c1(64)
    for k in range(0, 64, 8):
        r, w = hb.pair(k)
        say(f"plain pair {k}")
        e, t = X.band(r, w, ExtendParams(), (32, 32), buf)
        o, ot = oracle.band(r, w, ExtendParams(), (32, 32))
        assert e == o and t == ot


def test_config1_full_size(buf):
    """10,000 C1 pairs, band(32, 32), against the oracle."""
    from esperanto_amd.batch import DeviceBatch, workload_c1
    hb = workload_c1(10_000)
    say("config1 10k pairs")
    db = DeviceBatch(hb, buf)
    db.launch_gotoh(ExtendParams(), 0)
    db.sync()
    g = db.gotoh_results()
    o = oracle.band_batch(hb, ExtendParams(), 0, nthreads=8)
    assert np.array_equal(g[0], o[0])
    assert np.array_equal(g[1], o[1])
    assert np.array_equal(g[4], o[4])
    off = g[3]
    for k in range(hb.npairs):
        a = int(off[k])
        assert np.array_equal(g[2][a:a + g[4][k]], o[2][a:a + o[4][k]]), k
    db.free()


def test_extend_hint_band_route_and_escalation(buf):
    """extend_hint(offset 32) on 10,000 pairs of the 100/164-nt workload (band
    attempts in the half-30 band kernel), plus pairs built to escalate: a
    deletion or insertion of 35-60 bases shifts the diagonal past the band,
    and heavy substitution drops the score under 85% of perfect."""
    from esperanto_amd.batch import DeviceBatch, PairBatch, workload_c1
    hb = workload_c1(10_000)
    rng = np.random.default_rng(17)
    extra = []
    for _ in range(600):
        w = np.frombuffer(b"ACGT", dtype=np.uint8)[rng.integers(0, 4, 400)].tobytes()
        kind = int(rng.integers(0, 3))
        if kind == 0:  # deletion from the read: second half sits further right
            d = int(rng.integers(35, 61))
            r = w[60:120] + w[120 + d:180 + d]
        elif kind == 1:  # insertion into the read
            d = int(rng.integers(35, 61))
            r = w[60:120] + bytes(rng.integers(65, 69, d).astype(np.uint8)) + w[120:150]
        else:  # 20% substitutions
            r = bytearray(w[60:180])
            for q in rng.choice(120, 24, replace=False):
                r[q] = b"ACGT"[int(rng.integers(0, 4))]
            r = bytes(r)
        extra.append((r, w, 60, 0))
    from esperanto_amd import _native
    cases = [(hb, "hint_c1", 0), (PairBatch.from_pairs(extra), "hint_escalating", 0),
             (PairBatch.from_pairs(extra), "hint_escalating_lane_kernel", 1)]
    for part, label, kernel in cases:
        say(f"{label} {part.npairs} pairs")
        buf.set_option(_native.ESP_OPT_FULL_KERNEL, kernel)
        db = DeviceBatch(part, buf)
        db.launch_gotoh(ExtendParams(), 2)
        db.sync()
        buf.set_option(_native.ESP_OPT_FULL_KERNEL, 0)
        g = db.gotoh_results()
        o = oracle.band_batch(part, ExtendParams(), 2, nthreads=8)
        assert np.array_equal(g[0], o[0]), label
        assert np.array_equal(g[4], o[4]), label
        assert np.array_equal(g[5], o[5]), label
        off = g[3]
        for k in range(part.npairs):
            a = int(off[k])
            assert np.array_equal(g[2][a:a + g[4][k]], o[2][a:a + o[4][k]]), (label, k)
        if label.startswith("hint_escalating"):
            assert int((g[5] == 1).sum()) > 400
            rerun = [t for t in db.last_timings() if t[0] == 7]
            assert rerun and rerun[0][1] > 0  # the lane-per-pair full-matrix kernel ran
        kinds = {t[0] for t in db.last_timings()}
        assert 1 in kinds  # the band kernel ran the first attempts
        db.free()


@pytest.mark.skip(reason="wave-per-pair full-matrix kernel (DPP wave_shr:1 column pipeline) hung "
                         "on its first GPU run; its CPU emulation is bit-exact with the oracle and "
                         "no cause is shown in the code, so the kernel is not built")
def test_extend_hint_escalations_wave_per_pair_kernel(buf):
    """The escalation cases of the test above with ESP_OPT_FULL_KERNEL = 2."""
    from esperanto_amd import _native
    buf.set_option(_native.ESP_OPT_FULL_KERNEL, 2)


def test_config3_sample(buf):
    """20,000 pairs of the uneven-length workload (bands 16..64)."""
    from esperanto_amd.batch import DeviceBatch, workload_c3
    hb = workload_c3(20_000)
    say("config3 20k pairs")
    db = DeviceBatch(hb, buf)
    db.launch_gotoh(ExtendParams(), 0)
    db.sync()
    g = db.gotoh_results()
    o = oracle.band_batch(hb, ExtendParams(), 0, nthreads=8)
    assert np.array_equal(g[0], o[0])
    assert np.array_equal(g[1], o[1])
    assert np.array_equal(g[4], o[4])
    mism = [k for k in range(hb.npairs)
            if not np.array_equal(g[2][int(g[3][k]):int(g[3][k]) + g[4][k]],
                                  o[2][int(g[3][k]):int(g[3][k]) + o[4][k]])]
    assert not mism, mism[:10]
    db.free()


# ------------------------------------------------------------------ Myers

def _myers_cases():
    from seqgen import Rng, rand_seq, rand_window
    rng = Rng(0x3E7E75)
    cases = []
    for m in (1, 2, 7, 31, 63, 64, 65, 100, 127, 128, 129, 150, 191, 192, 193, 200, 255, 256, 257, 300):
        for n in (0, 1, 5, 40, 130, 214, 300, 1000):
            for _ in range(2):
                read = rand_seq(rng, m)
                text = rand_window(rng, read, n) if n > m else rand_seq(rng, n)
                if rng.below(5) == 0:
                    read = read.lower()
                cases.append((read, text, 0, 0))
    cases.append((b"", b"ACGT", 0, 0))
    cases.append((b"GGGG", b"AAAA", 0, 0))
    cases.append((b"CCCC", b"TTTT", 0, 0))
    cases.append((b"A" * 98 + b"G" * 52, b"A" * 1000, 0, 0))
    return cases


def _check_myers(buf, hb, label, bound=0):
    from esperanto_amd.batch import DeviceBatch
    say(f"myers {label} pairs={hb.npairs} bound={bound}")
    db = DeviceBatch(hb, buf)
    db.launch_myers(True, max_read_len=bound)
    db.sync()
    gd, ge, gs = db.myers_results()
    od, oe, os_ = oracle.myers_batch(hb, True, nthreads=8)
    bad = np.nonzero((gd != od) | (ge != oe) | (gs != os_))[0]
    assert bad.size == 0, [(int(k), int(hb.read_len[k]), int(hb.win_len[k]), gd[k], od[k], ge[k],
                            oe[k], gs[k], os_[k]) for k in bad[:5]]
    db.free()


def test_myers_sweep(buf):
    from esperanto_amd.batch import PairBatch
    _check_myers(buf, PairBatch.from_pairs(_myers_cases()), "sweep")


@pytest.mark.parametrize("bound", [300, 256, 150, 64, 1])
def test_myers_sweep_with_a_stated_bound(buf, bound):
    """esp_myers_batch with meta->max_read_len: the kernel's words come from
    the bound and nothing is read back; patterns past a short bound go to the
    8-word launch over the overflow list, so every answer equals the oracle's
    whatever the bound (300 and 256 need no list, 1 sends nearly every pair)."""
    from esperanto_amd.batch import PairBatch
    _check_myers(buf, PairBatch.from_pairs(_myers_cases()), "sweep", bound)


def test_myers_config2_sample(buf):
    from esperanto_amd.batch import workload_c2
    _check_myers(buf, workload_c2(50_000), "config2-50k")


def test_myers_plain(buf):
    from esperanto_amd import myers_ea as M
    for r, t, _, _ in _myers_cases()[::7]:
        assert M.infix_best_end(r, t, buf) == oracle.infix_best_end(r, t)
        assert M.infix_best_start(r, t, buf) == oracle.infix_best_start(r, t)
        assert M.infix(r, t, buf) == oracle.infix(r, t)


def test_out_of_envelope_scoring_is_rejected(buf):
    from esperanto_amd import extend as X
    from esperanto_amd._native import EspError
    with pytest.raises(EspError):
        X.extend(b"ACGT", b"ACGT", ExtendParams(gap_open=-1), buf)
    with pytest.raises(EspError):
        X.extend(b"ACGT", b"ACGT", ExtendParams(match_score=5000), buf)
