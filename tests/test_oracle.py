"""Pins the CPU restatement (oracle/) before it is trusted as the checker.

- every literal known-answer test of the reference (tests/kats.py), run
  through the dispatch path and through both accumulator widths of run_band;
- the score of every extension equals the maximum of an independent textbook
  local affine-gap DP over the same cell set (extend.rs:2033-2128 gives the
  closed form for the central-indel family; the brute force covers the rest);
- i16 and i32 run_band agree field for field on the reference's own seeded
  sweeps (extend.rs:2578-2913 require every kernel to equal run_band::<i16>;
  the i32 path is the reference's non-pure route);
- Myers infix_best_end / best_start equal a textbook semi-global EA edit
  distance (myers_ea.rs has no unit tests; the DP definition pins it);
- stitch_cigar literal cases of map/tests/splice_stitch.rs:42-142.
A planted-mistake check at the end shows the brute force catches a dropped
gap-extension term.
"""
from __future__ import annotations

import numpy as np
import pytest

from esperanto_amd.records import (DiagHint, Del, EditDir, ExtendParams, Ins, Match, RefSkip,
                                   SoftClip)
from oracle import oracle
from kats import ALL_EXTEND_KATS
from seqgen import Rng, rand_seq, rand_window


@pytest.fixture(scope="module", autouse=True)
def _lib():
    oracle.lib()


class OracleDispatch:
    extend = staticmethod(oracle.extend)
    extend_hint = staticmethod(oracle.extend_hint)
    band = staticmethod(oracle.band)


class OracleRunBand:
    def __init__(self, is16):
        self.is16 = is16

    def band(self, r, w, p, b):
        return oracle.run_band(r, w, p, self.is16, b)

    def extend(self, r, w, p):
        return oracle.run_band(r, w, p, self.is16, None)[0]

    def extend_hint(self, r, w, p, hint):
        return oracle.extend_hint(r, w, p, hint)


@pytest.mark.parametrize("kat", ALL_EXTEND_KATS, ids=lambda f: f.__name__)
@pytest.mark.parametrize("impl", [OracleDispatch(), OracleRunBand(True), OracleRunBand(False)],
                         ids=["dispatch", "run_band_i16", "run_band_i32"])
def test_reference_kats(kat, impl):
    kat(impl)


def _shapes():
    rng = Rng(0x5EED_0001)
    cases = []
    for m in (1, 2, 5, 8, 17, 33, 60):
        for n in (1, 3, 9, 20, 41, 80):
            for rep in range(2):
                read = rand_seq(rng, m)
                win = rand_window(rng, read, n) if rep == 0 else rand_seq(rng, n)
                cases.append((read, win, EditDir(rng.below(3))))
    return cases


def test_score_equals_textbook_local_affine_max():
    params_set = [ExtendParams(), ExtendParams(gap_open=400, gap_ext=400),
                  ExtendParams(match_score=1, mismatch=-3, gap_open=5, gap_ext=2)]
    rng = Rng(0x5EED_0002)
    for read, win, d in _shapes():
        for base in params_set:
            p = ExtendParams(base.match_score, base.mismatch, base.gap_open, base.gap_ext, d)
            assert oracle.extend(read, win, p).score == oracle.bf_local_affine(read, win, p)
            off, half = rng.below(len(win) + 20) - 10, rng.below(12)
            e, _ = oracle.band(read, win, p, (off, half))
            assert e.score == oracle.bf_local_affine(read, win, p, (off, half)), (read, win, off, half)


def test_cigar_consumes_read_and_rescores():
    """The CIGAR spans exactly m read bases, and re-walking it over the window
    (M: LUT score, gap of length k: open + k*ext) reproduces the score once
    the neutral flanks (score 0 by construction) are included."""
    lut = {}
    for read, win, d in _shapes():
        p = ExtendParams(edit_dir=d)
        e = oracle.extend(read, win, p)
        m = sum(c.len for c in e.cigar if c.op in (0, 1, 4))
        assert m == len(read)
        i, j, score = e.read_start, e.ref_start, 0
        for c in e.cigar:
            if c.op == 0:
                for _ in range(c.len):
                    q, r = read[i] & 0xDF, win[j] & 0xDF
                    if q == r and chr(q) in "ACGT":
                        score += p.match_score
                    elif d == EditDir.Ag and r == ord("A") and q == ord("G"):
                        pass
                    elif d == EditDir.Tc and r == ord("T") and q == ord("C"):
                        pass
                    else:
                        score += p.mismatch
                    i += 1
                    j += 1
            elif c.op == 1:
                score -= p.gap_open + p.gap_ext * c.len
                i += c.len
            elif c.op == 2:
                score -= p.gap_open + p.gap_ext * c.len
                j += c.len
        assert score == e.score, (read, win, e)


def _sweep_cases(seed, shapes, halfs_fn, offsets_fn, reps=2):
    rng = Rng(seed)
    for m, n in shapes:
        for rep in range(reps):
            read = rand_seq(rng, m)
            win = rand_seq(rng, n) if rep == 1 else rand_window(rng, read, n)
            for half in halfs_fn(m, n):
                for off in offsets_fn(m, n, half):
                    yield read, win, EditDir(rng.below(3)), off, half


def test_i16_equals_i32_on_reference_band_sweep():
    """Shapes/halves/offsets of avx2_band_matches_scalar (extend.rs:2723-2822)."""
    shapes = [(8, 9), (8, 40), (15, 16), (16, 16), (17, 33), (20, 200), (31, 31), (32, 64),
              (60, 61), (60, 300), (40, 20), (17, 9)]

    def halfs(m, n):
        mn = max(m, n)
        return [0, 1, 2, 30, mn + 5] if n <= 64 else [0, 1, 2, 30, 300]

    def offs(m, n, half):
        hh = min(half, max(m, n))
        return [-hh - 5, -1, 0, 5, n - m, n, n + hh + 5, 4 * n]

    for read, win, d, off, half in _sweep_cases(0xA5A55A5AC3C33C3C, shapes, halfs, offs):
        p = ExtendParams(edit_dir=d)
        a = oracle.run_band(read, win, p, True, (off, half))
        b = oracle.run_band(read, win, p, False, (off, half))
        assert a == b, (read, win, off, half)


def test_myers_matches_textbook_edit_distance():
    rng = Rng(0x5EED_0003)
    for m in (1, 2, 7, 31, 64, 65, 100, 127, 128, 129, 150, 200, 255, 256):
        for n in (0, 1, 5, 40, 130, 300):
            for _ in range(2):
                read = rand_seq(rng, m)
                text = rand_window(rng, read, n) if n > m else rand_seq(rng, n)
                if rng.below(4) == 0:
                    read = read.lower()
                assert oracle.infix_best_end(read, text) == oracle.bf_ea_edit(read, text)
                d, s = oracle.infix_best_start(read, text)
                rd, rend = oracle.bf_ea_edit(read[::-1], text[::-1])
                assert (d, s) == (rd, n - rend)


def test_myers_guards():
    assert oracle.infix_best_end(b"", b"ACGT") == (0, 0)
    assert oracle.infix_best_start(b"", b"ACGT") == (0, 0)
    assert oracle.infix_best_end(b"A" * 257, b"A" * 300) == (257, 0)
    assert oracle.infix_best_start(b"A" * 257, b"A" * 300) == (257, 0)
    assert oracle.infix(b"ACGT", b"") == 4


def test_myers_ea_equivalence():
    """Read G matches ref A, read C matches ref T, not the reverse (myers_ea.rs:28-38)."""
    assert oracle.infix(b"GGGG", b"AAAA") == 0
    assert oracle.infix(b"CCCC", b"TTTT") == 0
    assert oracle.infix(b"AAAA", b"GGGG") == 4
    assert oracle.infix(b"NNNN", b"NNNN") == 4


def test_mate_window_kat_myers_part():
    """align.rs:4694-4751: 1000 x A reference, read 98A+52G has EA distance 0."""
    win = b"A" * 1000
    assert oracle.infix_best_start(b"A" * 98 + b"G" * 52, win)[0] == 0
    e = oracle.extend_hint(b"A" * 150, win, ExtendParams(), DiagHint(oracle.infix_best_start(b"A" * 150, win)[1]))
    assert e.cigar == [Match(150)]


def test_stitch_cigar_literals():
    """map/tests/splice_stitch.rs:42-142"""
    bounds = [(0, 100, 110), (10, 210, 230)]
    cases = [
        ([Match(10), Ins(1), Match(10)], [Match(10), Ins(1), RefSkip(100), Match(10)]),
        ([Match(10), Del(2), Match(8)], [Match(10), RefSkip(100), Del(2), Match(8)]),
        ([Match(8), Del(2), Match(10)], [Match(8), Del(2), RefSkip(100), Match(10)]),
        ([Match(8), Del(2), Ins(1), Match(10)], [Match(8), Del(2), Ins(1), RefSkip(100), Match(10)]),
        ([SoftClip(3), Match(10), Ins(1), Match(10), SoftClip(2)],
         [SoftClip(3), Match(10), Ins(1), RefSkip(100), Match(10), SoftClip(2)]),
    ]
    for inp, exp in cases:
        assert oracle.stitch_cigar([c.word for c in inp], 0, bounds, [100]) == exp
    inp = [SoftClip(3), Match(10), Ins(1), Match(10), Del(2), Match(8), SoftClip(2)]
    b3 = [(0, 100, 110), (10, 210, 220), (20, 420, 440)]
    assert oracle.stitch_cigar([c.word for c in inp], 0, b3, [100, 200]) == [
        SoftClip(3), Match(10), Ins(1), RefSkip(100), Match(10), RefSkip(200), Del(2), Match(8),
        SoftClip(2)]
    for start, inp, exp in [
        (0, [Match(20)], [Match(10), RefSkip(100), Match(10)]),
        (8, [Del(4), Match(6)], [Del(2), RefSkip(100), Del(2), Match(6)]),
        (0, [Match(10), Ins(1)], [Match(10), Ins(1)]),
        (10, [Ins(1), Match(10)], [Ins(1), Match(10)]),
        (11, [Match(5)], [Match(5)]),
        (0, [Match(8)], [Match(8)]),
    ]:
        assert oracle.stitch_cigar([c.word for c in inp], start, bounds, [100]) == exp
    inp = [Match(5), Ins(1), Match(5)]
    assert oracle.stitch_cigar([c.word for c in inp], 0, [], []) == inp


def test_planted_mistake_is_caught():
    """A gap model missing its extension term disagrees with the restatement on
    the closed-form family, so the brute-force pin has teeth."""
    from seqgen import body
    ref = body(160)
    q = ref[:80] + ref[83:]
    p = ExtendParams()
    good = oracle.extend(q, ref, p).score
    assert good == 2 * len(q) - p.gap_open - 3 * p.gap_ext
    wrong = oracle.bf_local_affine(q, ref, ExtendParams(gap_open=p.gap_open + p.gap_ext, gap_ext=0))
    assert wrong != good


def test_path_score_checker_agrees_with_the_oracle():
    """The path checker the full-size GPU tests use (espo_path_scores) rescored
    every path the oracle itself reports, on the 100/164-nt band-32 workload
    and an uneven band sample: every path fits its pair and scores what the
    DP reported; a planted wrong score or shifted ref_start is caught."""
    from esperanto_amd.batch import workload_c1, workload_c3
    for hb in (workload_c1(3000), workload_c3(3000)):
        ext, _, cig, off, n, _ = oracle.band_batch(hb, ExtendParams(), 0, nthreads=8)
        score, bad = oracle.path_scores(hb, ext, cig, off, n, nthreads=8)
        assert not bad.any()
        assert np.array_equal(score, ext[:, 3])
        e2 = ext.copy()
        e2[::7, 3] += 1
        s2, _ = oracle.path_scores(hb, e2, cig, off, n, nthreads=8)
        assert (s2[::7] != e2[::7, 3]).all()
        e3 = ext.copy()
        e3[:, 2] += 4000
        _, b3 = oracle.path_scores(hb, e3, cig, off, n, nthreads=8)
        assert b3[n > 0].all()
