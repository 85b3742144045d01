"""The single-pair calls on one wavefront per pair (esp_wave.hip)
against the oracle and against the lane-per-pair kernel
(ESP_OPT_PLAIN_KERNEL = 1), bit-exact: score, read_start, read_end,
ref_start, CIGAR words, TouchRec.  Cases: the reference's banded, clamped and
full-matrix sweeps (empty reads and windows included), the i16 route, reads across
64-row block boundaries with the path on both band edges, traces that spill
from LDS to the wave's global scratch, and windows at the kernel's 4096-base
limit and past it.  Full-matrix and extend_hint calls (kept and escalated)
run the same sweep.
"""

from __future__ import annotations

import sys

import pytest

import sweeps
from esperanto_amd.records import DiagHint, ExtendParams
from oracle import oracle
from seqgen import Rng, rand_seq, rand_window

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-test] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def bufs():
    from esperanto_amd.extend import ExtendBuffer
    from esperanto_amd._native import ESP_OPT_PLAIN_KERNEL
    say("creating device contexts (wave, lane)")
    w = ExtendBuffer(0)
    l = ExtendBuffer(0)
    l.set_option(ESP_OPT_PLAIN_KERNEL, 1)
    yield w, l
    w.close()
    l.close()


def _same(a, b):
    assert (a.read_start, a.read_end, a.ref_start, a.score) == (b.read_start, b.read_end, b.ref_start, b.score)
    assert [c.word for c in a.cigar] == [c.word for c in b.cigar]


def _band_case(bufs, read, win, p, band):
    from esperanto_amd import extend as X
    ge, gt = X.band(read, win, p, band, bufs[0])
    oe, ot = oracle.band(read, win, p, band)
    _same(ge, oe)
    if band is not None:
        assert gt == ot
    le, lt = X.band(read, win, p, band, bufs[1])
    _same(ge, le)
    assert gt == lt


def test_wave_band_sweeps(bufs):
    cases = sweeps.band_blocks()[::3] + [c for c in sweeps.band_sweep() if len(c[1]) <= 1000][::2]
    cases += sweeps.band_dispatch() + sweeps.band_interior()[::2]
    say(f"band sweeps: {len(cases)} single calls")
    for k, (read, win, d, off, half) in enumerate(cases):
        if k % 250 == 0:
            say(f"case {k}: m={len(read)} n={len(win)} off={off} half={half}")
        _band_case(bufs, read, win, ExtendParams(edit_dir=d), (off, half))


def test_wave_full_sweeps(bufs):
    cases = [c for c in sweeps.full_sweep() if len(c[1]) <= 3000] + sweeps.full_blocks()[::2]
    cases += sweeps.full_dispatch()
    say(f"full sweeps: {len(cases)} single calls")
    for k, (read, win, d, _, _) in enumerate(cases):
        if k % 200 == 0 or not read or not win:
            say(f"case {k}: m={len(read)} n={len(win)}")
        _band_case(bufs, read, win, ExtendParams(edit_dir=d), None)


def _kept_hint_cases():
    """Reads on the hint diagonal whose first row ties no band edge (the
    window's first base and the row's right-edge base differ from the read's
    first base): extend_hint keeps the band attempt (path 0)."""
    rng = Rng(0x4B3E)

    def acgt(n):
        return bytes(b"ACGT"[rng.below(4)] for _ in range(n))

    out = []
    for m in (40, 100, 150, 250):
        read = bytearray(acgt(m))
        if read[30] == read[0]:
            read[30] = b"ACGT"[(b"ACGT".index(read[0]) + 1) % 4]
        pre = bytearray(acgt(20))
        if pre[0] == read[0]:
            pre[0] = b"ACGT"[(b"ACGT".index(read[0]) + 2) % 4]
        out.append((bytes(read), bytes(pre) + bytes(read) + acgt(40), sweeps.DIRS[m % 3], 20, 30))
    return out


def test_wave_extend_hint(bufs):
    from esperanto_amd import extend as X
    cases = sweeps.hint_cases()
    say(f"extend_hint: {len(cases)} single calls")
    paths = set()
    cases += _kept_hint_cases()
    for k, (read, win, d, off, _) in enumerate(cases):
        if k % 100 == 0:
            say(f"case {k}: m={len(read)} n={len(win)} hint={off}")
        p = ExtendParams(edit_dir=d)
        ge = X.extend_hint(read, win, p, bufs[0], DiagHint(off))
        oe, path = oracle.extend_hint(read, win, p, DiagHint(off), with_path=True)
        paths.add(path)
        _same(ge, oe)
    assert {0, 1, 3} <= paths


def test_wave_i16_route(bufs):
    cases = sweeps.i16_wrap_cases()
    say(f"i16 route: {len(cases)} pairs x 2 bands")
    for read, win, p in cases:
        say(f"m={len(read)} n={len(win)} match={p.match_score}")
        _band_case(bufs, read, win, p, (max(len(win) - len(read), 0) // 2, 12))
        if len(win) <= 1000:  # bands as wide as the window: the sizes the band sweeps cover
            _band_case(bufs, read, win, p, (0, len(win)))


@pytest.mark.parametrize("m", [63, 64, 65, 128, 129, 190, 300])
def test_wave_block_edges(bufs, m):
    rng = Rng(0xB10C + m)
    read = rand_seq(rng, m)
    win = rand_seq(rng, 20) + read[:60] + rand_seq(rng, 3) + read[60:] + rand_seq(rng, 25)
    say(f"block edges m={m}")
    for d in sweeps.DIRS:
        p = ExtendParams(edit_dir=d)
        _band_case(bufs, read, win, p, (20, 12))
        _band_case(bufs, read, win, p, (0, len(win)))
    for diag in (8, 32):
        w = rand_seq(rng, diag) + read + rand_seq(rng, 40)
        _band_case(bufs, read, w, ExtendParams(), (20, 12))


def test_wave_global_trace_and_limits(bufs):
    """Bands whose trace is larger than the LDS part (150 x 601, 400 x 161,
    1000 x 81 columns) go to the wave's global scratch; 4096 is the longest
    window the wave takes and 4097 goes to the lane kernel."""
    rng = Rng(0x7AC3)
    for m, n, half in ((150, 1000, 300), (400, 800, 80), (1000, 1300, 40), (100, 4096, 40), (100, 4097, 40)):
        read = rand_seq(rng, m)
        win = rand_window(rng, read, n)
        say(f"global trace / limit m={m} n={n} half={half}")
        for d in sweeps.DIRS:
            p = ExtendParams(edit_dir=d)
            _band_case(bufs, read, win, p, ((n - m) // 2, half))
            _band_case(bufs, read, win, p, (n // 3, 40))


def test_wave_empty_and_kernel_time(bufs):
    from esperanto_amd import extend as X
    _band_case(bufs, b"", b"ACGT", ExtendParams(), (0, 3))
    _band_case(bufs, b"ACGT", b"", ExtendParams(), (0, 3))
    rng = Rng(0x71AE)
    read = rand_seq(rng, 100)
    win = rand_window(rng, read, 164)
    X.band(read, win, ExtendParams(), (32, 32), bufs[0])
    ms = bufs[0].last_kernel_ms()
    say(f"one C1-shaped call on the wave kernel: {ms * 1e3:.1f} us")
    assert 0.0 < ms < 50.0
