"""The one-pair-per-wavefront sweep (esp_wave.hip) modelled on the CPU
(tests/wave_model.py: 64 lanes in lockstep, one-lane shifts, margins, the
block hand-off through the previous block's last row, per-block
reductions, row-window trace, traceback and finish), against the oracle on
the reference's seeded sweeps: banded, clamped, full-matrix, extend_hint
(kept and escalated) and the i16 route.  Reads cross one, two and three
64-row blocks, so the hand-off between blocks is exercised.
"""
from __future__ import annotations

import pytest

import sweeps
import wave_model
from esperanto_amd.records import DiagHint, ExtendParams
from oracle import oracle


def _words(cig):
    return [(ln << 4) | op for op, ln in cig]


def _check_band(read, win, d, off, half, params=None):
    p = params or ExtendParams(edit_dir=d)
    band = (off, half) if half is not None else None
    ext, touch = oracle.band(read, win, p, band)
    got = wave_model.wave_pair(read, win, p, "band" if band else "full", off, half or 0)
    rs, re, fs, sc, cig, tch, _ = got
    assert (rs, re, fs, sc) == (ext.read_start, ext.read_end, ext.ref_start, ext.score)
    assert _words(cig) == [c.word for c in ext.cigar]
    if band:
        assert tch == (int(touch.touched), touch.row, int(touch.left), int(touch.clamped))


def _subset(cases, every):
    return cases[::every]


def test_band_blocks_model():
    for read, win, d, off, half in _subset(sweeps.band_blocks(), 7):
        _check_band(read, win, d, off, half)


def test_band_sweep_model():
    for read, win, d, off, half in _subset([c for c in sweeps.band_sweep() if len(c[1]) <= 300], 5):
        _check_band(read, win, d, off, half)


def test_band_interior_model():
    for read, win, d, off, half in _subset([c for c in sweeps.band_interior() if len(c[0]) <= 200], 6):
        _check_band(read, win, d, off, half)


def test_full_model():
    cases = [c for c in sweeps.full_blocks()][::9] + [c for c in sweeps.full_dispatch() if len(c[1]) <= 300]
    cases += [c for c in sweeps.full_sweep() if len(c[1]) <= 257 and len(c[0]) <= 151][::11]
    for read, win, d, off, half in cases:
        _check_band(read, win, d, off, None)


def test_hint_model():
    for read, win, d, off, _ in sweeps.hint_cases()[:50] + sweeps.hint_cases()[-5:]:
        p = ExtendParams(edit_dir=d)
        ext, path = oracle.extend_hint(read, win, p, DiagHint(off), with_path=True)
        rs, re, fs, sc, cig, _, gpath = wave_model.wave_pair(read, win, p, "hint", off)
        assert (rs, re, fs, sc) == (ext.read_start, ext.read_end, ext.ref_start, ext.score)
        assert _words(cig) == [c.word for c in ext.cigar]
        assert gpath == path


def test_i16_route_model():
    for read, win, p in sweeps.i16_wrap_cases():
        if len(read) > 100:
            continue
        _check_band(read, win, p.edit_dir, 0, None, p)
        _check_band(read, win, p.edit_dir, len(win) - len(read), 8, p)


@pytest.mark.parametrize("m", [63, 64, 65, 128, 129, 190])
def test_block_boundaries_model(m):
    """Reads ending just before, at and after a block of 64 rows, on banded
    and full-matrix windows, with a gap straddling the block boundary."""
    from seqgen import Rng, rand_seq
    rng = Rng(0xB10C + m)
    read = rand_seq(rng, m)
    win = rand_seq(rng, 20) + read[:60] + rand_seq(rng, 3) + read[60:] + rand_seq(rng, 25)
    for d in sweeps.DIRS:
        _check_band(read, win, d, 20, 12)
        _check_band(read, win, d, 0, None)
    # the read on the band's left and right edge diagonals: the path runs
    # through the first and last cell of every row, across the block
    # boundary (the first lane's diagonal and row-above reads)
    for diag in (8, 32):
        w = rand_seq(rng, diag) + read + rand_seq(rng, 40)
        _check_band(read, w, sweeps.DIRS[0], 20, 12)
