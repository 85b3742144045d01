"""The restated AVX2 kernels (avx2_full, avx2_band: crates/map/src/extend.rs:
1094-2026) against the restated scalar run_band<i16>, field for field
(score, coordinates, CIGAR, TouchRec), on the reference's own seeded
equivalence sweeps -- the same comparison its tests make
(avx2_full_matrix_matches_scalar :2578, avx2_full_matrix_block_boundaries
:2632, avx2_dispatch_path_matches_scalar :2665, avx2_band_matches_scalar
:2723, avx2_band_block_boundaries :2828, avx2_band_dispatch_path_matches_scalar
:2869) -- plus the inputs where the i16 route passes 32767."""
from __future__ import annotations

import pytest

from esperanto_amd.records import ExtendParams
from oracle import oracle
import sweeps

pytestmark = pytest.mark.skipif(not oracle.avx2_available(), reason="host CPU without AVX2")


def _both(read, win, p, band):
    try:
        oracle.set_route(1)
        a = oracle.band(read, win, p, band)
    finally:
        oracle.set_route(0)
    return a, oracle.band(read, win, p, band)


@pytest.mark.parametrize("sweep", ["full_sweep", "full_blocks", "full_dispatch", "band_sweep", "band_blocks",
                                   "band_dispatch"])
def test_avx2_equals_scalar_on_reference_sweeps(sweep):
    n = 0
    for read, win, d, off, half in getattr(sweeps, sweep)():
        band = None if half is None else (off, half)
        a, s = _both(read, win, ExtendParams(edit_dir=d), band)
        assert a == s, (sweep, read, win, d, band)
        n += 1
    assert n >= 12


def test_avx2_equals_scalar_past_the_i16_range():
    for read, win, p in sweeps.i16_wrap_cases():
        for band in (None, (5, 30), (0, 8)):
            a, s = _both(read, win, p, band)
            assert a == s, (read, win, p, band)


def test_avx2_long_windows():
    cases, longs = sweeps.full_sweep(include_long=True)
    for read, win, d, off, half in longs:
        a, s = _both(read, win, ExtendParams(edit_dir=d), None)
        assert a == s
