"""The reference's i16 accumulator past the i16 range, on the device.

run_dispatch (crates/map/src/extend.rs:293-344) runs pure inputs (bytes in
{A, C, G, T, N}) through the i16 kernels, whose wrapping add clamped at
-30000 (:473-501) drops every diagonal candidate above 32767; other inputs
run in i32.  With scaled scores or long reads inside the accepted envelope
both routes are reachable, so the device must reproduce each exactly: the
pure pairs below and their lowercase copies, in band, full and extend_hint
modes, batched and through the plain one-call-per-pair entry points.
"""
from __future__ import annotations

import sys

import numpy as np
import pytest

from esperanto_amd.records import DiagHint
from oracle import oracle
import sweeps

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-test] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def buf():
    from esperanto_amd.extend import ExtendBuffer
    b = ExtendBuffer(0)
    yield b
    b.close()


def _cases():
    out = []
    for read, win, p in sweeps.i16_wrap_cases():
        out.append((read, win, p))
        out.append((read.lower(), win, p))
    return out


@pytest.mark.parametrize("mode,offset,half", [(0, 5, 30), (0, 0, 8), (1, 0, 0), (2, 5, 0)],
                         ids=["band30", "band8", "full", "hint"])
def test_i16_route_batched(buf, mode, offset, half):
    from esperanto_amd.batch import DeviceBatch, PairBatch
    by_params = {}
    for read, win, p in _cases():
        by_params.setdefault(p, []).append((read, win, offset, half))
    wrapped = 0
    for p, pairs in by_params.items():
        hb = PairBatch.from_pairs(pairs)
        say(f"i16 route mode={mode} band=({offset},{half}) params={p} pairs={hb.npairs}")
        db = DeviceBatch(hb, buf)
        db.launch_gotoh(p, mode)
        db.sync()
        g_ext, g_touch, g_cig, off, g_n, g_aux = db.gotoh_results()
        o_ext, o_touch, o_cig, _, o_n, o_aux = oracle.band_batch(hb, p, mode, nthreads=8)
        for k in range(hb.npairs):
            ctx = (mode, p, k, len(pairs[k][0]), len(pairs[k][1]))
            assert tuple(g_ext[k]) == tuple(o_ext[k]), ctx
            a = int(off[k])
            assert g_n[k] == o_n[k], ctx
            assert np.array_equal(g_cig[a:a + g_n[k]], o_cig[a:a + o_n[k]]), ctx
            if mode != 2:
                assert tuple(g_touch[k]) == tuple(o_touch[k]), ctx
            else:
                assert g_aux[k] == o_aux[k], ctx
        # the pure copy and the lowercase copy differ where the i16 route wrapped
        wrapped += sum(1 for k in range(0, hb.npairs, 2) if o_ext[k][3] != o_ext[k + 1][3])
        db.free()
    assert wrapped > 0


def test_i16_route_plain(buf):
    from esperanto_amd import extend as X
    say("i16 route plain entry points")
    for read, win, p in _cases():
        if len(read) > 100:
            continue
        assert X.extend(read, win, p, buf) == oracle.extend(read, win, p), (read, win, p)
        assert X.extend_hint(read, win, p, buf, DiagHint(5)) == oracle.extend_hint(read, win, p, DiagHint(5))
        assert X.band(read, win, p, (5, 30), buf) == oracle.band(read, win, p, (5, 30))


def test_i16_route_known_values(buf):
    """A 40-nt read inside a 50-nt window at match 1024 and 900."""
    from esperanto_amd import extend as X
    read = b"ACGTTGCAAG" * 4
    win = b"TTCAA" + read + b"GGTCA"
    for match in (900, 1024):
        from esperanto_amd.records import ExtendParams
        p = ExtendParams(match, -4, 4, 2)
        g = X.extend(read, win, p, buf)
        assert g == oracle.extend(read, win, p)
        assert g.score < 32768
        assert X.extend(read.lower(), win, p, buf).score == 40 * match
