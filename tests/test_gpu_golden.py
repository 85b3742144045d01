"""The HIP path against the committed golden fixtures, through the C ABI.

Every fixture of tests/golden/ runs twice: batched (esp_gotoh_batch /
esp_myers_batch, one call per (mode, scoring) group) and as plain
one-launch-per-call entry points (esp_band / esp_extend / esp_extend_hint,
esp_myers_best_end / esp_myers_best_start).  Every field must equal the
fixture: score, coordinates, CIGAR words, TouchRec or extend_hint path,
Myers distance, end and start.
"""
from __future__ import annotations

import sys

import numpy as np
import pytest

from esperanto_amd.records import DiagHint
import golden_cases as G

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


def test_golden_extend_batched(buf):
    from esperanto_amd.batch import DeviceBatch
    for (mode, prm, d), cases in G.groups(G.extend_cases()).items():
        say(f"golden batched mode={mode} params={prm} dir={d} pairs={len(cases)}")
        db = DeviceBatch(G.batch_of(cases), buf)
        db.launch_gotoh(cases[0]["prm"], mode)
        db.sync()
        ext, touch, cig, off, n, aux = db.gotoh_results()
        G.check_batch(cases, ext, touch, cig, off, n, aux, "gpu-batched")
        db.free()


def test_golden_extend_plain(buf):
    from esperanto_amd import extend as X
    cases = G.extend_cases()
    say(f"golden plain calls: {len(cases)}")
    for k, c in enumerate(cases):
        ctx = (k, c["mode"], c["params"], c["dir"], len(c["read"]), len(c["window"]))
        if c["mode"] == 2:
            e = X.extend_hint(c["read"], c["window"], c["prm"], buf, DiagHint(c["offset"]))
        elif c["mode"] == 1:
            e = X.extend(c["read"], c["window"], c["prm"], buf)
        else:
            e, t = X.band(c["read"], c["window"], c["prm"], (c["offset"], c["half"]), buf)
            assert [int(t.touched), t.row, int(t.left), int(t.clamped)] == c["touch"], ctx
        assert [e.read_start, e.read_end, e.ref_start, e.score] == c["ext"], ctx
        assert [int(op.word) for op in e.cigar] == c["cigar"], ctx


def test_golden_myers(buf):
    from esperanto_amd import myers_ea
    from esperanto_amd.batch import DeviceBatch
    cases = G.myers_cases()
    say(f"golden myers pairs={len(cases)}")
    db = DeviceBatch(G.myers_batch_of(cases), buf)
    db.launch_myers(True)
    db.sync()
    d, e, s = db.myers_results()
    wd, we, ws = G.myers_expected(cases)
    assert np.array_equal(d, wd) and np.array_equal(e, we) and np.array_equal(s, ws)
    db.free()
    for c in cases:
        assert myers_ea.infix_best_end(c["read"], c["text"], buf) == (c["dist"], c["end"]), c
        assert myers_ea.infix_best_start(c["read"], c["text"], buf) == (c["dist"], c["start"]), c
