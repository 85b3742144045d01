"""CPU-side checks of the C ABI library: it loads, exports every symbol the
headers declare, its host-only entry points (stitch_cigar, the synthetic
generator) behave, and device entry points fail loudly instead of falling
back to a CPU path when no GPU is visible.  No kernel is launched here."""
from __future__ import annotations

import ctypes
import os
import re

import numpy as np
import pytest

from esperanto_amd import _native
from esperanto_amd.records import Del, Ins, Match, RefSkip, SoftClip

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(esp_[a-z0-9_]+)\s*\(", text)))


@pytest.mark.parametrize("header", ["esp_align.h", "esp_synth.h"])
def test_library_exports_every_declared_symbol(header):
    L = _native.lib()
    names = _declared(header)
    assert names
    for name in names:
        assert hasattr(L, name), name


def test_python_binding_covers_the_abi():
    assert set(_native.EXPORTS) <= set(_declared("esp_align.h"))


def _stitch(words, start, bounds, skips):
    L = _native.lib()
    flat = [x for b in bounds for x in b]
    ops = (ctypes.c_uint32 * max(len(words), 1))(*words)
    bnd = (ctypes.c_uint32 * max(len(flat), 1))(*flat)
    sk = (ctypes.c_uint32 * max(len(skips), 1))(*skips)
    out = (ctypes.c_uint32 * 64)()
    n = ctypes.c_size_t()
    rc = L.esp_stitch_cigar(ops, len(words), start, bnd, len(bounds), sk, len(skips), out, 64,
                            ctypes.byref(n))
    assert rc == 0
    return list(out[: n.value])


def test_stitch_cigar_reference_cases():
    """map/tests/splice_stitch.rs:42-142 through the C ABI."""
    w = lambda ops: [o.word for o in ops]  # noqa: E731
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
        assert _stitch(w(inp), 0, bounds, [100]) == w(exp)
    inp = [SoftClip(3), Match(10), Ins(1), Match(10), Del(2), Match(8), SoftClip(2)]
    b3 = [(0, 100, 110), (10, 210, 220), (20, 420, 440)]
    assert _stitch(w(inp), 0, b3, [100, 200]) == w([SoftClip(3), Match(10), Ins(1), RefSkip(100),
                                                    Match(10), RefSkip(200), Del(2), Match(8),
                                                    SoftClip(2)])
    for start, inp, exp in [
        (0, [Match(20)], [Match(10), RefSkip(100), Match(10)]),
        (8, [Del(4), Match(6)], [Del(2), RefSkip(100), Del(2), Match(6)]),
        (0, [Match(10), Ins(1)], [Match(10), Ins(1)]),
        (10, [Ins(1), Match(10)], [Ins(1), Match(10)]),
        (11, [Match(5)], [Match(5)]),
        (0, [Match(8)], [Match(8)]),
    ]:
        assert _stitch(w(inp), start, bounds, [100]) == w(exp)
    assert _stitch(w([Match(5), Ins(1), Match(5)]), 0, [], []) == w([Match(5), Ins(1), Match(5)])


def test_stitch_cigar_matches_oracle_randomised():
    from oracle import oracle
    rng = np.random.default_rng(7)
    for _ in range(300):
        nops = int(rng.integers(1, 8))
        words = [int(rng.integers(1, 12)) << 4 | int(rng.choice([0, 1, 2, 4])) for _ in range(nops)]
        ne = int(rng.integers(0, 4))
        bounds, ps = [], 0
        rs = 100
        for _ in range(ne):
            ln = int(rng.integers(1, 15))
            bounds.append((ps, rs, rs + ln))
            ps += ln
            rs += ln + int(rng.integers(50, 200))
        skips = [int(rng.integers(50, 200)) for _ in range(max(ne - 1, 0))]
        start = int(rng.integers(0, ps + 3)) if ne else 0
        got = _stitch(words, start, bounds, skips)
        exp = [c.word for c in oracle.stitch_cigar(words, start, bounds, skips)]
        assert got == exp, (words, start, bounds, skips)


def test_synthetic_workloads_are_deterministic_and_well_formed():
    from esperanto_amd.batch import workload_c1, workload_c2, workload_c3
    a, b = workload_c1(500), workload_c1(500)
    assert np.array_equal(a.reads, b.reads) and np.array_equal(a.wins, b.wins)
    assert set(np.unique(a.reads)) <= set(b"ACGT")
    assert (a.read_len == 100).all() and (a.win_len == 164).all()
    assert (a.offset == 32).all() and (a.half == 32).all()
    c2 = workload_c2(300)
    assert (c2.read_len == 150).all() and (c2.win_len == 214).all()
    c3 = workload_c3(2000)
    assert c3.read_len.min() >= 50 and c3.read_len.max() <= 250
    assert set(np.unique(c3.half)) <= {16, 32, 48, 64}
    assert np.array_equal(c3.win_len, c3.read_len + 2 * c3.half.astype(np.uint32))
    # C1: reads mostly match their locus (1% subs, 0.1% indels)
    r, w = a.pair(0)
    assert sum(x == y for x, y in zip(r[:50], w[32:82])) >= 40


def test_device_entry_points_fail_loudly_without_gpu():
    """esp_create either succeeds (a GPU is visible) or reports ESP_E_DEVICE;
    it never silently falls back to a CPU implementation."""
    L = _native.lib()
    h = ctypes.c_void_p()
    rc = L.esp_create(0, ctypes.byref(h))
    if rc == _native.ESP_OK:
        L.esp_destroy(h)
        pytest.skip("a GPU is visible")
    assert rc == _native.ESP_E_DEVICE
    from esperanto_amd.extend import ExtendBuffer
    with pytest.raises(_native.EspError):
        ExtendBuffer(0)
