"""esp_window_myers_batch_noalloc against esp_window_myers_batch, through the
C ABI: the same jobs give the same distances, positions and statuses in every
mode, on two calls in a row whose second batch is smaller (fewer jobs,
shorter texts) so that it runs over the first call's leftovers in the shared
device scratch; a batch larger than that scratch is refused before anything
runs."""
from __future__ import annotations

import ctypes
import sys

import numpy as np
import pytest

from esperanto_amd import _native
from esperanto_amd.reference import Reference

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


def _jobs(rng, seqs, nreads, per_read, max_span):
    reads = []
    for _ in range(nreads):
        m = int(rng.integers(20, 300))
        P = int(rng.integers(0, len(seqs[0]) - m))
        r = bytearray(seqs[0][P:P + m])
        for _ in range(int(rng.integers(0, 6))):
            r[int(rng.integers(0, m))] = b"ACGTNacgt"[int(rng.integers(0, 9))]
        reads.append(bytes(r))
    jobs = []
    for k, r in enumerate(reads):
        for _ in range(per_read):
            qs = int(rng.integers(0, len(r)))
            qe = int(rng.integers(qs, min(len(r), qs + 260) + 1))
            c = int(rng.integers(0, len(seqs)))
            L = len(seqs[c])
            ts = int(rng.integers(0, L))
            te = min(L, ts + int(rng.integers(0, max_span)))
            jobs.append((k, qs, qe, c, ts, te, 0))
    jobs.append((0, 0, 10, len(seqs) + 3, 0, 10, 0))  # no such contig
    return reads, jobs


def _run(buf, dref, reads, jobs, mode, entry):
    from esperanto_amd.batch import DeviceArray
    from esperanto_amd.intron_chain import WINDOW_JOB_DTYPE
    J = np.array(jobs, dtype=WINDOW_JOB_DTYPE)
    lens = np.array([len(r) for r in reads], dtype=np.uint64)
    offs = np.zeros(len(reads), dtype=np.uint64)
    offs[1:] = np.cumsum(lens)[:-1]
    J["read_off"] = offs[J["read_off"].astype(np.int64)]
    rb = np.frombuffer(b"".join(reads), dtype=np.uint8)
    maxp = int((J["qend"] - J["qstart"]).max())
    maxt = int((J["tend"].astype(np.int64) - J["tstart"].astype(np.int64)).clip(0).max())
    n = J.shape[0]
    arrs = [DeviceArray.from_numpy(buf, J), DeviceArray.from_numpy(buf, rb), DeviceArray(buf, 4 * n),
            DeviceArray(buf, 8 * n), DeviceArray(buf, 4 * n)]
    d_jobs, d_reads, d_dist, d_pos, d_st = arrs
    try:
        fn = getattr(_native.lib(), entry)
        _native.check(fn(buf.handle, dref.handle, ctypes.c_void_p(d_jobs.ptr), n, ctypes.c_void_p(d_reads.ptr),
                         max(maxp, 1), max(maxt, 1), mode, ctypes.c_void_p(d_dist.ptr),
                         ctypes.c_void_p(d_pos.ptr), ctypes.c_void_p(d_st.ptr)), buf.handle)
        _native.check(_native.lib().esp_sync(buf.handle), buf.handle)
        return d_dist.to_numpy(np.int32, n), d_pos.to_numpy(np.uint64, n), d_st.to_numpy(np.int32, n)
    finally:
        for a in arrs:
            a.free()


def test_noalloc_window_myers_equals_allocating_call(buf):
    from esperanto_amd.reference import DeviceReference
    from oracle import oracle
    ref = Reference.synthetic(77, [40000, 900], n_runs=6, n_run_max=150)
    o = oracle.OracleRef(ref)
    seqs = [oracle.decode(o, k, 0, int(L)) for k, L in enumerate(ref.lens)]
    dref = DeviceReference(ref, buf)
    rng = np.random.default_rng(31)
    try:
        first = _jobs(rng, seqs, 300, 3, 400)
        second = _jobs(rng, seqs, 60, 2, 120)  # fewer jobs, shorter texts
        for mode in (0, 1, 2):
            for label, (reads, jobs) in (("first", first), ("second", second)):
                say(f"noalloc window myers mode {mode} {label}: {len(jobs)} jobs")
                want = _run(buf, dref, reads, jobs, mode, "esp_window_myers_batch")
                got = _run(buf, dref, reads, jobs, mode, "esp_window_myers_batch_noalloc")
                assert np.array_equal(got[0], want[0]), (mode, label)
                assert np.array_equal(got[2], want[2]), (mode, label)
                if mode != 2:
                    assert np.array_equal(got[1], want[1]), (mode, label)
                assert want[2][-1] != 0 and (want[2][:-1] == 0).all()
    finally:
        dref.close()


def test_noalloc_refuses_a_batch_larger_than_its_scratch(buf):
    from esperanto_amd.batch import DeviceArray
    from esperanto_amd.intron_chain import WINDOW_JOB_DTYPE
    from esperanto_amd.reference import DeviceReference
    ref = Reference.synthetic(78, [5000], n_runs=2, n_run_max=50)
    dref = DeviceReference(ref, buf)
    n = 1000
    J = np.zeros(n, dtype=WINDOW_JOB_DTYPE)  # empty jobs: nothing is decoded even if run
    arrs = [DeviceArray.from_numpy(buf, J), DeviceArray.from_numpy(buf, np.zeros(16, np.uint8)),
            DeviceArray(buf, 4 * n), DeviceArray(buf, 8 * n), DeviceArray(buf, 4 * n)]
    try:
        say("noalloc window myers: 1000 jobs at max_text 40000 (about 40 MB of scratch)")
        rc = _native.lib().esp_window_myers_batch_noalloc(
            buf.handle, dref.handle, ctypes.c_void_p(arrs[0].ptr), n, ctypes.c_void_p(arrs[1].ptr), 1, 40000, 0,
            ctypes.c_void_p(arrs[2].ptr), ctypes.c_void_p(arrs[3].ptr), ctypes.c_void_p(arrs[4].ptr))
        assert rc == 1  # ESP_E_INVAL
        msg = _native.lib().esp_last_error(buf.handle).decode()
        assert "scratch" in msg, msg
        _native.check(_native.lib().esp_sync(buf.handle), buf.handle)
    finally:
        for a in arrs:
            a.free()
        dref.close()
