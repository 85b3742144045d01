"""Serving loop (esp_server_*) on the GPU.

- a burst of mixed extension / mate-rescue requests of uneven sizes on the
  serving workload's 256 Mb reference, batched across requests by three
  workers, matches the oracle pair for pair
  (extend_hint over the decoded windows; mate_window_placement);
- every request resubmitted alone returns byte-identical records and CIGARs
  (the result of a pair does not depend on what shares its batch);
- requests larger than the device-memory budget are refused with
  ESP_E_NOMEM, and a full queue refuses with ESP_E_BUSY while every accepted
  request still completes.
"""
from __future__ import annotations

import sys

import numpy as np
import pytest

from esperanto_amd import _native
from esperanto_amd.batch import PairBatch
from esperanto_amd.records import ExtendParams
from esperanto_amd.rescue import MateRequests
from esperanto_amd.synth_window import serving_workload
from oracle import oracle

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


def submit(srv, r):
    return srv.submit_arrays(r["kind"], r["reads"], r["read_len"], r["contig"], r.get("win_start"),
                             r.get("win_end"), r.get("hint"), r.get("mate_pos"))


def oracle_result(o, r, params):
    n = r["read_len"].shape[0]
    off = np.zeros(n + 1, dtype=np.int64)
    off[1:] = np.cumsum(r["read_len"])
    reads = [r["reads"][off[k]:off[k + 1]].tobytes() for k in range(n)]
    if r["kind"] == 0:
        wins = [oracle.decode(o, int(r["contig"][k]), int(r["win_start"][k]), int(r["win_end"][k]))
                for k in range(n)]
        pb = PairBatch.from_pairs([(reads[k], wins[k], int(r["hint"][k]), 0) for k in range(n)])
        ext, _, cig, coff, cig_n, _ = oracle.band_batch(pb, params, 2, nthreads=8)
        return ext, cig, coff, cig_n
    req = MateRequests(r["reads"], off[:-1].astype(np.uint64), r["read_len"], r["contig"], r["mate_pos"])
    rec, cig, coff, cig_n = oracle.mate_batch(o, req, params, nthreads=8)
    return rec, cig, coff, cig_n


def assert_same(t, exp, kind, ctx):
    rec, cig, coff, cig_n = t.raw()
    erec, ecig, ecoff, ecig_n = exp
    assert np.array_equal(np.asarray(coff, dtype=np.uint64), np.asarray(ecoff, dtype=np.uint64)), ctx
    assert np.array_equal(cig_n, ecig_n), ctx
    if kind == 0:
        got = rec.astype(np.int64)
        got[:, 3] = rec[:, 3].view(np.int32)
        assert np.array_equal(got, erec), ctx
    else:
        assert [x.tolist() for x in rec] == [x.tolist() for x in erec], ctx
    for k in range(len(cig_n)):
        a = int(coff[k])
        assert np.array_equal(cig[a:a + int(cig_n[k])], ecig[a:a + int(ecig_n[k])]), (ctx, k)


def test_server_mix_matches_oracle_and_alone(buf):
    from esperanto_amd.reference import DeviceReference
    from esperanto_amd.server import Server
    params = ExtendParams()
    # the serving workload's 256 Mb reference resident in HBM
    say("serving workload on a 256 Mb reference")
    ref, reqs = serving_workload(90, ref_len=256_000_000, max_pairs=400, rescue_fraction=0.25)
    o = oracle.OracleRef(ref)
    dref = DeviceReference(ref, buf)
    srv = Server(dref, params, mem_budget=16 << 30, workers=3, max_batch_pairs=3000)
    say(f"server burst of {len(reqs)} requests")
    tickets = [submit(srv, r) for r in reqs]
    for t in tickets:
        info = t.wait()
        assert info.status == 0
    st = srv.stats()
    say(f"stats {st}")
    assert st["completed"] == len(reqs)
    assert st["max_batch_requests"] > 1 and st["batches"] < len(reqs)
    for j, (r, t) in enumerate(zip(reqs, tickets)):
        assert_same(t, oracle_result(o, r, params), r["kind"], ("mix", j))
    say("alone vs mix")
    for j in range(0, len(reqs), 4):
        t2 = submit(srv, reqs[j])
        t2.wait()
        a, b = tickets[j].raw(), t2.raw()
        for x, y in zip(a, b):
            assert np.array_equal(np.asarray(x), np.asarray(y)), ("alone", j)
    srv.close()
    dref.close()


def test_server_budget_and_backpressure(buf):
    from esperanto_amd.reference import DeviceReference
    from esperanto_amd.server import Server
    ref, reqs = serving_workload(60, ref_len=2_000_000, max_pairs=600, rescue_fraction=0.0, seed=7)
    dref = DeviceReference(ref, buf)
    reserve = (1 << 30) + (4 << 30) + (256 << 20)  # one worker's scratch
    srv = Server(dref, mem_budget=reserve + (2 << 20), workers=1, max_batch_pairs=4096, max_pending=3)
    big = max(reqs, key=lambda r: r["read_len"].shape[0])
    r = {k: (np.concatenate([v] * 8) if isinstance(v, np.ndarray) else v) for k, v in big.items()}
    with pytest.raises(_native.EspError) as e:  # ~2 KB per pair: over the 2 MB data budget
        submit(srv, r)
    assert e.value.code == _native.ESP_E_NOMEM
    say("back-pressure burst")
    accepted, busy = [], 0
    for r in reqs:
        try:
            accepted.append(submit(srv, r))
        except _native.EspError as err:
            assert err.code == _native.ESP_E_BUSY
            busy += 1
    for t in accepted:
        assert t.wait().status == 0
    say(f"accepted {len(accepted)} busy {busy}")
    assert busy > 0 and len(accepted) >= 3
    srv.close()
    dref.close()
