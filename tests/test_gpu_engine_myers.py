"""Parity of the transcriptome engine's EA-Myers on the device
(esp_engine_myers_batch, esp_engine_verify_batch) with the oracle's
restatement of crates/engine/src/myers.rs, through the C ABI: infix and
global distances on patterns of 1..256 bases on both sides of the 64- and
128-base word boundaries, with lower-case, N and IUPAC bytes, A>G / T>C
edits, empty texts; verify_dist's window geometry (crates/engine/src/lib.rs:
851-876) with anchors before the transcript start, past its end and empty
windows; the 1..256 length guard; and 200,000 pairs in one launch against a
sample."""
from __future__ import annotations

import sys

import numpy as np
import pytest

from esperanto_amd._native import EspError
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


def _pairs(rng, count):
    out = []
    lens = [1, 2, 63, 64, 65, 127, 128, 129, 191, 192, 193, 255, 256]
    for j in range(count):
        m = lens[j % len(lens)] if j < 3 * len(lens) else int(rng.integers(1, 257))
        pat = bytearray(rng.choice(list(b"ACGTACGTacgtNRY"), m).tolist())
        text = bytearray()
        text += bytes(rng.choice(list(b"ACGT"), int(rng.integers(0, 30))).tolist())
        for b in pat:
            r = rng.random()
            if r < 0.03:
                continue
            if r < 0.06:
                text.append(int(rng.choice(list(b"ACGT"))))
            text.append({71: 65, 67: 84}.get(b, b) if rng.random() < 0.05 else b)
        text += bytes(rng.choice(list(b"ACGTN"), int(rng.integers(0, 30))).tolist())
        if j % 97 == 0:
            text = bytearray()
        out.append((bytes(pat), bytes(text)))
    return out


@pytest.mark.parametrize("mode", [0, 1])
def test_engine_myers_matches_oracle(buf, mode):
    from esperanto_amd.engine_myers import engine_myers_batch
    rng = np.random.default_rng(7 + mode)
    pairs = _pairs(rng, 4000)
    say(f"engine myers mode {mode}: {len(pairs)} pairs")
    got = engine_myers_batch(pairs, mode, buf)
    for j, (p, t) in enumerate(pairs):
        assert int(got[j]) == oracle.engine_myers(p, t, mode == 1), (j, len(p), len(t))


def test_engine_myers_length_guard(buf):
    from esperanto_amd.engine_myers import engine_myers_batch, infix
    with pytest.raises(EspError):
        engine_myers_batch([(b"ACGT", b"ACGT"), (b"A" * 257, b"ACGT")], 0, buf)
    with pytest.raises(EspError):
        engine_myers_batch([(b"", b"ACGT")], 1, buf)
    assert infix(b"ACGT", b"TTACGTT", buf) == 0


def _verify_ref(read, tx, anchor, flank):
    wlo = max(anchor - flank, 0)
    whi = min(max(anchor + len(read) + flank, 0), len(tx))
    if whi <= wlo:
        return None
    return oracle.engine_myers(read, tx[wlo:whi], False)


def test_engine_verify_dist(buf):
    from esperanto_amd.engine_myers import verify_dist_batch
    rng = np.random.default_rng(21)
    txs = [bytes(rng.choice(list(b"ACGTacgtN"), int(L)).tolist()) for L in rng.integers(1, 3000, 40)]
    reads, ids, anchors = [], [], []
    for j in range(5000):
        t = int(rng.integers(0, len(txs)))
        tx = txs[t]
        m = int(rng.integers(1, 257))
        a = int(rng.integers(-m - 200, len(tx) + 200))
        if 0 <= a and a + m <= len(tx) and rng.random() < 0.7:
            read = bytearray(tx[a:a + m])
            for i in range(len(read)):
                if rng.random() < 0.03:
                    read[i] = int(rng.choice(list(b"ACGTN")))
        else:
            read = bytearray(rng.choice(list(b"ACGT"), m).tolist())
        reads.append(bytes(read)), ids.append(t), anchors.append(a)
    for flank in (0, 30, 100):
        say(f"verify_dist {len(reads)} candidates, flank {flank}")
        dist, status = verify_dist_batch(reads, txs, ids, anchors, flank, buf)
        for j in range(len(reads)):
            exp = _verify_ref(reads[j], txs[ids[j]], anchors[j], flank)
            if exp is None:
                assert status[j] == 0, j
            else:
                assert status[j] == 1 and int(dist[j]) == exp, (j, flank)


def test_engine_myers_large_batch(buf):
    from esperanto_amd.engine_myers import engine_myers_batch
    rng = np.random.default_rng(3)
    n, m = 200_000, 150
    reads = np.frombuffer(b"ACGT", dtype=np.uint8)[rng.integers(0, 4, (n, m))]
    texts = np.frombuffer(b"ACGT", dtype=np.uint8)[rng.integers(0, 4, (n, m + 60))]
    texts[:, 30:30 + m] = reads
    hit = rng.random(texts.shape) < 0.02
    texts[hit] = np.frombuffer(b"ACGTN", dtype=np.uint8)[rng.integers(0, 5, int(hit.sum()))]
    pairs = [(reads[j].tobytes(), texts[j].tobytes()) for j in range(n)]
    for mode in (0, 1):
        say(f"engine myers {n} pairs mode {mode}")
        got = engine_myers_batch(pairs, mode, buf)
        for j in rng.choice(n, 1500, replace=False).tolist():
            assert int(got[j]) == oracle.engine_myers(pairs[j][0], pairs[j][1], mode == 1), j
