"""Random chains for the fast-lane tests: reads copied along one diagonal of
a contig (reads overhanging either contig end included) with substitutions,
N, lower-case bytes and A>G / T>C edits, anchors every few bases on that
diagonal, and chains that break each acceptance condition (anchors off the
diagonal, an anchor gap above k + w + 5, uncovered read ends)."""
from __future__ import annotations

import numpy as np

from esperanto_amd.reference import Reference


def fast_reference(seed: int):
    """Two contigs (3000 and 700 bp) with N runs and soft-masked runs."""
    rng = np.random.default_rng(seed)
    seqs = []
    for L in (3000, 700):
        s = bytearray(rng.choice(list(b"ACGT"), L).tolist())
        for _ in range(3):
            a, r = int(rng.integers(0, L)), int(rng.integers(1, 20))
            s[a:a + r] = b"N" * len(s[a:a + r])
        for _ in range(2):
            a, r = int(rng.integers(0, L)), int(rng.integers(1, 50))
            s[a:a + r] = s[a:a + r].lower()
        seqs.append(bytes(s))
    return Reference.from_sequences(seqs), seqs


def random_chain(rng, seqs, k: int = 15, w: int = 5):
    """(query, contig, anchors as (qpos, rpos)) -- mostly acceptable chains."""
    contig = int(rng.integers(0, len(seqs)))
    ref = seqs[contig].upper()
    m = int(rng.integers(20, 200))
    diag = int(rng.integers(-10, len(ref) - m + 10))
    q = bytearray()
    for i in range(m):
        p = i + diag
        q.append(ref[p] if 0 <= p < len(ref) else int(rng.choice(list(b"ACGT"))))
    mode = rng.random()
    rate = 0.0 if mode < 0.5 else 0.01 if mode < 0.8 else 0.05
    for i in range(m):
        r = rng.random()
        if r < rate:
            q[i] = int(rng.choice(list(b"ACGT")))
        elif r < rate + 0.01:
            q[i] = ord("N")
        elif r < rate + 0.015:
            q[i] = ord(chr(q[i]).lower())
        elif r < rate + 0.025:
            q[i] = {ord("A"): ord("G"), ord("T"): ord("C")}.get(q[i], q[i])
    # anchors: qpos from 0..2 to m - k - (0..2), steps of 1..k + w + 5
    qs = int(rng.integers(0, 3))
    qe_last = m - k - int(rng.integers(0, 3))
    anchors, qp = [], qs
    while qp <= qe_last:
        anchors.append((qp, qp + diag))
        qp += int(rng.integers(1, k + w + 6))
    if anchors and anchors[-1][0] != qe_last and qe_last >= qs:
        anchors.append((qe_last, qe_last + diag))
    breaker = rng.random()
    if anchors and breaker < 0.05:
        j = int(rng.integers(0, len(anchors)))
        anchors[j] = (anchors[j][0], anchors[j][1] + 1)  # off the diagonal
    elif len(anchors) > 2 and breaker < 0.10:
        anchors.pop(len(anchors) // 2)  # may open a gap above k + w + 5
    elif anchors and breaker < 0.13:
        anchors.pop(0)  # uncovered start
    anchors = [(a, r) for a, r in anchors if r >= 0]
    anchors.sort(key=lambda x: (x[1], x[0]))
    return bytes(q), contig, anchors
