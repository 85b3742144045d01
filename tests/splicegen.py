"""Planted splice junctions for the junction-refinement, align_spliced and
reextend_with tests: one contig with a spliced read per locus, chains of
exact k-mer anchors on both exons, and the cases that make prepare_spliced
refuse (one segment, short or non-canonical introns without library support,
reads whose anchors disagree with the motif)."""
from __future__ import annotations

import numpy as np

from esperanto_amd.reference import Reference

K = 15
MOTIF = {False: (b"GT", b"AG"), True: (b"CT", b"AC")}  # plus / minus transcript strand


def _mutate(rng, s: bytearray, rate: float):
    for i in range(len(s)):
        if rng.random() < rate:
            s[i] = b"ACGT"[(b"ACGT".index(s[i]) + int(rng.integers(1, 4))) % 4] if s[i] in b"ACGT" else s[i]


def _anchors(rng, read_len: int, u: int, left_r: int, right_r: int, stride: int):
    """Exact k-mers at qpos multiples of stride on each exon, in (rpos, qpos) order."""
    out = []
    for q in range(0, max(u - K, -1) + 1, stride):
        out.append((q, left_r + q))
    for q in range(u, read_len - K + 1, stride):
        out.append((q, right_r + (q - u)))
    return sorted(out, key=lambda a: (a[1], a[0]))


def spliced_workload(seed: int, nreads: int = 200, read_len: int = 150):
    """Returns (reference, chains, library quads, truth) where each chain is
    (read, contig, minus, anchors) and truth[c] the planted (start, end, kind)."""
    rng = np.random.default_rng(seed)
    span = 14000
    genome = bytearray(np.frombuffer(b"ACGT", dtype=np.uint8)[rng.integers(0, 4, span * nreads + 2000)].tobytes())
    chains, truth, quads = [], [], []
    for c in range(nreads):
        base = 1000 + c * span
        kind = ["canon", "canon", "canon", "canon", "short", "noncanon", "onesegment", "three"][c % 8]
        minus = bool(rng.integers(0, 2))
        L = int(np.exp(rng.uniform(np.log(60), np.log(5000)))) if kind not in ("short",) else int(rng.integers(25, 50))
        s = base + 400
        e = s + L
        d, a = MOTIF[minus]
        if kind != "noncanon":
            genome[s:s + 2] = d
            genome[e - 2:e] = a
        else:
            genome[s:s + 2] = b"CA"
            genome[e - 2:e] = b"TT"
        u = int(rng.integers(30, read_len - (70 if kind == "three" else 30)))
        if kind == "three":
            # a second intron inside the right exon
            u2 = int(rng.integers(u + 30, read_len - 25))
            L2 = int(rng.integers(80, 800))
            s2 = e + (u2 - u)
            e2 = s2 + L2
            genome[s2:s2 + 2] = d
            genome[e2 - 2:e2] = a
            read = bytearray(genome[s - u:s] + genome[e:s2] + genome[e2:e2 + (read_len - u2)])
            anchors = _anchors(rng, u2, u, s - u, e, 7) + [(q, e2 + (q - u2)) for q in range(u2, read_len - K + 1, 7)]
            anchors = sorted(anchors, key=lambda x: (x[1], x[0]))
        elif kind == "onesegment":
            read = bytearray(genome[s - u:s - u + read_len])
            anchors = [(q, s - u + q) for q in range(0, read_len - K + 1, 9)]
        else:
            read = bytearray(genome[s - u:s] + genome[e:e + (read_len - u)])
            anchors = _anchors(rng, read_len, u, s - u, e, 7 if c % 3 else 11)
        _mutate(rng, read, 0.01)
        # keep only anchors still exact after mutation (as a seeder would)
        rb = bytes(read)
        anchors = [(q, r) for q, r in anchors if rb[q:q + K] == bytes(genome[r:r + K])]
        if len(read) and rng.random() < 0.15:
            read = bytearray(rb.lower())
        chains.append((bytes(read), 0, minus, anchors))
        truth.append((s, e, kind))
        if kind in ("canon", "short", "noncanon", "three") and rng.random() < 0.5:
            quads.append((0, s, e, int(minus)))
            if rng.random() < 0.3:
                quads.append((0, s, e, int(minus)))  # a duplicate counts twice
        if rng.random() < 0.2:
            quads.append((0, s + 3, e + 3, int(not minus)))  # a decoy nearby
    ref = Reference.from_sequences([bytes(genome)])
    return ref, chains, quads, truth


def refine_queries(seed: int, chains, truth, n: int = 600):
    """refine_junction_read queries around the planted junctions: naive
    breakpoints shifted by small amounts, read gaps of -3..25, both strands,
    some on a contig out of range, some with a fallback junction."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        c = int(rng.integers(0, len(chains)))
        read, contig, minus, _ = chains[c]
        s, e, _ = truth[c]
        shift = int(rng.integers(-4, 5))
        gap = int(rng.integers(-3, 26))
        qa = int(rng.integers(1, max(len(read) - 1, 2)))
        qb = qa + gap
        ns = s + shift
        ne = e + shift + gap
        q = (c, 5 if i % 97 == 0 else contig, max(ns, 0), max(ne, 0), qa, max(qb, 0), bool(rng.integers(0, 2)))
        if i % 5 == 0:
            q = q + ((s + int(rng.integers(-3, 4)), e),)
        out.append(q)
    return out


def intron_chain_workload(seed: int, nreads: int = 300, read_len: int = 150, n_runs: bool = False):
    """Reads over planted two-exon loci with their anchor pools (qpos, rpos,
    contig, minus): exact k-mers on both exons at a stride of 3..7, a few
    noise anchors elsewhere (and on the other strand), plus unspliced reads
    (one dominant fragment) and reads with a short exon side.  With n_runs,
    runs of N are then written into the genome next to and inside the exons
    (the anchors stay as planted, so some seed extensions stop at an N or
    start beside one) and a fifth of the reads get lower-case stretches.
    Returns (reference, reads, anchor pools, library quads)."""
    rng = np.random.default_rng(seed)
    span = 9000
    genome = bytearray(np.frombuffer(b"ACGT", dtype=np.uint8)[rng.integers(0, 4, span * nreads + 4000)].tobytes())
    reads, pools, quads = [], [], []
    for c in range(nreads):
        base = 2000 + c * span
        kind = c % 6
        minus = bool(rng.integers(0, 2))
        L = int(np.exp(rng.uniform(np.log(60), np.log(6000))))
        s = base + 700
        e = s + L
        d, a = MOTIF[minus]
        genome[s:s + 2] = d
        genome[e - 2:e] = a
        # the fast path places balanced splits (no fragment over 55% of the read)
        u = int(rng.integers(68, 83)) if kind < 4 else int(rng.integers(8, 40))
        if kind == 5:  # unspliced
            read = bytearray(genome[s - u:s - u + read_len])
            exons = [(0, s - u, read_len)]
        else:
            read = bytearray(genome[s - u:s] + genome[e:e + (read_len - u)])
            exons = [(0, s - u, u), (u, e, read_len - u)]
        _mutate(rng, read, 0.01)
        if rng.random() < 0.1:  # a technical clip
            nn = int(rng.integers(1, 12))
            read[:nn] = b"N" * nn
        stride = int(rng.integers(3, 8))
        pool = []
        rb = bytes(read)
        for q0, r0, ln in exons:
            for off in range(int(rng.integers(0, stride)), ln - K + 1, stride):
                q, r = q0 + off, r0 + off
                if rb[q:q + K] == bytes(genome[r:r + K]):
                    pool.append((q, r, 0, int(minus)))
        for _ in range(int(rng.integers(0, 4))):  # noise anchors
            q = int(rng.integers(0, read_len - K))
            pool.append((q, int(rng.integers(0, len(genome) - K)), 0, int(rng.integers(0, 2))))
        order = rng.permutation(len(pool))
        pools.append([pool[i] for i in order])
        if n_runs and rng.random() < 0.2:
            lb = bytearray(rb)
            a0 = int(rng.integers(0, read_len - 20))
            lb[a0:a0 + 20] = bytes(lb[a0:a0 + 20]).lower()
            rb = bytes(lb)
        reads.append(rb)
        if rng.random() < 0.3:
            quads.append((0, s, e, int(minus)))
        if n_runs:
            for _ in range(int(rng.integers(1, 4))):
                at = int(rng.choice([s - u - 30, s - u, s - 40, e, e + read_len - u - 30]))
                at = max(0, min(len(genome) - 40, at + int(rng.integers(-8, 9))))
                ln = int(rng.integers(1, 25))
                genome[at:at + ln] = b"N" * ln
    ref = Reference.from_sequences([bytes(genome)])
    return ref, reads, pools, quads
