"""Random placements for the placement-walk tests: CIGARs mixing M/I/D/N/S,
queries copied from the reference with substitutions, N and lowercase bytes,
positions up to past the contig end, queries shorter than their CIGAR."""
from __future__ import annotations

import numpy as np

from esperanto_amd.records import Del, Ins, Match, RefSkip, SoftClip
from esperanto_amd.reference import Reference


def words(ops):
    return [o.word for o in ops]


def random_placement(rng, ref: Reference, seq_bytes):
    """(query, contig, pos, BAM CIGAR words) near a true locus of seq_bytes."""
    contig = int(rng.integers(0, ref.ncontig))
    clen = int(ref.lens[contig])
    pos = int(rng.integers(0, clen + 5))
    ops, q = [], bytearray()
    rf = pos
    if rng.random() < 0.3:
        c = int(rng.integers(1, 8))
        ops.append(SoftClip(c))
        q += bytes(rng.choice(list(b"ACGTN"), c).tolist())
    for _ in range(int(rng.integers(1, 6))):
        kind = rng.random()
        if kind < 0.55:
            n = int(rng.integers(1, 60))
            ops.append(Match(n))
            s = bytearray(seq_bytes[contig][rf:rf + n].upper().ljust(n, b"A"))
            for t in range(n):
                r = rng.random()
                if r < 0.06:
                    s[t] = int(rng.choice(list(b"ACGT")))
                elif r < 0.08:
                    s[t] = ord("N")
                elif r < 0.10:
                    s[t] = ord(chr(s[t]).lower())
            q += s
            rf += n
        elif kind < 0.68:
            n = int(rng.integers(1, 6))
            ops.append(Ins(n))
            q += bytes(rng.choice(list(b"ACGT"), n).tolist())
        elif kind < 0.80:
            n = int(rng.integers(1, 6))
            ops.append(Del(n))
            rf += n
        else:
            n = int(rng.integers(0, 3000))
            ops.append(RefSkip(n))
            rf += n
    if rng.random() < 0.3:
        c = int(rng.integers(1, 8))
        ops.append(SoftClip(c))
        q += bytes(rng.choice(list(b"ACGT"), c).tolist())
    if rng.random() < 0.1 and len(q) > 3:
        q = q[: len(q) - int(rng.integers(1, 4))]  # query shorter than the CIGAR
    return bytes(q), contig, pos, words(ops)


def random_walk_reference(seed: int):
    """Three contigs (4000, 1500, 333 bp) with N runs and lower-case runs."""
    rng = np.random.default_rng(seed)
    seqs = []
    for L in (4000, 1500, 333):
        s = bytearray(rng.choice(list(b"ACGT"), L).tolist())
        for _ in range(4):
            a, r = int(rng.integers(0, L)), int(rng.integers(1, 30))
            s[a:a + r] = b"N" * len(s[a:a + r])
        for _ in range(3):  # soft-masked runs decode as upper case
            a, r = int(rng.integers(0, L)), int(rng.integers(1, 60))
            s[a:a + r] = s[a:a + r].lower()
        assert len(s) == L
        seqs.append(bytes(s))
    return Reference.from_sequences(seqs), seqs
