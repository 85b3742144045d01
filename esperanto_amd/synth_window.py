"""Seeded synthetic inputs for the spliced and mate-rescue window pipelines.

splice_rescue_workload: one uniform-ACGT contig (default 64 Mb); per read
pair, R1 spans an intron: split u ~ U{1..m-1}, intron length log-uniform in
[60, 50000] with GT at [s, s+2) and AG at [e-2, e) planted in the contig, R1
= contig[s-u, s) ++ contig[e, e+m-u).  Its spliced request is the two 300-nt
anchor windows [(s-300, s), (e, e+300)] with hint 300-u.  R2 is the reverse
complement of contig[p2, p2+m) at fragment length I ~ N(300, 50) (clipped
to [m+u, 600]), p2 = e + I - m - u; its mate-rescue request is (R2, contig
0, e).  Both reads carry 1% substitutions.  All positions come from numpy's
PCG64 stream seeded with `seed`.
"""
from __future__ import annotations

import numpy as np

from .reference import Reference
from .rescue import MateRequests
from .splice import SpliceRequests

_COMP = np.zeros(256, dtype=np.uint8)
for _a, _b in zip(b"ACGT", b"TGCA"):
    _COMP[_a] = _b


def _substitute(rng, reads: np.ndarray, rate: float) -> None:
    hit = rng.random(reads.shape) < rate
    alt = np.frombuffer(b"ACGT", dtype=np.uint8)[rng.integers(0, 4, int(hit.sum()))]
    reads[hit] = alt


def splice_rescue_workload(npairs: int, ref_len: int = 64_000_000, read_len: int = 150,
                           flank: int = 300, seed: int = 0xE5A00004, sub_rate: float = 0.01):
    """Returns (Reference, SpliceRequests, MateRequests, truth) where truth holds
    the planted s, e, u and R2 position per pair."""
    rng = np.random.default_rng(seed)
    m = read_len
    ref = np.frombuffer(b"ACGT", dtype=np.uint8)[rng.integers(0, 4, ref_len)].copy()
    u = rng.integers(1, m, npairs)
    intron = np.exp(rng.uniform(np.log(60), np.log(50_000), npairs)).astype(np.int64)
    margin = flank + 1000
    s = rng.integers(margin, ref_len - intron - margin - 1000)
    e = s + intron
    frag = np.clip(np.rint(rng.normal(300, 50, npairs)).astype(np.int64), m + u, 600)
    p2 = e + frag - m - u
    ref[s], ref[s + 1] = ord("G"), ord("T")
    ref[e - 2], ref[e - 1] = ord("A"), ord("G")
    t = np.arange(m, dtype=np.int64)[None, :]
    r1 = np.empty((npairs, m), dtype=np.uint8)
    r2 = np.empty((npairs, m), dtype=np.uint8)
    for lo in range(0, npairs, 65536):
        hi = min(npairs, lo + 65536)
        uu = u[lo:hi, None]
        idx = np.where(t < uu, s[lo:hi, None] - uu + t, e[lo:hi, None] + t - uu)
        r1[lo:hi] = ref[idx]
        fw = ref[p2[lo:hi, None] + t]
        r2[lo:hi] = _COMP[fw[:, ::-1]]
    _substitute(rng, r1, sub_rate)
    _substitute(rng, r2, sub_rate)
    reference = Reference.from_sequences([ref.tobytes()])
    rl = np.full(npairs, m, dtype=np.uint32)
    ro = (np.arange(npairs, dtype=np.uint64) * m)
    exons = np.stack([s - flank, s, e, e + flank], axis=1).astype(np.uint32).reshape(-1, 2)
    eo = (np.arange(npairs + 1, dtype=np.uint64) * 2)
    splice = SpliceRequests(r1.reshape(-1), ro, rl, np.zeros(npairs, dtype=np.uint32),
                            (flank - u).astype(np.int64), exons, eo)
    mate = MateRequests(r2.reshape(-1), ro.copy(), rl.copy(), np.zeros(npairs, dtype=np.uint32),
                        e.astype(np.uint32))
    truth = {"s": s, "e": e, "u": u, "p2": p2}
    return reference, splice, mate, truth


def serving_workload(n_requests: int, ref_len: int = 256_000_000, seed: int = 0xE5A00005,
                     rescue_fraction: float = 0.1, max_pairs: int = 4096, sub_rate: float = 0.01):
    """Requests for the serving loop on a uniform-ACGT contig of ref_len bases.

    Each request holds N ~ log-uniform[1, max_pairs] pairs with reads of
    U[50, 250] bases planted at uniform loci with 1% substitutions.  Extension
    requests (the rest) ask extend_hint over window [locus - 32, locus + m + 32)
    with hint 32; a rescue_fraction of requests are mate-rescue requests
    (mate position 0-300 bases before the planted locus, half of the reads
    reverse-complemented), whose reference-geometry windows span 1650-2250
    bases.  Returns (Reference, requests) with each request a dict of numpy
    arrays: kind, reads (concatenated), read_len, contig, and win_start /
    win_end / hint or mate_pos.
    """
    rng = np.random.default_rng(seed)
    ref = np.frombuffer(b"ACGT", dtype=np.uint8)[rng.integers(0, 4, ref_len)].copy()
    reference = Reference.from_sequences([ref.tobytes()])
    reqs = []
    for _ in range(n_requests):
        n = int(np.exp(rng.uniform(0.0, np.log(max_pairs))))
        n = max(1, min(max_pairs, n))
        rescue = rng.random() < rescue_fraction
        m = rng.integers(50, 251, n)
        loc = rng.integers(3000, ref_len - 3000, n)
        off = np.zeros(n + 1, dtype=np.int64)
        off[1:] = np.cumsum(m)
        idx = np.repeat(loc - off[:-1], m) + np.arange(off[-1])
        reads = ref[idx]
        if rescue:
            minus = rng.random(n) < 0.5
            for k in np.nonzero(minus)[0]:
                a, b = off[k], off[k + 1]
                reads[a:b] = _COMP[reads[a:b][::-1]]
        _substitute(rng, reads, sub_rate)
        req = {"kind": 1 if rescue else 0, "reads": reads, "read_len": m.astype(np.uint32),
               "contig": np.zeros(n, dtype=np.uint32)}
        if rescue:
            req["mate_pos"] = (loc - rng.integers(0, 301, n)).astype(np.uint32)
        else:
            req["win_start"] = (loc - 32).astype(np.uint32)
            req["win_end"] = (loc + m + 32).astype(np.uint32)
            req["hint"] = np.full(n, 32, dtype=np.int64)
        reqs.append(req)
    return reference, reqs


def placement_walk_workload(npairs: int, ref_len: int = 64_000_000, read_len: int = 150,
                            seed: int = 0xE5A00006, known_fraction: float = 0.5):
    """Placements of the pair workload above, as the walks see them after the
    DP: per pair, R1 spliced as uM (e-s)N (m-u)M at s-u, and R2 as mM at p2
    with its query in the alignment's (forward) frame.  Returns (Reference,
    Placements, junction triples) with the planted junctions of a
    known_fraction of the pairs in the library."""
    from .walk import Placements
    ref, sp, mt, truth = splice_rescue_workload(npairs, ref_len, read_len, seed=seed)
    m = read_len
    s, e, u, p2 = (truth[k].astype(np.int64) for k in ("s", "e", "u", "p2"))
    n = 2 * npairs
    r2 = mt.reads.reshape(npairs, m)
    q = np.empty((n, m), dtype=np.uint8)
    q[0::2] = sp.reads.reshape(npairs, m)
    q[1::2] = _COMP[r2[:, ::-1]]
    M, N = 0, 3
    cig = np.zeros((n, 3), dtype=np.uint32)
    cig[0::2, 0] = (u << 4 | M).astype(np.uint32)
    cig[0::2, 1] = ((e - s) << 4 | N).astype(np.uint32)
    cig[0::2, 2] = ((m - u) << 4 | M).astype(np.uint32)
    cig_n = np.tile(np.array([3, 1], dtype=np.uint32), npairs)
    flat = np.concatenate([cig[0::2], np.stack([np.full(npairs, m << 4 | M, dtype=np.uint32)], 1)],
                          axis=1).reshape(-1)
    cig_off = np.zeros(n, dtype=np.uint64)
    cig_off[1:] = np.cumsum(cig_n, dtype=np.uint64)[:-1]
    pos = np.empty(n, dtype=np.uint32)
    pos[0::2] = s - u
    pos[1::2] = p2
    pl = Placements(q.reshape(-1), np.arange(n, dtype=np.uint64) * m, np.full(n, m, dtype=np.uint32),
                    np.zeros(n, dtype=np.uint32), pos, flat, cig_off, cig_n, np.zeros(n, dtype=np.uint8))
    rng = np.random.default_rng(seed + 1)
    pick = rng.random(npairs) < known_fraction
    junctions = sorted(set(zip([0] * int(pick.sum()), s[pick].tolist(), e[pick].tolist())))
    return ref, pl, junctions


def junction_workload(nreads: int, ref_len: int = 64_000_000, read_len: int = 150, k: int = 15,
                      stride: int = 8, seed: int = 0xE5A00007, known_fraction: float = 0.5,
                      sub_rate: float = 0.01, clip: int = 12):
    """Inputs for align_spliced over chains, try_intron_chain_placement and
    reextend_with on one uniform-ACGT contig of ref_len bases.

    Per read an intron [s, e) of log-uniform length 60..50,000 with GT at
    [s, s+2) and AG at [e-2, e); split u ~ U{40..110}; read = contig[s-u, s)
    ++ contig[e, e+m-u) with sub_rate substitutions.  Its chain anchors are
    the exact k-mers at qpos 0, stride, 2*stride, ... lying wholly on one exon
    with no substitution in them, in (rpos, qpos) order; its intron-chain pool
    is the same anchors as (qpos, rpos, contig 0, plus).  known_fraction of
    the introns go into the junction library (plus strand).  Re-extension
    placements: a second read = contig[p, p+m) (substituted) at pos p + clip
    with CIGAR clip S, m - 2 clip M, clip S.  Returns a dict of numpy arrays.
    """
    rng = np.random.default_rng(seed)
    m = read_len
    ref = np.frombuffer(b"ACGT", dtype=np.uint8)[rng.integers(0, 4, ref_len)].copy()
    u = rng.integers(40, 111, nreads)
    intron = np.exp(rng.uniform(np.log(60), np.log(50_000), nreads)).astype(np.int64)
    s = rng.integers(2000, ref_len - intron - 2000)
    e = s + intron
    ref[s], ref[s + 1] = ord("G"), ord("T")
    ref[e - 2], ref[e - 1] = ord("A"), ord("G")
    p2 = rng.integers(2000, ref_len - 2000 - m, nreads)
    t = np.arange(m, dtype=np.int64)[None, :]
    r1 = np.empty((nreads, m), dtype=np.uint8)
    r2 = np.empty((nreads, m), dtype=np.uint8)
    for lo in range(0, nreads, 65536):
        hi = min(nreads, lo + 65536)
        uu = u[lo:hi, None]
        r1[lo:hi] = ref[np.where(t < uu, s[lo:hi, None] - uu + t, e[lo:hi, None] + t - uu)]
        r2[lo:hi] = ref[p2[lo:hi, None] + t]
    hit = rng.random(r1.shape) < sub_rate
    r1[hit] = np.frombuffer(b"ACGT", dtype=np.uint8)[rng.integers(0, 4, int(hit.sum()))]
    _substitute(rng, r2, sub_rate)
    # anchors: k-mers on one exon without a substituted base (a substitution
    # may redraw the same base; such k-mers are dropped all the same)
    q = np.arange(0, m - k + 1, stride, dtype=np.int64)[None, :]
    bad = np.concatenate([np.zeros((nreads, 1), np.int64), np.cumsum(hit, axis=1)], axis=1)
    clean = (bad[:, q[0] + k] - bad[:, q[0]]) == 0
    one_exon = (q + k <= u[:, None]) | (q >= u[:, None])
    keep = clean & one_exon
    rpos = np.where(q < u[:, None], s[:, None] - u[:, None] + q, e[:, None] + q - u[:, None])
    na = keep.sum(axis=1).astype(np.uint64)
    ao = np.zeros(nreads + 1, dtype=np.uint64)
    ao[1:] = np.cumsum(na)
    qq = np.broadcast_to(q, keep.shape)[keep].astype(np.uint32)
    rr = rpos[keep].astype(np.uint32)
    chain_flat = np.stack([qq, rr], axis=1).reshape(-1)
    zeros = np.zeros_like(qq)
    pool_flat = np.stack([qq, rr, zeros, zeros], axis=1).reshape(-1)
    known = rng.random(nreads) < known_fraction
    quads = np.stack([np.zeros(int(known.sum()), np.uint32), s[known].astype(np.uint32),
                      e[known].astype(np.uint32), np.zeros(int(known.sum()), np.uint32)], axis=1)
    rl = np.full(nreads, m, dtype=np.uint32)
    ro = np.arange(nreads, dtype=np.uint64) * m
    words = np.array([(clip << 4) | 4, ((m - 2 * clip) << 4) | 0, (clip << 4) | 4], dtype=np.uint32)
    return {"reference": Reference.from_sequences([ref.tobytes()]),
            "reads": r1.reshape(-1), "read_off": ro, "read_len": rl,
            "contig": np.zeros(nreads, np.uint32), "minus": np.zeros(nreads, np.uint32),
            "chain_anchors": chain_flat, "chain_off": ao, "pool_anchors": pool_flat, "pool_off": ao.copy(),
            "quads": quads, "s": s, "e": e, "u": u,
            "rx_reads": r2.reshape(-1), "rx_pos": (p2 + clip).astype(np.uint32),
            "rx_cigar": np.tile(words, nreads), "rx_cig_off": np.arange(nreads, dtype=np.uint64) * 3,
            "rx_cig_n": np.full(nreads, 3, dtype=np.uint32), "k": k}
