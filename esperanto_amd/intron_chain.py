"""Intron-chain fast path on the device: the EA-Myers breakpoint searches
(crates/map/src/intron_chain.rs:277-373) batched over many reads.

The geometry of each search (pattern window in the read, text window on the
contig) is the reference's own arithmetic, computed on the host; the searches
themselves run as one esp_window_myers_batch launch per kind (windows decoded
from the resident reference).  The fast path's final extension over the
two-exon pseudo-reference is `esperanto_amd.splice.spliced_batch` with the
hint `left.off_lo() - left_start` (intron_chain.rs:497-536).
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import List, Sequence, Tuple

import numpy as np

from . import _native
from .batch import DeviceArray
from .reference import DeviceReference

I32_MAX = 2**31 - 1

WINDOW_JOB_DTYPE = np.dtype([("read_off", "<u8"), ("qstart", "<u4"), ("qend", "<u4"),
                             ("contig", "<u4"), ("tstart", "<u4"), ("tend", "<u4"),
                             ("pad", "<u4")])
MODE_END, MODE_START, MODE_DIST = 0, 1, 2


@dataclass
class IntronParams:
    """intron_chain.rs:21-42"""

    min_frag_cov: int = 15
    refine_window: int = 30
    refine_pattern_len: int = 50
    intron_max: int = 105_000


def window_myers(dref: DeviceReference, reads: Sequence[bytes], jobs: np.ndarray, mode: int):
    """jobs: WINDOW_JOB_DTYPE rows whose read_off indexes `reads` (the job's read
    number, replaced here by its byte offset).  Returns (dist, pos, status)."""
    buf = dref.buf
    n = jobs.shape[0]
    if n == 0:
        return np.zeros(0, np.int32), np.zeros(0, np.uint64), np.zeros(0, np.int32)
    lens = np.array([len(r) for r in reads], dtype=np.uint64)
    offs = np.zeros(len(reads), dtype=np.uint64)
    if len(reads):
        offs[1:] = np.cumsum(lens)[:-1]
    J = jobs.copy()
    J["read_off"] = offs[jobs["read_off"].astype(np.int64)]
    rb = np.frombuffer(b"".join(reads) or b"\0", dtype=np.uint8)
    maxp = int((J["qend"] - J["qstart"]).max())
    maxt = int((J["tend"].astype(np.int64) - J["tstart"].astype(np.int64)).clip(0).max())
    d_jobs = DeviceArray.from_numpy(buf, J)
    d_reads = DeviceArray.from_numpy(buf, rb)
    d_dist = DeviceArray(buf, 4 * n)
    d_pos = DeviceArray(buf, 8 * n)
    d_st = DeviceArray(buf, 4 * n)
    try:
        L = _native.lib()
        _native.check(L.esp_window_myers_batch(buf.handle, dref.handle, ctypes.c_void_p(d_jobs.ptr), n,
                                               ctypes.c_void_p(d_reads.ptr), max(maxp, 1),
                                               max(maxt, 1), mode, ctypes.c_void_p(d_dist.ptr),
                                               ctypes.c_void_p(d_pos.ptr), ctypes.c_void_p(d_st.ptr)),
                      buf.handle)
        _native.check(L.esp_sync(buf.handle), buf.handle)
        return d_dist.to_numpy(np.int32, n), d_pos.to_numpy(np.uint64, n), d_st.to_numpy(np.int32, n)
    finally:
        for a in (d_jobs, d_reads, d_dist, d_pos, d_st):
            a.free()


def _job(k, qs, qe, contig, ts, te):
    return (k, qs, qe, contig, ts, te, 0)


def donor_plan(lens, reads: Sequence[bytes], contigs: Sequence[int], junction_off: Sequence[int],
               frag_ref_hi: Sequence[int], frag_implied_start: Sequence[int],
               params: IntronParams = IntronParams()):
    """The geometry of refine_donor_end (intron_chain.rs:285-309): (defaults,
    jobs, readers, text starts) -- defaults hold the answers that need no
    search; job j answers read readers[j] as text_start[j] + best end."""
    out: List[int] = [0] * len(reads)
    jobs, who, base = [], [], []
    w, plen = params.refine_window, params.refine_pattern_len
    for k, read in enumerate(reads):
        hi_ref = int(frag_ref_hi[k])
        off = min(int(junction_off[k]), len(read))
        lo = max(hi_ref - w, 0)
        hi = min(hi_ref + w, int(lens[contigs[k]]))
        if hi <= lo or off == 0:
            out[k] = max(hi_ref - 1, 0)
            continue
        text_lo_in_read = max(lo - int(frag_implied_start[k]) + 2, 0)
        pat_start = max(max(off - plen, 0), text_lo_in_read)
        if pat_start >= off:
            out[k] = max(hi_ref - 1, 0)
            continue
        jobs.append(_job(k, pat_start, off, contigs[k], lo, hi))
        who.append(k)
        base.append(lo)
    return out, np.array(jobs, dtype=WINDOW_JOB_DTYPE), who, base


def refine_donor_end(dref: DeviceReference, reads: Sequence[bytes], contigs: Sequence[int],
                     junction_off: Sequence[int], frag_ref_hi: Sequence[int],
                     frag_implied_start: Sequence[int], params: IntronParams = IntronParams()) -> List[int]:
    """intron_chain.rs:285-309 per read: infix_best_end of read[pat_start..off]
    over [frag_ref_hi - w, frag_ref_hi + w); empty window or pattern gives
    frag_ref_hi - 1 (saturating)."""
    out, jobs, who, base = donor_plan(dref.ref.lens, reads, contigs, junction_off, frag_ref_hi,
                                      frag_implied_start, params)
    if len(who):
        _, pos, _ = window_myers(dref, reads, jobs, MODE_END)
        for j, k in enumerate(who):
            out[k] = base[j] + int(pos[j])
    return out


def acceptor_plan(lens, reads: Sequence[bytes], contigs: Sequence[int], junction_off: Sequence[int],
                  frag_ref_lo: Sequence[int], frag_implied_start: Sequence[int],
                  params: IntronParams = IntronParams()):
    """The geometry of refine_acceptor_start (intron_chain.rs:318-342), laid out
    as donor_plan's."""
    out: List[int] = [0] * len(reads)
    jobs, who, base = [], [], []
    w, plen = params.refine_window, params.refine_pattern_len
    for k, read in enumerate(reads):
        lo_ref = int(frag_ref_lo[k])
        off = min(int(junction_off[k]), len(read))
        lo = max(lo_ref - w, 0)
        hi = min(lo_ref + w, int(lens[contigs[k]]))
        if hi <= lo:
            out[k] = lo_ref
            continue
        text_hi_in_read = max(hi - int(frag_implied_start[k]) - 2, 0)
        pat_end = min(off + plen, len(read), text_hi_in_read)
        if pat_end <= off:
            out[k] = lo_ref
            continue
        jobs.append(_job(k, off, pat_end, contigs[k], lo, hi))
        who.append(k)
        base.append(lo)
    return out, np.array(jobs, dtype=WINDOW_JOB_DTYPE), who, base


def refine_acceptor_start(dref: DeviceReference, reads: Sequence[bytes], contigs: Sequence[int],
                          junction_off: Sequence[int], frag_ref_lo: Sequence[int],
                          frag_implied_start: Sequence[int],
                          params: IntronParams = IntronParams()) -> List[int]:
    """intron_chain.rs:318-342 per read: infix_best_start of read[off..pat_end]
    over [frag_ref_lo - w, frag_ref_lo + w); empty window or pattern gives
    frag_ref_lo."""
    out, jobs, who, base = acceptor_plan(dref.ref.lens, reads, contigs, junction_off, frag_ref_lo,
                                         frag_implied_start, params)
    if len(who):
        _, pos, _ = window_myers(dref, reads, jobs, MODE_START)
        for j, k in enumerate(who):
            out[k] = base[j] + int(pos[j])
    return out


def verify_plan(lens, reads: Sequence[bytes], contigs: Sequence[int], frag1_off_lo: Sequence[int],
                frag2_r_hi: Sequence[int], donor_end: Sequence[int], acceptor_start: Sequence[int]):
    """The geometry of verify_split_read (intron_chain.rs:344-373): (jobs,
    readers); reader k owns jobs 2j (donor half) and 2j+1 (acceptor half)."""
    jobs, who = [], []
    for k, read in enumerate(reads):
        clen = int(lens[contigs[k]])
        mid = len(read) // 2
        d_start = max(int(frag1_off_lo[k]) - 30, 0)
        d_end = min(max(int(donor_end[k]) + 1 + 30, 1), clen)
        a_start = min(max(int(acceptor_start[k]) - 30, 0), clen)
        a_end = min(max(int(frag2_r_hi[k]) + 30, a_start + 1), clen)
        if d_start >= d_end or a_start >= a_end or mid == 0 or mid >= len(read):
            continue
        jobs.append(_job(k, 0, mid, contigs[k], d_start, d_end))
        jobs.append(_job(k, mid, len(read), contigs[k], a_start, a_end))
        who.append(k)
    return np.array(jobs, dtype=WINDOW_JOB_DTYPE), who


def verify_split_read(dref: DeviceReference, reads: Sequence[bytes], contigs: Sequence[int],
                      frag1_off_lo: Sequence[int], frag2_r_hi: Sequence[int],
                      donor_end: Sequence[int], acceptor_start: Sequence[int]) -> List[Tuple[int, int]]:
    """intron_chain.rs:344-373 per read: (long::infix of the first half over the
    donor window, of the second half over the acceptor window); degenerate
    windows give (i32::MAX, i32::MAX)."""
    out: List[Tuple[int, int]] = [(I32_MAX, I32_MAX)] * len(reads)
    jobs, who = verify_plan(dref.ref.lens, reads, contigs, frag1_off_lo, frag2_r_hi, donor_end,
                            acceptor_start)
    if len(who):
        dist, _, _ = window_myers(dref, reads, jobs, MODE_DIST)
        for j, k in enumerate(who):
            out[k] = (int(dist[2 * j]), int(dist[2 * j + 1]))
    return out
