"""Placement walks on the device (crates/map/src/align.rs:1060-1163, :1231-1288,
:2473-2516).

`walk_batch` runs, for many placements in one launch, the three CIGAR walks
the reference performs on every candidate after its DP:

* ``rescore_with`` (rescore_strict is dir = Off): the StrictScore record;
* ``recount_mm_ea_in``: (mm, ea_ag, ea_tc);
* ``exact_kmer_support``: non-overlapping exact k-mers along the M blocks.

Placements are (contig, pos, BAM CIGAR words, query in the alignment's frame,
EditDir).  The reference stays resident (`DeviceReference`); an annotated
junction library (`DeviceJunctions`, gtf.rs:120-169) is optional.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import List, Optional, Sequence, Tuple

import numpy as np

from . import _native
from .batch import ContextTimings, DeviceArray
from .extend import ExtendBuffer
from .records import ExtendParams
from .reference import DeviceReference

STRICT_DTYPE = np.dtype([("score", "<i4"), ("aligned", "<u4"), ("matches", "<u4"), ("mm", "<u4"),
                         ("ea_ag", "<u4"), ("ea_tc", "<u4"), ("ins", "<u4"), ("del", "<u4"),
                         ("clipped", "<u4"), ("n_splice", "<u4"), ("splice_score", "<i4")])
COUNTS_DTYPE = np.dtype([("mm", "<u4"), ("ea_ag", "<u4"), ("ea_tc", "<u4"), ("kmers", "<u4")])


@dataclass(frozen=True)
class WalkParams:
    """SpliceParams fields of rescore_with (splice.rs:40-52) and SeedParams k."""

    known_bonus: int = 6
    gt_ag_bonus: int = 4
    gc_ag_bonus: int = 2
    at_ac_bonus: int = 1
    noncanonical_penalty: int = 8
    seed_k: int = 15

    def to_c(self) -> _native.CWalkParams:
        return _native.CWalkParams(self.known_bonus, self.gt_ag_bonus, self.gc_ag_bonus,
                                   self.at_ac_bonus, self.noncanonical_penalty, self.seed_k)


def intron_length_penalty(n: int) -> int:
    """splice.rs:119 intron_length_penalty, as the walk kernel evaluates it."""
    return int(_native.lib().esp_intron_length_penalty(n))


class DeviceJunctions:
    """A JunctionLib's (contig, start, end) triples (any order, duplicates
    allowed), resident as the device hash set contains_any_strand is asked of."""

    def __init__(self, triples: Sequence[Tuple[int, int, int]], buf: ExtendBuffer):
        self.buf = buf
        t = np.ascontiguousarray(np.asarray(list(triples), dtype=np.uint32).reshape(-1, 3))
        h = ctypes.c_void_p()
        _native.check(_native.lib().esp_junctions_upload(buf.handle, t.shape[0],
                                                         ctypes.c_void_p(t.ctypes.data) if t.size else None,
                                                         ctypes.byref(h)), buf.handle)
        self.handle = h.value

    def close(self) -> None:
        if self.handle:
            _native.lib().esp_junctions_free(self.buf.handle, ctypes.c_void_p(self.handle))
            self.handle = None


@dataclass
class Placements:
    reads: np.ndarray
    read_off: np.ndarray
    read_len: np.ndarray
    contig: np.ndarray
    pos: np.ndarray
    cigar: np.ndarray
    cig_off: np.ndarray
    cig_n: np.ndarray
    dir: np.ndarray

    @property
    def n(self) -> int:
        return int(self.read_len.shape[0])

    @staticmethod
    def from_lists(queries: Sequence[bytes], contigs, positions, cigars, dirs=None) -> "Placements":
        """cigars: sequences of BAM words (len << 4 | op, M0 I1 D2 N3 S4)."""
        n = len(queries)
        rl = np.array([len(q) for q in queries], dtype=np.uint32)
        ro = np.zeros(n, dtype=np.uint64)
        if n:
            ro[1:] = np.cumsum(rl, dtype=np.uint64)[:-1]
        rb = np.frombuffer(b"".join(bytes(q) for q in queries) or b"\0", dtype=np.uint8).copy()
        cn = np.array([len(c) for c in cigars], dtype=np.uint32)
        co = np.zeros(n, dtype=np.uint64)
        if n:
            co[1:] = np.cumsum(cn, dtype=np.uint64)[:-1]
        words = [int(w) for c in cigars for w in c]
        cw = np.array(words or [0], dtype=np.uint32)
        d = np.zeros(n, dtype=np.uint8) if dirs is None else np.asarray(dirs, dtype=np.uint8)
        return Placements(rb, ro, rl, np.asarray(contigs, dtype=np.uint32),
                          np.asarray(positions, dtype=np.uint32), cw, co, cn, d)


FIELDS = ("reads", "read_off", "read_len", "contig", "pos", "cigar", "cig_off", "cig_n", "dir")


class DeviceWalkBatch(ContextTimings):
    def __init__(self, pl: Placements, dref: DeviceReference, buf: Optional[ExtendBuffer] = None):
        self.buf = buf or dref.buf
        self.pl = pl
        self.dref = dref
        b = self.buf
        n = pl.n
        self.d = {k: DeviceArray.from_numpy(b, getattr(pl, k)) for k in FIELDS}
        self.strict = DeviceArray(b, STRICT_DTYPE.itemsize * max(n, 1))
        self.counts = DeviceArray(b, COUNTS_DTYPE.itemsize * max(n, 1))
        self.status = DeviceArray(b, 4 * max(n, 1))
        self._cb = _native.CWalkBatch(n, *(self.d[k].ptr for k in FIELDS))

    def launch(self, params: Optional[ExtendParams] = None, wp: Optional[WalkParams] = None,
               junctions: Optional[DeviceJunctions] = None) -> None:
        p = (params or ExtendParams()).to_c()
        w = (wp or WalkParams()).to_c()
        _native.check(_native.lib().esp_walk_placements(
            self.buf.handle, self.dref.handle,
            ctypes.c_void_p(junctions.handle) if junctions is not None else None,
            ctypes.byref(self._cb), ctypes.byref(p), ctypes.byref(w), ctypes.c_void_p(self.strict.ptr),
            ctypes.c_void_p(self.counts.ptr), ctypes.c_void_p(self.status.ptr)), self.buf.handle)

    def results(self):
        n = self.pl.n
        strict = self.strict.to_numpy(np.uint8, STRICT_DTYPE.itemsize * n).view(STRICT_DTYPE)
        counts = self.counts.to_numpy(np.uint8, COUNTS_DTYPE.itemsize * n).view(COUNTS_DTYPE)
        return strict, counts, self.status.to_numpy(np.int32, n)

    def free(self) -> None:
        for a in list(self.d.values()) + [self.strict, self.counts, self.status]:
            a.free()


def walk_batch(dref: DeviceReference, pl: Placements, params: Optional[ExtendParams] = None,
               wp: Optional[WalkParams] = None, junctions: Optional[DeviceJunctions] = None):
    """(StrictScore records, (mm, ea_ag, ea_tc, kmers) records, status) per placement."""
    db = DeviceWalkBatch(pl, dref)
    try:
        db.launch(params, wp, junctions)
        db.sync()
        return db.results()
    finally:
        db.free()


def rescore_with(dref: DeviceReference, query: bytes, contig: int, pos: int, cigar, dir: int = 0,
                 params: Optional[ExtendParams] = None, wp: Optional[WalkParams] = None,
                 junctions: Optional[DeviceJunctions] = None):
    """align.rs:1060 Aligner::rescore_with for one placement: its StrictScore."""
    strict, _, st = walk_batch(dref, Placements.from_lists([query], [contig], [pos], [cigar], [dir]),
                               params, wp, junctions)
    if st[0]:
        raise ValueError("contig out of range")
    return strict[0]


def rescore_strict(dref: DeviceReference, query: bytes, contig: int, pos: int, cigar, **kw):
    """align.rs:1045 Aligner::rescore_strict = rescore_with(.., EditDir::Off)."""
    return rescore_with(dref, query, contig, pos, cigar, 0, **kw)


def recount_mm_ea_in(dref: DeviceReference, query: bytes, contig: int, pos: int, cigar):
    """align.rs:2473 recount_mm_ea_in: (mm, ea_ag, ea_tc)."""
    _, c, st = walk_batch(dref, Placements.from_lists([query], [contig], [pos], [cigar]))
    if st[0]:
        raise ValueError("contig out of range")
    return int(c[0]["mm"]), int(c[0]["ea_ag"]), int(c[0]["ea_tc"])


def exact_kmer_support(dref: DeviceReference, query: bytes, contig: int, pos: int, cigar,
                       ea_dir: int = 0, k: int = 15) -> int:
    """align.rs:1231 Aligner::exact_kmer_support (k = SeedParams k)."""
    _, c, st = walk_batch(dref, Placements.from_lists([query], [contig], [pos], [cigar], [ea_dir]),
                          wp=WalkParams(seed_k=k))
    if st[0]:
        raise ValueError("contig out of range")
    return int(c[0]["kmers"])
