"""Unspliced chain extension on the device (crates/map/src/align.rs:33-55, :1770-1919).

`unspliced_batch` runs extend_unspliced_chain after its fast lane for many
chains in one pass: the continuous extension window, the EA-Myers lane
(best start, per-base recount, Kadane soft clips) for reads of 1..256 bases,
and the banded extend_hint fallback where the caller's DP budget allows it
(esp_unspliced_batch).  `continuous_extension_window` is the geometry alone.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import List, Optional, Sequence, Tuple

import numpy as np

from . import _native
from .batch import ContextTimings, DeviceArray
from .extend import ExtendBuffer
from .records import CigarOp, ExtendParams, cigar_from_words
from .reference import DeviceReference

UNSPLICED_REC_DTYPE = np.dtype([("status", "<i4"), ("strand", "<i4"), ("pos", "<u4"),
                                ("score", "<i4"), ("dist", "<i4"), ("wstart", "<u4"),
                                ("wlen", "<u4")])
EA_LANE, BANDED_DP = 1, 2


def continuous_extension_window(ref_span: Tuple[int, int], read_span: Tuple[int, int],
                                read_len: int, contig_len: int) -> Tuple[int, int]:
    """align.rs:33-55 with Chain::ref_span / read_span given (u32 saturating)."""
    M = 0xFFFFFFFF
    sadd = lambda x, y: min(x + y, M)  # noqa: E731
    rs, re = ref_span
    qs, qe = read_span
    flank = read_len // 2 + 30
    head = max(flank, qs)
    tail = max(flank, max(read_len - qe, 0))
    start = min(max(rs - head, 0), contig_len)
    end = min(sadd(re, tail), contig_len, sadd(sadd(rs, read_len), flank),
              sadd(sadd(start, read_len), min(flank * 2, M)))
    return start, max(end, start)


@dataclass
class UnsplicedPlacement:
    """Placement::from_genomic fields (align.rs:1852-1862, :1910-1918)."""

    contig: int
    pos: int
    minus: bool
    cigar: List[CigarOp]
    score: int
    lane: int  # EA_LANE or BANDED_DP


@dataclass
class ChainRequests:
    reads: np.ndarray
    read_off: np.ndarray
    read_len: np.ndarray
    contig: np.ndarray
    minus: np.ndarray
    ref_start: np.ndarray
    ref_end: np.ndarray
    read_start: np.ndarray
    read_end: np.ndarray
    allow_dp: np.ndarray

    @property
    def npairs(self) -> int:
        return int(self.read_len.shape[0])

    @staticmethod
    def from_lists(reads: Sequence[bytes], contigs, minus, ref_spans, read_spans,
                   allow_dp) -> "ChainRequests":
        rl = np.array([len(r) for r in reads], dtype=np.uint32)
        ro = np.zeros(len(reads), dtype=np.uint64)
        if len(reads):
            ro[1:] = np.cumsum(rl, dtype=np.uint64)[:-1]
        rb = np.frombuffer(b"".join(bytes(r) for r in reads) or b"\0", dtype=np.uint8).copy()
        u = lambda x: np.asarray(x, dtype=np.uint32)  # noqa: E731
        rsp = np.asarray(ref_spans, dtype=np.uint32).reshape(-1, 2)
        qsp = np.asarray(read_spans, dtype=np.uint32).reshape(-1, 2)
        return ChainRequests(rb, ro, rl, u(contigs), u(minus), rsp[:, 0].copy(), rsp[:, 1].copy(),
                             qsp[:, 0].copy(), qsp[:, 1].copy(), u(allow_dp))

    def cigar_offsets(self) -> np.ndarray:
        caps = 2 * self.read_len.astype(np.uint64) + 5
        off = np.zeros(self.npairs + 1, dtype=np.uint64)
        off[1:] = np.cumsum(caps)
        return off


FIELDS = ("reads", "read_off", "read_len", "contig", "minus", "ref_start", "ref_end",
          "read_start", "read_end", "allow_dp")


class DeviceChainBatch(ContextTimings):
    def __init__(self, req: ChainRequests, dref: DeviceReference, buf: Optional[ExtendBuffer] = None):
        self.buf = buf or dref.buf
        self.req = req
        self.dref = dref
        b = self.buf
        n = req.npairs
        self.d = {k: DeviceArray.from_numpy(b, getattr(req, k)) for k in FIELDS}
        self.cig_off_host = req.cigar_offsets()
        self.cig_off = DeviceArray.from_numpy(b, self.cig_off_host)
        self.cigar = DeviceArray(b, 4 * max(int(self.cig_off_host[-1]), 1))
        self.cig_n = DeviceArray(b, 4 * max(n, 1))
        self.rec = DeviceArray(b, UNSPLICED_REC_DTYPE.itemsize * max(n, 1))
        maxm = int(req.read_len.max()) if n else 1
        self._cb = _native.CChainBatch(n, *(self.d[k].ptr for k in FIELDS), max(1, maxm))

    @property
    def npairs(self) -> int:
        return self.req.npairs

    def launch(self, params: Optional[ExtendParams] = None) -> None:
        p = (params or ExtendParams()).to_c()
        _native.check(_native.lib().esp_unspliced_batch(
            self.buf.handle, self.dref.handle, ctypes.byref(self._cb), ctypes.byref(p),
            ctypes.c_void_p(self.rec.ptr), ctypes.c_void_p(self.cigar.ptr),
            ctypes.c_void_p(self.cig_off.ptr), ctypes.c_void_p(self.cig_n.ptr)), self.buf.handle)

    def raw_results(self):
        n = self.npairs
        rec = self.rec.to_numpy(np.uint8, UNSPLICED_REC_DTYPE.itemsize * n).view(UNSPLICED_REC_DTYPE)
        cig = self.cigar.to_numpy(np.uint32, int(self.cig_off_host[-1]))
        return rec, cig, self.cig_off_host, self.cig_n.to_numpy(np.uint32, n)

    def placements(self) -> List[Optional[UnsplicedPlacement]]:
        rec, cig, off, cig_n = self.raw_results()
        out: List[Optional[UnsplicedPlacement]] = []
        for k in range(self.npairs):
            st = int(rec[k]["status"])
            if st < 0:
                raise ValueError(f"chain {k}: contig out of range or read longer than the bound")
            if st == 0:
                out.append(None)
                continue
            words = cig[int(off[k]):int(off[k]) + int(cig_n[k])]
            out.append(UnsplicedPlacement(int(self.req.contig[k]), int(rec[k]["pos"]),
                                          bool(rec[k]["strand"]), cigar_from_words(words),
                                          int(rec[k]["score"]), st))
        return out

    def free(self) -> None:
        for a in list(self.d.values()) + [self.cig_off, self.cigar, self.cig_n, self.rec]:
            a.free()


def unspliced_batch(dref: DeviceReference, req: ChainRequests,
                    params: Optional[ExtendParams] = None) -> List[Optional[UnsplicedPlacement]]:
    db = DeviceChainBatch(req, dref)
    try:
        db.launch(params)
        db.sync()
        return db.placements()
    finally:
        db.free()
