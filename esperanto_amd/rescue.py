"""Mate-window rescue on the device (crates/map/src/align.rs:851-921).

`mate_rescue_batch` runs mate_window_placement for many (read, mate contig,
mate position) requests in one pass over the resident reference: window
decode, EA-Myers best start of both orientations, selection, extend_hint and
placement (esp_mate_rescue_batch).  `mate_window_placement` is the one-read
form with the reference's signature minus `self`.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import List, Optional, Sequence

import numpy as np

from . import _native
from .batch import ContextTimings, DeviceArray
from .extend import ExtendBuffer, default_buffer
from .records import CigarOp, ExtendParams, cigar_capacity, cigar_from_words
from .reference import DeviceReference

MATE_WINDOW_INSERT = 500  # align.rs:181

MATE_REC_DTYPE = np.dtype([("placed", "<i4"), ("strand", "<i4"), ("pos", "<u4"), ("score", "<i4"),
                           ("dist", "<i4"), ("wstart", "<u4"), ("wlen", "<u4"),
                           ("read_start", "<u4"), ("read_end", "<u4")])


@dataclass
class MatePlacement:
    """Placement::from_genomic fields of a rescued read (align.rs:911-919)."""

    contig: int
    pos: int
    minus: bool
    cigar: List[CigarOp]
    score: int
    dist: int


@dataclass
class MateRequests:
    reads: np.ndarray     # u8, concatenated read bytes (alignment orientation as given)
    read_off: np.ndarray  # u64
    read_len: np.ndarray  # u32
    contig: np.ndarray    # u32 mate contig
    mate_pos: np.ndarray  # u32 mate position

    @property
    def npairs(self) -> int:
        return int(self.read_len.shape[0])

    @staticmethod
    def from_lists(reads: Sequence[bytes], contigs: Sequence[int], poss: Sequence[int]) -> "MateRequests":
        rl = np.array([len(r) for r in reads], dtype=np.uint32)
        ro = np.zeros(len(reads), dtype=np.uint64)
        if len(reads):
            ro[1:] = np.cumsum(rl, dtype=np.uint64)[:-1]
        rb = np.frombuffer(b"".join(bytes(r) for r in reads) or b"\0", dtype=np.uint8).copy()
        return MateRequests(rb, ro, rl, np.asarray(contigs, dtype=np.uint32),
                            np.asarray(poss, dtype=np.uint32))

    def cigar_offsets(self) -> np.ndarray:
        caps = 2 * self.read_len.astype(np.uint64) + 5
        off = np.zeros(self.npairs + 1, dtype=np.uint64)
        off[1:] = np.cumsum(caps)
        return off


class DeviceMateBatch(ContextTimings):
    """MateRequests resident in HBM plus the result arena."""

    def __init__(self, req: MateRequests, dref: DeviceReference, buf: Optional[ExtendBuffer] = None):
        self.buf = buf or dref.buf
        self.req = req
        self.dref = dref
        b = self.buf
        n = req.npairs
        self.d = {k: DeviceArray.from_numpy(b, getattr(req, k))
                  for k in ("reads", "read_off", "read_len", "contig", "mate_pos")}
        self.cig_off_host = req.cigar_offsets()
        self.cig_off = DeviceArray.from_numpy(b, self.cig_off_host)
        self.cigar = DeviceArray(b, 4 * max(int(self.cig_off_host[-1]), 1))
        self.cig_n = DeviceArray(b, 4 * max(n, 1))
        self.rec = DeviceArray(b, MATE_REC_DTYPE.itemsize * max(n, 1))
        maxm = int(req.read_len.max()) if n else 1
        self._cb = _native.CMateBatch(n, self.d["reads"].ptr, self.d["read_off"].ptr,
                                      self.d["read_len"].ptr, self.d["contig"].ptr,
                                      self.d["mate_pos"].ptr, max(1, min(maxm, 256)))

    @property
    def npairs(self) -> int:
        return self.req.npairs

    def launch(self, params: Optional[ExtendParams] = None) -> None:
        p = (params or ExtendParams()).to_c()
        _native.check(_native.lib().esp_mate_rescue_batch(
            self.buf.handle, self.dref.handle, ctypes.byref(self._cb), ctypes.byref(p),
            ctypes.c_void_p(self.rec.ptr), ctypes.c_void_p(self.cigar.ptr),
            ctypes.c_void_p(self.cig_off.ptr), ctypes.c_void_p(self.cig_n.ptr)), self.buf.handle)

    def capture(self, params: Optional[ExtendParams] = None):
        """launch() recorded as one HIP graph (esp_mate_rescue_graph_create;
        recording runs the pipeline once); the returned BatchGraph replays it."""
        from .batch import BatchGraph
        self._gp = (params or ExtendParams()).to_c()
        h = ctypes.c_void_p()
        _native.check(_native.lib().esp_mate_rescue_graph_create(
            self.buf.handle, self.dref.handle, ctypes.byref(self._cb), ctypes.byref(self._gp),
            ctypes.c_void_p(self.rec.ptr), ctypes.c_void_p(self.cigar.ptr),
            ctypes.c_void_p(self.cig_off.ptr), ctypes.c_void_p(self.cig_n.ptr), ctypes.byref(h)),
            self.buf.handle)
        return BatchGraph(self.buf, h.value)

    def raw_results(self):
        """(records, cigar words, cig_off, cig_n) as numpy arrays."""
        n = self.npairs
        rec = self.rec.to_numpy(np.uint8, MATE_REC_DTYPE.itemsize * n).view(MATE_REC_DTYPE)
        cig = self.cigar.to_numpy(np.uint32, int(self.cig_off_host[-1]))
        cig_n = self.cig_n.to_numpy(np.uint32, n)
        return rec, cig, self.cig_off_host, cig_n

    def placements(self) -> List[Optional[MatePlacement]]:
        rec, cig, off, cig_n = self.raw_results()
        out: List[Optional[MatePlacement]] = []
        for k in range(self.npairs):
            r = rec[k]
            if int(r["placed"]) < 0:
                raise ValueError(f"request {k}: mate contig out of range or read longer than "
                                 f"the batch bound")
            if int(r["placed"]) != 1:
                out.append(None)
                continue
            words = cig[int(off[k]):int(off[k]) + int(cig_n[k])]
            out.append(MatePlacement(int(self.req.contig[k]), int(r["pos"]), bool(r["strand"]),
                                     cigar_from_words(words), int(r["score"]), int(r["dist"])))
        return out

    def free(self) -> None:
        for a in list(self.d.values()) + [self.cig_off, self.cigar, self.cig_n, self.rec]:
            a.free()


def mate_rescue_batch(dref: DeviceReference, reads: Sequence[bytes], contigs: Sequence[int],
                      poss: Sequence[int], params: Optional[ExtendParams] = None
                      ) -> List[Optional[MatePlacement]]:
    db = DeviceMateBatch(MateRequests.from_lists(reads, contigs, poss), dref)
    try:
        db.launch(params)
        db.sync()
        return db.placements()
    finally:
        db.free()


def mate_window_placement(dref: DeviceReference, seq: bytes, mate_contig: int, mate_pos: int,
                          params: Optional[ExtendParams] = None) -> Optional[MatePlacement]:
    """align.rs:851 Aligner::mate_window_placement(seq, mate_contig, mate_pos)."""
    return mate_rescue_batch(dref, [seq], [mate_contig], [mate_pos], params)[0]
