"""The Kadane fast lane on the device: Aligner::try_fast_lane
(crates/map/src/align.rs:2274-2401) over a batch of chains.

A chain is (contig, its anchors as (qpos, rpos) pairs in the chain's
(rpos, qpos) order, the read in the chain's orientation).  The lane accepts a
decisive single-diagonal chain covering the read (anchor read gaps at most
k + w + 5, at most 2 uncovered bases at each end) whose best Kadane run along
the diagonal has no scored mismatch; it returns the Extension in extend_hint's
window coordinates and the window start.
"""
from __future__ import annotations

import ctypes
from typing import List, Optional, Sequence, Tuple

import numpy as np

from . import _native
from .batch import ContextTimings, DeviceArray
from .records import Extension, ExtendParams, cigar_from_words
from .reference import DeviceReference

FAST_DTYPE = np.dtype([("status", "<i4"), ("score", "<i4"), ("wstart", "<u4"), ("ref_start", "<u4"),
                       ("read_start", "<u4"), ("read_end", "<u4"), ("pos", "<u4"), ("n_ops", "<u4")])


class DeviceFastBatch(ContextTimings):
    """Chains resident in HBM (flat arrays as esp_fast_batch takes them) with
    their output records, for repeated launches."""

    FIELDS = ("reads", "read_off", "read_len", "contig", "anchors", "anc_off", "anc_n")

    def __init__(self, dref: DeviceReference, reads: np.ndarray, read_off: np.ndarray, read_len: np.ndarray,
                 contig: np.ndarray, anchors: np.ndarray, anc_off: np.ndarray, anc_n: np.ndarray):
        self.buf = dref.buf
        self.dref = dref
        self.n = int(read_len.shape[0])
        host = dict(reads=np.ascontiguousarray(reads, dtype=np.uint8),
                    read_off=np.ascontiguousarray(read_off, dtype=np.uint64),
                    read_len=np.ascontiguousarray(read_len, dtype=np.uint32),
                    contig=np.ascontiguousarray(contig, dtype=np.uint32),
                    anchors=np.ascontiguousarray(anchors, dtype=np.uint32).reshape(-1),
                    anc_off=np.ascontiguousarray(anc_off, dtype=np.uint64),
                    anc_n=np.ascontiguousarray(anc_n, dtype=np.uint32))
        self.d = {k: DeviceArray.from_numpy(self.buf, host[k]) for k in self.FIELDS}
        self.out = DeviceArray(self.buf, FAST_DTYPE.itemsize * max(self.n, 1))
        self.cig = DeviceArray(self.buf, 12 * max(self.n, 1))
        self._cb = _native.CFastBatch(self.n, *(self.d[k].ptr for k in self.FIELDS))

    def launch(self, params: Optional[ExtendParams] = None, k: int = 15, w: int = 5) -> None:
        _native.check(_native.lib().esp_fast_lane_batch(
            self.buf.handle, self.dref.handle, ctypes.byref(self._cb),
            ctypes.byref((params or ExtendParams()).to_c()), k, w, ctypes.c_void_p(self.out.ptr),
            ctypes.c_void_p(self.cig.ptr)), self.buf.handle)

    def results(self):
        rec = self.out.to_numpy(np.uint8, FAST_DTYPE.itemsize * self.n).view(FAST_DTYPE)
        return rec, self.cig.to_numpy(np.uint32, 3 * self.n).reshape(self.n, 3)

    def free(self) -> None:
        for a in list(self.d.values()) + [self.out, self.cig]:
            a.free()


def fast_lane_batch(dref: DeviceReference, reads: Sequence[bytes], contigs: Sequence[int],
                    anchors: Sequence[Sequence[Tuple[int, int]]], params: Optional[ExtendParams] = None,
                    k: int = 15, w: int = 5):
    """(records as FAST_DTYPE, CIGAR words (n, 3)) for each chain."""
    buf = dref.buf
    n = len(reads)
    if n == 0:
        return np.zeros(0, FAST_DTYPE), np.zeros((0, 3), np.uint32)
    rl = np.array([len(r) for r in reads], dtype=np.uint32)
    ro = np.zeros(n, dtype=np.uint64)
    ro[1:] = np.cumsum(rl, dtype=np.uint64)[:-1]
    rb = np.frombuffer(b"".join(bytes(r) for r in reads) or b"\0", dtype=np.uint8)
    an = np.array([len(a) for a in anchors], dtype=np.uint32)
    ao = np.zeros(n, dtype=np.uint64)
    ao[1:] = np.cumsum(an, dtype=np.uint64)[:-1]
    flat = np.array([x for a in anchors for pair in a for x in pair] or [0], dtype=np.uint32)
    host = [rb, ro, rl, np.asarray(contigs, dtype=np.uint32), flat, ao, an]
    dev = [DeviceArray.from_numpy(buf, h) for h in host]
    d_out = DeviceArray(buf, FAST_DTYPE.itemsize * n)
    d_cig = DeviceArray(buf, 12 * n)
    try:
        cb = _native.CFastBatch(n, *(d.ptr for d in dev))
        L = _native.lib()
        _native.check(L.esp_fast_lane_batch(buf.handle, dref.handle, ctypes.byref(cb),
                                            ctypes.byref((params or ExtendParams()).to_c()), k, w,
                                            ctypes.c_void_p(d_out.ptr), ctypes.c_void_p(d_cig.ptr)), buf.handle)
        _native.check(L.esp_sync(buf.handle), buf.handle)
        rec = d_out.to_numpy(np.uint8, FAST_DTYPE.itemsize * n).view(FAST_DTYPE)
        cig = d_cig.to_numpy(np.uint32, 3 * n).reshape(n, 3)
        return rec, cig
    finally:
        for d in dev + [d_out, d_cig]:
            d.free()


def try_fast_lane(dref: DeviceReference, query: bytes, contig: int, anchors: Sequence[Tuple[int, int]],
                  params: Optional[ExtendParams] = None, k: int = 15,
                  w: int = 5) -> Optional[Tuple[Extension, int]]:
    """align.rs:2274 for one chain: None when the lane refuses, else
    (Extension, wstart)."""
    rec, cig = fast_lane_batch(dref, [query], [contig], [anchors], params, k, w)
    r = rec[0]
    if r["status"] < 0:
        raise ValueError("contig out of range")
    if r["status"] != 1:
        return None
    ext = Extension(int(r["read_start"]), int(r["read_end"]), int(r["ref_start"]),
                    cigar_from_words(cig[0][: int(r["n_ops"])].tolist()), int(r["score"]))
    return ext, int(r["wstart"])
