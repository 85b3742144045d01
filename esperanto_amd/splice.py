"""Spliced window DP on the device (crates/map/src/splice.rs:435-664,
crates/map/src/intron_chain.rs:497-536).

`spliced_batch` takes prepared exon lists (what prepare_spliced or the
intron-chain fast path computes from anchors and the junction library) and
runs, per read: pseudo-reference concatenation from the resident reference,
extend_hint at the given hint offset, stitch_cigar back to genomic
coordinates and SplicedAlignment::genomic_start (esp_spliced_batch).
`stitch_cigar` is the host entry point of the same walk.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import List, Optional, Sequence, Tuple

import numpy as np

from . import _native
from .batch import ContextTimings, DeviceArray
from .extend import ExtendBuffer
from .records import CigarOp, Extension, ExtendParams, cigar_from_words
from .reference import DeviceReference

ALIGN_SPLICED_FLANK = 30  # splice.rs:356

SPLICE_REC_DTYPE = np.dtype([("status", "<i4"), ("genomic_start", "<u4"), ("pseudo_len", "<u4"),
                             ("n_stitched", "<u4")])


def stitch_cigar(cigar: Sequence[CigarOp], p_start: int, exon_bounds: Sequence[Tuple[int, int, int]],
                 junction_lens: Sequence[int]) -> List[CigarOp]:
    """splice.rs:435 stitch_cigar; junction_lens[k] = junctions[k].end - start."""
    L = _native.lib()
    words = (ctypes.c_uint32 * max(len(cigar), 1))(*[c.word for c in cigar])
    flat = [v for b in exon_bounds for v in b]
    bounds = (ctypes.c_uint32 * max(len(flat), 1))(*flat)
    skips = (ctypes.c_uint32 * max(len(junction_lens), 1))(*junction_lens)
    cap = len(cigar) + 2 * len(exon_bounds) + 2
    out = (ctypes.c_uint32 * cap)()
    n = ctypes.c_size_t()
    _native.check(L.esp_stitch_cigar(words, len(cigar), p_start, bounds, len(exon_bounds), skips,
                                     len(junction_lens), out, cap, ctypes.byref(n)))
    return cigar_from_words(list(out)[:n.value])


@dataclass
class SplicedAlignment:
    """splice.rs:397-409 minus the junction annotations (signal, support),
    which come from the caller's refinement."""

    extension: Extension
    cigar: List[CigarOp]
    exon_bounds: List[Tuple[int, int, int]]
    genomic_start: int


@dataclass
class SpliceRequests:
    reads: np.ndarray
    read_off: np.ndarray
    read_len: np.ndarray
    contig: np.ndarray    # u32 per read
    hint: np.ndarray      # i64 per read
    exons: np.ndarray     # u32 [E, 2] (ref_start, ref_end)
    exon_off: np.ndarray  # u64 [n + 1]

    @property
    def npairs(self) -> int:
        return int(self.read_len.shape[0])

    @staticmethod
    def from_lists(reads: Sequence[bytes], contigs: Sequence[int],
                   exon_lists: Sequence[Sequence[Tuple[int, int]]], hints: Sequence[int]) -> "SpliceRequests":
        rl = np.array([len(r) for r in reads], dtype=np.uint32)
        ro = np.zeros(len(reads), dtype=np.uint64)
        if len(reads):
            ro[1:] = np.cumsum(rl, dtype=np.uint64)[:-1]
        rb = np.frombuffer(b"".join(bytes(r) for r in reads) or b"\0", dtype=np.uint8).copy()
        ne = np.array([len(e) for e in exon_lists], dtype=np.uint64)
        eo = np.zeros(len(reads) + 1, dtype=np.uint64)
        eo[1:] = np.cumsum(ne)
        flat = [x for ex in exon_lists for pair in ex for x in pair]
        exons = np.array(flat, dtype=np.uint32).reshape(-1, 2) if flat else np.zeros((0, 2), np.uint32)
        return SpliceRequests(rb, ro, rl, np.asarray(contigs, dtype=np.uint32),
                              np.asarray(hints, dtype=np.int64), exons, eo)

    def pseudo_offsets(self) -> np.ndarray:
        """Prefix sums of each read's exon lengths (exons with re <= rs count 0;
        such reads are rejected before any DP)."""
        ln = self.exons[:, 1].astype(np.int64) - self.exons[:, 0].astype(np.int64)
        ln = np.maximum(ln, 0)
        cs = np.concatenate([[0], np.cumsum(ln)]).astype(np.uint64)
        return cs[self.exon_off.astype(np.int64)]

    def cigar_offsets(self) -> np.ndarray:
        caps = 2 * self.read_len.astype(np.uint64) + 5
        off = np.zeros(self.npairs + 1, dtype=np.uint64)
        off[1:] = np.cumsum(caps)
        return off

    def stitch_offsets(self) -> np.ndarray:
        ne = np.diff(self.exon_off.astype(np.int64)).astype(np.uint64)
        caps = 2 * self.read_len.astype(np.uint64) + 5 + 2 * ne
        off = np.zeros(self.npairs + 1, dtype=np.uint64)
        off[1:] = np.cumsum(caps)
        return off


class DeviceSpliceBatch(ContextTimings):
    def __init__(self, req: SpliceRequests, dref: DeviceReference, buf: Optional[ExtendBuffer] = None):
        self.buf = buf or dref.buf
        self.req = req
        self.dref = dref
        b = self.buf
        n = req.npairs
        self.pseudo_off_host = req.pseudo_offsets()
        exons = req.exons.reshape(-1) if req.exons.size else np.zeros(2, np.uint32)
        self.d = {k: DeviceArray.from_numpy(b, getattr(req, k))
                  for k in ("reads", "read_off", "read_len", "contig", "hint", "exon_off")}
        self.d["exons"] = DeviceArray.from_numpy(b, exons)
        self.d["pseudo_off"] = DeviceArray.from_numpy(b, self.pseudo_off_host)
        self.cig_off_host = req.cigar_offsets()
        self.stitch_off_host = req.stitch_offsets()
        self.cig_off = DeviceArray.from_numpy(b, self.cig_off_host)
        self.stitch_off = DeviceArray.from_numpy(b, self.stitch_off_host)
        self.ext = DeviceArray(b, 16 * max(n, 1))
        self.cigar = DeviceArray(b, 4 * max(int(self.cig_off_host[-1]), 1))
        self.cig_n = DeviceArray(b, 4 * max(n, 1))
        self.stitched = DeviceArray(b, 4 * max(int(self.stitch_off_host[-1]), 1))
        self.rec = DeviceArray(b, SPLICE_REC_DTYPE.itemsize * max(n, 1))
        self._cb = _native.CSpliceBatch(n, self.d["reads"].ptr, self.d["read_off"].ptr,
                                        self.d["read_len"].ptr, self.d["contig"].ptr,
                                        self.d["hint"].ptr, self.d["exons"].ptr,
                                        self.d["exon_off"].ptr, self.d["pseudo_off"].ptr,
                                        int(self.pseudo_off_host[-1]), int(req.exon_off[-1]))
        self._out = _native.CBatchOut(self.ext.ptr, None, self.cigar.ptr, self.cig_off.ptr,
                                      self.cig_n.ptr, None)

    @property
    def npairs(self) -> int:
        return self.req.npairs

    def launch(self, params: Optional[ExtendParams] = None) -> None:
        p = (params or ExtendParams()).to_c()
        _native.check(_native.lib().esp_spliced_batch(
            self.buf.handle, self.dref.handle, ctypes.byref(self._cb), ctypes.byref(p),
            ctypes.byref(self._out), ctypes.c_void_p(self.stitched.ptr),
            ctypes.c_void_p(self.stitch_off.ptr), ctypes.c_void_p(self.rec.ptr)), self.buf.handle)

    def capture(self, params: Optional[ExtendParams] = None, max_read_len: Optional[int] = None,
                max_pseudo_len: Optional[int] = None):
        """launch() recorded as one HIP graph (esp_spliced_graph_create;
        recording runs the pipeline once); bounds default to this batch's
        longest read and pseudo-reference.  The returned BatchGraph replays it."""
        from .batch import BatchGraph
        n = self.npairs
        if max_read_len is None:
            max_read_len = int(self.req.read_len.max()) if n else 1
        if max_pseudo_len is None:
            max_pseudo_len = int(np.diff(self.pseudo_off_host.astype(np.int64)).max()) if n else 1
        self._gp = (params or ExtendParams()).to_c()
        h = ctypes.c_void_p()
        _native.check(_native.lib().esp_spliced_graph_create(
            self.buf.handle, self.dref.handle, ctypes.byref(self._cb), ctypes.byref(self._gp),
            ctypes.byref(self._out), ctypes.c_void_p(self.stitched.ptr),
            ctypes.c_void_p(self.stitch_off.ptr), ctypes.c_void_p(self.rec.ptr),
            max(1, max_read_len), max(1, max_pseudo_len), ctypes.byref(h)), self.buf.handle)
        return BatchGraph(self.buf, h.value)

    def raw_results(self):
        """(records, ext[n,4] int64, cigar, cig_off, cig_n, stitched, stitch_off)."""
        n = self.npairs
        rec = self.rec.to_numpy(np.uint8, SPLICE_REC_DTYPE.itemsize * n).view(SPLICE_REC_DTYPE)
        e = self.ext.to_numpy(np.uint32, 4 * n).reshape(n, 4)
        ext = e.astype(np.int64)
        ext[:, 3] = e[:, 3].view(np.int32)
        cig = self.cigar.to_numpy(np.uint32, int(self.cig_off_host[-1]))
        cig_n = self.cig_n.to_numpy(np.uint32, n)
        st = self.stitched.to_numpy(np.uint32, int(self.stitch_off_host[-1]))
        return rec, ext, cig, self.cig_off_host, cig_n, st, self.stitch_off_host

    def alignments(self) -> List[Optional[SplicedAlignment]]:
        rec, ext, cig, off, cig_n, st, soff = self.raw_results()
        out: List[Optional[SplicedAlignment]] = []
        for k in range(self.npairs):
            status = int(rec[k]["status"])
            if status == 0:
                out.append(None)
                continue
            if status < 0:
                raise ValueError(f"read {k}: spliced request invalid (status {status})")
            e0, e1 = int(self.req.exon_off[k]), int(self.req.exon_off[k + 1])
            bounds, ps = [], 0
            for rs, re in self.req.exons[e0:e1]:
                bounds.append((ps, int(rs), int(re)))
                ps += int(re) - int(rs)
            words = cig[int(off[k]):int(off[k]) + int(cig_n[k])]
            x = Extension(int(ext[k, 0]), int(ext[k, 1]), int(ext[k, 2]), cigar_from_words(words),
                          int(ext[k, 3]))
            sw = st[int(soff[k]):int(soff[k]) + int(rec[k]["n_stitched"])]
            out.append(SplicedAlignment(x, cigar_from_words(sw), bounds, int(rec[k]["genomic_start"])))
        return out

    def free(self) -> None:
        for a in list(self.d.values()) + [self.cig_off, self.stitch_off, self.ext, self.cigar,
                                          self.cig_n, self.stitched, self.rec]:
            a.free()


def spliced_batch(dref: DeviceReference, reads: Sequence[bytes], contigs: Sequence[int],
                  exon_lists: Sequence[Sequence[Tuple[int, int]]], hints: Sequence[int],
                  params: Optional[ExtendParams] = None) -> List[Optional[SplicedAlignment]]:
    db = DeviceSpliceBatch(SpliceRequests.from_lists(reads, contigs, exon_lists, hints), dref)
    try:
        db.launch(params)
        db.sync()
        return db.alignments()
    finally:
        db.free()
