"""Read-aware junction refinement, align_spliced over chains, and the
re-extension of soft-clipped placements on the device.

Names follow the reference:
  refine_junction_read   crates/map/src/splice.rs:208-317  -> refine_junction_read_batch
  align_spliced          crates/map/src/splice.rs:625-664  -> align_spliced_batch
     (with prepare_spliced :545-621 and segment_chain :364-396 on the device)
  Aligner::reextend_with crates/map/src/align.rs:1170-1225 -> reextend_batch
The junction library is JunctionLib (crates/map/src/gtf.rs:72-149) uploaded
with strands and support counts (esp_junctions_upload_stranded).
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field
from typing import List, Optional, Sequence, Tuple

import numpy as np

from . import _native
from .batch import ContextTimings, DeviceArray
from .records import CigarOp, Extension, ExtendParams, cigar_from_words
from .reference import DeviceReference

SIGNALS = ("GtAg", "GcAg", "AtAc", "NonCanonical")
QUERY_DTYPE = np.dtype([("read_off", "<u8"), ("read_len", "<u4"), ("contig", "<u4"), ("naive_start", "<u4"),
                        ("naive_end", "<u4"), ("q_a", "<u4"), ("q_b", "<u4"), ("minus", "<u4"),
                        ("use_fallback", "<u4"), ("fallback_start", "<u4"), ("fallback_end", "<u4")])
REFINED_DTYPE = np.dtype([("found", "<i4"), ("start", "<u4"), ("end", "<u4"), ("minus", "<u4"),
                          ("signal", "<i4"), ("support", "<u4")])
CHAIN_REC_DTYPE = np.dtype([("status", "<i4"), ("refusal", "<i4"), ("n_segments", "<u4"),
                            ("genomic_start", "<u4"), ("pseudo_len", "<u4"), ("n_stitched", "<u4")])
REEXT_REC_DTYPE = np.dtype([("status", "<i4"), ("pos", "<u4"), ("score", "<i4"), ("wstart", "<u4"),
                            ("wlen", "<u4"), ("aligned", "<u4"), ("new_aligned", "<u4")])


@dataclass(frozen=True)
class SpliceParams:
    """splice.rs:21-53 (defaults :40-53)."""

    min_intron: int = 20
    max_intron: int = 1_000_000
    gt_ag_bonus: int = 4
    gc_ag_bonus: int = 2
    at_ac_bonus: int = 1
    noncanonical_penalty: int = 8
    known_bonus: int = 6
    refine_radius: int = 12

    def to_c(self) -> _native.CSpliceCfg:
        return _native.CSpliceCfg(self.min_intron, self.max_intron, self.gt_ag_bonus, self.gc_ag_bonus,
                                  self.at_ac_bonus, self.noncanonical_penalty, self.known_bonus,
                                  self.refine_radius)


@dataclass(frozen=True)
class RefinedJunction:
    """gtf.rs:52-61: the junction (contig implied), its signal and support."""

    start: int
    end: int
    minus: bool
    signal: str
    support: int


class JunctionLib:
    """JunctionLib::build of (contig, start, end, minus) quads, resident on the device."""

    def __init__(self, dref: DeviceReference, quads: Sequence[Tuple[int, int, int, int]]):
        self.buf = dref.buf
        q = np.ascontiguousarray(np.array(quads, dtype=np.uint32).reshape(-1, 4)) if len(quads) else \
            np.zeros((1, 4), np.uint32)
        h = ctypes.c_void_p()
        _native.check(_native.lib().esp_junctions_upload_stranded(self.buf.handle, len(quads),
                                                                   ctypes.c_void_p(q.ctypes.data), ctypes.byref(h)),
                      self.buf.handle)
        self.handle = h

    def close(self) -> None:
        if self.handle:
            _native.lib().esp_junctions_free(self.buf.handle, self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _i32(v: int) -> int:
    return v - (1 << 32) if v >= (1 << 31) else v


def _concat(reads: Sequence[bytes]):
    rl = np.array([len(r) for r in reads], dtype=np.uint32)
    ro = np.zeros(len(reads), dtype=np.uint64)
    if len(reads):
        ro[1:] = np.cumsum(rl, dtype=np.uint64)[:-1]
    rb = np.frombuffer(b"".join(bytes(r) for r in reads) or b"\0", dtype=np.uint8).copy()
    return rb, ro, rl


def _refined(r) -> Optional[RefinedJunction]:
    if int(r["found"]) == 0:
        return None
    return RefinedJunction(int(r["start"]), int(r["end"]), bool(r["minus"]), SIGNALS[int(r["signal"])],
                           int(r["support"]))


def refine_junction_read_batch(dref: DeviceReference, jlib: Optional[JunctionLib], reads: Sequence[bytes],
                               queries: Sequence[tuple], params: SpliceParams = SpliceParams()
                               ) -> List[Optional[RefinedJunction]]:
    """queries: (read index, contig, naive_start, naive_end, q_a, q_b, minus[, fallback (start, end)])."""
    buf = dref.buf
    n = len(queries)
    if n == 0:
        return []
    rb, ro, rl = _concat(reads)
    Q = np.zeros(n, dtype=QUERY_DTYPE)
    for i, q in enumerate(queries):
        k = q[0]
        fb = q[7] if len(q) > 7 and q[7] is not None else None
        Q[i] = (ro[k], rl[k], q[1], q[2], q[3], q[4], q[5], int(q[6]), int(fb is not None),
                fb[0] if fb else 0, fb[1] if fb else 0)
    arrs = [DeviceArray.from_numpy(buf, Q), DeviceArray.from_numpy(buf, rb), DeviceArray(buf, REFINED_DTYPE.itemsize * n)]
    try:
        p = params.to_c()
        _native.check(_native.lib().esp_refine_junction_batch(buf.handle, dref.handle, jlib.handle if jlib else None,
                                                               ctypes.c_void_p(arrs[0].ptr), n,
                                                               ctypes.c_void_p(arrs[1].ptr), ctypes.byref(p),
                                                               ctypes.c_void_p(arrs[2].ptr)), buf.handle)
        _native.check(_native.lib().esp_sync(buf.handle), buf.handle)
        out = arrs[2].to_numpy(np.uint8, REFINED_DTYPE.itemsize * n).view(REFINED_DTYPE)
        return [_refined(r) for r in out]
    finally:
        for a in arrs:
            a.free()


@dataclass
class SplicedChainResult:
    """align_spliced's Option<SplicedAlignment> plus why it refused."""

    status: int               # 1 aligned, 0 refused, -1 invalid, -2 capacity
    refusal: int
    n_segments: int
    extension: Optional[Extension] = None
    cigar: List[CigarOp] = field(default_factory=list)
    junctions: List[RefinedJunction] = field(default_factory=list)
    exon_bounds: List[Tuple[int, int, int]] = field(default_factory=list)
    genomic_start: int = 0


class DeviceChainBatch(ContextTimings):
    """align_spliced's chains resident in HBM with their result arenas
    (esp_align_spliced_batch): reads (concatenated, u64 offsets, u32
    lengths), contig and strand per chain, and the flat (qpos, rpos) anchor
    list with its u64 prefix offsets."""

    def __init__(self, dref: DeviceReference, rb, ro, rl, contigs, minus, flat, ao):
        self.buf = buf = dref.buf
        self.dref = dref
        n = self.npairs = int(rl.shape[0])
        self.ao = ao
        caps = 2 * rl.astype(np.uint64) + 5
        self.co = np.zeros(n + 1, dtype=np.uint64)
        self.co[1:] = np.cumsum(caps)
        self.so = np.zeros(n + 1, dtype=np.uint64)
        self.so[1:] = np.cumsum(caps + 2 * np.diff(ao.astype(np.uint64)))
        self.slots = slots = max(int(ao[-1]), 1)
        self.d = d = {name: DeviceArray.from_numpy(buf, arr) for name, arr in
                      (("reads", rb), ("read_off", ro), ("read_len", rl),
                       ("contig", np.asarray(contigs, dtype=np.uint32)),
                       ("minus", np.asarray(minus, dtype=np.uint32)), ("anchors", flat), ("anc_off", ao),
                       ("cig_off", self.co), ("stitch_off", self.so))}
        d["ext"] = DeviceArray(buf, 16 * n)
        d["cigar"] = DeviceArray(buf, 4 * int(self.co[-1]))
        d["cig_n"] = DeviceArray(buf, 4 * n)
        d["stitched"] = DeviceArray(buf, 4 * int(self.so[-1]))
        d["rec"] = DeviceArray(buf, CHAIN_REC_DTYPE.itemsize * n)
        d["junctions"] = DeviceArray(buf, REFINED_DTYPE.itemsize * slots)
        d["bounds"] = DeviceArray(buf, 12 * slots)
        self._cb = _native.CSplicedChainBatch(n, d["reads"].ptr, d["read_off"].ptr, d["read_len"].ptr,
                                              d["contig"].ptr, d["minus"].ptr, d["anchors"].ptr, d["anc_off"].ptr)
        self._out = _native.CBatchOut(d["ext"].ptr, None, d["cigar"].ptr, d["cig_off"].ptr, d["cig_n"].ptr,
                                      None)

    def launch(self, jlib: Optional[JunctionLib], k: int, params: SpliceParams = SpliceParams(),
               extend_params: Optional[ExtendParams] = None) -> None:
        d = self.d
        self._sp, self._ep = params.to_c(), (extend_params or ExtendParams()).to_c()
        _native.check(_native.lib().esp_align_spliced_batch(
            self.buf.handle, self.dref.handle, jlib.handle if jlib else None, ctypes.byref(self._cb), k,
            ctypes.byref(self._sp), ctypes.byref(self._ep), ctypes.byref(self._out),
            ctypes.c_void_p(d["stitched"].ptr), ctypes.c_void_p(d["stitch_off"].ptr), ctypes.c_void_p(d["rec"].ptr),
            ctypes.c_void_p(d["junctions"].ptr), ctypes.c_void_p(d["bounds"].ptr)), self.buf.handle)

    def records(self) -> np.ndarray:
        return self.d["rec"].to_numpy(np.uint8, CHAIN_REC_DTYPE.itemsize * self.npairs).view(CHAIN_REC_DTYPE)

    def results(self) -> List[SplicedChainResult]:
        d, n, ao, co, so = self.d, self.npairs, self.ao, self.co, self.so
        rec = self.records()
        e = d["ext"].to_numpy(np.uint32, 4 * n).reshape(n, 4)
        cig = d["cigar"].to_numpy(np.uint32, int(co[-1]))
        cn = d["cig_n"].to_numpy(np.uint32, n)
        st = d["stitched"].to_numpy(np.uint32, int(so[-1]))
        js = d["junctions"].to_numpy(np.uint8, REFINED_DTYPE.itemsize * self.slots).view(REFINED_DTYPE)
        bd = d["bounds"].to_numpy(np.uint32, 3 * self.slots)
        res = []
        for c in range(n):
            r = rec[c]
            x = SplicedChainResult(int(r["status"]), int(r["refusal"]), int(r["n_segments"]))
            if x.status in (1, -2):
                x.extension = Extension(int(e[c, 0]), int(e[c, 1]), int(e[c, 2]),
                                        cigar_from_words(cig[int(co[c]):int(co[c]) + int(cn[c])]),
                                        _i32(int(e[c, 3])))
                x.genomic_start = int(r["genomic_start"])
                a0, ns = int(ao[c]), int(r["n_segments"])
                x.junctions = [_refined(js[a0 + j]) for j in range(ns - 1)]
                x.exon_bounds = [tuple(int(v) for v in bd[3 * (a0 + s):3 * (a0 + s) + 3]) for s in range(ns)]
                if x.status == 1:
                    x.cigar = cigar_from_words(st[int(so[c]):int(so[c]) + int(r["n_stitched"])])
            res.append(x)
        return res

    def free(self) -> None:
        for a in self.d.values():
            a.free()


def align_spliced_batch(dref: DeviceReference, jlib: Optional[JunctionLib], reads: Sequence[bytes],
                        contigs: Sequence[int], minus: Sequence[bool], anchors: Sequence[Sequence[Tuple[int, int]]],
                        k: int, params: SpliceParams = SpliceParams(),
                        extend_params: Optional[ExtendParams] = None) -> List[SplicedChainResult]:
    """align_spliced per chain: anchors[c] is chain c's (qpos, rpos) list in (rpos, qpos) order."""
    n = len(reads)
    if n == 0:
        return []
    rb, ro, rl = _concat(reads)
    na = np.array([len(a) for a in anchors], dtype=np.uint64)
    ao = np.zeros(n + 1, dtype=np.uint64)
    ao[1:] = np.cumsum(na)
    flat = np.array([x for a in anchors for pr in a for x in pr], dtype=np.uint32) if int(ao[-1]) else \
        np.zeros(2, np.uint32)
    db = DeviceChainBatch(dref, rb, ro, rl, contigs, minus, flat, ao)
    try:
        db.launch(jlib, k, params, extend_params)
        db.sync()
        return db.results()
    finally:
        db.free()


@dataclass
class Reextension:
    status: int          # 1 replaced, 0 kept, 2 spliced, 3 empty window, -1 contig out of range
    pos: int
    score: int
    wstart: int
    wlen: int
    aligned: int
    new_aligned: int
    cigar: List[CigarOp]  # the new extension's CIGAR (empty when no DP ran)


class DeviceReextendBatch(ContextTimings):
    """reextend_with's placements resident in HBM (esp_reextend_batch_run):
    queries (concatenated), contig and pos per placement, and the flat BAM
    CIGAR words with their u64 offsets and u32 counts."""

    def __init__(self, dref: DeviceReference, rb, ro, rl, contigs, positions, cw, cof, cn):
        self.buf = buf = dref.buf
        self.dref = dref
        n = self.npairs = int(rl.shape[0])
        caps = 2 * rl.astype(np.uint64) + 5
        self.oo = np.zeros(n + 1, dtype=np.uint64)
        self.oo[1:] = np.cumsum(caps)
        self.d = d = {name: DeviceArray.from_numpy(buf, arr) for name, arr in
                      (("reads", rb), ("read_off", ro), ("read_len", rl),
                       ("contig", np.asarray(contigs, dtype=np.uint32)),
                       ("pos", np.asarray(positions, dtype=np.uint32)), ("cigar", cw), ("cig_off", cof),
                       ("cig_n", cn), ("out_off", self.oo))}
        d["rec"] = DeviceArray(buf, REEXT_REC_DTYPE.itemsize * n)
        d["out"] = DeviceArray(buf, 4 * int(self.oo[-1]))
        d["out_n"] = DeviceArray(buf, 4 * n)
        self._b = _native.CReextendBatch(n, d["reads"].ptr, d["read_off"].ptr, d["read_len"].ptr, d["contig"].ptr,
                                         d["pos"].ptr, d["cigar"].ptr, d["cig_off"].ptr, d["cig_n"].ptr)

    def launch(self, extend_params: Optional[ExtendParams], edit_dir: int) -> None:
        d = self.d
        self._ep = (extend_params or ExtendParams()).to_c()
        _native.check(_native.lib().esp_reextend_batch_run(
            self.buf.handle, self.dref.handle, ctypes.byref(self._b), ctypes.byref(self._ep), int(edit_dir),
            ctypes.c_void_p(d["rec"].ptr), ctypes.c_void_p(d["out"].ptr), ctypes.c_void_p(d["out_off"].ptr),
            ctypes.c_void_p(d["out_n"].ptr)), self.buf.handle)

    def records(self) -> np.ndarray:
        return self.d["rec"].to_numpy(np.uint8, REEXT_REC_DTYPE.itemsize * self.npairs).view(REEXT_REC_DTYPE)

    def results(self) -> List[Reextension]:
        rec = self.records()
        oo = self.oo
        out = self.d["out"].to_numpy(np.uint32, int(oo[-1]))
        on = self.d["out_n"].to_numpy(np.uint32, self.npairs)
        return [Reextension(int(r["status"]), int(r["pos"]), int(r["score"]), int(r["wstart"]), int(r["wlen"]),
                            int(r["aligned"]), int(r["new_aligned"]),
                            cigar_from_words(out[int(oo[k]):int(oo[k]) + int(on[k])])) for k, r in enumerate(rec)]

    def free(self) -> None:
        for a in self.d.values():
            a.free()


def reextend_batch(dref: DeviceReference, queries: Sequence[bytes], contigs: Sequence[int],
                   positions: Sequence[int], cigars: Sequence[Sequence[CigarOp]],
                   extend_params: Optional[ExtendParams], edit_dir: int) -> List[Reextension]:
    """reextend_with(query, placement, dir) per placement with one dir for the batch."""
    n = len(queries)
    if n == 0:
        return []
    rb, ro, rl = _concat(queries)
    cw = [c.word for cg in cigars for c in cg]
    cn = np.array([len(cg) for cg in cigars], dtype=np.uint32)
    cof = np.zeros(n, dtype=np.uint64)
    if n:
        cof[1:] = np.cumsum(cn, dtype=np.uint64)[:-1]
    db = DeviceReextendBatch(dref, rb, ro, rl, contigs, positions, np.array(cw or [0], dtype=np.uint32), cof, cn)
    try:
        db.launch(extend_params, edit_dir)
        db.sync()
        return db.results()
    finally:
        db.free()


IC_REC_DTYPE = np.dtype([("status", "<i4"), ("reason", "<i4"), ("genomic_start", "<u4"), ("score", "<i4"),
                         ("left_start", "<u4"), ("jstart", "<u4"), ("jend", "<u4"), ("right_end", "<u4"),
                         ("signal", "<i4"), ("support", "<u4"), ("minus", "<u4"), ("n_stitched", "<u4"),
                         ("donor_end", "<i8"), ("acceptor_start", "<i8"), ("vd", "<i4"), ("va", "<i4")])


@dataclass(frozen=True)
class IntronParams:
    """intron_chain.rs:20-42 (defaults :32-42)."""

    min_frag_cov: int = 15
    refine_window: int = 30
    refine_pattern_len: int = 50
    intron_max: int = 105_000


class DeviceIntronChainBatch(ContextTimings):
    """try_intron_chain_placement's reads resident in HBM
    (esp_intron_chain_batch_run): reads (concatenated) and the flat
    (qpos, rpos, contig, minus) anchor pools with their u64 offsets."""

    def __init__(self, dref: DeviceReference, rb, ro, rl, flat, ao):
        self.buf = buf = dref.buf
        self.dref = dref
        n = self.npairs = int(rl.shape[0])
        caps = 2 * rl.astype(np.uint64) + 5
        self.co = np.zeros(n + 1, dtype=np.uint64)
        self.co[1:] = np.cumsum(caps)
        self.so = np.zeros(n + 1, dtype=np.uint64)
        self.so[1:] = np.cumsum(caps + 4)
        self.d = d = {name: DeviceArray.from_numpy(buf, arr) for name, arr in
                      (("reads", rb), ("read_off", ro), ("read_len", rl), ("anchors", flat), ("anc_off", ao),
                       ("cig_off", self.co), ("stitch_off", self.so))}
        d["ext"] = DeviceArray(buf, 16 * n)
        d["cigar"] = DeviceArray(buf, 4 * int(self.co[-1]))
        d["cig_n"] = DeviceArray(buf, 4 * n)
        d["stitched"] = DeviceArray(buf, 4 * int(self.so[-1]))
        d["rec"] = DeviceArray(buf, IC_REC_DTYPE.itemsize * n)
        self._cb = _native.CIntronChainBatch(n, d["reads"].ptr, d["read_off"].ptr, d["read_len"].ptr,
                                             d["anchors"].ptr, d["anc_off"].ptr, int(rl.max()) if n else 1)
        self._out = _native.CBatchOut(d["ext"].ptr, None, d["cigar"].ptr, d["cig_off"].ptr, d["cig_n"].ptr,
                                      None)

    def launch(self, jlib: Optional[JunctionLib], k: int = 15, params: IntronParams = IntronParams(),
               extend_params: Optional[ExtendParams] = None) -> None:
        d = self.d
        self._ip = _native.CIntronParams(params.min_frag_cov, params.refine_window, params.refine_pattern_len,
                                         params.intron_max)
        self._ep = (extend_params or ExtendParams()).to_c()
        _native.check(_native.lib().esp_intron_chain_batch_run(
            self.buf.handle, self.dref.handle, jlib.handle if jlib else None, ctypes.byref(self._cb), k,
            ctypes.byref(self._ip), ctypes.byref(self._ep), ctypes.byref(self._out),
            ctypes.c_void_p(d["stitched"].ptr), ctypes.c_void_p(d["stitch_off"].ptr),
            ctypes.c_void_p(d["rec"].ptr)), self.buf.handle)

    def records(self) -> np.ndarray:
        return self.d["rec"].to_numpy(np.uint8, IC_REC_DTYPE.itemsize * self.npairs).view(IC_REC_DTYPE)

    def results(self):
        d, n, co, so = self.d, self.npairs, self.co, self.so
        rec = self.records()
        e = d["ext"].to_numpy(np.uint32, 4 * n).reshape(n, 4)
        cig = d["cigar"].to_numpy(np.uint32, int(co[-1]))
        cn = d["cig_n"].to_numpy(np.uint32, n)
        st = d["stitched"].to_numpy(np.uint32, int(so[-1]))
        exts, stitched = [], []
        for c in range(n):
            ran = int(rec[c]["left_start"]) or int(rec[c]["status"]) == 1
            exts.append(Extension(int(e[c, 0]), int(e[c, 1]), int(e[c, 2]),
                                  cigar_from_words(cig[int(co[c]):int(co[c]) + int(cn[c])]), _i32(int(e[c, 3])))
                        if ran else None)
            stitched.append(cigar_from_words(st[int(so[c]):int(so[c]) + int(rec[c]["n_stitched"])]))
        return rec, exts, stitched

    def free(self) -> None:
        for a in self.d.values():
            a.free()


def intron_chain_batch(dref: DeviceReference, jlib: Optional[JunctionLib], reads: Sequence[bytes],
                       anchors: Sequence[Sequence[Tuple[int, int, int, int]]], k: int = 15,
                       params: IntronParams = IntronParams(), extend_params: Optional[ExtendParams] = None):
    """try_intron_chain_placement per read; anchors[r] = the read's (qpos, rpos, contig, minus) pool.
    Returns (records, extensions, stitched CIGARs)."""
    n = len(reads)
    if n == 0:
        return np.zeros(0, IC_REC_DTYPE), [], []
    rb, ro, rl = _concat(reads)
    na = np.array([len(a) for a in anchors], dtype=np.uint64)
    ao = np.zeros(n + 1, dtype=np.uint64)
    ao[1:] = np.cumsum(na)
    flat = np.array([x for a in anchors for an in a for x in an], dtype=np.uint32) if int(ao[-1]) else \
        np.zeros(4, np.uint32)
    db = DeviceIntronChainBatch(dref, rb, ro, rl, flat, ao)
    try:
        db.launch(jlib, k, params, extend_params)
        db.sync()
        return db.results()
    finally:
        db.free()
