"""ctypes view of oracle/lib/libesp_oracle.so (the C restatement in esp_oracle.c).

Test infrastructure only: this is the checker, never the thing measured.
Function names follow the reference (crates/map/src/extend.rs,
crates/map/src/myers_ea.rs, crates/map/src/splice.rs).
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional, Sequence, Tuple

import numpy as np

from esperanto_amd.records import (CExt, CParams, CTouch, DiagHint, Extension, ExtendParams,
                                   TouchRec, cigar_capacity, cigar_from_words)

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libesp_oracle.so")

_lib = None


class CRef(ctypes.Structure):
    """espo_ref: the resident-reference layout of fasta.rs:62-71, flattened."""
    _fields_ = [("ncontig", ctypes.c_size_t), ("len", ctypes.c_void_p), ("packed", ctypes.c_void_p),
                ("packed_off", ctypes.c_void_p), ("iv", ctypes.c_void_p),
                ("iv_off", ctypes.c_void_p)]


class CBfRows(ctypes.Structure):
    """espo_bf_rows_out: what the brute-force matrices say of the row scan."""
    _fields_ = [("best", ctypes.c_int32), ("bi", ctypes.c_int64), ("bj", ctypes.c_int64),
                ("touched", ctypes.c_int32), ("row", ctypes.c_int64), ("left", ctypes.c_int32),
                ("clamped", ctypes.c_int32)]


class CSpliceParams(ctypes.Structure):
    """The SpliceParams fields rescore_with reads (splice.rs:21-52)."""
    _fields_ = [("known_bonus", ctypes.c_int32), ("gt_ag_bonus", ctypes.c_int32),
                ("gc_ag_bonus", ctypes.c_int32), ("at_ac_bonus", ctypes.c_int32),
                ("noncanonical_penalty", ctypes.c_int32)]


SPLICE_DEFAULTS = (6, 4, 2, 1, 8)  # splice.rs:40-52: known, GT-AG, GC-AG, AT-AC, non-canonical

FAST_DTYPE = np.dtype([("status", "<i4"), ("score", "<i4"), ("wstart", "<u4"), ("ref_start", "<u4"),
                       ("read_start", "<u4"), ("read_end", "<u4"), ("n_ops", "<u4"), ("cigar", "<u4", 3)])
STRICT_DTYPE = np.dtype([("score", "<i4"), ("aligned", "<u4"), ("matches", "<u4"), ("mm", "<u4"),
                         ("ea_ag", "<u4"), ("ea_tc", "<u4"), ("ins", "<u4"), ("del", "<u4"),
                         ("clipped", "<u4"), ("n_splice", "<u4"), ("splice_score", "<i4")])


def build() -> None:
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE, "-B"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library not built: {LIB_PATH} (run make -C oracle)")
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.POINTER
        u8p, u32p, u64p, i64p, i32p = (P(ctypes.c_uint8), P(ctypes.c_uint32), P(ctypes.c_uint64),
                                       P(ctypes.c_int64), P(ctypes.c_int32))
        sz = ctypes.c_size_t
        L.espo_run_band.argtypes = [u8p, sz, u8p, sz, P(CParams), ctypes.c_int, ctypes.c_int,
                                    ctypes.c_int64, ctypes.c_int64, P(CExt), P(CTouch), u32p, sz,
                                    P(sz)]
        L.espo_band.argtypes = [u8p, sz, u8p, sz, P(CParams), ctypes.c_int, ctypes.c_int64,
                                ctypes.c_int64, P(CExt), P(CTouch), u32p, sz, P(sz)]
        L.espo_extend.argtypes = [u8p, sz, u8p, sz, P(CParams), P(CExt), u32p, sz, P(sz)]
        L.espo_extend_hint.argtypes = [u8p, sz, u8p, sz, P(CParams), ctypes.c_int64, P(CExt), u32p,
                                       sz, P(sz), i32p]
        L.espo_myers_best_end.argtypes = [u8p, sz, u8p, sz, i32p, P(sz)]
        L.espo_myers_best_start.argtypes = [u8p, sz, u8p, sz, i32p, P(sz)]
        L.espo_stitch_cigar.argtypes = [u32p, sz, ctypes.c_uint32, u32p, sz, u32p, sz, u32p, sz,
                                        P(sz)]
        L.espo_bf_local_affine.argtypes = [u8p, sz, u8p, sz, P(CParams), ctypes.c_int,
                                           ctypes.c_int64, ctypes.c_int64]
        L.espo_bf_local_affine.restype = ctypes.c_int32
        L.espo_bf_ea_edit.argtypes = [u8p, sz, u8p, sz, i32p, P(sz)]
        L.espo_run_band_cell.argtypes = [u8p, sz, u8p, sz, P(CParams), ctypes.c_int, ctypes.c_int,
                                         ctypes.c_int64, ctypes.c_int64, P(CExt), P(CTouch), u32p, sz,
                                         P(sz), i64p]
        L.espo_set_mutation.argtypes = [ctypes.c_int]
        L.espo_set_mutation.restype = None
        L.espo_bf_band_rows.argtypes = [u8p, sz, u8p, sz, P(CParams), ctypes.c_int, ctypes.c_int64,
                                        ctypes.c_int64, ctypes.c_int64, P(CBfRows)]
        L.espo_bf_band_rows.restype = None
        L.espo_band_batch.argtypes = [sz, u8p, u64p, u32p, u8p, u64p, u32p, i64p, i64p,
                                      P(CParams), ctypes.c_int, P(CExt), P(CTouch), u32p, u64p,
                                      u32p, i32p, ctypes.c_int]
        L.espo_myers_batch.argtypes = [sz, u8p, u64p, u32p, u8p, u64p, u32p, i32p, u64p, u64p,
                                       ctypes.c_int]
        vp = ctypes.c_void_p
        L.espo_decode.argtypes = [P(CRef), ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                  ctypes.c_int, u8p]
        L.espo_revcomp.argtypes = [u8p, sz, u8p]
        L.espo_mate_window.argtypes = [P(CRef), u8p, sz, ctypes.c_uint32, ctypes.c_uint32,
                                       P(CParams), vp, u32p, sz, P(sz)]
        L.espo_spliced.argtypes = [P(CRef), u8p, sz, ctypes.c_uint32, u32p, sz, ctypes.c_int64,
                                   P(CParams), P(CExt), u32p, sz, P(sz), u32p, sz, vp]
        L.espo_ext_window.argtypes = [ctypes.c_uint32] * 6 + [u32p, u32p]
        L.espo_unspliced.argtypes = [P(CRef), u8p, sz, ctypes.c_uint32, ctypes.c_int,
                                     ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                     ctypes.c_uint32, ctypes.c_int, P(CParams), vp, u32p, sz,
                                     P(sz)]
        L.espo_mate_batch.argtypes = [P(CRef), sz, vp, vp, vp, vp, vp, P(CParams), vp, vp, vp, vp,
                                      ctypes.c_int]
        L.espo_spliced_batch.argtypes = [P(CRef), sz, vp, vp, vp, vp, vp, vp, vp, P(CParams), vp,
                                         vp, vp, vp, vp, vp, vp, ctypes.c_int]
        L.espo_intron_length_penalty.argtypes = [ctypes.c_uint32]
        L.espo_intron_length_penalty.restype = ctypes.c_int32
        L.espo_splice_signal_score.argtypes = [ctypes.c_uint8] * 4 + [ctypes.c_int, P(CSpliceParams)]
        L.espo_splice_signal_score.restype = ctypes.c_int32
        L.espo_rescore.argtypes = [P(CRef), u8p, sz, ctypes.c_uint32, ctypes.c_uint32, u32p, sz,
                                   ctypes.c_int, P(CParams), P(CSpliceParams), u32p, sz, vp]
        L.espo_recount.argtypes = [P(CRef), u8p, sz, ctypes.c_uint32, ctypes.c_uint32, u32p, sz, u32p]
        L.espo_kmer_support.argtypes = [P(CRef), u8p, sz, ctypes.c_uint32, ctypes.c_uint32, u32p, sz,
                                        ctypes.c_int, ctypes.c_uint32]
        L.espo_kmer_support.restype = ctypes.c_uint32
        L.espo_walk_batch.argtypes = [P(CRef), sz] + [vp] * 9 + [P(CParams), P(CSpliceParams), vp, sz,
                                                               ctypes.c_uint32, vp, vp, vp, ctypes.c_int]
        L.espo_walk_batch.restype = None
        L.espo_path_scores.argtypes = [sz, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, P(CParams), vp, vp,
                                       ctypes.c_int]
        L.espo_path_scores.restype = None
        L.espo_engine_myers.argtypes = [u8p, sz, u8p, sz, ctypes.c_int]
        L.espo_engine_myers.restype = ctypes.c_int64
        L.espo_fast_lane.argtypes = [P(CRef), u8p, ctypes.c_uint32, ctypes.c_uint32, u32p, sz,
                                     ctypes.c_uint32, ctypes.c_uint32, P(CParams), vp]
        _lib = L
    return _lib


def _buf(b: bytes):
    b = bytes(b)
    arr = (ctypes.c_uint8 * max(len(b), 1)).from_buffer_copy(b if b else b"\0")
    return arr, len(b)


def _params(p: Optional[ExtendParams]) -> CParams:
    return (p or ExtendParams()).to_c()


def run_band(read: bytes, window: bytes, params: Optional[ExtendParams] = None, is16: bool = True,
             band: Optional[Tuple[int, int]] = None) -> Tuple[Extension, TouchRec]:
    """extend.rs:606 run_band::<i16|i32>; band = (offset, half) or None (full)."""
    r, m = _buf(read)
    w, n = _buf(window)
    cap = cigar_capacity(m)
    cig = (ctypes.c_uint32 * cap)()
    e, t, nops = CExt(), CTouch(), ctypes.c_size_t()
    off, half = band if band is not None else (0, 0)
    rc = lib().espo_run_band(r, m, w, n, ctypes.byref(_params(params)), int(is16),
                             int(band is not None), off, half, ctypes.byref(e), ctypes.byref(t),
                             cig, cap, ctypes.byref(nops))
    assert rc == 0, rc
    return Extension.from_c(e, cig[: nops.value]), TouchRec.from_c(t)


def run_band_cell(read: bytes, window: bytes, params: Optional[ExtendParams] = None, is16: bool = True,
                  band: Optional[Tuple[int, int]] = None):
    """run_band plus the best cell (best_i, best_j) its traceback starts from."""
    r, m = _buf(read)
    w, n = _buf(window)
    cap = cigar_capacity(m)
    cig = (ctypes.c_uint32 * cap)()
    e, t, nops = CExt(), CTouch(), ctypes.c_size_t()
    cell = (ctypes.c_int64 * 2)()
    off, half = band if band is not None else (0, 0)
    rc = lib().espo_run_band_cell(r, m, w, n, ctypes.byref(_params(params)), int(is16),
                                  int(band is not None), off, half, ctypes.byref(e), ctypes.byref(t),
                                  cig, cap, ctypes.byref(nops), cell)
    assert rc == 0, rc
    return Extension.from_c(e, cig[: nops.value]), TouchRec.from_c(t), (cell[0], cell[1])


def set_mutation(k: int) -> None:
    """Plants mistake k in run_band (0 restores it); for tests that must catch them."""
    lib().espo_set_mutation(int(k))


def bf_band_rows(read: bytes, window: bytes, params: Optional[ExtendParams] = None,
                 band: Optional[Tuple[int, int]] = None, diag_cap: int = -1) -> dict:
    """Brute-force matrices: best score, first row-major best cell, TouchRec."""
    r, m = _buf(read)
    w, n = _buf(window)
    o = CBfRows()
    off, half = band if band is not None else (0, 0)
    lib().espo_bf_band_rows(r, m, w, n, ctypes.byref(_params(params)), int(band is not None), off, half,
                            int(diag_cap), ctypes.byref(o))
    return {"best": o.best, "cell": (o.bi, o.bj),
            "touch": TouchRec(bool(o.touched), o.row, bool(o.left), bool(o.clamped))}


def band(read: bytes, window: bytes, params: Optional[ExtendParams] = None,
         band: Optional[Tuple[int, int]] = None) -> Tuple[Extension, TouchRec]:
    """extend.rs:293 run_dispatch(.., Some((offset,1,1,half)) | None)."""
    r, m = _buf(read)
    w, n = _buf(window)
    cap = cigar_capacity(m)
    cig = (ctypes.c_uint32 * cap)()
    e, t, nops = CExt(), CTouch(), ctypes.c_size_t()
    off, half = band if band is not None else (0, 0)
    rc = lib().espo_band(r, m, w, n, ctypes.byref(_params(params)), int(band is not None), off,
                         half, ctypes.byref(e), ctypes.byref(t), cig, cap, ctypes.byref(nops))
    assert rc == 0, rc
    return Extension.from_c(e, cig[: nops.value]), TouchRec.from_c(t)


def extend(read: bytes, window: bytes, params: Optional[ExtendParams] = None) -> Extension:
    """extend.rs:347"""
    return band(read, window, params, None)[0]


def extend_hint(read: bytes, window: bytes, params: Optional[ExtendParams] = None,
                hint: DiagHint = DiagHint(0), with_path: bool = False):
    """extend.rs:392.  with_path also returns 0 band kept / 1 escalated / 2 empty / 3 band>=n."""
    r, m = _buf(read)
    w, n = _buf(window)
    cap = cigar_capacity(m)
    cig = (ctypes.c_uint32 * cap)()
    e, nops, esc = CExt(), ctypes.c_size_t(), ctypes.c_int32()
    rc = lib().espo_extend_hint(r, m, w, n, ctypes.byref(_params(params)), int(hint.offset),
                                ctypes.byref(e), cig, cap, ctypes.byref(nops), ctypes.byref(esc))
    assert rc == 0, rc
    ext = Extension.from_c(e, cig[: nops.value])
    return (ext, esc.value) if with_path else ext


def infix_best_end(read: bytes, text: bytes) -> Tuple[int, int]:
    """myers_ea.rs:271 long::infix_best_end"""
    r, m = _buf(read)
    t, n = _buf(text)
    d, e = ctypes.c_int32(), ctypes.c_size_t()
    lib().espo_myers_best_end(r, m, t, n, ctypes.byref(d), ctypes.byref(e))
    return d.value, e.value


def infix(read: bytes, text: bytes) -> int:
    """myers_ea.rs:255 long::infix"""
    return infix_best_end(read, text)[0]


def infix_best_start(read: bytes, text: bytes) -> Tuple[int, int]:
    """myers_ea.rs:288 long::infix_best_start"""
    r, m = _buf(read)
    t, n = _buf(text)
    d, s = ctypes.c_int32(), ctypes.c_size_t()
    lib().espo_myers_best_start(r, m, t, n, ctypes.byref(d), ctypes.byref(s))
    return d.value, s.value


def stitch_cigar(cigar_words: Sequence[int], p_start: int, exon_bounds: Sequence[Tuple[int, int, int]],
                 skips: Sequence[int]) -> list:
    """splice.rs:435 stitch_cigar; skips[k] = junction k's end - start."""
    ops = (ctypes.c_uint32 * max(len(cigar_words), 1))(*cigar_words)
    flat = [x for b in exon_bounds for x in b]
    bnd = (ctypes.c_uint32 * max(len(flat), 1))(*flat)
    sk = (ctypes.c_uint32 * max(len(skips), 1))(*skips)
    cap = 2 * len(cigar_words) + 2 * len(exon_bounds) + 4
    out = (ctypes.c_uint32 * cap)()
    n = ctypes.c_size_t()
    rc = lib().espo_stitch_cigar(ops, len(cigar_words), p_start, bnd, len(exon_bounds), sk,
                                 len(skips), out, cap, ctypes.byref(n))
    assert rc == 0
    return cigar_from_words(out[: n.value])


def bf_local_affine(read: bytes, window: bytes, params: Optional[ExtendParams] = None,
                    band: Optional[Tuple[int, int]] = None) -> int:
    r, m = _buf(read)
    w, n = _buf(window)
    off, half = band if band is not None else (0, 0)
    return lib().espo_bf_local_affine(r, m, w, n, ctypes.byref(_params(params)),
                                      int(band is not None), off, half)


def bf_ea_edit(read: bytes, text: bytes) -> Tuple[int, int]:
    r, m = _buf(read)
    t, n = _buf(text)
    d, e = ctypes.c_int32(), ctypes.c_size_t()
    lib().espo_bf_ea_edit(r, m, t, n, ctypes.byref(d), ctypes.byref(e))
    return d.value, e.value


def _ptr(a: np.ndarray, ct):
    return a.ctypes.data_as(ctypes.POINTER(ct))


def band_batch(batch, params: Optional[ExtendParams] = None, mode: int = 0, nthreads: int = 1):
    """Runs a PairBatch (esperanto_amd.batch) on the CPU restatement.

    mode 0 = band (run_dispatch Some((offset,1,1,half))), 1 = full matrix
    (extend), 2 = extend_hint with offset as the hint.  Returns
    (ext[n,4] int64 view, touch[n,4], cig_words, cig_off, cig_n, aux).
    """
    n = batch.npairs
    out = (CExt * max(n, 1))()
    touch = (CTouch * max(n, 1))()
    cig_off = batch.cigar_offsets()
    cig = np.zeros(int(cig_off[-1]) + 1, dtype=np.uint32)
    cig_n = np.zeros(n, dtype=np.uint32)
    aux = np.zeros(n, dtype=np.int32)
    rc = lib().espo_band_batch(n, _ptr(batch.reads, ctypes.c_uint8), _ptr(batch.read_off, ctypes.c_uint64),
                               _ptr(batch.read_len, ctypes.c_uint32), _ptr(batch.wins, ctypes.c_uint8),
                               _ptr(batch.win_off, ctypes.c_uint64), _ptr(batch.win_len, ctypes.c_uint32),
                               _ptr(batch.offset, ctypes.c_int64), _ptr(batch.half, ctypes.c_int64),
                               ctypes.byref(_params(params)), mode, out, touch,
                               _ptr(cig, ctypes.c_uint32), _ptr(cig_off, ctypes.c_uint64),
                               _ptr(cig_n, ctypes.c_uint32), _ptr(aux, ctypes.c_int32), nthreads)
    assert rc == 0, rc
    ext = np.ctypeslib.as_array(out)[:n] if n else np.zeros(0)
    ext = np.array([(e.read_start, e.read_end, e.ref_start, e.score) for e in out[:n]],
                   dtype=np.int64).reshape(n, 4)
    tch = np.array([(t.touched, t.row, t.left, t.clamped) for t in touch[:n]],
                   dtype=np.int64).reshape(n, 4)
    return ext, tch, cig, cig_off, cig_n, aux


def myers_batch(batch, with_start: bool = True, nthreads: int = 1):
    """long::infix_best_end (+ infix_best_start) over a PairBatch (read, text=window)."""
    n = batch.npairs
    dist = np.zeros(n, dtype=np.int32)
    end = np.zeros(n, dtype=np.uint64)
    start = np.zeros(n, dtype=np.uint64) if with_start else None
    lib().espo_myers_batch(n, _ptr(batch.reads, ctypes.c_uint8), _ptr(batch.read_off, ctypes.c_uint64),
                           _ptr(batch.read_len, ctypes.c_uint32), _ptr(batch.wins, ctypes.c_uint8),
                           _ptr(batch.win_off, ctypes.c_uint64), _ptr(batch.win_len, ctypes.c_uint32),
                           _ptr(dist, ctypes.c_int32), _ptr(end, ctypes.c_uint64),
                           _ptr(start, ctypes.c_uint64) if with_start else None, nthreads)
    return dist, end, start


# ------------------------------------------------------------ reference windows

class OracleRef:
    """Keeps the numpy arrays of an esperanto_amd.reference.Reference alive
    behind the espo_ref view."""

    def __init__(self, ref):
        self.arrs = [np.ascontiguousarray(ref.lens, dtype=np.uint32),
                     np.ascontiguousarray(ref.packed, dtype=np.uint8) if ref.packed.size else np.zeros(1, np.uint8),
                     np.ascontiguousarray(ref.packed_off, dtype=np.uint64),
                     np.ascontiguousarray(ref.iv.reshape(-1), dtype=np.uint32) if ref.iv.size else np.zeros(2, np.uint32),
                     np.ascontiguousarray(ref.iv_off, dtype=np.uint64)]
        self.c = CRef(ref.ncontig, *(a.ctypes.data for a in self.arrs))


def decode(oref: OracleRef, contig: int, start: int, end: int, rc: bool = False) -> Optional[bytes]:
    """Contig::decode_into (fasta.rs:77-82), revcomp_into of it when rc; None out of range."""
    n = max(end - start, 0)
    out = (ctypes.c_uint8 * max(n, 1))()
    bad = lib().espo_decode(ctypes.byref(oref.c), contig, start, end, int(rc), out)
    return None if bad else bytes(out[:n])


def revcomp(seq: bytes) -> bytes:
    """align.rs:2525 revcomp_into"""
    s, m = _buf(seq)
    out = (ctypes.c_uint8 * max(m, 1))()
    lib().espo_revcomp(s, m, out)
    return bytes(out[:m])


MATE_REC_DTYPE = np.dtype([("placed", "<i4"), ("strand", "<i4"), ("pos", "<u4"), ("score", "<i4"),
                           ("dist", "<i4"), ("wstart", "<u4"), ("wlen", "<u4"),
                           ("read_start", "<u4"), ("read_end", "<u4")])
SPLICE_REC_DTYPE = np.dtype([("status", "<i4"), ("genomic_start", "<u4"), ("pseudo_len", "<u4"),
                             ("n_stitched", "<u4")])


def mate_window(oref: OracleRef, read: bytes, contig: int, mate_pos: int,
                params: Optional[ExtendParams] = None):
    """align.rs:851 mate_window_placement: (record, cigar ops)."""
    r, m = _buf(read)
    rec = np.zeros(1, dtype=MATE_REC_DTYPE)
    cap = cigar_capacity(m)
    cig = (ctypes.c_uint32 * cap)()
    n = ctypes.c_size_t()
    rc = lib().espo_mate_window(ctypes.byref(oref.c), r, m, contig, mate_pos,
                                ctypes.byref(_params(params)), rec.ctypes.data, cig, cap,
                                ctypes.byref(n))
    assert rc == 0, rc
    return rec[0], cigar_from_words(cig[: n.value])


def spliced(oref: OracleRef, read: bytes, contig: int, exons: Sequence[Tuple[int, int]], hint: int,
            params: Optional[ExtendParams] = None):
    """Spliced window DP (splice.rs:589-664): (record, Extension, stitched ops)."""
    r, m = _buf(read)
    flat = [x for e in exons for x in e]
    ex = (ctypes.c_uint32 * max(len(flat), 1))(*flat)
    cap = cigar_capacity(m)
    cig = (ctypes.c_uint32 * cap)()
    scap = cap + 2 * len(exons)
    st = (ctypes.c_uint32 * scap)()
    e, n = CExt(), ctypes.c_size_t()
    rec = np.zeros(1, dtype=SPLICE_REC_DTYPE)
    rc = lib().espo_spliced(ctypes.byref(oref.c), r, m, contig, ex, len(exons), hint,
                            ctypes.byref(_params(params)), ctypes.byref(e), cig, cap,
                            ctypes.byref(n), st, scap, rec.ctypes.data)
    assert rc == 0, rc
    return rec[0], Extension.from_c(e, cig[: n.value]), cigar_from_words(st[: rec[0]["n_stitched"]])


def mate_batch(oref: OracleRef, req, params: Optional[ExtendParams] = None, nthreads: int = 1):
    """mate_window_placement over esperanto_amd.rescue.MateRequests:
    (records, cigar words, cig_off, cig_n)."""
    n = req.npairs
    rec = np.zeros(max(n, 1), dtype=MATE_REC_DTYPE)
    off = req.cigar_offsets()
    cig = np.zeros(int(off[-1]) + 1, dtype=np.uint32)
    cig_n = np.zeros(max(n, 1), dtype=np.uint32)
    arrs = [np.ascontiguousarray(a) for a in (req.reads, req.read_off, req.read_len, req.contig,
                                              req.mate_pos)]
    lib().espo_mate_batch(ctypes.byref(oref.c), n, *(a.ctypes.data for a in arrs),
                          ctypes.byref(_params(params)), rec.ctypes.data, cig.ctypes.data,
                          off.ctypes.data, cig_n.ctypes.data, nthreads)
    return rec[:n], cig, off, cig_n[:n]


def spliced_batch(oref: OracleRef, req, params: Optional[ExtendParams] = None, nthreads: int = 1):
    """Spliced window DP over esperanto_amd.splice.SpliceRequests:
    (records, ext[n,4], cigar, cig_off, cig_n, stitched, stitch_off)."""
    n = req.npairs
    rec = np.zeros(max(n, 1), dtype=SPLICE_REC_DTYPE)
    off = req.cigar_offsets()
    soff = req.stitch_offsets()
    cig = np.zeros(int(off[-1]) + 1, dtype=np.uint32)
    st = np.zeros(int(soff[-1]) + 1, dtype=np.uint32)
    cig_n = np.zeros(max(n, 1), dtype=np.uint32)
    ext = np.zeros((max(n, 1), 4), dtype=np.uint32)
    exons = np.ascontiguousarray(req.exons.reshape(-1), dtype=np.uint32)
    if exons.size == 0:
        exons = np.zeros(2, np.uint32)
    arrs = [np.ascontiguousarray(a) for a in (req.reads, req.read_off, req.read_len, req.contig,
                                              req.hint)]
    lib().espo_spliced_batch(ctypes.byref(oref.c), n, *(a.ctypes.data for a in arrs),
                             exons.ctypes.data, np.ascontiguousarray(req.exon_off).ctypes.data,
                             ctypes.byref(_params(params)), ext.ctypes.data, cig.ctypes.data,
                             off.ctypes.data, cig_n.ctypes.data, st.ctypes.data, soff.ctypes.data,
                             rec.ctypes.data, nthreads)
    e = ext[:n].astype(np.int64)
    e[:, 3] = ext[:n, 3].view(np.int32)
    return rec[:n], e, cig, off, cig_n[:n], st, soff


UNSPLICED_REC_DTYPE = np.dtype([("status", "<i4"), ("strand", "<i4"), ("pos", "<u4"),
                                ("score", "<i4"), ("dist", "<i4"), ("wstart", "<u4"),
                                ("wlen", "<u4")])


def ext_window(ref_span, read_span, read_len: int, contig_len: int):
    """align.rs:33 continuous_extension_window."""
    a, b = ctypes.c_uint32(), ctypes.c_uint32()
    lib().espo_ext_window(ref_span[0], ref_span[1], read_span[0], read_span[1], read_len,
                          contig_len, ctypes.byref(a), ctypes.byref(b))
    return a.value, b.value


def unspliced(oref: OracleRef, read: bytes, contig: int, minus: bool, ref_span, read_span,
              allow_dp: bool, params: Optional[ExtendParams] = None):
    """extend_unspliced_chain after the fast lane (align.rs:1770-1919): (record, ops)."""
    r, m = _buf(read)
    rec = np.zeros(1, dtype=UNSPLICED_REC_DTYPE)
    cap = cigar_capacity(m)
    cig = (ctypes.c_uint32 * cap)()
    n = ctypes.c_size_t()
    rc = lib().espo_unspliced(ctypes.byref(oref.c), r, m, contig, int(minus), ref_span[0],
                              ref_span[1], read_span[0], read_span[1], int(allow_dp),
                              ctypes.byref(_params(params)), rec.ctypes.data, cig, cap,
                              ctypes.byref(n))
    assert rc == 0, rc
    return rec[0], cigar_from_words(cig[: n.value])


def unspliced_batch(oref: OracleRef, req, params: Optional[ExtendParams] = None):
    """unspliced() over esperanto_amd.unspliced.ChainRequests: (records, cigars)."""
    recs, cigs = [], []
    for k in range(req.npairs):
        a = int(req.read_off[k])
        read = req.reads[a:a + int(req.read_len[k])].tobytes()
        rec, cig = unspliced(oref, read, int(req.contig[k]), bool(req.minus[k]),
                             (int(req.ref_start[k]), int(req.ref_end[k])),
                             (int(req.read_start[k]), int(req.read_end[k])),
                             bool(req.allow_dp[k]), params)
        recs.append(rec)
        cigs.append(cig)
    return recs, cigs


# ------------------------------------------------------------ placement walks

def intron_length_penalty(n: int) -> int:
    """splice.rs:119 intron_length_penalty"""
    return int(lib().espo_intron_length_penalty(n))


def splice_signal_score(donor: bytes, acceptor: bytes, minus: bool, sp=SPLICE_DEFAULTS) -> int:
    """splice.rs:78 splice_signal(..).score(&SpliceParams)"""
    return int(lib().espo_splice_signal_score(donor[0], donor[1], acceptor[0], acceptor[1], int(minus),
                                              ctypes.byref(CSpliceParams(*sp))))


def _cigar_arr(cigar):
    words = [int(w) for w in cigar]
    return (ctypes.c_uint32 * max(len(words), 1))(*words), len(words)


def _junction_arr(junctions):
    flat = [int(x) for t in sorted(junctions or []) for x in t]
    return (ctypes.c_uint32 * max(len(flat), 1))(*flat), len(flat) // 3


def rescore(oref: OracleRef, query: bytes, contig: int, pos: int, cigar, dir: int = 0,
            params: Optional[ExtendParams] = None, sp=SPLICE_DEFAULTS, junctions=None):
    """align.rs:1060 Aligner::rescore_with -> StrictScore record (STRICT_DTYPE);
    cigar is BAM words, junctions (contig, start, end) triples."""
    q, m = _buf(query)
    c, nc = _cigar_arr(cigar)
    j, nj = _junction_arr(junctions)
    out = np.zeros(1, dtype=STRICT_DTYPE)
    rc = lib().espo_rescore(ctypes.byref(oref.c), q, m, contig, pos, c, nc, dir,
                            ctypes.byref(_params(params)), ctypes.byref(CSpliceParams(*sp)), j, nj,
                            out.ctypes.data)
    assert rc == 0, rc
    return out[0]


def recount(oref: OracleRef, query: bytes, contig: int, pos: int, cigar):
    """align.rs:2473 recount_mm_ea_in -> (mm, ea_ag, ea_tc)"""
    q, m = _buf(query)
    c, nc = _cigar_arr(cigar)
    out = (ctypes.c_uint32 * 3)()
    rc = lib().espo_recount(ctypes.byref(oref.c), q, m, contig, pos, c, nc, out)
    assert rc == 0, rc
    return tuple(out)


def kmer_support(oref: OracleRef, query: bytes, contig: int, pos: int, cigar, dir: int = 0,
                 k: int = 15) -> int:
    """align.rs:1231 Aligner::exact_kmer_support (k from SeedParams, 15 by default)"""
    q, m = _buf(query)
    c, nc = _cigar_arr(cigar)
    return int(lib().espo_kmer_support(ctypes.byref(oref.c), q, len(query), contig, pos, c, nc, dir, k))


def walk_batch(oref: OracleRef, pl, params: Optional[ExtendParams] = None, sp=SPLICE_DEFAULTS,
               junctions=None, k: int = 15, nthreads: int = 1):
    """rescore_with + recount + exact_kmer_support over a Placements batch
    (esperanto_amd.walk.Placements): (STRICT_DTYPE records, (n, 4) counts
    mm/ea_ag/ea_tc/k-mers, status)."""
    n = pl.n
    arrs = [np.ascontiguousarray(getattr(pl, f)) for f in
            ("reads", "read_off", "read_len", "contig", "pos", "cigar", "cig_off", "cig_n", "dir")]
    flat = np.array([x for t in sorted(junctions or []) for x in t] or [0], dtype=np.uint32)
    strict = np.zeros(max(n, 1), dtype=STRICT_DTYPE)
    counts = np.zeros((max(n, 1), 4), dtype=np.uint32)
    status = np.zeros(max(n, 1), dtype=np.int32)
    lib().espo_walk_batch(ctypes.byref(oref.c), n, *(a.ctypes.data for a in arrs), ctypes.byref(_params(params)),
                          ctypes.byref(CSpliceParams(*sp)), flat.ctypes.data, len(flat) // 3 if junctions else 0,
                          k, strict.ctypes.data, counts.ctypes.data, status.ctypes.data, nthreads)
    return strict[:n], counts[:n], status[:n]


def fast_lane(oref: OracleRef, query: bytes, contig: int, anchors, k: int = 15, w: int = 5,
              params: Optional[ExtendParams] = None):
    """align.rs:2274 Aligner::try_fast_lane: None when the lane refuses, else
    (Extension, wstart).  anchors: the chain's (qpos, rpos) pairs in its
    (rpos, qpos) order; query in the chain's orientation."""
    q, m = _buf(query)
    flat = [int(x) for a in anchors for x in a]
    arr = (ctypes.c_uint32 * max(len(flat), 1))(*flat)
    out = np.zeros(1, dtype=FAST_DTYPE)
    rc = lib().espo_fast_lane(ctypes.byref(oref.c), q, len(query), contig, arr, len(flat) // 2, k, w,
                              ctypes.byref(_params(params)), out.ctypes.data)
    assert rc == 0, rc
    r = out[0]
    if r["status"] != 1:
        return None
    ext = Extension(int(r["read_start"]), int(r["read_end"]), int(r["ref_start"]),
                    cigar_from_words(r["cigar"][: int(r["n_ops"])].tolist()), int(r["score"]))
    return ext, int(r["wstart"])


def engine_myers(pattern: bytes, text: bytes, global_: bool = False) -> int:
    """engine/src/myers.rs: infix (infix_min) or global EA distance; -1 for a
    pattern length outside 1..=256."""
    p, m = _buf(pattern)
    t, n = _buf(text)
    return int(lib().espo_engine_myers(p, m, t, n, int(global_)))


EXT_DTYPE = np.dtype([("read_start", "<u4"), ("read_end", "<u4"), ("ref_start", "<u4"), ("score", "<i4")])


def path_scores(batch, ext, cig, cig_off, cig_n, params: Optional[ExtendParams] = None, nthreads: int = 1):
    """Affine score of every reported alignment path, rescored from its CIGAR
    over the pair's read and window, and whether the path fits the pair
    (espo_path_scores).  ext: (n, 4) read_start, read_end, ref_start, score."""
    n = batch.npairs
    e = np.zeros(max(n, 1), dtype=EXT_DTYPE)
    ext = np.asarray(ext)
    if n:
        e["read_start"][:n], e["read_end"][:n], e["ref_start"][:n] = ext[:, 0], ext[:, 1], ext[:, 2]
        e["score"][:n] = ext[:, 3]
    cig = np.ascontiguousarray(cig, dtype=np.uint32)
    cig_off = np.ascontiguousarray(cig_off, dtype=np.uint64)
    cig_n = np.ascontiguousarray(cig_n, dtype=np.uint32)
    score = np.zeros(max(n, 1), dtype=np.int64)
    bad = np.zeros(max(n, 1), dtype=np.int32)
    lib().espo_path_scores(n, batch.reads.ctypes.data, batch.read_off.ctypes.data, batch.read_len.ctypes.data,
                           batch.wins.ctypes.data, batch.win_off.ctypes.data, batch.win_len.ctypes.data,
                           e.ctypes.data, cig.ctypes.data, cig_off.ctypes.data, cig_n.ctypes.data,
                           ctypes.byref(_params(params)), score.ctypes.data, bad.ctypes.data, nthreads)
    return score[:n], bad[:n]


# ---- junction refinement, align_spliced, reextend_with (splice.rs, align.rs) ----

class CSpliceCfg(ctypes.Structure):
    """SpliceParams (splice.rs:21-53)."""
    _fields_ = [("min_intron", ctypes.c_uint32), ("max_intron", ctypes.c_uint32),
                ("gt_ag_bonus", ctypes.c_int32), ("gc_ag_bonus", ctypes.c_int32), ("at_ac_bonus", ctypes.c_int32),
                ("noncanonical_penalty", ctypes.c_int32), ("known_bonus", ctypes.c_int32),
                ("refine_radius", ctypes.c_uint32)]


SPLICE_CFG_DEFAULTS = (20, 1_000_000, 4, 2, 1, 8, 6, 12)  # SpliceParams::default, splice.rs:40-53


class CRefined(ctypes.Structure):
    _fields_ = [("found", ctypes.c_int32), ("start", ctypes.c_uint32), ("end", ctypes.c_uint32),
                ("minus", ctypes.c_uint32), ("signal", ctypes.c_int32), ("support", ctypes.c_uint32)]


class CAlignSpliceRec(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int32), ("refusal", ctypes.c_int32), ("n_segments", ctypes.c_uint32),
                ("genomic_start", ctypes.c_uint32), ("pseudo_len", ctypes.c_uint32),
                ("n_stitched", ctypes.c_uint32)]


class CReextRec(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int32), ("pos", ctypes.c_uint32), ("score", ctypes.c_int32),
                ("wstart", ctypes.c_uint32), ("wlen", ctypes.c_uint32), ("aligned", ctypes.c_uint32),
                ("new_aligned", ctypes.c_uint32)]


def _splice_lib():
    L = lib()
    if getattr(L, "_splice_bound", False):
        return L
    P, vp, sz = ctypes.POINTER, ctypes.c_void_p, ctypes.c_size_t
    u32p = P(ctypes.c_uint32)
    L.espo_jlib_build.argtypes = [u32p, sz]
    L.espo_jlib_build.restype = vp
    L.espo_jlib_free.argtypes = [vp]
    L.espo_jlib_free.restype = None
    L.espo_jlib_support.argtypes = [vp, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32]
    L.espo_jlib_support.restype = ctypes.c_uint32
    L.espo_refine_junction_read.argtypes = [P(CRef), P(ctypes.c_uint8), sz, ctypes.c_uint32, ctypes.c_uint32,
                                            ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int, vp,
                                            P(CSpliceCfg), P(CRefined)]
    L.espo_segment_chain.argtypes = [u32p, sz, ctypes.c_uint32, ctypes.c_uint32, u32p]
    L.espo_segment_chain.restype = sz
    L.espo_align_spliced.argtypes = [P(CRef), vp, P(ctypes.c_uint8), sz, ctypes.c_uint32, ctypes.c_int, u32p, sz,
                                     ctypes.c_uint32, P(CSpliceCfg), P(CParams), P(CExt), u32p, sz, P(sz), u32p,
                                     sz, P(CAlignSpliceRec), P(CRefined), u32p]
    L.espo_reextend_with.argtypes = [P(CRef), P(ctypes.c_uint8), sz, ctypes.c_uint32, ctypes.c_uint32, u32p, sz,
                                     P(CParams), ctypes.c_int, P(CReextRec), u32p, sz, P(sz)]
    L._splice_bound = True
    return L


def splice_cfg(min_intron=20, max_intron=1_000_000, gt_ag=4, gc_ag=2, at_ac=1, noncanonical=8, known=6,
               refine_radius=12) -> CSpliceCfg:
    return CSpliceCfg(min_intron, max_intron, gt_ag, gc_ag, at_ac, noncanonical, known, refine_radius)


class JLib:
    """JunctionLib::build (gtf.rs:84-114) of (contig, start, end, minus) quads."""

    def __init__(self, quads: Sequence[Tuple[int, int, int, int]]):
        L = _splice_lib()
        q = np.ascontiguousarray(np.array(quads, dtype=np.uint32).reshape(-1, 4) if len(quads) else
                                 np.zeros((0, 4), np.uint32))
        self.quads = q
        self.h = L.espo_jlib_build(q.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), q.shape[0])

    def support(self, contig, start, end, minus) -> int:
        return int(_splice_lib().espo_jlib_support(self.h, contig, start, end, int(minus)))

    def close(self):
        if self.h:
            _splice_lib().espo_jlib_free(self.h)
            self.h = None

    def __del__(self):
        self.close()


SIGNALS = ("GtAg", "GcAg", "AtAc", "NonCanonical")


def refine_junction_read(oref: OracleRef, read: bytes, contig: int, naive_start: int, naive_end: int,
                         q_a: int, q_b: int, minus: bool, jlib: Optional[JLib], cfg: Optional[CSpliceCfg] = None):
    """splice.rs:208 refine_junction_read -> (start, end, minus, signal, support) or None."""
    L = _splice_lib()
    r, m = _buf(read)
    o = CRefined()
    found = L.espo_refine_junction_read(ctypes.byref(oref.c), r, m, contig, naive_start, naive_end, q_a, q_b,
                                        int(minus), jlib.h if jlib else None, ctypes.byref(cfg or splice_cfg()),
                                        ctypes.byref(o))
    return (o.start, o.end, bool(o.minus), SIGNALS[o.signal], o.support) if found else None


def segment_chain(anchors: Sequence[Tuple[int, int]], k: int, min_intron: int = 20):
    """splice.rs:364 segment_chain over (qpos, rpos) anchors -> [(q_start, q_end, r_start, r_end)]."""
    L = _splice_lib()
    a = np.ascontiguousarray(np.array(anchors, dtype=np.uint32).reshape(-1))
    out = np.zeros(max(len(anchors), 1) * 4, np.uint32)
    ns = L.espo_segment_chain(a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), len(anchors), k, min_intron,
                              out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)))
    return [tuple(int(x) for x in out[4 * s:4 * s + 4]) for s in range(ns)]


def align_spliced(oref: OracleRef, jlib: Optional[JLib], read: bytes, contig: int, minus: bool,
                  anchors: Sequence[Tuple[int, int]], k: int, cfg: Optional[CSpliceCfg] = None,
                  params: Optional[ExtendParams] = None) -> dict:
    """splice.rs:625 align_spliced (after prepare_spliced, :545)."""
    L = _splice_lib()
    r, m = _buf(read)
    na = len(anchors)
    a = np.ascontiguousarray(np.array(anchors, dtype=np.uint32).reshape(-1)) if na else np.zeros(2, np.uint32)
    cap = cigar_capacity(m)
    cig = (ctypes.c_uint32 * cap)()
    scap = cap + 2 * max(na, 1)
    st = (ctypes.c_uint32 * scap)()
    js = (CRefined * max(na, 1))()
    bounds = (ctypes.c_uint32 * (3 * max(na, 1)))()
    e, rec, nops = CExt(), CAlignSpliceRec(), ctypes.c_size_t()
    u32 = ctypes.POINTER(ctypes.c_uint32)
    L.espo_align_spliced(ctypes.byref(oref.c), jlib.h if jlib else None, r, m, contig, int(minus), a.ctypes.data_as(u32),
                         na, k, ctypes.byref(cfg or splice_cfg()), ctypes.byref(_params(params)), ctypes.byref(e),
                         cig, cap, ctypes.byref(nops), st, scap, ctypes.byref(rec), js, bounds)
    ns = rec.n_segments if rec.status != 0 else 0
    return {"status": rec.status, "refusal": rec.refusal, "n_segments": rec.n_segments,
            "genomic_start": rec.genomic_start, "pseudo_len": rec.pseudo_len,
            "ext": Extension.from_c(e, cig[: nops.value]) if rec.status != 0 else None,
            "stitched": cigar_from_words(st[: rec.n_stitched]) if rec.status == 1 else [],
            "junctions": [(js[j].start, js[j].end, bool(js[j].minus), SIGNALS[js[j].signal], js[j].support)
                          for j in range(max(ns - 1, 0))],
            "bounds": [tuple(bounds[3 * s:3 * s + 3]) for s in range(ns)]}


def reextend_with(oref: OracleRef, query: bytes, contig: int, pos: int, cigar_words: Sequence[int],
                  params: Optional[ExtendParams], edit_dir: int) -> dict:
    """align.rs:1170 Aligner::reextend_with."""
    L = _splice_lib()
    q, m = _buf(query)
    cw = (ctypes.c_uint32 * max(len(cigar_words), 1))(*cigar_words)
    cap = cigar_capacity(m)
    out = (ctypes.c_uint32 * cap)()
    rec, n = CReextRec(), ctypes.c_size_t()
    L.espo_reextend_with(ctypes.byref(oref.c), q, m, contig, pos, cw, len(cigar_words), ctypes.byref(_params(params)),
                         int(edit_dir), ctypes.byref(rec), out, cap, ctypes.byref(n))
    return {"status": rec.status, "pos": rec.pos, "score": rec.score, "wstart": rec.wstart, "wlen": rec.wlen,
            "aligned": rec.aligned, "new_aligned": rec.new_aligned, "cigar": list(out[: n.value])}


class CIntronCfg(ctypes.Structure):
    """IntronParams (intron_chain.rs:20-42)."""
    _fields_ = [("min_frag_cov", ctypes.c_uint32), ("refine_window", ctypes.c_uint32),
                ("refine_pattern_len", ctypes.c_uint32), ("intron_max", ctypes.c_uint32)]


INTRON_CFG_DEFAULTS = (15, 30, 50, 105_000)  # IntronParams::default, intron_chain.rs:32-42


class CIcRec(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int32), ("reason", ctypes.c_int32), ("genomic_start", ctypes.c_uint32),
                ("score", ctypes.c_int32), ("left_start", ctypes.c_uint32), ("jstart", ctypes.c_uint32),
                ("jend", ctypes.c_uint32), ("right_end", ctypes.c_uint32), ("signal", ctypes.c_int32),
                ("support", ctypes.c_uint32), ("minus", ctypes.c_uint32), ("n_stitched", ctypes.c_uint32),
                ("donor_end", ctypes.c_int64), ("acceptor_start", ctypes.c_int64), ("vd", ctypes.c_int32),
                ("va", ctypes.c_int32)]


def intron_chain(oref: OracleRef, jlib: Optional[JLib], read: bytes, anchors: Sequence[Tuple[int, int, int, int]],
                 k: int = 15, cfg: Tuple[int, int, int, int] = INTRON_CFG_DEFAULTS,
                 params: Optional[ExtendParams] = None) -> dict:
    """intron_chain.rs:392 try_intron_chain_placement; anchors (qpos, rpos, contig, minus)."""
    L = _splice_lib()
    if not getattr(L, "_ic_bound", False):
        P, vp, sz = ctypes.POINTER, ctypes.c_void_p, ctypes.c_size_t
        u32p = P(ctypes.c_uint32)
        L.espo_intron_chain.argtypes = [P(CRef), vp, P(ctypes.c_uint8), sz, u32p, sz, ctypes.c_uint32, P(CIntronCfg),
                                        P(CParams), P(CExt), u32p, sz, P(sz), u32p, sz, P(CIcRec)]
        L._ic_bound = True
    r, m = _buf(read)
    a = np.ascontiguousarray(np.array(anchors, dtype=np.uint32).reshape(-1)) if len(anchors) else \
        np.zeros(4, np.uint32)
    cap = cigar_capacity(m)
    cig = (ctypes.c_uint32 * cap)()
    st = (ctypes.c_uint32 * (cap + 4))()
    e, rec, n = CExt(), CIcRec(), ctypes.c_size_t()
    c = CIntronCfg(*cfg)
    L.espo_intron_chain(ctypes.byref(oref.c), jlib.h if jlib else None, r, m,
                        a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), len(anchors), k, ctypes.byref(c),
                        ctypes.byref(_params(params)), ctypes.byref(e), cig, cap, ctypes.byref(n), st, cap + 4,
                        ctypes.byref(rec))
    out = {f: getattr(rec, f) for f, _ in CIcRec._fields_}
    out["signal"] = SIGNALS[rec.signal]
    out["ext"] = Extension.from_c(e, cig[: n.value]) if rec.left_start or rec.status else None
    out["stitched"] = cigar_from_words(st[: rec.n_stitched])
    return out


def set_route(route: int) -> None:
    """0: pure inputs run the scalar run_band<i16> restatement; 1: the AVX2
    restatements where run_dispatch takes them (extend.rs:293-344)."""
    L = lib()
    L.espo_set_route.argtypes = [ctypes.c_int]
    L.espo_set_route.restype = None
    L.espo_set_route(int(route))


def avx2_available() -> bool:
    L = lib()
    L.espo_avx2_available.restype = ctypes.c_int
    return bool(L.espo_avx2_available())
