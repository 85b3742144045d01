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
        L.espo_band_batch.argtypes = [sz, u8p, u64p, u32p, u8p, u64p, u32p, i64p, i64p,
                                      P(CParams), ctypes.c_int, P(CExt), P(CTouch), u32p, u64p,
                                      u32p, i32p, ctypes.c_int]
        L.espo_myers_batch.argtypes = [sz, u8p, u64p, u32p, u8p, u64p, u32p, i32p, u64p, u64p,
                                       ctypes.c_int]
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
