"""The transcriptome engine's EA-Myers on the device (crates/engine/src/myers.rs)
and its caller's verification step (verify_dist, crates/engine/src/lib.rs:851-876).

Edit distance with identity and the editing pairs (read G over text A, read C
over text T) free, N matching nothing, case folded, indels 1: `infix` frees
the text overhang on both sides, `global_` forces both ends.  Patterns are
1..256 bases (the reference's single- and two-block forms run as one kernel);
other lengths raise, as the reference asserts.
"""
from __future__ import annotations

import ctypes
from typing import Optional, Sequence, Tuple

import numpy as np

from . import _native
from .batch import DeviceArray, DeviceBatch, PairBatch
from .extend import ExtendBuffer

MODE_INFIX, MODE_GLOBAL = 0, 1


def engine_myers_batch(pairs: Sequence[Tuple[bytes, bytes]], mode: int = MODE_INFIX,
                       buf: Optional[ExtendBuffer] = None) -> np.ndarray:
    """Distances of (read, text) pairs, one launch."""
    pb = PairBatch.from_pairs([(r, t, 0, 0) for r, t in pairs])
    db = DeviceBatch(pb, buf)
    try:
        db.launch_engine_myers(mode)
        db.sync()
        return db.myers_results()[0][: pb.npairs].copy()
    finally:
        db.free()


def infix(read: bytes, text: bytes, buf: Optional[ExtendBuffer] = None) -> int:
    """myers::infix (myers.rs:31-37)."""
    return int(engine_myers_batch([(read, text)], MODE_INFIX, buf)[0])


def global_(read: bytes, text: bytes, buf: Optional[ExtendBuffer] = None) -> int:
    """myers::global (myers.rs:41-43)."""
    return int(engine_myers_batch([(read, text)], MODE_GLOBAL, buf)[0])


class long:  # noqa: N801 -- the reference's module name
    """myers::long (myers.rs:46-61): the same distances for 128 < m <= 256."""

    infix = staticmethod(infix)
    global_ = staticmethod(global_)


def verify_dist_batch(reads: Sequence[bytes], tx_seqs: Sequence[bytes], tx_ids: Sequence[int],
                      anchors: Sequence[int], flank: int, buf: ExtendBuffer):
    """verify_dist (lib.rs:851-876) per candidate: read k (in the candidate's
    orientation) against transcript tx_ids[k]'s window around anchor k
    (ext.tx_lo - ext.read_lo).  Returns (dist, status); status 0 is None."""
    n = len(reads)
    if n == 0:
        return np.zeros(0, np.int32), np.zeros(0, np.int32)
    rl = np.array([len(r) for r in reads], dtype=np.uint32)
    ro = np.zeros(n, dtype=np.uint64)
    ro[1:] = np.cumsum(rl, dtype=np.uint64)[:-1]
    tl_all = np.array([len(s) for s in tx_seqs], dtype=np.uint64)
    to_all = np.zeros(len(tx_seqs), dtype=np.uint64)
    if len(tx_seqs):
        to_all[1:] = np.cumsum(tl_all, dtype=np.uint64)[:-1]
    ids = np.asarray(tx_ids, dtype=np.int64)
    host = [np.frombuffer(b"".join(bytes(r) for r in reads) or b"\0", dtype=np.uint8), ro, rl,
            np.frombuffer(b"".join(bytes(s) for s in tx_seqs) or b"\0", dtype=np.uint8), to_all[ids],
            tl_all[ids].astype(np.uint32), np.asarray(anchors, dtype=np.int64)]
    dev = [DeviceArray.from_numpy(buf, h) for h in host]
    d_dist = DeviceArray(buf, 4 * n)
    d_st = DeviceArray(buf, 4 * n)
    try:
        v = _native.CEngineVerify(n, *(d.ptr for d in dev))
        L = _native.lib()
        _native.check(L.esp_engine_verify_batch(buf.handle, ctypes.byref(v), ctypes.c_void_p(rl.ctypes.data),
                                                flank, ctypes.c_void_p(d_dist.ptr), ctypes.c_void_p(d_st.ptr)),
                      buf.handle)
        _native.check(L.esp_sync(buf.handle), buf.handle)
        return d_dist.to_numpy(np.int32, n), d_st.to_numpy(np.int32, n)
    finally:
        for d in dev + [d_dist, d_st]:
            d.free()
