"""Serving loop with continuous batching (esp_server_*), over the resident reference.

Requests are extension requests -- per pair extend_hint(read, contig[ws, we),
hint) as the mapper's extension callers issue it (align.rs:906, :1907) -- or
mate-rescue requests -- per pair mate_window_placement(read, contig,
mate_pos) (align.rs:851).  `Server.submit_*` copies the inputs and returns a
`Ticket` that owns the output buffers; `Ticket.wait()` blocks until the
request's batch completed.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import List, Optional, Sequence

import numpy as np

from . import _native
from .records import ExtendParams, Extension, cigar_from_words
from .reference import DeviceReference
from .rescue import MATE_REC_DTYPE, MatePlacement

REQ_EXTEND, REQ_RESCUE = 0, 1


def _u32(a):
    return np.ascontiguousarray(a, dtype=np.uint32)


@dataclass
class RequestInfo:
    status: int
    done: bool
    submit_ns: int
    done_ns: int
    batch: int

    @property
    def latency_s(self) -> float:
        return (self.done_ns - self.submit_ns) / 1e9


class Ticket:
    """One submitted request and the host buffers its results land in."""

    def __init__(self, server: "Server", kind: int, read_len: np.ndarray, contig: np.ndarray):
        self.server = server
        self.kind = kind
        self.read_len = read_len
        self.contig = contig
        n = read_len.shape[0]
        caps = 2 * read_len.astype(np.uint64) + 5
        self.cig_off = np.zeros(n + 1, dtype=np.uint64)
        self.cig_off[1:] = np.cumsum(caps)
        self.cigar = np.zeros(max(int(self.cig_off[-1]), 1), dtype=np.uint32)
        self.cig_n = np.zeros(max(n, 1), dtype=np.uint32)
        self.ext = np.zeros((max(n, 1), 4), dtype=np.uint32) if kind == REQ_EXTEND else None
        self.mate = np.zeros(max(n, 1), dtype=MATE_REC_DTYPE) if kind == REQ_RESCUE else None
        self.id = 0
        self.info: Optional[RequestInfo] = None

    def out_struct(self) -> _native.CRequestOut:
        return _native.CRequestOut(self.ext.ctypes.data if self.ext is not None else None,
                                   self.mate.ctypes.data if self.mate is not None else None,
                                   self.cigar.ctypes.data, self.cig_n.ctypes.data)

    def wait(self, timeout_us: int = -1) -> Optional[RequestInfo]:
        """Blocks until done (or the timeout); returns the info once done."""
        if self.info is not None:
            return self.info
        ci = _native.CRequestInfo()
        L = _native.server_lib()
        _native.check(L.esp_server_wait(self.server.handle, self.id, timeout_us, ctypes.byref(ci)))
        if not ci.done:
            return None
        self.info = RequestInfo(ci.status, True, ci.submit_ns, ci.done_ns, ci.batch)
        return self.info

    @property
    def npairs(self) -> int:
        return int(self.read_len.shape[0])

    def raw(self):
        """(records, cigar words, cig_off, cig_n); records are ext[n,4] u32 or mate records."""
        n = self.npairs
        rec = self.ext[:n] if self.kind == REQ_EXTEND else self.mate[:n]
        return rec, self.cigar, self.cig_off, self.cig_n[:n]

    def extensions(self) -> List[Extension]:
        out = []
        for k in range(self.npairs):
            e = self.ext[k]
            words = self.cigar[int(self.cig_off[k]):int(self.cig_off[k]) + int(self.cig_n[k])]
            out.append(Extension(int(e[0]), int(e[1]), int(e[2]), cigar_from_words(words),
                                 int(np.int32(e[3]))))
        return out

    def placements(self) -> List[Optional[MatePlacement]]:
        out: List[Optional[MatePlacement]] = []
        for k in range(self.npairs):
            r = self.mate[k]
            if int(r["placed"]) != 1:
                out.append(None)
                continue
            words = self.cigar[int(self.cig_off[k]):int(self.cig_off[k]) + int(self.cig_n[k])]
            out.append(MatePlacement(int(self.contig[k]), int(r["pos"]), bool(r["strand"]),
                                     cigar_from_words(words), int(r["score"]), int(r["dist"])))
        return out


class Server:
    def __init__(self, dref: DeviceReference, params: Optional[ExtendParams] = None,
                 mem_budget: int = 32 << 30, workers: int = 3, max_batch_pairs: int = 65536,
                 max_pending: int = 1 << 20, max_read_len: int = 256):
        self.dref = dref
        cfg = _native.CServerConfig((params or ExtendParams()).to_c(), mem_budget, workers,
                                    max_batch_pairs, max_pending, max_read_len)
        h = ctypes.c_void_p()
        L = _native.server_lib()
        _native.check(L.esp_server_create(dref.buf.device, dref.handle, ctypes.byref(cfg),
                                          ctypes.byref(h)))
        self.handle = h

    def _submit(self, kind, reads: Sequence[bytes], contigs, win_start=None, win_end=None,
                hints=None, mate_pos=None) -> Ticket:
        rl = np.array([len(r) for r in reads], dtype=np.uint32)
        rb = np.frombuffer(b"".join(bytes(r) for r in reads) or b"\0", dtype=np.uint8)
        return self.submit_arrays(kind, rb, rl, contigs, win_start, win_end, hints, mate_pos)

    def submit_arrays(self, kind, reads: np.ndarray, read_len: np.ndarray, contigs,
                      win_start=None, win_end=None, hints=None, mate_pos=None) -> Ticket:
        """As submit_extend / submit_rescue with the reads already concatenated."""
        rb = np.ascontiguousarray(reads, dtype=np.uint8)
        if rb.shape[0] == 0:
            rb = np.zeros(1, np.uint8)
        rl = _u32(read_len)
        ctg = _u32(contigs)
        t = Ticket(self, kind, rl, ctg)
        keep = [rb, rl, ctg]
        req = _native.CRequest(kind, rl.shape[0], rb.ctypes.data, rl.ctypes.data, ctg.ctypes.data)
        if kind == REQ_EXTEND:
            ws, we = _u32(win_start), _u32(win_end)
            hi = np.ascontiguousarray(hints, dtype=np.int64)
            keep += [ws, we, hi]
            req.win_start, req.win_end, req.hint = ws.ctypes.data, we.ctypes.data, hi.ctypes.data
        else:
            mp = _u32(mate_pos)
            keep.append(mp)
            req.mate_pos = mp.ctypes.data
        out = t.out_struct()
        tid = ctypes.c_uint64()
        _native.check(_native.server_lib().esp_server_submit(self.handle, ctypes.byref(req),
                                                             ctypes.byref(out), ctypes.byref(tid)))
        t.id = tid.value
        return t

    def submit_extend(self, reads: Sequence[bytes], contigs, win_start, win_end, hints) -> Ticket:
        """extend_hint(read, contig[win_start, win_end), hint) per pair."""
        return self._submit(REQ_EXTEND, reads, contigs, win_start, win_end, hints)

    def submit_rescue(self, reads: Sequence[bytes], contigs, mate_pos) -> Ticket:
        """mate_window_placement(read, contig, mate_pos) per pair."""
        return self._submit(REQ_RESCUE, reads, contigs, mate_pos=mate_pos)

    def stats(self) -> dict:
        st = _native.CServerStats()
        _native.check(_native.server_lib().esp_server_stats_get(self.handle, ctypes.byref(st)))
        return {f: getattr(st, f) for f, _ in _native.CServerStats._fields_}

    def close(self) -> None:
        if getattr(self, "handle", None) is not None and self.handle.value:
            _native.server_lib().esp_server_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def now_ns() -> int:
    return int(_native.server_lib().esp_now_ns())
