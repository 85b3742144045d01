"""Batched alignment over device-resident pairs (the batched form).

The reference aligns reads one by one on rayon workers
(crates/map/src/pipeline.rs:393-404); here a batch of read-window pairs is
uploaded once and aligned by one persistent kernel launch per call.

PairBatch      host-side concatenated pairs (numpy)
DeviceBatch    the same pairs resident in HBM plus a result arena
synth_pairs    seeded synthetic workloads (include/esp_synth.h)
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import List, Optional, Sequence

import numpy as np

from . import _native
from .extend import ExtendBuffer, default_buffer
from .records import (CExt, CTouch, Extension, ExtendParams, TouchRec, cigar_from_words)

MODE_BAND, MODE_FULL, MODE_HINT = 0, 1, 2


@dataclass
class PairBatch:
    reads: np.ndarray     # uint8, concatenated
    read_off: np.ndarray  # uint64
    read_len: np.ndarray  # uint32
    wins: np.ndarray
    win_off: np.ndarray
    win_len: np.ndarray
    offset: np.ndarray    # int64: band offset or hint offset
    half: np.ndarray      # int64: band half-width

    @property
    def npairs(self) -> int:
        return int(self.read_len.shape[0])

    @staticmethod
    def from_pairs(pairs: Sequence[tuple]) -> "PairBatch":
        """pairs: (read, window, offset, half) tuples of bytes/ints."""
        reads = [bytes(p[0]) for p in pairs]
        wins = [bytes(p[1]) for p in pairs]
        rl = np.array([len(r) for r in reads], dtype=np.uint32)
        wl = np.array([len(w) for w in wins], dtype=np.uint32)
        ro = np.zeros(len(pairs), dtype=np.uint64)
        wo = np.zeros(len(pairs), dtype=np.uint64)
        if len(pairs):
            ro[1:] = np.cumsum(rl, dtype=np.uint64)[:-1]
            wo[1:] = np.cumsum(wl, dtype=np.uint64)[:-1]
        rb = np.frombuffer(b"".join(reads) or b"\0", dtype=np.uint8).copy()
        wb = np.frombuffer(b"".join(wins) or b"\0", dtype=np.uint8).copy()
        off = np.array([int(p[2]) for p in pairs], dtype=np.int64)
        half = np.array([int(p[3]) if len(p) > 3 else 0 for p in pairs], dtype=np.int64)
        return PairBatch(rb, ro, rl, wb, wo, wl, off, half)

    def pair(self, k: int):
        ro, rl = int(self.read_off[k]), int(self.read_len[k])
        wo, wl = int(self.win_off[k]), int(self.win_len[k])
        return (self.reads[ro:ro + rl].tobytes(), self.wins[wo:wo + wl].tobytes())

    def subset(self, idx: Sequence[int]) -> "PairBatch":
        return PairBatch.from_pairs([(*self.pair(k), int(self.offset[k]), int(self.half[k]))
                                     for k in idx])

    def cigar_offsets(self) -> np.ndarray:
        cap = 2 * self.read_len.astype(np.uint64) + 5
        off = np.zeros(self.npairs + 1, dtype=np.uint64)
        off[1:] = np.cumsum(cap, dtype=np.uint64)
        return off

    def cells_band(self) -> int:
        """In-band cell census sum_i (hi_i - lo_i + 1) of run_dispatch(Some(..)) (extend.rs:664-668)."""
        total = 0
        for k in range(self.npairs):
            total += band_cells(int(self.read_len[k]), int(self.win_len[k]), int(self.offset[k]),
                                int(self.half[k]))
        return total


def band_cells(m: int, n: int, offset: int, half: int) -> int:
    if m == 0 or n == 0:
        return 0
    half = max(half, 0)
    i = np.arange(1, m + 1, dtype=np.int64)
    c = np.clip(offset + i, 1, n)
    lo = np.maximum(c - half, 1)
    hi = np.minimum(c + half, n)
    return int(np.sum(hi - lo + 1))


class _SynthSpec(ctypes.Structure):
    _fields_ = [("seed", ctypes.c_uint64), ("npairs", ctypes.c_uint32),
                ("len_min", ctypes.c_uint32), ("len_max", ctypes.c_uint32),
                ("nbands", ctypes.c_uint32), ("bands", ctypes.c_uint32 * 8),
                ("sub_rate", ctypes.c_double), ("ag_weight", ctypes.c_double),
                ("ag_fraction", ctypes.c_double), ("indel_rate", ctypes.c_double),
                ("indel_max", ctypes.c_uint32), ("edits_max", ctypes.c_uint32)]


def synth_pairs(seed: int, npairs: int, len_min: int, len_max: int, bands: Sequence[int],
                sub_rate: float = 0.0, ag_weight: float = 1.0, ag_fraction: float = 0.0,
                indel_rate: float = 0.0, indel_max: int = 1, edits_max: int = 0) -> PairBatch:
    L = _native.lib()
    spec = _SynthSpec()
    spec.seed, spec.npairs, spec.len_min, spec.len_max = seed, npairs, len_min, len_max
    spec.nbands = len(bands)
    for k, b in enumerate(bands):
        spec.bands[k] = b
    spec.sub_rate, spec.ag_weight, spec.ag_fraction = sub_rate, ag_weight, ag_fraction
    spec.indel_rate, spec.indel_max, spec.edits_max = indel_rate, indel_max, edits_max
    rb, wb = ctypes.c_uint64(), ctypes.c_uint64()
    L.esp_synth_sizes(ctypes.byref(spec), ctypes.byref(rb), ctypes.byref(wb))
    reads = np.zeros(max(rb.value, 1), dtype=np.uint8)
    wins = np.zeros(max(wb.value, 1), dtype=np.uint8)
    ro = np.zeros(npairs, dtype=np.uint64)
    wo = np.zeros(npairs, dtype=np.uint64)
    rl = np.zeros(npairs, dtype=np.uint32)
    wl = np.zeros(npairs, dtype=np.uint32)
    off = np.zeros(npairs, dtype=np.int64)
    half = np.zeros(npairs, dtype=np.int64)
    vp = ctypes.c_void_p
    L.esp_synth_pairs(ctypes.byref(spec), vp(reads.ctypes.data), vp(ro.ctypes.data),
                      vp(rl.ctypes.data), vp(wins.ctypes.data), vp(wo.ctypes.data),
                      vp(wl.ctypes.data), vp(off.ctypes.data), vp(half.ctypes.data))
    return PairBatch(reads, ro, rl, wins, wo, wl, off, half)


# Workloads (seeds as stated for each configuration).
def workload_c1(npairs: int = 10_000) -> PairBatch:
    """164-nt windows (100-nt locus + 32-nt flanks), 100-nt reads, 1% subs
    (half A->G), 0.1% indels of 1-3, band offset = half = 32."""
    return synth_pairs(0xE5A00001, npairs, 100, 100, [32], sub_rate=0.01, ag_fraction=0.5,
                       indel_rate=0.001, indel_max=3)


def workload_c2(npairs: int = 1_000_000) -> PairBatch:
    """150-nt reads in 214-nt windows (32 each side), 0-10 random edits."""
    return synth_pairs(0xE5A00002, npairs, 150, 150, [32], edits_max=10)


def workload_c3(npairs: int = 2_000_000) -> PairBatch:
    """Reads U{50..250}, band U{16,32,48,64}, window = read + 2*band, 1% subs
    with A->G weighted 3x, 0.2% indels of 1-5."""
    return synth_pairs(0xE5A00003, npairs, 50, 250, [16, 32, 48, 64], sub_rate=0.01,
                       ag_weight=3.0, indel_rate=0.002, indel_max=5)


# Stripe half-width and head pad (rows) of the pruned full-matrix passes that
# the tuners try, the default first (esp_autotune_prune tries the same list).
PRUNE_SHAPES = ((32, 16), (24, 16), (48, 16), (24, 12), (32, 12), (24, 8), (32, 8), (32, 24))


def tune_prune_shape(buf: ExtendBuffer, launch, timings, reps: int = 2) -> dict:
    """esp_autotune_prune for a whole pipeline (mate rescue, spliced DP, ...):
    runs launch() (which must wait for the device) 1 + reps times at each shape
    of PRUNE_SHAPES, sums the full-matrix stages (escalation sort, pruned
    passes, one-wave and unpruned reruns) of timings() (the last call's
    [(kind, half, ms)]), and keeps the fastest shape in the context when it
    beats the default by more than 2%.  Results do not depend on the shape.
    Returns {"chosen": (half, pad), "ms_default", "ms_best", "candidates"}."""
    cands = []
    for half, pad in PRUNE_SHAPES:
        buf.set_option(_native.ESP_OPT_FULL_PRUNE_SHAPE, half | pad << 8)
        total = 0.0
        for r in range(reps + 1):
            launch()
            if r:
                total += sum(ms for kind, _, ms in timings() if kind in (7, 11, 12, 13))
        cands.append((half, pad, round(total / reps, 3)))
    best = min(range(len(cands)), key=lambda k: cands[k][2])
    if not cands[0][2] or cands[best][2] > 0.98 * cands[0][2]:
        best = 0
    half, pad, _ = cands[best]
    buf.set_option(_native.ESP_OPT_FULL_PRUNE_SHAPE, half | pad << 8)
    return {"chosen": (half, pad), "ms_default": cands[0][2], "ms_best": cands[best][2], "candidates": cands}


class DeviceArray:
    """Device allocation owned through the library (esp_device_alloc)."""

    def __init__(self, buf: ExtendBuffer, nbytes: int):
        self.buf = buf
        self.nbytes = int(nbytes)
        p = ctypes.c_void_p()
        _native.check(_native.lib().esp_device_alloc(buf.handle, max(self.nbytes, 1),
                                                     ctypes.byref(p)), buf.handle)
        self.ptr = p.value

    @staticmethod
    def from_numpy(buf: ExtendBuffer, a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = DeviceArray(buf, a.nbytes)
        if a.nbytes:
            _native.check(_native.lib().esp_copy_to_device(buf.handle, ctypes.c_void_p(d.ptr),
                                                           ctypes.c_void_p(a.ctypes.data),
                                                           a.nbytes), buf.handle)
        return d

    def to_numpy(self, dtype, count: int) -> np.ndarray:
        out = np.zeros(count, dtype=dtype)
        if out.nbytes:
            _native.check(_native.lib().esp_copy_to_host(self.buf.handle,
                                                         ctypes.c_void_p(out.ctypes.data),
                                                         ctypes.c_void_p(self.ptr), out.nbytes),
                          self.buf.handle)
        return out

    def free(self):
        if self.ptr:
            _native.lib().esp_device_free(self.buf.handle, ctypes.c_void_p(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class ContextTimings:
    """Device times of the last batch call on self.buf's context."""

    def sync(self) -> None:
        _native.check(_native.lib().esp_sync(self.buf.handle), self.buf.handle)

    def last_kernel_ms(self) -> float:
        ms = ctypes.c_double()
        _native.check(_native.lib().esp_last_kernel_ms(self.buf.handle, ctypes.byref(ms)),
                      self.buf.handle)
        return ms.value

    def last_plan_stats(self) -> dict:
        """Routing of the last batch: pairs per band half, strip pairs, escalations,
        and of those the pairs the pruned full-matrix passes finished, the pairs
        the unpruned kernel ran and the pairs the pruned pass could not prove
        (rerun one wavefront per pair); with
        the census option on, the full-matrix passes' computed cells, issued
        lane-cells and the cells of the full matrices they stand for; and the
        pruned passes' pairs with small heads (at most 3 row groups) and the rest."""
        st = _native.CPlanStats()
        _native.check(_native.lib().esp_last_plan_stats(self.buf.handle, ctypes.byref(st)),
                      self.buf.handle)
        return {"band_pairs": dict(zip((16, 30, 32, 48, 64), list(st.band_pairs))),
                "strip_pairs": st.strip_pairs, "escalated": st.escalated,
                "full_pruned": st.full_pruned, "full_unpruned": st.full_unpruned,
                "full_second": st.full_second,
                "full_cells_computed": st.full_cells_computed,
                "full_lane_cells_issued": st.full_lane_cells_issued,
                "full_cells_matrix": st.full_cells_matrix,
                "full_small_head": st.full_small_head, "full_large_head": st.full_large_head}

    def last_timings(self):
        """[(kind, half, ms)] of the last batch call: 0 planner, 1 band, 2 strip,
        3 Myers, 5 window preparation, 6 selection / stitching, 7 full-matrix
        rerun of escalated extend_hint attempts (the first pruned pass when
        pruning is on), 8 placement walks, 9 fast lane, 10 a whole graph launch,
        11 the unpruned pass over what the pruned passes could not prove, 12 the
        sort of the escalation list, 13 the one-wavefront-per-pair rerun of what
        the pruned pass could not prove."""
        arr = (_native.CTiming * 16)()
        n = ctypes.c_size_t()
        _native.check(_native.lib().esp_last_timings(self.buf.handle, arr, 16, ctypes.byref(n)),
                      self.buf.handle)
        return [(arr[k].kind, arr[k].half, arr[k].ms) for k in range(min(n.value, 16))]


class DeviceBatch(ContextTimings):
    """A PairBatch resident in HBM plus its result arena."""

    def __init__(self, host: PairBatch, buf: Optional[ExtendBuffer] = None):
        self.buf = buf or default_buffer()
        self.host = host
        b = self.buf
        self.d = {k: DeviceArray.from_numpy(b, getattr(host, k))
                  for k in ("reads", "read_off", "read_len", "wins", "win_off", "win_len",
                            "offset", "half")}
        n = host.npairs
        self.cig_off_host = host.cigar_offsets()
        self.cig_off = DeviceArray.from_numpy(b, self.cig_off_host)
        self.ext = DeviceArray(b, 16 * max(n, 1))
        self.touch = DeviceArray(b, 24 * max(n, 1))
        self.cigar = DeviceArray(b, 4 * max(int(self.cig_off_host[-1]), 1))
        self.cig_n = DeviceArray(b, 4 * max(n, 1))
        self.aux = DeviceArray(b, 4 * max(n, 1))
        self.dist = DeviceArray(b, 4 * max(n, 1))
        self.end = DeviceArray(b, 8 * max(n, 1))
        self.start = DeviceArray(b, 8 * max(n, 1))
        self._cb = _native.CBatch(n, *(self.d[k].ptr for k in ("reads", "read_off", "read_len",
                                                                 "wins", "win_off", "win_len",
                                                                 "offset", "half")))
        self._meta = _native.CHostMeta(host.read_len.ctypes.data, host.win_len.ctypes.data,
                                       host.offset.ctypes.data, host.half.ctypes.data)
        self._out = _native.CBatchOut(self.ext.ptr, self.touch.ptr, self.cigar.ptr,
                                      self.cig_off.ptr, self.cig_n.ptr, self.aux.ptr)

    @property
    def npairs(self) -> int:
        return self.host.npairs

    # ---- launches (asynchronous on the context stream) ----
    def launch_gotoh(self, params: Optional[ExtendParams] = None, mode: int = MODE_BAND) -> None:
        p = (params or ExtendParams()).to_c()
        _native.check(_native.lib().esp_gotoh_batch(self.buf.handle, ctypes.byref(self._cb),
                                                    ctypes.byref(self._meta), ctypes.byref(p),
                                                    mode, ctypes.byref(self._out)), self.buf.handle)

    def autotune_band(self, params: Optional[ExtendParams] = None, reps: int = 2) -> dict:
        """Tunes the band kernels' grids on this batch (esp_autotune_band); the
        context keeps the choice.  Returns {half: (chosen blocks per CU,
        occupancy limit, ms at the limit, ms at the choice)}."""
        p = (params or ExtendParams()).to_c()
        r = _native.CTuneReport()
        _native.check(_native.lib().esp_autotune_band(self.buf.handle, ctypes.byref(self._cb),
                                                      ctypes.byref(self._meta), ctypes.byref(p),
                                                      ctypes.byref(self._out), reps, ctypes.byref(r)),
                      self.buf.handle)
        return {int(r.half[k]): (int(r.blocks_per_cu[k]), int(r.occupancy[k]), round(float(r.ms_default[k]), 3),
                                 round(float(r.ms_best[k]), 3)) for k in range(5)}

    def autotune_prune(self, params: Optional[ExtendParams] = None, reps: int = 2) -> dict:
        """Tunes the pruned full-matrix passes' stripe half-width and head pad on
        this batch in extend_hint mode (esp_autotune_prune); the context keeps
        the choice.  Returns {"chosen": (half, pad), "ms_default", "ms_best",
        "candidates": [(half, pad, ms), ...]}."""
        p = (params or ExtendParams()).to_c()
        r = _native.CPruneTuneReport()
        _native.check(_native.lib().esp_autotune_prune(self.buf.handle, ctypes.byref(self._cb),
                                                       ctypes.byref(self._meta), ctypes.byref(p),
                                                       ctypes.byref(self._out), reps, ctypes.byref(r)),
                      self.buf.handle)
        return {"chosen": (int(r.chosen_half), int(r.chosen_pad)), "ms_default": round(float(r.ms_default), 3),
                "ms_best": round(float(r.ms_best), 3),
                "candidates": [(int(r.half[k]), int(r.pad[k]), round(float(r.ms[k]), 3))
                               for k in range(int(r.candidates))]}

    def capture_gotoh(self, params: Optional[ExtendParams] = None, mode: int = MODE_BAND) -> "GotohGraph":
        """launch_gotoh recorded once as a HIP graph; GotohGraph.launch() replays it."""
        return GotohGraph(self, params, mode)

    def launch_myers(self, with_start: bool = True, max_read_len: int = 0) -> None:
        """esp_myers_batch; max_read_len > 0 states a bound on the read lengths
        (the call then reads nothing back; longer reads are still exact)."""
        self._meta.max_read_len = int(max_read_len)
        try:
            _native.check(_native.lib().esp_myers_batch(
                self.buf.handle, ctypes.byref(self._cb), ctypes.byref(self._meta),
                ctypes.c_void_p(self.dist.ptr), ctypes.c_void_p(self.end.ptr),
                ctypes.c_void_p(self.start.ptr) if with_start else None), self.buf.handle)
        finally:
            self._meta.max_read_len = 0

    def launch_engine_myers(self, mode: int = 0) -> None:
        """The transcriptome engine's EA-Myers (engine/src/myers.rs) per pair:
        mode 0 infix, 1 global; distances land in myers_results()[0]."""
        _native.check(_native.lib().esp_engine_myers_batch(
            self.buf.handle, ctypes.byref(self._cb), ctypes.byref(self._meta), mode,
            ctypes.c_void_p(self.dist.ptr)), self.buf.handle)

    # ---- results ----
    def gotoh_results(self):
        """(ext[n,4] int64, touch[n,4] int64, cigar words, cig_off, cig_n, aux)."""
        n = self.npairs
        e = self.ext.to_numpy(np.uint32, 4 * n).reshape(n, 4)
        ext = e.astype(np.int64)
        ext[:, 3] = e[:, 3].view(np.int32)
        traw = self.touch.to_numpy(np.uint8, 24 * n).reshape(n, 24)
        touch = np.zeros((n, 4), dtype=np.int64)
        touch[:, 0] = traw[:, 0:4].copy().view(np.int32)[:, 0]
        touch[:, 1] = traw[:, 8:16].copy().view(np.int64)[:, 0]
        touch[:, 2] = traw[:, 16:20].copy().view(np.int32)[:, 0]
        touch[:, 3] = traw[:, 20:24].copy().view(np.int32)[:, 0]
        cig = self.cigar.to_numpy(np.uint32, int(self.cig_off_host[-1]))
        cig_n = self.cig_n.to_numpy(np.uint32, n)
        aux = self.aux.to_numpy(np.int32, n)
        return ext, touch, cig, self.cig_off_host, cig_n, aux

    def myers_results(self):
        n = self.npairs
        return (self.dist.to_numpy(np.int32, n), self.end.to_numpy(np.uint64, n),
                self.start.to_numpy(np.uint64, n))

    def extensions(self) -> List[Extension]:
        ext, _, cig, off, cig_n, _ = self.gotoh_results()
        out = []
        for k in range(self.npairs):
            words = cig[int(off[k]):int(off[k]) + int(cig_n[k])]
            out.append(Extension(int(ext[k, 0]), int(ext[k, 1]), int(ext[k, 2]),
                                 cigar_from_words(words), int(ext[k, 3])))
        return out

    def free(self):
        for a in list(self.d.values()) + [self.cig_off, self.ext, self.touch, self.cigar,
                                          self.cig_n, self.aux, self.dist, self.end, self.start]:
            a.free()


class BatchGraph:
    """A batched call captured as a HIP graph (esp_*_graph_create): launch()
    redoes the call on the device arrays' current contents with one graph
    launch; check() waits and raises when a launch met pairs past the bounds
    the graph was recorded with."""

    def __init__(self, buf, handle):
        self.buf = buf
        self.handle = handle

    def launch(self) -> None:
        _native.check(_native.lib().esp_graph_launch(self.buf.handle, ctypes.c_void_p(self.handle)),
                      self.buf.handle)

    def check(self) -> None:
        _native.check(_native.lib().esp_graph_check(self.buf.handle, ctypes.c_void_p(self.handle)),
                      self.buf.handle)

    def destroy(self) -> None:
        if self.handle:
            _native.lib().esp_graph_destroy(self.buf.handle, ctypes.c_void_p(self.handle))
            self.handle = None


class GotohGraph(BatchGraph):
    """One batched Gotoh call of a DeviceBatch (any mode) captured as a HIP
    graph (esp_gotoh_graph_create); recording runs the call once."""

    def __init__(self, db: DeviceBatch, params: Optional[ExtendParams], mode: int):
        self.db = db
        self._p = (params or ExtendParams()).to_c()
        h = ctypes.c_void_p()
        _native.check(_native.lib().esp_gotoh_graph_create(
            db.buf.handle, ctypes.byref(db._cb), ctypes.byref(db._meta), ctypes.byref(self._p), mode,
            ctypes.byref(db._out), ctypes.byref(h)), db.buf.handle)
        super().__init__(db.buf, h.value)
