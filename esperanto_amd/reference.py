"""Resident 2-bit reference (crates/map/src/fasta.rs) and the device window gather.

`Reference` holds contigs in the reference's own packed layout: ceil(len/4)
bytes per contig, base at pos in byte pos/4 at shift 6 - 2*(pos%4), N runs as
merged half-open intervals (fasta.rs:1-135, push_base at :291-310).
`DeviceReference` uploads it once (esp_ref_upload) and decodes windows on the
GPU (esp_gather_batch): Contig::decode_into, optionally reverse-complemented
as revcomp_into (align.rs:2525-2533) does.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field
from typing import List, Optional, Sequence, Tuple

import numpy as np

from . import _native
from .extend import ExtendBuffer, default_buffer

# Base::from_ascii (fasta.rs:30-39): A/a 0, C/c 1, G/g 2, T/t 3, anything else N.
_CODE = np.full(256, 255, dtype=np.uint8)
for _b, _c in ((b"A", 0), (b"C", 1), (b"G", 2), (b"T", 3)):
    _CODE[_b[0]] = _c
    _CODE[_b.lower()[0]] = _c


def pack_contig(seq: bytes) -> Tuple[np.ndarray, np.ndarray]:
    """push_base over a whole contig (fasta.rs:291-310) plus the tail padding of
    parse_fasta_bytes (:199-201): (packed u8[ceil(len/4)], intervals u32[k, 2])."""
    s = np.frombuffer(bytes(seq), dtype=np.uint8)
    L = s.shape[0]
    codes = _CODE[s]
    isn = codes == 255
    c2 = np.where(isn, 0, codes).astype(np.uint8)
    pad = (-L) % 4
    c2 = np.concatenate([c2, np.zeros(pad, dtype=np.uint8)]).reshape(-1, 4)
    packed = (c2[:, 0] << 6) | (c2[:, 1] << 4) | (c2[:, 2] << 2) | c2[:, 3]
    packed = packed.astype(np.uint8)
    if not isn.any():
        return packed, np.zeros((0, 2), dtype=np.uint32)
    d = np.diff(np.concatenate([[0], isn.astype(np.int8), [0]]))
    starts = np.nonzero(d == 1)[0]
    ends = np.nonzero(d == -1)[0]
    return packed, np.stack([starts, ends], axis=1).astype(np.uint32)


@dataclass
class Reference:
    """Contigs in file order (fasta.rs:159-165), flat for the C ABI."""

    names: List[str]
    lens: np.ndarray        # u32 per contig
    packed: np.ndarray      # u8, contigs concatenated
    packed_off: np.ndarray  # u64 per contig
    iv: np.ndarray          # u32 [total_intervals, 2]
    iv_off: np.ndarray      # u64 per contig + 1

    @staticmethod
    def from_sequences(seqs: Sequence[bytes], names: Optional[Sequence[str]] = None) -> "Reference":
        packs, ivs = zip(*[pack_contig(s) for s in seqs]) if seqs else ((), ())
        lens = np.array([len(s) for s in seqs], dtype=np.uint32)
        psz = np.array([p.shape[0] for p in packs], dtype=np.uint64)
        poff = np.zeros(len(seqs), dtype=np.uint64)
        if len(seqs):
            poff[1:] = np.cumsum(psz)[:-1]
        ivn = np.array([v.shape[0] for v in ivs], dtype=np.uint64)
        ivoff = np.zeros(len(seqs) + 1, dtype=np.uint64)
        ivoff[1:] = np.cumsum(ivn)
        packed = np.concatenate(packs) if packs else np.zeros(0, np.uint8)
        iv = np.concatenate(ivs) if ivs else np.zeros((0, 2), np.uint32)
        return Reference(list(names) if names else [f"c{k}" for k in range(len(seqs))], lens,
                         packed.astype(np.uint8), poff, iv.astype(np.uint32), ivoff)

    @staticmethod
    def synthetic(seed: int, contig_lens: Sequence[int], n_runs: int = 0,
                  n_run_max: int = 200) -> "Reference":
        """Uniform ACGT contigs from a seeded generator, with n_runs N runs of
        1..n_run_max bases planted per contig (in place, sequence length kept)."""
        rng = np.random.default_rng(seed)
        seqs = []
        for L in contig_lens:
            s = np.frombuffer(b"ACGT", dtype=np.uint8)[rng.integers(0, 4, int(L))].copy()
            for _ in range(n_runs):
                a = int(rng.integers(0, max(1, L)))
                b = min(int(L), a + int(rng.integers(1, n_run_max + 1)))
                s[a:b] = ord("N")
            seqs.append(s.tobytes())
        return Reference.from_sequences(seqs)

    @property
    def ncontig(self) -> int:
        return int(self.lens.shape[0])


class CGatherJob(ctypes.Structure):
    _fields_ = [("contig", ctypes.c_uint32), ("start", ctypes.c_uint32), ("end", ctypes.c_uint32),
                ("flags", ctypes.c_uint32), ("dst_off", ctypes.c_uint64)]


GATHER_JOB_DTYPE = np.dtype([("contig", "<u4"), ("start", "<u4"), ("end", "<u4"),
                             ("flags", "<u4"), ("dst_off", "<u8")])


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


class DeviceReference:
    """The reference resident in HBM (esp_ref_upload), released by close()."""

    def __init__(self, ref: Reference, buf: Optional[ExtendBuffer] = None):
        self.buf = buf or default_buffer()
        self.ref = ref
        L = _native.lib()
        h = ctypes.c_void_p()
        lens = np.ascontiguousarray(ref.lens, dtype=np.uint32)
        packed = np.ascontiguousarray(ref.packed, dtype=np.uint8)
        if packed.shape[0] == 0:
            packed = np.zeros(1, np.uint8)
        poff = np.ascontiguousarray(ref.packed_off, dtype=np.uint64)
        iv = np.ascontiguousarray(ref.iv.reshape(-1), dtype=np.uint32)
        if iv.shape[0] == 0:
            iv = np.zeros(2, np.uint32)
        ivoff = np.ascontiguousarray(ref.iv_off, dtype=np.uint64)
        _native.check(L.esp_ref_upload(self.buf.handle, ref.ncontig, _ptr(lens), _ptr(packed),
                                       _ptr(poff), _ptr(iv), _ptr(ivoff), ctypes.byref(h)),
                      self.buf.handle)
        self.handle = h

    def device_bytes(self) -> int:
        n = ctypes.c_size_t()
        _native.check(_native.lib().esp_ref_device_bytes(self.handle, ctypes.byref(n)))
        return int(n.value)

    def gather(self, jobs: Sequence[Tuple[int, int, int, bool]]) -> Tuple[List[bytes], np.ndarray]:
        """Decodes (contig, start, end, revcomp) windows on the device; returns the
        windows and the per-job status (0 ok, ESP_E_INVAL out of range)."""
        from .batch import DeviceArray
        n = len(jobs)
        if n == 0:
            return [], np.zeros(0, np.int32)
        J = np.zeros(n, dtype=GATHER_JOB_DTYPE)
        lens = np.zeros(n, dtype=np.uint64)
        for k, (c, s, e, rc) in enumerate(jobs):
            J[k] = (c, s, e, 1 if rc else 0, 0)
            lens[k] = max(0, int(e) - int(s))
        off = np.zeros(n, dtype=np.uint64)
        off[1:] = np.cumsum(lens)[:-1]
        J["dst_off"] = off
        total = int(lens.sum())
        dj = DeviceArray.from_numpy(self.buf, J)
        dd = DeviceArray(self.buf, max(total, 1))
        ds = DeviceArray(self.buf, 4 * n)
        L = _native.lib()
        _native.check(L.esp_gather_batch(self.buf.handle, self.handle, ctypes.c_void_p(dj.ptr), n,
                                         ctypes.c_void_p(dd.ptr), ctypes.c_void_p(ds.ptr)),
                      self.buf.handle)
        _native.check(L.esp_sync(self.buf.handle), self.buf.handle)
        data = dd.to_numpy(np.uint8, max(total, 1)).tobytes()
        status = ds.to_numpy(np.int32, n)
        for a in (dj, dd, ds):
            a.free()
        out = [data[int(off[k]):int(off[k] + lens[k])] if status[k] == 0 else b"" for k in range(n)]
        return out, status

    def close(self) -> None:
        if getattr(self, "handle", None) is not None and self.handle.value:
            _native.lib().esp_ref_free(self.buf.handle, self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
