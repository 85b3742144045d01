"""Host mirror of crates/map/src/extend.rs over the HIP library.

Same names, argument order and results as the reference:
  extend(read, ref_window, params, buf) -> Extension              extend.rs:347
  extend_hint(read, ref_window, params, buf, hint) -> Extension   extend.rs:392
  run_dispatch(read, ref_window, params, buf, band_spec)          extend.rs:293
      -> (Extension, TouchRec), band_spec = (offset, num, den, half) | None
ExtendBuffer (extend.rs:231) is a device context: one HIP stream plus
scratch; like the reference's buffer it has no semantic effect.  Every call
here is the plain form: one kernel launch per call, one wavefront sweeping
the pair (windows up to 4096 bases; longer ones, or ESP_OPT_PLAIN_KERNEL = 1,
take the lane-per-pair kernel).
"""
from __future__ import annotations

import ctypes
from typing import Optional, Tuple

from . import _native
from .records import (CExt, CParams, CTouch, DiagHint, Extension, ExtendParams, TouchRec,
                      cigar_capacity)


class ExtendBuffer:
    """Per-thread device context (esp_ctx)."""

    def __init__(self, device: int = 0):
        L = _native.lib()
        h = ctypes.c_void_p()
        rc = L.esp_create(int(device), ctypes.byref(h))
        if rc != _native.ESP_OK:
            raise _native.EspError(rc, "esp_create failed (no GPU visible?)")
        self._h = h
        self.device = device

    @property
    def handle(self):
        return self._h

    def set_option(self, option: int, value: int) -> None:
        """esp_ctx_set_option (ESP_OPT_FULL_KERNEL, ESP_OPT_TRACE_BUDGET,
        ESP_OPT_BAND_GRID, ESP_OPT_FULL_PRUNE, ESP_OPT_FULL_PRUNE_SHAPE,
        ESP_OPT_PLAIN_KERNEL, ESP_OPT_FULL_CENSUS)."""
        _native.check(_native.lib().esp_ctx_set_option(self._h, option, value), self._h)

    def last_kernel_ms(self) -> float:
        """esp_last_kernel_ms: device time of the last call's kernel (or batch)."""
        ms = ctypes.c_double()
        _native.check(_native.lib().esp_last_kernel_ms(self._h, ctypes.byref(ms)), self._h)
        return ms.value

    def close(self) -> None:
        if self._h:
            _native.lib().esp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


_default = None


def default_buffer() -> ExtendBuffer:
    global _default
    if _default is None:
        _default = ExtendBuffer(0)
    return _default


def _bytes(b: bytes):
    b = bytes(b)
    return (ctypes.c_uint8 * max(len(b), 1)).from_buffer_copy(b or b"\0"), len(b)


def _call_single(fn, buf, read, window, params, *extra, touch=False):
    buf = buf or default_buffer()
    r, m = _bytes(read)
    w, n = _bytes(window)
    cap = cigar_capacity(m)
    cig = (ctypes.c_uint32 * cap)()
    e, nops = CExt(), ctypes.c_size_t()
    p = (params or ExtendParams()).to_c()
    t = CTouch()
    args = [buf.handle, r, m, w, n, ctypes.byref(p), *extra, ctypes.byref(e)]
    if touch:
        args.append(ctypes.byref(t))
    args += [cig, cap, ctypes.byref(nops)]
    _native.check(fn(*args), buf.handle)
    ext = Extension.from_c(e, cig[: nops.value])
    return (ext, TouchRec.from_c(t)) if touch else ext


def extend(read: bytes, ref_window: bytes, params: Optional[ExtendParams] = None,
           buf: Optional[ExtendBuffer] = None) -> Extension:
    """Full-window extension (extend.rs:347)."""
    return _call_single(_native.lib().esp_extend, buf, read, ref_window, params)


def extend_hint(read: bytes, ref_window: bytes, params: Optional[ExtendParams] = None,
                buf: Optional[ExtendBuffer] = None, hint: DiagHint = DiagHint(0)) -> Extension:
    """Banded extension along a diagonal hint with full-matrix escalation (extend.rs:392)."""
    return _call_single(_native.lib().esp_extend_hint, buf, read, ref_window, params,
                        ctypes.c_int64(int(hint.offset)))


def run_dispatch(read: bytes, ref_window: bytes, params: Optional[ExtendParams] = None,
                 buf: Optional[ExtendBuffer] = None,
                 band_spec: Optional[Tuple[int, int, int, int]] = None) -> Tuple[Extension, TouchRec]:
    """run_dispatch (extend.rs:293); band_spec = (offset, num, den, half), num == den == 1."""
    if band_spec is not None:
        offset, num, den, half = band_spec
        if num != 1 or max(den, 1) != 1:
            raise ValueError("only the unit-slope band (num == den == 1) exists at every call site")
        extra = (ctypes.c_int(1), ctypes.c_int64(int(offset)), ctypes.c_int64(int(half)))
    else:
        extra = (ctypes.c_int(0), ctypes.c_int64(0), ctypes.c_int64(0))
    return _call_single(_native.lib().esp_band, buf, read, ref_window, params, *extra, touch=True)


def band(read: bytes, ref_window: bytes, params: Optional[ExtendParams] = None,
         band: Optional[Tuple[int, int]] = None,
         buf: Optional[ExtendBuffer] = None) -> Tuple[Extension, TouchRec]:
    """run_dispatch with band = (offset, half) or None (full matrix)."""
    spec = None if band is None else (band[0], 1, 1, band[1])
    return run_dispatch(read, ref_window, params, buf, spec)
