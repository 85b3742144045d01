"""Host mirror of crates/map/src/myers_ea.rs over the HIP library.

The device kernel carries the pattern in 32-bit words (ceil(m/32) <= 8) with
a full carry chain between words, so one code path serves both the
reference's single-block (m <= 128) and long:: two-block (m <= 256)
functions; results depend only on the DP.  Guards are kept: m == 0 -> (0, 0),
m > 256 -> (m, 0) (myers_ea.rs:123-125, :254-256).
"""
from __future__ import annotations

import ctypes
from typing import Optional, Tuple

from . import _native
from .extend import ExtendBuffer, default_buffer


def _b(x: bytes):
    x = bytes(x)
    return (ctypes.c_uint8 * max(len(x), 1)).from_buffer_copy(x or b"\0"), len(x)


def infix_best_end(read: bytes, text: bytes, buf: Optional[ExtendBuffer] = None) -> Tuple[int, int]:
    """myers_ea.rs:141 / :271: (distance, earliest best end)."""
    buf = buf or default_buffer()
    r, m = _b(read)
    t, n = _b(text)
    d, e = ctypes.c_int32(), ctypes.c_size_t()
    _native.check(_native.lib().esp_myers_best_end(buf.handle, r, m, t, n, ctypes.byref(d),
                                                   ctypes.byref(e)), buf.handle)
    return d.value, e.value


def infix(read: bytes, text: bytes, buf: Optional[ExtendBuffer] = None) -> int:
    """myers_ea.rs:118 / :255"""
    buf = buf or default_buffer()
    r, m = _b(read)
    t, n = _b(text)
    d = ctypes.c_int32()
    _native.check(_native.lib().esp_myers_infix(buf.handle, r, m, t, n, ctypes.byref(d)), buf.handle)
    return d.value


def infix_best_start(read: bytes, text: bytes, buf: Optional[ExtendBuffer] = None) -> Tuple[int, int]:
    """myers_ea.rs:160 / :288: (distance, rightmost optimal start)."""
    buf = buf or default_buffer()
    r, m = _b(read)
    t, n = _b(text)
    d, s = ctypes.c_int32(), ctypes.c_size_t()
    _native.check(_native.lib().esp_myers_best_start(buf.handle, r, m, t, n, ctypes.byref(d),
                                                     ctypes.byref(s)), buf.handle)
    return d.value, s.value


class long:  # noqa: N801 -- mirrors the reference module name myers_ea::long
    infix = staticmethod(infix)
    infix_best_end = staticmethod(infix_best_end)
    infix_best_start = staticmethod(infix_best_start)
