"""Value types of the alignment core, mirroring the reference's crates/map API.

Reference definitions:
  EditDir        crates/map/src/extend.rs:43-93
  ExtendParams   crates/map/src/extend.rs:96-122 (default 2, -4, open 4, ext 2, Off)
  INITIAL_BAND   crates/map/src/extend.rs:126
  CigarOp        crates/map/src/extend.rs:129-140 (BAM codes M0 I1 D2 N3 S4,
                 crates/bamio/src/encode.rs:169-177)
  DiagHint       crates/map/src/extend.rs:191-199
  Extension      crates/map/src/extend.rs:202-214
  TouchRec       crates/map/src/extend.rs:220-229

CIGARs travel across the C ABI as BAM u32 words (``len << 4 | op``); the
Python ``CigarOp`` is a thin view over one such word.
"""
from __future__ import annotations

import ctypes
import enum
from dataclasses import dataclass, field
from typing import List, Sequence

INITIAL_BAND = 30

OP_M, OP_I, OP_D, OP_N, OP_S = 0, 1, 2, 3, 4
_OP_CHARS = "MIDNS"


class EditDir(enum.IntEnum):
    """Which A-to-I substitution scores 0 (extend.rs:43-50)."""

    Off = 0
    Ag = 1  # ref A / read G free (plus-strand transcript)
    Tc = 2  # ref T / read C free (minus-strand transcript)

    def is_on(self) -> bool:
        return self != EditDir.Off

    @staticmethod
    def from_plus(plus: bool) -> "EditDir":
        return EditDir.Ag if plus else EditDir.Tc


def exempt_dir(ref_b: int, read_b: int):
    """extend.rs:82-93: True for ref A/read G, False for ref T/read C, else None."""
    if ref_b == ord("A") and read_b == ord("G"):
        return True
    if ref_b == ord("T") and read_b == ord("C"):
        return False
    return None


@dataclass(frozen=True)
class CigarOp:
    """One CIGAR operation (extend.rs:129-140) carried as a BAM word."""

    word: int

    @staticmethod
    def make(op: int, n: int) -> "CigarOp":
        return CigarOp((int(n) << 4) | int(op))

    @property
    def op(self) -> int:
        return self.word & 0xF

    @property
    def len(self) -> int:
        return self.word >> 4

    def __repr__(self) -> str:  # e.g. "M23"
        return f"{_OP_CHARS[self.op]}{self.len}"


def Match(n: int) -> CigarOp:
    return CigarOp.make(OP_M, n)


def Ins(n: int) -> CigarOp:
    return CigarOp.make(OP_I, n)


def Del(n: int) -> CigarOp:
    return CigarOp.make(OP_D, n)


def RefSkip(n: int) -> CigarOp:
    return CigarOp.make(OP_N, n)


def SoftClip(n: int) -> CigarOp:
    return CigarOp.make(OP_S, n)


def cigar_from_words(words: Sequence[int]) -> List[CigarOp]:
    return [CigarOp(int(w)) for w in words]


def cigar_str(cigar: Sequence[CigarOp]) -> str:
    return "".join(f"{c.len}{_OP_CHARS[c.op]}" for c in cigar)


class CParams(ctypes.Structure):
    """C layout of ExtendParams shared by the C ABI (esp_params)."""

    _fields_ = [
        ("match_score", ctypes.c_int32),
        ("mismatch", ctypes.c_int32),
        ("gap_open", ctypes.c_int32),
        ("gap_ext", ctypes.c_int32),
        ("edit_dir", ctypes.c_int32),
    ]


class CExt(ctypes.Structure):
    _fields_ = [
        ("read_start", ctypes.c_uint32),
        ("read_end", ctypes.c_uint32),
        ("ref_start", ctypes.c_uint32),
        ("score", ctypes.c_int32),
    ]


class CTouch(ctypes.Structure):
    _fields_ = [
        ("touched", ctypes.c_int32),
        ("row", ctypes.c_int64),
        ("left", ctypes.c_int32),
        ("clamped", ctypes.c_int32),
    ]


@dataclass(frozen=True)
class ExtendParams:
    """Extension scoring (extend.rs:96-122); penalties are stored positive."""

    match_score: int = 2
    mismatch: int = -4
    gap_open: int = 4
    gap_ext: int = 2
    edit_dir: EditDir = EditDir.Off

    def to_c(self) -> CParams:
        return CParams(self.match_score, self.mismatch, self.gap_open, self.gap_ext,
                       int(self.edit_dir))


@dataclass(frozen=True)
class DiagHint:
    """Band centre j - i plus anchor support num/den (extend.rs:191-199)."""

    offset: int
    num: int = 1
    den: int = 1


@dataclass
class Extension:
    """Result of one extension (extend.rs:202-214)."""

    read_start: int
    read_end: int
    ref_start: int
    cigar: List[CigarOp] = field(default_factory=list)
    score: int = 0

    @staticmethod
    def from_c(e: CExt, words: Sequence[int]) -> "Extension":
        return Extension(int(e.read_start), int(e.read_end), int(e.ref_start),
                         cigar_from_words(words), int(e.score))


@dataclass
class TouchRec:
    """First strict band-edge touch (extend.rs:220-229)."""

    touched: bool = False
    row: int = 0
    left: bool = False
    clamped: bool = False

    @staticmethod
    def from_c(t: CTouch) -> "TouchRec":
        return TouchRec(bool(t.touched), int(t.row), bool(t.left), bool(t.clamped))


def cigar_capacity(m: int) -> int:
    """Upper bound on CIGAR ops for a read of m bases.

    A path has at most m read-consuming (M/I) steps; runs of D are separated
    by at least one M/I step, so M/I/D runs number at most 2m + 1, plus two
    neutral-flank M and two soft clips (the flank M's merge into runs).
    """
    return 2 * int(m) + 5
