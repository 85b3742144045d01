"""MI355X-native alignment core of ESPERANTO's crates/map.

Drop-in for the reference's extend / extend_hint / run_dispatch, myers_ea
and stitch_cigar, backed by hand-written gfx950 HIP kernels in
esperanto_amd/lib/libesp_align.so (C ABI: include/esp_align.h).
"""
from .records import (INITIAL_BAND, CigarOp, DiagHint, EditDir, Extension, ExtendParams,
                      TouchRec, Del, Ins, Match, RefSkip, SoftClip, cigar_str)

__all__ = ["INITIAL_BAND", "CigarOp", "DiagHint", "EditDir", "Extension", "ExtendParams",
           "TouchRec", "Del", "Ins", "Match", "RefSkip", "SoftClip", "cigar_str"]
