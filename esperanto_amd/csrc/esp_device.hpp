// esp_device.hpp -- device-side building blocks shared by the alignment kernels.
//
// Semantics restated from the reference (crates/map/src/extend.rs):
//   base_code            :151-159   A/a 0, C/c 1, G/g 2, T/t 3, anything else 4
//   SubstLut             :145-186   match on equal codes < 4, EditDir-exempt pairs score 0
//   row windows          :664-668   c = clamp(offset + i, 1, n), [max(c-half,1), min(c+half,n)]
//   cell recurrence      :715-793   E/F strict '>' extension, H cascade diag > E > F > start
//   margins              :718-759   out-of-window reads H = 0, E = F = NEG
//   finish_extension     :864-1068  traceback, neutral flanks, soft clips
//
// Everything is int32.  Inside the scoring envelope the host validates
// (esp_align.h), the reference's i16 accumulator differs from its i32 one
// only by dropping diagonal candidates above 32767 on pure pairs; the strip
// kernel reproduces that (i16_wraps), and the band and full-matrix kernels
// only admit scores and read lengths where it cannot happen.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/esp_align.h"

namespace espd {

constexpr int32_t NEG = -1000000000;  // i32 sentinel of the reference (extend.rs:504)
constexpr uint32_t OP_M = 0, OP_I = 1, OP_D = 2, OP_N = 3, OP_S = 4;

// Trace nibble per cell: bits 0-1 hdir (0 diag, 1 from E, 2 from F, 3 start),
// bit 2 E-from-extension, bit 3 F-from-extension.
constexpr uint32_t HD_DIAG = 0, HD_E = 1, HD_F = 2, HD_START = 3;

struct GParams {
    int32_t match;     // LUT diagonal (codes equal, < 4)
    int32_t mismatch;  // every other pair
    int32_t oe;        // open_ext = -(gap_open + gap_ext)
    int32_t ext;       // -gap_ext
    int32_t edit;      // 0 Off, 1 Ag (read G / ref A free), 2 Tc (read C / ref T free)
};

// Base code A0 C1 G2 T3 (either case), `other` for every other byte, with
// no branch: after folding case (b & 0xDF) the four bases are the bytes
// 0x40 + {1, 3, 7, 20}, and bits 1-2 of those bytes (0, 1, 3, 2) are a Gray
// code whose decode t ^ (t >> 1) is the base code.
__device__ __forceinline__ int32_t base_code_or(uint32_t b, int32_t other) {
    const uint32_t u = b & 0xDFu;
    const bool acgt = (u & 0xE0u) == 0x40u && ((0x10008Au >> (u & 31u)) & 1u);
    const uint32_t t = (u >> 1) & 3u;
    return acgt ? (int32_t)(t ^ (t >> 1)) : other;
}
// Read-side code: A0 C1 G2 T3, other 4.  Window-side code maps "other" to 5
// so a read N never equals a window N (N scores mismatch, extend.rs:166-170).
__device__ __forceinline__ int32_t read_code(uint32_t b) { return base_code_or(b, 4); }
__device__ __forceinline__ int32_t win_code(uint32_t b) { return base_code_or(b, 5); }

// The window code that scores 0 against this read code under the active
// EditDir (extend.rs:176-183), or 6 (matches nothing) when none does.
__device__ __forceinline__ int32_t exempt_partner(int32_t rc, int32_t edit) {
    return (edit == 1 && rc == 2) ? 0 : (edit == 2 && rc == 1) ? 3 : 6;
}

// LUT score for codes (read, window) with window "other" = 5.
__device__ __forceinline__ int32_t lut_score(const GParams &P, int32_t rc, int32_t fc) {
    if (rc == fc) return P.match;
    if (fc == exempt_partner(rc, P.edit)) return 0;
    return P.mismatch;
}

// Row window of the unit-slope band.  off and half arrive pre-clamped to
// [-(m+1), n] and [0, n]; those clamps leave every row window unchanged.
__device__ __forceinline__ void row_window(int32_t off, int32_t half, int32_t n, int32_t i,
                                           int32_t &lo, int32_t &hi) {
    int32_t c = min(max(off + i, 1), n);
    lo = max(c - half, 1);
    hi = min(c + half, n);
}

// neutral_column of extend.rs:997-999 on (read code, window code) pairs:
// identical ACGT or an editing pair (ref A/read G, ref T/read C) in either
// EditDir.
__device__ __forceinline__ bool neutral_column(int32_t q, int32_t r) {
    return q < 4 && r < 4 && (q == r || (r == 0 && q == 2) || (r == 3 && q == 1));
}

// Writes the walked op runs, neutral flanks and soft clips into the pair's
// CIGAR slot [0, cap).  The traceback stores its t-th run (walk order, i.e.
// reverse) at slot[cap - 1 - t], so forward run f sits at
// slot[cap - nrev + f]; assembling left to right never overtakes an unread
// run because nrev + 1 < cap.  Mirrors extend.rs:1036-1059.
__device__ __forceinline__ uint32_t assemble_cigar(uint32_t *slot, uint32_t cap, uint32_t nrev,
                                                   uint32_t m, uint32_t read_start,
                                                   uint32_t neutral_left, uint32_t read_end,
                                                   uint32_t neutral_right) {
    uint32_t k = 0;
    auto push = [&](uint32_t kind, uint32_t len) {
        if (k > 0 && (slot[k - 1] & 0xFu) == kind) {
            slot[k - 1] = (((slot[k - 1] >> 4) + len) << 4) | kind;
        } else {
            slot[k++] = (len << 4) | kind;
        }
    };
    if (read_start > 0) push(OP_S, read_start);
    if (neutral_left > 0) push(OP_M, neutral_left);
    for (uint32_t f = 0; f < nrev; f++) {
        uint32_t w = slot[cap - nrev + f];
        push(w & 0xFu, w >> 4);
    }
    if (neutral_right > 0) push(OP_M, neutral_right);
    if (read_end < m) push(OP_S, m - read_end);
    return k;
}

struct PairView {
    const uint8_t *read;
    const uint8_t *win;
    int32_t m, n, off, half;
};

__device__ __forceinline__ int32_t clamp_off(int64_t off, int32_t m, int32_t n) {
    int64_t lo = -(int64_t)m - 1, hi = n;
    return (int32_t)(off < lo ? lo : (off > hi ? hi : off));
}
__device__ __forceinline__ int32_t clamp_half(int64_t half, int32_t n) {
    return (int32_t)(half < 0 ? 0 : (half > n ? n : half));
}

struct TbResult {
    uint32_t read_start, ref_start, nm, ni, nd, nrev;
};

// End trimming and score rebuild (extend.rs:932-990) reduce to the identity
// here: with gap_open, gap_ext >= 0 every prefix of the walked path ends on
// a cell whose H (or gap value) is positive, and every proper prefix stays
// below the first row-major maximum, so neither trim fires and the score is
// the best cell's H.  Neutral flanks and soft clips follow extend.rs:991-1059.
static __device__ void finish_pair(const PairView &pv, const GParams &P, int32_t best,
                            const TbResult &tb, uint32_t *slot, uint32_t cap, esp_ext *ext,
                            uint32_t *n_ops) {
    uint32_t read_start = tb.read_start, ref_start = tb.ref_start;
    uint32_t read_end = read_start + tb.nm + tb.ni;
    uint32_t nl = 0, nr = 0;
    if (best > 0) {
        int32_t delta = 0;
        const uint32_t lim = min(read_start, ref_start);
        for (uint32_t k = 1; k <= lim; k++) {
            const int32_t q = read_code(pv.read[read_start - k]);
            const int32_t r = win_code(pv.win[ref_start - k]);
            if (!neutral_column(q, r)) break;
            delta += lut_score(P, q, r);
            if (delta == 0) nl = k;
        }
        const uint32_t ref_end = ref_start + tb.nm + tb.nd;
        const uint32_t avail = min((uint32_t)pv.m - read_end, (uint32_t)pv.n - ref_end);
        delta = 0;
        for (uint32_t k = 0; k < avail; k++) {
            const int32_t q = read_code(pv.read[read_end + k]);
            const int32_t r = win_code(pv.win[ref_end + k]);
            if (!neutral_column(q, r)) break;
            delta += lut_score(P, q, r);
            if (delta == 0) nr = k + 1;
        }
        read_start -= nl;
        ref_start -= nl;
        read_end += nr;
    }
    *n_ops = assemble_cigar(slot, cap, tb.nrev, (uint32_t)pv.m, read_start, nl, read_end, nr);
    ext->read_start = read_start;
    ext->read_end = read_end;
    ext->ref_start = ref_start;
    ext->score = best;
}

// Work-queue pop by lane 0 of the wave, broadcast to all lanes.  Every
// caller runs it with the whole wave active; the scalar copy lets the
// compiler see that the unit, and every bound derived from it, is uniform.
__device__ __forceinline__ uint32_t wave_pop(uint32_t *counter, uint32_t lane, uint32_t step = 64u) {
    uint32_t base = 0;
    if (lane == 0) base = atomicAdd(counter, step);
    return __builtin_amdgcn_readfirstlane(__shfl(base, 0));
}

}  // namespace espd
