// esp_myers.hip -- EA-Myers bit-parallel filter (crates/map/src/myers_ea.rs),
// also the transcriptome engine's EA-Myers (crates/engine/src/myers.rs: the
// same EA alphabet and recurrence, plus its global mode).
//
// One lane owns one (read, text) pair.  The pattern is carried in W 64-bit
// words (W = ceil(m / 64) <= 4) with the edlib-form carry between words
// (myers_ea.rs:64-91 restated per word); the reference's u128 blocks are a
// different word size over the same DP, and only D[m][j] is read out, so the
// results are identical.  Bits above m in the last word are left unmasked:
// carries and shifts only move information upward, so they never reach bit
// m - 1, the only bit read.
//
// LDS per wave: the pattern-equality table peq[code][word][lane] (text codes
// A, C, G, T and "other", which matches nothing) and a 128-column staging
// chunk of text codes chunk[col][lane].  Text bytes reach LDS through aligned
// dword loads, eight in flight, so the column loop itself issues no global
// load.
#include "esp_kernels.hpp"

using namespace espd;

namespace espk {

namespace {

// Text bases staged per pass (LDS [MY_CHUNK][64] bytes per wave).  LDS sets
// the occupancy: 64 rather than 128 takes the 5-word kernel from 2.5 to 3.75
// waves per SIMD and its 1 M-pair launch from 0.94 to 0.83 ms; 32 is no
// faster, and prefetching the next column's Peq words costs more registers
// than it saves.
constexpr int MY_CHUNK = 64;

// Text code for the equality table: A0 C1 G2 T3, anything else 4 (no match);
// case-insensitive like the reference's raw-byte table (myers_ea.rs:42-56).
__device__ __forceinline__ uint32_t text_code(uint32_t b) {
    const uint32_t u = b & 0xDFu;
    return u == 'A' ? 0u : u == 'C' ? 1u : u == 'G' ? 2u : u == 'T' ? 3u : 4u;
}

// Writes text_code(src[q]) to chunk[q * 64] for q in [0, len): head and tail
// bytewise, the aligned body as dwords, eight loads in flight.
__device__ void stage_codes(const uint8_t *src, int32_t len, uint8_t *chunk) {
    const int32_t head = min(len, (int32_t)((4u - ((uint32_t)(uintptr_t)src & 3u)) & 3u));
    for (int32_t q = 0; q < head; q++) chunk[q * 64] = (uint8_t)text_code(src[q]);
    const uint32_t *d = (const uint32_t *)(src + head);
    const int32_t nd = (len - head) >> 2;
    for (int32_t q = 0; q < nd; q += 8) {
        uint32_t w[8];
#pragma unroll
        for (int u = 0; u < 8; u++) w[u] = q + u < nd ? d[q + u] : 0u;
#pragma unroll
        for (int u = 0; u < 8; u++) {
            if (q + u < nd) {
                const int32_t r = head + 4 * (q + u);
#pragma unroll
                for (int b = 0; b < 4; b++) chunk[(r + b) * 64] = (uint8_t)text_code((w[u] >> (8 * b)) & 0xFFu);
            }
        }
    }
    for (int32_t q = head + 4 * nd; q < len; q++) chunk[q * 64] = (uint8_t)text_code(src[q]);
}

// ea_matches (myers_ea.rs:28-38): read A = {A}, C = {C, T}, G = {G, A}, T = {T}.
// Builds peq rows for the read (reversed when `reverse`), staging the read
// through the chunk area MY_CHUNK bases at a time.
template <int W>
__device__ void build_peq(uint64_t *peq, uint8_t *chunk, const uint8_t *read, int32_t m,
                          bool reverse) {
    uint64_t A[W], C[W], G[W], T[W];
#pragma unroll
    for (int w = 0; w < W; w++) A[w] = C[w] = G[w] = T[w] = 0;
    for (int32_t q0 = 0; q0 < m; q0 += MY_CHUNK) {
        const int32_t len = min(MY_CHUNK, m - q0);
        stage_codes(read + (reverse ? m - q0 - len : q0), len, chunk);
        for (int32_t q = 0; q < len; q++) {
            const uint32_t rc = chunk[(reverse ? len - 1 - q : q) * 64];
            const int32_t i = q0 + q;  // pattern position
            const uint64_t bit = 1ull << (i & 63);
            const int32_t wi = i >> 6;
#pragma unroll
            for (int w = 0; w < W; w++) {
                if (w == wi) {
                    if (rc == 0) A[w] |= bit;
                    if (rc == 1) { C[w] |= bit; T[w] |= bit; }
                    if (rc == 2) { G[w] |= bit; A[w] |= bit; }
                    if (rc == 3) T[w] |= bit;
                }
            }
        }
    }
#pragma unroll
    for (int w = 0; w < W; w++) {
        peq[(0 * W + w) * 64] = A[w];
        peq[(1 * W + w) * 64] = C[w];
        peq[(2 * W + w) * 64] = G[w];
        peq[(3 * W + w) * 64] = T[w];
        peq[(4 * W + w) * 64] = 0;
    }
}

// Column scan over the text (reversed when `reverse`): (best, best_at) with
// the strict-< earliest-end rule of infix_best_end (myers_ea.rs:141-156).
// With `global` (myers::global, engine/src/myers.rs:149-168) every column
// enters row 0 with a +1 delta (D[0][j] = j) and the result is the last
// column's D[m][n], floored at 0.
template <int W>
__device__ void myers_scan(const uint64_t *peq, uint8_t *chunk, int32_t m, const uint8_t *text,
                           int32_t n, bool reverse, bool global, int32_t &best_out, uint32_t &at_out) {
    const int32_t lastw = (m - 1) >> 6;
    const uint32_t lastb = (uint32_t)(m - 1) & 63u;
    uint64_t pv[W], mv[W];
#pragma unroll
    for (int w = 0; w < W; w++) {
        pv[w] = w < lastw ? ~0ull : (w == lastw ? (lastb == 63 ? ~0ull : ((1ull << (lastb + 1)) - 1)) : 0);
        mv[w] = 0;
    }
    int32_t score = m, best = m;
    uint32_t best_at = 0;
    for (int32_t c0 = 0; c0 < n; c0 += MY_CHUNK) {
        const int32_t len = min(MY_CHUNK, n - c0);
        stage_codes(text + (reverse ? n - c0 - len : c0), len, chunk);
        for (int32_t q = 0; q < len; q++) {
            const uint32_t code = chunk[(reverse ? len - 1 - q : q) * 64];
            const uint64_t *row = peq + code * (W * 64);
            int32_t hin = global ? 1 : 0;
#pragma unroll
            for (int w = 0; w < W; w++) {
                if (w <= lastw) {
                    const uint64_t eq_in = row[w * 64];
                    const uint64_t hneg = hin < 0 ? 1ull : 0ull;
                    const uint64_t hpos = hin > 0 ? 1ull : 0ull;
                    const uint64_t p = pv[w], mm = mv[w];
                    const uint64_t xv = eq_in | mm;
                    const uint64_t eq = eq_in | hneg;
                    const uint64_t xh = (((eq & p) + p) ^ p) | eq;
                    const uint64_t ph = mm | ~(xh | p);
                    const uint64_t mh = p & xh;
                    const uint32_t hb = w < lastw ? 63u : lastb;
                    hin = (int32_t)((ph >> hb) & 1ull) - (int32_t)((mh >> hb) & 1ull);
                    const uint64_t phs = (ph << 1) | hpos;
                    const uint64_t mhs = (mh << 1) | hneg;
                    pv[w] = mhs | ~(xv | phs);
                    mv[w] = phs & xv;
                }
            }
            score += hin;
            if (score < best) {
                best = score;
                best_at = (uint32_t)(c0 + q + 1);
            }
        }
    }
    if (global) {
        best = max(score, 0);
        best_at = (uint32_t)n;
    }
    best_out = best;
    at_out = best_at;
}

// ---- 32-bit words with a full carry chain (infix modes) ----------------
//
// The pattern sits left-aligned in NW 32-bit words: row i at bit
// off + i, off = 32 NW - m, so row m - 1 is always bit 31 of the last word.
// The off rows below it match every text byte and start with zero vertical
// deltas; by induction every one of them keeps D = 0 in every column, which
// is exactly the free top row of the infix recurrence, so D[m][j] is
// unchanged.  (Global distances, whose top row is D[0][j] = j, keep the
// 64-bit kernel above.)  Each column is one multi-word step of the textbook
// recurrence: the addition carries across words and the two shifts are
// funnel shifts, so no per-word horizontal carry is needed.

// Peq rows (A, C, G, T, other) for the read, left-aligned, padding rows set.
// Word by word: the bases that land in word w set their bit in each row whose
// letter they match (ea_matches, myers_ea.rs:28-38), read from the nibble
// table 0x85A1 (A: A; C: C,T; G: G,A; T: T; other: none; rows A C G T =
// bits 0 1 2 3).
template <int NW>
__device__ void build_peq32(uint32_t *peq, uint8_t *chunk, const uint8_t *read, int32_t m, bool reverse) {
    const int32_t off = 32 * NW - m;
    uint32_t A[NW], C[NW], G[NW], T[NW];
#pragma unroll
    for (int w = 0; w < NW; w++) A[w] = C[w] = G[w] = T[w] = 0;
    for (int32_t q0 = 0; q0 < m; q0 += MY_CHUNK) {
        const int32_t len = min(MY_CHUNK, m - q0);
        stage_codes(read + (reverse ? m - q0 - len : q0), len, chunk);
#pragma unroll
        for (int w = 0; w < NW; w++) {
            // pattern indices of this pass that land in word w
            const int32_t lo = max(32 * w - off, q0), hi = min(32 * w + 32 - off, q0 + len);
            for (int32_t i = lo; i < hi; i++) {
                const uint32_t rc = chunk[(reverse ? len - 1 - (i - q0) : i - q0) * 64];
                const uint32_t mk = (0x85A1u >> (4 * rc)) & 15u;
                const uint32_t b = (uint32_t)(i + off - 32 * w);
                A[w] |= (mk & 1u) << b;
                C[w] |= ((mk >> 1) & 1u) << b;
                G[w] |= ((mk >> 2) & 1u) << b;
                T[w] |= (mk >> 3) << b;
            }
        }
    }
#pragma unroll
    for (int w = 0; w < NW; w++) {
        const int32_t lo = 32 * w;
        const uint32_t pad = off >= lo + 32 ? 0xFFFFFFFFu : (off <= lo ? 0u : ((1u << (off - lo)) - 1u));
        peq[(0 * NW + w) * 64] = A[w] | pad;
        peq[(1 * NW + w) * 64] = C[w] | pad;
        peq[(2 * NW + w) * 64] = G[w] | pad;
        peq[(3 * NW + w) * 64] = T[w] | pad;
        peq[(4 * NW + w) * 64] = pad;
    }
}

// Column scan: (best, best_at) with the strict-< earliest-end rule of
// infix_best_end (myers_ea.rs:141-156).
template <int NW>
__device__ void myers_scan32(const uint32_t *peq, uint8_t *chunk, int32_t m, const uint8_t *text, int32_t n,
                             bool reverse, int32_t &best_out, uint32_t &at_out) {
    const int32_t off = 32 * NW - m;
    uint32_t pv[NW], mv[NW];
#pragma unroll
    for (int w = 0; w < NW; w++) {
        const int32_t lo = 32 * w;
        const uint32_t pad = off >= lo + 32 ? 0xFFFFFFFFu : (off <= lo ? 0u : ((1u << (off - lo)) - 1u));
        pv[w] = ~pad;
        mv[w] = 0;
    }
    int32_t score = m, best = m;
    uint32_t best_at = 0;
    for (int32_t c0 = 0; c0 < n; c0 += MY_CHUNK) {
        const int32_t len = min(MY_CHUNK, n - c0);
        stage_codes(text + (reverse ? n - c0 - len : c0), len, chunk);
        for (int32_t q = 0; q < len; q++) {
            const uint32_t code = chunk[(reverse ? len - 1 - q : q) * 64];
            const uint32_t *row = peq + code * (NW * 64);
            uint32_t ph[NW], mh[NW], xv[NW];
            uint32_t carry = 0;
#pragma unroll
            for (int w = 0; w < NW; w++) {
                const uint32_t eq = row[w * 64], p = pv[w];
                xv[w] = eq | mv[w];
                uint32_t c1 = 0;
                const uint32_t sum = __builtin_addc(eq & p, p, carry, &c1);
                carry = c1;
                const uint32_t xh = (sum ^ p) | eq;
                ph[w] = mv[w] | ~(xh | p);
                mh[w] = p & xh;
            }
            score += (int32_t)(ph[NW - 1] >> 31) - (int32_t)(mh[NW - 1] >> 31);
#pragma unroll
            for (int w = 0; w < NW; w++) {
                const uint32_t phs = w ? __builtin_amdgcn_alignbit(ph[w], ph[w - 1], 31) : ph[0] << 1;
                const uint32_t mhs = w ? __builtin_amdgcn_alignbit(mh[w], mh[w - 1], 31) : mh[0] << 1;
                pv[w] = mhs | ~(xv[w] | phs);
                mv[w] = phs & xv[w];
            }
            if (score < best) {
                best = score;
                best_at = (uint32_t)(c0 + q + 1);
            }
        }
    }
    best_out = best;
    at_out = best_at;
}

template <int NW>
__global__ __launch_bounds__(MYERS_BLOCK) void k_myers32(MyersArgs a) {
    extern __shared__ uint32_t lds32[];  // peq [5][NW][64] u32, then chunk [MY_CHUNK][64] bytes
    const uint32_t lane = threadIdx.x & 63u;
    uint32_t *peq = lds32 + lane;
    uint8_t *chunk = (uint8_t *)(lds32 + 5 * NW * 64) + lane;
    const uint32_t first = a.range ? a.range[0] : 0u;
    const uint32_t count = a.range ? a.range[1] : a.npairs;
    for (;;) {
        const uint32_t base = wave_pop(a.counter, lane);
        if (base >= count) break;
        const uint32_t idx = base + lane;
        if (idx < count) {
            const uint32_t p = a.order ? a.order[first + idx] : idx;
            const int32_t m = (int32_t)a.read_len[p], n = (int32_t)a.text_len[p];
            const uint8_t *read = a.reads + a.read_off[p];
            const uint8_t *text = a.texts + a.text_off[p];
            if (m > 32 * NW && m <= 256 && a.ovf_list) {
                // a wider instance takes it (the host's bound was short)
                a.ovf_list[atomicAdd(&a.ovf_range[1], 1u)] = p;
            } else if (m == 0 || m > 32 * NW) {
                // long::infix_best_end guards (myers_ea.rs:275-280); without
                // an overflow list the host sizes NW for every m <= 256, so
                // only the guards land here
                a.dist[p] = m == 0 ? 0 : m;
                if (a.end) a.end[p] = 0;
                if (a.start) a.start[p] = 0;
            } else {
                int32_t d;
                uint32_t e;
                if (a.end || !a.start) {
                    build_peq32<NW>(peq, chunk, read, m, false);
                    myers_scan32<NW>(peq, chunk, m, text, n, false, d, e);
                    a.dist[p] = d;
                    if (a.end) a.end[p] = e;
                }
                if (a.start) {
                    build_peq32<NW>(peq, chunk, read, m, true);
                    myers_scan32<NW>(peq, chunk, m, text, n, true, d, e);
                    a.dist[p] = d;
                    a.start[p] = (uint64_t)(n - (int32_t)e);
                }
            }
        }
    }
}

size_t myers32_lds(int nw) { return (size_t)5 * nw * 64 * 4 + (size_t)MY_CHUNK * 64; }

template <int W>
__global__ __launch_bounds__(MYERS_BLOCK) void k_myers(MyersArgs a) {
    extern __shared__ uint64_t lds[];  // peq [5][W][64] u64, then chunk [MY_CHUNK][64] bytes
    const uint32_t lane = threadIdx.x & 63u;
    uint64_t *peq = lds + lane;
    uint8_t *chunk = (uint8_t *)(lds + 5 * W * 64) + lane;
    const uint32_t first = a.range ? a.range[0] : 0u;
    const uint32_t count = a.range ? a.range[1] : a.npairs;
    for (;;) {
        const uint32_t base = wave_pop(a.counter, lane);
        if (base >= count) break;
        const uint32_t idx = base + lane;
        if (idx < count) {
            const uint32_t p = a.order ? a.order[first + idx] : idx;
            const int32_t m = (int32_t)a.read_len[p], n = (int32_t)a.text_len[p];
            const uint8_t *read = a.reads + a.read_off[p];
            const uint8_t *text = a.texts + a.text_off[p];
            if (m == 0 || m > 256) {
                // long::infix_best_end guards (myers_ea.rs:275-280)
                a.dist[p] = m == 0 ? 0 : m;
                if (a.end) a.end[p] = 0;
                if (a.start) a.start[p] = 0;
            } else {
                int32_t d;
                uint32_t e;
                // Forward scan for the end (and for infix, which asks for
                // neither position); the distance of the reversed pair is
                // the same minimum, so a start-only call scans once.
                if (a.end || !a.start) {
                    build_peq<W>(peq, chunk, read, m, false);
                    myers_scan<W>(peq, chunk, m, text, n, false, a.global != 0, d, e);
                    a.dist[p] = d;
                    if (a.end) a.end[p] = e;
                }
                if (a.start) {
                    // infix_best_start: best end of the doubly reversed pair
                    // (myers_ea.rs:166-185); start = n - rev_end.
                    build_peq<W>(peq, chunk, read, m, true);
                    myers_scan<W>(peq, chunk, m, text, n, true, false, d, e);
                    a.dist[p] = d;
                    a.start[p] = (uint64_t)(n - (int32_t)e);
                }
            }
        }
    }
}

__global__ void k_engine_prep(EnginePrepArgs a) {
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < a.n; k += gridDim.x * blockDim.x) {
        const int64_t anchor = a.anchor[k];
        const int64_t lo = anchor - (int64_t)a.flank;
        const int64_t hi = anchor + (int64_t)a.read_len[k] + (int64_t)a.flank;
        const uint64_t wlo = lo > 0 ? (uint64_t)lo : 0u;
        const uint64_t whi = min(hi > 0 ? (uint64_t)hi : 0u, (uint64_t)a.tx_len[k]);
        const bool ok = whi > wlo;
        a.text_off[k] = a.tx_off[k] + (ok ? wlo : 0u);
        a.text_len[k] = ok ? (uint32_t)(whi - wlo) : 0u;
        a.status[k] = ok ? 1 : 0;
    }
}

size_t myers_lds(int words) { return (size_t)5 * words * 64 * 8 + (size_t)MY_CHUNK * 64; }

}  // namespace

hipError_t launch_engine_prep(const EnginePrepArgs &a, hipStream_t s) {
    const uint32_t blocks = std::max<uint32_t>(1u, std::min<uint32_t>((a.n + 255) / 256, 4096u));
    hipLaunchKernelGGL(k_engine_prep, dim3(blocks), dim3(256), 0, s, a);
    return hipGetLastError();
}

// Pattern lengths up to maxm (clamped to 1..256): infix modes run the 32-bit
// kernel with ceil(maxm / 32) words, global mode the 64-bit kernel.
hipError_t launch_myers(const MyersArgs &a, int maxm, int blocks, hipStream_t s) {
    maxm = std::min(std::max(maxm, 1), 256);
    if (a.global) {
        const int words = (maxm + 63) / 64;
        const size_t lds = myers_lds(words);
        switch (words) {
        case 1: hipLaunchKernelGGL(k_myers<1>, dim3(blocks), dim3(MYERS_BLOCK), lds, s, a); break;
        case 2: hipLaunchKernelGGL(k_myers<2>, dim3(blocks), dim3(MYERS_BLOCK), lds, s, a); break;
        case 3: hipLaunchKernelGGL(k_myers<3>, dim3(blocks), dim3(MYERS_BLOCK), lds, s, a); break;
        default: hipLaunchKernelGGL(k_myers<4>, dim3(blocks), dim3(MYERS_BLOCK), lds, s, a); break;
        }
        return hipGetLastError();
    }
    const int nw = (maxm + 31) / 32;
    const size_t lds = myers32_lds(nw);
    switch (nw) {
    case 1: hipLaunchKernelGGL(k_myers32<1>, dim3(blocks), dim3(MYERS_BLOCK), lds, s, a); break;
    case 2: hipLaunchKernelGGL(k_myers32<2>, dim3(blocks), dim3(MYERS_BLOCK), lds, s, a); break;
    case 3: hipLaunchKernelGGL(k_myers32<3>, dim3(blocks), dim3(MYERS_BLOCK), lds, s, a); break;
    case 4: hipLaunchKernelGGL(k_myers32<4>, dim3(blocks), dim3(MYERS_BLOCK), lds, s, a); break;
    case 5: hipLaunchKernelGGL(k_myers32<5>, dim3(blocks), dim3(MYERS_BLOCK), lds, s, a); break;
    case 6: hipLaunchKernelGGL(k_myers32<6>, dim3(blocks), dim3(MYERS_BLOCK), lds, s, a); break;
    case 7: hipLaunchKernelGGL(k_myers32<7>, dim3(blocks), dim3(MYERS_BLOCK), lds, s, a); break;
    default: hipLaunchKernelGGL(k_myers32<8>, dim3(blocks), dim3(MYERS_BLOCK), lds, s, a); break;
    }
    return hipGetLastError();
}

template <typename K>
static int occupancy(K kernel, size_t lds) {
    int nb = 0;
    const hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kernel, MYERS_BLOCK, lds);
    return (e == hipSuccess && nb > 0) ? nb : 1;
}

int myers_blocks_per_cu(int maxm, bool global) {
    maxm = std::min(std::max(maxm, 1), 256);
    if (global) {
        const int words = (maxm + 63) / 64;
        const size_t lds = myers_lds(words);
        switch (words) {
        case 1: return occupancy(k_myers<1>, lds);
        case 2: return occupancy(k_myers<2>, lds);
        case 3: return occupancy(k_myers<3>, lds);
        default: return occupancy(k_myers<4>, lds);
        }
    }
    const int nw = (maxm + 31) / 32;
    const size_t lds = myers32_lds(nw);
    switch (nw) {
    case 1: return occupancy(k_myers32<1>, lds);
    case 2: return occupancy(k_myers32<2>, lds);
    case 3: return occupancy(k_myers32<3>, lds);
    case 4: return occupancy(k_myers32<4>, lds);
    case 5: return occupancy(k_myers32<5>, lds);
    case 6: return occupancy(k_myers32<6>, lds);
    case 7: return occupancy(k_myers32<7>, lds);
    default: return occupancy(k_myers32<8>, lds);
    }
}

}  // namespace espk
