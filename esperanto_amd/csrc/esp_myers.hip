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

constexpr int MY_CHUNK = 128;

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
// through the chunk area 128 bases at a time.
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

template <int W>
__global__ __launch_bounds__(MYERS_BLOCK) void k_myers(MyersArgs a) {
    extern __shared__ uint64_t lds[];  // peq [5][W][64] u64, then chunk [128][64] bytes
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

hipError_t launch_myers(const MyersArgs &a, int words, int blocks, hipStream_t s) {
    const size_t lds = myers_lds(words);
    switch (words) {
    case 1: hipLaunchKernelGGL(k_myers<1>, dim3(blocks), dim3(MYERS_BLOCK), lds, s, a); break;
    case 2: hipLaunchKernelGGL(k_myers<2>, dim3(blocks), dim3(MYERS_BLOCK), lds, s, a); break;
    case 3: hipLaunchKernelGGL(k_myers<3>, dim3(blocks), dim3(MYERS_BLOCK), lds, s, a); break;
    default: hipLaunchKernelGGL(k_myers<4>, dim3(blocks), dim3(MYERS_BLOCK), lds, s, a); break;
    }
    return hipGetLastError();
}

int myers_blocks_per_cu(int words) {
    int nb = 0;
    const size_t lds = myers_lds(words);
    hipError_t e = hipSuccess;
    switch (words) {
    case 1: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_myers<1>, MYERS_BLOCK, lds); break;
    case 2: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_myers<2>, MYERS_BLOCK, lds); break;
    case 3: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_myers<3>, MYERS_BLOCK, lds); break;
    default: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_myers<4>, MYERS_BLOCK, lds); break;
    }
    return (e == hipSuccess && nb > 0) ? nb : 1;
}

}  // namespace espk
