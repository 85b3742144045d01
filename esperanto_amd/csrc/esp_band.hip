// esp_band.hip -- diagonal-frame banded Gotoh kernel for interior bands.
//
// A pair is "interior" when every DP row window [c-half, c+half] with
// c = offset + i lies inside [1, n] (offset >= half, offset + m + half <= n).
// Then row i covers columns cb + i + k, k in [0, W), W = 2*half + 1, and the
// band shifts right by exactly one column per row (extend.rs:664-668), so in
// band coordinates k the three DP dependencies are fixed register moves:
//   up   (i-1, j)   -> band index k+1 of the previous row
//   diag (i-1, j-1) -> band index k   of the previous row
//   left (i, j-1)   -> band index k-1 of this row.
// One lane owns one pair; the previous row's H and E live in VGPR arrays of
// W entries (static indices after unrolling, updated in place), so a cell is
// pure VALU work with no masking.  The window codes of the wave's 64 pairs
// sit in LDS as codes[j - cb][lane], so the load for cell (i, k) is at the
// wave-uniform row (i + k) and is bank-conflict free.
//
// Stored values are biased: U = H + open_ext and B = E + ext, which is what
// the next row's E/F recurrences consume, so each cell needs one add per
// stored value.  The trace keeps 4 bits per cell, (D < H, E < H, E from
// extension, F from extension), each the sign bit of one subtraction shifted
// in with v_alignbit; the traceback carries the cell value along the path, so
// H == 0 (the reference's H_START) needs no stored bit.
// Pairs that are not interior, or whose half has no specialisation, run the
// general strip kernel (esp_kernels.hip).
#include "esp_kernels.hpp"

using namespace espd;

namespace espk {

namespace {

// Traceback over the band trace layout (extend.rs:884-930) with value
// tracking; runs are written in walk order to slot[cap-1], slot[cap-2], ...
template <int W>
__device__ TbResult traceback_band(const PairView &pv, const GParams &P, const uint8_t *codes,
                                   uint32_t lane, const uint32_t *tr, int32_t cb,
                                   int32_t best, int32_t bi, int32_t bj, uint32_t *slot,
                                   uint32_t cap) {
    constexpr int WW = (W + 7) / 8;
    int32_t i = bi, j = bj, V = best;
    uint32_t state = 0;
    uint32_t cur_op = 0xFF, cur_len = 0, nrev = 0, nm = 0, ni = 0, nd = 0;
    auto emit = [&](uint32_t op) {
        if (op == cur_op) {
            cur_len++;
        } else {
            if (cur_len) slot[cap - 1 - nrev++] = (cur_len << 4) | cur_op;
            cur_op = op;
            cur_len = 1;
        }
    };
    // Column cache: the nibbles of band cell ck for rows ci0-31 .. ci0, row r
    // at bit 4*(ci0 - r) of the 128-bit pair (lo, hi).  Diagonal steps keep
    // k, so one batch of 32 independent loads serves up to 32 steps.
    int32_t ci0 = -1, ck = -1;
    uint64_t nlo = 0, nhi = 0;
    while (i > 0 && j > 0) {
        const int32_t k = j - (cb + i);
        if (k < 0 || k >= W) break;
        if (k != ck || i > ci0 || i <= ci0 - 32) {
            const int32_t t = k >> 3;
            const int32_t c = min(8, W - 8 * t);
            const int32_t sh = 4 * (c - 1 - (k & 7));
            uint32_t w[32];
#pragma unroll
            for (int q = 0; q < 32; q++)
                w[q] = (i - q >= 1) ? tr[((size_t)(i - 1 - q) * WW + t) * 64] : 0u;
            nlo = 0;
            nhi = 0;
#pragma unroll
            for (int q = 0; q < 16; q++) {
                nlo |= (uint64_t)((w[q] >> sh) & 0xFu) << (4 * q);
                nhi |= (uint64_t)((w[q + 16] >> sh) & 0xFu) << (4 * q);
            }
            ci0 = i;
            ck = k;
        }
        const int32_t d = ci0 - i;
        const uint32_t nib = (uint32_t)((d < 16 ? (nlo >> (4 * d)) : (nhi >> (4 * (d - 16)))) & 0xFu);
        if (state == 0) {
            if (V == 0) break;  // H == 0: the reference's H_START
            if (!(nib & 8u)) {
                const int32_t rc = codes[(size_t)i * 64 + lane] >> 5;
                const int32_t fc = (codes[(size_t)(j - cb) * 64 + lane] & 31) / 5;
                V -= lut_score(P, rc, fc);
                emit(OP_M);
                nm++;
                i--;
                j--;
            } else if (!(nib & 4u)) {
                emit(OP_I);
                ni++;
                if (nib & 2u) {
                    state = 1;
                    V -= P.ext;
                } else {
                    V -= P.oe;
                }
                i--;
            } else {
                emit(OP_D);
                nd++;
                if (nib & 1u) {
                    state = 2;
                    V -= P.ext;
                } else {
                    V -= P.oe;
                }
                j--;
            }
        } else if (state == 1) {
            emit(OP_I);
            ni++;
            if (nib & 2u) {
                V -= P.ext;
            } else {
                state = 0;
                V -= P.oe;
            }
            i--;
        } else {
            emit(OP_D);
            nd++;
            if (nib & 1u) {
                V -= P.ext;
            } else {
                state = 0;
                V -= P.oe;
            }
            j--;
        }
    }
    if (cur_len) slot[cap - 1 - nrev++] = (cur_len << 4) | cur_op;
    TbResult r;
    r.read_start = (uint32_t)max(i, 0);
    r.ref_start = (uint32_t)max(j, 0);
    r.nm = nm;
    r.ni = ni;
    r.nd = nd;
    r.nrev = nrev;
    return r;
}

template <int HALF>
__global__ __launch_bounds__(BAND_BLOCK) void k_gotoh_band(GotohArgs a) {
    constexpr int W = 2 * HALF + 1;
    constexpr int WW = (W + 7) / 8;
    extern __shared__ uint8_t codes[];  // [span][64]: (read code << 5) | 5 * window code
    uint32_t *lut_lds = (uint32_t *)codes;  // row 0 (never a code row): the 5 row LUTs
    const uint32_t lane = threadIdx.x;
    const uint32_t slot = blockIdx.x * BAND_BLOCK + lane;
    const GParams P = a.P;
    const int32_t oe = P.oe, ext = P.ext;
    const int32_t noe = -oe;
    const int32_t negx = NEG + ext;
    const uint32_t first = a.range[0], total = a.range[1];
    const uint32_t nunits = (total + 63) / 64;
    for (;;) {
        // A unit is 64 consecutive pairs of this class in `order`, which the
        // planner sorted by read length descending: the first is the longest.
        const uint32_t u = wave_pop(a.counter, lane, 1u);
        if (u >= nunits) break;
        const uint32_t ubegin = first + 64 * u;
        const uint32_t ucount = min(64u, total - 64 * u);
        const int32_t umaxm = (int32_t)a.read_len[a.order[ubegin]];
        const bool active = lane < ucount;
        const uint32_t p = active ? a.order[ubegin + lane] : a.order[ubegin];
        PairView pv;
        pv.read = a.reads + a.read_off[p];
        pv.win = a.wins + a.win_off[p];
        pv.m = active ? (int32_t)a.read_len[p] : 0;
        pv.n = (int32_t)a.win_len[p];
        pv.off = (int32_t)a.offset[p];
        pv.half = HALF;
        const int32_t m = pv.m, n = pv.n;
        const int32_t cb = pv.off - HALF;
        // Window codes of columns cb + 1 .. cb + m + W - 1 (interior: inside
        // [1, n]) into LDS rows 1 .. m + W - 1.  Aligned dwords carry the
        // body, eight loads in flight per round; head and tail go bytewise,
        // so no load leaves the window.
        {
            const int32_t L = m > 0 ? m + W - 1 : 0;
            const uint8_t *s0 = pv.win + cb;
            const int32_t head = min(L, (int32_t)((4u - ((uint32_t)(uintptr_t)s0 & 3u)) & 3u));
            for (int32_t q = 0; q < head; q++)
                codes[(size_t)(1 + q) * 64 + lane] = (uint8_t)(5 * win_code(s0[q]));
            const uint32_t *d = (const uint32_t *)(s0 + head);
            const int32_t nd = (L - head) >> 2;
            for (int32_t q = 0; q < nd; q += 8) {
                uint32_t w[8];
#pragma unroll
                for (int u = 0; u < 8; u++) w[u] = q + u < nd ? d[q + u] : 0u;
#pragma unroll
                for (int u = 0; u < 8; u++) {
                    if (q + u < nd) {
                        const int32_t r = 1 + head + 4 * (q + u);
#pragma unroll
                        for (int b = 0; b < 4; b++)
                            codes[(size_t)(r + b) * 64 + lane] =
                                (uint8_t)(5 * win_code((w[u] >> (8 * b)) & 0xFFu));
                    }
                }
            }
            for (int32_t q = head + 4 * nd; q < L; q++)
                codes[(size_t)(1 + q) * 64 + lane] = (uint8_t)(5 * win_code(s0[q]));
            // Read code of row r (read base r-1) in bits 5-7 of LDS row r: the
            // row loop then needs no global load (a vector-memory wait there
            // would also wait for every trace store in flight).
            for (int32_t q = 0; q < m; q += 8) {
                uint32_t rb[8];
#pragma unroll
                for (int u = 0; u < 8; u++) rb[u] = q + u < m ? pv.read[q + u] : 0u;
#pragma unroll
                for (int u = 0; u < 8; u++) {
                    if (q + u < m) {
                        uint8_t *cell = codes + (size_t)(1 + q + u) * 64 + lane;
                        *cell = (uint8_t)(*cell | (read_code(rb[u]) << 5));
                    }
                }
            }
        }
        if (lane < 5) lut_lds[lane] = a.lut[lane];
        __syncthreads();  // one wave per block: orders the LUT writes before any lane reads them

        // Previous row, band index k: U = H + open_ext and B = E + ext.  Wide
        // bands keep both as int16 halves of one VGPR (U low, B high): the
        // reads fold into SDWA operands and one v_perm repacks, which halves
        // the register arrays (the planner admits only pairs whose values fit
        // int16; the rest run the int32 strip kernel).
        constexpr bool PACKED = HALF >= 48;
        constexpr int32_t NEG16 = -20000;  // packed "E = NEG": below every reachable U
        int32_t Ur[PACKED ? 1 : W], Bx[PACKED ? 1 : W];
        uint32_t Pk[PACKED ? W : 1];
        const uint32_t pmargin = ((uint32_t)oe & 0xFFFFu) | ((uint32_t)NEG16 << 16);
#pragma unroll
        for (int k = 0; k < W; k++) {
            if constexpr (PACKED) {
                Pk[k] = pmargin;  // row 0: H = 0, E = NEG
            } else {
                Ur[k] = oe;
                Bx[k] = negx;
            }
        }
        int32_t best = 0, bi = 0, bj = 0;
        int32_t touched = 0, trow = 0, tleft = 0, tclamped = 0;
        // Per-wave trace slab: word w of this lane at tr[w * 64].
        uint32_t *tr = a.trace + (size_t)blockIdx.x * a.maxw * 64 + lane;
        // Per-read-code row LUT: 5-bit signed scores for window codes 0..5
        // (A C G T, -, other) at bit 5*code; the window codes in LDS are
        // stored pre-multiplied by 5.
        for (int32_t i = 1; i <= umaxm; i++) {
            if (i <= m) {
                const uint8_t *crow = codes + (size_t)i * 64 + lane;
                const uint32_t rowlut = lut_lds[crow[0] >> 5];
                int32_t Ul = oe, Fx = negx, key = 0, H0 = 0, Hl = 0;
                uint32_t acc = 0;
                uint32_t *trow_p = tr + (size_t)(i - 1) * WW * 64;
#pragma unroll
                for (int k = 0; k < W; k++) {
                    int32_t Uup, Bup, Ud;
                    if constexpr (PACKED) {
                        const uint32_t pu = (k + 1 < W) ? Pk[k + 1] : pmargin;
                        Uup = (int32_t)(int16_t)(pu & 0xFFFFu);
                        Bup = ((int32_t)pu) >> 16;
                        Ud = (int32_t)(int16_t)(Pk[k] & 0xFFFFu);
                    } else {
                        Uup = (k + 1 < W) ? Ur[k + 1] : oe;
                        Bup = (k + 1 < W) ? Bx[k + 1] : negx;
                        Ud = Ur[k];
                    }
                    const uint32_t code5 = crow[k * 64];
                    const int32_t sc = __builtin_amdgcn_sbfe((int32_t)rowlut, code5, 5);
                    int32_t D;
                    asm("v_add3_u32 %0, %1, %2, %3" : "=v"(D) : "v"(Ud), "v"(sc), "s"(noe));
                    const int32_t E = max(Uup, Bup);
                    const int32_t F = max(Ul, Fx);
                    const int32_t H = max(max(max(D, E), F), 0);
                    // Sign bits: D < H, E < H, E from extension, F from extension,
                    // shifted into the trace word with v_alignbit.
                    const uint32_t x0 = (uint32_t)(D - H), x1 = (uint32_t)(E - H);
                    const uint32_t x2 = (uint32_t)(Uup - Bup), x3 = (uint32_t)(Ul - Fx);
                    asm("v_alignbit_b32 %0, %0, %1, 31\n\t"
                                 "v_alignbit_b32 %0, %0, %2, 31\n\t"
                                 "v_alignbit_b32 %0, %0, %3, 31\n\t"
                                 "v_alignbit_b32 %0, %0, %4, 31"
                                 : "+v"(acc) : "v"(x0), "v"(x1), "v"(x2), "v"(x3));
                    const int32_t Un = H + oe;
                    if constexpr (PACKED) {
                        Pk[k] = __builtin_amdgcn_perm((uint32_t)(E + ext), (uint32_t)Un, 0x05040100u);
                    } else {
                        Ur[k] = Un;
                        Bx[k] = E + ext;
                    }
                    Ul = Un;
                    Fx = F + ext;
                    // Running row key (H << 8 | 255 - k): kept as a sequential chain
                    // so the compiler cannot defer it into a tree that keeps every
                    // cell's H live until the row ends.
                    int32_t kk;
                    asm("v_lshl_or_b32 %0, %1, 8, %2" : "=v"(kk) : "v"(H), "s"(255 - k));
                    asm("v_max_i32 %0, %1, %2" : "=v"(key) : "v"(kk), "v"(key));
                    if (k == 0) H0 = H;
                    if (k == W - 1) Hl = H;
                    if ((k & 7) == 7 || k == W - 1) {
                        trow_p[(k >> 3) * 64] = acc;
                        acc = 0;
                    }
                }
                const int32_t rmax = key >> 8;
                const int32_t rarg = 255 - (key & 255);
                if (!touched) {
                    if (H0 == rmax) {
                        touched = 1;
                        trow = i;
                        tleft = 1;
                        tclamped = cb + i <= 1;
                    } else if (cb + i + W - 1 < n && Hl == rmax) {
                        touched = 1;
                        trow = i;
                        tleft = 0;
                        tclamped = 0;
                    }
                }
                if (rmax > best) {
                    best = rmax;
                    bi = i;
                    bj = cb + i + rarg;
                }
            }
        }
        // extend_hint (extend.rs:420-426): a band maximum on the band edge, or
        // a score under 85% of perfect, sends the pair to the full matrix.
        const bool escalate =
            active && a.mode == ESP_MODE_HINT && (touched || best * 20 < (m * P.match) * 17);
        if (escalate) {
            a.esc_list[atomicAdd(&a.esc_range[1], 1u)] = p;
            if (a.aux) a.aux[p] = 1;
        }
        if (active && !escalate) {
            const uint64_t c0 = a.cig_off[p];
            const uint32_t cap = (uint32_t)(a.cig_off[p + 1] - c0);
            uint32_t *cslot = a.cigar + c0;
            TbResult tb = traceback_band<W>(pv, P, codes, lane, tr, cb, best, bi, bj, cslot, cap);
            esp_ext e;
            uint32_t nops;
            finish_pair(pv, P, best, tb, cslot, cap, &e, &nops);
            a.ext[p] = e;
            a.cig_n[p] = nops;
            if (a.touch) {
                esp_touch t;
                t.touched = touched;
                t.row = trow;
                t.left = tleft;
                t.clamped = tclamped;
                a.touch[p] = t;
            }
            if (a.aux) a.aux[p] = 0;
        }
    }
}

template <int HALF>
hipError_t launch_half(const GotohArgs &a, int blocks, size_t lds, hipStream_t s) {
    hipLaunchKernelGGL((k_gotoh_band<HALF>), dim3(blocks), dim3(BAND_BLOCK), lds, s, a);
    return hipGetLastError();
}

template <int HALF>
int occupancy_half(size_t lds) {
    int nb = 0;
    hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_gotoh_band<HALF>, BAND_BLOCK, lds);
    return nb > 0 ? nb : 1;
}

}  // namespace

bool band_half_supported(int half) {
    return half == 16 || half == 30 || half == 32 || half == 48 || half == 64;
}

static size_t band_lds(int half) { return (size_t)band_span(half) * 64; }

hipError_t launch_gotoh_band(const GotohArgs &a, int half, int blocks, hipStream_t s) {
    const size_t lds = band_lds(half);
    switch (half) {
    case 16: return launch_half<16>(a, blocks, lds, s);
    case 30: return launch_half<30>(a, blocks, lds, s);
    case 32: return launch_half<32>(a, blocks, lds, s);
    case 48: return launch_half<48>(a, blocks, lds, s);
    case 64: return launch_half<64>(a, blocks, lds, s);
    default: return hipErrorInvalidValue;
    }
}

int gotoh_band_blocks_per_cu(int half) {
    const size_t lds = band_lds(half);
    switch (half) {
    case 16: return occupancy_half<16>(lds);
    case 30: return occupancy_half<30>(lds);
    case 32: return occupancy_half<32>(lds);
    case 48: return occupancy_half<48>(lds);
    default: return occupancy_half<64>(lds);
    }
}

}  // namespace espk
