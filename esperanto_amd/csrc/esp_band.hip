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
    while (i > 0 && j > 0) {
        const int32_t k = j - (cb + i);
        if (k < 0 || k >= W) break;
        const int32_t t = k >> 3;
        const int32_t c = min(8, W - 8 * t);
        const uint32_t word = tr[((size_t)(i - 1) * WW + t) * 64];
        const uint32_t nib = (word >> (4 * (c - 1 - (k & 7)))) & 0xFu;
        if (state == 0) {
            if (V == 0) break;  // H == 0: the reference's H_START
            if (!(nib & 8u)) {
                const int32_t rc = read_code(pv.read[i - 1]);
                const int32_t fc = codes[(size_t)(j - cb) * 64 + lane] / 5;
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
    extern __shared__ uint8_t codes[];  // [m + W][64]
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
        // Window codes of columns cb + 1 .. cb + m + W - 1 (interior: inside [1, n]).
        for (int32_t jr = 1; jr <= m + W - 1; jr++)
            codes[(size_t)jr * 64 + lane] = (uint8_t)(5 * win_code(pv.win[cb + jr - 1]));

        int32_t Ur[W], Bx[W];
#pragma unroll
        for (int k = 0; k < W; k++) {
            Ur[k] = oe;  // row 0: H = 0
            Bx[k] = negx;  // row 0: E = NEG
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
                const int32_t rc = read_code(pv.read[i - 1]);
                const uint32_t rowlut = rc == 0 ? a.lut[0] : rc == 1 ? a.lut[1] : rc == 2 ? a.lut[2]
                                      : rc == 3 ? a.lut[3] : a.lut[4];
                const uint8_t *crow = codes + (size_t)i * 64 + lane;
                int32_t Ul = oe, Fx = negx, key = 0, H0 = 0, Hl = 0;
                uint32_t acc = 0;
                uint32_t *trow_p = tr + (size_t)(i - 1) * WW * 64;
#pragma unroll
                for (int k = 0; k < W; k++) {
                    const int32_t Uup = (k + 1 < W) ? Ur[k + 1] : oe;
                    const int32_t Bup = (k + 1 < W) ? Bx[k + 1] : negx;
                    const uint32_t code5 = crow[k * 64];
                    const int32_t sc = __builtin_amdgcn_sbfe((int32_t)rowlut, code5, 5);
                    int32_t D;
                    asm("v_add3_u32 %0, %1, %2, %3" : "=v"(D) : "v"(Ur[k]), "v"(sc), "s"(noe));
                    const int32_t E = max(Uup, Bup);
                    const int32_t F = max(Ul, Fx);
                    const int32_t H = max(max(max(D, E), F), 0);
                    // Sign bits: D < H, E < H, E from extension, F from extension.
                    acc = __builtin_amdgcn_alignbit(acc, (uint32_t)(D - H), 31);
                    acc = __builtin_amdgcn_alignbit(acc, (uint32_t)(E - H), 31);
                    acc = __builtin_amdgcn_alignbit(acc, (uint32_t)(Uup - Bup), 31);
                    acc = __builtin_amdgcn_alignbit(acc, (uint32_t)(Ul - Fx), 31);
                    const int32_t Un = H + oe;
                    Ur[k] = Un;
                    Bx[k] = E + ext;
                    Ul = Un;
                    Fx = F + ext;
                    key = max(key, (H << 8) | (255 - k));
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
        if (active) {
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

static size_t band_lds(const GotohArgs &, int) { return (size_t)BAND_SPAN * 64; }

hipError_t launch_gotoh_band(const GotohArgs &a, int half, int blocks, hipStream_t s) {
    const size_t lds = band_lds(a, half);
    switch (half) {
    case 16: return launch_half<16>(a, blocks, lds, s);
    case 30: return launch_half<30>(a, blocks, lds, s);
    case 32: return launch_half<32>(a, blocks, lds, s);
    case 48: return launch_half<48>(a, blocks, lds, s);
    case 64: return launch_half<64>(a, blocks, lds, s);
    default: return hipErrorInvalidValue;
    }
}

int gotoh_band_blocks_per_cu(int half, uint32_t maxm) {
    GotohArgs a{};
    a.maxm = maxm;
    const size_t lds = band_lds(a, half);
    switch (half) {
    case 16: return occupancy_half<16>(lds);
    case 30: return occupancy_half<30>(lds);
    case 32: return occupancy_half<32>(lds);
    case 48: return occupancy_half<48>(lds);
    default: return occupancy_half<64>(lds);
    }
}

}  // namespace espk
