// esp_wave.hip -- one pair per wavefront: the plain single-pair Gotoh path.
//
// The 64 lanes of a wave share one read-window pair.  Lane l owns the rows
// i0 + l of a block of 64 rows and the block is swept as a wavefront: at
// step t lane l computes cell (i0 + l, J0 + t - l), one column behind the
// lane above it.  The values a cell needs from the row above, H(i-1, j) and
// E(i-1, j), are what the lane above produced one step earlier, so every
// step moves H and E one lane down (a DPP row shift plus the three lanes
// that start a DPP row); the diagonal H(i-1, j-1) is the value that arrived
// the step before, and H(i, j-1), F(i, j-1) are the lane's own last outputs.
// The first lane of a block reads the row above from an LDS copy of the
// previous block's last row, which that block's last lane wrote as it went.
//
// Every loop bound is the same on all 64 lanes (block count, step count),
// so the lane shift always runs with the whole wave active and every lane
// reaches every exit.  A lane whose cell lies outside its row's band window
// (or past the read) outputs the reference's margin values H = 0,
// E = F = NEG, which is exactly what the reference reads out of window
// (extend.rs:718-759), so banded, clamped and full-matrix calls share one
// sweep.  Per row the lane keeps the first strict maximum and the H values at
// its window's two edges; per block the wave reduces them to the first
// row-major maximum (extend.rs:789-792) and the first touching row
// (extend.rs:803-821).  The 4-bit trace (hdir, E-extend, F-extend) goes to
// LDS when it fits and to the wave's global scratch otherwise; one lane then
// walks it (extend.rs:884-930) and finishes the CIGAR as the strip kernel
// does (extend.rs:932-1059).
#include "esp_kernels.hpp"

using namespace espd;

namespace espk {

namespace {

// Moves x one lane down the wave: lane l returns lane l-1's x, and lane 0
// returns first (a value of its own).  Two DPP moves: row_bcast:15 (rows 1-3)
// puts the last lane of each 16-lane row into every lane of the next row,
// then row_shr:1 (bound_ctrl off, so the first lane of each row keeps that
// value) moves every other lane's neighbour in; lane 0 is set by one select.
// Called with all 64 lanes active, so no source lane is disabled and no
// branch splits the wave.
__device__ __forceinline__ int32_t lane_down1(int32_t x, int32_t first, bool lane0) {
    const int32_t b = __builtin_amdgcn_update_dpp(0, x, 0x142, 0xE, 0xF, false);
    const int32_t r = __builtin_amdgcn_update_dpp(b, x, 0x111, 0xF, 0xF, false);
    return lane0 ? first : r;
}

__device__ __forceinline__ int32_t wave_max(int32_t v) {
#pragma unroll
    for (int o = 32; o; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
    return v;
}

struct WaveFill {
    int32_t best, bi, bj;
    int32_t touched, trow, tleft, tclamped;
};

// Trace words per row for a pair's band: the widest row window.
__device__ __forceinline__ uint32_t wave_wpr(int32_t n, int32_t half) {
    const int64_t w = min((int64_t)2 * half + 1, (int64_t)n);
    return (uint32_t)((w + 7) / 8);
}

// Fills one pair.  win: the window's codes in LDS; bh/be: LDS copy of the
// row above the current block (columns 1..n); tr: trace, wpr words per row,
// in LDS (TRL) or in the wave's global scratch.  Each step's LDS reads (the
// lane's window code, lane 0's row above) are issued one step ahead, so no
// step waits on LDS.
template <bool EDIT, bool TRL>
__device__ __forceinline__ WaveFill wave_fill(const PairView &pv, const GParams P, bool wrap16, const uint8_t *win,
                              int32_t *bh, int32_t *be, uint32_t *tr, uint32_t wpr) {
    const int32_t lane = (int32_t)(threadIdx.x & 63u);
    const bool lane0 = lane == 0, lane63 = lane == 63;
    const int32_t m = pv.m, n = pv.n, off = pv.off, half = pv.half;
    WaveFill r;
    r.best = 0;
    r.bi = 0;
    r.bj = 0;
    r.touched = 0;
    r.trow = 0;
    r.tleft = 0;
    r.tclamped = 0;
    for (int32_t i0 = 1; i0 <= m; i0 += 64) {
        const int32_t i1 = min(i0 + 63, m);
        const int32_t i = i0 + lane;
        const bool rowok = i <= m;
        int32_t lo, hi, lo0, hi0, lo1, hi1, plo = 1, phi = 0;
        row_window(off, half, n, rowok ? i : m, lo, hi);
        row_window(off, half, n, i0, lo0, hi0);
        row_window(off, half, n, i1, lo1, hi1);
        if (i0 > 1) row_window(off, half, n, i0 - 1, plo, phi);
        if (!rowok) {
            lo = 1;
            hi = 0;  // no valid cell
        }
        const int32_t J0 = lo0;
        const int32_t T = hi1 - lo0 + (i1 - i0) + 1;
        const int32_t rc = rowok ? read_code(pv.read[(rowok ? i : 1) - 1]) : 4;
        const int32_t partner = EDIT ? exempt_partner(rc, P.edit) : 6;
        // outputs of the previous step: cell (i, J0 + t - 1 - lane), a margin
        // before the first step
        int32_t oH = 0, oE = NEG, oF = NEG;
        // H(i-1, j-1) for the first step: lane 0 reads it from the row above;
        // every other lane's first cell lies left of its window
        int32_t dH = 0;
        if (lane0 && i0 > 1 && plo <= J0 - 1 && J0 - 1 <= phi) dH = bh[max(J0 - 1, 0)];
        int32_t rmax = -1, rarg = 0, hl = -1, hr = -1;
        uint32_t word = 0;
        uint32_t *trow = tr + (size_t)(rowok ? i - 1 : 0) * wpr;
        // step 0's loads: the lane's window code at column J0 - lane, lane
        // 0's row above at column J0 (every index read is in range)
        auto wcode = [&](int32_t j) -> int32_t { return (j >= 1 && j <= n) ? (int32_t)win[(j >= 1 && j <= n) ? j - 1 : 0] : 7; };
        auto above = [&](int32_t j, const int32_t *v, int32_t dflt) -> int32_t {
            return (i0 > 1 && plo <= j && j <= phi) ? v[(i0 > 1 && plo <= j && j <= phi) ? j : 0] : dflt;
        };
        int32_t fcn = wcode(J0 - lane);
        int32_t bhn = above(J0, bh, 0), ben = above(J0, be, NEG);
        for (int32_t t = 0; t < T; t++) {
            const int32_t j = J0 + t - lane;
            const int32_t fc = fcn, ah = bhn, ae = ben;
            fcn = wcode(j + 1);
            bhn = above(J0 + t + 1, bh, 0);
            ben = above(J0 + t + 1, be, NEG);
            const int32_t uH = lane_down1(oH, ah, lane0), uE = lane_down1(oE, ae, lane0);
            const bool valid = j >= lo && j <= hi;
            const int32_t eo = uH + P.oe, ee = uE + P.ext;
            const int32_t E = max(eo, ee);
            const uint32_t edir = (uint32_t)(eo - ee) >> 31;  // ee > eo
            const int32_t fo = oH + P.oe, fe = oF + P.ext;
            const int32_t F = max(fo, fe);
            const uint32_t fdir = (uint32_t)(fo - fe) >> 31;  // fe > fo
            int32_t sc = rc == fc ? P.match : P.mismatch;
            if (EDIT) sc = fc == partner ? 0 : sc;
            const int32_t D0 = dH + sc;
            const int32_t D = wrap16 && D0 > 32767 ? NEG : D0;
            const int32_t H = max(max(D, E), max(F, 0));
            // the cell's nibble from sign bits (H >= D, E and H >= 0):
            // HD_DIAG when D == H, else HD_E when E == H, else HD_F; HD_START
            // when H == 0; then the E and F extension bits
            const uint32_t a = (uint32_t)(D - H) >> 31, b = (uint32_t)(E - H) >> 31;
            const uint32_t z = (uint32_t)(H - 1) >> 31;
            const uint32_t hd = (a + (a & b)) | (z * 3u);
            const uint32_t nib = hd | (edir << 2) | (fdir << 3);
            dH = uH;
            oH = valid ? H : 0;
            oE = valid ? E : NEG;
            oF = valid ? F : NEG;
            if (valid) {
                const int32_t k = j - lo;
                word |= nib << (4 * (k & 7));
                if ((k & 7) == 7 || j == hi) {
                    trow[k >> 3] = word;
                    word = 0;
                }
                if (H > rmax) {
                    rmax = H;
                    rarg = j;
                }
                if (j == lo) hl = H;
                if (j == hi) hr = H;
                // the block's last row is the next block's row above
                if (lane63) {
                    bh[j] = H;
                    be[j] = E;
                }
            }
        }
        // first row-major strict maximum and first touching row of the block
        const int32_t vmax = wave_max(rowok ? rmax : -1);
        const uint64_t atmax = __ballot(rowok && rmax == vmax);
        if (vmax > r.best) {
            const int32_t f = __ffsll((unsigned long long)atmax) - 1;
            r.best = vmax;
            r.bi = i0 + f;
            r.bj = __shfl(rarg, f, 64);
        }
        if (!r.touched) {
            const uint64_t L = __ballot(rowok && hl == rmax);
            const uint64_t R = __ballot(rowok && hi < n && hr == rmax);
            const uint64_t C = L | R;
            if (C) {
                const int32_t f = __ffsll((unsigned long long)C) - 1;
                int32_t flo, fhi;
                row_window(off, half, n, i0 + f, flo, fhi);
                r.touched = 1;
                r.trow = i0 + f;
                r.tleft = (int32_t)((L >> f) & 1ull);
                r.tclamped = r.tleft ? (flo <= 1) : 0;
            }
        }
        // the block's LDS writes (row above, trace) are complete for everyone
        __syncthreads();
    }
    return r;
}

// State-machine traceback (extend.rs:884-930) over the row-window trace.
__device__ TbResult wave_traceback(const PairView &pv, const uint32_t *tr, uint32_t wpr, int32_t bi, int32_t bj,
                                   uint32_t *slot, uint32_t cap) {
    int32_t i = bi, j = bj;
    uint32_t state = 0;  // 0 H, 1 E, 2 F
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
        int32_t lo, hi;
        row_window(pv.off, pv.half, pv.n, i, lo, hi);
        if (j < lo || j > hi) break;
        const int32_t k = j - lo;
        const uint32_t nib = (tr[(size_t)(i - 1) * wpr + (k >> 3)] >> (4 * (k & 7))) & 0xFu;
        if (state == 0) {
            const uint32_t hd = nib & 3u;
            if (hd == HD_DIAG) {
                emit(OP_M);
                nm++;
                i--;
                j--;
            } else if (hd == HD_E) {
                emit(OP_I);
                ni++;
                state = (nib >> 2) & 1u ? 1u : 0u;
                i--;
            } else if (hd == HD_F) {
                emit(OP_D);
                nd++;
                state = (nib >> 3) & 1u ? 2u : 0u;
                j--;
            } else {
                break;
            }
        } else if (state == 1) {
            emit(OP_I);
            ni++;
            if (((nib >> 2) & 1u) == 0) state = 0;
            i--;
        } else {
            emit(OP_D);
            nd++;
            if (((nib >> 3) & 1u) == 0) state = 0;
            j--;
        }
    }
    if (cur_len) slot[cap - 1 - nrev++] = (cur_len << 4) | cur_op;
    TbResult t;
    t.read_start = (uint32_t)max(i, 0);
    t.ref_start = (uint32_t)max(j, 0);
    t.nm = nm;
    t.ni = ni;
    t.nd = nd;
    t.nrev = nrev;
    return t;
}

// True when some byte of s[0, n) is not one of A, C, G, T, N (the wave
// reads it 64 bytes at a time).
__device__ bool wave_impure(const uint8_t *s, int32_t n) {
    bool bad = false;
    for (int32_t k = (int32_t)(threadIdx.x & 63u); k < n; k += 64) {
        const uint32_t b = s[k];
        bad = bad || !(b == 'A' || b == 'C' || b == 'G' || b == 'T' || b == 'N');
    }
    return __ballot(bad) != 0ull;
}

// One pair: run_dispatch with Some((offset, 1, 1, half)) in ESP_MODE_BAND,
// extend (the full matrix) in ESP_MODE_FULL, extend_hint (extend.rs:392-428:
// the (hint, 30) band, the full matrix when the band attempt touched its
// edge or scored below 85% of perfect) in ESP_MODE_HINT.
template <bool EDIT>
__device__ void wave_one(const GotohArgs &a, uint32_t p, uint32_t mu, uint32_t nu, uint8_t *s_win, int32_t *s_bh,
                         int32_t *s_be, uint32_t *s_tr, uint32_t tr_lds, uint32_t *g_tr) {
    const uint32_t lane = threadIdx.x & 63u;
    const int32_t m = (int32_t)mu, n = (int32_t)nu;
    PairView pv;
    pv.read = a.reads + a.read_off[p];
    pv.win = a.wins + a.win_off[p];
    pv.m = m;
    pv.n = n;
    const uint64_t c0 = a.cig_off[p];
    const uint32_t cap = (uint32_t)(a.cig_off[p + 1] - c0);
    uint32_t *slot = a.cigar + c0;
    if (m == 0 || n == 0) {
        // empty_extension (extend.rs:430-442)
        if (lane == 0) {
            esp_ext e;
            e.read_start = e.read_end = e.ref_start = 0;
            e.score = 0;
            uint32_t k = 0;
            if (m > 0) slot[k++] = ((uint32_t)m << 4) | OP_S;
            a.ext[p] = e;
            a.cig_n[p] = k;
            if (a.touch) a.touch[p] = esp_touch{0, 0, 0, 0};
            if (a.aux) a.aux[p] = 2;
        }
        return;
    }
    int32_t path = 0;
    bool band;
    if (a.mode == ESP_MODE_FULL) {
        band = false;
    } else if (a.mode == ESP_MODE_HINT) {
        band = 30 < n;  // INITIAL_BAND >= n: full matrix directly
        path = band ? 0 : 3;
    } else {
        band = true;
    }
    pv.off = band ? clamp_off(a.offset[p], m, n) : 0;
    pv.half = band ? clamp_half(a.mode == ESP_MODE_HINT ? 30 : a.half[p], n) : n;
    // i16 route of pure pairs whose values can pass 32767 (see i16_wraps)
    const int64_t smax = max(max(a.P.match, a.P.mismatch), 0);
    const bool wrap16 = (int64_t)m * smax > 32767 && !wave_impure(pv.read, m) && !wave_impure(pv.win, n);
    for (int32_t k = (int32_t)lane; k < n; k += 64) s_win[k] = (uint8_t)win_code(pv.win[k]);
    __syncthreads();
    // the trace in LDS when it fits (each instance of the fill then stores it
    // with LDS or global instructions, never flat ones)
    auto fill = [&](uint32_t w) -> WaveFill {
        if ((uint64_t)m * w <= tr_lds) return wave_fill<EDIT, true>(pv, a.P, wrap16, s_win, s_bh, s_be, s_tr, w);
        return wave_fill<EDIT, false>(pv, a.P, wrap16, s_win, s_bh, s_be, g_tr, w);
    };
    uint32_t wpr = wave_wpr(n, pv.half);
    uint32_t *tr = (uint64_t)m * wpr <= tr_lds ? s_tr : g_tr;
    WaveFill fr = fill(wpr);
    if (a.mode == ESP_MODE_HINT && band) {
        // extend_hint's escalation test in the reference's wrapping i32
        const int32_t lhs = (int32_t)((uint32_t)fr.best * 20u);
        const int32_t rhs = (int32_t)((uint32_t)m * (uint32_t)a.P.match * 17u);
        if (fr.touched || lhs < rhs) {
            pv.off = 0;
            pv.half = n;
            wpr = wave_wpr(n, n);
            tr = (uint64_t)m * wpr <= tr_lds ? s_tr : g_tr;
            fr = fill(wpr);
            path = 1;
        }
    }
    if (lane == 0) {
        const TbResult tb = wave_traceback(pv, tr, wpr, fr.bi, fr.bj, slot, cap);
        esp_ext e;
        uint32_t nops;
        finish_pair(pv, a.P, fr.best, tb, slot, cap, &e, &nops);
        a.ext[p] = e;
        a.cig_n[p] = nops;
        if (a.touch) {
            esp_touch t;
            t.touched = fr.touched;
            t.row = fr.trow;
            t.left = fr.tleft;
            t.clamped = fr.tclamped;
            a.touch[p] = t;
        }
        if (a.aux) a.aux[p] = path;
    }
}

// One block of 64 threads (one wave) per pair, pair = blockIdx.x, and no
// work queue: a persistent loop in which lane 0 alone pops the queue let the
// compiled code run lanes 1-63 back into the loop without lane 0, where they
// re-tested the same pair forever.  Pairs longer than the kernel takes
// (window over WAVE_NMAX, or a trace larger than the block's scratch) are
// left alone; the host sends only pairs that fit, and gives each block
// maxw words of global trace scratch.
template <bool EDIT>
__global__ __launch_bounds__(64) void k_gotoh_wave(GotohArgs a) {
    __shared__ uint8_t s_win[WAVE_NMAX];
    __shared__ int32_t s_bh[WAVE_NMAX + 1];
    __shared__ int32_t s_be[WAVE_NMAX + 1];
    __shared__ uint32_t s_tr[WAVE_TR_LDS];
    const uint32_t p = blockIdx.x;
    if (p >= a.npairs) return;
    const uint32_t m = __builtin_amdgcn_readfirstlane(a.read_len[p]);
    const uint32_t n = __builtin_amdgcn_readfirstlane(a.win_len[p]);
    const uint64_t need = wave_trace_words(m, n, a.mode, a.mode == ESP_MODE_BAND ? a.half[p] : 0);
    if (n > (uint32_t)WAVE_NMAX || (need > WAVE_TR_LDS && need > a.maxw)) return;
    wave_one<EDIT>(a, p, m, n, s_win, s_bh, s_be, s_tr, WAVE_TR_LDS, a.trace + (size_t)blockIdx.x * a.maxw);
}

// extend_hint's full-matrix reruns of the pairs the pruned pass could not
// prove (the list order[range[0] ..], its count at range[1] on the device),
// one wave per pair: block b takes entries b, b + gridDim.x, ... below the
// count.  There is no work queue and every loop bound is uniform (the count
// and each pair's lengths are read once into scalars).  Entries from
// list_cap on, pairs whose window is longer than WAVE_NMAX and pairs whose
// trace fits neither LDS nor the block's maxw words are appended to fb_list
// (the unpruned lane kernel's list) through fb_range[1].  With a.census set: census[0] += m * n, census[1] +=
// the lane-cells the steps issue.
// LDS is sized per launch (wave_list_lds): windows up to a.wave_lds_n bases
// and a.wave_lds_tr trace words, so a batch of short windows keeps many
// waves per CU.
template <bool EDIT>
__global__ __launch_bounds__(64) void k_gotoh_wave_list(GotohArgs a) {
    extern __shared__ uint8_t wlds[];
    const uint32_t ln = a.wave_lds_n;  // a multiple of 16
    int32_t *s_bh = (int32_t *)wlds;
    int32_t *s_be = s_bh + (ln + 1);
    uint32_t *s_tr = (uint32_t *)(s_be + (ln + 1));
    uint8_t *s_win = (uint8_t *)(s_tr + a.wave_lds_tr);
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t first = __builtin_amdgcn_readfirstlane(a.range[0]);
    const uint32_t count = __builtin_amdgcn_readfirstlane(a.range[1]);
    for (uint32_t u = blockIdx.x; u < count; u += gridDim.x) {
        const uint32_t p = __builtin_amdgcn_readfirstlane(a.order[first + u]);
        const uint32_t m = __builtin_amdgcn_readfirstlane(a.read_len[p]);
        const uint32_t n = __builtin_amdgcn_readfirstlane(a.win_len[p]);
        const uint64_t need = wave_trace_words(m, n, ESP_MODE_FULL, 0);
        if (u >= a.list_cap || n > ln || (need > a.wave_lds_tr && need > a.maxw)) {
            if (lane == 0) a.fb_list[atomicAdd(&a.fb_range[1], 1u)] = p;
            continue;
        }
        wave_one<EDIT>(a, p, m, n, s_win, s_bh, s_be, s_tr, a.wave_lds_tr, a.trace + (size_t)blockIdx.x * a.maxw);
        if (a.census && lane == 0) {
            const uint64_t blocks = (m + 63) / 64;
            atomicAdd(&a.census[0], (unsigned long long)m * n);
            atomicAdd(&a.census[1], 64ull * (blocks * n + (m - blocks)));
            if (a.census_first) atomicAdd(&a.census[2], (unsigned long long)m * n);
        }
        __syncthreads();  // the next pair reuses the LDS
    }
}

}  // namespace

size_t wave_list_lds(uint32_t ln, uint32_t tr) { return (size_t)8 * (ln + 1) + (size_t)4 * tr + ln; }

hipError_t launch_gotoh_wave_list(const GotohArgs &a, int blocks, hipStream_t s) {
    const size_t lds = wave_list_lds(a.wave_lds_n, a.wave_lds_tr);
    if (a.P.edit)
        hipLaunchKernelGGL(k_gotoh_wave_list<true>, dim3(blocks), dim3(64), lds, s, a);
    else
        hipLaunchKernelGGL(k_gotoh_wave_list<false>, dim3(blocks), dim3(64), lds, s, a);
    return hipGetLastError();
}

int gotoh_wave_blocks_per_cu(uint32_t ln, uint32_t tr) {
    int nb = 0;
    const hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_gotoh_wave_list<true>, 64,
                                                                      wave_list_lds(ln, tr));
    return (e == hipSuccess && nb > 0) ? nb : 1;
}

hipError_t launch_gotoh_wave(const GotohArgs &a, int blocks, hipStream_t s) {
    if (a.P.edit)
        hipLaunchKernelGGL(k_gotoh_wave<true>, dim3(blocks), dim3(64), 0, s, a);
    else
        hipLaunchKernelGGL(k_gotoh_wave<false>, dim3(blocks), dim3(64), 0, s, a);
    return hipGetLastError();
}

}  // namespace espk
