// esp_kernels.hip -- hand-written gfx950 kernels of the alignment core.
//
// Gotoh (general form): one lane owns one read-window pair and sweeps the
// DP in vertical strips of STRIP_T columns.  The strip's previous-row H/E and
// window codes live in VGPRs (static indices after unrolling), so a cell
// costs only VALU work; the strip-to-strip hand-off (H and F of the strip's
// last column per row) and the per-row maximum go through lane-interleaved
// global scratch, which makes every such access a coalesced 256-byte wave
// transaction.  Cells outside a row's band window take the reference's
// margin values (H = 0, E = F = NEG), which is exactly what the reference
// reads out of window (extend.rs:718-759), so banded, clamped and
// full-matrix calls share one code path.
//
// The grid is persistent: each wave pops 64 pairs from a device work queue
// until it is empty, so every wave reaches the exit.
#include "esp_kernels.hpp"

#include <algorithm>

using namespace espd;

namespace espk {

namespace {

struct FillResult {
    int32_t best, bi, bj;
    int32_t touched, trow, tleft, tclamped;
};

// Lane-interleaved scratch accessors.
struct Slot {
    int32_t *rows;
    uint32_t *trace;
    uint32_t maxm, nl, slot;
    __device__ __forceinline__ int32_t *row(int arr, int32_t i) const {
        return rows + ((size_t)arr * maxm + (size_t)(i - 1)) * nl + slot;
    }
    __device__ __forceinline__ uint32_t *tword(size_t w) const { return trace + w * nl + slot; }
};

enum { R_BNDH = 0, R_BNDF = 1, R_MAX = 2, R_HL = 3, R_HR = 4 };

// Fills the DP of one pair strip by strip.  Returns the best cell (first
// row-major strict maximum, extend.rs:789-792, :823-827) and the first touch
// (extend.rs:803-821).  WRAP16 reproduces the reference's i16 accumulator
// (extend.rs:473-501) for pure pairs whose values can pass i16's range: a
// diagonal candidate above 32767 wraps there to a negative value (or to
// the -30000 sentinel) and never wins, so it is dropped; inside the scoring
// envelope no other add of the i16 route leaves the i16 range.
template <int T, bool EDIT, bool WRAP16>
__device__ FillResult fill_strips(const PairView &pv, const GParams &P, const Slot &S) {
    const int32_t m = pv.m, n = pv.n, off = pv.off, half = pv.half;
    for (int32_t i = 1; i <= m; i++) {
        *S.row(R_BNDH, i) = 0;
        *S.row(R_BNDF, i) = NEG;
        *S.row(R_MAX, i) = -1;
    }
    int32_t bestv = 0, besti = 0, bestj = 0;
    const int32_t nstrips = (n + T - 1) / T;
    int32_t ia = 1;
    for (int32_t s = 0; s < nstrips; s++) {
        const int32_t J0 = 1 + s * T, J1 = J0 + T - 1;
        int32_t lo, hi;
        // First row reaching this strip (hi non-decreasing in i).
        while (ia <= m) {
            row_window(off, half, n, ia, lo, hi);
            if (hi >= J0) break;
            ia++;
        }
        if (ia > m) break;
        int32_t fc[T], Hr[T], Er[T];
#pragma unroll
        for (int c = 0; c < T; c++) {
            int32_t j = J0 + c;
            fc[c] = j <= n ? win_code(pv.win[j - 1]) : 7;
            Hr[c] = 0;
            Er[c] = NEG;
        }
        // H(ia-1, J0-1) inside row ia-1's window, else the margin 0.
        int32_t Hd = 0;
        if (ia > 1) {
            int32_t plo, phi;
            row_window(off, half, n, ia - 1, plo, phi);
            if (plo <= J0 - 1 && J0 - 1 <= phi) Hd = *S.row(R_BNDH, ia - 1);
        }
        for (int32_t i = ia; i <= m; i++) {
            row_window(off, half, n, i, lo, hi);
            if (lo > J1) break;
            const int32_t rc = read_code(pv.read[i - 1]);
            const int32_t partner = EDIT ? exempt_partner(rc, P.edit) : 6;
            const bool lin = lo <= J0 - 1 && J0 - 1 <= hi;
            const int32_t Hl = lin ? *S.row(R_BNDH, i) : 0;
            const int32_t Fl = lin ? *S.row(R_BNDF, i) : NEG;
            const int32_t clo = lo - J0, chi = hi - J0;
            int32_t hdiag = Hd, hleft = Hl, fleft = Fl;
            int32_t rmax = -1, rarg = 0, hlo = -1, hhi = -1;
            uint32_t word = 0;
            const size_t wbase = ((size_t)s * m + (size_t)(i - 1)) * (T / 8);
#pragma unroll
            for (int c = 0; c < T; c++) {
                const int32_t hup = Hr[c], eup = Er[c];
                const int32_t eo = hup + P.oe, ee = eup + P.ext;
                const int32_t E = max(eo, ee);
                const uint32_t edir = ee > eo;
                const int32_t fo = hleft + P.oe, fe = fleft + P.ext;
                const int32_t F = max(fo, fe);
                const uint32_t fdir = fe > fo;
                int32_t sc = rc == fc[c] ? P.match : P.mismatch;
                if (EDIT) sc = fc[c] == partner ? 0 : sc;
                const int32_t D0 = hdiag + sc;
                const int32_t D = WRAP16 && D0 > 32767 ? NEG : D0;
                const int32_t H = max(max(D, E), max(F, 0));
                const uint32_t hd = H == 0 ? HD_START : (D == H ? HD_DIAG : (E == H ? HD_E : HD_F));
                const bool valid = c >= clo && c <= chi;
                const int32_t Hv = valid ? H : 0;
                const int32_t Ev = valid ? E : NEG;
                const int32_t Fv = valid ? F : NEG;
                hdiag = hup;
                Hr[c] = Hv;
                Er[c] = Ev;
                hleft = Hv;
                fleft = Fv;
                word |= (hd | (edir << 2) | (fdir << 3)) << (4 * (c & 7));
                if ((c & 7) == 7) {
                    *S.tword(wbase + (c >> 3)) = word;
                    word = 0;
                }
                if (Hv > rmax) {
                    rmax = Hv;
                    rarg = c;
                }
                hlo = c == clo ? Hv : hlo;
                hhi = c == chi ? Hv : hhi;
            }
            // Hand-off to the next strip: H and F of column J1 (margins when
            // J1 lies outside the row window).
            *S.row(R_BNDH, i) = hleft;
            *S.row(R_BNDF, i) = fleft;
            Hd = Hl;
            int32_t *rm = S.row(R_MAX, i);
            if (rmax > *rm) *rm = rmax;
            if (clo >= 0 && clo < T) *S.row(R_HL, i) = hlo;
            if (chi >= 0 && chi < T) *S.row(R_HR, i) = hhi;
            if (rmax > bestv || (rmax == bestv && rmax > 0 && i < besti)) {
                bestv = rmax;
                besti = i;
                bestj = J0 + rarg;
            }
        }
    }
    FillResult r;
    r.best = bestv;
    r.bi = besti;
    r.bj = bestj;
    r.touched = 0;
    r.trow = 0;
    r.tleft = 0;
    r.tclamped = 0;
    for (int32_t i = 1; i <= m; i++) {
        int32_t lo, hi;
        row_window(off, half, n, i, lo, hi);
        const int32_t rm = *S.row(R_MAX, i);
        if (*S.row(R_HL, i) == rm) {
            r.touched = 1;
            r.trow = i;
            r.tleft = 1;
            r.tclamped = lo <= 1;
            break;
        }
        if (hi < n && *S.row(R_HR, i) == rm) {
            r.touched = 1;
            r.trow = i;
            r.tleft = 0;
            r.tclamped = 0;
            break;
        }
    }
    return r;
}

// State-machine traceback (extend.rs:884-930) over the strip trace layout.
// Runs are written in walk order to slot[cap-1], slot[cap-2], ...
template <int T>
__device__ TbResult traceback_strips(const PairView &pv, const Slot &S, int32_t bi, int32_t bj,
                                     uint32_t *slot, uint32_t cap) {
    const int32_t m = pv.m, n = pv.n;
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
        row_window(pv.off, pv.half, n, i, lo, hi);
        if (j < lo || j > hi) break;
        const int32_t s = (j - 1) / T, c = (j - 1) % T;
        const size_t w = ((size_t)s * m + (size_t)(i - 1)) * (T / 8) + (c >> 3);
        const uint32_t nib = (*S.tword(w) >> (4 * (c & 7))) & 0xFu;
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

// run_dispatch (extend.rs:293-344) sends pairs whose every byte is one of
// A, C, G, T, N (upper case) to the i16 kernels and every other pair to the
// i32 one.
__device__ bool pure_bytes(const uint8_t *s, uint32_t n) {
    bool ok = true;
    for (uint32_t k = 0; k < n && ok; k++) {
        const uint32_t b = s[k];
        ok = b == 'A' || b == 'C' || b == 'G' || b == 'T' || b == 'N';
    }
    return ok;
}

// True when the reference's i16 route can leave the i16 range on this pair:
// it is pure and a diagonal candidate can exceed 32767 (a path gains at most
// max(match, mismatch, 0) a row, so m times that bounds every H).
__device__ bool i16_wraps(const PairView &pv, const GParams &P) {
    const int64_t s = max(max(P.match, P.mismatch), 0);
    if ((int64_t)pv.m * s <= 32767) return false;
    return pure_bytes(pv.read, (uint32_t)pv.m) && pure_bytes(pv.win, (uint32_t)pv.n);
}

template <int T, bool EDIT>
__device__ FillResult fill_route(const PairView &pv, const GParams &P, const Slot &S, bool wrap16) {
    if (wrap16) return fill_strips<T, EDIT, true>(pv, P, S);
    return fill_strips<T, EDIT, false>(pv, P, S);
}

// extend_hint's escalation test `score * 20 < perfect * 17` (extend.rs:424)
// in the reference's wrapping i32 arithmetic.
__device__ __forceinline__ bool hint_escalates(int32_t touched, int32_t best, int32_t m, int32_t match) {
    const int32_t lhs = (int32_t)((uint32_t)best * 20u);
    const int32_t rhs = (int32_t)((uint32_t)m * (uint32_t)match * 17u);
    return touched || lhs < rhs;
}

template <int T, bool EDIT>
__device__ void gotoh_one(const GotohArgs &a, uint32_t p, const Slot &S) {
    const uint32_t m = a.read_len[p], n = a.win_len[p];
    PairView pv;
    pv.read = a.reads + a.read_off[p];
    pv.win = a.wins + a.win_off[p];
    pv.m = (int32_t)m;
    pv.n = (int32_t)n;
    const uint64_t c0 = a.cig_off[p];
    const uint32_t cap = (uint32_t)(a.cig_off[p + 1] - c0);
    uint32_t *slot = a.cigar + c0;
    esp_touch tr;
    tr.touched = 0;
    tr.row = 0;
    tr.left = 0;
    tr.clamped = 0;
    int32_t path = 0;
    if (m == 0 || n == 0) {
        // empty_extension (extend.rs:430-442)
        esp_ext e;
        e.read_start = e.read_end = e.ref_start = 0;
        e.score = 0;
        uint32_t k = 0;
        if (m > 0) slot[k++] = (m << 4) | OP_S;
        a.ext[p] = e;
        a.cig_n[p] = k;
        if (a.touch) a.touch[p] = tr;
        if (a.aux) a.aux[p] = 2;
        return;
    }
    bool band;
    if (a.mode == ESP_MODE_FULL) {
        band = false;
    } else if (a.mode == ESP_MODE_HINT) {
        band = 30 < (int64_t)n;  // INITIAL_BAND >= n: full matrix directly
        path = band ? 0 : 3;
    } else {
        band = true;
    }
    pv.off = band ? clamp_off(a.offset[p], pv.m, pv.n) : 0;
    pv.half = band ? clamp_half(a.mode == ESP_MODE_HINT ? 30 : a.half[p], pv.n) : pv.n;
    const bool wrap16 = i16_wraps(pv, a.P);
    FillResult fr = fill_route<T, EDIT>(pv, a.P, S, wrap16);
    if (a.mode == ESP_MODE_HINT && band) {
        if (hint_escalates(fr.touched, fr.best, pv.m, a.P.match)) {
            pv.off = 0;
            pv.half = pv.n;
            fr = fill_route<T, EDIT>(pv, a.P, S, wrap16);
            path = 1;
        }
    }
    TbResult tb = traceback_strips<T>(pv, S, fr.bi, fr.bj, slot, cap);
    esp_ext e;
    uint32_t nops;
    finish_pair(pv, a.P, fr.best, tb, slot, cap, &e, &nops);
    a.ext[p] = e;
    a.cig_n[p] = nops;
    if (a.touch) {
        tr.touched = fr.touched;
        tr.row = fr.trow;
        tr.left = fr.tleft;
        tr.clamped = fr.tclamped;
        a.touch[p] = tr;
    }
    if (a.aux) a.aux[p] = path;
}

template <int T, bool EDIT>
__global__ __launch_bounds__(GOTOH_BLOCK) void k_gotoh_strip(GotohArgs a) {
    const uint32_t lane = threadIdx.x & 63u;
    Slot S;
    S.rows = a.rows;
    S.trace = a.trace;
    S.maxm = a.maxm;
    S.nl = a.nl;
    S.slot = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t first = a.range ? a.range[0] : 0u;
    const uint32_t count = a.range ? a.range[1] : a.npairs;
    for (;;) {
        const uint32_t base = wave_pop(a.counter, lane);
        if (base >= count) break;
        const uint32_t idx = base + lane;
        if (idx < count) {
            const uint32_t p = a.order ? a.order[first + idx] : idx;
            // A pair larger than the scratch was sized for is left alone (a
            // captured graph replayed on longer pairs; the planner counts
            // them as bad and the caller's check reports it).
            if (a.read_len[p] <= a.maxm && strip_trace_words(a.read_len[p], a.win_len[p]) <= a.maxw)
                gotoh_one<T, EDIT>(a, p, S);
        }
    }
}

// ----------------------------------------------------------------- planner

constexpr int kPlanHalves[5] = {16, 30, 32, 48, 64};

__device__ __forceinline__ uint32_t plan_bin(const PlanArgs &a, uint32_t p, uint32_t &cls) {
    const uint32_t m = a.read_len[p], n = a.win_len[p];
    cls = PLAN_STRIP;
    // The band kernel takes bands whose every row window lies inside [1, n]
    // (offset >= half and offset + m + half <= n) and, with left_clamp_ok,
    // bands clamped only on the left (0 <= offset < half): there the row
    // centres offset + i stay inside [1, n] and the columns left of the
    // window become 'N' columns that keep H = 0.
    if (a.hint_mode && m >= 1 && n > 30u) {
        // extend_hint attempt 1 is the band (hint, 30) (extend.rs:408-419)
        const int64_t off = a.offset[p];
        if (off >= (a.left_clamp_ok ? 0 : 30) && off + (int64_t)m + 30 <= (int64_t)n &&
            (int64_t)m <= band_max_read(30))
            cls = 1;
    } else if (a.band_ok && m >= 1) {
        const int64_t off = a.offset[p], half = a.half[p];
        int c = -1;
#pragma unroll
        for (int t = 0; t < 5; t++)
            if (half == kPlanHalves[t]) c = t;
        if (c >= 0 && off >= (a.left_clamp_ok ? 0 : half) && off + (int64_t)m + half <= (int64_t)n &&
            (int64_t)m <= band_max_read((int)half))
            cls = (uint32_t)c;
    }
    return cls * PLAN_NBIN + (PLAN_NBIN - 1 - min(m, (uint32_t)PLAN_NBIN - 1));
}

// Wave-aggregated increment of counter[b]: lanes that share a bin add once
// and take consecutive slots; returns this lane's slot.  Up to 4 bins per
// wave are served that way (a batch dominated by one read length puts whole
// waves on one address); lanes of further bins add on their own.  Every lane
// of the wave calls it.
__device__ __forceinline__ uint32_t agg_inc(uint32_t *counter, uint32_t b, bool active) {
    const uint32_t lane = threadIdx.x & 63u;
    uint64_t pending = __ballot(active);
    uint32_t mine = 0;
    for (int round = 0; round < 4 && pending; round++) {
        const int leader = __ffsll((unsigned long long)pending) - 1;
        const uint32_t lb = __shfl(b, leader);
        const uint64_t same = __ballot(active && b == lb) & pending;
        uint32_t old = 0;
        if ((int)lane == leader) old = atomicAdd(&counter[lb], (uint32_t)__popcll(same));
        old = __shfl(old, leader);
        if ((same >> lane) & 1ull) mine = old + (uint32_t)__popcll(same & ((1ull << lane) - 1ull));
        pending &= ~same;
    }
    if ((pending >> lane) & 1ull) mine = atomicAdd(&counter[b], 1u);
    return mine;
}

// Planner blocks: PLAN_BT threads each, every block over one contiguous
// chunk of the pairs.  Global atomics are the planner's cost (they serialize
// on their address), so each block reduces in LDS first and then adds once
// per bin it holds (the histogram) or reserves its share of each bin with one
// add (the scatter): a few hundred global adds per block instead of one per
// pair, and one per statistic per block instead of one per thread.
constexpr int PLAN_BT = 1024;

__device__ __forceinline__ void plan_chunk(uint32_t npairs, uint32_t &p0, uint32_t &p1) {
    const uint64_t per = ((uint64_t)npairs + gridDim.x - 1) / gridDim.x;
    p0 = (uint32_t)min((uint64_t)blockIdx.x * per, (uint64_t)npairs);
    p1 = (uint32_t)min((uint64_t)p0 + per, (uint64_t)npairs);
}

__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v) {
#pragma unroll
    for (int o = 32; o; o >>= 1) v = max(v, (uint32_t)__shfl_xor((int)v, o, 64));
    return v;
}

__device__ __forceinline__ unsigned long long wave_max_u64(unsigned long long v) {
#pragma unroll
    for (int o = 32; o; o >>= 1) {
        const unsigned long long u = __shfl_xor(v, o, 64);
        v = u > v ? u : v;
    }
    return v;
}

__global__ __launch_bounds__(PLAN_BT) void k_plan_hist(PlanArgs a) {
    __shared__ uint32_t h[PLAN_NCLASS * PLAN_NBIN];
    __shared__ uint32_t work, bsmax, bemax, benmax, bsnmax;
    __shared__ unsigned long long bwmax, bewmax;
    for (int t = threadIdx.x; t < PLAN_NCLASS * PLAN_NBIN; t += blockDim.x) h[t] = 0;
    if (threadIdx.x == 0) {
        work = bsmax = bemax = benmax = bsnmax = 0;
        bwmax = bewmax = 0;
    }
    __syncthreads();
    uint32_t smax = 0, emax = 0, enmax = 0, snmax = 0, swork = 0;
    unsigned long long wmax = 0, ewmax = 0;
    uint32_t p0, p1;
    plan_chunk(a.npairs, p0, p1);
    // The loop bound is uniform per wave, so agg_inc sees whole waves.
    for (uint32_t base = p0; base < p1; base += blockDim.x) {
        const uint32_t p = base + threadIdx.x;
        const bool active = p < p1;
        uint32_t cls = PLAN_STRIP, b = 0;
        if (active) {
            b = plan_bin(a, p, cls);
            const uint32_t m = a.read_len[p];
            const uint64_t sw = strip_trace_words(m, a.win_len[p]);
            if (m > a.max_read || a.win_len[p] > a.max_win || (cls == PLAN_STRIP && sw > a.max_strip_w))
                atomicAdd(&a.plan->sum.bad, 1u);
            if (cls == PLAN_STRIP) {
                smax = max(smax, m);
                wmax = max(wmax, (unsigned long long)sw);
                if (m >= 1 && a.win_len[p] >= 1) {
                    swork++;
                    snmax = max(snmax, a.win_len[p]);
                }
            } else if (a.hint_mode) {
                // a band attempt that escalates reruns on the full matrix
                emax = max(emax, m);
                enmax = max(enmax, a.win_len[p]);
                ewmax = max(ewmax, (unsigned long long)strip_trace_words(m, a.win_len[p]));
            }
        }
        (void)agg_inc(h, b, active);
    }
    // block reductions: one LDS update per wave, one global update per block
    smax = wave_max_u32(smax);
    snmax = wave_max_u32(snmax);
    emax = wave_max_u32(emax);
    enmax = wave_max_u32(enmax);
    wmax = wave_max_u64(wmax);
    ewmax = wave_max_u64(ewmax);
    for (int o = 32; o; o >>= 1) swork += (uint32_t)__shfl_xor((int)swork, o, 64);
    if ((threadIdx.x & 63u) == 0) {
        if (swork) atomicAdd(&work, swork);
        atomicMax(&bsmax, smax);
        atomicMax(&bsnmax, snmax);
        atomicMax(&bemax, emax);
        atomicMax(&benmax, enmax);
        atomicMax(&bwmax, wmax);
        atomicMax(&bewmax, ewmax);
    }
    __syncthreads();
    for (int t = threadIdx.x; t < PLAN_NCLASS * PLAN_NBIN; t += blockDim.x)
        if (h[t]) atomicAdd(&a.plan->hist[t], h[t]);
    if (threadIdx.x == 0) {
        if (work) atomicAdd(&a.plan->sum.strip_work, work);
        atomicMax(&a.plan->sum.strip_maxm, bsmax);
        atomicMax(&a.plan->sum.strip_maxn, bsnmax);
        atomicMax(&a.plan->sum.strip_maxw, bwmax);
        if (a.hint_mode) {
            atomicMax(&a.plan->sum.esc_maxm, bemax);
            atomicMax(&a.plan->sum.esc_maxn, benmax);
            atomicMax(&a.plan->sum.esc_maxw, bewmax);
        }
    }
}

// Exclusive scan of the bins (one block), class ranges, cursors.
__global__ __launch_bounds__(1024) void k_plan_scan(PlanArgs a) {
    __shared__ uint32_t part[1024];
    constexpr int NB = PLAN_NCLASS * PLAN_NBIN;
    constexpr int PER = (NB + 1023) / 1024;
    const int t = threadIdx.x;
    uint32_t local[PER];
    uint32_t sum = 0;
#pragma unroll
    for (int k = 0; k < PER; k++) {
        const int b = t * PER + k;
        local[k] = b < NB ? a.plan->hist[b] : 0u;
        sum += local[k];
    }
    part[t] = sum;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        uint32_t v = t >= off ? part[t - off] : 0u;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    uint32_t run = part[t] - sum;
#pragma unroll
    for (int k = 0; k < PER; k++) {
        const int b = t * PER + k;
        if (b < NB) {
            a.plan->cursor[b] = run;
            if (b % PLAN_NBIN == 0) a.plan->sum.range[b / PLAN_NBIN][0] = run;
        }
        run += local[k];
    }
    __syncthreads();
    if (t < PLAN_NCLASS) {
        uint32_t c = 0;
        for (int b = 0; b < PLAN_NBIN; b++) c += a.plan->hist[t * PLAN_NBIN + b];
        a.plan->sum.range[t][1] = c;
    }
    if (t == 0) a.plan->esc_range[0] = a.npairs;  // escalation list follows the order
}

// Each block counts its chunk per bin in LDS, reserves that many slots of
// each bin with one global add, and hands them out with LDS adds.  Pairs of
// one bin share class and read length; their order inside the bin is free.
__global__ __launch_bounds__(PLAN_BT) void k_plan_scatter(PlanArgs a) {
    __shared__ uint32_t lh[PLAN_NCLASS * PLAN_NBIN];
    for (int t = threadIdx.x; t < PLAN_NCLASS * PLAN_NBIN; t += blockDim.x) lh[t] = 0;
    __syncthreads();
    uint32_t p0, p1;
    plan_chunk(a.npairs, p0, p1);
    for (uint32_t base = p0; base < p1; base += blockDim.x) {
        const uint32_t p = base + threadIdx.x;
        const bool active = p < p1;
        uint32_t cls = PLAN_STRIP, b = 0;
        if (active) b = plan_bin(a, p, cls);
        (void)agg_inc(lh, b, active);
    }
    __syncthreads();
    for (int t = threadIdx.x; t < PLAN_NCLASS * PLAN_NBIN; t += blockDim.x)
        if (lh[t]) lh[t] = atomicAdd(&a.plan->cursor[t], lh[t]);
    __syncthreads();
    for (uint32_t base = p0; base < p1; base += blockDim.x) {
        const uint32_t p = base + threadIdx.x;
        const bool active = p < p1;
        uint32_t cls = PLAN_STRIP, b = 0;
        if (active) b = plan_bin(a, p, cls);
        const uint32_t pos = agg_inc(lh, b, active);
        if (active) a.order[pos] = p;
    }
}

__global__ void k_zero_plan(PlanDev *d) {
    uint32_t *w = (uint32_t *)d;
    const size_t n = sizeof(PlanDev) / 4;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        w[i] = 0;
}

__global__ void k_zero_u32(uint32_t *p, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// The longest pattern of at most 256 bases (longer ones take the guard) and
// the longest text of a Myers batch: out[0] and out[1], zeroed beforehand.
// The launch sizes its words from the first and checks the second.
__global__ __launch_bounds__(256) void k_len_max(const uint32_t *__restrict__ rl,
                                                 const uint32_t *__restrict__ wl, size_t n,
                                                 uint32_t *out) {
    uint32_t mm = 0, mn = 0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const uint32_t m = rl[i];
        mm = max(mm, m <= 256u ? m : 0u);
        mn = max(mn, wl[i]);
    }
    for (int o = 32; o > 0; o >>= 1) {
        mm = max(mm, (uint32_t)__shfl_xor((int)mm, o));
        mn = max(mn, (uint32_t)__shfl_xor((int)mn, o));
    }
    // One atomic pair per block: same-address atomics serialize in the L2,
    // so a pair per wave over a full grid cost about 0.1 ms.
    __shared__ uint32_t part[2][4];
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        part[0][wave] = mm;
        part[1][wave] = mn;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        atomicMax(out, max(max(part[0][0], part[0][1]), max(part[0][2], part[0][3])));
        atomicMax(out + 1, max(max(part[1][0], part[1][1]), max(part[1][2], part[1][3])));
    }
}

__global__ __launch_bounds__(256) void k_copy16(uint4 *dst, const uint4 *src, size_t n16) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

__global__ void k_copy1(uint8_t *dst, const uint8_t *src, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

}  // namespace

hipError_t launch_copy(void *dst, const void *src, size_t bytes, hipStream_t s) {
    const size_t n16 = bytes / 16;
    if (n16) hipLaunchKernelGGL(k_copy16, dim3(2048), dim3(256), 0, s, (uint4 *)dst, (const uint4 *)src, n16);
    const size_t tail = bytes - n16 * 16;
    if (tail)
        hipLaunchKernelGGL(k_copy1, dim3(1), dim3(64), 0, s, (uint8_t *)dst + n16 * 16,
                           (const uint8_t *)src + n16 * 16, tail);
    return hipGetLastError();
}

hipError_t launch_gotoh_strip(const GotohArgs &a, int blocks, hipStream_t s) {
    if (a.P.edit) {
        hipLaunchKernelGGL((k_gotoh_strip<STRIP_T, true>), dim3(blocks), dim3(GOTOH_BLOCK), 0, s, a);
    } else {
        hipLaunchKernelGGL((k_gotoh_strip<STRIP_T, false>), dim3(blocks), dim3(GOTOH_BLOCK), 0, s, a);
    }
    return hipGetLastError();
}

hipError_t launch_plan(const PlanArgs &a, int ncu, hipStream_t s) {
    hipLaunchKernelGGL(k_zero_plan, dim3(8), dim3(256), 0, s, a.plan);
    // half a block per CU: few enough blocks that the per-block global adds
    // stay cheap, with 16 waves each to keep the loads in flight
    const int blocks = (int)std::max<uint64_t>(
        1, std::min<uint64_t>((uint64_t)std::max(ncu / 2, 1), ((uint64_t)a.npairs + PLAN_BT - 1) / PLAN_BT));
    hipLaunchKernelGGL(k_plan_hist, dim3(blocks), dim3(PLAN_BT), 0, s, a);
    hipLaunchKernelGGL(k_plan_scan, dim3(1), dim3(1024), 0, s, a);
    hipLaunchKernelGGL(k_plan_scatter, dim3(blocks), dim3(PLAN_BT), 0, s, a);
    return hipGetLastError();
}

int plan_class_half(int c) { return c >= 0 && c < 5 ? kPlanHalves[c] : -1; }

hipError_t launch_len_max(const uint32_t *rl, const uint32_t *wl, size_t n, uint32_t *out, int ncu,
                          hipStream_t s) {
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, s, out, (size_t)2);
    if (n == 0) return hipGetLastError();
    const uint64_t blocks = std::max<uint64_t>(1, std::min<uint64_t>((uint64_t)ncu, (n + 255) / 256));
    hipLaunchKernelGGL(k_len_max, dim3((unsigned)blocks), dim3(256), 0, s, rl, wl, n, out);
    return hipGetLastError();
}

hipError_t launch_zero_u32(uint32_t *p, size_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_zero_u32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n);
    return hipGetLastError();
}

int gotoh_strip_blocks_per_cu(int edit) {
    int nb = 0;
    if (edit)
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_gotoh_strip<STRIP_T, true>, GOTOH_BLOCK, 0);
    else
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_gotoh_strip<STRIP_T, false>, GOTOH_BLOCK, 0);
    return nb > 0 ? nb : 1;
}

}  // namespace espk
