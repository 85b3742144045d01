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
// Bands clamped on the left (0 <= offset < half, offset + m + half <= n,
// mismatch score <= 0) run here too.  The reference's row window is then
// [max(c - half, 1), c + half]; the band's columns cb + i + k <= 0 read the
// window code of 'N'.  By induction those cells hold H = 0: their diagonal
// is such a cell or row 0, so D = 0 + mismatch <= 0, their E comes from the
// cell above (such a cell, U = open_ext <= 0) and F is floored at 0.  A real
// cell therefore reads exactly the reference's margins through them (H = 0
// left and diagonal; the floored F is 0 either way), the traceback stops at
// column 0 as the reference's loop does, and the row maximum is positive
// only at real cells, so the left touch is "the first maximum lies at or
// before the first real cell" (with a zero row, that cell holds 0 as well).
// Other pairs, and halves with no specialisation, run the general strip
// kernel (esp_kernels.hip).
#include "esp_kernels.hpp"

using namespace espd;

namespace espk {

namespace {

// Bytes p[lo .. lo + 4*NW - 1] as NW words in which byte t is p[lo + t],
// for the bytes up to index hi (later ones are unspecified; with hi < lo
// all are).  Every load is an aligned dword holding at least one byte of
// p[0 .. max(hi, 0)] (the caller guarantees p[0] is readable), so none
// leaves the sequence's pages and no load needs a branch; one round trip.
template <int NW>
__device__ __forceinline__ void load_span(const uint8_t *p, int32_t lo, int32_t hi, uint32_t out[NW]) {
    const int32_t last = max(hi, 0);
    const uintptr_t a0 = (uintptr_t)(p + min(lo, last));
    const uintptr_t d0 = a0 & ~(uintptr_t)3, dl = (uintptr_t)(p + last) & ~(uintptr_t)3;
    const uint32_t sh = (uint32_t)(a0 & 3u);
    uint32_t w[NW + 1];
#pragma unroll
    for (int k = 0; k <= NW; k++) w[k] = *(const uint32_t *)(d0 + 4 * k <= dl ? d0 + 4 * k : dl);
#pragma unroll
    for (int k = 0; k < NW; k++) out[k] = __builtin_amdgcn_alignbyte(w[k + 1], w[k], sh);
}

__device__ __forceinline__ void load_span16(const uint8_t *p, int32_t lo, int32_t hi, uint32_t out[4]) {
    load_span<4>(p, lo, hi, out);
}

// Byte t of the words; t is a compile-time constant where the ring uses it.
template <int NW>
__device__ __forceinline__ uint32_t byte_at(const uint32_t w[NW], int32_t t) {
    if constexpr (NW == 4) {
        const uint64_t lo = (uint64_t)w[1] << 32 | w[0], hi = (uint64_t)w[3] << 32 | w[2];
        return (uint32_t)((t < 8 ? lo >> (8 * t) : hi >> (8 * (t - 8))) & 0xFFu);
    } else {
        return (w[t >> 2] >> (8 * (t & 3))) & 0xFFu;
    }
}

__device__ __forceinline__ uint32_t byte16(const uint32_t w[4], int32_t t) { return byte_at<4>(w, t); }

// Ring byte of logical row c: read code of base c-1 in bits 5-7 (rows up to
// m), 5 * window code below.
__device__ __forceinline__ uint8_t ring_byte(uint32_t wbyte, uint32_t rbyte, bool has_read) {
    const uint32_t rcode = has_read ? (uint32_t)read_code(rbyte) << 5 : 0u;
    return (uint8_t)(rcode | (uint32_t)(5 * win_code(wbyte)));
}

// Logical rows c0 .. c0 + nrows - 1 (1-based, nrows <= 4*NW and uniform)
// into the LDS ring from physical row ph on: row c is the byte (read code of
// base c-1 << 5 when c <= m) | 5 * window code of s0[c-1], at physical row
// 1 + (c-1) mod RING and again at RING + that row for the first W-1
// physical rows, so that a DP row reads its W codes from consecutive
// physical rows.  Rows past this lane's last row L get unspecified bytes
// (the caller only writes slots no row reads any more).
template <int W, int RING, int NW>
__device__ __forceinline__ void ring_batch(uint8_t *codes, uint32_t lane, const uint8_t *s0,
                                           const uint8_t *read, int32_t m, int32_t L, int32_t c0,
                                           int32_t ph, int32_t nrows) {
    uint32_t ww[NW], rw[NW];
    load_span<NW>(s0, c0 - 1, L - 1, ww);
    load_span<NW>(read, c0 - 1, m - 1, rw);
#pragma unroll
    for (int t = 0; t < 4 * NW; t++) {
        if (t < nrows) {
            const uint8_t v = ring_byte(byte_at<NW>(ww, t), byte_at<NW>(rw, t), c0 + t <= m);
            codes[(size_t)ph * 64 + lane] = v;
            if (ph <= W - 1) codes[(size_t)(ph + RING) * 64 + lane] = v;
            ph = ph == RING ? 1 : ph + 1;
        }
    }
}

// ring_batch for a band clamped on the left (cb < 0): logical row c's window
// byte is win[cb + c - 1] when that index is >= 0 and c <= L, else 'N'.
template <int W, int RING>
__device__ void ring_batch_clamped(uint8_t *codes, uint32_t lane, const uint8_t *win, int32_t cb,
                                   const uint8_t *read, int32_t m, int32_t L, int32_t c0, int32_t nrows) {
    uint32_t rw[4];
    load_span<4>(read, c0 - 1, m - 1, rw);
    uint32_t wb[16];
#pragma unroll
    for (int t = 0; t < 16; t++) {
        const int32_t c = c0 + t, idx = cb + c - 1;
        const bool in = t < nrows && idx >= 0 && c <= L;
        wb[t] = in ? (uint32_t)win[in ? idx : 0] : (uint32_t)'N';
    }
    int32_t ph = c0;
#pragma unroll
    for (int t = 0; t < 16; t++) {
        if (t < nrows) {
            const uint8_t v = ring_byte(wb[t], byte_at<4>(rw, t), c0 + t <= m);
            codes[(size_t)ph * 64 + lane] = v;
            if (ph <= W - 1) codes[(size_t)(ph + RING) * 64 + lane] = v;
            ph = ph == RING ? 1 : ph + 1;
        }
    }
}

// Traceback over the band trace layout (extend.rs:884-930) with value
// tracking; runs are written in walk order to slot[cap-1], slot[cap-2], ...
// A step is branch-free apart from the cache refills and the run writes:
// lanes of a wave walk in lockstep whatever their moves.
// Bytes p[lo .. lo + 63] (valid up to index hi) into this lane's dwords
// q0 .. q0 + 15 of the wave's LDS, dword q of lane at uint32 index q*64+lane.
__device__ __forceinline__ void stage64(uint32_t *lds, uint32_t lane, const uint8_t *p, int32_t lo,
                                        int32_t hi, int32_t q0) {
#pragma unroll
    for (int h = 0; h < 4; h++) {
        uint32_t w[4];
        load_span<4>(p, lo + 16 * h, hi, w);
#pragma unroll
        for (int k = 0; k < 4; k++) lds[(size_t)(q0 + 4 * h + k) * 64 + lane] = w[k];
    }
}

template <int W>
__device__ TbResult traceback_band(const PairView &pv, const GParams &P, const uint32_t *tr,
                                   int32_t cb, int32_t best, int32_t bi, int32_t bj,
                                   uint32_t *slot, uint32_t cap, uint8_t *lds, uint32_t lane) {
    constexpr int WW = (W + 7) / 8;
    int32_t i = bi, j = bj, V = best;
    uint32_t state = 0;  // 0 H, 1 E (insertion run), 2 F (deletion run)
    uint32_t cur_op = 0xFF, cur_len = 0, nrev = 0, nm = 0, ni = 0, nd = 0;
    // Column cache: the nibbles of band cell ck for rows ci0-31 .. ci0, row r
    // at bit 4*(ci0 - r) of the 128-bit pair (lo, hi).  Diagonal steps keep
    // k, so one batch of 32 independent loads serves up to 32 steps; loads
    // of rows above row 1 are clamped to row 1 (never used).
    int32_t ci0 = -1, ck = -1;
    uint64_t nlo = 0, nhi = 0;
    // Read and window bytes of the diagonal steps: every 64 iterations (the
    // same iteration for every lane still walking) the 64 bytes of each that
    // the next 64 steps can reach are staged in this lane's LDS words (the
    // window-code ring is free once the DP is done): read bytes in dwords
    // 0-15, window bytes in dwords 16-31.
    uint32_t *ldsw = (uint32_t *)lds;
    int32_t it = 0, rb = 0, wb = 0;
    while (i > 0 && j > 0) {
        const int32_t k = j - (cb + i);
        if (k < 0 || k >= W) break;
        if ((it++ & 63) == 0) {
            rb = max(i - 64, 0);
            wb = max(j - 64, 0);
            stage64(ldsw, lane, pv.read, rb, i - 1, 0);
            stage64(ldsw, lane, pv.win, wb, j - 1, 16);
        }
        if (k != ck || i > ci0 || i <= ci0 - 32) {
            const int32_t t = k >> 3;
            const int32_t c = min(8, W - 8 * t);
            const int32_t sh = 4 * (c - 1 - (k & 7));
            uint32_t w[32];
#pragma unroll
            for (int q = 0; q < 32; q++) w[q] = tr[((size_t)max(i - 1 - q, 0) * WW + t) * 64];
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
        if (state == 0 && V == 0) break;  // H == 0: the reference's H_START
        const int32_t d = ci0 - i;
        const uint32_t nib = (uint32_t)((d < 16 ? (nlo >> (4 * d)) : (nhi >> (4 * (d - 16)))) & 0xFu);
        // State 0 takes the diagonal when D == H, else E when E == H, else F;
        // a gap run continues while its cell's extension bit is set.
        const bool isM = state == 0 && !(nib & 8u);
        const bool isI = state == 1 || (state == 0 && (nib & 12u) == 8u);
        const bool isD = !isM && !isI;
        const bool gext = isI ? (nib & 2u) != 0 : (isD && (nib & 1u) != 0);
        const int32_t rq = i - 1 - rb, wq = j - 1 - wb;  // 0 .. 63 inside the stage
        const int32_t rc = read_code(lds[(((size_t)(rq >> 2) * 64 + lane) << 2) + (rq & 3)]);
        const int32_t fc = win_code(lds[(((size_t)(16 + (wq >> 2)) * 64 + lane) << 2) + (wq & 3)]);
        V -= isM ? lut_score(P, rc, fc) : (gext ? P.ext : P.oe);
        state = gext ? (isI ? 1u : 2u) : 0u;
        const uint32_t op = isM ? OP_M : (isI ? OP_I : OP_D);
        nm += isM;
        ni += isI;
        nd += isD;
        i -= isM || isI;
        j -= isM || isD;
        if (op != cur_op) {
            if (cur_len) slot[cap - 1 - nrev++] = (cur_len << 4) | cur_op;
            cur_op = op;
            cur_len = 1;
        } else {
            cur_len++;
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

constexpr int BAND_PF = 4;  // window-code prefetch distance in cells
// Halves from this one up keep U and B packed as int16 pairs (half the
// registers, one SDWA op more per cell).
#ifndef BAND_PACK_FROM
#define BAND_PACK_FROM 48
#endif

// pk = (low 16 bits of e + x) << 16 | (pk & 0xFFFF), in one VALU op.
__device__ __forceinline__ void pack_hi(uint32_t &pk, int32_t e, int32_t x) {
    asm("v_add_u32_sdwa %0, %1, %2 dst_sel:WORD_1 dst_unused:UNUSED_PRESERVE src0_sel:DWORD src1_sel:DWORD"
        : "+v"(pk)
        : "v"(e), "v"(x));
}

// Waves per SIMD each specialisation is built for (the register budget:
// 512 / waves VGPRs); the LDS ring is sized to match.
__host__ __device__ constexpr int band_waves(int half) {
    return half <= 16 ? 4 : (half <= 32 ? 2 : (half <= 48 ? 3 : 2));
}

template <int HALF, bool FOLD>
__global__ __launch_bounds__(BAND_BLOCK, band_waves(HALF)) void k_gotoh_band(GotohArgs a) {
    constexpr int W = 2 * HALF + 1;
    constexpr int WW = (W + 7) / 8;
    constexpr int RING = band_ring(HALF);
    static_assert((RING + W) * 64 >= 32 * 256, "traceback staging needs 32 dwords per lane");
    extern __shared__ uint8_t codes[];  // [RING + W][64]: (read code << 5) | 5 * window code
    uint32_t *lut_lds = (uint32_t *)codes;  // row 0 (never a code row): the 5 row LUTs
    const uint32_t lane = threadIdx.x;
    const uint32_t slot = blockIdx.x * BAND_BLOCK + lane;
    const GParams P = a.P;
    const int32_t oe = P.oe, ext = P.ext;
    const int32_t noe = -oe;
    const int32_t negx = NEG + ext;
    const int32_t zr = a.zero;
    // VGPR copies of the penalties: v_add with two VGPR operands issues at a
    // higher rate than with an SGPR operand.
    int32_t oev, extv;
    asm("v_mov_b32 %0, %1" : "=v"(oev) : "s"(oe));
    asm("v_mov_b32 %0, %1" : "=v"(extv) : "s"(ext));
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
        // Window codes of columns cb + 1 .. cb + L (interior: inside [1, n]),
        // logical rows 1 .. L, pass through an LDS ring of RING rows; the
        // first RING now, the rest as the row loop needs them.
        const int32_t L = m > 0 ? m + W - 1 : 0;
        const uint8_t *s0 = pv.win + cb;
        if (cb >= 0) {
            for (int32_t c0 = 1; c0 <= RING; c0 += 16)
                ring_batch<W, RING, 4>(codes, lane, s0, pv.read, m, L, c0, c0, min(16, RING - c0 + 1));
        } else {
            // left-clamped band: logical rows c <= -cb are 'N' columns, the
            // rest are read byte by byte from the window (every row the ring
            // refills later lies past RING > half >= -cb, inside the window)
            for (int32_t c0 = 1; c0 <= RING; c0 += 16)
                ring_batch_clamped<W, RING>(codes, lane, pv.win, cb, pv.read, m, L, c0, min(16, RING - c0 + 1));
        }
        if (lane < 5) lut_lds[lane] = FOLD ? a.lutf[lane] : a.lut[lane];
        __syncthreads();  // one wave per block: orders the LUT writes before any lane reads them

        // Previous row, band index k: U = H + open_ext and B = E + ext.  Wide
        // bands keep both as int16 halves of one VGPR (U low, B high), which
        // halves the register arrays (the planner admits only pairs whose
        // values fit int16; the rest run the int32 strip kernel): the reads
        // fold into SDWA operands, and the cell after k, once its F has read
        // U_k as a full word, writes E_k + ext into the high half in place
        // (an SDWA add with WORD_1 destination), so packing costs no op.
        constexpr bool PACKED = HALF >= BAND_PACK_FROM;
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
        // The ring advances RF logical rows at a time: row i = RF + 1,
        // 2 RF + 1, ... loads logical rows RING + i - RF .. RING + i - 1 (one
        // memory round trip) into the slots of rows i - RF .. i - 1, which no
        // later row reads.  Rows i .. i + RF - 1 read up to i + RF - 1 + W - 1
        // <= RING + i - 1 (RING >= W + RF - 1).  Refilling many rows at once
        // touches each cache line of the sequences in one burst instead of
        // once per few rows (when it may have left the L2 in between).
        constexpr int RF = (RING - W + 1) >= 32 ? 32 : 16;
        static_assert(RING >= W + RF - 1, "ring too small for the refill");
        int32_t pb = 1;  // physical ring row of logical row i
        for (int32_t i = 1; i <= umaxm; i++, pb = pb == RING ? 1 : pb + 1) {
            if (i > 1 && ((i - 1) & (RF - 1)) == 0 && RING + i - RF <= umaxm + W - 1) {
                int32_t ph = pb > RF ? pb - RF : pb - RF + RING;  // slot of row i - RF
                // 8 rows per round trip keeps the refill's registers below the
                // cell loop's; the later rounds hit the lines the first fetched
#pragma unroll
                for (int q = 0; q < RF; q += 8) {
                    ring_batch<W, RING, 2>(codes, lane, s0, pv.read, m, L, RING + i - RF + q, ph, 8);
                    ph = ph + 8 > RING ? ph + 8 - RING : ph + 8;
                }
            }
            if (i <= m) {
                const uint8_t *crow = codes + (size_t)pb * 64 + lane;
                const uint32_t rowlut = lut_lds[crow[0] >> 5];
                int32_t Ul = oe, Fx = negx, Hl = 0;
                // Row keys H << 8 | (64 - k') over segments of 65 cells
                // (k' = k mod 65): the inline constant 64 - k' needs no scalar
                // register, and the larger key is the larger H, then the
                // smaller k.  Keys are folded pairwise with v_max3.
                int32_t key1 = 0, key2 = 0, kpend = 0;
                int32_t Eprev = 0;
                uint32_t acc = 0;
                uint32_t *trow_p = tr + (size_t)(i - 1) * WW * 64;
                // Window codes are read BAND_PF cells ahead of their use (a
                // rotating register buffer), and a scheduling barrier closes
                // every cell, so the LDS latency hides behind the cells in
                // between and no value lives longer than the pipeline needs.
                uint32_t cbuf[BAND_PF];
#pragma unroll
                for (int q = 0; q < BAND_PF; q++) cbuf[q] = q < W ? crow[q * 64] : 0u;
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
                    const uint32_t code5 = cbuf[k % BAND_PF];
                    if (k + BAND_PF < W) cbuf[k % BAND_PF] = crow[(k + BAND_PF) * 64];
                    const int32_t sc = __builtin_amdgcn_sbfe((int32_t)rowlut, code5, 5);
                    int32_t D;
                    if constexpr (FOLD && PACKED) {  // sc is score - oe
                        asm("v_add_u32_sdwa %0, sext(%1), %2 dst_sel:DWORD dst_unused:UNUSED_PAD "
                            "src0_sel:WORD_0 src1_sel:DWORD"
                            : "=v"(D) : "v"(Pk[k]), "v"(sc));
                    } else if constexpr (FOLD) {
                        D = Ud + sc;
                    } else {
                        asm("v_add3_u32 %0, %1, %2, %3" : "=v"(D) : "v"(Ud), "v"(sc), "s"(noe));
                    }
                    // F is kept floored at 0, which folds H's floor into one
                    // v_max3: H is unchanged, and a traceback only follows an F
                    // run through cells whose F is positive, where the floored
                    // value and its extension bit equal the unfloored ones.
                    const int32_t E = max(Uup, Bup);
                    const int32_t F = max(max(Ul, Fx), zr);
                    const int32_t H = max(max(D, E), F);
                    // Sign bits: D < H, E < H, E from extension, F from extension,
                    // shifted into the trace word with v_alignbit.
                    const uint32_t x0 = (uint32_t)(D - H), x1 = (uint32_t)(E - H);
                    const uint32_t x2 = (uint32_t)(Uup - Bup), x3 = (uint32_t)(Ul - Fx);
                    asm("v_alignbit_b32 %0, %0, %1, 31\n\t"
                                 "v_alignbit_b32 %0, %0, %2, 31\n\t"
                                 "v_alignbit_b32 %0, %0, %3, 31\n\t"
                                 "v_alignbit_b32 %0, %0, %4, 31"
                                 : "+v"(acc) : "v"(x0), "v"(x1), "v"(x2), "v"(x3));
                    const int32_t Un = H + oev;
                    if constexpr (PACKED) {
                        // Ul (= U_{k-1}) has been read: pack B_{k-1} beside it.
                        if (k > 0) pack_hi(Pk[k - 1], Eprev, extv);
                        Pk[k] = (uint32_t)Un;
                        if (k == W - 1) pack_hi(Pk[k], E, extv);
                        Eprev = E;
                    } else {
                        Ur[k] = Un;
                        Bx[k] = E + extv;
                    }
                    Ul = Un;
                    Fx = F + extv;
                    // Running row keys, kept as sequential chains so the compiler
                    // cannot defer them into a tree that keeps every cell's H
                    // live until the row ends.
                    {
                        const int kq = k % 65;  // compile-time after unrolling
                        const bool last = (k == W - 1) || (kq == 64);
                        int32_t &key = k < 65 ? key1 : key2;
                        // one asm statement per fold: a dependent pair of asm
                        // statements gets a conservative s_nop between them
                        int32_t kk;
                        if ((kq & 1) == 1) {
                            asm("v_lshl_or_b32 %1, %2, 8, %4\n\t"
                                "v_max3_i32 %0, %0, %3, %1"
                                : "+v"(key), "=&v"(kk) : "v"(H), "v"(kpend), "I"(64 - kq));
                        } else if (last) {
                            asm("v_lshl_or_b32 %1, %2, 8, %3\n\t"
                                "v_max_i32 %0, %0, %1"
                                : "+v"(key), "=&v"(kk) : "v"(H), "I"(64 - kq));
                        } else {
                            asm("v_lshl_or_b32 %0, %1, 8, %2" : "=v"(kpend) : "v"(H), "I"(64 - kq));
                        }
                    }
                    if (k == W - 1) Hl = H;
                    if ((k & 7) == 7 || k == W - 1) {
                        trow_p[(k >> 3) * 64] = acc;
                        acc = 0;
                    }
                    __builtin_amdgcn_sched_barrier(0);
                }
                int32_t rmax = key1 >> 8;
                int32_t rarg = 64 - (key1 & 255);
                if constexpr (W > 65) {
                    if ((key2 >> 8) > rmax) {  // ties keep the first segment
                        rmax = key2 >> 8;
                        rarg = 65 + 64 - (key2 & 255);
                    }
                }
                if (!touched) {
                    // the first maximum is the window's first cell (band cell
                    // 0, or the first real one of a left-clamped row)
                    if (rarg <= -min(cb + i - 1, 0)) {
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
            if (a.esc_info) {
                a.esc_info[2 * (size_t)p] = best;
                a.esc_info[2 * (size_t)p + 1] = bj - bi;
            }
        }
        if (active && !escalate) {
            const uint64_t c0 = a.cig_off[p];
            const uint32_t cap = (uint32_t)(a.cig_off[p + 1] - c0);
            uint32_t *cslot = a.cigar + c0;
            TbResult tb = traceback_band<W>(pv, P, tr, cb, best, bi, bj, cslot, cap, codes, lane);
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

template <int HALF, bool FOLD>
hipError_t launch_half(const GotohArgs &a, int blocks, size_t lds, hipStream_t s) {
    hipLaunchKernelGGL((k_gotoh_band<HALF, FOLD>), dim3(blocks), dim3(BAND_BLOCK), lds, s, a);
    return hipGetLastError();
}

// The folded LUT when its fields fit, else the plain one plus -oe per cell.
template <int HALF>
hipError_t launch_half(const GotohArgs &a, int blocks, size_t lds, hipStream_t s) {
    if (a.lut_folded) return launch_half<HALF, true>(a, blocks, lds, s);
    return launch_half<HALF, false>(a, blocks, lds, s);
}

template <int HALF>
int occupancy_half(size_t lds) {
    int nb = 0;
    hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_gotoh_band<HALF, true>, BAND_BLOCK, lds);
    return nb > 0 ? nb : 1;
}

}  // namespace

bool band_half_supported(int half) {
    return half == 16 || half == 30 || half == 32 || half == 48 || half == 64;
}

static size_t band_lds(int half) { return (size_t)(band_ring(half) + 2 * half + 1) * 64; }

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
