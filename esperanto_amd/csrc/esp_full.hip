// esp_full.hip -- full-matrix Gotoh kernel for extend_hint's second attempt.
//
// extend_hint (extend.rs:392-428) reruns a band attempt on the full matrix
// [1, m] x [1, n] when the band's maximum touched its edge or scored under
// 85% of perfect; with the reference's own edge rule this is frequent (a
// first-row maximum tied on the band edge already counts).  These windows are
// wide (600-2000 columns for spliced and mate-rescue windows) and reads short
// (at most 259 bases: the band attempt admitted them), so the matrix is swept
// column by column with the whole read held in registers:
//   * lane per pair; row i's state (U = H(i, j-1) + open_ext, Fx = F(i, j-1)
//     + ext) lives in one VGPR as two int16 halves (read lengths and the
//     band-kernel scoring envelope keep every value inside int16);
//   * a column is MAXM unrolled rows top-down, the vertical carries (H + oe,
//     E + ext of the cell above) in registers, so a cell is pure VALU work;
//   * window codes stream through an LDS chunk per lane, read codes sit in
//     LDS per row, and the substitution score is one v_bfe_i32 of a per-
//     column LUT word indexed by the read code;
//   * trace: the band kernel's 4 bits per cell, column-major, 8 rows a word.
// Rows past m compute with a read code that scores -16 against everything,
// so no such cell can beat or tie a real row's maximum at an earlier row.
// The best cell is the first row-major strict maximum (extend.rs:751-772):
// a column key (H << 16 | 0xFFFF - i) taken strictly across columns.
#include <algorithm>

#include "esp_kernels.hpp"

using namespace espd;

namespace espk {

namespace {

constexpr int FULL_CHUNK = 64;  // window columns staged per LDS chunk
constexpr int32_t NEG16 = -20000;

// Writes win_code(src[q]) to chunk[q * 64] for q in [0, len): head and
// tail bytewise, the aligned body as dwords, eight loads in flight.
__device__ void stage_win(const uint8_t *src, int32_t len, uint8_t *chunk) {
    const int32_t head = min(len, (int32_t)((4u - ((uint32_t)(uintptr_t)src & 3u)) & 3u));
    for (int32_t q = 0; q < head; q++) chunk[q * 64] = (uint8_t)win_code(src[q]);
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
                for (int b = 0; b < 4; b++) chunk[(r + b) * 64] = (uint8_t)win_code((w[u] >> (8 * b)) & 0xFFu);
            }
        }
    }
    for (int32_t q = head + 4 * nd; q < len; q++) chunk[q * 64] = (uint8_t)win_code(src[q]);
}

// Traceback over the column-major trace (extend.rs:884-930) with value
// tracking; runs are written in walk order to slot[cap-1], slot[cap-2], ...
// A diagonal cache holds the nibbles and window codes of cells (i-q, j-q),
// q < 32: one batch of independent loads serves up to 32 diagonal steps.
template <int WR>
__device__ TbResult traceback_full(const PairView &pv, const GParams &P, const uint8_t *rcode,
                                   const uint32_t *tr, int32_t best, int32_t bi, int32_t bj,
                                   uint32_t *slot, uint32_t cap) {
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
    int32_t ci0 = -1, cdg = 0x7FFFFFFF;
    uint64_t nlo = 0, nhi = 0, flo = 0, fhi = 0;
    while (i > 0 && j > 0) {
        if (j - i != cdg || i > ci0 || i <= ci0 - 32) {
            uint32_t w[32], f[32];
#pragma unroll
            for (int q = 0; q < 32; q++) {
                const bool ok = i - q >= 1 && j - q >= 1;
                w[q] = ok ? tr[((size_t)(j - 1 - q) * WR + ((i - 1 - q) >> 3)) * 64] : 0u;
                f[q] = ok ? pv.win[j - 1 - q] : 0u;
            }
            nlo = nhi = flo = fhi = 0;
#pragma unroll
            for (int q = 0; q < 32; q++) {
                const uint32_t r = (uint32_t)(i - 1 - q) & 7u;
                const uint64_t nib = (w[q] >> (4 * (7 - r))) & 0xFu;
                const uint64_t fc = (uint64_t)win_code(f[q] & 0xFFu);
                if (q < 16) {
                    nlo |= nib << (4 * q);
                    flo |= fc << (4 * q);
                } else {
                    nhi |= nib << (4 * (q - 16));
                    fhi |= fc << (4 * (q - 16));
                }
            }
            ci0 = i;
            cdg = j - i;
        }
        const int32_t d = ci0 - i;
        const uint32_t nib = (uint32_t)((d < 16 ? (nlo >> (4 * d)) : (nhi >> (4 * (d - 16)))) & 0xFu);
        if (state == 0) {
            if (V == 0) break;  // H == 0: the reference's H_START
            if (!(nib & 8u)) {
                const int32_t rc = rcode[(size_t)i * 64] / 5;
                const uint32_t fc5 = (uint32_t)((d < 16 ? (flo >> (4 * d)) : (fhi >> (4 * (d - 16)))) & 0xFu);
                V -= lut_score(P, rc, (int32_t)fc5);
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

// pk = (low 16 bits of e + x) << 16 | (pk & 0xFFFF), in one VALU op.
__device__ __forceinline__ void pack_hi(uint32_t &pk, int32_t e, int32_t x) {
    asm("v_add_u32_sdwa %0, %1, %2 dst_sel:WORD_1 dst_unused:UNUSED_PRESERVE src0_sel:DWORD src1_sel:DWORD"
        : "+v"(pk)
        : "v"(e), "v"(x));
}

// Pruned mode (PRUNE).  Let L be the band attempt's best score (a lower
// bound on the full matrix's best: band cells see margins no higher than the
// full matrix's values) and S the largest substitution score.  A value V at
// row i can lie on, or decide a tie on, the optimal path only if
// V + S * (m - i) >= L (call it live): a path gains at most S a row.  Every
// maximiser of a live value is live, so skipping cells that are not live,
// with margins (H = 0, E = F = NEG) standing in for them, leaves every live
// value, every trace bit the traceback reads and the first row-major
// maximum unchanged.  Per column the wave computes the head groups (every
// row from which a path starting at 0 can reach L, plus a pad) and the
// groups meeting rows [j - dmax - h, j - dmin + h] (its pairs' band-maximum
// diagonals dmin..dmax, h = prune_half); the stripe only moves down, so
// a group enters it once, with its initial state.  The proof that nothing
// skipped was live: every computed cell with an uncomputed down, right or
// diagonal neighbour must hold H, E and F below L - S * (m - i).  A pair that
// fails it goes to the fallback list (the unpruned kernel) untraced.
template <int MAXM, bool FOLD, bool PRUNE>
__global__ __launch_bounds__(64) void k_gotoh_full(GotohArgs a) {
    static_assert(MAXM % 8 == 0, "rows come in groups of 8");
    constexpr int WR = MAXM / 8;  // trace words per column
    extern __shared__ uint8_t lds[];
    // LDS: column LUTs (8 words), read codes x5 of rows 0..MAXM, window chunk
    uint32_t *flut = (uint32_t *)lds;
    const uint32_t lane = threadIdx.x;
    uint8_t *rcode = lds + 32 + lane;                      // row i at rcode[i * 64]
    uint8_t *chunk = lds + 32 + (MAXM + 1) * 64 + lane;    // column q at chunk[q * 64]
    const GParams P = a.P;
    const int32_t oe = P.oe, ext = P.ext, noe = -oe;
    const int32_t zr = a.zero;  // 0 the compiler cannot see through (one v_max3)
    // VGPR copies of the penalties: v_add with two VGPR operands issues at a
    // higher rate than with an SGPR operand.
    int32_t oev, extv, noev;
    asm("v_mov_b32 %0, %1" : "=v"(oev) : "s"(oe));
    asm("v_mov_b32 %0, %1" : "=v"(extv) : "s"(ext));
    asm("v_mov_b32 %0, %1" : "=v"(noev) : "s"(noe));
    const uint32_t first = a.range[0], total = a.range[1];
    const uint32_t usz = a.unit ? a.unit : 64u;
    const uint32_t nunits = (total + usz - 1) / usz;
    if (lane < 8) flut[lane] = FOLD ? a.flutf[lane < 6 ? lane : 5] : a.flut[lane < 6 ? lane : 5];
    for (int q = 0; q < FULL_CHUNK; q++) chunk[q * 64] = 0;
    __syncthreads();
    uint32_t *tr = a.trace + (size_t)blockIdx.x * a.maxw * 64 + lane;
    for (;;) {
        const uint32_t u = wave_pop(a.counter, lane, 1u);
        if (u >= nunits) break;
        const uint32_t ubegin = first + usz * u;
        const uint32_t ucount = min(usz, total - usz * u);
        const uint32_t p = a.order[ubegin + (lane < ucount ? lane : 0u)];
        // A pair longer than the tile or the trace slab was sized for (a
        // graph replayed on pairs past its bounds, which the planner counts
        // as bad) is left alone.
        const bool active = lane < ucount && a.read_len[p] <= (uint32_t)MAXM &&
                            (uint64_t)a.win_len[p] * WR <= (uint64_t)a.maxw;
        PairView pv;
        pv.read = a.reads + a.read_off[p];
        pv.win = a.wins + a.win_off[p];
        pv.m = active ? (int32_t)a.read_len[p] : 0;
        pv.n = active ? (int32_t)a.win_len[p] : 0;
        pv.off = 0;
        pv.half = pv.n;
        const int32_t m = pv.m, n = pv.n;
        // wave maxima (uniform loop bounds)
        int32_t umaxm = m, umaxn = n;
#pragma unroll
        for (int sh = 32; sh >= 1; sh >>= 1) {
            umaxm = max(umaxm, __shfl_xor(umaxm, sh));
            umaxn = max(umaxn, __shfl_xor(umaxn, sh));
        }
        // uniform after the reductions: scalar copies keep the column loop and
        // the group branches scalar
        const int32_t ngroups = __builtin_amdgcn_readfirstlane((umaxm + 7) >> 3);
        umaxn = __builtin_amdgcn_readfirstlane(umaxn);
        // Pruned mode: the lane's threshold base L - S * m, and the wave's head
        // groups and diagonal range (uniform).
        int32_t tbase = 0, nhead = ngroups, dmin = 0, dmax = 0;
        if constexpr (PRUNE) {
            const int32_t S = a.prune_s;
            const int32_t L = active ? a.esc_info[2 * (size_t)p] : 0;
            const int32_t d = active ? a.esc_info[2 * (size_t)p + 1] : 0;
            tbase = L - S * m;
            int32_t nh = active ? esc_head_groups(m, L, S, a.prune_pad) : 0;
            int32_t lo = active ? d : 0x7FFFFFFF, hi = active ? d : -0x7FFFFFFF;
#pragma unroll
            for (int sh = 32; sh >= 1; sh >>= 1) {
                nh = max(nh, __shfl_xor(nh, sh));
                lo = min(lo, __shfl_xor(lo, sh));
                hi = max(hi, __shfl_xor(hi, sh));
            }
            // uniform after the reductions: scalar copies keep the group
            // branches scalar
            if (lo > hi) lo = hi = 0;  // no active lane: no column runs
            nhead = __builtin_amdgcn_readfirstlane(min(nh, ngroups));
            dmin = __builtin_amdgcn_readfirstlane(lo);
            dmax = __builtin_amdgcn_readfirstlane(hi);
        }
        bool fail = false;
        // Pruned: bit g of gmask(jj) says group g is computed at column jj
        // (uniform): the head groups, and the groups meeting rows
        // [jj - dmax - h, jj - dmin + h].  The masks of columns j - 1, j, j + 1
        // roll along the sweep.
        auto gmask = [&](int32_t jj) -> uint32_t {
            const int32_t rlo = jj - dmax - a.prune_half, rhi = jj - dmin + a.prune_half;
            const int32_t glo = rlo <= 1 ? 0 : (rlo - 1) >> 3;
            const int32_t ghi = min(rhi < 1 ? -1 : (rhi - 1) >> 3, ngroups - 1);
            uint32_t mk = (1u << nhead) - 1u;
            if (glo <= ghi) mk |= (2u << ghi) - (1u << glo);
            return mk;
        };
        uint32_t mprev = 0, mcur = 0, mnext = 0;
        if constexpr (PRUNE) {
            mcur = gmask(0);
            mnext = gmask(1);
        }
        // Read codes x5 into LDS rows 1..m; rows past m get code 5 (-16 vs all).
        for (int32_t i = 1; i <= 8 * ngroups; i++)
            rcode[(size_t)i * 64] = (uint8_t)(5 * (i <= m ? read_code(pv.read[i - 1]) : 5u));
        uint32_t Pk[MAXM];
        const uint32_t pinit = ((uint32_t)oe & 0xFFFFu) | ((uint32_t)NEG16 << 16);  // H = 0, F = NEG
#pragma unroll
        for (int r = 0; r < MAXM; r++) Pk[r] = pinit;
        int32_t bestkey = 0xFFFF, bj = 0;
        // census (a.census set): cells this lane computed inside its matrix,
        // and the wave's computed group-columns
        uint32_t ccells = 0, cgroups = 0;
        for (int32_t j0 = 0; j0 < umaxn; j0 += FULL_CHUNK) {
            const int32_t len = max(0, min(FULL_CHUNK, n - j0));
            stage_win(pv.win + j0, len, chunk);
            const int32_t qend = min(FULL_CHUNK, umaxn - j0);
            for (int32_t q = 0; q < qend; q++) {
                const int32_t j = j0 + q + 1;
                const uint32_t clut = flut[chunk[q * 64]];
                if constexpr (PRUNE) {
                    mprev = mcur;
                    mcur = mnext;
                    mnext = gmask(j + 1);
                }
                // group g is computed at column j - 1 + t3
                auto computed = [&](int g, int t3) {
                    const uint32_t mk = t3 == 0 ? mprev : (t3 == 1 ? mcur : mnext);
                    return g < 32 && ((mk >> g) & 1u);
                };
                // m and the threshold base, opaque per column: otherwise every
                // row's exit predicate is hoisted out of the column loop and
                // held in registers.
                int32_t mc = m, tbc = tbase;
                if constexpr (PRUNE) asm volatile("" : "+v"(mc), "+v"(tbc));
                // row 0: H(0, j) = 0, E(0, j) = NEG; the diagonal of row 1 is
                // H(0, j-1) = 0, i.e. U = oe in the low half of pdiag
                int32_t Uc = oe, Bc = NEG + ext, ckey = 0, kpend = 0;
                uint32_t pdiag = (uint32_t)oe & 0xFFFFu;
                uint32_t *tcol = tr + (size_t)(j - 1) * WR * 64;
#pragma unroll
                for (int g = 0; g < WR; g++) {
                    // scalar masks, opaque per group: every bit test stays at its
                    // use instead of being hoisted into a lane mask per group
                    if constexpr (PRUNE) asm volatile("" : "+s"(mprev), "+s"(mcur), "+s"(mnext));
                    if (PRUNE ? computed(g, 1) : g < ngroups) {
                        // Pruned: after a skipped group the carries come from margins,
                        // and the diagonal from column j - 1 only if row 8g was
                        // computed there.
                        if constexpr (PRUNE) {
                            if (g > 0 && !computed(g - 1, 1)) {
                                Uc = oe;
                                Bc = NEG + ext;
                                pdiag = (j > 1 && computed(g - 1, 0)) ? Pk[8 * g - 1] : ((uint32_t)oe & 0xFFFFu);
                            }
                        }
                        uint32_t acc = 0;
#pragma unroll
                        for (int r = 0; r < 8; r++) {
                            const int i = 8 * g + r + 1;
                            // pk = (F(i, j-1) + ext, H(i, j-1) + oe) as int16 halves;
                            // SDWA operands read them sign-extended.
                            const uint32_t pk = Pk[i - 1];
                            const uint32_t rc5 = rcode[(size_t)i * 64];
                            int32_t sc = __builtin_amdgcn_sbfe((int32_t)clut, rc5, 5);
                            if constexpr (!FOLD) sc += noev;  // the folded LUT holds score - oe
                            int32_t D, F, x3;
                            asm("v_add_u32_sdwa %0, sext(%1), %2 dst_sel:DWORD dst_unused:UNUSED_PAD "
                                "src0_sel:WORD_0 src1_sel:DWORD"
                                : "=v"(D) : "v"(pdiag), "v"(sc));
                            // E is kept floored at 0, which folds H's floor into one
                            // v_max3: H is unchanged, and a traceback only follows an
                            // E run through cells whose E is positive, where the
                            // floored value and its extension bit equal the
                            // unfloored ones.
                            const int32_t E = max(max(Uc, Bc), zr);
                            asm("v_max_i32_sdwa %0, sext(%1), sext(%1) dst_sel:DWORD dst_unused:UNUSED_PAD "
                                "src0_sel:WORD_0 src1_sel:WORD_1"
                                : "=v"(F) : "v"(pk));
                            asm("v_sub_u32_sdwa %0, sext(%1), sext(%1) dst_sel:DWORD dst_unused:UNUSED_PAD "
                                "src0_sel:WORD_0 src1_sel:WORD_1"
                                : "=v"(x3) : "v"(pk));
                            const int32_t H = max(max(D, E), F);
                            const uint32_t x0 = (uint32_t)(D - H), x1 = (uint32_t)(E - H);
                            const uint32_t x2 = (uint32_t)(Uc - Bc);
                            asm("v_alignbit_b32 %0, %0, %1, 31\n\t"
                                "v_alignbit_b32 %0, %0, %2, 31\n\t"
                                "v_alignbit_b32 %0, %0, %3, 31\n\t"
                                "v_alignbit_b32 %0, %0, %4, 31"
                                : "+v"(acc) : "v"(x0), "v"(x1), "v"(x2), "v"((uint32_t)x3));
                            const int32_t Un = H + oev;
                            pdiag = pk;  // H(i, j-1) + oe is the next row's diagonal
                            uint32_t pn = (uint32_t)Un;
                            pack_hi(pn, F, extv);
                            Pk[i - 1] = pn;
                            Uc = Un;
                            Bc = E + extv;
                            // column key H << 16 | 0xFFFF - i, folded pairwise with
                            // v_max3 (one asm statement per fold)
                            if (r & 1) {
                                int32_t kk;
                                asm("v_lshl_or_b32 %1, %2, 16, %4\n\t"
                                    "v_max3_i32 %0, %0, %3, %1"
                                    : "+v"(ckey), "=&v"(kk) : "v"(H), "v"(kpend), "s"(0xFFFF - i));
                            } else {
                                asm("v_lshl_or_b32 %0, %1, 16, %2" : "=v"(kpend) : "v"(H), "s"(0xFFFF - i));
                            }
                        }
                        tcol[g * 64] = acc;
                        // Pruned exit checks, on H(i, j) = the low half of Pk[i - 1]
                        // minus oe (H bounds E and F): every row when the group
                        // is dropped at j + 1 (right and diagonal exits, j < n),
                        // the last row when the group below is not computed at j
                        // or j + 1 (down and diagonal exits, i < m).
                        if constexpr (PRUNE) {
                            const bool chk_all = !computed(g, 2);
                            const bool chk_last = !computed(g + 1, 1) || !computed(g + 1, 2);
                            if (chk_all || chk_last) {
#pragma unroll
                                for (int r = 0; r < 8; r++) {
                                    const int i = 8 * g + r + 1;
                                    if (chk_all || r == 7) {
                                        const int32_t H = (int32_t)(int16_t)(Pk[i - 1] & 0xFFFFu) - oe;
                                        const bool side = chk_all && j < n && i <= mc;
                                        const bool down = r == 7 && chk_last && j <= n && i < mc;
                                        fail |= (side || down) && H >= tbc + a.prune_s * i;
                                    }
                                }
                            }
                        }
                    }
                }
                if (j <= n && ckey > bestkey) {
                    bestkey = ckey;
                    bj = j;
                }
                if constexpr (PRUNE) {
                    if (a.census) {
                        // rows 1..m of the computed groups: whole groups below
                        // m >> 3, then the m & 7 rows of the partial one
                        const uint32_t gm = (uint32_t)m >> 3;
                        const uint32_t rows = 8u * (uint32_t)__builtin_popcount(mcur & ((1u << gm) - 1u)) +
                                              (((mcur >> gm) & 1u) ? ((uint32_t)m & 7u) : 0u);
                        ccells += j <= n ? rows : 0u;
                        cgroups += (uint32_t)__builtin_popcount(mcur);
                    }
                }
            }
        }
        if (a.census) {
            if constexpr (!PRUNE) {
                ccells = (uint32_t)m * (uint32_t)n;
                cgroups = (uint32_t)ngroups * (uint32_t)umaxn;
            }
            if (active) atomicAdd(&a.census[0], (unsigned long long)ccells);
            if (lane == 0) atomicAdd(&a.census[1], (unsigned long long)cgroups * 8ull * 64ull);
            if (a.census_first && active) atomicAdd(&a.census[2], (unsigned long long)m * (unsigned long long)n);
        }
        if (PRUNE && active && fail) {
            a.fb_list[atomicAdd(&a.fb_range[1], 1u)] = p;
        } else if (active) {
            int32_t best = bestkey >> 16, bi = 0xFFFF - (bestkey & 0xFFFF);
            if (best <= 0) {
                best = 0;
                bi = 0;
                bj = 0;
            }
            const uint64_t c0 = a.cig_off[p];
            const uint32_t cap = (uint32_t)(a.cig_off[p + 1] - c0);
            uint32_t *cslot = a.cigar + c0;
            TbResult tb = traceback_full<WR>(pv, P, rcode, tr, best, bi, bj, cslot, cap);
            esp_ext e;
            uint32_t nops;
            finish_pair(pv, P, best, tb, cslot, cap, &e, &nops);
            a.ext[p] = e;
            a.cig_n[p] = nops;
        }
    }
}

// ---- Pruned pass over a register window ---------------------------------
//
// k_gotoh_full keeps every row of the tile in registers (one VGPR a row),
// which holds it at two waves per SIMD.  A pair whose head is at most PR_HG
// groups (the escalation sort's first class) only ever computes its head
// groups and the stripe, which moves down at most one group a column, so
// k_gotoh_pruned keeps just those: the head rows in H[8 * PR_HG] and the
// stripe's groups in a window S[8 * PR_SG] whose base group sb (uniform)
// follows the stripe.  When sb moves down one group, the window shifts by
// one group (the leaving group's last row is kept as the next group's
// diagonal) and the entering group starts from margins, as every group does
// when it first enters the stripe in the whole-tile pass.  The computed set
// is the whole-tile pass's (head groups, and the stripe groups from the
// first non-head one, at most PR_SG of them), the stripe's half-width is
// narrowed so its rows span at most PR_SG groups, and the exit checks are
// the same: a pair that fails them goes to the fallback list untraced.  The
// whole-tile pass runs the other head classes at the same time on a second
// stream.
constexpr int PR_HG = 3, PR_SG = 8, PR_CHUNK = 32;

// One group of 8 rows (i0 + 1 .. i0 + 8) of a column: P[r] holds row i0 + r
// + 1's packed (F + ext, H + oe) of the previous column and receives this
// column's; Uc, Bc carry H + oe and E + ext of the cell above and pdiag the
// row above's packed value of the previous column; rc points at row i0 + 1's
// read code.  Folds the rows' keys into ckey and returns the trace word.
template <bool FOLD>
__device__ __forceinline__ uint32_t group8(uint32_t *P, const uint8_t *rc, uint32_t clut, int32_t i0, int32_t zr,
                                           int32_t noev, int32_t oev, int32_t extv, int32_t &Uc, int32_t &Bc,
                                           uint32_t &pdiag, int32_t &ckey) {
    uint32_t acc = 0;
    int32_t kpend = 0;
    // the group's read codes load together (one LDS round trip a group)
    uint32_t rcs[8];
#pragma unroll
    for (int r = 0; r < 8; r++) rcs[r] = rc[r * 64];
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int r = 0; r < 8; r++) {
        const int32_t i = i0 + r + 1;
        const uint32_t pk = P[r];
        const uint32_t rc5 = rcs[r];
        int32_t sc = __builtin_amdgcn_sbfe((int32_t)clut, rc5, 5);
        if constexpr (!FOLD) sc += noev;
        int32_t D, F, x3;
        asm("v_add_u32_sdwa %0, sext(%1), %2 dst_sel:DWORD dst_unused:UNUSED_PAD "
            "src0_sel:WORD_0 src1_sel:DWORD"
            : "=v"(D) : "v"(pdiag), "v"(sc));
        const int32_t E = max(max(Uc, Bc), zr);
        asm("v_max_i32_sdwa %0, sext(%1), sext(%1) dst_sel:DWORD dst_unused:UNUSED_PAD "
            "src0_sel:WORD_0 src1_sel:WORD_1"
            : "=v"(F) : "v"(pk));
        asm("v_sub_u32_sdwa %0, sext(%1), sext(%1) dst_sel:DWORD dst_unused:UNUSED_PAD "
            "src0_sel:WORD_0 src1_sel:WORD_1"
            : "=v"(x3) : "v"(pk));
        const int32_t H = max(max(D, E), F);
        const uint32_t x0 = (uint32_t)(D - H), x1 = (uint32_t)(E - H);
        const uint32_t x2 = (uint32_t)(Uc - Bc);
        asm("v_alignbit_b32 %0, %0, %1, 31\n\t"
            "v_alignbit_b32 %0, %0, %2, 31\n\t"
            "v_alignbit_b32 %0, %0, %3, 31\n\t"
            "v_alignbit_b32 %0, %0, %4, 31"
            : "+v"(acc) : "v"(x0), "v"(x1), "v"(x2), "v"((uint32_t)x3));
        const int32_t Un = H + oev;
        pdiag = pk;
        uint32_t pn = (uint32_t)Un;
        pack_hi(pn, F, extv);
        P[r] = pn;
        Uc = Un;
        Bc = E + extv;
        if (r & 1) {
            int32_t kk;
            asm("v_lshl_or_b32 %1, %2, 16, %4\n\t"
                "v_max3_i32 %0, %0, %3, %1"
                : "+v"(ckey), "=&v"(kk) : "v"(H), "v"(kpend), "s"(0xFFFF - i));
        } else {
            asm("v_lshl_or_b32 %0, %1, 16, %2" : "=v"(kpend) : "v"(H), "s"(0xFFFF - i));
        }
    }
    return acc;
}

template <int MAXM, bool FOLD>
__global__ __launch_bounds__(64, 3) void k_gotoh_pruned(GotohArgs a) {
    static_assert(MAXM % 8 == 0 && MAXM <= 8 * 32, "rows come in groups of 8");
    constexpr int WR = MAXM / 8;  // trace words per column (the whole-tile layout)
    extern __shared__ uint8_t lds[];
    // LDS: column LUTs (8 words), read codes x5 of rows 0..MAXM, window chunk
    uint32_t *flut = (uint32_t *)lds;
    const uint32_t lane = threadIdx.x;
    uint8_t *rcode = lds + 32 + lane;                    // row i at rcode[i * 64]
    uint8_t *chunk = lds + 32 + (MAXM + 1) * 64 + lane;  // column q at chunk[q * 64]
    const GParams P = a.P;
    const int32_t oe = P.oe, ext = P.ext, noe = -oe;
    const int32_t zr = a.zero;
    int32_t oev, extv, noev;
    asm("v_mov_b32 %0, %1" : "=v"(oev) : "s"(oe));
    asm("v_mov_b32 %0, %1" : "=v"(extv) : "s"(ext));
    asm("v_mov_b32 %0, %1" : "=v"(noev) : "s"(noe));
    const uint32_t first = a.range[0], total = a.range[1];
    const uint32_t nunits = (total + 63) / 64;
    if (lane < 8) flut[lane] = FOLD ? a.flutf[lane < 6 ? lane : 5] : a.flut[lane < 6 ? lane : 5];
    for (int q = 0; q < PR_CHUNK; q++) chunk[q * 64] = 0;
    __syncthreads();
    uint32_t *tr = a.trace + (size_t)blockIdx.x * a.maxw * 64 + lane;
    const uint32_t pinit = ((uint32_t)oe & 0xFFFFu) | ((uint32_t)NEG16 << 16);  // H = 0, F = NEG
    for (;;) {
        const uint32_t u = wave_pop(a.counter, lane, 1u);
        if (u >= nunits) break;
        const uint32_t ubegin = first + 64 * u;
        const uint32_t ucount = min(64u, total - 64 * u);
        const uint32_t p = a.order[ubegin + (lane < ucount ? lane : 0u)];
        const bool active = lane < ucount && a.read_len[p] <= (uint32_t)MAXM &&
                            (uint64_t)a.win_len[p] * WR <= (uint64_t)a.maxw;
        PairView pv;
        pv.read = a.reads + a.read_off[p];
        pv.win = a.wins + a.win_off[p];
        pv.m = active ? (int32_t)a.read_len[p] : 0;
        pv.n = active ? (int32_t)a.win_len[p] : 0;
        pv.off = 0;
        pv.half = pv.n;
        const int32_t m = pv.m, n = pv.n;
        const int32_t S = a.prune_s;
        const int32_t L = active ? a.esc_info[2 * (size_t)p] : 0;
        const int32_t d = active ? a.esc_info[2 * (size_t)p + 1] : 0;
        const int32_t tbase = L - S * m;
        int32_t umaxm = m, umaxn = n;
        int32_t nh = active ? esc_head_groups(m, L, S, a.prune_pad) : 0;
        int32_t lo = active ? d : 0x7FFFFFFF, hi = active ? d : -0x7FFFFFFF;
#pragma unroll
        for (int sh = 32; sh >= 1; sh >>= 1) {
            umaxm = max(umaxm, __shfl_xor(umaxm, sh));
            umaxn = max(umaxn, __shfl_xor(umaxn, sh));
            nh = max(nh, __shfl_xor(nh, sh));
            lo = min(lo, __shfl_xor(lo, sh));
            hi = max(hi, __shfl_xor(hi, sh));
        }
        // uniform after the reductions: scalar copies keep the column loop
        // and the group branches scalar
        const int32_t ngroups = __builtin_amdgcn_readfirstlane((umaxm + 7) >> 3);
        umaxn = __builtin_amdgcn_readfirstlane(umaxn);
        if (lo > hi) lo = hi = 0;  // no active lane: no column runs
        const int32_t nhead = __builtin_amdgcn_readfirstlane(min(nh, ngroups));
        const int32_t dmin = __builtin_amdgcn_readfirstlane(lo);
        const int32_t dmax = __builtin_amdgcn_readfirstlane(hi);
        // the stripe's half-width: at most prune_half, and narrow enough that
        // its 2 hw + 1 + (dmax - dmin) rows span at most PR_SG groups (a wave
        // whose diagonals spread further, as when few pairs share each
        // diagonal, computes a narrower stripe than its pairs need, and the
        // pairs the exit checks cannot prove fail over)
        const int64_t room = ((int64_t)8 * PR_SG - 9 - ((int64_t)dmax - (int64_t)dmin)) / 2;
        const int32_t hw = (int32_t)max((int64_t)0, min((int64_t)a.prune_half, room));
        // a head wider than the registers (the sort's class rules it out)
        // fails over whole
        bool fail = nhead > PR_HG;
        // the stripe's groups at column jj: [sbase, stop] (empty when stop < sbase)
        auto stripe = [&](int32_t jj, int32_t &sbase, int32_t &stop) {
            const int32_t rlo = jj - dmax - hw, rhi = jj - dmin + hw;
            const int32_t glo = rlo <= 1 ? 0 : (rlo - 1) >> 3;
            const int32_t ghi = min(rhi < 1 ? -1 : (rhi - 1) >> 3, ngroups - 1);
            sbase = max(glo, nhead);
            stop = min(ghi, sbase + PR_SG - 1);
        };
        for (int32_t i = 1; i <= 8 * ngroups; i++)
            rcode[(size_t)i * 64] = (uint8_t)(5 * (i <= m ? read_code(pv.read[i - 1]) : 5u));
        uint32_t Hr[8 * PR_HG], Sr[8 * PR_SG];
#pragma unroll
        for (int r = 0; r < 8 * PR_HG; r++) Hr[r] = pinit;
#pragma unroll
        for (int r = 0; r < 8 * PR_SG; r++) Sr[r] = pinit;
        // Columns run in stretches over which the window stays put and no
        // chunk is staged, so the shift and the staging sit outside the
        // per-column loop.
        int32_t sbc, topc;  // the stripe of the next column to compute
        stripe(1, sbc, topc);
        int32_t sbw = sbc;  // the group Sr[0..8) holds
        int32_t bestkey = 0xFFFF, bj = 0;
        uint32_t ccells = 0, cgroups = 0;
        int32_t j = fail ? umaxn + 1 : 1;
        while (j <= umaxn) {
            // the window follows the stripe down: sb grows by at most one
            // group a column (the stripe's first row grows by one)
            uint32_t drop = (uint32_t)oe & 0xFFFFu;
            if (sbw < sbc) {
                drop = Sr[7];
#pragma unroll
                for (int r = 0; r < 8 * (PR_SG - 1); r++) Sr[r] = Sr[r + 8];
#pragma unroll
                for (int r = 0; r < 8; r++) Sr[8 * (PR_SG - 1) + r] = pinit;
                sbw = sbc;
            }
            if (((j - 1) & (PR_CHUNK - 1)) == 0) {
                const int32_t j0 = j - 1;
                stage_win(pv.win + j0, max(0, min(PR_CHUNK, n - j0)), chunk);
            }
            bool more;
            do {
                const uint32_t clut = flut[chunk[((j - 1) & (PR_CHUNK - 1)) * 64]];
                const int32_t sb = sbc, top = topc;
                int32_t sbn, topn;
                stripe(j + 1, sbn, topn);
                int32_t mc = m, tbc = tbase;
                asm volatile("" : "+v"(mc), "+v"(tbc));
                // row 0: H(0, j) = 0, E(0, j) = NEG
                int32_t Uc = oe, Bc = NEG + ext, ckey = 0;
                uint32_t pdiag = (uint32_t)oe & 0xFFFFu;
                uint32_t *tcol = tr + (size_t)(j - 1) * WR * 64;
                // exit check of row i (H = the low half of its new packed
                // value minus oe): a right or diagonal exit when `side`, a
                // down or diagonal exit when `down`
                auto check = [&](uint32_t pk, int32_t i, bool side, bool down) {
                    const int32_t H = (int32_t)(int16_t)(pk & 0xFFFFu) - oe;
                    fail |= ((side && j < n && i <= mc) || (down && j <= n && i < mc)) && H >= tbc + S * i;
                };
                // head groups: computed at every column; the last one's bottom
                // row exits down unless the stripe continues it at j and j + 1
                const bool below = sb == nhead && top >= sb && sbn == nhead && topn >= sbn;
#pragma unroll
                for (int g = 0; g < PR_HG; g++) {
                    if (g < nhead) {
                        tcol[g * 64] = group8<FOLD>(&Hr[8 * g], rcode + (size_t)(8 * g + 1) * 64, clut, 8 * g, zr,
                                                    noev, oev, extv, Uc, Bc, pdiag, ckey);
                        if (g == nhead - 1 && !below) check(Hr[8 * g + 7], 8 * g + 8, false, true);
                    }
                }
                if (top >= sb) {
                    if (sb != nhead) {
                        // the group above is not computed here: margins, and
                        // the diagonal from its last row if it was computed at
                        // j - 1 (it has just left the window)
                        Uc = oe;
                        Bc = NEG + ext;
                        pdiag = drop;
                    }
#pragma unroll
                    for (int t = 0; t < PR_SG; t++) {
                        if (sb + t <= top) {
                            const int32_t g = sb + t;
                            tcol[(size_t)g * 64] = group8<FOLD>(&Sr[8 * t], rcode + (size_t)(8 * g + 1) * 64, clut,
                                                                8 * g, zr, noev, oev, extv, Uc, Bc, pdiag, ckey);
                            // the top group leaving at j + 1: every row exits
                            // right and diagonally; the bottom group: its last
                            // row exits down
                            const bool all = t == 0 && sbn > sb, last = g == top;
                            if (all || last) {
#pragma unroll
                                for (int r = 0; r < 8; r++)
                                    if (all || r == 7) check(Sr[8 * t + r], 8 * g + r + 1, all, last && r == 7);
                            }
                        }
                    }
                }
                if (j <= n && ckey > bestkey) {
                    bestkey = ckey;
                    bj = j;
                }
                if (a.census) {
                    const uint32_t mk = ((1u << nhead) - 1u) | (top >= sb ? (2u << top) - (1u << sb) : 0u);
                    const uint32_t gm = (uint32_t)m >> 3;
                    const uint32_t rows = 8u * (uint32_t)__builtin_popcount(mk & ((1u << gm) - 1u)) +
                                          (((mk >> gm) & 1u) ? ((uint32_t)m & 7u) : 0u);
                    ccells += j <= n ? rows : 0u;
                    cgroups += (uint32_t)__builtin_popcount(mk);
                }
                drop = (uint32_t)oe & 0xFFFFu;
                sbc = sbn;
                topc = topn;
                j++;
                more = j <= umaxn && sbc == sbw && ((j - 1) & (PR_CHUNK - 1)) != 0;
            } while (more);
        }
        if (a.census) {
            if (active) atomicAdd(&a.census[0], (unsigned long long)ccells);
            if (lane == 0) atomicAdd(&a.census[1], (unsigned long long)cgroups * 8ull * 64ull);
        }
        if (active && fail) {
            a.fb_list[atomicAdd(&a.fb_range[1], 1u)] = p;
        } else if (active) {
            int32_t best = bestkey >> 16, bi = 0xFFFF - (bestkey & 0xFFFF);
            if (best <= 0) {
                best = 0;
                bi = 0;
                bj = 0;
            }
            const uint64_t c0 = a.cig_off[p];
            const uint32_t cap = (uint32_t)(a.cig_off[p + 1] - c0);
            uint32_t *cslot = a.cigar + c0;
            TbResult tb = traceback_full<WR>(pv, P, rcode, tr, best, bi, bj, cslot, cap);
            esp_ext e;
            uint32_t nops;
            finish_pair(pv, P, best, tb, cslot, cap, &e, &nops);
            a.ext[p] = e;
            a.cig_n[p] = nops;
        }
    }
}

template <int MAXM>
size_t full_lds() { return 32 + (size_t)(MAXM + 1) * 64 + (size_t)FULL_CHUNK * 64; }

// The folded LUT when its fields fit, else the plain one plus -oe per cell.
template <int MAXM, bool PRUNE>
hipError_t launch_full_p(const GotohArgs &a, int blocks, hipStream_t s) {
    if (a.flut_folded)
        hipLaunchKernelGGL((k_gotoh_full<MAXM, true, PRUNE>), dim3(blocks), dim3(64), full_lds<MAXM>(), s, a);
    else
        hipLaunchKernelGGL((k_gotoh_full<MAXM, false, PRUNE>), dim3(blocks), dim3(64), full_lds<MAXM>(), s, a);
    return hipGetLastError();
}

// The pruned pass exists for tiles up to 192 rows (at 264 its checks stop the
// group loop from unrolling and the row state would leave the registers).
template <int MAXM>
hipError_t launch_full_t(const GotohArgs &a, bool prune, int blocks, hipStream_t s) {
    if constexpr (MAXM <= FULL_PRUNE_MAX_ROWS)
        if (prune) return launch_full_p<MAXM, true>(a, blocks, s);
    if (prune) return hipErrorInvalidValue;
    return launch_full_p<MAXM, false>(a, blocks, s);
}

template <int MAXM>
size_t pruned_lds() { return 32 + (size_t)(MAXM + 1) * 64 + (size_t)PR_CHUNK * 64; }

template <int MAXM>
hipError_t launch_pruned_t(const GotohArgs &a, int blocks, hipStream_t s) {
    if (a.flut_folded)
        hipLaunchKernelGGL((k_gotoh_pruned<MAXM, true>), dim3(blocks), dim3(64), pruned_lds<MAXM>(), s, a);
    else
        hipLaunchKernelGGL((k_gotoh_pruned<MAXM, false>), dim3(blocks), dim3(64), pruned_lds<MAXM>(), s, a);
    return hipGetLastError();
}

template <int MAXM>
int pruned_blocks_t() {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_gotoh_pruned<MAXM, true>, 64, pruned_lds<MAXM>()) !=
        hipSuccess)
        return 1;
    return nb > 0 ? nb : 1;
}

template <int MAXM>
int full_blocks_t() {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_gotoh_full<MAXM, true, false>, 64, full_lds<MAXM>()) !=
        hipSuccess)
        return 1;
    return nb > 0 ? nb : 1;
}

}  // namespace

int full_rows(uint32_t maxm) {
    return maxm <= 64 ? 64 : maxm <= 128 ? 128 : maxm <= 160 ? 160 : maxm <= 192 ? 192 : 264;
}

hipError_t launch_gotoh_full(const GotohArgs &a, int rows, bool prune, int blocks, hipStream_t s) {
    switch (rows) {
    case 64: return launch_full_t<64>(a, prune, blocks, s);
    case 128: return launch_full_t<128>(a, prune, blocks, s);
    case 160: return launch_full_t<160>(a, prune, blocks, s);
    case 192: return launch_full_t<192>(a, prune, blocks, s);
    default: return launch_full_t<264>(a, prune, blocks, s);
    }
}

hipError_t launch_gotoh_pruned(const GotohArgs &a, int rows, int blocks, hipStream_t s) {
    switch (rows) {
    case 64: return launch_pruned_t<64>(a, blocks, s);
    case 128: return launch_pruned_t<128>(a, blocks, s);
    case 160: return launch_pruned_t<160>(a, blocks, s);
    case 192: return launch_pruned_t<192>(a, blocks, s);
    default: return hipErrorInvalidValue;
    }
}

int gotoh_pruned_blocks_per_cu(int rows) {
    switch (rows) {
    case 64: return pruned_blocks_t<64>();
    case 128: return pruned_blocks_t<128>();
    case 160: return pruned_blocks_t<160>();
    default: return pruned_blocks_t<192>();
    }
}

namespace {

// Escalation sort key: head-group class * ESC_NDIAG + diagonal bucket, or
// ESC_NB for a pair the pruned pass cannot help (no positive band score, no
// positive substitution score, or a head of more than 12 groups or covering
// the read).
__device__ __forceinline__ uint32_t esc_key(const EscSortArgs &a, uint32_t p) {
    const int32_t L = a.esc_info[2 * (size_t)p], d = a.esc_info[2 * (size_t)p + 1];
    const int32_t m = (int32_t)a.read_len[p];
    if (a.s <= 0 || L <= 0) return ESC_NB;
    const int32_t nh = esc_head_groups(m, L, a.s, a.pad);
    if (nh > 12 || 8 * nh >= m) return ESC_NB;
    const int32_t cls = nh <= 3 ? 0 : (nh <= 6 ? 1 : 2);
    const int32_t db = min(max((d + 256) >> 2, 0), ESC_NDIAG - 1);
    return (uint32_t)(cls * ESC_NDIAG + db);
}

// Block b owns the list entries [b * C, (b + 1) * C), C = ceil(count / grid):
// its bucket counts gather in LDS, and each touched bucket costs one global
// atomic per block (the keys of a pipeline crowd into few buckets).
__device__ __forceinline__ void esc_chunk(uint32_t n, uint32_t &lo, uint32_t &hi) {
    const uint32_t c = (n + gridDim.x - 1) / gridDim.x;
    lo = min(n, blockIdx.x * c);
    hi = min(n, lo + c);
}

__global__ __launch_bounds__(256) void k_esc_hist(EscSortArgs a) {
    __shared__ uint32_t h[ESC_NB];
    for (uint32_t b = threadIdx.x; b < ESC_NB; b += blockDim.x) h[b] = 0;
    __syncthreads();
    __shared__ unsigned long long cells;
    if (threadIdx.x == 0) cells = 0;
    __syncthreads();
    uint32_t lo, hi;
    esc_chunk(*a.count, lo, hi);
    unsigned long long mine = 0;
    for (uint32_t t = lo + threadIdx.x; t < hi; t += blockDim.x) {
        const uint32_t p = a.list[t];
        const uint32_t k = esc_key(a, p);
        if (k < ESC_NB) atomicAdd(&h[k], 1u);
        if (a.census) mine += (unsigned long long)a.read_len[p] * a.win_len[p];
    }
    if (a.census && mine) atomicAdd(&cells, mine);
    __syncthreads();
    for (uint32_t b = threadIdx.x; b < ESC_NB; b += blockDim.x)
        if (h[b]) atomicAdd(&a.esc->hist[b], h[b]);
    if (a.census && threadIdx.x == 0 && cells) atomicAdd(&a.census[2], cells);
}

// One block of 1024 threads: exclusive scan of the ESC_NB bucket counts into
// the scatter cursors, three buckets a thread.
__global__ __launch_bounds__(1024) void k_esc_scan(EscSortArgs a) {
    static_assert(ESC_NB == 3 * 1024, "three buckets a thread");
    __shared__ uint32_t part[1024];
    const uint32_t t = threadIdx.x;
    const uint32_t h0 = a.esc->hist[3 * t], h1 = a.esc->hist[3 * t + 1], h2 = a.esc->hist[3 * t + 2];
    part[t] = h0 + h1 + h2;
    __syncthreads();
    for (uint32_t sh = 1; sh < 1024; sh <<= 1) {
        const uint32_t v = t >= sh ? part[t - sh] : 0u;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    const uint32_t base = part[t] - (h0 + h1 + h2);
    a.esc->cursor[3 * t] = base;
    a.esc->cursor[3 * t + 1] = base + h0;
    a.esc->cursor[3 * t + 2] = base + h0 + h1;
    // the first class (heads of at most 3 groups) ends where bucket
    // ESC_NDIAG begins
    __shared__ uint32_t n0;
    static_assert(ESC_NDIAG % 3 == 1, "bucket ESC_NDIAG is a thread's second");
    if (t == ESC_NDIAG / 3) n0 = base + h0;
    __syncthreads();
    if (t == 1023) {
        a.esc->prange[1] = part[t];
        a.esc->crange0[0] = 0;
        a.esc->crange0[1] = n0;
        a.esc->crange1[0] = n0;
        a.esc->crange1[1] = part[t] - n0;
    }
}

// The same chunks as k_esc_hist: count per bucket in LDS, reserve each
// touched bucket's run with one global atomic, then place each entry at its
// run's base plus its rank in LDS.
__global__ __launch_bounds__(256) void k_esc_scatter(EscSortArgs a) {
    __shared__ uint32_t h[ESC_NB];
    __shared__ uint32_t fb_count, fb_base;
    for (uint32_t b = threadIdx.x; b < ESC_NB; b += blockDim.x) h[b] = 0;
    if (threadIdx.x == 0) fb_count = 0;
    __syncthreads();
    uint32_t lo, hi;
    esc_chunk(*a.count, lo, hi);
    for (uint32_t t = lo + threadIdx.x; t < hi; t += blockDim.x) {
        const uint32_t k = esc_key(a, a.list[t]);
        atomicAdd(k < ESC_NB ? &h[k] : &fb_count, 1u);
    }
    __syncthreads();
    for (uint32_t b = threadIdx.x; b < ESC_NB; b += blockDim.x)
        if (h[b]) h[b] = atomicAdd(&a.esc->cursor[b], h[b]);
    if (threadIdx.x == 0) fb_base = fb_count ? atomicAdd(&a.esc->fbrange[1], fb_count) : 0u;
    __syncthreads();
    for (uint32_t t = lo + threadIdx.x; t < hi; t += blockDim.x) {
        const uint32_t p = a.list[t];
        const uint32_t k = esc_key(a, p);
        if (k < ESC_NB)
            a.plist[atomicAdd(&h[k], 1u)] = p;
        else
            a.fblist[atomicAdd(&fb_base, 1u)] = p;
    }
}

}  // namespace

hipError_t launch_esc_sort(const EscSortArgs &a, uint32_t bound, hipStream_t s) {
    const uint32_t blocks = std::max(1u, std::min(512u, (bound + 1023) / 1024));
    hipLaunchKernelGGL(k_esc_hist, dim3(blocks), dim3(256), 0, s, a);
    hipLaunchKernelGGL(k_esc_scan, dim3(1), dim3(1024), 0, s, a);
    hipLaunchKernelGGL(k_esc_scatter, dim3(blocks), dim3(256), 0, s, a);
    return hipGetLastError();
}

int gotoh_full_blocks_per_cu(int rows) {
    switch (rows) {
    case 64: return full_blocks_t<64>();
    case 128: return full_blocks_t<128>();
    case 160: return full_blocks_t<160>();
    case 192: return full_blocks_t<192>();
    default: return full_blocks_t<264>();
    }
}

}  // namespace espk
