// esp_junction.hip -- read-aware junction refinement and what is built on it:
// the chain preparation of align_spliced and the re-extension of soft-clipped
// placements.
//
// Semantics restated from the reference (crates/map/src/):
//   refine_junction_read  splice.rs:208-317  boundary shift over the read gap, flank match, signal, library
//   splice_signal         splice.rs:75-104   plus tests and their minus-strand mirrors
//   dinuc                 splice.rs:330-347  donor [p, p+2), acceptor [p-2, p), N past the contig
//   splice_support_pass   splice.rs:107-115
//   segment_chain         splice.rs:364-396
//   prepare_spliced       splice.rs:545-621  guards and the pseudo-reference bounds
//   junction_signal       splice.rs:864-868  (the intron-chain fallback, intron_chain.rs:475-486)
//   reextend_with         align.rs:1170-1225 window, hint, acceptance
//
// One lane owns one query, chain or placement: each is a short serial walk
// (at most 25 boundary shifts of 8 base comparisons; a chain's anchors), so
// the cost is a few memory round trips per item; window decodes run a wave
// per window as in the window pipelines.
#include <algorithm>

#include "esp_kernels.hpp"
#include "esp_refdev.hpp"

using namespace espd;

namespace espk {

namespace {

constexpr int JBLOCK = 256;

inline unsigned lane_blocks(size_t n) { return (unsigned)((n + JBLOCK - 1) / JBLOCK); }
inline unsigned wave_blocks(size_t n) { return (unsigned)((n + JBLOCK / 64 - 1) / (JBLOCK / 64)); }

__device__ __forceinline__ uint32_t up_ascii(uint32_t b) { return (b >= 'a' && b <= 'z') ? b - 32 : b; }

// splice_signal: 0 GT-AG, 1 GC-AG, 2 AT-AC, 3 non-canonical.
__device__ int splice_sig(uint32_t d0, uint32_t d1, uint32_t a0, uint32_t a1, bool minus) {
    d0 = up_ascii(d0);
    d1 = up_ascii(d1);
    a0 = up_ascii(a0);
    a1 = up_ascii(a1);
    if (!minus) {
        if (d0 == 'G' && d1 == 'T' && a0 == 'A' && a1 == 'G') return 0;
        if (d0 == 'G' && d1 == 'C' && a0 == 'A' && a1 == 'G') return 1;
        if (d0 == 'A' && d1 == 'T' && a0 == 'A' && a1 == 'C') return 2;
        return 3;
    }
    if (d0 == 'C' && d1 == 'T' && a0 == 'A' && a1 == 'C') return 0;
    if (d0 == 'C' && d1 == 'T' && a0 == 'G' && a1 == 'C') return 1;
    if (d0 == 'G' && d1 == 'T' && a0 == 'A' && a1 == 'T') return 2;
    return 3;
}

__device__ __forceinline__ int32_t sig_bonus(int sig, const SpliceCfgDev &sp) {
    return sig == 0 ? sp.gt_ag : sig == 1 ? sp.gc_ag : sig == 2 ? sp.at_ac : -sp.noncanon;
}

// dinuc: the donor dinucleotide at p or the acceptor one ending at p.
__device__ void dinuc(const RefView &R, uint32_t ctg, uint32_t p, bool donor, uint32_t &c0, uint32_t &c1) {
    c0 = c1 = 'N';
    if (!donor && p < 2) return;
    const uint32_t len = R.len[ctg];
    const uint32_t a = donor ? p : p - 2;
    if (a >= len) return;
    const uint64_t b = min((uint64_t)a + 2, (uint64_t)len);
    c0 = base_ascii(R, ctg, a);
    if ((uint64_t)a + 1 < b) c1 = base_ascii(R, ctg, a + 1);
}

// Base::from_ascii code (A0 C1 G2 T3, every other byte N = 4) and edit_equiv
// (splice.rs:320-325): equal bases (N equals N) or A/G, T/C either way.
__device__ __forceinline__ int32_t base4(uint32_t b) { return base_code_or(b, 4); }
__device__ __forceinline__ bool edit_equiv(int32_t a, int32_t b) {
    return a == b || (a ^ b) == 2;  // A0^G2 and C1^T3
}

// JunctionLib::support through the hash set: the stranded count, 0 if absent.
__device__ uint32_t jsupport(const JunctionView &J, uint32_t c, uint32_t s, uint32_t e, bool minus) {
    if (!J.jn) return 0u;
    uint32_t slot = junction_hash(c, s, e) & J.jmask;
    while (true) {
        const uint4 v = J.jn[slot];
        if (!v.w) return 0u;
        if (v.x == c && v.y == s && v.z == e) {
            const uint2 n = J.cnt[slot];
            return minus ? n.y : n.x;
        }
        slot = (slot + 1) & J.jmask;
    }
}

// refine_junction_read.  The reference indexes read[q1] only through its
// own bounds; a query position past the read (an anchor beyond it) is
// skipped here instead of read.
__device__ esp_refined_junction refine_read(const RefView &R, const JunctionView &J, const SpliceCfgDev &sp,
                                            const uint8_t *read, uint32_t read_len, uint32_t ctg, uint32_t ns,
                                            uint32_t ne, uint32_t qa, uint32_t qbv, bool minus) {
    esp_refined_junction best;
    best.found = 0;
    best.start = best.end = best.minus = best.support = 0;
    best.signal = 0;
    if (ctg >= R.ncontig) return best;
    const uint32_t len = R.len[ctg];
    const int64_t ref_gap = (int64_t)ne - (int64_t)ns, read_gap = (int64_t)qbv - (int64_t)qa;
    if (read_gap > 20) return best;
    const int64_t il = ref_gap - read_gap;
    if (il < (int64_t)sp.min_intron || il > (int64_t)sp.max_intron) return best;
    const int64_t lo = min(read_gap, (int64_t)0) - 2, hi = max(read_gap, (int64_t)0) + 2;
    int32_t bs = 0;
    for (int64_t delta = lo; delta <= hi; delta++) {
        const int64_t s = (int64_t)ns + delta;
        if (s < 0) continue;
        const int64_t e = s + il;
        if (e < 2 || e > (int64_t)len) continue;
        const int64_t qb = (int64_t)qa + delta;
        if (qb < 1 || (uint32_t)qb >= read_len) continue;
        int32_t mt = 0;
#pragma unroll
        for (int64_t i = 1; i <= 4; i++) {
            const int64_t q1 = qb - i, g1 = s - i;
            if (q1 >= 0 && g1 >= 0 && (uint64_t)q1 < read_len &&
                edit_equiv(base4(read[q1]), base4(base_ascii(R, ctg, (uint32_t)g1))))
                mt++;
            const int64_t q2 = qb + i - 1, g2 = e + i - 1;
            if ((uint32_t)q2 < read_len && (uint64_t)q2 < read_len && g2 < (int64_t)len &&
                edit_equiv(base4(read[q2]), base4(base_ascii(R, ctg, (uint32_t)g2))))
                mt++;
        }
        uint32_t d0, d1, a0, a1;
        dinuc(R, ctg, (uint32_t)s, true, d0, d1);
        dinuc(R, ctg, (uint32_t)e, false, a0, a1);
        const int32_t ad = (int32_t)(delta < 0 ? -delta : delta);
        for (int t = 0; t < 2; t++) {
            const bool tm = t == 0 ? minus : !minus;
            const int sig = splice_sig(d0, d1, a0, a1, tm);
            const uint32_t support = jsupport(J, ctg, (uint32_t)s, (uint32_t)e, tm);
            int32_t score = mt * 50 + sig_bonus(sig, sp) * 5 - ad * 5;
            if (support > 0) score += sp.known;
            if (!best.found || score > bs) {
                bs = score;
                best.found = 1;
                best.start = (uint32_t)s;
                best.end = (uint32_t)e;
                best.minus = tm ? 1u : 0u;
                best.signal = sig;
                best.support = support;
            }
        }
    }
    return best;
}

__global__ __launch_bounds__(JBLOCK) void k_refine(RefineArgs a) {
    const uint32_t k = blockIdx.x * JBLOCK + threadIdx.x;
    if (k >= a.n) return;
    const esp_junction_query q = a.q[k];
    esp_refined_junction r = refine_read(a.R, a.J, a.sp, a.reads + q.read_off, q.read_len, q.contig,
                                         q.naive_start, q.naive_end, q.q_a, q.q_b, q.minus != 0);
    if (!r.found && q.use_fallback && q.contig < a.R.ncontig) {
        // intron_chain.rs:475-486: the fragment pair's own junction, with its
        // junction_signal (splice.rs:864-868) and library support.
        uint32_t d0, d1, a0, a1;
        dinuc(a.R, q.contig, q.fallback_start, true, d0, d1);
        dinuc(a.R, q.contig, q.fallback_end, false, a0, a1);
        r.found = 2;
        r.start = q.fallback_start;
        r.end = q.fallback_end;
        r.minus = q.minus ? 1u : 0u;
        r.signal = splice_sig(d0, d1, a0, a1, q.minus != 0);
        r.support = jsupport(a.J, q.contig, q.fallback_start, q.fallback_end, q.minus != 0);
    }
    a.out[k] = r;
}

// prepare_spliced up to the pseudo-reference bounds, per chain.
__global__ __launch_bounds__(JBLOCK) void k_chain_prep(ChainPrepArgs a) {
    const uint32_t k = blockIdx.x * JBLOCK + threadIdx.x;
    if (k >= a.npairs) return;
    const uint64_t a0 = a.anc_off[k], na = a.anc_off[k + 1] - a0;
    const uint32_t ctg = a.contig[k], m = a.read_len[k], K = a.k;
    const bool minus = a.minus[k] != 0;
    const uint32_t *anc = a.anchors + 2 * a0;
    uint32_t *segs = a.segs + 4 * a0;
    // segment_chain
    uint32_t ns = 0;
    if (na > 0) {
        uint32_t c0 = anc[0], c1 = anc[0] + K, c2 = anc[1], c3 = anc[1] + K;
        for (uint64_t t = 1; t < na; t++) {
            const uint32_t q1 = anc[2 * t - 2], r1 = anc[2 * t - 1], q2 = anc[2 * t], r2 = anc[2 * t + 1];
            const uint32_t ref_gap = r2 > r1 + K ? r2 - (r1 + K) : 0u;
            const uint32_t read_gap = q2 > q1 + K ? q2 - (q1 + K) : 0u;
            if (ref_gap > a.sp.min_intron && read_gap <= 10u) {
                segs[4 * ns] = c0;
                segs[4 * ns + 1] = c1;
                segs[4 * ns + 2] = c2;
                segs[4 * ns + 3] = c3;
                ns++;
                c0 = q2;
                c1 = q2 + K;
                c2 = r2;
                c3 = r2 + K;
            } else {
                c1 = q2 + K;
                c3 = r2 + K;
            }
        }
        segs[4 * ns] = c0;
        segs[4 * ns + 1] = c1;
        segs[4 * ns + 2] = c2;
        segs[4 * ns + 3] = c3;
        ns++;
    }
    int32_t state = 1, refusal = 0;
    if (ns < 2) {
        state = 0;
        refusal = 1;
    } else if (ctg >= a.R.ncontig) {
        state = -1;
    }
    const uint8_t *read = a.reads + a.read_off[k];
    esp_refined_junction *jr = a.jrec + a0;
    for (uint32_t w = 0; state == 1 && w + 1 < ns; w++) {
        const esp_refined_junction r = refine_read(a.R, a.J, a.sp, read, m, ctg, segs[4 * w + 3],
                                                   segs[4 * (w + 1) + 2], segs[4 * w + 1], segs[4 * (w + 1)], minus);
        jr[w] = r;
        if (!r.found) {
            state = 0;
            refusal = 2;
            break;
        }
        const uint32_t il = r.end - r.start;
        if (il < a.sp.min_intron || il > a.sp.max_intron) {
            state = 0;
            refusal = 3;
        } else if ((il < 50u && r.support == 0) || (r.signal == 3 && r.support == 0)) {
            state = 0;
            refusal = 4;
        }
    }
    uint64_t total = 0;
    uint32_t *bnd = a.bounds + 3 * a0;
    for (uint32_t i = 0; state == 1 && i < ns; i++) {
        const uint32_t rs = i == 0 ? (segs[2] > 30u ? segs[2] - 30u : 0u) : jr[i - 1].end;
        const uint32_t re = i + 1 == ns ? (uint32_t)min((uint64_t)segs[4 * i + 3] + 30u, (uint64_t)a.R.len[ctg])
                                        : jr[i].start;
        if (re <= rs) {
            state = 0;
            refusal = 5;
            break;
        }
        bnd[3 * i] = (uint32_t)total;
        bnd[3 * i + 1] = rs;
        bnd[3 * i + 2] = re;
        total += re - rs;
    }
    a.state[k] = state;
    a.refusal[k] = refusal;
    a.nseg[k] = ns;
    a.nex[k] = state == 1 ? ns : 0u;
    a.plen[k] = state == 1 ? total : 0ull;
}

// Compacted exon list and zero hints for the spliced pipeline.
__global__ __launch_bounds__(JBLOCK) void k_chain_exons(ChainPrepArgs a) {
    const uint32_t k = blockIdx.x * JBLOCK + threadIdx.x;
    if (k >= a.npairs) return;
    a.hint[k] = 0;  // align_spliced extends at DiagHint offset 0
    const uint64_t e0 = a.exon_off[k], ne = a.exon_off[k + 1] - e0;
    const uint32_t *bnd = a.bounds + 3 * a.anc_off[k];
    for (uint64_t i = 0; i < ne; i++) {
        a.exons[2 * (e0 + i)] = bnd[3 * i + 1];
        a.exons[2 * (e0 + i) + 1] = bnd[3 * i + 2];
    }
}

__global__ __launch_bounds__(JBLOCK) void k_chain_rec(ChainPrepArgs a) {
    const uint32_t k = blockIdx.x * JBLOCK + threadIdx.x;
    if (k >= a.npairs) return;
    esp_spliced_chain_rec r;
    r.status = a.state[k];
    r.refusal = a.refusal[k];
    r.n_segments = a.nseg[k];
    r.genomic_start = r.pseudo_len = r.n_stitched = 0;
    if (r.status == 1) {
        const esp_splice_rec s = a.srec[k];
        r.status = s.status;
        r.genomic_start = s.genomic_start;
        r.pseudo_len = s.pseudo_len;
        r.n_stitched = s.n_stitched;
    }
    a.rec[k] = r;
}

// reextend_with, stage 1: the CIGAR's clips, span and aligned count, the
// window and the hint.
__global__ __launch_bounds__(JBLOCK) void k_reext_prep(ReextArgs a) {
    const uint32_t k = blockIdx.x * JBLOCK + threadIdx.x;
    if (k >= a.n) return;
    const uint32_t *cg = a.cigar + a.cig_off[k];
    const uint32_t nops = a.cig_n[k], ctg = a.contig[k], pos = a.pos[k];
    uint32_t lead = 0, tail = 0, span = 0, aligned = 0;
    bool seen_m = false, spliced = false;
    for (uint32_t t = 0; t < nops; t++) {
        const uint32_t w = cg[t], op = w & 0xFu, n = w >> 4;
        if (op == OP_N) spliced = true;
        else if (op == OP_S) (seen_m ? tail : lead) += n;
        else if (op == OP_M) {
            seen_m = true;
            span += n;
            aligned += n;
        } else if (op == OP_I) aligned += n;
        else if (op == OP_D) span += n;
    }
    esp_reextend_rec r;
    r.status = 0;
    r.pos = pos;
    r.score = 0;
    r.wstart = r.wlen = 0;
    r.aligned = 0;
    r.new_aligned = 0;
    int32_t state = 1;
    if (spliced) state = 2;
    else if (ctg >= a.R.ncontig) state = -1;
    uint32_t ws = 0, we = 0;
    if (state == 1) {
        const uint32_t back = lead + 30u;
        ws = pos > back ? pos - back : 0u;
        we = min(pos + span + tail + 30u, a.R.len[ctg]);  // u32 arithmetic, as the reference's
        if (we <= ws) state = 3;
        r.aligned = aligned;
    }
    if (state == 1) {
        r.wstart = ws;
        r.wlen = we - ws;
    }
    r.status = state == 1 ? 0 : state;
    a.rec[k] = r;
    a.state[k] = state;
    a.wlen[k] = state == 1 ? we - ws : 0u;
    a.g_read_len[k] = state == 1 ? a.read_len[k] : 0u;
    a.hint[k] = state == 1 ? (int64_t)(pos - ws) - (int64_t)lead : 0;
}

__global__ __launch_bounds__(JBLOCK) void k_reext_gather(ReextArgs a) {
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t k = blockIdx.x * (JBLOCK / 64) + (threadIdx.x >> 6);
    if (k >= a.n) return;
    const uint32_t wl = a.wlen[k];
    if (wl == 0) return;
    const uint32_t ws = a.rec[k].wstart;
    wave_decode(a.R, a.contig[k], ws, ws + wl, false, a.arena + a.win_off[k], lane);
}

// reextend_with, stage 3: the acceptance rule on the new extension.
__global__ __launch_bounds__(JBLOCK) void k_reext_finish(ReextArgs a) {
    const uint32_t k = blockIdx.x * JBLOCK + threadIdx.x;
    if (k >= a.n) return;
    if (a.state[k] != 1) {
        a.cig_out_n[k] = 0;
        return;
    }
    esp_reextend_rec r = a.rec[k];
    const esp_ext e = a.ext[k];
    const uint32_t *cg = a.cig_out + a.cig_out_off[k];
    const uint32_t nops = a.cig_out_n[k];
    uint32_t na = 0;
    for (uint32_t t = 0; t < nops; t++) {
        const uint32_t op = cg[t] & 0xFu;
        if (op == OP_M || op == OP_I) na += cg[t] >> 4;
    }
    r.score = e.score;
    r.new_aligned = na;
    if (e.score > 0 && na >= r.aligned && na <= a.read_len[k]) {
        r.status = 1;
        r.pos = r.wstart + e.ref_start;
    }
    a.rec[k] = r;
}

// ---- try_intron_chain_placement (intron_chain.rs:392-538) -------------------

// base_eq (intron_chain.rs:113-117): the same upper-case base, A C G T only.
__device__ __forceinline__ bool ic_base_eq(uint32_t a, uint32_t b) {
    a &= 0xDFu;
    b &= 0xDFu;
    return a == b && (a == 'A' || a == 'C' || a == 'G' || a == 'T');
}

// MEM fields: contig, strand, q_lo, q_hi, r_lo, r_hi.  Order key of
// cluster_frags (:174-183): (contig, strand, q_lo - r_lo, q_lo, r_lo).
__device__ __forceinline__ bool mem_after(const uint32_t *x, const uint32_t *y) {
    if (x[0] != y[0]) return x[0] > y[0];
    if (x[1] != y[1]) return x[1] > y[1];
    const int64_t dx = (int64_t)x[2] - (int64_t)x[4], dy = (int64_t)y[2] - (int64_t)y[4];
    if (dx != dy) return dx > dy;
    if (x[2] != y[2]) return x[2] > y[2];
    return x[4] > y[4];
}

// First N interval of contig ctg ending after p (its index, or the end of
// the contig's intervals).
__device__ __forceinline__ uint64_t iv_after(const RefView &R, uint32_t ctg, uint32_t p) {
    uint64_t lo = R.iv_off[ctg], hi = R.iv_off[ctg + 1];
    while (lo < hi) {
        const uint64_t mid = (lo + hi) >> 1;
        if (R.iv[2 * mid + 1] <= p) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// The extension loops of extend_mems (intron_chain.rs:137-152), 16 bases a
// step.  A step loads the read's 16 bytes as five dwords and the contig's 16
// bases as two packed dwords, spells the bases as ASCII through an LDS table
// (packed byte -> four ASCII bytes) and compares four bytes at a time.  The
// contig's N intervals bound the extension once per side (an N base never
// equals a read base), so the bases compared are A, C, G or T and base_eq is
// the read byte's upper case equal to them.
constexpr uint32_t MEM_STEP = 16;

struct MemSeq {
    const uint32_t *rw;  // the read's bytes as the dwords holding them
    uint32_t o0;         // byte offset of the read's first byte in rw[0]
    uint32_t rjmax;      // last dword of rw holding a read byte
    const uint32_t *pk;  // the contig's packed bytes as dwords
    uint32_t pklast;     // last dword of the contig's region (its length / 4 + 16 bytes)
    const uint32_t *lut; // LDS: packed byte -> its four bases in ASCII
};

// Read bytes [q, q + 16) (dwords past the read's last repeat it).
__device__ __forceinline__ void mem_read16(const MemSeq &M, uint32_t q, uint32_t (&w)[4]) {
    const uint32_t b = M.o0 + q, j0 = b >> 2;
    uint32_t d[5];
#pragma unroll
    for (uint32_t u = 0; u < 5; u++) d[u] = M.rw[min(j0 + u, M.rjmax)];
#pragma unroll
    for (uint32_t k = 0; k < 4; k++) w[k] = __builtin_amdgcn_alignbyte(d[k + 1], d[k], b & 3u);
}

// Contig bases [p, p + 16) in ASCII.
__device__ __forceinline__ void mem_ref16(const MemSeq &M, uint32_t p, uint32_t (&w)[4]) {
    const uint32_t pb = p >> 2;  // packed byte of base p
    const uint32_t D0 = M.pk[min(pb >> 2, M.pklast)], D1 = M.pk[min((pb >> 2) + 1, M.pklast)];
    const uint32_t sh = pb & 3u;
    const uint32_t B = __builtin_amdgcn_alignbyte(D1, D0, sh);  // packed bytes pb .. pb + 3
    const uint32_t B4 = (D1 >> (8 * sh)) & 0xFFu;                // packed byte pb + 4
    uint32_t a[5];
#pragma unroll
    for (uint32_t k = 0; k < 4; k++) a[k] = M.lut[(B >> (8 * k)) & 0xFFu];
    a[4] = M.lut[B4];
#pragma unroll
    for (uint32_t k = 0; k < 4; k++) w[k] = __builtin_amdgcn_alignbyte(a[k + 1], a[k], p & 3u);
}

// Right: while q < rl, r < win_hi and read[q] equals base r.
__device__ void mem_extend_right(const RefView &R, const MemSeq &M, uint32_t ctg, uint32_t rl, uint32_t win_hi,
                                 uint32_t &q, uint32_t &r) {
    uint32_t stop = win_hi;
    const uint64_t t = iv_after(R, ctg, r);
    if (t < R.iv_off[ctg + 1]) stop = min(stop, max(R.iv[2 * t], r));  // an N base at or after r
    while (q < rl && r < stop) {
        const uint32_t n = min(min(rl - q, stop - r), MEM_STEP);
        uint32_t rb[4], cb[4];
        mem_read16(M, q, rb);
        mem_ref16(M, r, cb);
        uint32_t k = n;
        bool found = false;
#pragma unroll
        for (uint32_t w = 0; w < 4; w++) {
            const uint32_t x = (rb[w] & 0xDFDFDFDFu) ^ cb[w];
            if (!found && x) {
                k = min(k, 4 * w + ((uint32_t)__builtin_ctz(x) >> 3));
                found = true;
            }
        }
        q += k;
        r += k;
        if (k < MEM_STEP) break;
    }
}

// Left: while q > 0, r > win_lo and read[q - 1] equals base r - 1.
__device__ void mem_extend_left(const RefView &R, const MemSeq &M, uint32_t ctg, uint32_t win_lo, uint32_t &q,
                                uint32_t &r) {
    if (q == 0 || r <= win_lo) return;
    const uint64_t a0 = R.iv_off[ctg];
    const uint64_t t = iv_after(R, ctg, r - 1);
    if (t < R.iv_off[ctg + 1] && R.iv[2 * t] <= r - 1) return;  // base r - 1 is N
    // bases from the end of the last interval before r - 1 on are not N
    const uint32_t floor_r = max(win_lo, t > a0 ? R.iv[2 * (t - 1) + 1] : 0u);
    while (q > 0 && r > floor_r) {
        // the L bytes before q against the L bases before r, compared from
        // the top byte down (bytes of the window from L on are masked)
        const uint32_t L = min(min(q, r), MEM_STEP), n = min(L, r - floor_r);
        uint32_t rb[4], cb[4];
        mem_read16(M, q - L, rb);
        mem_ref16(M, r - L, cb);
        uint32_t k = n;
        bool found = false;
#pragma unroll
        for (int w = 3; w >= 0; w--) {
            const uint32_t lo = 4u * (uint32_t)w;
            const uint32_t mk = L >= lo + 4 ? 0xFFFFFFFFu : (L <= lo ? 0u : (1u << (8 * (L - lo))) - 1u);
            const uint32_t x = ((rb[w] & 0xDFDFDFDFu) ^ cb[w]) & mk;
            if (!found && x) {
                const uint32_t hb = lo + ((31u - (uint32_t)__builtin_clz(x)) >> 3);
                k = min(k, L - 1 - hb);
                found = true;
            }
        }
        q -= k;
        r -= k;
        if (k < MEM_STEP) break;
    }
}

// Frag fields: contig, strand, q_lo, q_hi, r_lo, r_hi, first member, members.
__device__ uint32_t ic_cov(const uint32_t *f, const uint32_t *mems) {
    uint32_t total = 0, end = 0;
    for (uint32_t i = 0; i < f[7]; i++) {
        const uint32_t *m = mems + 6 * (f[6] + i);
        if (m[3] > end) {
            total += m[3] - max(m[2], end);
            end = m[3];
        }
    }
    return total;
}

// Stage 0: MEMs (extend_mems :122-166), fragments (cluster_frags :171-214),
// the min_frag_cov filter and dominant check (:405-418), the best pair
// (best_frag_pair :233-277) and the two refinement searches' geometry
// (refine_donor_end :285-309, refine_acceptor_start :318-342).
__global__ __launch_bounds__(JBLOCK) void k_ic_frags(IcArgs a) {
    __shared__ uint32_t lut[256];
    static_assert(JBLOCK >= 256, "one table entry a thread");
    if (threadIdx.x < 256) {
        const uint32_t b = threadIdx.x;
        uint32_t v = 0;
        for (int t = 0; t < 4; t++) v |= ((ACGT_ASCII >> (8 * ((b >> (6 - 2 * t)) & 3u))) & 0xFFu) << (8 * t);
        lut[b] = v;
    }
    __syncthreads();
    const uint32_t k = blockIdx.x * JBLOCK + threadIdx.x;
    if (k >= a.n) return;
    const uint64_t a0 = a.anc_off[k], na = a.anc_off[k + 1] - a0;
    const uint32_t rl = a.read_len[k], K = a.k;
    const uint8_t *read = a.reads + a.read_off[k];
    uint32_t *mems = a.mems + 6 * a0;
    uint32_t *frags = a.frags + 8 * a0;
    IcState S;
    S.reason = 0;
    S.ctg = S.minus = S.junction_off = 0;
    S.l_q_lo = S.l_q_hi = S.l_r_lo = S.l_r_hi = S.l_nm = 0;
    S.r_q_lo = S.r_q_hi = S.r_r_lo = S.r_r_hi = S.r_nm = 0;
    S.istart = S.iend = 0;
    S.d_run = S.a_run = S.d_lo = S.a_lo = S.v_run = 0;
    S.donor_end = S.acceptor_start = 0;
    // both searches default to empty jobs
    a.p_off[k] = a.p_off[a.n + k] = a.read_off[k];
    a.p_len[k] = a.p_len[a.n + k] = 0;
    a.t_ctg[k] = a.t_ctg[a.n + k] = 0;
    a.t_lo[k] = a.t_lo[a.n + k] = 0;
    a.t_len[k] = a.t_len[a.n + k] = 0;
    if (na == 0 || rl == 0 || rl > a.max_read) {
        S.reason = rl > a.max_read ? -1 : 1;  // -1: past the batch's max_read_len
        a.st[k] = S;
        return;
    }
    uint32_t nm = 0;
    for (uint64_t t = 0; t < na; t++) {
        const uint32_t *an = a.anchors + 4 * (a0 + t);
        const uint32_t qpos = an[0], rpos = an[1], ctg = an[2];
        if (ctg >= a.R.ncontig) continue;
        const uint32_t clen = a.R.len[ctg];
        const uint32_t win_lo = rpos > rl ? rpos - rl : 0u;
        const uint32_t win_hi = (uint32_t)min((uint64_t)rpos + K + rl, (uint64_t)clen);
        if (win_hi <= win_lo) continue;
        MemSeq M;
        M.rw = (const uint32_t *)((uintptr_t)read & ~(uintptr_t)3);
        M.o0 = (uint32_t)((uintptr_t)read & 3u);
        M.rjmax = (M.o0 + rl - 1) >> 2;
        M.pk = (const uint32_t *)(a.R.packed + a.R.pk_off[ctg]);
        M.pklast = ((((clen + 3u) >> 2) + 16u + 15u) & ~15u) / 4u - 1u;
        M.lut = lut;
        uint32_t q_lo = qpos, r_lo = rpos;
        // q_lo <= rl, r_lo <= win_hi: an anchor past the read or the contig
        // does not extend (the reference would index out of its slices)
        if (q_lo <= rl && r_lo <= win_hi) mem_extend_left(a.R, M, ctg, win_lo, q_lo, r_lo);
        uint32_t q_hi = qpos + K, r_hi = rpos + K;
        mem_extend_right(a.R, M, ctg, rl, win_hi, q_hi, r_hi);
        if (q_hi > q_lo) {
            uint32_t x[6] = {ctg, an[3] ? 1u : 0u, q_lo, q_hi, r_lo, r_hi};
            // stable insertion into the sorted prefix
            uint32_t j = nm;
            while (j > 0 && mem_after(mems + 6 * (j - 1), x)) {
                for (int f = 0; f < 6; f++) mems[6 * j + f] = mems[6 * (j - 1) + f];
                j--;
            }
            for (int f = 0; f < 6; f++) mems[6 * j + f] = x[f];
            nm++;
        }
    }
    uint32_t nf = 0;
    for (uint32_t i = 0; i < nm; i++) {
        const uint32_t *x = mems + 6 * i;
        bool start_new = true;
        if (nf > 0) {
            const uint32_t *f = frags + 8 * (nf - 1);
            if (f[0] == x[0] && f[1] == x[1])
                start_new = ((int64_t)x[2] - (int64_t)x[4]) - ((int64_t)f[2] - (int64_t)f[4]) > 5;
        }
        if (start_new) {
            uint32_t *f = frags + 8 * nf++;
            for (int q = 0; q < 6; q++) f[q] = x[q];
            f[6] = i;
            f[7] = 1;
        } else {
            uint32_t *f = frags + 8 * (nf - 1);
            f[2] = min(f[2], x[2]);
            f[3] = max(f[3], x[3]);
            f[4] = min(f[4], x[4]);
            f[5] = max(f[5], x[5]);
            f[7]++;
        }
    }
    for (uint32_t g = 0; g < nf; g++) {  // members by q_lo, stable
        const uint32_t m0 = frags[8 * g + 6], mn = frags[8 * g + 7];
        for (uint32_t i = m0 + 1; i < m0 + mn; i++) {
            uint32_t x[6];
            for (int q = 0; q < 6; q++) x[q] = mems[6 * i + q];
            uint32_t j = i;
            while (j > m0 && mems[6 * (j - 1) + 2] > x[2]) {
                for (int q = 0; q < 6; q++) mems[6 * j + q] = mems[6 * (j - 1) + q];
                j--;
            }
            for (int q = 0; q < 6; q++) mems[6 * j + q] = x[q];
        }
    }
    uint32_t nk = 0, dom = 0;
    for (uint32_t g = 0; g < nf; g++) {
        const uint32_t cv = ic_cov(frags + 8 * g, mems);
        if (cv >= a.min_frag_cov) {
            for (int q = 0; q < 8; q++) frags[8 * nk + q] = frags[8 * g + q];
            nk++;
            dom = max(dom, cv);
        }
    }
    if (nk == 0) S.reason = 2;
    else if ((uint64_t)dom * 20 >= (uint64_t)rl * 11) S.reason = 3;
    bool have = false;
    uint32_t bi = 0, bj = 0;
    int32_t bscore = 0;
    for (uint32_t i = 0; S.reason == 0 && i < nk; i++) {
        const uint32_t *fa = frags + 8 * i;
        const uint32_t ca = ic_cov(fa, mems);
        for (uint32_t j = 0; j < nk; j++) {
            if (i == j) continue;
            const uint32_t *fb = frags + 8 * j;
            if (fa[0] != fb[0] || fa[1] != fb[1] || fa[3] > fb[2] + 3u) continue;
            const uint32_t sp = min(fa[3], fb[2]);
            const int64_t off_a = (int64_t)fa[5] - (int64_t)fa[3], off_b = (int64_t)fb[4] - (int64_t)fb[2];
            const int64_t len0 = off_b - off_a;
            if (len0 < 20 || len0 > (int64_t)a.intron_max) continue;
            const uint32_t cb = ic_cov(fb, mems);
            const int64_t bal = (int64_t)ca - (int64_t)cb;
            const int32_t score = (int32_t)((ca + cb) * 4u) + (int32_t)((fa[7] + fb[7]) * 3u) -
                                  (int32_t)(bal < 0 ? -bal : bal);
            if (!have || score > bscore) {
                have = true;
                bscore = score;
                bi = i;
                bj = j;
                S.istart = (uint32_t)((int64_t)sp + off_a);
                S.iend = (uint32_t)((int64_t)sp + off_b);
            }
        }
    }
    if (S.reason == 0 && !have) S.reason = 4;
    if (S.reason == 0 && frags[8 * bi] >= a.R.ncontig) S.reason = 5;
    if (S.reason) {
        S.istart = S.iend = 0;
        a.st[k] = S;
        return;
    }
    const uint32_t *L = frags + 8 * bi, *Rf = frags + 8 * bj;
    S.ctg = L[0];
    S.minus = L[1];
    S.l_q_lo = L[2];
    S.l_q_hi = L[3];
    S.l_r_lo = L[4];
    S.l_r_hi = L[5];
    S.l_nm = L[7];
    S.r_q_lo = Rf[2];
    S.r_q_hi = Rf[3];
    S.r_r_lo = Rf[4];
    S.r_r_hi = Rf[5];
    S.r_nm = Rf[7];
    const uint32_t j_est = rl / 2;
    S.junction_off = S.l_q_hi >= S.r_q_lo ? j_est : min(max(j_est, S.l_q_hi), S.r_q_lo);
    const uint32_t clen = a.R.len[S.ctg];
    const int64_t w = a.refine_window;
    const uint32_t plen = a.refine_pattern_len;
    const uint32_t off = min(S.junction_off, rl);
    {  // refine_donor_end
        const uint32_t lo = (uint32_t)max((int64_t)S.l_r_hi - w, (int64_t)0);
        const uint32_t hi = (uint32_t)min((int64_t)S.l_r_hi + w, (int64_t)clen);
        S.donor_end = S.l_r_hi > 0 ? S.l_r_hi - 1 : 0;
        if (!(hi <= lo || off == 0)) {
            const int64_t tl = (int64_t)lo - ((int64_t)S.l_r_lo - (int64_t)S.l_q_lo) + 2;
            const uint64_t text_lo_in_read = (uint64_t)max(tl, (int64_t)0);
            const uint64_t pat_start = max((uint64_t)(off > plen ? off - plen : 0u), text_lo_in_read);
            if (pat_start < off) {
                S.d_run = 1;
                S.d_lo = lo;
                a.p_off[k] = a.read_off[k] + pat_start;
                a.p_len[k] = (uint32_t)(off - pat_start);
                a.t_ctg[k] = S.ctg;
                a.t_lo[k] = lo;
                a.t_len[k] = hi - lo;
            }
        }
    }
    {  // refine_acceptor_start
        const uint32_t lo = (uint32_t)max((int64_t)S.r_r_lo - w, (int64_t)0);
        const uint32_t hi = (uint32_t)min((int64_t)S.r_r_lo + w, (int64_t)clen);
        S.acceptor_start = S.r_r_lo;
        if (hi > lo) {
            const int64_t th = (int64_t)hi - ((int64_t)S.r_r_lo - (int64_t)S.r_q_lo) - 2;
            const uint64_t text_hi_in_read = (uint64_t)max(th, (int64_t)0);
            const uint64_t pat_end = min(min((uint64_t)off + plen, (uint64_t)rl), text_hi_in_read);
            if (pat_end > off) {
                S.a_run = 1;
                S.a_lo = lo;
                a.p_off[a.n + k] = a.read_off[k] + off;
                a.p_len[a.n + k] = (uint32_t)(pat_end - off);
                a.t_ctg[a.n + k] = S.ctg;
                a.t_lo[a.n + k] = lo;
                a.t_len[a.n + k] = hi - lo;
            }
        }
    }
    a.st[k] = S;
}

// Decodes job j's text (contig t_ctg[j], [t_lo, t_lo + t_len)) to arena + t_off[j].
__global__ __launch_bounds__(JBLOCK) void k_ic_gather(IcArgs a) {
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t j = blockIdx.x * (JBLOCK / 64) + (threadIdx.x >> 6);
    if (j >= 2 * a.n) return;
    const uint32_t tl = a.t_len[j];
    if (tl == 0) return;
    wave_decode(a.R, a.t_ctg[j], a.t_lo[j], a.t_lo[j] + tl, false, a.arena + a.t_off[j], lane);
}

// Stage 2: the refined breakpoints, then verify_split_read's two windows
// (:346-373) as the next two jobs.
__global__ __launch_bounds__(JBLOCK) void k_ic_verify_jobs(IcArgs a) {
    const uint32_t k = blockIdx.x * JBLOCK + threadIdx.x;
    if (k >= a.n) return;
    IcState S = a.st[k];
    const uint32_t rl = a.read_len[k];
    a.p_off[k] = a.p_off[a.n + k] = a.read_off[k];
    a.p_len[k] = a.p_len[a.n + k] = 0;
    a.t_len[k] = a.t_len[a.n + k] = 0;
    if (S.reason) return;
    if (S.d_run) S.donor_end = (int64_t)S.d_lo + (int64_t)a.end[k];
    if (S.a_run) S.acceptor_start = (int64_t)S.a_lo + (int64_t)a.start[a.n + k];
    const int64_t clen = a.R.len[S.ctg];
    const uint32_t mid = rl / 2;
    const uint32_t d_start = (uint32_t)max((int64_t)S.l_r_lo - (int64_t)S.l_q_lo - 30, (int64_t)0);
    const uint32_t d_end = (uint32_t)min(max(S.donor_end + 1 + 30, (int64_t)1), clen);
    const uint32_t a_start = (uint32_t)min(max(S.acceptor_start - 30, (int64_t)0), clen);
    const uint32_t a_end = (uint32_t)min(max((int64_t)S.r_r_hi + 30, (int64_t)a_start + 1), clen);
    S.v_run = !(d_start >= d_end || a_start >= a_end || mid == 0 || mid >= rl);
    if (S.v_run) {
        a.p_len[k] = mid;
        a.t_ctg[k] = S.ctg;
        a.t_lo[k] = d_start;
        a.t_len[k] = d_end - d_start;
        a.p_off[a.n + k] = a.read_off[k] + mid;
        a.p_len[a.n + k] = rl - mid;
        a.t_ctg[a.n + k] = S.ctg;
        a.t_lo[a.n + k] = a_start;
        a.t_len[a.n + k] = a_end - a_start;
    }
    a.st[k] = S;
}

// Stage 3: the gates (:436-446), the junction (:452-486) and its checks
// (:487-496), the exon bounds and hint (:497-511).  Every read owns two exon
// slots; a refused read gets exons the spliced pipeline rejects (end <= start).
__global__ __launch_bounds__(JBLOCK) void k_ic_junction(IcArgs a) {
    const uint32_t k = blockIdx.x * JBLOCK + threadIdx.x;
    if (k >= a.n) return;
    IcState S = a.st[k];
    const uint32_t rl = a.read_len[k];
    esp_intron_chain_rec r;
    r.status = S.reason == -1 ? -1 : 0;
    r.reason = S.reason == -1 ? 0 : S.reason;
    r.genomic_start = 0;
    r.score = 0;
    r.left_start = r.jstart = r.jend = r.right_end = 0;
    r.signal = 0;
    r.support = r.minus = r.n_stitched = 0;
    r.donor_end = r.acceptor_start = 0;
    r.vd = r.va = 0;
    uint32_t *ex = a.exons + 4 * k;
    ex[0] = 1;
    ex[1] = 0;
    ex[2] = 1;
    ex[3] = 0;
    a.plen[k] = 0;
    a.hint[k] = 0;
    if (S.reason == 0) {
        r.donor_end = S.donor_end;
        r.acceptor_start = S.acceptor_start;
        r.minus = S.minus;
        const int32_t vd = S.v_run ? a.dist[k] : INT32_MAX, va = S.v_run ? a.dist[a.n + k] : INT32_MAX;
        r.vd = vd;
        r.va = va;
        const int64_t intron_len = S.acceptor_start - S.donor_end - 1;
        const int32_t max_v = max(vd, va);
        const int32_t sum_v = (int32_t)((uint32_t)vd + (uint32_t)va);  // release i32 wraps
        const int32_t max_thr = (int32_t)((double)rl * 0.17), sum_thr = (int32_t)((double)rl * 0.24);
        if (S.l_nm + S.r_nm < 2) r.reason = 6;
        else if (intron_len < 20 || intron_len > (int64_t)a.intron_max) r.reason = 7;
        else if (max_v > max_thr || sum_v > sum_thr) r.reason = 8;
    }
    if (r.status == 0 && S.reason == 0 && r.reason == 0) {
        SpliceCfgDev sp = a.sp;
        sp.max_intron = a.intron_max;
        const uint8_t *read = a.reads + a.read_off[k];
        esp_refined_junction j = refine_read(a.R, a.J, sp, read, rl, S.ctg, S.l_r_hi, S.r_r_lo, S.l_q_hi, S.r_q_lo,
                                             S.minus != 0);
        if (!j.found) {
            uint32_t d0, d1, a0, a1;
            dinuc(a.R, S.ctg, S.istart, true, d0, d1);
            dinuc(a.R, S.ctg, S.iend, false, a0, a1);
            j.found = 2;
            j.start = S.istart;
            j.end = S.iend;
            j.minus = S.minus;
            j.signal = splice_sig(d0, d1, a0, a1, S.minus != 0);
            j.support = jsupport(a.J, S.ctg, S.istart, S.iend, S.minus != 0);
        }
        r.jstart = j.start;
        r.jend = j.end;
        r.signal = j.signal;
        r.support = j.support;
        if (j.end <= j.start || j.end - j.start < 20u || j.end - j.start > a.intron_max) r.reason = 9;
        if (r.reason == 0) {
            const uint32_t left_start = S.l_r_lo > rl ? S.l_r_lo - rl : 0u;
            const uint32_t right_end = (uint32_t)min((uint64_t)S.r_r_hi + rl, (uint64_t)a.R.len[S.ctg]);
            if (left_start >= j.start || j.end >= right_end) {
                r.reason = 10;
            } else {
                r.left_start = left_start;
                r.right_end = right_end;
                ex[0] = left_start;
                ex[1] = j.start;
                ex[2] = j.end;
                ex[3] = right_end;
                a.plen[k] = (uint64_t)(j.start - left_start) + (right_end - j.end);
                a.hint[k] = ((int64_t)S.l_r_lo - (int64_t)S.l_q_lo) - (int64_t)left_start;
            }
        }
    }
    a.nex[k] = 2;
    a.ctg_out[k] = S.reason == 0 ? S.ctg : 0u;
    a.exon_off[k] = 2ull * k;
    if (k + 1 == a.n) a.exon_off[a.n] = 2ull * a.n;
    a.rec[k] = r;
}

// Stage 4: the acceptance (:513-530).
__global__ __launch_bounds__(JBLOCK) void k_ic_finish(IcArgs a) {
    const uint32_t k = blockIdx.x * JBLOCK + threadIdx.x;
    if (k >= a.n) return;
    esp_intron_chain_rec r = a.rec[k];
    if (r.reason != 0 || r.status != 0) return;
    const esp_splice_rec s = a.srec[k];
    const esp_ext e = a.ext[k];
    r.genomic_start = s.genomic_start;
    r.score = e.score;
    r.n_stitched = s.n_stitched;
    const uint32_t *cg = a.stitched + a.stitch_off[k];
    uint64_t q = 0;
    bool has_n = false;
    for (uint32_t t = 0; s.status == 1 && t < s.n_stitched; t++) {
        const uint32_t op = cg[t] & 0xFu;
        if (op == OP_M || op == OP_I || op == OP_S) q += cg[t] >> 4;
        if (op == OP_N) has_n = true;
    }
    if (s.status != 1) r.reason = 12;
    else if (e.score <= 0) r.reason = 11;
    else if (q != a.read_len[k]) r.reason = 12;
    else if (!has_n) r.reason = 13;
    r.status = r.reason ? 0 : 1;
    a.rec[k] = r;
}

// Exclusive scans of up to two arrays (u32 counts, u64 lengths) in three
// launches: per-tile sums, one block scanning the tile sums (and writing the
// totals at [n]), then each tile's own scan from its offset.  A tile is
// SCAN_TILE elements, SCAN_PER consecutive ones per thread.
constexpr uint32_t SCAN_THREADS = 256, SCAN_PER = 16, SCAN_TILE = SCAN_THREADS * SCAN_PER;

// Block-wide exclusive scan of one u64 per thread (all SCAN_THREADS threads).
__device__ uint64_t block_excl_scan(uint64_t v, uint64_t *sh, uint64_t &total) {
    const uint32_t t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (uint32_t off = 1; off < SCAN_THREADS; off <<= 1) {
        const uint64_t x = t >= off ? sh[t - off] : 0ull;
        __syncthreads();
        sh[t] += x;
        __syncthreads();
    }
    total = sh[SCAN_THREADS - 1];
    const uint64_t r = sh[t] - v;
    __syncthreads();
    return r;
}

__global__ __launch_bounds__(SCAN_THREADS) void k_scan_tiles(const uint32_t *cnt, const uint64_t *len, uint64_t n,
                                                            uint64_t *tmp) {
    __shared__ uint64_t sh[SCAN_THREADS];
    const uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE;
    uint64_t c = 0, l = 0;
    for (uint32_t j = 0; j < SCAN_PER; j++) {  // coalesced: consecutive threads, consecutive elements
        const uint64_t i = base + (uint64_t)j * SCAN_THREADS + threadIdx.x;
        if (i < n) {
            if (cnt) c += cnt[i];
            if (len) l += len[i];
        }
    }
    uint64_t tc, tl;
    (void)block_excl_scan(c, sh, tc);
    (void)block_excl_scan(l, sh, tl);
    if (threadIdx.x == 0) {
        tmp[2 * blockIdx.x] = tc;
        tmp[2 * blockIdx.x + 1] = tl;
    }
}

__global__ __launch_bounds__(SCAN_THREADS) void k_scan_top(uint64_t ntiles, uint64_t n, uint64_t *tmp,
                                                          uint64_t *cnt_off, uint64_t *len_off) {
    __shared__ uint64_t sh[SCAN_THREADS];
    uint64_t carry_c = 0, carry_l = 0;
    for (uint64_t b = 0; b < ntiles; b += SCAN_THREADS) {
        const uint64_t i = b + threadIdx.x;
        const uint64_t c = i < ntiles ? tmp[2 * i] : 0ull, l = i < ntiles ? tmp[2 * i + 1] : 0ull;
        uint64_t tc, tl;
        const uint64_t ec = block_excl_scan(c, sh, tc), el = block_excl_scan(l, sh, tl);
        if (i < ntiles) {
            tmp[2 * i] = carry_c + ec;
            tmp[2 * i + 1] = carry_l + el;
        }
        carry_c += tc;
        carry_l += tl;
    }
    if (threadIdx.x == 0) {
        if (cnt_off) cnt_off[n] = carry_c;
        if (len_off) len_off[n] = carry_l;
    }
}

__global__ __launch_bounds__(SCAN_THREADS) void k_scan_apply(const uint32_t *cnt, const uint64_t *len, uint64_t n,
                                                            const uint64_t *tmp, uint64_t *cnt_off,
                                                            uint64_t *len_off) {
    __shared__ uint64_t sh[SCAN_THREADS];
    const uint64_t b = (uint64_t)blockIdx.x * SCAN_TILE + (uint64_t)threadIdx.x * SCAN_PER;
    uint64_t vc[SCAN_PER], vl[SCAN_PER];
    uint64_t c = 0, l = 0;
#pragma unroll
    for (uint32_t j = 0; j < SCAN_PER; j++) {
        const uint64_t i = b + j;
        vc[j] = (cnt && i < n) ? cnt[i] : 0ull;
        vl[j] = (len && i < n) ? len[i] : 0ull;
        c += vc[j];
        l += vl[j];
    }
    uint64_t tc, tl;
    c = block_excl_scan(c, sh, tc) + tmp[2 * blockIdx.x];
    l = block_excl_scan(l, sh, tl) + tmp[2 * blockIdx.x + 1];
#pragma unroll
    for (uint32_t j = 0; j < SCAN_PER; j++) {
        const uint64_t i = b + j;
        if (i < n) {
            if (cnt_off) cnt_off[i] = c;
            if (len_off) len_off[i] = l;
        }
        c += vc[j];
        l += vl[j];
    }
}

}  // namespace

hipError_t launch_refine(const RefineArgs &a, hipStream_t s) {
    if (a.n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_refine, dim3(lane_blocks(a.n)), dim3(JBLOCK), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_chain_prep(const ChainPrepArgs &a, hipStream_t s) {
    hipLaunchKernelGGL(k_chain_prep, dim3(lane_blocks(a.npairs)), dim3(JBLOCK), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_chain_exons(const ChainPrepArgs &a, hipStream_t s) {
    hipLaunchKernelGGL(k_chain_exons, dim3(lane_blocks(a.npairs)), dim3(JBLOCK), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_chain_rec(const ChainPrepArgs &a, hipStream_t s) {
    hipLaunchKernelGGL(k_chain_rec, dim3(lane_blocks(a.npairs)), dim3(JBLOCK), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_reext_prep(const ReextArgs &a, hipStream_t s) {
    hipLaunchKernelGGL(k_reext_prep, dim3(lane_blocks(a.n)), dim3(JBLOCK), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_reext_gather(const ReextArgs &a, hipStream_t s) {
    hipLaunchKernelGGL(k_reext_gather, dim3(wave_blocks(a.n)), dim3(JBLOCK), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_reext_finish(const ReextArgs &a, hipStream_t s) {
    hipLaunchKernelGGL(k_reext_finish, dim3(lane_blocks(a.n)), dim3(JBLOCK), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_ic(const IcArgs &a, int stage, hipStream_t s) {
    switch (stage) {
    case 0: hipLaunchKernelGGL(k_ic_frags, dim3(lane_blocks(a.n)), dim3(JBLOCK), 0, s, a); break;
    case 1: hipLaunchKernelGGL(k_ic_gather, dim3(wave_blocks(2 * (size_t)a.n)), dim3(JBLOCK), 0, s, a); break;
    case 2: hipLaunchKernelGGL(k_ic_verify_jobs, dim3(lane_blocks(a.n)), dim3(JBLOCK), 0, s, a); break;
    case 3: hipLaunchKernelGGL(k_ic_junction, dim3(lane_blocks(a.n)), dim3(JBLOCK), 0, s, a); break;
    default: hipLaunchKernelGGL(k_ic_finish, dim3(lane_blocks(a.n)), dim3(JBLOCK), 0, s, a); break;
    }
    return hipGetLastError();
}

size_t scan2_tmp_bytes(size_t n) { return 16 * std::max<size_t>((n + SCAN_TILE - 1) / SCAN_TILE, 1); }

hipError_t launch_scan2(const uint32_t *cnt, const uint64_t *len, size_t n, uint64_t *cnt_off, uint64_t *len_off,
                        uint64_t *tmp, hipStream_t s) {
    const uint64_t ntiles = std::max<uint64_t>((n + SCAN_TILE - 1) / SCAN_TILE, 1);
    hipLaunchKernelGGL(k_scan_tiles, dim3((uint32_t)ntiles), dim3(SCAN_THREADS), 0, s, cnt, len, (uint64_t)n, tmp);
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(SCAN_THREADS), 0, s, ntiles, (uint64_t)n, tmp, cnt_off, len_off);
    hipLaunchKernelGGL(k_scan_apply, dim3((uint32_t)ntiles), dim3(SCAN_THREADS), 0, s, cnt, len, (uint64_t)n, tmp,
                       cnt_off, len_off);
    return hipGetLastError();
}

}  // namespace espk
