// esp_walk.hip -- placement walks over the resident reference.
//
// One lane per placement walks its CIGAR against the 2-bit reference and the
// query (align.rs:1060-1163 rescore_with, :2473-2516 recount_mm_ea_in,
// :1231-1288 exact_kmer_support) in one pass.
//
// A lane runs a flat event loop: each iteration consumes one tile of up to 64
// columns of its current M op (its query and reference words loaded at once,
// whole lines rather than a few words at a time), handles every I/D/S/N op
// that precedes the next tile, and scores the tile's 32-column segments from
// registers.  Columns are scored four at a time in byte lanes of a dword:
//   * reference bases come from the resident 2-bit contig through a 256-entry
//     LDS table (packed byte -> four ASCII bytes) and v_alignbyte to the
//     segment's base offset; N intervals and bases past the contig end are
//     substituted on a slow path that only segments meeting one take;
//   * the query's four bytes come from aligned dwords and v_alignbyte; bytes
//     past the query end read as N;
//   * byte equality is exact zero-byte detection on q ^ r, and the strict
//     counts are popcounts of those byte masks.  The strict score is rebuilt
//     from the counts at the end (it is linear in them: match x matches +
//     mismatch x non-exempt substitutions - gap costs + splice terms);
//   * exact_kmer_support's scan (restart one column on at a failing k-mer)
//     counts the same k-mers as a run length over good columns that counts one
//     k-mer and restarts each time k good columns accumulate; four columns at a
//     time this is a first-bad / last-bad update when k >= 4, and a per-column
//     loop otherwise.
// RefSkips look the junction library up in its device hash set, else score
// the donor / acceptor signal on both
// strands minus the intron-length penalty, read from a host-built threshold
// table (exact against the f64 formula) held in LDS.
#include "esp_kernels.hpp"

using namespace espd;

namespace espk {

namespace {

constexpr uint32_t OPM = 0, OPI = 1, OPD = 2, OPN = 3, OPS = 4;
// Columns per segment; NQ query dwords and NR reference dwords hold one.
constexpr uint32_t SEG = 32, NG = SEG / 4, NQ = NG + 1, NR = SEG / 16 + 1;
typedef uint32_t u32x4 __attribute__((ext_vector_type(4), aligned(4)));
constexpr uint32_t N4 = 0x4E4E4E4Eu, A4 = 0x41414141u, C4 = 0x43434343u, G4 = 0x47474747u,
                   T4 = 0x54545454u, HI = 0x80808080u;

// 0x80 in every byte of x that is zero, 0 in the others (exact for all bytes).
__device__ __forceinline__ uint32_t zbytes(uint32_t x) {
    const uint32_t t = (x & 0x7F7F7F7Fu) + 0x7F7F7F7Fu;
    return ~(t | x) & HI;
}

// Four bytes of the eight {hi, lo} starting at byte sh (0..3).
__device__ __forceinline__ uint32_t align4(uint32_t hi, uint32_t lo, uint32_t sh) {
    return __builtin_amdgcn_alignbyte(hi, lo, sh);
}

// Byte mask of the first c (clamped to 0..4) bytes of a dword.
__device__ __forceinline__ uint32_t first_bytes(int32_t c) {
    return c >= 4 ? 0xFFFFFFFFu : c <= 0 ? 0u : (0xFFFFFFFFu >> (32 - 8 * c));
}

__device__ __forceinline__ uint32_t up_ascii(uint32_t b) { return (b >= 'a' && b <= 'z') ? b - 32 : b; }

// splice_signal(..).score(sp) (splice.rs:61-103).
__device__ int32_t signal_score(const WalkArgs &a, uint32_t d0, uint32_t d1, uint32_t a0, uint32_t a1,
                                bool minus) {
    d0 = up_ascii(d0), d1 = up_ascii(d1), a0 = up_ascii(a0), a1 = up_ascii(a1);
    const uint32_t d = d0 << 8 | d1, ac = a0 << 8 | a1;
    constexpr uint32_t GT = 'G' << 8 | 'T', GC = 'G' << 8 | 'C', AT = 'A' << 8 | 'T', AG = 'A' << 8 | 'G',
                       AC = 'A' << 8 | 'C', CT = 'C' << 8 | 'T';
    int sig;
    if (!minus) sig = (d == GT && ac == AG) ? 0 : (d == GC && ac == AG) ? 1 : (d == AT && ac == AC) ? 2 : 3;
    else sig = (d == CT && ac == AC) ? 0 : (d == CT && ac == GC) ? 1 : (d == GT && ac == AT) ? 2 : 3;
    return sig == 0 ? a.gtag : sig == 1 ? a.gcag : sig == 2 ? a.atac : -a.noncanon;
}

// Number of entries of the ascending t[0..n) that are <= v.
__device__ __forceinline__ int32_t count_le(const uint32_t *t, int32_t n, uint32_t v) {
    int32_t lo = 0, hi = n;
    while (lo < hi) {
        const int32_t mid = (lo + hi) >> 1;
        if (t[mid] <= v) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// Number of entries of the descending t[0..n) that are >= v.
__device__ __forceinline__ int32_t count_ge(const uint32_t *t, int32_t n, uint32_t v) {
    int32_t lo = 0, hi = n;
    while (lo < hi) {
        const int32_t mid = (lo + hi) >> 1;
        if (t[mid] >= v) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// contains_any_strand (gtf.rs:165-169): probe the library's hash set from
// the triple's slot (v0 holds that slot, loaded ahead) until the triple or an
// empty slot (the set is at most a quarter full, so an empty slot exists).
__device__ bool junction_known(const WalkArgs &a, uint4 v, uint32_t c, uint32_t s, uint32_t e) {
    uint32_t slot = junction_hash(c, s, e) & a.jmask;
    while (true) {
        if (!v.w) return false;
        if (v.x == c && v.y == s && v.z == e) return true;
        slot = (slot + 1) & a.jmask;
        v = a.jn[slot];
    }
}

// First N interval of the contig whose end is > p.
__device__ __forceinline__ uint32_t iv_seek(const uint32_t *iv, uint32_t niv, uint32_t p) {
    uint32_t lo = 0, hi = niv;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (iv[2 * mid + 1] <= p) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// ASCII reference base at p given the dword pk[p >> 4]: N inside an N
// interval or past the contig end.
__device__ uint32_t ref_base(uint32_t word, const uint32_t *iv, uint32_t niv, uint32_t len, uint32_t p) {
    if (p >= len) return 'N';
    if (niv) {
        const uint32_t j = iv_seek(iv, niv, p);
        if (j < niv && iv[2 * j] <= p) return 'N';
    }
    const uint32_t byte = (word >> (8 * ((p >> 2) & 3u))) & 0xFFu;
    const uint32_t code = (byte >> (6 - 2 * (p & 3u))) & 3u;
    return (0x54474341u >> (8 * code)) & 0xFFu;
}

struct Lane {
    // placement
    uint32_t idx;
    const uint32_t *qw;  // the query's bytes as aligned dwords
    uint32_t qoff0;      // byte offset of the query's first byte in qw[0]
    uint32_t qlen, qjmax;
    const uint32_t *pk;  // the contig's packed bases as dwords
    const uint32_t *iv;
    uint32_t niv, len, pkmax, contig;
    const uint32_t *cig;
    uint32_t nops, o, rem, ro, rf;
    uint32_t nw;  // cig[o] when o < nops (loaded one op ahead)
    // for a RefSkip that follows the current M op, loaded when the M op
    // starts: its library slot and the dwords of its four signal bases
    bool pre;
    uint4 pslot;
    uint32_t pw[4];
    uint32_t exag, extc;  // all-ones when dir exempts A>G / T>C
    bool fresh;           // the next segment starts an M op
    // N-interval cursor: first interval whose end > the last scored position
    uint32_t ci, ivs, ive;
    // tallies
    uint32_t matches, subs, ag, tc, kmers, run;
};

// Tallies that only non-M ops change, kept in LDS per lane (column
// threadIdx.x): I/D lengths, I count, D count, clipped, RefSkip count, gap
// cost, RefSkip score.
enum { T_GAPLEN, T_INS, T_DEL, T_CLIP, T_NSPL, T_GAPS, T_SPL, NT };
typedef uint32_t Tally[NT][256];

__device__ __forceinline__ void iv_load(Lane &L) {
    if (L.ci < L.niv) {
        L.ivs = L.iv[2 * L.ci];
        L.ive = L.iv[2 * L.ci + 1];
    } else {
        L.ivs = L.ive = 0xFFFFFFFFu;
    }
}

// Library slot and signal-base dwords of a RefSkip of length n at rf.
__device__ __forceinline__ void refskip_load(const WalkArgs &a, const Lane &L, uint32_t rf, uint32_t n,
                                             uint4 &slot, uint32_t (&w)[4]) {
    slot = a.jn ? a.jn[junction_hash(L.contig, rf, rf + n) & a.jmask] : make_uint4(0, 0, 0, 0);
    const uint32_t last = L.pkmax + NR - 1;
    const uint32_t p1 = rf == 0xFFFFFFFFu ? rf : rf + 1;
    w[0] = L.pk[min(rf >> 4, last)];
    w[1] = L.pk[min(p1 >> 4, last)];
    w[2] = L.pk[min((rf + (n >= 2 ? n - 2 : 0u)) >> 4, last)];
    w[3] = L.pk[min((rf + (n >= 1 ? n - 1 : 0u)) >> 4, last)];
}

// Handles every non-M op up to the next M op of nonzero length (rem > 0) or
// the end of the CIGAR (rem == 0).  Entering an M op that a RefSkip follows
// issues that RefSkip's loads.
__device__ void advance(const WalkArgs &a, Lane &L, const uint32_t *s_up, const uint32_t *s_down, Tally &T) {
    const uint32_t me = threadIdx.x;
    while (L.rem == 0 && L.o < L.nops) {
        const uint32_t w = L.nw;
        L.o++;
        if (L.o < L.nops) L.nw = L.cig[L.o];
        const uint32_t op = w & 15u, n = w >> 4;
        if (op == OPM) {
            L.rem = n;
            L.fresh = true;
            if (n && L.o < L.nops && (L.nw & 15u) == OPN) {
                refskip_load(a, L, L.rf + n, L.nw >> 4, L.pslot, L.pw);
                L.pre = true;
            }
        } else if (op == OPI || op == OPD) {
            T[T_GAPS][me] += (uint32_t)(a.gap_open + a.gap_ext * (int32_t)n);
            T[T_GAPLEN][me] += n;
            if (op == OPI) {
                T[T_INS][me]++;
                L.ro += n;
            } else {
                T[T_DEL][me]++;
                L.rf += n;
            }
        } else if (op == OPS) {
            T[T_CLIP][me] += n;
            L.ro += n;
        } else if (op == OPN) {
            const uint32_t rf = L.rf;
            uint4 slot;
            uint32_t wd[4];
            if (L.pre) {
                slot = L.pslot;
                for (int i = 0; i < 4; i++) wd[i] = L.pw[i];
                L.pre = false;
            } else {
                refskip_load(a, L, rf, n, slot, wd);
            }
            int32_t d;
            if (a.jn && junction_known(a, slot, L.contig, rf, rf + n)) {
                d = a.known;
            } else {
                const uint32_t p1 = rf == 0xFFFFFFFFu ? rf : rf + 1;
                const uint32_t p2 = rf + (n >= 2 ? n - 2 : 0u), p3 = rf + (n >= 1 ? n - 1 : 0u);
                const uint32_t d0 = ref_base(wd[0], L.iv, L.niv, L.len, rf);
                const uint32_t d1 = ref_base(wd[1], L.iv, L.niv, L.len, p1);
                const uint32_t a0 = ref_base(wd[2], L.iv, L.niv, L.len, p2);
                const uint32_t a1 = ref_base(wd[3], L.iv, L.niv, L.len, p3);
                const int32_t sp = signal_score(a, d0, d1, a0, a1, false);
                const int32_t sm = signal_score(a, d0, d1, a0, a1, true);
                const int32_t pen = n >= 1000 ? count_le(s_up, a.nup, n) : count_ge(s_down, a.ndown, n);
                d = max(sp, sm) - pen;
            }
            T[T_SPL][me] += (uint32_t)d;
            T[T_NSPL][me]++;
            L.rf += n;
        }
    }
}

// A placement's own fields, loaded one placement ahead; the fields of its
// contig and its first CIGAR word follow once those have arrived.
struct Meta {
    uint64_t read_off, cig_off;
    uint32_t contig, read_len, pos, cig_n, dir;
    uint64_t pk_off, iv0, iv1;
    uint32_t len, w0;
};

__device__ __forceinline__ void meta_load2(const WalkArgs &a, Meta &m) {
    if (m.contig < a.R.ncontig) {
        m.pk_off = a.R.pk_off[m.contig];
        m.len = a.R.len[m.contig];
        m.iv0 = a.R.iv_off[m.contig];
        m.iv1 = a.R.iv_off[m.contig + 1];
        m.w0 = m.cig_n ? a.cigar[m.cig_off] : 0u;
    }
}

__device__ __forceinline__ void meta_load(const WalkArgs &a, uint32_t idx, Meta &m) {
    m.contig = a.contig[idx];
    m.read_off = a.read_off[idx];
    m.read_len = a.read_len[idx];
    m.pos = a.pos[idx];
    m.cig_off = a.cig_off[idx];
    m.cig_n = a.cig_n[idx];
    m.dir = a.dir ? a.dir[idx] : 0u;
}

// Columns per tile of the walk loop, and the tile's query and reference words.
constexpr uint32_t WT = 64;
// Dwords of LDS per wave for its round's queries: 64 queries of up to 150
// bytes, and three blocks of four waves still fit a CU's 160 KB.
constexpr uint32_t STAGE_DW = 2432;
// u32x4 loads per lane that cover STAGE_DW dwords over a wave.
constexpr uint32_t STAGE_IT = (STAGE_DW + 255) / 256;
static_assert(STAGE_DW % 4 == 0, "the copy writes whole u32x4 groups");
typedef const __attribute__((address_space(1))) uint32_t *gdw;
typedef const __attribute__((address_space(1))) u32x4 *gdw4;

// Copies the ndw (<= STAGE_DW) dwords from g0 into the wave's LDS stage, B
// u32x4 loads per lane in flight at a time: a lane issues a batch's loads
// before it writes any of them, so the copy waits for memory once per batch
// rather than once per 256 dwords.  Each load lies in [g0, g0 + 4 ndw): the
// lane whose group runs past ndw loads its dwords with the index clamped to
// the last one (the extra LDS dwords, still inside the stage, are never read).
template <uint32_t B>
__device__ __forceinline__ void stage_copy(uint32_t *stage, uintptr_t g0, uint32_t ndw, uint32_t lane) {
    const gdw src = (gdw)g0;
#pragma unroll
    for (uint32_t i0 = 0; i0 < STAGE_IT; i0 += B) {
        u32x4 v[B];
#pragma unroll
        for (uint32_t i = 0; i < B; i++) {
            const uint32_t d = 4 * lane + 256 * (i0 + i);
            if (d + 4 <= ndw) {
                v[i] = *(gdw4)(src + d);
            } else if (d < ndw) {
                v[i].x = src[d];
                v[i].y = src[min(d + 1, ndw - 1)];
                v[i].z = src[min(d + 2, ndw - 1)];
                v[i].w = src[min(d + 3, ndw - 1)];
            }
        }
#pragma unroll
        for (uint32_t i = 0; i < B; i++) {
            const uint32_t d = 4 * lane + 256 * (i0 + i);
            if (d < ndw) *(u32x4 *)(stage + d) = v[i];
        }
    }
}
// Loads in flight per lane in each kernel's copy.
constexpr uint32_t WALK_STAGE_B = STAGE_IT, FAST_STAGE_B = STAGE_IT;
// sbase: the (wave-uniform) address the wave's LDS copy of its round's
// queries starts at, 0 when the round's queries were not copied.
__device__ __forceinline__ void tile_load(const Lane &L, const uint32_t *stage, uintptr_t sbase,
                                          uint32_t (&q)[WT / 4 + 1], uint32_t (&r)[WT / 16 + 1]) {
    if (L.qlen && sbase) {
        // from the wave's LDS copy (every dword up to qjmax is there)
        const uint32_t j0 = (L.qoff0 + L.ro) >> 2;
        const uint32_t *sp = stage + (uint32_t)(((uintptr_t)L.qw - sbase) >> 2);
        if (j0 + WT / 4 <= L.qjmax) {
#pragma unroll
            for (uint32_t j = 0; j <= WT / 4; j++) q[j] = sp[j0 + j];
        } else {
#pragma unroll
            for (uint32_t j = 0; j <= WT / 4; j++) q[j] = sp[min(j0 + j, L.qjmax)];
        }
    } else if (L.qlen) {
        const uint32_t j0 = (L.qoff0 + L.ro) >> 2;
        if (j0 + WT / 4 <= L.qjmax) {
#pragma unroll
            for (uint32_t j = 0; j < WT / 16; j++) {
                const u32x4 v = *(const u32x4 *)(L.qw + j0 + 4 * j);
                q[4 * j] = v.x, q[4 * j + 1] = v.y, q[4 * j + 2] = v.z, q[4 * j + 3] = v.w;
            }
            q[WT / 4] = L.qw[j0 + WT / 4];
        } else {
#pragma unroll
            for (uint32_t j = 0; j <= WT / 4; j++) q[j] = L.qw[min(j0 + j, L.qjmax)];
        }
    } else {
#pragma unroll
        for (uint32_t j = 0; j <= WT / 4; j++) q[j] = N4;
    }
    const uint32_t last = L.pkmax + NR - 1, d = L.rf >> 4;
    if (d + WT / 16 <= last) {
#pragma unroll
        for (uint32_t j = 0; j < WT / 64; j++) {
            const u32x4 v = *(const u32x4 *)(L.pk + d + 4 * j);
            r[4 * j] = v.x, r[4 * j + 1] = v.y, r[4 * j + 2] = v.z, r[4 * j + 3] = v.w;
        }
        r[WT / 16] = L.pk[d + WT / 16];
    } else {
#pragma unroll
        for (uint32_t j = 0; j <= WT / 16; j++) r[j] = L.pk[min(d + j, last)];
    }
}

__device__ void place_begin(const WalkArgs &a, Lane &L, uint32_t idx, const Meta &m, const uint32_t *s_up,
                            const uint32_t *s_down, Tally &T) {
    L.idx = idx;
    L.contig = m.contig;
    L.qlen = m.read_len;
    const uintptr_t qa = (uintptr_t)(a.reads + m.read_off);
    L.qw = (const uint32_t *)(qa & ~(uintptr_t)3);
    L.qoff0 = (uint32_t)(qa & 3u);
    L.qjmax = L.qlen ? (L.qoff0 + L.qlen - 1) >> 2 : 0;
    L.pk = (const uint32_t *)(a.R.packed + m.pk_off);
    L.len = m.len;
    // first dword of the last NR the contig's region holds: the upload gives
    // it ceil(len / 4) + 16 bytes, rounded up to 16
    L.pkmax = ((((L.len + 3u) >> 2) + 16u + 15u) & ~15u) / 4u - NR;
    L.iv = a.R.iv + 2 * m.iv0;
    L.niv = (uint32_t)(m.iv1 - m.iv0);
    L.cig = a.cigar + m.cig_off;
    L.nops = m.cig_n;
    L.exag = m.dir == 1 ? 0xFFFFFFFFu : 0u;
    L.extc = m.dir == 2 ? 0xFFFFFFFFu : 0u;
    L.o = L.rem = L.ro = 0;
    L.nw = m.w0;
    L.rf = m.pos;
    L.fresh = false;
    L.pre = false;
    L.ci = L.niv ? iv_seek(L.iv, L.niv, L.rf) : 0u;
    iv_load(L);
    L.matches = L.subs = L.ag = L.tc = L.kmers = L.run = 0;
    for (int f = 0; f < NT; f++) T[f][threadIdx.x] = 0;
    advance(a, L, s_up, s_down, T);
}

__device__ void place_end(const WalkArgs &a, const Lane &L, const Tally &T) {
    const uint32_t me = threadIdx.x;
    esp_strict_score s;
    const uint32_t ex = (L.ag & L.exag) | (L.tc & L.extc);
    s.score = (int32_t)((uint32_t)(a.match * (int32_t)L.matches) + (uint32_t)(a.mismatch * (int32_t)(L.subs - ex)) -
                        T[T_GAPS][me] + T[T_SPL][me]);
    s.aligned = L.matches + L.subs + T[T_GAPLEN][me];
    s.matches = L.matches;
    s.mm = L.subs - L.ag - L.tc;
    s.ea_ag = L.ag;
    s.ea_tc = L.tc;
    s.ins = T[T_INS][me];
    s.del = T[T_DEL][me];
    s.clipped = T[T_CLIP][me];
    s.n_splice = T[T_NSPL][me];
    s.splice_score = (int32_t)T[T_SPL][me];
    esp_walk_counts c;
    c.mm = s.mm;
    c.ea_ag = L.ag;
    c.ea_tc = L.tc;
    c.kmers = (a.k && L.qlen) ? L.kmers : 0u;
    if (a.strict) a.strict[L.idx] = s;
    if (a.counts) a.counts[L.idx] = c;
    if (a.status) a.status[L.idx] = 0;
}

// Byte mask of the set bits of a 4-bit mask.
__device__ __forceinline__ uint32_t nibble_bytes(uint32_t m4) {
    return 0xFFu * ((m4 & 1u) | ((m4 & 2u) << 7) | ((m4 & 4u) << 14) | ((m4 & 8u) << 21));
}

// Scores one segment: ncol columns at query index qa, reference position p.
__device__ __forceinline__ void seg_score(const WalkArgs &a, Lane &L, const uint32_t *lut,
                                          const uint32_t *q, const uint32_t *r, uint32_t qa,
                                          uint32_t p, uint32_t ncol, bool fresh) {
    // reference: NG + 1 packed bytes from byte (p >> 2) on, as ASCII groups
    const uint32_t bo = (p >> 2) & 3u;
    uint32_t A[NG + 1];
#pragma unroll
    for (uint32_t i = 0; i + 1 < NR; i++) {
        const uint32_t W = align4(r[i + 1], r[i], bo);
#pragma unroll
        for (uint32_t t = 0; t < 4; t++) A[4 * i + t] = lut[(W >> (8 * t)) & 0xFFu];
    }
    A[NG] = lut[(r[NR - 1] >> (8 * bo)) & 0xFFu];
    // N intervals and the contig end: the cursor moves to the first interval
    // ending after p; a segment that meets an interval or the end marks its
    // reference N columns in nref
    while (L.ci < L.niv && L.ive <= p) {
        L.ci++;
        iv_load(L);
    }
    uint32_t nref = 0;
    if (p + ncol > L.len || L.ivs < p + ncol) {
        for (uint32_t c = 0; c < ncol; c++) {
            const uint32_t x = p + c;
            while (L.ci < L.niv && L.ive <= x) {
                L.ci++;
                iv_load(L);
            }
            if (x >= L.len || L.ivs <= x) nref |= 1u << c;
        }
    }
    // query bytes from qlen on read as N
    const int32_t qv = (int32_t)(L.qlen > qa ? L.qlen - qa : 0u);
    const bool slow = nref != 0 || qa + ncol > L.qlen;
    const uint32_t qsh = (L.qoff0 + qa) & 3u, rsh = p & 3u;
    const uint32_t k = a.k;
    uint32_t run = fresh ? 0u : L.run;
#pragma unroll
    for (uint32_t t = 0; t < NG; t++) {
        // a segment shorter than SEG ends its M op, so the run it leaves is
        // never read again: groups wholly past ncol change nothing that is
        if (4 * t >= ncol) break;
        const uint32_t cm = first_bytes((int32_t)ncol - 4 * (int32_t)t) & HI;
        uint32_t q4 = L.qlen ? align4(q[t + 1], q[t], qsh) : N4;
        uint32_t r4 = align4(A[t + 1], A[t], rsh);
        uint32_t rnh = 0;  // 0x80 in the bytes whose reference base is N
        if (slow) {
            const uint32_t qm = first_bytes(qv - 4 * (int32_t)t);
            const uint32_t rn = nibble_bytes((nref >> (4 * t)) & 15u);
            q4 = (q4 & qm) | (N4 & ~qm);
            r4 = (r4 & ~rn) | (N4 & rn);
            rnh = rn & HI;
        }
        const uint32_t eq = zbytes(q4 ^ r4);
        const uint32_t valid = cm & ~(zbytes(q4 ^ N4) | rnh);
        const uint32_t agm = zbytes((r4 ^ A4) | (q4 ^ G4)) & cm;
        const uint32_t tcm = zbytes((r4 ^ T4) | (q4 ^ C4)) & cm;
        L.matches += __builtin_popcount(eq & valid);
        L.subs += __builtin_popcount(valid & ~eq);
        L.ag += __builtin_popcount(agm);
        L.tc += __builtin_popcount(tcm);
        const uint32_t bh = ~((eq | (agm & L.exag) | (tcm & L.extc)) & cm) & HI;
        if (k >= 4) {
            // run of good columns: the ones before the first bad column may
            // complete a k-mer (at most one, k >= 4); after a bad column the
            // run is the good columns after the last one
            const uint32_t f = bh ? ((uint32_t)__builtin_ctz(bh) >> 3) : 4u;
            run += f;
            const bool hit = run >= k;
            L.kmers += hit ? 1u : 0u;
            run = bh ? ((uint32_t)__builtin_clz(bh) >> 3) : (hit ? run - k : run);
        } else if (k) {
            for (uint32_t c = 0; c < 4; c++) {
                if ((bh >> (8 * c + 7)) & 1u) {
                    run = 0;
                } else if (++run == k) {
                    L.kmers++;
                    run = 0;
                }
            }
        }
    }
    L.run = run;
}

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_walk(WalkArgs a) {
    __shared__ uint32_t lut[256];
    __shared__ uint32_t s_up[PEN_UP];
    __shared__ uint32_t s_down[PEN_DOWN];
    __shared__ Tally T;
    __shared__ __align__(16) uint32_t s_stage[4][STAGE_DW];
    {
        // packed byte (four bases, first in bits 7:6) -> four ASCII bytes
        const uint32_t b = threadIdx.x;
        uint32_t v = 0;
        for (int t = 0; t < 4; t++) v |= ((0x54474341u >> (8 * ((b >> (6 - 2 * t)) & 3u))) & 0xFFu) << (8 * t);
        lut[b] = v;
        if (b < PEN_UP) s_up[b] = a.up[b];
        if (b < PEN_DOWN) s_down[b] = a.down[b];
    }
    __syncthreads();

    // Rounds: every lane of a wave begins its next placement together, with
    // that placement's fields loaded during the previous one, so a wave pays
    // each placement's set-up latency once rather than once per lane.  The
    // wave's 64 placements are consecutive, and so are their queries in a
    // batch laid out in placement order: when the round's queries span at most
    // STAGE_DW dwords the wave copies them into LDS with coalesced loads first,
    // so each query byte leaves HBM once instead of once per tile.
    const uint32_t stride = gridDim.x * blockDim.x;
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    uint32_t *const stage = s_stage[wave];
    Meta m;
    {
        const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
        if (idx < a.n) {
            meta_load(a, idx, m);
            meta_load2(a, m);
        }
    }
    for (uint32_t first = blockIdx.x * blockDim.x + 64 * wave; first < a.n; first += stride) {
        const uint32_t idx = first + lane;
        const bool act = idx < a.n;
        const Meta cur = m;
        if (act && idx + stride < a.n) meta_load(a, idx + stride, m);
        const bool walk = act && cur.contig < a.R.ncontig;
        uint64_t lo = walk && cur.read_len ? cur.read_off : ~0ull;
        uint64_t hi = walk && cur.read_len ? cur.read_off + cur.read_len : 0ull;
#pragma unroll
        for (int o = 32; o; o >>= 1) {
            lo = min(lo, (uint64_t)__shfl_xor((unsigned long long)lo, o, 64));
            hi = max(hi, (uint64_t)__shfl_xor((unsigned long long)hi, o, 64));
        }
        // every lane holds the same span now; keep it in scalar registers
        lo = (uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)lo) |
             (uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)(lo >> 32)) << 32;
        hi = (uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)hi) |
             (uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)(hi >> 32)) << 32;
        uintptr_t sbase = 0;
        if (hi > lo) {
            const uintptr_t g0 = (uintptr_t)(a.reads + lo) & ~(uintptr_t)3;
            const uintptr_t g1 = ((uintptr_t)(a.reads + hi) + 3) & ~(uintptr_t)3;
            const uint32_t ndw = (uint32_t)min((uint64_t)(g1 - g0) >> 2, (uint64_t)STAGE_DW + 1);
            if (ndw <= STAGE_DW) {
                // every dword copied lies in [g0, g1): the dwords holding the
                // round's query bytes, which the global path reads too
                stage_copy<WALK_STAGE_B>(stage, g0, ndw, lane);
                sbase = g0;
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            }
        }
        if (act && !walk) {
            // a placement on a contig out of range gets zeroed records
            if (a.strict) a.strict[idx] = esp_strict_score{};
            if (a.counts) a.counts[idx] = esp_walk_counts{};
            if (a.status) a.status[idx] = ESP_E_INVAL;
        }
        if (walk) {
            Lane L;
            place_begin(a, L, idx, cur, s_up, s_down, T);
            bool second = idx + stride < a.n;  // the next placement's contig fields are due
            while (L.rem) {
                if (second) {
                    meta_load2(a, m);
                    second = false;
                }
                // one tile: up to WT columns of the M op, its words loaded at once
                uint32_t tq[WT / 4 + 1], tr[WT / 16 + 1];
                tile_load(L, stage, sbase, tq, tr);
                const uint32_t tcol = min(L.rem, WT), qa = L.ro, p = L.rf;
                const bool fresh = L.fresh;
                L.fresh = false;
                L.rem -= tcol;
                L.ro += tcol;
                L.rf += tcol;
                if (L.rem == 0) advance(a, L, s_up, s_down, T);
#pragma unroll
                for (uint32_t sg = 0; sg < WT / SEG; sg++) {
                    const uint32_t c0 = sg * SEG;
                    if (c0 < tcol)
                        seg_score(a, L, lut, tq + sg * (SEG / 4), tr + sg * (SEG / 16), qa + c0, p + c0,
                                  min(tcol - c0, SEG), fresh && sg == 0);
                }
            }
            if (second) meta_load2(a, m);
            place_end(a, L, T);
        }
        // the next round's copy overwrites this one only after every lane's
        // reads of it (same-wave LDS operations complete in order)
        __builtin_amdgcn_wave_barrier();
    }
}

// ------------------------------------------------------------ fast lane

// Columns per fast-lane tile.  The tile is unrolled over its columns in two
// forms (whole and partial): 64 columns keep the kernel's code near 32 KB,
// inside the instruction cache two CUs share, where 128 columns ran slower
// for their code size alone.
constexpr uint32_t FT = 64;

// Kadane state of try_fast_lane (align.rs:2346-2366).
struct FastKadane {
    int32_t best = 0, cur = 0;
    uint32_t bs = 0, be = 0, cs = 0, cmm = 0, bmm = 0;
};

// Sixteen ASCII reference bases from x0 on, base by base: N inside an N
// interval, before the contig start or past its end.  Out of line, so that
// the tiles unrolled over their columns hold one call each rather than a
// copy of the interval search per column.
struct Ref16 {
    uint32_t w[4];
};
__device__ __attribute__((noinline)) Ref16 ref16_slow(int64_t x0, const uint32_t *pk, uint32_t pklast,
                                                      const uint32_t *iv, uint32_t niv, uint32_t clen) {
    Ref16 out;
    for (int t = 0; t < 4; t++) {
        uint32_t v = 0;
        for (int c = 0; c < 4; c++) {
            const int64_t x = x0 + 4 * t + c;
            uint32_t base = 'N';
            if (x >= 0 && x < (int64_t)clen)
                base = ref_base(pk[min((uint32_t)x >> 4, pklast)], iv, niv, clen, (uint32_t)x);
            v |= base << (8 * c);
        }
        out.w[t] = v;
    }
    return out;
}

// One tile of FT columns from registers: query words q (aligned, shifted by
// qsh), reference words r (dwords from (p0 >> 4), used when fast), else the
// bases come one by one (tile off the contig or meeting an N interval).
// FULL: every column of the tile is in the read.
template <bool FULL>
__device__ __forceinline__ void fast_tile(const FastArgs &a, FastKadane &K, const uint32_t *lut,
                                          const uint32_t (&q)[FT / 4 + 1], const uint32_t (&r)[FT / 16 + 1],
                                          uint32_t qsh, bool fast, int64_t p0, uint32_t i0, uint32_t ncol,
                                          const uint32_t *pk, uint32_t pklast, const uint32_t *iv, uint32_t niv,
                                          uint32_t clen, uint32_t exag, uint32_t extc) {
    const uint32_t bo = fast ? ((uint32_t)p0 >> 2) & 3u : 0u, rsh = fast ? (uint32_t)p0 & 3u : 0u;
#pragma unroll
    for (uint32_t g = 0; g < FT / 16; g++) {
        // columns from ncol on leave the Kadane state as it is
        if (!FULL && 16 * g >= ncol) break;
        // reference: four ASCII words of this 16-column group
        uint32_t r4s[4];
        if (fast) {
            const uint32_t W = align4(r[g + 1], r[g], bo);
            uint32_t A[5];
#pragma unroll
            for (int t = 0; t < 4; t++) A[t] = lut[(W >> (8 * t)) & 0xFFu];
            A[4] = lut[(r[g + 1] >> (8 * bo)) & 0xFFu];
#pragma unroll
            for (int t = 0; t < 4; t++) r4s[t] = align4(A[t + 1], A[t], rsh);
        } else {
            const Ref16 w = ref16_slow(p0 + 16 * g, pk, pklast, iv, niv, clen);
#pragma unroll
            for (int t = 0; t < 4; t++) r4s[t] = w.w[t];
        }
#pragma unroll
        for (uint32_t t = 0; t < 4; t++) {
            const uint32_t q4 = align4(q[4 * g + t + 1], q[4 * g + t], qsh), r4 = r4s[t];
            const uint32_t eq = zbytes(q4 ^ r4);
            const uint32_t nb = zbytes(q4 ^ N4) | zbytes(r4 ^ N4);
            const uint32_t ex = (zbytes((r4 ^ A4) | (q4 ^ G4)) & exag) | (zbytes((r4 ^ T4) | (q4 ^ C4)) & extc);
#pragma unroll
            for (uint32_t c = 0; c < 4; c++) {
                const uint32_t col = 16 * g + 4 * t + c;
                const uint32_t i = i0 + col;
                const uint32_t bit = 8 * c + 7;
                const bool isn = (nb >> bit) & 1u, ise = (eq >> bit) & 1u, isx = (ex >> bit) & 1u;
                const int32_t s = isn ? 0 : ise ? a.match : isx ? 0 : a.mismatch;
                const uint32_t mm = (!isn && !ise && !isx) ? 1u : 0u;
                const bool on = FULL || col < ncol;
                // Kadane with <= 0 reset (align.rs:2346-2366)
                const bool reset = K.cur <= 0;
                const int32_t ncur = reset ? s : K.cur + s;
                const uint32_t ncs = reset ? i : K.cs, ncmm = reset ? mm : K.cmm + mm;
                if (on) {
                    K.cur = ncur;
                    K.cs = ncs;
                    K.cmm = ncmm;
                    if (ncur > K.best) {
                        K.best = ncur;
                        K.bs = ncs;
                        K.be = i + 1;
                        K.bmm = ncmm;
                    }
                }
            }
        }
    }
}


// try_fast_lane (align.rs:2274-2401), one lane per chain: the anchor
// conditions, then Kadane along the chain diagonal four columns at a time
// (column scores from exact byte masks as in the walks; the running sum's
// updates stay per column), then the zero-mismatch gate and the at most
// three-op CIGAR.
// One lane per chain, in rounds of 64 consecutive chains per wave: the
// round's queries (consecutive in the batch) are copied into the wave's LDS
// with coalesced loads first when they span at most STAGE_DW dwords, so each
// query byte leaves HBM once instead of once per tile.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_fast_lane(FastArgs a) {
    __shared__ uint32_t lut[256];
    __shared__ __align__(16) uint32_t s_stage[4][STAGE_DW];
    {
        const uint32_t b = threadIdx.x;
        uint32_t v = 0;
        for (int t = 0; t < 4; t++) v |= ((0x54474341u >> (8 * ((b >> (6 - 2 * t)) & 3u))) & 0xFFu) << (8 * t);
        lut[b] = v;
    }
    __syncthreads();
    const uint32_t stride = gridDim.x * blockDim.x;
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    uint32_t *const stage = s_stage[wave];
    for (uint32_t first = blockIdx.x * blockDim.x + 64 * wave; first < a.n; first += stride) {
        const uint32_t idx = first + lane;
        const bool act = idx < a.n;
        const uint32_t contig = act ? a.contig[idx] : 0u;
        const bool incontig = act && contig < a.R.ncontig;
        const uint32_t m = incontig ? a.read_len[idx] : 0u;
        const uint64_t roff = incontig ? a.read_off[idx] : 0ull;
        // the round's query span, uniform over the wave
        uint64_t lo = m ? roff : ~0ull, hi = m ? roff + m : 0ull;
#pragma unroll
        for (int o = 32; o; o >>= 1) {
            lo = min(lo, (uint64_t)__shfl_xor((unsigned long long)lo, o, 64));
            hi = max(hi, (uint64_t)__shfl_xor((unsigned long long)hi, o, 64));
        }
        lo = (uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)lo) |
             (uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)(lo >> 32)) << 32;
        hi = (uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)hi) |
             (uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)(hi >> 32)) << 32;
        uintptr_t sbase = 0;
        if (hi > lo) {
            const uintptr_t g0 = (uintptr_t)(a.reads + lo) & ~(uintptr_t)3;
            const uintptr_t g1 = ((uintptr_t)(a.reads + hi) + 3) & ~(uintptr_t)3;
            const uint32_t ndw = (uint32_t)min((uint64_t)(g1 - g0) >> 2, (uint64_t)STAGE_DW + 1);
            if (ndw <= STAGE_DW) {
                // every dword copied lies in [g0, g1): the dwords holding the
                // round's query bytes, which the global path reads too
                stage_copy<FAST_STAGE_B>(stage, g0, ndw, lane);
                sbase = g0;
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            }
        }
        do {
            if (!act) break;
            esp_fast_rec rec = {};
            if (!incontig) {
                rec.status = -1;
                a.out[idx] = rec;
                break;
            }
            const uint32_t na = a.anc_n[idx];
            const uint32_t *anc = a.anchors + 2 * a.anc_off[idx];
            // conditions 2-4 (align.rs:2284-2305)
            bool ok = na > 0;
            int64_t diag = 0;
            uint32_t qstart = 0, qend = 0, rs = 0;
            if (ok) {
                qstart = anc[0];
                rs = anc[1];
                diag = (int64_t)rs - (int64_t)qstart;
                uint32_t prevq = qstart;
                for (uint32_t j = 1; j < na && ok; j++) {
                    const uint32_t q = anc[2 * j], r = anc[2 * j + 1];
                    ok = (int64_t)r - (int64_t)q == diag && q - prevq <= a.k + a.w + 5u;
                    prevq = q;
                }
                qend = prevq + a.k;
                ok = ok && qstart <= 2 && m - qend <= 2;
            }
            if (!ok) {
                a.out[idx] = rec;
                break;
            }
            const uint32_t flank = m / 2 + 30;
            const uint32_t wstart = rs > flank ? rs - flank : 0u;
            const uint32_t *pk = (const uint32_t *)(a.R.packed + a.R.pk_off[contig]);
            const uint32_t clen = a.R.len[contig];
            const uint32_t *iv = a.R.iv + 2 * a.R.iv_off[contig];
            const uint32_t niv = (uint32_t)(a.R.iv_off[contig + 1] - a.R.iv_off[contig]);
            const uintptr_t qaddr = (uintptr_t)(a.reads + roff);
            const uint32_t *qw = (const uint32_t *)(qaddr & ~(uintptr_t)3);
            // the lane's query dwords in the wave's LDS copy, when the round was copied
            const uint32_t *sq = sbase ? stage + (uint32_t)(((uintptr_t)qw - sbase) >> 2) : nullptr;
            const uint32_t qoff0 = (uint32_t)(qaddr & 3u);
            const uint32_t qjmax = m ? (qoff0 + m - 1) >> 2 : 0u;
            // dwords the contig's region holds: ceil(len / 4) + 16 bytes, rounded up to 16
            const uint32_t pklast = ((((clen + 3u) >> 2) + 16u + 15u) & ~15u) / 4u - 1u;
            const uint32_t exag = a.edit_dir == 1 ? 0xFFFFFFFFu : 0u, extc = a.edit_dir == 2 ? 0xFFFFFFFFu : 0u;
            uint32_t ci = 0;
            {
                const int64_t p0 = diag > 0 ? diag : 0;
                ci = niv ? iv_seek(iv, niv, (uint32_t)min(p0, (int64_t)0xFFFFFFFF)) : 0u;
            }
            FastKadane K;
            for (uint32_t i0 = 0; i0 < m; i0 += FT) {
                // one tile: its query and reference words load at once, then the
                // FT columns run from registers
                const uint32_t ncol = min(m - i0, FT);
                uint32_t q[FT / 4 + 1], r[FT / 16 + 1];
                const uint32_t j0 = (qoff0 + i0) >> 2;
                if (sq) {
#pragma unroll
                    for (uint32_t j = 0; j <= FT / 4; j++) q[j] = sq[min(j0 + j, qjmax)];
                } else if (j0 + FT / 4 <= qjmax) {
#pragma unroll
                    for (uint32_t j = 0; j < FT / 16; j++) {
                        const u32x4 v = *(const u32x4 *)(qw + j0 + 4 * j);
                        q[4 * j] = v.x, q[4 * j + 1] = v.y, q[4 * j + 2] = v.z, q[4 * j + 3] = v.w;
                    }
                    q[FT / 4] = qw[j0 + FT / 4];
                } else {
#pragma unroll
                    for (uint32_t j = 0; j <= FT / 4; j++) q[j] = qw[min(j0 + j, qjmax)];
                }
                const uint32_t qsh = (qoff0 + i0) & 3u;
                const int64_t p0 = (int64_t)i0 + diag;
                while (ci < niv && (int64_t)iv[2 * ci + 1] <= p0) ci++;
                const bool inside = p0 >= 0 && p0 + FT <= (int64_t)clen;
                const bool clear = ci >= niv || (int64_t)iv[2 * ci] >= p0 + FT;
                const bool fast = inside && clear;
                const uint32_t d0 = fast ? (uint32_t)p0 >> 4 : 0u;
#pragma unroll
                for (uint32_t j = 0; j <= FT / 16; j++) r[j] = pk[min(d0 + j, pklast)];
                if (ncol == FT)
                    fast_tile<true>(a, K, lut, q, r, qsh, fast, p0, i0, FT, pk, pklast, iv, niv, clen, exag, extc);
                else
                    fast_tile<false>(a, K, lut, q, r, qsh, fast, p0, i0, ncol, pk, pklast, iv, niv, clen, exag, extc);
            }
            const int32_t best = K.best;
            const uint32_t bs = K.bs, be = K.be, bmm = K.bmm;
            if (best > 0 && bmm == 0) {
                uint32_t *cg = a.cigar + 3 * (size_t)idx;
                uint32_t n = 0;
                if (bs > 0) cg[n++] = bs << 4 | OPS;
                cg[n++] = (be - bs) << 4 | OPM;
                if (m > be) cg[n++] = (m - be) << 4 | OPS;
                rec.status = 1;
                rec.score = best;
                rec.wstart = wstart;
                rec.ref_start = (uint32_t)((int64_t)bs + diag - (int64_t)wstart);
                rec.read_start = bs;
                rec.read_end = be;
                rec.pos = wstart + rec.ref_start;
                rec.n_ops = n;
            }
            a.out[idx] = rec;
        } while (0);
        // the next round's copy overwrites this one only after every lane's
        // reads of it (same-wave LDS operations complete in order)
        __builtin_amdgcn_wave_barrier();
    }
}

}  // namespace

int walk_blocks_per_cu() {
    static int nb = 0;
    if (!nb && (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_walk, 256, 0) != hipSuccess || nb < 1))
        nb = 1;
    return nb;
}

int fast_lane_blocks_per_cu() {
    static int nb = 0;
    if (!nb && (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_fast_lane, 256, 0) != hipSuccess || nb < 1))
        nb = 1;
    return nb;
}

hipError_t launch_walk(const WalkArgs &a, int blocks, hipStream_t s) {
    hipLaunchKernelGGL(k_walk, dim3(blocks), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_fast_lane(const FastArgs &a, int blocks, hipStream_t s) {
    hipLaunchKernelGGL(k_fast_lane, dim3(blocks), dim3(256), 0, s, a);
    return hipGetLastError();
}

}  // namespace espk
