// esp_walk.hip -- placement walks over the resident reference.
//
// One lane per placement walks its CIGAR against the 2-bit reference and the
// query, in the reference's order (align.rs:1060-1163 rescore_with,
// :2473-2516 recount_mm_ea_in, :1231-1288 exact_kmer_support), in one pass:
//   * M columns: a reference byte from a per-lane cursor (a cached dword of
//     16 packed bases, and a forward cursor over the contig's sorted N
//     intervals, so no per-base binary search), a query byte from a cached
//     dword of the read; strict score and counts per column; the k-mer scan
//     as a streaming window (a bad column restarts the window after it,
//     which is where the reference's one-step rescan first succeeds);
//   * I/D/S: the gap and clip bookkeeping;
//   * N: the junction library (binary search of sorted triples), else the
//     donor/acceptor signal on both strands minus the intron-length penalty
//     (a host-built threshold table, exact against the f64 formula).
#include "esp_kernels.hpp"

using namespace espd;

namespace espk {

namespace {

constexpr uint32_t OPM = 0, OPI = 1, OPD = 2, OPN = 3, OPS = 4;

// Reference bytes of one contig read mostly in ascending position order.
struct RefCursor {
    const uint32_t *pk;  // the contig's packed bytes as dwords (16-byte aligned)
    const uint32_t *iv;  // (start, end) of its N intervals
    uint32_t len, niv;
    uint32_t ci = 0;            // first interval whose end > the last position
    uint32_t last = 0;
    uint32_t wi = 0xFFFFFFFFu;  // cached dword index
    uint32_t w = 0;

    __device__ void seek(uint32_t p) {
        uint32_t lo = 0, hi = niv;
        while (lo < hi) {
            const uint32_t mid = (lo + hi) >> 1;
            if (iv[2 * mid + 1] <= p) lo = mid + 1;
            else hi = mid;
        }
        ci = lo;
        last = p;
    }
    // ASCII base at p ('N' inside an N interval or past the contig end).
    __device__ uint32_t get(uint32_t p) {
        if (p >= len) return 'N';
        if (p < last) seek(p);
        while (ci < niv && iv[2 * ci + 1] <= p) ci++;
        last = p;
        if (ci < niv && iv[2 * ci] <= p) return 'N';
        const uint32_t d = p >> 4;
        if (d != wi) {
            wi = d;
            w = pk[d];
        }
        const uint32_t byte = (w >> (8 * ((p >> 2) & 3u))) & 0xFFu;
        const uint32_t code = (byte >> (6 - 2 * (p & 3u))) & 3u;
        return (0x54474341u >> (8 * code)) & 0xFFu;  // "ACGT"
    }
};

// Query bytes read in ascending order through a cached aligned dword; the
// dword holds the byte asked for, so no load leaves the read's pages.
struct QueryCursor {
    const uint8_t *q;
    uint32_t len;
    uintptr_t wa = 0;
    uint32_t w = 0;
    __device__ uint32_t get(uint32_t i) {
        if (i >= len) return 'N';
        const uintptr_t addr = (uintptr_t)(q + i);
        const uintptr_t a = addr & ~(uintptr_t)3;
        if (a != wa) {
            wa = a;
            w = *(const uint32_t *)a;
        }
        return (w >> (8 * (uint32_t)(addr & 3u))) & 0xFFu;
    }
};

// exempt_dir (extend.rs:86-93): 1 ref A / read G, 0 ref T / read C, -1 else.
__device__ __forceinline__ int exempt_dir_b(uint32_t rb, uint32_t qb) {
    return (rb == 'A' && qb == 'G') ? 1 : (rb == 'T' && qb == 'C') ? 0 : -1;
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

__device__ int32_t intron_penalty(const WalkArgs &a, uint32_t len) {
    int32_t v = 0;
    if (len >= 1000) {
        for (int t = 0; t < a.nup; t++) v += a.up[t] <= len;
    } else {
        for (int t = 0; t < a.ndown; t++) v += a.down[t] >= len;
    }
    return v;
}

__device__ bool junction_known(const WalkArgs &a, uint32_t c, uint32_t s, uint32_t e) {
    uint32_t lo = 0, hi = a.nj;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        const uint32_t *t = a.jn + 3 * (size_t)mid;
        const bool less = t[0] != c ? t[0] < c : (t[1] != s ? t[1] < s : t[2] < e);
        if (!less && t[0] == c && t[1] == s && t[2] == e) return true;
        if (less) lo = mid + 1;
        else hi = mid;
    }
    return false;
}

__global__ __launch_bounds__(256) void k_walk(WalkArgs a) {
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x; idx < a.n; idx += stride) {
        esp_strict_score s = {};
        esp_walk_counts cnt = {};
        const uint32_t contig = a.contig[idx];
        if (contig >= a.R.ncontig) {
            if (a.strict) a.strict[idx] = s;
            if (a.counts) a.counts[idx] = cnt;
            if (a.status) a.status[idx] = ESP_E_INVAL;
            continue;
        }
        RefCursor ref;
        ref.pk = (const uint32_t *)(a.R.packed + a.R.pk_off[contig]);
        ref.iv = a.R.iv + 2 * a.R.iv_off[contig];
        ref.niv = (uint32_t)(a.R.iv_off[contig + 1] - a.R.iv_off[contig]);
        ref.len = a.R.len[contig];
        QueryCursor q;
        q.q = a.reads + a.read_off[idx];
        q.len = a.read_len[idx];
        const int dir = a.dir ? a.dir[idx] : 0;
        const uint32_t k = a.k;
        uint32_t rf = a.pos[idx], ro = 0;
        ref.seek(rf);
        int32_t score = 0;
        const uint32_t *cig = a.cigar + a.cig_off[idx];
        const uint32_t nops = a.cig_n[idx];
        for (uint32_t o = 0; o < nops; o++) {
            const uint32_t op = cig[o] & 15u, n = cig[o] >> 4;
            if (op == OPM) {
                uint32_t kstart = 0;  // candidate k-mer start inside this block
                for (uint32_t x = 0; x < n; x++) {
                    const uint32_t qb = q.get(ro + x), rb = ref.get(rf + x);
                    const int ed = exempt_dir_b(rb, qb);
                    const bool exempt = (dir == 1 && ed == 1) || (dir == 2 && ed == 0);
                    if (qb != 'N' && rb != 'N') {
                        s.aligned++;
                        if (qb == rb) {
                            score += a.match;
                            s.matches++;
                        } else {
                            if (!exempt) score += a.mismatch;
                            if (ed == 1) s.ea_ag++;
                            else if (ed == 0) s.ea_tc++;
                            else s.mm++;
                        }
                    }
                    // k-mer scan: exact in the reduced alphabet of dir
                    if (k) {
                        if (qb != rb && !exempt) {
                            kstart = x + 1;
                        } else if (x + 1 - kstart == k) {
                            cnt.kmers++;
                            kstart = x + 1;
                        }
                    }
                }
                ro += n;
                rf += n;
            } else if (op == OPI) {
                score -= a.gap_open + a.gap_ext * (int32_t)n;
                s.aligned += n;
                s.ins++;
                ro += n;
            } else if (op == OPD) {
                score -= a.gap_open + a.gap_ext * (int32_t)n;
                s.aligned += n;
                s.del++;
                rf += n;
            } else if (op == OPS) {
                s.clipped += n;
                ro += n;
            } else if (op == OPN) {
                const int32_t before = score;
                if (junction_known(a, contig, rf, rf + n)) {
                    score += a.known;
                } else {
                    const uint32_t d0 = ref.get(rf), d1 = ref.get(rf == 0xFFFFFFFFu ? rf : rf + 1);
                    const uint32_t a0 = ref.get(rf + (n >= 2 ? n - 2 : 0u));
                    const uint32_t a1 = ref.get(rf + (n >= 1 ? n - 1 : 0u));
                    const int32_t sp = signal_score(a, d0, d1, a0, a1, false);
                    const int32_t sm = signal_score(a, d0, d1, a0, a1, true);
                    score += max(sp, sm) - intron_penalty(a, n);
                }
                s.n_splice++;
                s.splice_score += score - before;
                rf += n;
            }
        }
        s.score = score;
        cnt.mm = s.mm;
        cnt.ea_ag = s.ea_ag;
        cnt.ea_tc = s.ea_tc;
        if (!q.len) cnt.kmers = 0;  // exact_kmer_support of an empty query
        if (a.strict) a.strict[idx] = s;
        if (a.counts) a.counts[idx] = cnt;
        if (a.status) a.status[idx] = 0;
    }
}

}  // namespace

hipError_t launch_walk(const WalkArgs &a, int blocks, hipStream_t s) {
    hipLaunchKernelGGL(k_walk, dim3(blocks), dim3(256), 0, s, a);
    return hipGetLastError();
}

}  // namespace espk
