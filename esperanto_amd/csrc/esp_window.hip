// esp_window.hip -- window gather from the resident 2-bit reference and the
// per-pair stages of the mate-rescue and spliced window pipelines.
//
// Gather: one wave per window; each lane decodes 16 consecutive bases per
// pass from a 5-byte big-endian window of the packed contig (base at pos in
// byte pos/4 at shift 6 - 2*(pos%4), fasta.rs:1-6), masks N intervals found
// by one binary search per 16-base run (the reference searches per base,
// fasta.rs:119-135; the intervals are sorted and disjoint, so the masks are
// the same), and stores the run as 4 dwords when aligned.
#include "esp_kernels.hpp"
#include "esp_refdev.hpp"
#include "esp_stitch.hpp"

using namespace espd;

namespace espk {

namespace {

constexpr int WIN_BLOCK = 256;  // 4 waves, one window (or pair) each

// Complement of one raw read byte through Base::from_ascii (fasta.rs:30-39):
// A/a -> T, C/c -> G, G/g -> C, T/t -> A, anything else -> N.
__device__ __forceinline__ uint32_t comp_ascii(uint32_t b) {
    const uint32_t u = b & 0xDFu;
    return u == 'A' ? 'T' : u == 'C' ? 'G' : u == 'G' ? 'C' : u == 'T' ? 'A' : 'N';
}

// Wave-cooperative copy of the read and of its reverse complement
// (revcomp_into, align.rs:2525-2533).
__device__ void wave_read_pair(const uint8_t *src, uint32_t m, uint8_t *fwd, uint8_t *rcd,
                               uint32_t lane) {
    for (uint32_t k0 = lane * 16; k0 < m; k0 += 1024) {
        const uint32_t cnt = min(16u, m - k0);
        uint32_t wf[4] = {0u, 0u, 0u, 0u}, wr[4] = {0u, 0u, 0u, 0u};
#pragma unroll
        for (uint32_t t = 0; t < 16; t++) {
            if (t < cnt) {
                wf[t >> 2] |= (uint32_t)src[k0 + t] << (8 * (t & 3));
                wr[t >> 2] |= comp_ascii(src[m - 1 - (k0 + t)]) << (8 * (t & 3));
            }
        }
        store_run(fwd + k0, wf, cnt);
        store_run(rcd + k0, wr, cnt);
    }
}

__global__ __launch_bounds__(WIN_BLOCK) void k_gather(GatherArgs a) {
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t k = blockIdx.x * (WIN_BLOCK / 64) + (threadIdx.x >> 6);
    if (k >= a.njobs) return;
    const esp_gather_job j = a.jobs[k];
    const bool ok = j.contig < a.R.ncontig && j.start <= j.end && j.end <= a.R.len[j.contig];
    if (lane == 0 && a.status) a.status[k] = ok ? 0 : (int32_t)ESP_E_INVAL;
    if (!ok) return;
    wave_decode(a.R, j.contig, j.start, j.end, (j.flags & 1u) != 0, a.dst + j.dst_off, lane);
}

// Mate rescue, stage 1 (align.rs:857-874): geometry, window decode, read and
// reverse-complement slots, Myers descriptors.
__global__ __launch_bounds__(WIN_BLOCK) void k_mate_prep(MateArgs a) {
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t k = blockIdx.x * (WIN_BLOCK / 64) + (threadIdx.x >> 6);
    if (k >= a.npairs) return;
    const uint32_t m = a.read_len[k];
    const uint32_t ctg = a.mate_contig[k];
    int32_t state = 1;
    uint32_t ws = 0, wl = 0;
    if (m == 0 || m > 256) {
        state = 0;
    } else if (m > a.max_read || ctg >= a.R.ncontig) {
        state = -1;
    } else {
        const uint32_t mp = a.mate_pos[k];
        const uint32_t span = 750u + m;  // MATE_WINDOW_INSERT * 3/2 + read_len
        const uint32_t wstart = mp > span ? mp - span : 0u;
        const uint64_t sat = min((uint64_t)mp + span + m, (uint64_t)0xFFFFFFFFu);
        const uint32_t wend = (uint32_t)min(sat, (uint64_t)a.R.len[ctg]);
        if (wend <= wstart || wend - wstart < m) {
            state = 0;
        } else {
            ws = wstart;
            wl = wend - wstart;
        }
    }
    const uint64_t qo = (uint64_t)k * a.qstride;
    if (lane == 0) {
        a.state[k] = state;
        a.wstart[k] = ws;
        a.w_len[k] = wl;
        a.w_off[k] = (uint64_t)k * a.wstride;
        a.q_off_plus[k] = qo;
        a.q_off_minus[k] = qo + a.qstride / 2;
        a.q_len[k] = state == 1 ? m : 0u;
    }
    if (state != 1) return;
    wave_decode(a.R, ctg, ws, ws + wl, false, a.warena + (uint64_t)k * a.wstride, lane);
    wave_read_pair(a.reads + a.read_off[k], m, a.qarena + qo, a.qarena + qo + a.qstride / 2, lane);
}

// Mate rescue, stage 2 (align.rs:875-902): threshold max(m/3, 10), the
// start + m <= window guard, Plus first and Minus only when strictly closer;
// descriptors of the extend_hint batch (read_len 0 when nothing qualifies).
__global__ __launch_bounds__(WIN_BLOCK) void k_mate_select(MateArgs a) {
    const uint32_t k = blockIdx.x * WIN_BLOCK + threadIdx.x;
    if (k >= a.npairs) return;
    int32_t sel = 0, d = 0;
    uint64_t start = 0;
    if (a.state[k] == 1) {
        const uint32_t m = a.read_len[k];
        const int32_t thr = max((int32_t)m / 3, 10);
        const uint64_t wl = a.w_len[k];
        if (a.d_plus[k] <= thr && a.s_plus[k] + m <= wl) {
            sel = 1;
            d = a.d_plus[k];
            start = a.s_plus[k];
        }
        if (a.d_minus[k] <= thr && a.s_minus[k] + m <= wl && (sel == 0 || a.d_minus[k] < d)) {
            sel = 2;
            d = a.d_minus[k];
            start = a.s_minus[k];
        }
    }
    a.sel[k] = sel;
    a.dist[k] = d;
    a.g_read_off[k] = sel == 2 ? a.q_off_minus[k] : a.q_off_plus[k];
    a.g_read_len[k] = sel ? a.read_len[k] : 0u;
    a.g_offset[k] = (int64_t)start;
}

// Mate rescue, stage 3 (align.rs:903-920): placement when the score is positive.
__global__ __launch_bounds__(WIN_BLOCK) void k_mate_finish(MateArgs a) {
    const uint32_t k = blockIdx.x * WIN_BLOCK + threadIdx.x;
    if (k >= a.npairs) return;
    const int32_t state = a.state[k], sel = a.sel[k];
    esp_mate_rec r;
    r.strand = sel == 2 ? 1 : 0;
    r.dist = a.dist[k];
    r.wstart = a.wstart[k];
    r.wlen = a.w_len[k];
    if (sel) {
        const esp_ext e = a.ext[k];
        r.placed = e.score > 0 ? 1 : 0;
        r.pos = r.wstart + e.ref_start;
        r.score = e.score;
        r.read_start = e.read_start;
        r.read_end = e.read_end;
    } else {
        r.placed = state < 0 ? -1 : 0;
        r.pos = 0;
        r.score = 0;
        r.read_start = 0;
        r.read_end = 0;
    }
    a.out[k] = r;
    if (r.placed != 1) a.cig_n[k] = 0;
}

// Spliced pipeline, stage 1 (splice.rs:589-608, intron_chain.rs:497-505):
// exon checks, exon bounds and the concatenated pseudo-reference.
__global__ __launch_bounds__(WIN_BLOCK) void k_splice_prep(SpliceArgs a) {
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t k = blockIdx.x * (WIN_BLOCK / 64) + (threadIdx.x >> 6);
    if (k >= a.npairs) return;
    const uint32_t ctg = a.contig[k];
    const uint64_t e0 = a.exon_off[k], e1 = a.exon_off[k + 1];
    int32_t state = 1;
    uint64_t total = 0;
    if (ctg >= a.R.ncontig || e1 <= e0 || e1 > a.n_exons ||
        a.pseudo_off[k + 1] > a.pseudo_bytes || a.pseudo_off[k + 1] < a.pseudo_off[k]) {
        state = -1;
    } else {
        const uint32_t clen = a.R.len[ctg];
        bool rejected = false;
        for (uint64_t e = e0; e < e1; e++) {
            const uint32_t rs = a.exons[2 * e], re = a.exons[2 * e + 1];
            if (re <= rs) rejected = true;
            else if (re > clen) state = -1;
            else total += re - rs;
        }
        if (state == 1 && rejected) state = 0;
        if (state == 1 && a.pseudo_off[k + 1] - a.pseudo_off[k] != total) state = -1;
    }
    if (lane == 0) {
        a.state[k] = state;
        a.p_len[k] = state == 1 ? (uint32_t)total : 0u;
        a.g_read_len[k] = state == 1 ? a.read_len[k] : 0u;
    }
    if (state != 1) return;
    uint8_t *dst = a.parena + a.pseudo_off[k];
    uint32_t ps = 0;
    for (uint64_t e = e0; e < e1; e++) {
        const uint32_t rs = a.exons[2 * e], re = a.exons[2 * e + 1];
        if (lane == 0) {
            a.bounds[3 * e] = ps;
            a.bounds[3 * e + 1] = rs;
            a.bounds[3 * e + 2] = re;
        }
        wave_decode(a.R, ctg, rs, re, false, dst + ps, lane);
        ps += re - rs;
    }
}

// Spliced pipeline, stage 2 (splice.rs:651-663): stitch the pseudo-reference
// CIGAR into genomic coordinates and map the start back.
__global__ __launch_bounds__(WIN_BLOCK) void k_splice_finish(SpliceArgs a) {
    const uint32_t k = blockIdx.x * WIN_BLOCK + threadIdx.x;
    if (k >= a.npairs) return;
    esp_splice_rec r;
    r.status = a.state[k];
    r.pseudo_len = a.p_len[k];
    r.genomic_start = 0;
    r.n_stitched = 0;
    if (r.status == 1) {
        const uint64_t e0 = a.exon_off[k];
        const uint32_t nex = (uint32_t)(a.exon_off[k + 1] - e0);
        const esp_ext e = a.ext[k];
        const uint32_t *bounds = a.bounds + 3 * e0;
        const uint64_t so = a.stitch_off[k];
        const uint32_t cap = (uint32_t)min(a.stitch_off[k + 1] - so, (uint64_t)0xFFFFFFFFu);
        const uint32_t n = esps::stitch_walk(a.cigar + a.cig_off[k], a.cig_n[k], e.ref_start,
                                             bounds, nex, nullptr, nex - 1, a.stitched + so, cap);
        r.n_stitched = n;
        if (n > cap) r.status = -2;
        r.genomic_start = esps::genomic_start(e.ref_start, bounds, nex);
    }
    a.rec[k] = r;
}

// Window EA-Myers jobs: checks, the text decoded into the job's slot, and
// the descriptors of the Myers launch (an invalid job gets a 0-length
// pattern, which the Myers kernel answers with (0, 0)).
__global__ __launch_bounds__(WIN_BLOCK) void k_winmyers_prep(WinMyersArgs a) {
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t k = blockIdx.x * (WIN_BLOCK / 64) + (threadIdx.x >> 6);
    if (k >= a.njobs) return;
    const esp_window_job j = a.jobs[k];
    const bool ok = j.contig < a.R.ncontig && j.tstart <= j.tend && j.tend <= a.R.len[j.contig] &&
                    j.qstart <= j.qend && j.qend - j.qstart <= a.max_pattern &&
                    j.tend - j.tstart <= a.max_text;
    const uint64_t to = (uint64_t)k * a.tstride;
    if (lane == 0) {
        if (a.status) a.status[k] = ok ? 0 : (int32_t)ESP_E_INVAL;
        a.p_off[k] = j.read_off + j.qstart;
        a.p_len[k] = ok ? j.qend - j.qstart : 0u;
        a.t_off[k] = to;
        a.t_len[k] = ok ? j.tend - j.tstart : 0u;
    }
    if (ok) wave_decode(a.R, j.contig, j.tstart, j.tend, false, a.tarena + to, lane);
}

// continuous_extension_window (align.rs:33-55), u32 saturating arithmetic.
__device__ void ext_window(uint32_t rs, uint32_t re, uint32_t qs, uint32_t qe, uint32_t read_len,
                           uint32_t clen, uint32_t &start, uint32_t &end) {
    auto sadd = [](uint32_t x, uint32_t y) { return (uint32_t)min((uint64_t)x + y, (uint64_t)0xFFFFFFFFu); };
    const uint32_t flank = read_len / 2 + 30;
    const uint32_t head = max(flank, qs);
    const uint32_t tail = max(flank, read_len > qe ? read_len - qe : 0u);
    const uint32_t st = min(rs > head ? rs - head : 0u, clen);
    uint32_t en = min(sadd(re, tail), clen);
    en = min(en, sadd(sadd(rs, read_len), flank));
    en = min(en, sadd(sadd(st, read_len), sadd(flank, flank)));
    start = st;
    end = max(en, st);
}

// Unspliced chain, stage 1 (align.rs:1776-1790, :1895-1906): window, decoded
// window, the read in alignment orientation, Myers descriptors.
__global__ __launch_bounds__(WIN_BLOCK) void k_unsp_prep(UnspArgs a) {
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t k = blockIdx.x * (WIN_BLOCK / 64) + (threadIdx.x >> 6);
    if (k >= a.npairs) return;
    const uint32_t m = a.read_len[k], ctg = a.contig[k];
    int32_t state = 1;
    uint32_t ws = 0, we = 0;
    if (ctg >= a.R.ncontig || m > a.max_read) {
        state = -1;
    } else {
        ext_window(a.rs[k], a.re[k], a.qs[k], a.qe[k], m, a.R.len[ctg], ws, we);
    }
    const uint64_t qo = (uint64_t)k * a.qstride, wo = (uint64_t)k * a.wstride;
    if (lane == 0) {
        a.state[k] = state;
        a.wstart[k] = ws;
        a.w_len[k] = we - ws;
        a.w_off[k] = wo;
        a.q_off[k] = qo;
        a.q_len[k] = (state == 1 && m >= 1 && m <= 256 && we - ws >= m) ? m : 0u;
    }
    if (state != 1) return;
    wave_decode(a.R, ctg, ws, we, false, a.warena + wo, lane);
    const uint8_t *src = a.reads + a.read_off[k];
    uint8_t *q = a.qarena + qo;
    const bool rc = a.minus[k] != 0;
    for (uint32_t k0 = lane * 16; k0 < m; k0 += 1024) {
        const uint32_t cnt = min(16u, m - k0);
        uint32_t wq[4] = {0u, 0u, 0u, 0u};
#pragma unroll
        for (uint32_t t = 0; t < 16; t++)
            if (t < cnt)
                wq[t >> 2] |= (rc ? comp_ascii(src[m - 1 - (k0 + t)]) : (uint32_t)src[k0 + t]) << (8 * (t & 3));
        store_run(q + k0, wq, cnt);
    }
}

// extend.rs:87-93 on raw bytes: 1 plus (ref A / read G), 0 minus (ref T / read C), -1 none.
__device__ __forceinline__ int32_t exempt_dir_b(uint32_t ref_b, uint32_t read_b) {
    return (ref_b == 'A' && read_b == 'G') ? 1 : (ref_b == 'T' && read_b == 'C') ? 0 : -1;
}

// Unspliced chain, stage 2 (align.rs:1807-1877): the distance gate, the
// per-base recount, Kadane over the diagonal; else the DP descriptors.
__global__ __launch_bounds__(WIN_BLOCK) void k_unsp_ea(UnspArgs a) {
    const uint32_t k = blockIdx.x * WIN_BLOCK + threadIdx.x;
    if (k >= a.npairs) return;
    const uint32_t m = a.read_len[k];
    int32_t sum = 0;
    uint32_t bi = 0, bj = 0;
    bool ea = false;
    if (a.q_len[k]) {
        const int32_t thr = max((int32_t)m / 3, 10), d = a.dist[k];
        const uint64_t start = a.s_start[k];
        const uint8_t *q = a.qarena + a.q_off[k], *w = a.warena + a.w_off[k] + start;
        bool usable = d <= thr && start + m <= a.w_len[k];
        if (usable) {
            int32_t mm = 0;
            for (uint32_t t = 0; t < m; t++) {
                const uint32_t qb = q[t], rb = w[t];
                if (qb == 'N' || rb == 'N') {
                    usable = false;
                    break;
                }
                if (qb != rb && exempt_dir_b(rb, qb) < 0) mm++;
            }
            if (mm != d) usable = false;
        }
        if (usable) {
            int32_t cur = 0;
            uint32_t ci = 0;
            for (uint32_t t = 0; t < m; t++) {
                const uint32_t qb = q[t], rb = w[t];
                const int32_t ed = exempt_dir_b(rb, qb);
                const bool ex = (a.P.edit == 1 && ed == 1) || (a.P.edit == 2 && ed == 0);
                const int32_t s = qb == rb ? a.P.match : (ex ? 0 : a.P.mismatch);
                if (cur <= 0) {
                    cur = s;
                    ci = t;
                } else {
                    cur += s;
                }
                if (cur > sum) {
                    sum = cur;
                    bi = ci;
                    bj = t + 1;
                }
            }
            ea = sum > 0 && (bj - bi) * 2 >= m;
        }
    }
    a.ea_sum[k] = ea ? sum : 0;
    a.ea_bi[k] = bi;
    a.ea_bj[k] = bj;
    const bool dp = !ea && a.state[k] == 1 && a.allow_dp[k] != 0;
    a.g_read_off[k] = a.q_off[k];
    a.g_read_len[k] = dp ? m : 0u;
    a.g_hint[k] = (int64_t)(a.rs[k] - a.wstart[k]) - (int64_t)a.qs[k];
}

// Unspliced chain, stage 3 (align.rs:1850-1918): the placement.
__global__ __launch_bounds__(WIN_BLOCK) void k_unsp_finish(UnspArgs a) {
    const uint32_t k = blockIdx.x * WIN_BLOCK + threadIdx.x;
    if (k >= a.npairs) return;
    esp_unspliced_rec r;
    r.status = a.state[k] < 0 ? -1 : 0;
    r.strand = a.minus[k] ? 1 : 0;
    r.pos = 0;
    r.score = 0;
    r.dist = 0;
    r.wstart = a.wstart[k];
    r.wlen = a.w_len[k];
    uint32_t nops = 0;
    const uint64_t c0 = a.cig_off[k];
    if (a.ea_sum[k] > 0) {
        const uint32_t m = a.read_len[k], bi = a.ea_bi[k], bj = a.ea_bj[k];
        uint32_t *cg = a.cigar + c0;
        if (bi > 0) cg[nops++] = (bi << 4) | OP_S;
        cg[nops++] = ((bj - bi) << 4) | OP_M;
        if (m > bj) cg[nops++] = ((m - bj) << 4) | OP_S;
        r.status = 1;
        r.pos = r.wstart + (uint32_t)a.s_start[k] + bi;
        r.score = a.ea_sum[k];
        r.dist = a.dist[k];
    } else if (a.g_read_len[k] && a.ext[k].score > 0) {
        r.status = 2;
        r.pos = r.wstart + a.ext[k].ref_start;
        r.score = a.ext[k].score;
        nops = a.cig_n[k];
    }
    a.out[k] = r;
    a.cig_n[k] = nops;
}

uint32_t wave_blocks(uint32_t n) { return (n + WIN_BLOCK / 64 - 1) / (WIN_BLOCK / 64); }
uint32_t lane_blocks(uint32_t n) { return (n + WIN_BLOCK - 1) / WIN_BLOCK; }

}  // namespace

hipError_t launch_gather(const GatherArgs &a, hipStream_t s) {
    if (a.njobs == 0) return hipSuccess;
    hipLaunchKernelGGL(k_gather, dim3(wave_blocks(a.njobs)), dim3(WIN_BLOCK), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_mate_prep(const MateArgs &a, hipStream_t s) {
    hipLaunchKernelGGL(k_mate_prep, dim3(wave_blocks(a.npairs)), dim3(WIN_BLOCK), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_mate_select(const MateArgs &a, hipStream_t s) {
    hipLaunchKernelGGL(k_mate_select, dim3(lane_blocks(a.npairs)), dim3(WIN_BLOCK), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_mate_finish(const MateArgs &a, hipStream_t s) {
    hipLaunchKernelGGL(k_mate_finish, dim3(lane_blocks(a.npairs)), dim3(WIN_BLOCK), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_winmyers_prep(const WinMyersArgs &a, hipStream_t s) {
    if (a.njobs == 0) return hipSuccess;
    hipLaunchKernelGGL(k_winmyers_prep, dim3(wave_blocks(a.njobs)), dim3(WIN_BLOCK), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_unsp_prep(const UnspArgs &a, hipStream_t s) {
    hipLaunchKernelGGL(k_unsp_prep, dim3(wave_blocks(a.npairs)), dim3(WIN_BLOCK), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_unsp_ea(const UnspArgs &a, hipStream_t s) {
    hipLaunchKernelGGL(k_unsp_ea, dim3(lane_blocks(a.npairs)), dim3(WIN_BLOCK), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_unsp_finish(const UnspArgs &a, hipStream_t s) {
    hipLaunchKernelGGL(k_unsp_finish, dim3(lane_blocks(a.npairs)), dim3(WIN_BLOCK), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_splice_prep(const SpliceArgs &a, hipStream_t s) {
    hipLaunchKernelGGL(k_splice_prep, dim3(wave_blocks(a.npairs)), dim3(WIN_BLOCK), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_splice_finish(const SpliceArgs &a, hipStream_t s) {
    hipLaunchKernelGGL(k_splice_finish, dim3(lane_blocks(a.npairs)), dim3(WIN_BLOCK), 0, s, a);
    return hipGetLastError();
}

}  // namespace espk
