// esp_stitch.hpp -- stitch_cigar (crates/map/src/splice.rs:435-494), shared by
// the host entry point esp_stitch_cigar and the spliced-pipeline kernel.
//
// Walks a pseudo-reference CIGAR from pseudo coordinate p_start: Match/Del
// split at exon ends (take = min(left, remain_in_exon)), a RefSkip of the
// junction length goes in at every crossing, Ins/SoftClip/RefSkip pass
// through, and runs of one kind merge as push_op does (extend.rs:254-282).
// Ops are emitted through a one-op pending register, so the returned count is
// exact even when it exceeds `cap` (only the first `cap` ops are written).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace esps {

struct OpSink {
    uint32_t *out;
    uint32_t cap;
    uint32_t n = 0;       // ops flushed
    uint32_t kind = 16;   // pending op kind (16 = none)
    uint64_t len = 0;
    __host__ __device__ void flush() {
        if (kind < 16) {
            if (n < cap) out[n] = ((uint32_t)len << 4) | kind;
            n++;
        }
    }
    __host__ __device__ void push(uint32_t k, uint64_t l) {
        if (k == kind) {
            len += l;
        } else {
            flush();
            kind = k;
            len = l;
        }
    }
    __host__ __device__ uint32_t finish() {
        flush();
        kind = 16;
        return n;
    }
};

// bounds: 3 u32 per exon (pseudo_start, ref_start, ref_end).  Junction k's
// skip is skips[k] when skips != nullptr (k < n_junctions), else derived
// from adjacent exons as ref_start[k+1] - ref_end[k] (the junction of
// prepare_spliced / intron_chain sits exactly between them).
__host__ __device__ inline uint32_t stitch_walk(const uint32_t *cigar, uint32_t n_ops,
                                                uint32_t p_start, const uint32_t *bounds,
                                                uint32_t n_exons, const uint32_t *skips,
                                                uint32_t n_junctions, uint32_t *out,
                                                uint32_t cap) {
    OpSink s{out, cap};
    if (n_exons == 0) {
        // exon_bounds empty: the CIGAR is returned as is (no merging).
        for (uint32_t k = 0; k < n_ops; k++)
            if (k < cap) out[k] = cigar[k];
        return n_ops;
    }
    auto ps = [&](uint32_t e) -> uint64_t { return bounds[3 * e]; };
    auto exon_end = [&](uint32_t e) -> uint64_t {
        return (uint64_t)bounds[3 * e] + (uint64_t)(bounds[3 * e + 2] - bounds[3 * e + 1]);
    };
    auto skip = [&](uint32_t e) -> uint32_t {
        return skips ? skips[e] : bounds[3 * (e + 1) + 1] - bounds[3 * e + 2];
    };
    uint64_t p = p_start;
    uint32_t ei = 0;
    while (ei + 1 < n_exons && p >= exon_end(ei)) ei++;
    for (uint32_t k = 0; k < n_ops; k++) {
        const uint32_t kind = cigar[k] & 0xFu;
        const uint32_t len = cigar[k] >> 4;
        if (kind != 0u && kind != 2u) {  // Ins, RefSkip, SoftClip pass through
            s.push(kind, len);
            continue;
        }
        uint64_t remaining = len;
        while (remaining > 0) {
            if (ei >= n_exons) {
                s.push(kind, remaining);
                p += remaining;
                break;
            }
            if (p < ps(ei)) p = ps(ei);  // snap into the exon slice
            const uint64_t ee = exon_end(ei);
            const uint64_t in_exon = ee > p ? ee - p : 0;
            const uint64_t take = remaining < in_exon ? remaining : in_exon;
            if (take > 0) s.push(kind, take);
            p += take;
            remaining -= take;
            if (remaining > 0) {
                if (ei + 1 >= n_exons) {
                    s.push(kind, remaining);
                    p += remaining;
                    break;
                }
                if (ei < n_junctions) s.push(3u, skip(ei));
                ei++;
            }
        }
    }
    return s.finish();
}

// SplicedAlignment::genomic_start (splice.rs:414-425): the pseudo
// coordinate p mapped through the exon bounds; past the last bound it maps
// off the last exon.
__host__ __device__ inline uint32_t genomic_start(uint32_t p, const uint32_t *bounds,
                                                  uint32_t n_exons) {
    for (uint32_t e = 0; e < n_exons; e++) {
        const uint32_t q = bounds[3 * e], rs = bounds[3 * e + 1], re = bounds[3 * e + 2];
        if (p >= q && p < q + (re - rs)) return rs + (p - q);
    }
    const uint32_t q = bounds[3 * (n_exons - 1)], rs = bounds[3 * (n_exons - 1) + 1];
    return rs + (p > q ? p - q : 0u);
}

}  // namespace esps
