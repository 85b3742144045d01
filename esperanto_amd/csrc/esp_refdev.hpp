// esp_refdev.hpp -- device reads of the resident 2-bit reference (fasta.rs:
// base at pos in byte pos/4 at shift 6 - 2*(pos%4); N over the contig's
// sorted, disjoint N intervals), shared by the window kernels.
#pragma once

#include "esp_kernels.hpp"

namespace espk {

constexpr uint32_t ACGT_ASCII = 0x54474341u;  // 'A' 'C' 'G' 'T' by code

// Writes cnt (<= 16) bytes held little-endian in w[4] to d.
__device__ __forceinline__ inline void store_run(uint8_t *d, const uint32_t w[4], uint32_t cnt) {
    if (cnt == 16 && ((uintptr_t)d & 3u) == 0) {
        uint32_t *d4 = (uint32_t *)d;
#pragma unroll
        for (int q = 0; q < 4; q++) d4[q] = w[q];
    } else {
#pragma unroll
        for (uint32_t t = 0; t < 16; t++)
            if (t < cnt) d[t] = (uint8_t)(w[t >> 2] >> (8 * (t & 3)));
    }
}

// Bit mask over [lo, lo + cnt) of the positions an N interval covers.
__device__ inline uint32_t n_mask(const RefView &R, uint32_t ctg, uint32_t lo, uint32_t cnt) {
    uint64_t a = R.iv_off[ctg], b = R.iv_off[ctg + 1];
    const uint64_t hi = b;
    if (a == b) return 0u;
    while (a < b) {  // first interval ending after lo
        const uint64_t mid = (a + b) >> 1;
        if (R.iv[2 * mid + 1] > lo) b = mid;
        else a = mid + 1;
    }
    uint32_t mask = 0;
    for (uint64_t j = a; j < hi && R.iv[2 * j] < lo + cnt; j++) {
        const uint32_t s = max(R.iv[2 * j], lo) - lo;
        const uint32_t e = min(R.iv[2 * j + 1], lo + cnt) - lo;
        mask |= ((1u << e) - 1u) & ~((1u << s) - 1u);
    }
    return mask;
}

// Wave-cooperative decode of contig[start, end) to dst, reverse-complemented
// when rc (output byte k is the complement of base end - 1 - k; N stays N).
__device__ inline void wave_decode(const RefView &R, uint32_t ctg, uint32_t start, uint32_t end, bool rc,
                            uint8_t *dst, uint32_t lane) {
    const uint8_t *pk = R.packed + R.pk_off[ctg];
    const uint32_t L = end - start;
    for (uint32_t k0 = lane * 16; k0 < L; k0 += 1024) {
        const uint32_t cnt = min(16u, L - k0);
        const uint32_t lo = rc ? end - k0 - cnt : start + k0;
        const uint32_t b0 = lo >> 2;
        uint64_t win = 0;
#pragma unroll
        for (int q = 0; q < 5; q++) win = (win << 8) | pk[b0 + q];
        const uint32_t nm = n_mask(R, ctg, lo, cnt);
        uint32_t w[4] = {0u, 0u, 0u, 0u};
#pragma unroll
        for (uint32_t t = 0; t < 16; t++) {
            if (t < cnt) {
                const uint32_t r = rc ? cnt - 1 - t : t;   // run position of output byte t
                const uint32_t rel = (lo & 3u) + r;        // base index from byte b0
                uint32_t code = (uint32_t)(win >> (38 - 2 * rel)) & 3u;
                if (rc) code ^= 3u;                        // A<->T, C<->G
                uint32_t ch = (ACGT_ASCII >> (8 * code)) & 0xFFu;
                if ((nm >> r) & 1u) ch = 'N';
                w[t >> 2] |= ch << (8 * (t & 3));
            }
        }
        store_run(dst + k0, w, cnt);
    }
}

// ASCII base of contig ctg at p < its length (Contig::base, fasta.rs:105-117):
// N inside an N interval (one binary search), else A C G T.
__device__ inline uint32_t base_ascii(const RefView &R, uint32_t ctg, uint32_t p) {
    const uint64_t a1 = R.iv_off[ctg + 1];
    uint64_t lo = R.iv_off[ctg], hi = a1;
    while (lo < hi) {  // first interval ending after p
        const uint64_t mid = (lo + hi) >> 1;
        if (R.iv[2 * mid + 1] <= p) lo = mid + 1;
        else hi = mid;
    }
    if (lo < a1 && R.iv[2 * lo] <= p) return 'N';
    const uint32_t byte = R.packed[R.pk_off[ctg] + (p >> 2)];
    return (ACGT_ASCII >> (8 * ((byte >> (6 - 2 * (p & 3u))) & 3u))) & 0xFFu;
}

}  // namespace espk
