// esp_kernels.hpp -- launch interface between the host runtime (esp_host.cpp)
// and the device kernels (esp_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/esp_align.h"
#include "esp_device.hpp"

namespace espk {

// Per-launch description of a Gotoh batch.  Sequences are ASCII bytes.
struct GotohArgs {
    uint32_t npairs;
    const uint32_t *order;  // visiting order (length-bucketed), or nullptr for identity
    const uint32_t *range;  // device {start, count} into order for this launch, or nullptr
    const uint8_t *reads;
    const uint64_t *read_off;
    const uint32_t *read_len;
    const uint8_t *wins;
    const uint64_t *win_off;
    const uint32_t *win_len;
    const int64_t *offset;  // band offset or extend_hint hint offset
    const int64_t *half;    // band half-width (band mode)
    int32_t mode;           // esp_mode
    espd::GParams P;
    esp_ext *ext;
    esp_touch *touch;
    uint32_t *cigar;
    const uint64_t *cig_off;
    uint32_t *cig_n;
    int32_t *aux;
    // scratch
    uint32_t *counter;  // work-queue head, zeroed by the launcher
    int32_t *rows;      // 5 arrays of maxm x NL int32, lane-interleaved
    uint32_t *trace;    // maxw x NL trace words, lane-interleaved
    uint32_t maxm;
    uint32_t maxw;
    uint32_t nl;        // lanes in the persistent grid
    // Band kernel substitution rows: for read code 0..4, 5-bit signed scores
    // of window codes 0..5 packed at bit 5*code.
    uint32_t lut[5];
};

struct MyersArgs {
    uint32_t npairs;
    const uint32_t *order;
    const uint32_t *range;  // device {start, count} into order, or nullptr
    const uint8_t *reads;
    const uint64_t *read_off;
    const uint32_t *read_len;
    const uint8_t *texts;
    const uint64_t *text_off;
    const uint32_t *text_len;
    int32_t *dist;
    uint64_t *end;
    uint64_t *start;  // may be nullptr
    uint32_t *counter;
};

// Device-side batch planner: pairs are classified (one class per band
// kernel specialisation, plus the general strip class) and counting-sorted
// by (class, read length descending) into `order`.
constexpr int PLAN_NCLASS = 6;  // halves 16, 30, 32, 48, 64, then the strip class
constexpr int PLAN_STRIP = 5;
constexpr int PLAN_NBIN = 512;
struct PlanSummary {
    uint32_t range[PLAN_NCLASS][2];  // {start, count} per class
    uint32_t strip_maxm;
    uint32_t bad;                   // pairs outside the length limits
    unsigned long long strip_maxw;  // largest strip-layout trace of the strip class (words)
};
struct PlanDev {
    PlanSummary sum;  // first: the host reads back only this part
    uint32_t counters[PLAN_NCLASS + 2];
    uint32_t hist[PLAN_NCLASS * PLAN_NBIN];
    uint32_t cursor[PLAN_NCLASS * PLAN_NBIN];
};
struct PlanArgs {
    uint32_t npairs;
    const uint32_t *read_len;
    const uint32_t *win_len;
    const int64_t *offset;
    const int64_t *half;
    int32_t band_ok;  // 1 when mode is band and the scoring fits the band kernel
    uint32_t max_read;
    uint32_t max_win;
    PlanDev *plan;
    uint32_t *order;
};
hipError_t launch_plan(const PlanArgs &a, int ncu, hipStream_t s);
int plan_class_half(int c);

// Strip width of the general kernel (columns per register tile).
constexpr int STRIP_T = 32;
constexpr int GOTOH_BLOCK = 64;
constexpr int MYERS_BLOCK = 64;

// Trace words one pair needs in the strip layout.
__host__ __device__ inline uint64_t strip_trace_words(uint64_t m, uint64_t n) {
    return ((n + STRIP_T - 1) / STRIP_T) * m * (STRIP_T / 8);
}

hipError_t launch_gotoh_strip(const GotohArgs &a, int blocks, hipStream_t s);

// Diagonal-frame band kernel for interior bands (every row window inside
// [1, n]); half must be one of band_halves().  LDS holds the window codes of
// the unit's 64 pairs: (max m + 2*half) * 64 bytes.
constexpr int BAND_BLOCK = 64;
// LDS rows of window codes per wave (64 bytes each): a pair fits when
// m + 2*half + 1 <= span, i.e. reads up to 255 bases at half 64.  The spans
// keep LDS at 20-24 KB per wave so that LDS does not cut occupancy below
// what the register arrays allow.
__host__ __device__ constexpr int band_span(int half) {
    return half <= 32 ? 320 : (half <= 48 ? 352 : 384);
}
bool band_half_supported(int half);
hipError_t launch_gotoh_band(const GotohArgs &a, int half, int blocks, hipStream_t s);
int gotoh_band_blocks_per_cu(int half);
__host__ __device__ inline uint64_t band_trace_words(uint64_t m, uint64_t half) { return m * ((2 * half + 1 + 7) / 8); }
hipError_t launch_myers(const MyersArgs &a, int words, int blocks, hipStream_t s);
hipError_t launch_zero_u32(uint32_t *p, size_t n, hipStream_t s);
hipError_t launch_copy(void *dst, const void *src, size_t bytes, hipStream_t s);
int gotoh_strip_blocks_per_cu(int edit);
int myers_blocks_per_cu(int words);

}  // namespace espk
