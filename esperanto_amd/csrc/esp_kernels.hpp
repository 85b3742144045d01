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
};

struct MyersArgs {
    uint32_t npairs;
    const uint32_t *order;
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

// Strip width of the general kernel (columns per register tile).
constexpr int STRIP_T = 32;
constexpr int GOTOH_BLOCK = 64;
constexpr int MYERS_BLOCK = 64;

// Trace words one pair needs in the strip layout.
inline uint64_t strip_trace_words(uint64_t m, uint64_t n) {
    return ((n + STRIP_T - 1) / STRIP_T) * m * (STRIP_T / 8);
}

hipError_t launch_gotoh_strip(const GotohArgs &a, int blocks, hipStream_t s);
hipError_t launch_myers(const MyersArgs &a, int words, int blocks, hipStream_t s);
hipError_t launch_zero_u32(uint32_t *p, size_t n, hipStream_t s);
int gotoh_strip_blocks_per_cu(int edit);
int myers_blocks_per_cu(int words);

}  // namespace espk
