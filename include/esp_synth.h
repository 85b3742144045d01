/*
 * esp_synth.h -- seeded synthetic read/window pairs for benchmarks and tests.
 *
 * Not part of the drop-in boundary (the reference has no counterpart; its
 * tests build inputs in-memory).  Generator: splitmix64.  Each pair: a window
 * of uniform ACGT of length L + 2*b, a read of length L sampled from
 * window[b, b+L) with substitution / indel events, re-trimmed or extended
 * from the right flank to exactly L, band offset = half = b.
 */
#ifndef ESP_SYNTH_H
#define ESP_SYNTH_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    uint64_t seed;
    uint32_t npairs;
    uint32_t len_min, len_max;  /* read length ~ U{len_min..len_max} */
    uint32_t nbands;            /* band choices, uniform */
    uint32_t bands[8];
    double sub_rate;            /* per read base */
    double ag_weight;           /* weight of A->G among the 12 substitution types (others 1) */
    double ag_fraction;         /* if > 0: this fraction of events is A->G at a random A */
    double indel_rate;          /* per read base */
    uint32_t indel_max;         /* indel length ~ U{1..indel_max} */
    uint32_t edits_max;         /* if > 0: e ~ U{0..edits_max} events, type uniform sub/ins/del, length 1 */
} esp_synth_spec;

/* Sizes needed for the spec (total read bytes, total window bytes). */
void esp_synth_sizes(const esp_synth_spec *s, uint64_t *read_bytes, uint64_t *win_bytes);

/* Fills host arrays: reads/wins (concatenated ASCII), *_off (npairs), *_len
 * (npairs), offset/half (npairs).  Returns 0. */
int esp_synth_pairs(const esp_synth_spec *s, uint8_t *reads, uint64_t *read_off,
                    uint32_t *read_len, uint8_t *wins, uint64_t *win_off, uint32_t *win_len,
                    int64_t *offset, int64_t *half);

#ifdef __cplusplus
}
#endif

#endif
