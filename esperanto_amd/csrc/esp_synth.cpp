// esp_synth.cpp -- seeded synthetic pairs (include/esp_synth.h).
#include <stdint.h>
#include <string.h>

#include <thread>
#include <vector>

#include "../../include/esp_synth.h"

namespace {

struct SplitMix {
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    uint32_t below(uint32_t n) { return (uint32_t)(next() % n); }
    double unit() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

const char kACGT[4] = {'A', 'C', 'G', 'T'};

int code_of(uint8_t b) { return b == 'A' ? 0 : b == 'C' ? 1 : b == 'G' ? 2 : 3; }

// Pair-local RNG stream: seed mixed with the pair index, so any pair can be
// regenerated alone and the generator parallelises trivially.
SplitMix pair_rng(uint64_t seed, uint64_t k) {
    SplitMix r{seed ^ (0xD1B54A32D192ED03ull * (k + 1))};
    r.next();
    return r;
}

void substitute(SplitMix &r, const esp_synth_spec *s, std::vector<uint8_t> &rd) {
    if (rd.empty()) return;
    // Event type: A->G with weight ag_weight, the 11 other (src,dst) pairs 1.
    if (s->ag_fraction > 0.0) {
        if (r.unit() < s->ag_fraction) {
            for (int t = 0; t < 16; t++) {
                uint32_t p = r.below((uint32_t)rd.size());
                if (rd[p] == 'A') {
                    rd[p] = 'G';
                    return;
                }
            }
            return;
        }
        uint32_t p = r.below((uint32_t)rd.size());
        int c = code_of(rd[p]);
        rd[p] = (uint8_t)kACGT[(c + 1 + r.below(3)) & 3];
        return;
    }
    const double total = s->ag_weight + 11.0;
    double u = r.unit() * total;
    int src, dst;
    if (u < s->ag_weight) {
        src = 0;
        dst = 2;
    } else {
        int k = (int)(u - s->ag_weight);
        if (k > 10) k = 10;
        // enumerate the 11 remaining (src,dst) pairs in a fixed order
        int idx = 0;
        src = dst = -1;
        for (int a = 0; a < 4 && src < 0; a++)
            for (int b = 0; b < 4; b++) {
                if (a == b || (a == 0 && b == 2)) continue;
                if (idx++ == k) {
                    src = a;
                    dst = b;
                    break;
                }
            }
    }
    for (int t = 0; t < 16; t++) {
        uint32_t p = r.below((uint32_t)rd.size());
        if (rd[p] == (uint8_t)kACGT[src]) {
            rd[p] = (uint8_t)kACGT[dst];
            return;
        }
    }
}

void make_pair(const esp_synth_spec *s, uint64_t k, uint32_t L, uint32_t b, uint8_t *read,
               uint8_t *win) {
    SplitMix r = pair_rng(s->seed, k);
    const uint32_t n = L + 2 * b;
    for (uint32_t j = 0; j < n; j++) win[j] = (uint8_t)kACGT[r.below(4)];
    std::vector<uint8_t> rd;
    rd.reserve(L + 64);
    if (s->edits_max > 0) {
        // e ~ U{0..edits_max} single-base events on the locus, type uniform.
        rd.assign(win + b, win + b + L);
        uint32_t e = r.below(s->edits_max + 1);
        for (uint32_t t = 0; t < e; t++) {
            uint32_t kind = r.below(3);
            uint32_t p = r.below((uint32_t)rd.size());
            if (kind == 0) {
                int c = code_of(rd[p]);
                rd[p] = (uint8_t)kACGT[(c + 1 + r.below(3)) & 3];
            } else if (kind == 1) {
                rd.insert(rd.begin() + p, (uint8_t)kACGT[r.below(4)]);
            } else if (rd.size() > 1) {
                rd.erase(rd.begin() + p);
            }
        }
    } else {
        uint32_t j = b;
        while (rd.size() < L + 16 && j < n) {
            if (s->indel_rate > 0.0 && r.unit() < s->indel_rate) {
                uint32_t len = 1 + r.below(s->indel_max ? s->indel_max : 1);
                if (r.below(2) == 0) {
                    for (uint32_t t = 0; t < len; t++) rd.push_back((uint8_t)kACGT[r.below(4)]);
                } else {
                    j += len;
                }
                continue;
            }
            rd.push_back(win[j++]);
        }
        const uint32_t nsub = (uint32_t)(s->sub_rate * (double)L + r.unit());
        for (uint32_t t = 0; t < nsub; t++) substitute(r, s, rd);
    }
    // Re-trim or extend from the right flank to exactly L bases.
    if (rd.size() > L) rd.resize(L);
    uint32_t j = b + L;
    while (rd.size() < L) rd.push_back(j < n ? win[j++] : (uint8_t)kACGT[r.below(4)]);
    memcpy(read, rd.data(), L);
}

uint32_t pair_len(const esp_synth_spec *s, uint64_t k) {
    SplitMix r{s->seed * 0x2545F4914F6CDD1Dull + k};
    r.next();
    return s->len_min + (s->len_max > s->len_min ? r.below(s->len_max - s->len_min + 1) : 0);
}

uint32_t pair_band(const esp_synth_spec *s, uint64_t k) {
    if (s->nbands <= 1) return s->nbands ? s->bands[0] : 0;
    SplitMix r{s->seed * 0x9E3779B97F4A7C15ull + 7 * k + 3};
    r.next();
    return s->bands[r.below(s->nbands)];
}

}  // namespace

extern "C" {

void esp_synth_sizes(const esp_synth_spec *s, uint64_t *read_bytes, uint64_t *win_bytes) {
    uint64_t rb = 0, wb = 0;
    for (uint64_t k = 0; k < s->npairs; k++) {
        uint32_t L = pair_len(s, k), b = pair_band(s, k);
        rb += L;
        wb += L + 2ull * b;
    }
    *read_bytes = rb;
    *win_bytes = wb;
}

int esp_synth_pairs(const esp_synth_spec *s, uint8_t *reads, uint64_t *read_off,
                    uint32_t *read_len, uint8_t *wins, uint64_t *win_off, uint32_t *win_len,
                    int64_t *offset, int64_t *half) {
    uint64_t ro = 0, wo = 0;
    for (uint64_t k = 0; k < s->npairs; k++) {
        uint32_t L = pair_len(s, k), b = pair_band(s, k);
        read_off[k] = ro;
        win_off[k] = wo;
        read_len[k] = L;
        win_len[k] = L + 2 * b;
        offset[k] = b;
        half[k] = b;
        ro += L;
        wo += L + 2ull * b;
    }
    // Pairs are independent: fill them in parallel chunks.
    const uint64_t N = s->npairs;
    unsigned nt = std::thread::hardware_concurrency();
    if (nt == 0) nt = 1;
    if (nt > 16) nt = 16;
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; t++) {
        th.emplace_back([=]() {
            for (uint64_t k = t; k < N; k += nt)
                make_pair(s, k, read_len[k], (uint32_t)offset[k], reads + read_off[k],
                          wins + win_off[k]);
        });
    }
    for (auto &x : th) x.join();
    return 0;
}

}  // extern "C"
