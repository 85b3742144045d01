// esp_walk.cpp -- junction library upload and the placement-walk entry point
// (rescore_with / recount_mm_ea_in / exact_kmer_support) behind the C ABI.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/esp_align.h"
#include "esp_ctx.hpp"
#include "esp_kernels.hpp"

using namespace esph;

// The library as an open-addressing hash set of its distinct triples: cap
// (a power of two, at least four times the count) slots of (contig, start,
// end, used), probed linearly from espk::junction_hash.  Membership is what
// contains_any_strand asks of the sorted library, one probe or two per query.
namespace {

// intron_length_penalty (splice.rs:119-121) evaluated directly.
int32_t penalty_formula(uint32_t len) {
    return (int32_t)(8.0 * std::fabs(std::log10((double)std::max<uint32_t>(len, 1)) - 3.0));
}

// Threshold form of the penalty: it is non-increasing on [0, 1000] and
// non-decreasing from 1000 up, so each side is a count of thresholds passed.
struct PenaltyTable {
    uint32_t up[espk::PEN_UP] = {};
    uint32_t down[espk::PEN_DOWN] = {};
    int nup = 0, ndown = 0;
    PenaltyTable() {
        // up[v-1]: smallest len >= 1000 with penalty >= v
        for (int v = 1; v <= espk::PEN_UP && penalty_formula(0xFFFFFFFFu) >= v; v++) {
            uint64_t lo = 1000, hi = 0xFFFFFFFFull;
            while (lo < hi) {
                const uint64_t mid = (lo + hi) / 2;
                if (penalty_formula((uint32_t)mid) >= v) hi = mid;
                else lo = mid + 1;
            }
            up[nup++] = (uint32_t)lo;
        }
        // down[v-1]: largest len <= 1000 with penalty >= v
        for (int v = 1; v <= espk::PEN_DOWN && penalty_formula(0) >= v; v++) {
            int64_t lo = 0, hi = 1000;
            while (lo < hi) {
                const int64_t mid = (lo + hi + 1) / 2;
                if (penalty_formula((uint32_t)mid) >= v) lo = mid;
                else hi = mid - 1;
            }
            down[ndown++] = (uint32_t)lo;
        }
    }
    int32_t eval(uint32_t len) const {
        int32_t v = 0;
        if (len >= 1000) {
            for (int t = 0; t < nup; t++) v += up[t] <= len;
        } else {
            for (int t = 0; t < ndown; t++) v += down[t] >= len;
        }
        return v;
    }
};

const PenaltyTable &penalty_table() {
    static const PenaltyTable t;
    return t;
}

}  // namespace

// Builds the library's hash set from n junctions of `width` u32 each:
// (contig, start, end) triples, or (contig, start, end, minus) quads whose
// per-strand counts (JunctionLib::build merges identical junctions into a
// count, gtf.rs:84-114) go into a parallel uint2 array.
static int junctions_upload(esp_ctx *c, size_t n, const uint32_t *t, int width, esp_junctions **out) {
    if (!c) return ESP_E_INVAL;
    if (!out || (n && !t)) return fail(c, ESP_E_INVAL, "NULL argument");
    *out = nullptr;
    if (n > (1ull << 30)) return fail(c, ESP_E_INVAL, "too many junctions");
    uint64_t cap = 16;
    while (cap < 4 * (uint64_t)n) cap <<= 1;
    std::vector<uint32_t> slots(4 * cap, 0);
    std::vector<uint32_t> counts(width == 4 ? 2 * cap : 0, 0);
    uint32_t distinct = 0;
    for (size_t k = 0; k < n; k++) {
        const uint32_t *x = t + (size_t)width * k;
        uint64_t s = espk::junction_hash(x[0], x[1], x[2]) & (cap - 1);
        while (slots[4 * s + 3]) {
            if (slots[4 * s] == x[0] && slots[4 * s + 1] == x[1] && slots[4 * s + 2] == x[2]) break;
            s = (s + 1) & (cap - 1);
        }
        if (!slots[4 * s + 3]) {
            slots[4 * s] = x[0];
            slots[4 * s + 1] = x[1];
            slots[4 * s + 2] = x[2];
            slots[4 * s + 3] = 1;
            distinct++;
        }
        if (width == 4) counts[2 * s + (x[3] ? 1 : 0)]++;
    }
    HIP_TRY(c, hipSetDevice(c->device));
    esp_junctions *j = new esp_junctions;
    j->device = c->device;
    j->n = distinct;
    j->cap = (uint32_t)cap;
    if (distinct) {
        const size_t bytes = slots.size() * sizeof(uint32_t);
        const size_t cbytes = counts.size() * sizeof(uint32_t);
        hipError_t e = hipMalloc(&j->d, bytes);
        if (e == hipSuccess && cbytes) e = hipMalloc(&j->counts, cbytes);
        if (e != hipSuccess) {
            esp_junctions_free(c, j);
            return fail(c, ESP_E_NOMEM, "junction library allocation failed");
        }
        e = hipMemcpy(j->d, slots.data(), bytes, hipMemcpyHostToDevice);
        if (e == hipSuccess && cbytes) e = hipMemcpy(j->counts, counts.data(), cbytes, hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            esp_junctions_free(c, j);
            return hip_fail(c, e, "junction library upload");
        }
    }
    *out = j;
    return ESP_OK;
}

extern "C" {

int32_t esp_intron_length_penalty(uint32_t len) { return penalty_table().eval(len); }

int esp_junctions_upload(esp_ctx *c, size_t n, const uint32_t *t, esp_junctions **out) {
    return junctions_upload(c, n, t, 3, out);
}

int esp_junctions_upload_stranded(esp_ctx *c, size_t n, const uint32_t *quads, esp_junctions **out) {
    return junctions_upload(c, n, quads, 4, out);
}

int esp_junctions_free(esp_ctx *c, esp_junctions *j) {
    if (!j) return ESP_OK;
    if (c) (void)hipSetDevice(j->device);
    if (j->d) (void)hipFree(j->d);
    if (j->counts) (void)hipFree(j->counts);
    delete j;
    return ESP_OK;
}

int esp_walk_placements(esp_ctx *c, const esp_ref *r, const esp_junctions *jl, const esp_walk_batch *b,
                        const esp_params *p, const esp_walk_params *wp, esp_strict_score *strict,
                        esp_walk_counts *counts, int32_t *status) {
    if (!c) return ESP_E_INVAL;
    int rc = check_params(c, p);
    if (rc) return rc;
    if (!r || !b || !wp) return fail(c, ESP_E_INVAL, "NULL argument");
    if (r->device != c->device) return fail(c, ESP_E_INVAL, "reference lives on another device");
    if (jl && jl->device != c->device) return fail(c, ESP_E_INVAL, "junctions live on another device");
    if (b->nplace == 0) return ESP_OK;
    if (b->nplace > 0x7FFFFFFFull) return fail(c, ESP_E_INVAL, "too many placements");
    if (!b->reads || !b->read_off || !b->read_len || !b->contig || !b->pos || !b->cigar || !b->cig_off ||
        !b->cig_n)
        return fail(c, ESP_E_INVAL, "batch arrays are required");
    HIP_TRY(c, hipSetDevice(c->device));
    espk::WalkArgs a{};
    a.R = r->view();
    a.jn = jl && jl->n ? (const uint4 *)jl->d : nullptr;
    a.jmask = jl && jl->n ? jl->cap - 1 : 0;
    a.n = (uint32_t)b->nplace;
    a.reads = b->reads;
    a.read_off = b->read_off;
    a.read_len = b->read_len;
    a.contig = b->contig;
    a.pos = b->pos;
    a.cigar = b->cigar;
    a.cig_off = b->cig_off;
    a.cig_n = b->cig_n;
    a.dir = b->dir;
    a.match = p->match_score;
    a.mismatch = p->mismatch;
    a.gap_open = p->gap_open;
    a.gap_ext = p->gap_ext;
    a.known = wp->known_bonus;
    a.gtag = wp->gt_ag_bonus;
    a.gcag = wp->gc_ag_bonus;
    a.atac = wp->at_ac_bonus;
    a.noncanon = wp->noncanonical_penalty;
    a.k = wp->seed_k;
    const PenaltyTable &pt = penalty_table();
    std::memcpy(a.up, pt.up, sizeof a.up);
    std::memcpy(a.down, pt.down, sizeof a.down);
    a.nup = pt.nup;
    a.ndown = pt.ndown;
    a.strict = strict;
    a.counts = counts;
    a.status = status;
    // one resident generation of blocks: each lane then walks its share of
    // the placements in rounds, with no partial last generation
    const uint64_t blocks =
        std::min<uint64_t>((b->nplace + 255) / 256, (uint64_t)c->ncu * (uint64_t)espk::walk_blocks_per_cu());
    HIP_TRY(c, ev_record(c, c->ev0));
    tim_begin(c);
    HIP_TRY(c, espk::launch_walk(a, (int)std::max<uint64_t>(blocks, 1), c->stream));
    tim_mark(c, 8, 0);
    HIP_TRY(c, ev_record(c, c->ev1));
    return ESP_OK;
}

int esp_fast_lane_batch(esp_ctx *c, const esp_ref *r, const esp_fast_batch *b, const esp_params *p, uint32_t k,
                        uint32_t w, esp_fast_rec *out, uint32_t *cigar) {
    if (!c) return ESP_E_INVAL;
    int rc = check_params(c, p);
    if (rc) return rc;
    if (!r || !b) return fail(c, ESP_E_INVAL, "NULL argument");
    if (r->device != c->device) return fail(c, ESP_E_INVAL, "reference lives on another device");
    if (b->nchain == 0) return ESP_OK;
    if (b->nchain > 0x7FFFFFFFull) return fail(c, ESP_E_INVAL, "too many chains");
    if (!b->reads || !b->read_off || !b->read_len || !b->contig || !b->anchors || !b->anc_off || !b->anc_n ||
        !out || !cigar)
        return fail(c, ESP_E_INVAL, "batch arrays are required");
    HIP_TRY(c, hipSetDevice(c->device));
    espk::FastArgs a{};
    a.R = r->view();
    a.n = (uint32_t)b->nchain;
    a.reads = b->reads;
    a.read_off = b->read_off;
    a.read_len = b->read_len;
    a.contig = b->contig;
    a.anchors = b->anchors;
    a.anc_off = b->anc_off;
    a.anc_n = b->anc_n;
    a.k = k;
    a.w = w;
    a.match = p->match_score;
    a.mismatch = p->mismatch;
    a.edit_dir = p->edit_dir;
    a.out = out;
    a.cigar = cigar;
    const uint64_t blocks = std::min<uint64_t>((b->nchain + 255) / 256,
                                              (uint64_t)c->ncu * (uint64_t)espk::fast_lane_blocks_per_cu());
    HIP_TRY(c, ev_record(c, c->ev0));
    tim_begin(c);
    HIP_TRY(c, espk::launch_fast_lane(a, (int)std::max<uint64_t>(blocks, 1), c->stream));
    tim_mark(c, 9, 0);
    HIP_TRY(c, ev_record(c, c->ev1));
    return ESP_OK;
}

}  // extern "C"
