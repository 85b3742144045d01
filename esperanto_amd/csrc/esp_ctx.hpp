// esp_ctx.hpp -- host-side context shared by the C ABI translation units
// (esp_host.cpp: extension and Myers calls; esp_window.cpp: resident
// reference, window gather and the mate-rescue / spliced pipelines).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <utility>
#include <vector>

#include "../../include/esp_align.h"
#include "esp_kernels.hpp"

namespace esph {

struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    hipError_t ensure(size_t need) {
        if (need <= bytes) return hipSuccess;
        if (p) {
            hipError_t e = hipFree(p);
            p = nullptr;
            bytes = 0;
            if (e != hipSuccess) return e;
        }
        size_t grow = std::max(need, (size_t)4096);
        hipError_t e = hipMalloc(&p, grow);
        if (e == hipSuccess) bytes = grow;
        return e;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
};

struct HostBuf {
    void *p = nullptr;
    size_t bytes = 0;
    hipError_t ensure(size_t need) {
        if (need <= bytes) return hipSuccess;
        if (p) (void)hipHostFree(p);
        p = nullptr;
        bytes = 0;
        size_t grow = std::max(need, (size_t)4096);
        hipError_t e = hipHostMalloc(&p, grow, hipHostMallocDefault);
        if (e == hipSuccess) bytes = grow;
        return e;
    }
    void release() {
        if (p) (void)hipHostFree(p);
        p = nullptr;
        bytes = 0;
    }
};

inline constexpr size_t kTraceBudget = (size_t)8 << 30;  // bytes of trace scratch per launch
inline constexpr size_t kRowBudget = (size_t)2 << 30;
inline constexpr size_t kFullTraceBudget = (size_t)16 << 30;  // full-matrix kernel trace

}  // namespace esph

namespace esph {

// Host-side bounds of a batch recorded into a HIP graph (esp_graph.cpp): the
// longest read and window any replay may hold, and the strip kernel's trace
// words per pair.  While a context captures, the Gotoh core sizes every
// launch from these instead of reading the planner's summary back.
struct CaptureBounds {
    uint32_t maxm = 1, maxn = 0;
    uint64_t strip_maxw = 1;
};

// Device scratch of the batch calls, as one set: a captured graph owns one
// and swaps it into its context while it records.
struct ScratchSet {
    DevBuf counter, rows, trace, btrace, ftrace, plan, order, escbuf, censusbuf;
    DevBuf wquery, warena, wdesc, wmyers, eng, jscratch, jarena, mlist, scantmp;
    void release() {
        for (DevBuf *b : {&counter, &rows, &trace, &btrace, &ftrace, &plan, &order, &escbuf, &censusbuf, &wquery,
                          &warena, &wdesc, &wmyers, &eng, &jscratch, &jarena, &mlist, &scantmp})
            b->release();
    }
};

}  // namespace esph

struct esp_ctx {
    int device = 0;
    int ncu = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::string err;
    esph::DevBuf counter, rows, trace, btrace, ftrace, stage, stage_out, plan, order;
    int full_blocks[512] = {};  // occupancy of the full-matrix kernel per row tile
    // Trace scratch caps per launch (grids shrink to fit); a server lowers them
    // so that several contexts share one device-memory budget.
    size_t trace_budget = esph::kTraceBudget;
    size_t full_trace_budget = esph::kFullTraceBudget;
    int full_kernel = 0;  // extend_hint reruns: 0 by batch size, 1 lane per pair
    int full_prune = 1;   // extend_hint reruns: pruned passes first (1; 2 without the register-window pass) or the full matrix only (0)
    int plain_kernel = 0;  // single-pair calls: 0 one wave per pair (windows up to 4096), 1 one lane
    bool last_pruned = false;  // the last batch ran the pruned reruns (their counts are in escbuf)
    bool last_wave_direct = false;  // ... with every escalation sent to the one-wave rerun first
    // full-matrix cell census (ESP_OPT_FULL_CENSUS): on/off, its device
    // counters, and whether the last batch ran the full-matrix stage with it
    int census = 0;
    esph::DevBuf censusbuf;
    bool last_census = false;
    // stripe half-width and head pad of the pruned pass
    int prune_half_a = espk::ESC_HALF_A, prune_pad_a = espk::ESC_PAD_A;
    // pruned reruns: EscDev, per-pair band results, the sorted and fallback lists
    esph::DevBuf escbuf;
    // window pipelines: query arena, window / pseudo-reference arena,
    // per-pair descriptors and Myers results
    esph::DevBuf wquery, warena, wdesc, wmyers;
    // transcriptome-engine verification windows (text offsets and lengths)
    esph::DevBuf eng;
    // chain preparation and re-extension: per-item state, offsets and arenas
    esph::DevBuf jscratch, jarena;
    // EA-Myers batch: pairs longer than the caller's bound, for the 8-word launch
    esph::DevBuf mlist;
    // tile sums of the prefix scans that lay out the junction-side arenas
    esph::DevBuf scantmp;
    hipEvent_t ev_plan = nullptr;
    // A second stream (default priority) for the whole-tile pruned pass,
    // which runs beside the register-window pass; fork and join events.
    hipStream_t stream2 = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    int band_blocks[espk::PLAN_STRIP] = {};
    esph::HostBuf hstage, hout, hplan;
    // Set while a graph records a call on this context (see CaptureBounds).
    const esph::CaptureBounds *cap = nullptr;
    // Per-launch timing of the last batch call.
    static constexpr int kMaxEv = 16;
    hipEvent_t evs[kMaxEv + 1] = {};
    esp_timing tim[kMaxEv] = {};
    int nev = 0;
};

// The resident reference (esp_ref_upload): device tables plus a host copy of
// the contig lengths for argument checks.
struct esp_ref {
    int device = 0;
    uint32_t ncontig = 0;
    std::vector<uint32_t> len;  // host copy for argument checks
    void *packed = nullptr, *pk_off = nullptr, *dlen = nullptr, *iv = nullptr, *iv_off = nullptr;
    size_t bytes = 0;
    espk::RefView view() const {
        espk::RefView v;
        v.packed = (const uint8_t *)packed;
        v.pk_off = (const uint64_t *)pk_off;
        v.len = (const uint32_t *)dlen;
        v.iv = (const uint32_t *)iv;
        v.iv_off = (const uint64_t *)iv_off;
        v.ncontig = ncontig;
        return v;
    }
};

// The junction library (esp_junctions_upload*): an open-addressing hash set
// of (contig, start, end) slots, at most a quarter full; a stranded upload
// also keeps, per slot, the support count of the plus and of the minus
// junction.
struct esp_junctions {
    int device = 0;
    uint32_t n = 0;    // distinct triples
    uint32_t cap = 0;  // slots
    void *d = nullptr;       // uint4 per slot: contig, start, end, used
    void *counts = nullptr;  // uint2 per slot (stranded uploads only)
};

namespace esph {

inline int fail(esp_ctx *c, int code, const std::string &msg) {
    if (c) c->err = msg;
    return code;
}

inline int hip_fail(esp_ctx *c, hipError_t e, const char *where) {
    return fail(c, ESP_E_DEVICE, std::string(where) + ": " + hipGetErrorString(e));
}

#define HIP_TRY(ctx, expr)                                   \
    do {                                                     \
        hipError_t _e = (expr);                              \
        if (_e != hipSuccess) return hip_fail(ctx, _e, #expr); \
    } while (0)

// Scoring envelope (include/esp_align.h).  Inside it the int32 device
// datapath cannot overflow, and the only place where the reference's i16
// accumulator leaves the i16 range is a diagonal candidate above 32767 (the
// strip kernel drops it for pure pairs; no other kernel admits such pairs).
inline int check_params(esp_ctx *c, const esp_params *p) {
    if (!p) return fail(c, ESP_E_INVAL, "params is NULL");
    if (p->gap_open < 0 || p->gap_ext < 0 || p->gap_open + p->gap_ext > 16384)
        return fail(c, ESP_E_INVAL, "gap penalties outside [0, 16384]");
    if (p->match_score < -1024 || p->match_score > 1024)
        return fail(c, ESP_E_INVAL, "match_score outside [-1024, 1024]");
    if (p->mismatch < -16384 || p->mismatch > 16384)
        return fail(c, ESP_E_INVAL, "mismatch outside [-16384, 16384]");
    if (p->edit_dir < 0 || p->edit_dir > 2) return fail(c, ESP_E_INVAL, "edit_dir not in {0,1,2}");
    return ESP_OK;
}

inline espd::GParams dev_params(const esp_params *p) {
    espd::GParams g;
    g.match = p->match_score;
    g.mismatch = p->mismatch;
    g.oe = -(p->gap_open + p->gap_ext);
    g.ext = -p->gap_ext;
    g.edit = p->edit_dir;
    return g;
}

inline constexpr uint32_t kMaxRead = 16000;
inline constexpr uint64_t kMaxWin = (uint64_t)1 << 24;

inline int check_lengths(esp_ctx *c, uint64_t m, uint64_t n) {
    if (m > kMaxRead) return fail(c, ESP_E_INVAL, "read longer than 16000");
    if (n > kMaxWin) return fail(c, ESP_E_INVAL, "window longer than 1<<24");
    return ESP_OK;
}

inline size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

// espk::launch_scan2 on the context's scan scratch.
inline int scan2(esp_ctx *c, const uint32_t *cnt, const uint64_t *len, size_t n, uint64_t *cnt_off,
                 uint64_t *len_off) {
    HIP_TRY(c, c->scantmp.ensure(espk::scan2_tmp_bytes(n)));
    HIP_TRY(c, espk::launch_scan2(cnt, len, n, cnt_off, len_off, (uint64_t *)c->scantmp.p, c->stream));
    return ESP_OK;
}

// Window stride of the mate-rescue pipeline for reads of at most maxm
// bases: its windows span 2 * 750 + m bases at most.
inline uint64_t mate_window_stride(uint64_t maxm) { return align_up(1500 + 3 * maxm, 16); }

// The call's start / end events (esp_last_kernel_ms); not recorded while a
// graph captures.
inline hipError_t ev_record(esp_ctx *c, hipEvent_t e) { return c->cap ? hipSuccess : hipEventRecord(e, c->stream); }

// Swaps the context's batch scratch with a graph's set.
inline void swap_scratch(esp_ctx *c, ScratchSet &s) {
    std::swap(c->counter, s.counter);
    std::swap(c->rows, s.rows);
    std::swap(c->trace, s.trace);
    std::swap(c->btrace, s.btrace);
    std::swap(c->ftrace, s.ftrace);
    std::swap(c->plan, s.plan);
    std::swap(c->order, s.order);
    std::swap(c->escbuf, s.escbuf);
    std::swap(c->censusbuf, s.censusbuf);
    std::swap(c->wquery, s.wquery);
    std::swap(c->warena, s.warena);
    std::swap(c->wdesc, s.wdesc);
    std::swap(c->wmyers, s.wmyers);
    std::swap(c->eng, s.eng);
    std::swap(c->jscratch, s.jscratch);
    std::swap(c->jarena, s.jarena);
    std::swap(c->mlist, s.mlist);
    std::swap(c->scantmp, s.scantmp);
}

// Launch timing: mark() before the first launch and after every launch of
// a batch call; esp_last_timings() reports the differences.
// Under capture nothing is timed: a graph launch is timed as one unit.
inline void tim_begin(esp_ctx *c) {
    if (c->cap) return;
    c->nev = 0;
    (void)hipEventRecord(c->evs[0], c->stream);
}
inline void tim_mark(esp_ctx *c, int kind, int half) {
    if (c->cap || c->nev >= esp_ctx::kMaxEv) return;
    c->tim[c->nev].kind = kind;
    c->tim[c->nev].half = half;
    c->nev++;
    (void)hipEventRecord(c->evs[c->nev], c->stream);
}

// Substitution rows of the band kernel: for read code rc = 0..4, the
// 5-bit signed LUT score of window code f = 0..5 at bit 5*f (window code 4
// never occurs; 5 is any non-ACGT byte).  Mirrors SubstLut (extend.rs:161-186).
// The band kernel's scores are 5-bit LUT fields, and its wide-band
// specialisations keep U = H + open_ext and E + ext as int16: with reads of
// at most band_max_read(half) <= 287 bases and |score| <= 16, every such value lies in
// (-20000, 16384) when gap_open + gap_ext <= 8000.
inline bool band_lut_fits(const esp_params *p) {
    return p->match_score >= -16 && p->match_score <= 15 && p->mismatch >= -16 &&
           p->mismatch <= 15 && p->gap_open + p->gap_ext <= 8000;
}

// Columns of the full-matrix kernel: for window code f = 0..5 the 5-bit
// score of read code rc = 0..4 at bit 5*rc, and -16 at bit 25 for the rows
// past the read (the transpose of band_lut).
inline void full_lut(const esp_params *p, uint32_t out[6]) {
    for (int f = 0; f < 6; f++) {
        uint32_t w = (uint32_t)(-16 & 31) << 25;
        for (int rc = 0; rc < 5; rc++) {
            const int partner = (p->edit_dir == 1 && rc == 2) ? 0 : (p->edit_dir == 2 && rc == 1) ? 3 : 6;
            const int s = (rc < 4 && f == rc) ? p->match_score : (f == partner ? 0 : p->mismatch);
            w |= ((uint32_t)s & 31u) << (5 * rc);
        }
        out[f] = w;
    }
}

// full_lut with -open_ext added to every field (the rows past the read keep
// -16 + gap_open + gap_ext); false when a field leaves the 5-bit range.
inline bool full_lut_folded(const esp_params *p, uint32_t out[6]) {
    const int noe = p->gap_open + p->gap_ext;
    bool fits = -16 + noe <= 15;
    for (int f = 0; f < 6; f++) {
        uint32_t w = (uint32_t)((-16 + noe) & 31) << 25;
        for (int rc = 0; rc < 5; rc++) {
            const int partner = (p->edit_dir == 1 && rc == 2) ? 0 : (p->edit_dir == 2 && rc == 1) ? 3 : 6;
            const int s = ((rc < 4 && f == rc) ? p->match_score : (f == partner ? 0 : p->mismatch)) + noe;
            fits = fits && s >= -16 && s <= 15;
            w |= ((uint32_t)s & 31u) << (5 * rc);
        }
        out[f] = w;
    }
    return fits;
}

// The largest score of the substitution table (match, the EditDir-exempt 0,
// or mismatch): no path gains more on one read row.
inline int max_subst(const esp_params *p) {
    int s = p->match_score > p->mismatch ? p->match_score : p->mismatch;
    if (p->edit_dir == 1 || p->edit_dir == 2) s = s > 0 ? s : 0;
    return s;
}

inline void band_lut(const esp_params *p, uint32_t out[5]) {
    for (int rc = 0; rc < 5; rc++) {
        const int partner = (p->edit_dir == 1 && rc == 2) ? 0 : (p->edit_dir == 2 && rc == 1) ? 3 : 6;
        uint32_t w = 0;
        for (int f = 0; f < 6; f++) {
            int s = (rc < 4 && f == rc) ? p->match_score : (f == partner ? 0 : p->mismatch);
            w |= ((uint32_t)s & 31u) << (5 * f);
        }
        out[rc] = w;
    }
}

// band_lut with -open_ext added to every field (score + gap_open + gap_ext);
// false when a field leaves the 5-bit signed range.
inline bool band_lut_folded(const esp_params *p, uint32_t out[5]) {
    const int noe = p->gap_open + p->gap_ext;
    bool fits = true;
    for (int rc = 0; rc < 5; rc++) {
        const int partner = (p->edit_dir == 1 && rc == 2) ? 0 : (p->edit_dir == 2 && rc == 1) ? 3 : 6;
        uint32_t w = 0;
        for (int f = 0; f < 6; f++) {
            const int s = ((rc < 4 && f == rc) ? p->match_score : (f == partner ? 0 : p->mismatch)) + noe;
            fits = fits && s >= -16 && s <= 15;
            w |= ((uint32_t)s & 31u) << (5 * f);
        }
        out[rc] = w;
    }
    return fits;
}

// Launch geometry of the persistent strip grid for `count` pairs whose
// largest has maxm rows and maxw trace words.
inline int strip_launch(esp_ctx *c, espk::GotohArgs &a, uint64_t count, uint64_t maxm, uint64_t maxw,
                        size_t trace_budget) {
    const int per_cu = espk::gotoh_strip_blocks_per_cu(a.P.edit);
    uint64_t lanes = (uint64_t)c->ncu * per_cu * espk::GOTOH_BLOCK;
    lanes = std::min<uint64_t>(lanes, align_up(count, 64));
    const uint64_t tw = std::max<uint64_t>(maxw, 1), rw = 5 * std::max<uint64_t>(maxm, 1);
    lanes = std::min<uint64_t>(lanes, trace_budget / (tw * 4));
    lanes = std::min<uint64_t>(lanes, kRowBudget / (rw * 4));
    lanes = std::max<uint64_t>(lanes / 64 * 64, 64);
    a.nl = (uint32_t)lanes;
    a.maxm = (uint32_t)std::max<uint64_t>(maxm, 1);
    a.maxw = (uint32_t)tw;
    HIP_TRY(c, c->rows.ensure(rw * lanes * 4));
    HIP_TRY(c, c->trace.ensure(tw * lanes * 4));
    a.rows = (int32_t *)c->rows.p;
    a.trace = (uint32_t *)c->trace.p;
    HIP_TRY(c, espk::launch_gotoh_strip(a, (int)(lanes / espk::GOTOH_BLOCK), c->stream));
    return ESP_OK;
}


// Window EA-Myers batch: argument checks, scratch sizes and the launches on
// caller-provided scratch (esp_window.cpp).
int window_myers_check(esp_ctx *c, const esp_ref *r, const esp_window_job *jobs, size_t njobs,
                       const uint8_t *reads, uint32_t max_text, int mode, int32_t *dist, uint64_t *pos);
size_t window_myers_desc_bytes(size_t njobs);
size_t window_myers_arena_bytes(size_t njobs, uint32_t max_text);
int window_myers_core(esp_ctx *c, const esp_ref *r, const esp_window_job *jobs, size_t njobs,
                      const uint8_t *reads, uint32_t max_pattern, uint32_t max_text, int mode, int32_t *dist,
                      uint64_t *pos, int32_t *status, uint8_t *desc, uint8_t *tarena, uint32_t *counter);

// The spliced window pipeline (esp_window.cpp) on caller-checked arguments.
int spliced_core(esp_ctx *c, const esp_ref *r, const esp_splice_batch *b, const esp_params *p,
                 esp_batch_out *o, uint32_t *stitched, const uint64_t *stitch_off, esp_splice_rec *rec);

// Set-up of the allocation-free entry points' device scratch on the current
// device (esp_noalloc.hip); called by esp_create.
void noalloc_setup(int device);

// Gotoh batch on the context stream without resetting the launch timings
// (the pipelines of esp_window.cpp time their own stages around it).
int gotoh_batch_core(esp_ctx *c, const esp_batch *b, const esp_params *p, int mode,
                     esp_batch_out *o);

}  // namespace esph
