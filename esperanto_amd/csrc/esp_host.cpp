// esp_host.cpp -- host runtime behind the C ABI (include/esp_align.h).
//
// A context owns one HIP stream (default priority; a second one, forked and
// joined inside a call, carries one of the pruned passes), grow-only device scratch
// and pinned staging for the single-pair calls.  Batched calls are
// asynchronous on the context stream; single-pair calls stage the pair,
// launch one kernel and copy the result back (the plain form).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/esp_align.h"
#include "esp_ctx.hpp"
#include "esp_kernels.hpp"
#include "esp_stitch.hpp"

namespace {

using namespace esph;

// Single-pair staging layout (host pinned -> device), 16-byte aligned fields.
struct PairStage {
    size_t o_read, o_win, o_meta, total;
};

PairStage pair_layout(size_t m, size_t n) {
    PairStage s;
    s.o_read = 0;
    s.o_win = align_up(m, 16);
    s.o_meta = s.o_win + align_up(n, 16);
    // meta: read_off, win_off (u64), read_len, win_len (u32, padded), offset,
    // half (i64), cig_off[2] (u64)
    s.total = s.o_meta + 8 * 8;
    return s;
}

struct PairOutLayout {
    size_t o_ext, o_touch, o_n, o_aux, o_cig, total;
};

PairOutLayout out_layout(size_t cap) {
    PairOutLayout o;
    o.o_ext = 0;
    o.o_touch = 16;
    o.o_n = 48;
    o.o_aux = 52;
    o.o_cig = 64;
    o.total = o.o_cig + cap * 4;
    return o;
}

int run_single(esp_ctx *c, const uint8_t *read, size_t m, const uint8_t *win, size_t n,
               const esp_params *p, int mode, int64_t offset, int64_t half, esp_ext *out,
               esp_touch *touch, uint32_t *cigar, size_t cap, size_t *n_ops, int32_t *aux_out) {
    if (!c) return ESP_E_INVAL;
    int rc = check_params(c, p);
    if (rc) return rc;
    if ((m && !read) || (n && !win) || !out || !n_ops) return fail(c, ESP_E_INVAL, "NULL argument");
    rc = check_lengths(c, m, n);
    if (rc) return rc;
    HIP_TRY(c, hipSetDevice(c->device));
    const size_t slot_cap = esp_cigar_capacity(m);
    PairStage L = pair_layout(m, n);
    PairOutLayout O = out_layout(slot_cap);
    HIP_TRY(c, c->hstage.ensure(L.total));
    HIP_TRY(c, c->stage.ensure(L.total));
    HIP_TRY(c, c->hout.ensure(O.total));
    HIP_TRY(c, c->stage_out.ensure(O.total));
    uint8_t *h = (uint8_t *)c->hstage.p;
    if (m) std::memcpy(h + L.o_read, read, m);
    if (n) std::memcpy(h + L.o_win, win, n);
    uint64_t *meta = (uint64_t *)(h + L.o_meta);
    meta[0] = 0;  // read_off
    meta[1] = 0;  // win_off
    uint32_t *lens = (uint32_t *)(meta + 2);
    lens[0] = (uint32_t)m;
    lens[1] = (uint32_t)n;
    int64_t *oh = (int64_t *)(meta + 3);
    oh[0] = offset;
    oh[1] = half;
    uint64_t *co = meta + 5;
    co[0] = 0;
    co[1] = slot_cap;
    HIP_TRY(c, hipMemcpyAsync(c->stage.p, h, L.total, hipMemcpyHostToDevice, c->stream));
    uint8_t *d = (uint8_t *)c->stage.p;
    uint8_t *dout = (uint8_t *)c->stage_out.p;
    espk::GotohArgs a{};
    a.npairs = 1;
    a.order = nullptr;
    a.reads = d + L.o_read;
    a.wins = d + L.o_win;
    a.read_off = (const uint64_t *)(d + L.o_meta);
    a.win_off = (const uint64_t *)(d + L.o_meta) + 1;
    a.read_len = (const uint32_t *)(d + L.o_meta + 16);
    a.win_len = (const uint32_t *)(d + L.o_meta + 20);
    a.offset = (const int64_t *)(d + L.o_meta + 24);
    a.half = (const int64_t *)(d + L.o_meta + 32);
    a.cig_off = (const uint64_t *)(d + L.o_meta + 40);
    a.mode = mode;
    a.P = dev_params(p);
    a.ext = (esp_ext *)(dout + O.o_ext);
    a.touch = (esp_touch *)(dout + O.o_touch);
    a.cig_n = (uint32_t *)(dout + O.o_n);
    a.aux = (int32_t *)(dout + O.o_aux);
    a.cigar = (uint32_t *)(dout + O.o_cig);
    HIP_TRY(c, c->counter.ensure(256));
    a.counter = (uint32_t *)c->counter.p;
    HIP_TRY(c, espk::launch_zero_u32(a.counter, 1, c->stream));
    if (c->plain_kernel == 0 && n <= (size_t)espk::WAVE_NMAX) {
        // one wave sweeps the pair; its trace stays in LDS when small enough
        const uint64_t need = espk::wave_trace_words(m, n, mode, half);
        const uint64_t maxw = need > espk::WAVE_TR_LDS ? need : 1;
        HIP_TRY(c, c->trace.ensure(maxw * 4));
        a.trace = (uint32_t *)c->trace.p;
        a.maxw = (uint32_t)maxw;
        HIP_TRY(c, ev_record(c, c->ev0));
        HIP_TRY(c, espk::launch_gotoh_wave(a, 1, c->stream));
        HIP_TRY(c, ev_record(c, c->ev1));
    } else {
        const uint64_t maxw = espk::strip_trace_words(m, n);
        HIP_TRY(c, c->rows.ensure(5 * std::max<size_t>(m, 1) * 64 * 4));
        HIP_TRY(c, c->trace.ensure(std::max<uint64_t>(maxw, 1) * 64 * 4));
        a.rows = (int32_t *)c->rows.p;
        a.trace = (uint32_t *)c->trace.p;
        a.maxm = (uint32_t)std::max<size_t>(m, 1);
        a.maxw = (uint32_t)std::max<uint64_t>(maxw, 1);
        a.nl = 64;
        HIP_TRY(c, ev_record(c, c->ev0));
        HIP_TRY(c, espk::launch_gotoh_strip(a, 1, c->stream));
        HIP_TRY(c, ev_record(c, c->ev1));
    }
    HIP_TRY(c, hipMemcpyAsync(c->hout.p, c->stage_out.p, O.total, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    const uint8_t *ho = (const uint8_t *)c->hout.p;
    std::memcpy(out, ho + O.o_ext, sizeof(esp_ext));
    if (touch) std::memcpy(touch, ho + O.o_touch, sizeof(esp_touch));
    if (aux_out) std::memcpy(aux_out, ho + O.o_aux, sizeof(int32_t));
    uint32_t nops;
    std::memcpy(&nops, ho + O.o_n, 4);
    *n_ops = nops;
    if (nops > cap || (nops && !cigar)) return fail(c, ESP_E_CAP, "CIGAR capacity too small");
    if (nops) std::memcpy(cigar, ho + O.o_cig, nops * 4);
    return ESP_OK;
}

}  // namespace

// Planner, band kernels and strip kernel of one Gotoh batch (asynchronous).
int esph::gotoh_batch_core(esp_ctx *c, const esp_batch *b, const esp_params *p, int mode,
                           esp_batch_out *o) {
    int rc = check_params(c, p);
    if (rc) return rc;
    if (!b || !o) return fail(c, ESP_E_INVAL, "batch and output are required");
    if (mode < ESP_MODE_BAND || mode > ESP_MODE_HINT) return fail(c, ESP_E_INVAL, "bad mode");
    if (b->npairs == 0) return ESP_OK;
    if (b->npairs > 0xFFFFFF00ull) return fail(c, ESP_E_INVAL, "too many pairs");
    if ((mode != ESP_MODE_FULL && !b->offset) || (mode == ESP_MODE_BAND && !b->half))
        return fail(c, ESP_E_INVAL, "offset/half arrays required for this mode");
    HIP_TRY(c, hipSetDevice(c->device));
    espk::GotohArgs a{};
    a.npairs = (uint32_t)b->npairs;
    a.reads = b->reads;
    a.read_off = b->read_off;
    a.read_len = b->read_len;
    a.wins = b->wins;
    a.win_off = b->win_off;
    a.win_len = b->win_len;
    a.offset = b->offset;
    a.half = b->half;
    a.mode = mode;
    a.P = dev_params(p);
    band_lut(p, a.lut);
    a.lut_folded = band_lut_folded(p, a.lutf) ? 1 : 0;
    full_lut(p, a.flut);
    a.flut_folded = full_lut_folded(p, a.flutf) ? 1 : 0;
    a.ext = o->ext;
    a.touch = o->touch;
    a.cigar = o->cigar;
    a.cig_off = o->cig_off;
    a.cig_n = o->cig_n;
    a.aux = o->aux;

    // 1. Device planner: classify, bucket by (class, read length), order.
    HIP_TRY(c, c->plan.ensure(sizeof(espk::PlanDev)));
    const bool hint_band = mode == ESP_MODE_HINT && band_lut_fits(p);
    // order: the planner's visiting order, then (extend_hint) the escalation list
    HIP_TRY(c, c->order.ensure(b->npairs * 4 * (hint_band ? 2 : 1)));
    HIP_TRY(c, c->hplan.ensure(sizeof(espk::PlanSummary)));
    espk::PlanArgs pa{};
    pa.npairs = a.npairs;
    pa.read_len = b->read_len;
    pa.win_len = b->win_len;
    pa.offset = b->offset;
    pa.half = b->half;
    pa.band_ok = mode == ESP_MODE_BAND && band_lut_fits(p);
    pa.hint_mode = hint_band;
    pa.left_clamp_ok = p->mismatch <= 0;
    // Under capture, pairs past the graph's bounds are counted as bad and
    // left alone (esp_graph_check reports them).
    const esph::CaptureBounds *capb = c->cap;
    pa.max_read = capb ? capb->maxm : kMaxRead;
    pa.max_win = capb ? capb->maxn : (uint32_t)kMaxWin;
    pa.max_strip_w = capb ? capb->strip_maxw : ~0ull;
    pa.plan = (espk::PlanDev *)c->plan.p;
    pa.order = (uint32_t *)c->order.p;
    HIP_TRY(c, espk::launch_plan(pa, c->ncu, c->stream));
    tim_mark(c, 0, 0);
    // The plan summary travels back while the band kernels run: they read
    // their class ranges from device memory and need nothing from the host.
    // A graph sizes from its bounds instead and reads nothing back.
    if (!capb) {
        HIP_TRY(c, hipMemcpyAsync(c->hplan.p, c->plan.p, sizeof(espk::PlanSummary),
                                  hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(c, hipEventRecord(c->ev_plan, c->stream));
    }
    a.order = pa.order;
    espk::PlanDev *dp = pa.plan;
    a.esc_range = dp->esc_range;
    a.esc_list = pa.order + b->npairs;
    // Pruned reruns: EscDev, then per-pair band results (2 x int32), the sorted
    // list and the two fallback lists (npairs entries each).
    const int prune_s = max_subst(p);
    const bool prune = hint_band && c->full_prune && prune_s > 0;
    c->last_pruned = false;
    c->last_wave_direct = false;
    c->last_census = false;
    espk::EscDev *esc = nullptr;
    uint32_t *plist = nullptr, *fblist = nullptr, *fb2list = nullptr;
    if (prune) {
        const size_t head = align_up(sizeof(espk::EscDev), 256);
        HIP_TRY(c, c->escbuf.ensure(head + (size_t)b->npairs * 20));
        uint8_t *base = (uint8_t *)c->escbuf.p;
        esc = (espk::EscDev *)base;
        a.esc_info = (int32_t *)(base + head);
        plist = (uint32_t *)(base + head + (size_t)b->npairs * 8);
        fblist = plist + b->npairs;
        fb2list = fblist + b->npairs;
        HIP_TRY(c, espk::launch_zero_u32((uint32_t *)esc, sizeof(espk::EscDev) / 4, c->stream));
    }

    // 2. Band kernel per half specialisation (an empty class exits at once);
    //    extend_hint's band attempt uses only the half-30 class.
    if (pa.band_ok || pa.hint_mode) {
        for (int cls = 0; cls < espk::PLAN_STRIP; cls++) {
            const int half = espk::plan_class_half(cls);
            if (pa.hint_mode && half != 30) continue;
            const uint64_t maxw = espk::band_trace_words(espk::band_max_read(half), half);
            if (!c->band_blocks[cls]) c->band_blocks[cls] = espk::gotoh_band_blocks_per_cu(half);
            uint64_t blocks = (uint64_t)c->ncu * c->band_blocks[cls];
            blocks = std::min<uint64_t>(blocks, (b->npairs + 63) / 64);
            blocks = std::min<uint64_t>(blocks, c->trace_budget / (maxw * 64 * 4));
            blocks = std::max<uint64_t>(blocks, 1);
            HIP_TRY(c, c->btrace.ensure(maxw * 64 * 4 * blocks));
            a.trace = (uint32_t *)c->btrace.p;
            a.maxw = (uint32_t)maxw;
            a.range = &dp->sum.range[cls][0];
            a.counter = &dp->counters[cls];
            HIP_TRY(c, espk::launch_gotoh_band(a, half, (int)blocks, c->stream));
            tim_mark(c, 1, half);
        }
    }

    // 3. Strip kernel for everything else, sized from the plan summary (or,
    //    under capture, from the graph's bounds: its grid then drains
    //    whatever the class holds at each replay, possibly nothing).
    uint64_t strip_count, strip_maxm, strip_maxw, esc_bound = 0, esc_maxm = 0, esc_maxn = 0;
    if (capb) {
        strip_count = b->npairs;
        strip_maxm = capb->maxm;
        strip_maxw = capb->strip_maxw;
        if (hint_band) {
            esc_bound = b->npairs;
            esc_maxm = std::min<uint64_t>(capb->maxm, espk::band_max_read(30));
            esc_maxn = capb->maxn;
        }
    } else {
        HIP_TRY(c, hipEventSynchronize(c->ev_plan));
        const espk::PlanSummary *hp = (const espk::PlanSummary *)c->hplan.p;
        if (hp->bad) {
            // The band kernels already queued read this batch's arrays: let them
            // finish so that the caller may free them as soon as this returns.
            (void)hipStreamSynchronize(c->stream);
            return fail(c, ESP_E_INVAL, "read longer than 16000 or window longer than 1<<24");
        }
        strip_count = hp->range[espk::PLAN_STRIP][1];
        strip_maxm = hp->strip_maxm;
        strip_maxw = hp->strip_maxw;
        if (pa.hint_mode) {
            esc_bound = hp->range[1][1];
            esc_maxm = hp->esc_maxm;
            esc_maxn = hp->esc_maxn;
        }
    }
    // A strip class of at most ncu * 64 pairs with a read and a window, all
    // windows within the wave kernel's limit (the eager call knows both from
    // the plan summary), runs one wave per pair: a lane-per-pair launch
    // takes as long as one lane's whole DP, which a few heavy pairs set.
    const espk::PlanSummary *hs = capb ? nullptr : (const espk::PlanSummary *)c->hplan.p;
    const bool wave_strip = hs && strip_count && c->plain_kernel == 0 &&
                            hs->strip_work <= (uint64_t)c->ncu * 64 && hs->strip_maxn <= (uint32_t)espk::WAVE_NMAX;
    if (wave_strip) {
        // LDS for the class's longest window, the trace in global scratch
        const uint32_t ln = (uint32_t)align_up(std::max<uint32_t>(hs->strip_maxn, 1), 16);
        // every pair's wave trace fits its strip-layout trace (m * ceil(n/8)
        // words against m * 4 * ceil(n/32)), so nothing overflows; the lane
        // kernel below still drains whatever would (64 lanes, an empty list
        // exits at once)
        const uint64_t tw = std::max<uint64_t>(strip_maxw, 1);
        const uint64_t blocks_w = std::max<uint64_t>(
            1, std::min<uint64_t>({(uint64_t)c->ncu * espk::gotoh_wave_blocks_per_cu(ln, 0), strip_count,
                                   c->trace_budget / (tw * 4)}));
        HIP_TRY(c, c->trace.ensure(tw * 4 * std::max<uint64_t>(blocks_w, 64)));
        HIP_TRY(c, c->mlist.ensure(strip_count * 4 + 64));
        espk::GotohArgs w = a;
        w.range = &dp->sum.range[espk::PLAN_STRIP][0];
        w.trace = (uint32_t *)c->trace.p;
        w.maxw = (uint32_t)tw;
        w.list_cap = 0xFFFFFFFFu;
        w.wave_lds_n = ln;
        w.wave_lds_tr = 0;
        w.fb_list = (uint32_t *)c->mlist.p;
        w.fb_range = dp->wave_ovf;
        HIP_TRY(c, espk::launch_gotoh_wave_list(w, (int)blocks_w, c->stream));
        a.order = (const uint32_t *)c->mlist.p;
        a.range = dp->wave_ovf;
        a.counter = &dp->wave_ovf_counter;
        rc = strip_launch(c, a, 64, strip_maxm, strip_maxw, c->trace_budget);
        if (rc) return rc;
        a.order = pa.order;
        tim_mark(c, 2, 0);
    } else if (strip_count) {
        a.range = &dp->sum.range[espk::PLAN_STRIP][0];
        a.counter = &dp->counters[espk::PLAN_STRIP];
        // a graph's strip grid is sized for every pair but holds at most a
        // quarter of the trace budget
        rc = strip_launch(c, a, strip_count, strip_maxm, strip_maxw, capb ? c->trace_budget / 4 : c->trace_budget);
        if (rc) return rc;
        tim_mark(c, 2, 0);
    }
    // 4. extend_hint attempt 2: the full matrix for every band pair that
    //    escalated (its count is known only on the device; the grid drains
    //    the list and exits).  aux already holds 1 for these pairs.
    //    With pruning the list is first sorted by (head groups, band-maximum
    //    diagonal).  The pruned pass takes the sorted list 64 pairs a wave;
    //    its failures and the pairs it cannot help rerun one wave per pair
    //    over the whole matrix (up to a list cap); the unpruned lane kernel
    //    drains the rest (usually nothing: it then exits at once).
    if (pa.hint_mode && esc_bound) {
        espk::GotohArgs f = a;
        f.mode = ESP_MODE_FULL;
        f.aux = nullptr;
        f.touch = nullptr;  // extend_hint returns no TouchRec
        f.esc_info = nullptr;
        f.range = dp->esc_range;
        f.counter = &dp->counters[espk::PLAN_NCLASS];
        const uint64_t bound = esc_bound;
        const int rows = espk::full_rows((uint32_t)esc_maxm);
        const uint64_t maxw = std::max<uint64_t>(1, esc_maxn * (rows / 8));
        if (!c->full_blocks[rows & 511]) c->full_blocks[rows & 511] = espk::gotoh_full_blocks_per_cu(rows);
        const uint64_t resident = (uint64_t)c->ncu * c->full_blocks[rows & 511];
        const uint64_t cap = std::max<uint64_t>(1, c->full_trace_budget / (maxw * 64 * 4));
        // 64 pairs a wave
        const uint64_t blocks = std::max<uint64_t>(1, std::min({resident, (bound + 63) / 64, cap}));
        const bool pruned = prune && rows <= espk::FULL_PRUNE_MAX_ROWS;
        // the second pass: one wave per pair the pruned pass could not prove,
        // each block with the full-matrix trace of the longest pair
        const uint64_t wave_w = esc_maxm * ((esc_maxn + 7) / 8);
        const uint32_t wave_ln = (uint32_t)align_up(std::min<uint64_t>(std::max<uint64_t>(esc_maxn, 1),
                                                                       (uint64_t)espk::WAVE_NMAX), 16);
        const uint64_t blocks_w = std::max<uint64_t>(
            1, std::min<uint64_t>({(uint64_t)c->ncu * espk::gotoh_wave_blocks_per_cu(wave_ln, 0), bound,
                                   std::max<uint64_t>(1, c->full_trace_budget / (std::max<uint64_t>(wave_w, 1) * 4))}));
        // The register-window pass takes the sort's first head class (more
        // waves per SIMD than the whole-tile pass), the whole-tile pass the
        // other classes at the same time on the second stream: each pass's
        // waves sweep whole windows, so one after the other would cost the
        // second pass's one-wave latency on top of the first's.  Their trace
        // slabs are disjoint: the whole-tile pass's follow the first pass's.
        // Short windows keep the whole-tile pass alone: there the stripe is
        // most of each matrix, and the window's bookkeeping costs more than
        // the extra waves per SIMD return.
        const bool windowed = pruned && c->full_prune == 1 && esc_maxn >= espk::PRUNE_WINDOW_MIN_N;
        uint64_t blocks0 = 0, blocks1 = blocks;
        if (windowed) {
            blocks1 = std::max<uint64_t>(1, std::min<uint64_t>(blocks, std::max<uint64_t>(1, cap / 3)));
            blocks0 = std::max<uint64_t>(
                1, std::min({(uint64_t)c->ncu * espk::gotoh_pruned_blocks_per_cu(rows), (bound + 63) / 64,
                             cap > blocks1 ? cap - blocks1 : (uint64_t)1}));
        }
        HIP_TRY(c, c->ftrace.ensure(std::max({maxw * 64 * 4 * (blocks0 + (windowed ? blocks1 : blocks)),
                                              prune ? wave_w * 4 * blocks_w : 0})));
        f.trace = (uint32_t *)c->ftrace.p;
        f.maxw = (uint32_t)maxw;
        if (c->census && !c->cap) {
            HIP_TRY(c, c->censusbuf.ensure(64));
            HIP_TRY(c, espk::launch_zero_u32((uint32_t *)c->censusbuf.p, 16, c->stream));
            f.census = (unsigned long long *)c->censusbuf.p;
            f.census_first = 1;
            c->last_census = true;
        }
        if (pruned) {
            espk::EscSortArgs sa{};
            sa.list = a.esc_list;
            sa.count = &dp->esc_range[1];
            sa.read_len = b->read_len;
            sa.esc_info = a.esc_info;
            sa.s = prune_s;
            sa.pad = c->prune_pad_a;
            sa.esc = esc;
            sa.plist = plist;
            sa.fblist = fblist;
            sa.win_len = b->win_len;
            sa.census = f.census;
            HIP_TRY(c, espk::launch_esc_sort(sa, (uint32_t)bound, c->stream));
            tim_mark(c, 12, 0);
            c->last_pruned = true;
            espk::GotohArgs g = f;
            g.order = plist;
            g.range = esc->prange;
            g.counter = &esc->pcounter;
            g.esc_info = a.esc_info;
            g.prune_s = prune_s;
            g.prune_half = c->prune_half_a;
            g.prune_pad = c->prune_pad_a;
            g.fb_list = fblist;
            g.fb_range = esc->fbrange;
            g.census_first = 0;  // the sort counted the matrices
            if (windowed) {
                espk::GotohArgs g0 = g;
                g0.range = esc->crange0;
                g0.counter = &esc->c0counter;
                g.range = esc->crange1;
                g.trace = (uint32_t *)c->ftrace.p + blocks0 * maxw * 64;
                HIP_TRY(c, hipEventRecord(c->ev_fork, c->stream));
                HIP_TRY(c, hipStreamWaitEvent(c->stream2, c->ev_fork, 0));
                // the join is queued whatever fails in between, so that no
                // work of this call is left on the second stream
                const hipError_t e1 = espk::launch_gotoh_full(g, rows, true, (int)blocks1, c->stream2);
                const hipError_t e2 = hipEventRecord(c->ev_join, c->stream2);
                const hipError_t e3 =
                    e1 == hipSuccess ? espk::launch_gotoh_pruned(g0, rows, (int)blocks0, c->stream) : hipSuccess;
                const hipError_t e4 = e2 == hipSuccess ? hipStreamWaitEvent(c->stream, c->ev_join, 0) : hipSuccess;
                if (e2 != hipSuccess) (void)hipStreamSynchronize(c->stream2);
                HIP_TRY(c, e1);
                HIP_TRY(c, e2);
                HIP_TRY(c, e3);
                HIP_TRY(c, e4);
            } else {
                HIP_TRY(c, espk::launch_gotoh_full(g, rows, true, (int)blocks, c->stream));
            }
            tim_mark(c, 7, rows);
            // What the pass could not prove, and the pairs the sort found no
            // pass could help: one wave per pair over the whole matrix, up to
            // the list cap (windows over the wave kernel's limit, and the
            // rest of a long list, go on to the unpruned lane kernel).
            espk::GotohArgs w = f;
            w.census_first = 0;  // the sort counted the matrices
            w.order = fblist;
            w.range = esc->fbrange;
            w.fb_list = fb2list;
            w.fb_range = esc->fb2range;
            w.trace = (uint32_t *)c->ftrace.p;
            w.maxw = (uint32_t)std::max<uint64_t>(wave_w, 1);
            // two rounds of the grid: a longer list runs faster in the lane
            // kernel, whose one-wave latency is a whole window sweep
            w.list_cap = (uint32_t)std::min<uint64_t>(2 * blocks_w, 0xFFFFFFFFull);
            w.wave_lds_n = wave_ln;
            w.wave_lds_tr = 0;
            HIP_TRY(c, espk::launch_gotoh_wave_list(w, (int)blocks_w, c->stream));
            tim_mark(c, 13, rows);
            f.order = fb2list;
            f.range = esc->fb2range;
            f.counter = &esc->fb2counter;
            f.census_first = 0;
            HIP_TRY(c, espk::launch_gotoh_full(f, rows, false, (int)blocks, c->stream));
            tim_mark(c, 11, rows);
        } else if (prune) {
            // Reads past the pruned pass's tiles: the escalations run one wave
            // per pair first (up to the list cap), the rest in the unpruned
            // lane kernel, whose one-wave latency is a whole window sweep.
            c->last_pruned = true;
            c->last_wave_direct = true;
            espk::GotohArgs w = f;
            w.fb_list = fb2list;
            w.fb_range = esc->fb2range;
            w.trace = (uint32_t *)c->ftrace.p;
            w.maxw = (uint32_t)std::max<uint64_t>(wave_w, 1);
            w.list_cap = (uint32_t)std::min<uint64_t>(2 * blocks_w, 0xFFFFFFFFull);
            w.wave_lds_n = wave_ln;
            w.wave_lds_tr = 0;
            HIP_TRY(c, espk::launch_gotoh_wave_list(w, (int)blocks_w, c->stream));
            tim_mark(c, 13, rows);
            f.order = fb2list;
            f.range = esc->fb2range;
            f.counter = &esc->fb2counter;
            HIP_TRY(c, espk::launch_gotoh_full(f, rows, false, (int)blocks, c->stream));
            tim_mark(c, 11, rows);
        } else {
            HIP_TRY(c, espk::launch_gotoh_full(f, rows, false, (int)blocks, c->stream));
            tim_mark(c, 7, rows);
        }
    }
    return ESP_OK;
}


extern "C" {

size_t esp_cigar_capacity(size_t m) { return 2 * m + 5; }

int esp_create(int device, esp_ctx **out) {
    if (!out) return ESP_E_INVAL;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) return ESP_E_DEVICE;
    esp_ctx *c = new esp_ctx();
    c->device = device;
    if (hipSetDevice(device) != hipSuccess) {
        delete c;
        return ESP_E_DEVICE;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) {
        delete c;
        return ESP_E_DEVICE;
    }
    c->ncu = prop.multiProcessorCount;
    // the allocation-free entry points' device scratch, once per device
    esph::noalloc_setup(device);
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess ||
        hipEventCreateWithFlags(&c->ev_plan, hipEventDisableTiming) != hipSuccess ||
        hipStreamCreateWithFlags(&c->stream2, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming) != hipSuccess) {
        delete c;
        return ESP_E_DEVICE;
    }
    for (int k = 0; k <= esp_ctx::kMaxEv; k++) {
        if (hipEventCreate(&c->evs[k]) != hipSuccess) {
            delete c;
            return ESP_E_DEVICE;
        }
    }
    *out = c;
    return ESP_OK;
}

int esp_destroy(esp_ctx *c) {
    if (!c) return ESP_OK;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    c->counter.release();
    c->plan.release();
    c->order.release();
    c->hplan.release();
    c->rows.release();
    c->trace.release();
    c->btrace.release();
    c->ftrace.release();
    if (c->ev_plan) (void)hipEventDestroy(c->ev_plan);
    if (c->stream2) (void)hipStreamSynchronize(c->stream2);
    if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
    if (c->ev_join) (void)hipEventDestroy(c->ev_join);
    if (c->stream2) (void)hipStreamDestroy(c->stream2);
    c->stage.release();
    c->stage_out.release();
    c->wquery.release();
    c->warena.release();
    c->wdesc.release();
    c->wmyers.release();
    c->eng.release();
    c->jscratch.release();
    c->jarena.release();
    c->mlist.release();
    c->scantmp.release();
    c->escbuf.release();
    c->hstage.release();
    c->hout.release();
    for (int k = 0; k <= esp_ctx::kMaxEv; k++)
        if (c->evs[k]) (void)hipEventDestroy(c->evs[k]);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return ESP_OK;
}

const char *esp_last_error(const esp_ctx *c) { return c ? c->err.c_str() : "no context"; }

int esp_sync(esp_ctx *c) {
    if (!c) return ESP_E_INVAL;
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return ESP_OK;
}

int esp_extend(esp_ctx *c, const uint8_t *read, size_t m, const uint8_t *win, size_t n,
               const esp_params *p, esp_ext *out, uint32_t *cigar, size_t cap, size_t *n_ops) {
    return run_single(c, read, m, win, n, p, ESP_MODE_FULL, 0, 0, out, nullptr, cigar, cap, n_ops,
                      nullptr);
}

int esp_extend_hint(esp_ctx *c, const uint8_t *read, size_t m, const uint8_t *win, size_t n,
                    const esp_params *p, int64_t hint_offset, esp_ext *out, uint32_t *cigar,
                    size_t cap, size_t *n_ops) {
    return run_single(c, read, m, win, n, p, ESP_MODE_HINT, hint_offset, 0, out, nullptr, cigar,
                      cap, n_ops, nullptr);
}

int esp_band(esp_ctx *c, const uint8_t *read, size_t m, const uint8_t *win, size_t n,
             const esp_params *p, int has_band, int64_t offset, int64_t half, esp_ext *out,
             esp_touch *touch, uint32_t *cigar, size_t cap, size_t *n_ops) {
    return run_single(c, read, m, win, n, p, has_band ? ESP_MODE_BAND : ESP_MODE_FULL, offset,
                      half, out, touch, cigar, cap, n_ops, nullptr);
}

int esp_gotoh_batch(esp_ctx *c, const esp_batch *b, const esp_host_meta *meta,
                    const esp_params *p, int mode, esp_batch_out *o) {
    (void)meta;  // the device planner reads the lengths itself
    if (!c) return ESP_E_INVAL;
    HIP_TRY(c, hipSetDevice(c->device));
    HIP_TRY(c, ev_record(c, c->ev0));
    tim_begin(c);
    const int rc = gotoh_batch_core(c, b, p, mode, o);
    if (rc) return rc;
    HIP_TRY(c, ev_record(c, c->ev1));
    return ESP_OK;
}

int esp_last_kernel_ms(esp_ctx *c, double *ms) {
    if (!c || !ms) return ESP_E_INVAL;
    float f = 0.f;
    HIP_TRY(c, hipEventSynchronize(c->ev1));
    HIP_TRY(c, hipEventElapsedTime(&f, c->ev0, c->ev1));
    *ms = f;
    return ESP_OK;
}

int esp_last_timings(esp_ctx *c, esp_timing *out, size_t cap, size_t *n) {
    if (!c || !n) return ESP_E_INVAL;
    *n = (size_t)c->nev;
    if (c->nev == 0) return ESP_OK;
    HIP_TRY(c, hipEventSynchronize(c->evs[c->nev]));
    for (int k = 0; k < c->nev && (size_t)k < cap; k++) {
        float f = 0.f;
        HIP_TRY(c, hipEventElapsedTime(&f, c->evs[k], c->evs[k + 1]));
        out[k] = c->tim[k];
        out[k].ms = f;
    }
    return ESP_OK;
}

int esp_ctx_set_option(esp_ctx *c, int option, int64_t value) {
    if (!c) return ESP_E_INVAL;
    switch (option) {
    case ESP_OPT_FULL_KERNEL:
        // 0 (by batch size) and 1 (lane per pair) run the lane-per-pair kernel;
        // the wave-per-pair kernel is not built.
        if (value < 0 || value > 1) return fail(c, ESP_E_INVAL, "full kernel choice not in {0,1}");
        c->full_kernel = (int)value;
        return ESP_OK;
    case ESP_OPT_FULL_CENSUS:
        if (value < 0 || value > 1) return fail(c, ESP_E_INVAL, "full-matrix census not in {0,1}");
        c->census = (int)value;
        return ESP_OK;
    case ESP_OPT_PLAIN_KERNEL:
        if (value < 0 || value > 1) return fail(c, ESP_E_INVAL, "plain kernel choice not in {0,1}");
        c->plain_kernel = (int)value;
        return ESP_OK;
    case ESP_OPT_FULL_PRUNE:
        if (value < 0 || value > 2) return fail(c, ESP_E_INVAL, "full-matrix pruning not in {0,1,2}");
        c->full_prune = (int)value;
        return ESP_OK;
    case ESP_OPT_FULL_PRUNE_SHAPE: {
        if (value == 0) {
            c->prune_half_a = espk::ESC_HALF_A;
            c->prune_pad_a = espk::ESC_PAD_A;
            return ESP_OK;
        }
        const int f[2] = {(int)(value & 255), (int)((value >> 8) & 255)};
        if (value < 0 || value > 0xFFFFll || !f[0] || !f[1])
            return fail(c, ESP_E_INVAL, "pruning shape: half | pad << 8, each 1..255");
        c->prune_half_a = f[0];
        c->prune_pad_a = f[1];
        return ESP_OK;
    }
    case ESP_OPT_TRACE_BUDGET:
        if (value < (1 << 20)) return fail(c, ESP_E_INVAL, "trace budget below 1 MiB");
        c->trace_budget = (size_t)value;
        c->full_trace_budget = (size_t)value * 2;
        return ESP_OK;
    case ESP_OPT_BAND_GRID: {
        const int half = (int)(value >> 16), k = (int)(value & 0xFFFF);
        for (int cls = 0; cls < espk::PLAN_STRIP; cls++) {
            if (espk::plan_class_half(cls) != half) continue;
            const int occ = espk::gotoh_band_blocks_per_cu(half);
            if (k < 0 || k > occ) return fail(c, ESP_E_INVAL, "blocks per CU above the occupancy limit");
            c->band_blocks[cls] = k ? k : occ;
            return ESP_OK;
        }
        return fail(c, ESP_E_INVAL, "no band kernel for that half");
    }
    default:
        return fail(c, ESP_E_INVAL, "unknown option");
    }
}

int esp_autotune_band(esp_ctx *c, const esp_batch *b, const esp_host_meta *meta, const esp_params *p,
                      esp_batch_out *o, int reps, esp_tune_report *rep) {
    if (!c) return ESP_E_INVAL;
    if (!b || !o || !rep || b->npairs == 0) return fail(c, ESP_E_INVAL, "probe batch, output and report are required");
    if (reps < 1 || reps > 64) return fail(c, ESP_E_INVAL, "reps must be 1..64");
    HIP_TRY(c, hipSetDevice(c->device));
    std::memset(rep, 0, sizeof *rep);
    int occ[espk::PLAN_STRIP], maxocc = 1;
    for (int cls = 0; cls < espk::PLAN_STRIP; cls++) {
        occ[cls] = espk::gotoh_band_blocks_per_cu(espk::plan_class_half(cls));
        maxocc = std::max(maxocc, occ[cls]);
        rep->half[cls] = espk::plan_class_half(cls);
        rep->occupancy[cls] = occ[cls];
    }
    double best[espk::PLAN_STRIP];
    for (int cls = 0; cls < espk::PLAN_STRIP; cls++) best[cls] = 1e30;
    // Candidate k: every class with occupancy >= k runs k blocks (waves) per
    // CU; each class keeps its fastest k (its kernel's time is its own).
    for (int k = 1; k <= maxocc; k++) {
        for (int cls = 0; cls < espk::PLAN_STRIP; cls++) c->band_blocks[cls] = std::min(k, occ[cls]);
        double sum[espk::PLAN_STRIP] = {};
        for (int r = 0; r <= reps; r++) {  // the first run warms up
            int rc = esp_gotoh_batch(c, b, meta, p, ESP_MODE_BAND, o);
            if (rc) return rc;
            esp_timing tim[esp_ctx::kMaxEv];
            size_t nt = 0;
            rc = esp_last_timings(c, tim, esp_ctx::kMaxEv, &nt);
            if (rc) return rc;
            if (r == 0) continue;
            for (size_t e = 0; e < nt; e++)
                for (int cls = 0; cls < espk::PLAN_STRIP; cls++)
                    if (tim[e].kind == 1 && tim[e].half == espk::plan_class_half(cls)) sum[cls] += tim[e].ms;
        }
        for (int cls = 0; cls < espk::PLAN_STRIP; cls++) {
            if (k > occ[cls]) continue;
            const double ms = sum[cls] / reps;
            if (k == occ[cls]) rep->ms_default[cls] = ms;
            if (ms < best[cls]) {
                best[cls] = ms;
                rep->blocks_per_cu[cls] = k;
                rep->ms_best[cls] = ms;
            }
        }
        rep->candidates++;
    }
    // A half the probe does not use keeps its occupancy limit, and a choice
    // within 2% of the limit's time is taken as noise.
    esp_plan_stats st;
    int rc = esp_last_plan_stats(c, &st);
    if (rc) return rc;
    for (int cls = 0; cls < espk::PLAN_STRIP; cls++) {
        if (st.band_pairs[cls] == 0 || rep->ms_best[cls] > 0.98 * rep->ms_default[cls]) {
            rep->blocks_per_cu[cls] = occ[cls];
            rep->ms_best[cls] = rep->ms_default[cls];
        }
        c->band_blocks[cls] = rep->blocks_per_cu[cls];
    }
    return ESP_OK;
}

int esp_autotune_prune(esp_ctx *c, const esp_batch *b, const esp_host_meta *meta, const esp_params *p,
                       esp_batch_out *o, int reps, esp_prune_tune_report *rep) {
    if (!c) return ESP_E_INVAL;
    if (!b || !o || !rep || b->npairs == 0) return fail(c, ESP_E_INVAL, "probe batch, output and report are required");
    if (reps < 1 || reps > 64) return fail(c, ESP_E_INVAL, "reps must be 1..64");
    HIP_TRY(c, hipSetDevice(c->device));
    std::memset(rep, 0, sizeof *rep);
    // the default first, then narrower and wider stripes and heads (the
    // pipeline tuner, esperanto_amd.batch.tune_prune_shape, tries the same)
    static const int kShapes[][2] = {{espk::ESC_HALF_A, espk::ESC_PAD_A}, {24, 16}, {48, 16}, {24, 12},
                                     {32, 12}, {24, 8},  {32, 8},  {32, 24}};
    const int nshape = (int)(sizeof kShapes / sizeof kShapes[0]);
    const int keep_half = c->prune_half_a, keep_pad = c->prune_pad_a;
    bool escalated = false;
    for (int s = 0; s < nshape; s++) {
        c->prune_half_a = kShapes[s][0];
        c->prune_pad_a = kShapes[s][1];
        double sum = 0.0;
        for (int r = 0; r <= reps; r++) {  // the first run warms up
            int rc = esp_gotoh_batch(c, b, meta, p, ESP_MODE_HINT, o);
            if (rc) {
                c->prune_half_a = keep_half;
                c->prune_pad_a = keep_pad;
                return rc;
            }
            esp_timing tim[esp_ctx::kMaxEv];
            size_t nt = 0;
            rc = esp_last_timings(c, tim, esp_ctx::kMaxEv, &nt);
            if (rc) return rc;
            if (r == 0) continue;
            for (size_t e = 0; e < nt; e++)
                if (tim[e].kind == 7 || tim[e].kind == 11 || tim[e].kind == 12 || tim[e].kind == 13) {
                    sum += tim[e].ms;
                    escalated = true;
                }
        }
        rep->half[s] = kShapes[s][0];
        rep->pad[s] = kShapes[s][1];
        rep->ms[s] = sum / reps;
        rep->candidates++;
    }
    rep->ms_default = rep->ms[0];
    int best = 0;
    for (int s = 1; s < nshape; s++)
        if (rep->ms[s] < rep->ms[best]) best = s;
    // a choice within 2% of the default's time is taken as noise
    if (!escalated || rep->ms[best] > 0.98 * rep->ms_default) best = 0;
    rep->chosen_half = c->prune_half_a = kShapes[best][0];
    rep->chosen_pad = c->prune_pad_a = kShapes[best][1];
    rep->ms_best = rep->ms[best];
    return ESP_OK;
}

int esp_last_plan_stats(esp_ctx *c, esp_plan_stats *out) {
    if (!c || !out) return ESP_E_INVAL;
    std::memset(out, 0, sizeof(*out));
    if (!c->plan.p) return ESP_OK;
    HIP_TRY(c, hipSetDevice(c->device));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    espk::PlanDev *dp = (espk::PlanDev *)c->plan.p;
    espk::PlanSummary sum;
    uint32_t esc[2];
    HIP_TRY(c, hipMemcpy(&sum, &dp->sum, sizeof(sum), hipMemcpyDeviceToHost));
    HIP_TRY(c, hipMemcpy(esc, dp->esc_range, sizeof(esc), hipMemcpyDeviceToHost));
    for (int k = 0; k < espk::PLAN_STRIP; k++) out->band_pairs[k] = sum.range[k][1];
    out->strip_pairs = sum.range[espk::PLAN_STRIP][1];
    out->escalated = esc[1];
    out->full_unpruned = esc[1];
    if (c->last_pruned && c->escbuf.p) {
        uint32_t fb[2], fa[2];
        const espk::EscDev *ed = (const espk::EscDev *)c->escbuf.p;
        HIP_TRY(c, hipMemcpy(fb, ed->fb2range, sizeof(fb), hipMemcpyDeviceToHost));
        HIP_TRY(c, hipMemcpy(fa, ed->fbrange, sizeof(fa), hipMemcpyDeviceToHost));
        out->full_unpruned = fb[1];
        out->full_pruned = esc[1] - fb[1];
        // the pairs the one-wave rerun was given: the pruned pass's failures
        // and the sort's hopeless pairs, or every escalation of a long read
        out->full_second = c->last_wave_direct ? esc[1] : fa[1];
        if (!c->last_wave_direct) {
            uint32_t c0[2], c1[2];
            HIP_TRY(c, hipMemcpy(c0, ed->crange0, sizeof(c0), hipMemcpyDeviceToHost));
            HIP_TRY(c, hipMemcpy(c1, ed->crange1, sizeof(c1), hipMemcpyDeviceToHost));
            out->full_small_head = c0[1];
            out->full_large_head = c1[1];
        }
    }
    if (c->last_census && c->censusbuf.p) {
        unsigned long long cs[3];
        HIP_TRY(c, hipMemcpy(cs, c->censusbuf.p, sizeof(cs), hipMemcpyDeviceToHost));
        out->full_cells_computed = cs[0];
        out->full_lane_cells_issued = cs[1];
        out->full_cells_matrix = cs[2];
    }
    return ESP_OK;
}

int esp_device_copy(esp_ctx *c, void *dst, const void *src, size_t bytes) {
    if (!c || (bytes && (!dst || !src))) return ESP_E_INVAL;
    HIP_TRY(c, hipSetDevice(c->device));
    HIP_TRY(c, espk::launch_copy(dst, src, bytes, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return ESP_OK;
}

int esp_myers_batch(esp_ctx *c, const esp_batch *b, const esp_host_meta *meta, int32_t *dist,
                    uint64_t *end, uint64_t *start) {
    if (!c) return ESP_E_INVAL;
    if (!b || !meta || !meta->read_len || !dist)
        return fail(c, ESP_E_INVAL, "batch, host lengths and outputs are required");
    if (b->npairs == 0) return ESP_OK;
    if (b->npairs > 0xFFFFFFFFull) return fail(c, ESP_E_INVAL, "too many pairs");
    HIP_TRY(c, hipSetDevice(c->device));
    HIP_TRY(c, c->counter.ensure(256));
    HIP_TRY(c, c->hout.ensure(64));
    espk::MyersArgs a{};
    a.npairs = (uint32_t)b->npairs;
    a.order = nullptr;
    a.reads = b->reads;
    a.read_off = b->read_off;
    a.read_len = b->read_len;
    a.texts = b->wins;
    a.text_off = b->win_off;
    a.text_len = b->win_len;
    a.dist = dist;
    a.end = end;
    a.start = start;
    uint32_t *cnt = (uint32_t *)c->counter.p;  // [0] queue, [16..17] lengths, [32] overflow queue, [48..49] its range
    a.counter = cnt;
    uint32_t maxm;
    if (meta->max_read_len) {
        // The caller's bound sizes the words; longer patterns (up to 256)
        // are listed on the device and an 8-word launch drains the list
        // (an empty list ends it at once).  Nothing is read back.
        maxm = std::min<uint32_t>(meta->max_read_len, 256);
        HIP_TRY(c, c->mlist.ensure(b->npairs * 4));
        a.ovf_list = (uint32_t *)c->mlist.p;
        a.ovf_range = cnt + 48;
        HIP_TRY(c, espk::launch_zero_u32(cnt, 64, c->stream));
    } else {
        // The longest pattern of at most 256 bases sizes the kernel's words
        // (longer ones take the guard) and the longest text is checked: both
        // come from one reduction over the device length arrays, 8 bytes read
        // back, which costs far less than scanning the host copies.
        uint32_t *lmax = cnt + 16;
        HIP_TRY(c, espk::launch_len_max(b->read_len, b->win_len, b->npairs, lmax, c->ncu, c->stream));
        HIP_TRY(c, hipMemcpyAsync(c->hout.p, lmax, 8, hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(c, hipStreamSynchronize(c->stream));
        maxm = ((const uint32_t *)c->hout.p)[0];
        const uint32_t maxn = ((const uint32_t *)c->hout.p)[1];
        if (maxn > kMaxWin) return fail(c, ESP_E_INVAL, "text too long");
        HIP_TRY(c, espk::launch_zero_u32(cnt, 1, c->stream));
    }
    const int per_cu = espk::myers_blocks_per_cu((int)maxm, false);
    uint64_t blocks = (uint64_t)c->ncu * per_cu;
    blocks = std::min<uint64_t>(blocks, (b->npairs + 63) / 64);
    blocks = std::max<uint64_t>(blocks, 1);
    HIP_TRY(c, ev_record(c, c->ev0));
    tim_begin(c);
    HIP_TRY(c, espk::launch_myers(a, (int)maxm, (int)blocks, c->stream));
    tim_mark(c, 3, 0);
    if (a.ovf_list && maxm < 256) {
        espk::MyersArgs o = a;
        o.order = a.ovf_list;
        o.range = a.ovf_range;
        o.ovf_list = nullptr;
        o.ovf_range = nullptr;
        o.counter = cnt + 32;
        const uint64_t ob = std::max<uint64_t>(1, std::min<uint64_t>((uint64_t)c->ncu * espk::myers_blocks_per_cu(256, false),
                                                                   (b->npairs + 63) / 64));
        HIP_TRY(c, espk::launch_myers(o, 256, (int)ob, c->stream));
        tim_mark(c, 3, 256);
    }
    HIP_TRY(c, ev_record(c, c->ev1));
    return ESP_OK;
}

// The engine's EA-Myers over (pattern, text) pairs: texts at
// texts + text_off[k], text_len[k]; pattern lengths (host copy) in 1..=256.
static int engine_myers_launch(esp_ctx *c, size_t n, const uint8_t *reads, const uint64_t *read_off,
                               const uint32_t *read_len, const uint32_t *host_read_len, const uint8_t *texts,
                               const uint64_t *text_off, const uint32_t *text_len, int global, int32_t *dist) {
    uint32_t maxm = 1;
    for (size_t k = 0; k < n; k++) {
        const uint32_t m = host_read_len[k];
        if (m == 0 || m > 256) return fail(c, ESP_E_INVAL, "engine EA-Myers patterns must be 1..256 bases");
        maxm = std::max(maxm, m);
    }
    HIP_TRY(c, c->counter.ensure(256));
    espk::MyersArgs a{};
    a.npairs = (uint32_t)n;
    a.reads = reads;
    a.read_off = read_off;
    a.read_len = read_len;
    a.texts = texts;
    a.text_off = text_off;
    a.text_len = text_len;
    a.dist = dist;
    a.counter = (uint32_t *)c->counter.p;
    a.global = global;
    uint64_t blocks = (uint64_t)c->ncu * espk::myers_blocks_per_cu((int)maxm, global != 0);
    blocks = std::max<uint64_t>(std::min<uint64_t>(blocks, (n + 63) / 64), 1);
    HIP_TRY(c, espk::launch_zero_u32(a.counter, 1, c->stream));
    HIP_TRY(c, espk::launch_myers(a, (int)maxm, (int)blocks, c->stream));
    tim_mark(c, 3, 0);
    return ESP_OK;
}

int esp_engine_myers_batch(esp_ctx *c, const esp_batch *b, const esp_host_meta *meta, int mode, int32_t *dist) {
    if (!c) return ESP_E_INVAL;
    if (!b || !meta || !meta->read_len || !dist)
        return fail(c, ESP_E_INVAL, "batch, host lengths and outputs are required");
    if (mode != 0 && mode != 1) return fail(c, ESP_E_INVAL, "mode must be 0 (infix) or 1 (global)");
    if (b->npairs == 0) return ESP_OK;
    if (meta->win_len)
        for (size_t k = 0; k < b->npairs; k++)
            if (meta->win_len[k] > kMaxWin) return fail(c, ESP_E_INVAL, "text too long");
    if (b->npairs > 0xFFFFFFFFull) return fail(c, ESP_E_INVAL, "too many pairs");
    HIP_TRY(c, hipSetDevice(c->device));
    HIP_TRY(c, ev_record(c, c->ev0));
    tim_begin(c);
    int rc = engine_myers_launch(c, b->npairs, b->reads, b->read_off, b->read_len, meta->read_len, b->wins,
                                 b->win_off, b->win_len, mode, dist);
    if (rc) return rc;
    HIP_TRY(c, ev_record(c, c->ev1));
    return ESP_OK;
}

int esp_engine_verify_batch(esp_ctx *c, const esp_engine_verify *v, const uint32_t *host_read_len, uint32_t flank,
                            int32_t *dist, int32_t *status) {
    if (!c) return ESP_E_INVAL;
    if (!v || !host_read_len || !dist || !status) return fail(c, ESP_E_INVAL, "NULL argument");
    if (v->n == 0) return ESP_OK;
    if (v->n > 0xFFFFFFFFull) return fail(c, ESP_E_INVAL, "too many candidates");
    if (!v->reads || !v->read_off || !v->read_len || !v->tx || !v->tx_off || !v->tx_len || !v->anchor)
        return fail(c, ESP_E_INVAL, "candidate arrays are required");
    HIP_TRY(c, hipSetDevice(c->device));
    const size_t n = v->n;
    HIP_TRY(c, c->eng.ensure(n * 12 + 16));
    espk::EnginePrepArgs pa{};
    pa.n = (uint32_t)n;
    pa.read_len = v->read_len;
    pa.tx_off = v->tx_off;
    pa.tx_len = v->tx_len;
    pa.anchor = v->anchor;
    pa.flank = flank;
    pa.text_off = (uint64_t *)c->eng.p;
    pa.text_len = (uint32_t *)((uint8_t *)c->eng.p + n * 8);
    pa.status = status;
    HIP_TRY(c, ev_record(c, c->ev0));
    tim_begin(c);
    HIP_TRY(c, espk::launch_engine_prep(pa, c->stream));
    tim_mark(c, 5, 0);
    int rc = engine_myers_launch(c, n, v->reads, v->read_off, v->read_len, host_read_len, v->tx, pa.text_off,
                                 pa.text_len, 0, dist);
    if (rc) return rc;
    HIP_TRY(c, ev_record(c, c->ev1));
    return ESP_OK;
}

static int myers_single(esp_ctx *c, const uint8_t *read, size_t m, const uint8_t *text, size_t n,
                        int32_t *dist, size_t *end, size_t *start) {
    if (!c) return ESP_E_INVAL;
    if ((m && !read) || (n && !text) || !dist) return fail(c, ESP_E_INVAL, "NULL argument");
    if (n > kMaxWin) return fail(c, ESP_E_INVAL, "text too long");
    HIP_TRY(c, hipSetDevice(c->device));
    PairStage L = pair_layout(m, n);
    HIP_TRY(c, c->hstage.ensure(L.total));
    HIP_TRY(c, c->stage.ensure(L.total));
    HIP_TRY(c, c->hout.ensure(64));
    HIP_TRY(c, c->stage_out.ensure(64));
    uint8_t *h = (uint8_t *)c->hstage.p;
    if (m) std::memcpy(h + L.o_read, read, m);
    if (n) std::memcpy(h + L.o_win, text, n);
    uint64_t *meta = (uint64_t *)(h + L.o_meta);
    meta[0] = 0;
    meta[1] = 0;
    uint32_t *lens = (uint32_t *)(meta + 2);
    lens[0] = (uint32_t)m;
    lens[1] = (uint32_t)n;
    HIP_TRY(c, hipMemcpyAsync(c->stage.p, h, L.total, hipMemcpyHostToDevice, c->stream));
    uint8_t *d = (uint8_t *)c->stage.p;
    uint8_t *dout = (uint8_t *)c->stage_out.p;
    HIP_TRY(c, c->counter.ensure(256));
    espk::MyersArgs a{};
    a.npairs = 1;
    a.reads = d + L.o_read;
    a.texts = d + L.o_win;
    a.read_off = (const uint64_t *)(d + L.o_meta);
    a.text_off = (const uint64_t *)(d + L.o_meta) + 1;
    a.read_len = (const uint32_t *)(d + L.o_meta + 16);
    a.text_len = (const uint32_t *)(d + L.o_meta + 20);
    a.dist = (int32_t *)dout;
    a.end = (uint64_t *)(dout + 8);
    a.start = start ? (uint64_t *)(dout + 16) : nullptr;
    a.counter = (uint32_t *)c->counter.p;
    HIP_TRY(c, espk::launch_zero_u32(a.counter, 1, c->stream));
    HIP_TRY(c, espk::launch_myers(a, m > 256 ? 1 : (int)m, 1, c->stream));
    HIP_TRY(c, hipMemcpyAsync(c->hout.p, dout, 24, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    const uint8_t *ho = (const uint8_t *)c->hout.p;
    std::memcpy(dist, ho, 4);
    uint64_t e, s;
    std::memcpy(&e, ho + 8, 8);
    std::memcpy(&s, ho + 16, 8);
    if (end) *end = (size_t)e;
    if (start) *start = (size_t)s;
    return ESP_OK;
}

int esp_myers_infix(esp_ctx *c, const uint8_t *read, size_t m, const uint8_t *text, size_t n, int32_t *dist) {
    return myers_single(c, read, m, text, n, dist, nullptr, nullptr);
}

int esp_myers_best_end(esp_ctx *c, const uint8_t *read, size_t m, const uint8_t *text, size_t n,
                       int32_t *dist, size_t *end) {
    if (!end) return fail(c, ESP_E_INVAL, "NULL argument");
    return myers_single(c, read, m, text, n, dist, end, nullptr);
}

int esp_myers_best_start(esp_ctx *c, const uint8_t *read, size_t m, const uint8_t *text,
                         size_t n, int32_t *dist, size_t *start) {
    if (!start) return fail(c, ESP_E_INVAL, "NULL argument");
    return myers_single(c, read, m, text, n, dist, nullptr, start);
}

int esp_device_alloc(esp_ctx *c, size_t bytes, void **out) {
    if (!c || !out) return ESP_E_INVAL;
    *out = nullptr;
    HIP_TRY(c, hipSetDevice(c->device));
    hipError_t e = hipMalloc(out, std::max<size_t>(bytes, 1));
    if (e != hipSuccess) return fail(c, ESP_E_NOMEM, hipGetErrorString(e));
    return ESP_OK;
}

int esp_device_free(esp_ctx *c, void *p) {
    if (!c) return ESP_E_INVAL;
    if (!p) return ESP_OK;
    HIP_TRY(c, hipSetDevice(c->device));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    HIP_TRY(c, hipFree(p));
    return ESP_OK;
}

int esp_copy_to_device(esp_ctx *c, void *dst, const void *src, size_t bytes) {
    if (!c || (bytes && (!dst || !src))) return ESP_E_INVAL;
    if (!bytes) return ESP_OK;
    HIP_TRY(c, hipSetDevice(c->device));
    HIP_TRY(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return ESP_OK;
}

int esp_copy_to_host(esp_ctx *c, void *dst, const void *src, size_t bytes) {
    if (!c || (bytes && (!dst || !src))) return ESP_E_INVAL;
    if (!bytes) return ESP_OK;
    HIP_TRY(c, hipSetDevice(c->device));
    HIP_TRY(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return ESP_OK;
}

// stitch_cigar (splice.rs:435-494); the walk is shared with the spliced
// pipeline kernel (esp_stitch.hpp).
int esp_stitch_cigar(const uint32_t *cigar, size_t n_ops, uint32_t p_start,
                     const uint32_t *bounds, size_t n_exons, const uint32_t *skips,
                     size_t n_junctions, uint32_t *out, size_t cap, size_t *n_out) {
    if ((n_ops && !cigar) || !n_out || (n_exons && !bounds) || (n_junctions && !skips))
        return ESP_E_INVAL;
    if (n_ops > 0xFFFFFFFFull || n_exons > 0xFFFFFFFFull || n_junctions > 0xFFFFFFFFull)
        return ESP_E_INVAL;
    const uint32_t cap32 = (uint32_t)std::min<size_t>(cap, 0xFFFFFFFFull);
    const uint32_t n = esps::stitch_walk(cigar, (uint32_t)n_ops, p_start, bounds,
                                         (uint32_t)n_exons, skips, (uint32_t)n_junctions,
                                         out ? out : nullptr, out ? cap32 : 0u);
    *n_out = n;
    if (n > cap || (n && !out)) return ESP_E_CAP;
    return ESP_OK;
}

}  // extern "C"
