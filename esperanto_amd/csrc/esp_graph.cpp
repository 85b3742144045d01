// esp_graph.cpp -- a batched Gotoh call captured as one HIP graph.
//
// esp_gotoh_batch in band or full mode is a fixed kernel sequence: the
// planner (zero, histogram, scan, scatter), one persistent band kernel per
// half specialisation and the persistent strip kernel, each draining its
// class from device counters.  Only the strip kernel's scratch size comes
// from the host (the eager call reads it back from the planner); here it is
// sized from the host copies of the lengths at capture time, so the sequence
// needs nothing from the host and is captured once.  The graph owns its
// planner state and scratch, so eager calls on the same context cannot move
// memory under it.  A replay runs the whole batch again on whatever the
// batch's device arrays hold: same pair count, reads and windows no longer
// than at capture (longer pairs are counted as bad and left alone; the check
// entry point reports them).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>

#include "../../include/esp_align.h"
#include "esp_ctx.hpp"
#include "esp_kernels.hpp"

using namespace esph;

struct esp_graph {
    int device = 0;
    uint32_t npairs = 0;
    DevBuf plan, order, btrace, rows, trace;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace {

void graph_release(esp_graph *g) {
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    g->plan.release();
    g->order.release();
    g->btrace.release();
    g->rows.release();
    g->trace.release();
    delete g;
}

// Enqueues the batch on the context stream (under capture): planner, band
// kernels, strip kernel, all on the graph's own buffers.
int enqueue_batch(esp_ctx *c, esp_graph *g, const esp_batch *b, const esp_params *p, int mode,
                  esp_batch_out *o, uint32_t maxm, uint32_t maxn, uint64_t maxw, int strip_per_cu) {
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

    espk::PlanArgs pa{};
    pa.npairs = a.npairs;
    pa.read_len = b->read_len;
    pa.win_len = b->win_len;
    pa.offset = b->offset;
    pa.half = b->half;
    pa.band_ok = mode == ESP_MODE_BAND && band_lut_fits(p);
    pa.hint_mode = 0;
    // pairs longer than at capture, and strip pairs whose trace would not fit
    // the captured strip scratch (a longer read with a longer window than any
    // one captured pair), count as bad
    pa.max_read = maxm;
    pa.max_win = maxn;
    pa.max_strip_w = maxw;
    pa.plan = (espk::PlanDev *)g->plan.p;
    pa.order = (uint32_t *)g->order.p;
    HIP_TRY(c, espk::launch_plan(pa, c->ncu, c->stream));
    a.order = pa.order;
    espk::PlanDev *dp = pa.plan;

    if (pa.band_ok) {
        for (int cls = 0; cls < espk::PLAN_STRIP; cls++) {
            const int half = espk::plan_class_half(cls);
            const uint64_t bw = espk::band_trace_words(espk::band_max_read(half), half);
            if (!c->band_blocks[cls]) c->band_blocks[cls] = espk::gotoh_band_blocks_per_cu(half);
            uint64_t blocks = (uint64_t)c->ncu * c->band_blocks[cls];
            blocks = std::min<uint64_t>(blocks, (b->npairs + 63) / 64);
            blocks = std::min<uint64_t>(blocks, c->trace_budget / (bw * 64 * 4));
            blocks = std::max<uint64_t>(blocks, 1);
            if (bw * 64 * 4 * blocks > g->btrace.bytes) return fail(c, ESP_E_NOMEM, "band trace scratch");
            a.trace = (uint32_t *)g->btrace.p;
            a.maxw = (uint32_t)bw;
            a.range = &dp->sum.range[cls][0];
            a.counter = &dp->counters[cls];
            HIP_TRY(c, espk::launch_gotoh_band(a, half, (int)blocks, c->stream));
        }
    }

    // Strip kernel over its class (empty classes exit at once), sized for
    // the longest read and the largest trace any pair of the batch could need.
    uint64_t lanes = (uint64_t)c->ncu * strip_per_cu * espk::GOTOH_BLOCK;
    lanes = std::min<uint64_t>(lanes, align_up(b->npairs, 64));
    const uint64_t tw = std::max<uint64_t>(maxw, 1), rw = 5 * (uint64_t)std::max<uint32_t>(maxm, 1);
    lanes = std::min<uint64_t>(lanes, c->trace_budget / (tw * 4));
    lanes = std::min<uint64_t>(lanes, kRowBudget / (rw * 4));
    lanes = std::max<uint64_t>(lanes / 64 * 64, 64);
    if (rw * lanes * 4 > g->rows.bytes || tw * lanes * 4 > g->trace.bytes)
        return fail(c, ESP_E_NOMEM, "strip scratch");
    a.nl = (uint32_t)lanes;
    a.maxm = std::max<uint32_t>(maxm, 1);
    a.maxw = (uint32_t)tw;
    a.rows = (int32_t *)g->rows.p;
    a.trace = (uint32_t *)g->trace.p;
    a.range = &dp->sum.range[espk::PLAN_STRIP][0];
    a.counter = &dp->counters[espk::PLAN_STRIP];
    HIP_TRY(c, espk::launch_gotoh_strip(a, (int)(lanes / espk::GOTOH_BLOCK), c->stream));
    return ESP_OK;
}

}  // namespace

extern "C" {

int esp_gotoh_graph_create(esp_ctx *c, const esp_batch *b, const esp_host_meta *meta, const esp_params *p,
                           int mode, esp_batch_out *o, esp_graph **out) {
    if (!c) return ESP_E_INVAL;
    if (!out) return fail(c, ESP_E_INVAL, "NULL argument");
    *out = nullptr;
    int rc = check_params(c, p);
    if (rc) return rc;
    if (!b || !o || !meta || !meta->read_len || !meta->win_len)
        return fail(c, ESP_E_INVAL, "batch, output and host lengths are required");
    if (mode != ESP_MODE_BAND && mode != ESP_MODE_FULL)
        return fail(c, ESP_E_INVAL, "graphs capture band or full mode");
    if (b->npairs == 0 || b->npairs > 0xFFFFFF00ull) return fail(c, ESP_E_INVAL, "pair count out of range");
    if (mode == ESP_MODE_BAND && (!b->offset || !b->half))
        return fail(c, ESP_E_INVAL, "offset/half arrays required for this mode");
    uint32_t maxm = 1, maxn = 0;
    uint64_t maxw = 1;
    for (size_t k = 0; k < b->npairs; k++) {
        const uint32_t m = meta->read_len[k], n = meta->win_len[k];
        if (m > kMaxRead || n > kMaxWin) return fail(c, ESP_E_INVAL, "read longer than 16000 or window longer than 1<<24");
        maxm = std::max(maxm, m);
        maxn = std::max(maxn, n);
        maxw = std::max<uint64_t>(maxw, espk::strip_trace_words(m, n));
    }
    HIP_TRY(c, hipSetDevice(c->device));
    esp_graph *g = new esp_graph();
    g->device = c->device;
    g->npairs = (uint32_t)b->npairs;
    // scratch: the largest band trace of any half, the strip rows and trace
    uint64_t btrace = 0;
    for (int cls = 0; cls < espk::PLAN_STRIP; cls++) {
        const int half = espk::plan_class_half(cls);
        const uint64_t bw = espk::band_trace_words(espk::band_max_read(half), half);
        if (!c->band_blocks[cls]) c->band_blocks[cls] = espk::gotoh_band_blocks_per_cu(half);
        uint64_t blocks = (uint64_t)c->ncu * c->band_blocks[cls];
        blocks = std::min<uint64_t>(blocks, (b->npairs + 63) / 64);
        blocks = std::min<uint64_t>(blocks, c->trace_budget / (bw * 64 * 4));
        blocks = std::max<uint64_t>(blocks, 1);
        btrace = std::max<uint64_t>(btrace, bw * 64 * 4 * blocks);
    }
    const int per_cu = espk::gotoh_strip_blocks_per_cu(dev_params(p).edit);
    uint64_t lanes = (uint64_t)c->ncu * per_cu * espk::GOTOH_BLOCK;
    lanes = std::min<uint64_t>(lanes, align_up(b->npairs, 64));
    lanes = std::max<uint64_t>(lanes, 64);
    const uint64_t rw = 5 * (uint64_t)maxm;
    const uint64_t rows = std::min<uint64_t>(rw * lanes * 4, kRowBudget) + rw * 64 * 4;
    const uint64_t trace = std::min<uint64_t>(maxw * lanes * 4, c->trace_budget) + maxw * 64 * 4;
    hipError_t e = g->plan.ensure(sizeof(espk::PlanDev));
    if (e == hipSuccess) e = g->order.ensure(b->npairs * 4);
    if (e == hipSuccess) e = g->btrace.ensure(btrace);
    if (e == hipSuccess) e = g->rows.ensure(rows);
    if (e == hipSuccess) e = g->trace.ensure(trace);
    if (e != hipSuccess) {
        graph_release(g);
        return fail(c, ESP_E_NOMEM, "graph scratch allocation failed");
    }
    // Finish pending work first: capture records, it does not run.
    e = hipStreamSynchronize(c->stream);
    if (e == hipSuccess) e = hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) {
        graph_release(g);
        return hip_fail(c, e, "begin capture");
    }
    rc = enqueue_batch(c, g, b, p, mode, o, maxm, maxn, maxw, per_cu);
    hipGraph_t graph = nullptr;
    e = hipStreamEndCapture(c->stream, &graph);
    g->graph = graph;
    if (rc) {
        graph_release(g);
        return rc;
    }
    if (e == hipSuccess) e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        graph_release(g);
        return hip_fail(c, e, "graph capture");
    }
    *out = g;
    return ESP_OK;
}

int esp_graph_launch(esp_ctx *c, esp_graph *g) {
    if (!c) return ESP_E_INVAL;
    if (!g) return fail(c, ESP_E_INVAL, "NULL graph");
    if (g->device != c->device) return fail(c, ESP_E_INVAL, "graph belongs to another device");
    HIP_TRY(c, hipSetDevice(c->device));
    HIP_TRY(c, hipEventRecord(c->ev0, c->stream));
    tim_begin(c);
    HIP_TRY(c, hipGraphLaunch(g->exec, c->stream));
    tim_mark(c, 10, 0);
    HIP_TRY(c, hipEventRecord(c->ev1, c->stream));
    return ESP_OK;
}

int esp_graph_check(esp_ctx *c, esp_graph *g) {
    if (!c) return ESP_E_INVAL;
    if (!g) return fail(c, ESP_E_INVAL, "NULL graph");
    HIP_TRY(c, hipSetDevice(c->device));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    espk::PlanSummary sum;
    HIP_TRY(c, hipMemcpy(&sum, g->plan.p, sizeof(sum), hipMemcpyDeviceToHost));
    if (sum.bad) return fail(c, ESP_E_INVAL, "pairs longer than at capture were left alone");
    return ESP_OK;
}

int esp_graph_destroy(esp_ctx *c, esp_graph *g) {
    if (!g) return ESP_OK;
    if (c) (void)hipSetDevice(g->device);
    if (c) (void)hipStreamSynchronize(c->stream);
    graph_release(g);
    return ESP_OK;
}

}  // extern "C"
