// esp_graph.cpp -- batched calls captured as HIP graphs.
//
// A batch call is a fixed sequence of launches whose work lists live on the
// device: the planner fills class ranges and counters that the persistent
// band, strip and full-matrix kernels drain, and extend_hint's escalation
// list is built and sorted on the device.  The eager call reads one thing
// back, the planner's summary, to size the strip and full-matrix launches.
// A graph instead takes host bounds (the longest read and window any replay
// may hold) and sizes every launch from them, so the whole call needs
// nothing from the host and is recorded once.  It owns a full set of batch
// scratch, swapped into the context while it records, so eager calls on the
// same context cannot move memory under it.  Recording runs the call twice:
// once eagerly with the graph's scratch and bounds in place (this sizes
// every buffer and writes the outputs), then under stream capture, where
// every buffer is already large enough and nothing is allocated.  A replay
// redoes the whole call on whatever the call's device arrays hold; pairs
// past the bounds are counted as bad and left alone (esp_graph_check).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>

#include "../../include/esp_align.h"
#include "esp_ctx.hpp"
#include "esp_kernels.hpp"

using namespace esph;

struct esp_graph {
    int device = 0;
    ScratchSet s;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace {

void graph_release(esp_graph *g) {
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    g->s.release();
    delete g;
}

// Records what fn enqueues on the context stream into a new graph (see the
// file comment).  fn returns an esp_status.
template <class F>
int record(esp_ctx *c, const CaptureBounds &bounds, F fn, esp_graph **out) {
    HIP_TRY(c, hipSetDevice(c->device));
    // the call's own work only: finish what is queued before the graph's
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    esp_graph *g = new esp_graph();
    g->device = c->device;
    swap_scratch(c, g->s);
    c->cap = &bounds;
    int rc = fn();
    hipError_t e = rc ? hipSuccess : hipStreamSynchronize(c->stream);
    hipGraph_t graph = nullptr;
    if (!rc && e == hipSuccess) {
        e = hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal);
        if (e == hipSuccess) {
            rc = fn();
            const hipError_t e2 = hipStreamEndCapture(c->stream, &graph);
            if (e == hipSuccess) e = e2;
        }
    }
    c->cap = nullptr;
    swap_scratch(c, g->s);
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
    if (mode < ESP_MODE_BAND || mode > ESP_MODE_HINT) return fail(c, ESP_E_INVAL, "bad mode");
    if (b->npairs == 0 || b->npairs > 0xFFFFFF00ull) return fail(c, ESP_E_INVAL, "pair count out of range");
    if ((mode != ESP_MODE_FULL && !b->offset) || (mode == ESP_MODE_BAND && !b->half))
        return fail(c, ESP_E_INVAL, "offset/half arrays required for this mode");
    CaptureBounds bounds;
    for (size_t k = 0; k < b->npairs; k++) {
        const uint32_t m = meta->read_len[k], n = meta->win_len[k];
        if (m > kMaxRead || n > kMaxWin) return fail(c, ESP_E_INVAL, "read longer than 16000 or window longer than 1<<24");
        bounds.maxm = std::max(bounds.maxm, m);
        bounds.maxn = std::max(bounds.maxn, n);
        bounds.strip_maxw = std::max<uint64_t>(bounds.strip_maxw, espk::strip_trace_words(m, n));
    }
    return record(c, bounds, [&] { return gotoh_batch_core(c, b, p, mode, o); }, out);
}

int esp_mate_rescue_graph_create(esp_ctx *c, const esp_ref *r, const esp_mate_batch *b, const esp_params *p,
                                 esp_mate_rec *rec, uint32_t *cigar, const uint64_t *cig_off, uint32_t *cig_n,
                                 esp_graph **out) {
    if (!c) return ESP_E_INVAL;
    if (!out) return fail(c, ESP_E_INVAL, "NULL argument");
    *out = nullptr;
    if (!b || b->npairs == 0) return fail(c, ESP_E_INVAL, "a graph needs a non-empty batch");
    // the pipeline's extend_hint batch: reads of at most max_read_len (256
    // at most; longer ones never place) in windows inside its window stride
    CaptureBounds bounds;
    bounds.maxm = std::max<uint32_t>(1, std::min<uint32_t>(b->max_read_len, 256));
    bounds.maxn = (uint32_t)mate_window_stride(bounds.maxm);
    bounds.strip_maxw = espk::strip_trace_words(bounds.maxm, bounds.maxn);
    return record(c, bounds, [&] { return esp_mate_rescue_batch(c, r, b, p, rec, cigar, cig_off, cig_n); }, out);
}

int esp_spliced_graph_create(esp_ctx *c, const esp_ref *r, const esp_splice_batch *b, const esp_params *p,
                             esp_batch_out *o, uint32_t *stitched, const uint64_t *stitch_off,
                             esp_splice_rec *rec, uint32_t max_read_len, uint32_t max_pseudo_len,
                             esp_graph **out) {
    if (!c) return ESP_E_INVAL;
    if (!out) return fail(c, ESP_E_INVAL, "NULL argument");
    *out = nullptr;
    if (!b || b->npairs == 0) return fail(c, ESP_E_INVAL, "a graph needs a non-empty batch");
    if (max_read_len == 0 || max_read_len > kMaxRead || max_pseudo_len == 0 || max_pseudo_len > kMaxWin)
        return fail(c, ESP_E_INVAL, "graph bounds out of range");
    CaptureBounds bounds;
    bounds.maxm = max_read_len;
    bounds.maxn = max_pseudo_len;
    bounds.strip_maxw = espk::strip_trace_words(max_read_len, max_pseudo_len);
    return record(c, bounds, [&] { return esp_spliced_batch(c, r, b, p, o, stitched, stitch_off, rec); }, out);
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
    HIP_TRY(c, hipMemcpy(&sum, g->s.plan.p, sizeof(sum), hipMemcpyDeviceToHost));
    if (sum.bad) return fail(c, ESP_E_INVAL, "pairs past the graph's bounds were left alone");
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
