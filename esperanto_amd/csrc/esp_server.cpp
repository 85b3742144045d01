// esp_server.cpp -- serving loop with continuous batching over the resident
// reference (include/esp_align.h, esp_server_*).
//
// Submitted requests are copied and queued in arrival order.  Each worker
// thread owns a context (its own HIP stream and scratch) and loops: take the
// longest prefix of the queue that fits its pair cap and the device-memory
// budget still free, stage it in one pinned buffer, run the batch on its
// stream (window gather + extend_hint for extension pairs, the mate-rescue
// pipeline for rescue pairs), copy results back, complete the requests and
// return the budget.  With several workers, batches overlap on the device,
// and a request that arrives while others run is taken by the next worker
// that frees up: batching is continuous, not a fixed queue drain.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/esp_align.h"
#include "esp_ctx.hpp"

using namespace esph;

namespace {

uint64_t now_ns() {
    return (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(
               std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

struct Request {
    uint64_t ticket = 0;
    uint32_t kind = 0;
    uint32_t npairs = 0;
    std::vector<uint8_t> reads;
    std::vector<uint32_t> read_len, contig, win_start, win_end, mate_pos;
    std::vector<int64_t> hint;
    std::vector<uint64_t> read_off;  // within reads
    uint64_t win_bytes = 0;          // EXTEND: sum of window lengths
    uint64_t cig_words = 0;          // sum of CIGAR capacities
    uint64_t bytes = 0;              // device bytes charged to the budget
    esp_request_out out{};
    esp_request_info info{};
};

// Device bytes a request occupies while its batch runs: inputs, outputs and
// the per-pair scratch of the pipelines (descriptors, planner order, mate
// windows of at most 2 * (750 + m) + m bases and the read / reverse
// complement slots).
uint64_t request_bytes(const Request &r) {
    uint64_t b = r.reads.size() + 4 * r.cig_words + 256ull * r.npairs;
    if (r.kind == ESP_REQ_EXTEND) {
        b += r.win_bytes + 64ull * r.npairs;
    } else {
        uint32_t maxm = 1;
        for (uint32_t m : r.read_len) maxm = std::max(maxm, std::min<uint32_t>(m, 256));
        b += (uint64_t)r.npairs * (1500 + 3ull * maxm + 4ull * maxm + 160);
    }
    return b;
}

// Sections of a batch in the staging buffers, 256-byte aligned.
struct Layout {
    size_t off = 0;
    size_t take(size_t bytes) {
        off = align_up(off, 256);
        const size_t o = off;
        off += bytes;
        return o;
    }
};

}  // namespace

struct esp_server {
    int device = 0;
    const esp_ref *ref = nullptr;
    esp_server_config cfg{};
    uint64_t data_budget = 0;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Request *> pending;
    std::unordered_map<uint64_t, Request *> live;
    uint64_t next_ticket = 1;
    uint64_t inflight_bytes = 0;
    bool stop = false;
    esp_server_stats st{};
    std::vector<std::thread> threads;
    std::vector<esp_ctx *> ctxs;
};

namespace {

struct Worker {
    esp_server *s;
    esp_ctx *c;
    DevBuf din, dout;
    HostBuf hin, hout;
};

// Marks requests done (status rc) and returns their budget.
void complete(esp_server *s, const std::vector<Request *> &reqs, int rc, uint64_t seq) {
    if (reqs.empty()) return;
    const uint64_t t = now_ns();
    {
        std::lock_guard<std::mutex> lk(s->mu);
        for (Request *r : reqs) {
            s->inflight_bytes -= r->bytes;
            r->info.status = rc;
            r->info.done = 1;
            r->info.done_ns = t;
            r->info.batch = seq;
            s->st.completed++;
            s->st.pairs += r->npairs;
        }
    }
    s->cv_done.notify_all();
    s->cv_work.notify_all();
}

// Copies one request's records, counts and valid CIGAR words out of the
// host result buffer.  The callers' CIGAR arenas are laid out like the
// batch's (capacities in pair order); only each pair's cig_n words are
// copied, so no word of another request's pairs (left in a reused device
// slot) ever reaches a caller.
void scatter(Request *r, const uint8_t *rec, size_t rec_size, const uint8_t *cn, const uint8_t *cg) {
    std::memcpy(r->kind == ESP_REQ_EXTEND ? (void *)r->out.ext : (void *)r->out.mate, rec,
                rec_size * r->npairs);
    std::memcpy(r->out.cig_n, cn, 4 * r->npairs);
    uint64_t at = 0;
    for (uint32_t k = 0; k < r->npairs; k++) {
        uint32_t nops;
        std::memcpy(&nops, cn + 4 * k, 4);
        const uint64_t cap = esp_cigar_capacity(r->read_len[k]);
        std::memcpy(r->out.cigar + at, cg + 4 * at, 4 * std::min<uint64_t>(nops, cap));
        at += cap;
    }
}

// Runs one batch on the worker's stream and completes its requests: the
// extension pairs first (their results go out before the mate-rescue phase
// starts, so short extension requests never wait for rescue windows), then
// the rescue pairs.
void run_batch(Worker &w, std::vector<Request *> &batch, uint64_t seq) {
    esp_server *s = w.s;
    esp_ctx *c = w.c;
    uint64_t nE = 0, nR = 0, eread = 0, rread = 0, ewin = 0, ecig = 0, rcig = 0;
    uint32_t rmaxm = 1;
    for (Request *r : batch) {
        if (r->kind == ESP_REQ_EXTEND) {
            nE += r->npairs;
            eread += r->reads.size();
            ewin += r->win_bytes;
            ecig += r->cig_words;
        } else {
            nR += r->npairs;
            rread += r->reads.size();
            rcig += r->cig_words;
            for (uint32_t m : r->read_len) rmaxm = std::max(rmaxm, std::min<uint32_t>(m, 256));
        }
    }
    const uint64_t t0 = now_ns();
    hipError_t e0 = hipSetDevice(c->device);  // per host thread
    // Inputs (staged on the host, one copy to the device).
    Layout L;
    const size_t o_ereads = L.take(eread), o_eroff = L.take(8 * nE), o_erlen = L.take(4 * nE),
                 o_jobs = L.take(sizeof(esp_gather_job) * nE), o_ewoff = L.take(8 * nE),
                 o_ewlen = L.take(4 * nE), o_hint = L.take(8 * nE), o_ecoff = L.take(8 * (nE + 1)),
                 o_rreads = L.take(rread), o_rroff = L.take(8 * nR), o_rrlen = L.take(4 * nR),
                 o_rctg = L.take(4 * nR), o_rpos = L.take(4 * nR), o_rcoff = L.take(8 * (nR + 1));
    const size_t in_bytes = align_up(L.off, 256);
    // Outputs (device-written, one copy back); the decoded windows sit after them.
    Layout O;
    const size_t o_eext = O.take(sizeof(esp_ext) * nE), o_ecn = O.take(4 * nE), o_ecig = O.take(4 * ecig),
                 o_rrec = O.take(sizeof(esp_mate_rec) * nR), o_rcn = O.take(4 * nR),
                 o_rcig = O.take(4 * rcig);
    const size_t out_bytes = align_up(O.off, 256);
    const size_t o_ewins = O.take(ewin + 16);
    if (e0 == hipSuccess) e0 = w.hin.ensure(in_bytes);
    if (e0 == hipSuccess) e0 = w.din.ensure(in_bytes);
    if (e0 == hipSuccess) e0 = w.hout.ensure(out_bytes);
    if (e0 == hipSuccess) e0 = w.dout.ensure(align_up(O.off, 256));
    if (e0 != hipSuccess) {
        complete(s, batch, e0 == hipErrorOutOfMemory ? ESP_E_NOMEM : ESP_E_DEVICE, seq);
        return;
    }
    uint8_t *h = (uint8_t *)w.hin.p;
    uint64_t ke = 0, kr = 0, pe = 0, pr = 0, we = 0, ce = 0, cr = 0;
    uint64_t *ecoff = (uint64_t *)(h + o_ecoff), *rcoff = (uint64_t *)(h + o_rcoff);
    for (Request *r : batch) {
        for (uint32_t k = 0; k < r->npairs; k++) {
            const uint32_t m = r->read_len[k];
            const uint8_t *src = r->reads.data() + r->read_off[k];
            const uint64_t cap = esp_cigar_capacity(m);
            if (r->kind == ESP_REQ_EXTEND) {
                std::memcpy(h + o_ereads + pe, src, m);
                ((uint64_t *)(h + o_eroff))[ke] = pe;
                ((uint32_t *)(h + o_erlen))[ke] = m;
                esp_gather_job j;
                j.contig = r->contig[k];
                j.start = r->win_start[k];
                j.end = r->win_end[k];
                j.flags = 0;
                j.dst_off = we;
                ((esp_gather_job *)(h + o_jobs))[ke] = j;
                ((uint64_t *)(h + o_ewoff))[ke] = we;
                ((uint32_t *)(h + o_ewlen))[ke] = j.end - j.start;
                ((int64_t *)(h + o_hint))[ke] = r->hint[k];
                ecoff[ke] = ce;
                pe += m;
                we += j.end - j.start;
                ce += cap;
                ke++;
            } else {
                std::memcpy(h + o_rreads + pr, src, m);
                ((uint64_t *)(h + o_rroff))[kr] = pr;
                ((uint32_t *)(h + o_rrlen))[kr] = m;
                ((uint32_t *)(h + o_rctg))[kr] = r->contig[k];
                ((uint32_t *)(h + o_rpos))[kr] = r->mate_pos[k];
                rcoff[kr] = cr;
                pr += m;
                cr += cap;
                kr++;
            }
        }
    }
    ecoff[nE] = ce;
    rcoff[nR] = cr;
    std::vector<Request *> ereq, rreq;
    for (Request *r : batch) (r->kind == ESP_REQ_EXTEND ? ereq : rreq).push_back(r);
    uint64_t t1 = now_ns(), t2 = t1, t3 = t1, t4 = t1, t5 = t1, t6 = t1;
    hipError_t he = hipMemcpyAsync(w.din.p, w.hin.p, in_bytes, hipMemcpyHostToDevice, c->stream);
    int rc = he == hipSuccess ? ESP_OK : hip_fail(c, he, "staging copy");
    uint8_t *d = (uint8_t *)w.din.p, *q = (uint8_t *)w.dout.p;
    const uint8_t *ho = (const uint8_t *)w.hout.p;
    if (nE && rc == ESP_OK) {
        rc = esp_gather_batch(c, s->ref, (const esp_gather_job *)(d + o_jobs), nE, q + o_ewins, nullptr);
        if (rc == ESP_OK) {
            esp_batch b{};
            b.npairs = nE;
            b.reads = d + o_ereads;
            b.read_off = (const uint64_t *)(d + o_eroff);
            b.read_len = (const uint32_t *)(d + o_erlen);
            b.wins = q + o_ewins;
            b.win_off = (const uint64_t *)(d + o_ewoff);
            b.win_len = (const uint32_t *)(d + o_ewlen);
            b.offset = (const int64_t *)(d + o_hint);
            b.half = nullptr;
            esp_batch_out bo{};
            bo.ext = (esp_ext *)(q + o_eext);
            bo.cigar = (uint32_t *)(q + o_ecig);
            bo.cig_off = (const uint64_t *)(d + o_ecoff);
            bo.cig_n = (uint32_t *)(q + o_ecn);
            rc = esp_gotoh_batch(c, &b, nullptr, &s->cfg.params, ESP_MODE_HINT, &bo);
        }
        if (rc == ESP_OK) {
            he = hipMemcpyAsync(w.hout.p, q, o_rrec, hipMemcpyDeviceToHost, c->stream);
            if (he == hipSuccess) he = hipStreamSynchronize(c->stream);
            if (he != hipSuccess) rc = hip_fail(c, he, "extension results");
        }
        t2 = now_ns();
        if (rc == ESP_OK) {
            uint64_t k0 = 0, c0 = 0;
            for (Request *r : ereq) {
                scatter(r, ho + o_eext + sizeof(esp_ext) * k0, sizeof(esp_ext), ho + o_ecn + 4 * k0,
                        ho + o_ecig + 4 * c0);
                k0 += r->npairs;
                c0 += r->cig_words;
            }
        }
    }
    t3 = now_ns();
    complete(s, ereq, rc, seq);
    t4 = t5 = t6 = now_ns();
    if (nR && rc == ESP_OK) {
        esp_mate_batch mb{};
        mb.npairs = nR;
        mb.reads = d + o_rreads;
        mb.read_off = (const uint64_t *)(d + o_rroff);
        mb.read_len = (const uint32_t *)(d + o_rrlen);
        mb.mate_contig = (const uint32_t *)(d + o_rctg);
        mb.mate_pos = (const uint32_t *)(d + o_rpos);
        mb.max_read_len = rmaxm;
        rc = esp_mate_rescue_batch(c, s->ref, &mb, &s->cfg.params, (esp_mate_rec *)(q + o_rrec),
                                   (uint32_t *)(q + o_rcig), (const uint64_t *)(d + o_rcoff),
                                   (uint32_t *)(q + o_rcn));
        if (rc == ESP_OK) {
            he = hipMemcpyAsync((uint8_t *)w.hout.p + o_rrec, q + o_rrec, out_bytes - o_rrec,
                                hipMemcpyDeviceToHost, c->stream);
            if (he == hipSuccess) he = hipStreamSynchronize(c->stream);
            if (he != hipSuccess) rc = hip_fail(c, he, "rescue results");
        }
        t5 = now_ns();
        if (rc == ESP_OK) {
            uint64_t k0 = 0, c0 = 0;
            for (Request *r : rreq) {
                scatter(r, ho + o_rrec + sizeof(esp_mate_rec) * k0, sizeof(esp_mate_rec),
                        ho + o_rcn + 4 * k0, ho + o_rcig + 4 * c0);
                k0 += r->npairs;
                c0 += r->cig_words;
            }
        }
    }
    t6 = now_ns();
    complete(s, rreq, rc, seq);
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->st.stage_ns += t1 - t0;
        s->st.extend_ns += t2 - t1;
        s->st.scatter_ns += (t3 - t2) + (t6 - t5);
        s->st.rescue_ns += t5 - t4;
    }
}

void worker_loop(esp_server *s, esp_ctx *c) {
    Worker w{s, c};
    for (;;) {
        std::vector<Request *> batch;
        uint64_t bytes = 0, pairs = 0, seq = 0;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            for (;;) {
                if (s->pending.empty()) {
                    if (s->stop) {
                        w.din.release();
                        w.dout.release();
                        w.hin.release();
                        w.hout.release();
                        return;
                    }
                    s->cv_work.wait(lk);
                    continue;
                }
                // Longest queue prefix within the pair cap and the free budget.
                while (!s->pending.empty()) {
                    Request *r = s->pending.front();
                    if (!batch.empty() && pairs + r->npairs > s->cfg.max_batch_pairs) break;
                    if (s->inflight_bytes + bytes + r->bytes > s->data_budget) break;
                    s->pending.pop_front();
                    batch.push_back(r);
                    bytes += r->bytes;
                    pairs += r->npairs;
                }
                if (!batch.empty()) break;
                s->cv_work.wait(lk);  // the head waits for budget to free up
            }
            s->inflight_bytes += bytes;
            s->st.peak_inflight_bytes = std::max(s->st.peak_inflight_bytes, s->inflight_bytes);
            s->st.max_batch_requests = std::max<uint32_t>(s->st.max_batch_requests, (uint32_t)batch.size());
            seq = ++s->st.batches;
        }
        run_batch(w, batch, seq);
    }
}

}  // namespace

extern "C" {

uint64_t esp_now_ns(void) { return now_ns(); }

int esp_server_create(int device, const esp_ref *ref, const esp_server_config *cfg,
                      esp_server **out) {
    if (!out || !ref || !cfg) return ESP_E_INVAL;
    *out = nullptr;
    if (cfg->workers < 1 || cfg->workers > 8 || cfg->max_batch_pairs < 1 || cfg->max_pending < 1 ||
        cfg->max_read_len < 1 || cfg->max_read_len > kMaxRead || ref->device != device)
        return ESP_E_INVAL;
    // The batch calls check the same envelope; refusing here keeps a server
    // from accepting requests that every batch would fail.
    if (check_params(nullptr, &cfg->params) != ESP_OK) return ESP_E_INVAL;
    // Scratch per context: band trace, full-matrix trace, planner and slack.
    const size_t trace = (size_t)1 << 30, full_trace = (size_t)4 << 30, slack = (size_t)256 << 20;
    const uint64_t reserve = (uint64_t)cfg->workers * (trace + full_trace + slack);
    if (cfg->mem_budget <= reserve) return ESP_E_NOMEM;
    esp_server *s = new esp_server();
    s->device = device;
    s->ref = ref;
    s->cfg = *cfg;
    s->data_budget = cfg->mem_budget - reserve;
    for (uint32_t k = 0; k < cfg->workers; k++) {
        esp_ctx *c = nullptr;
        const int rc = esp_create(device, &c);
        if (rc) {
            for (esp_ctx *x : s->ctxs) esp_destroy(x);
            delete s;
            return rc;
        }
        c->trace_budget = trace;
        c->full_trace_budget = full_trace;
        s->ctxs.push_back(c);
    }
    for (esp_ctx *c : s->ctxs) s->threads.emplace_back(worker_loop, s, c);
    *out = s;
    return ESP_OK;
}

int esp_server_destroy(esp_server *s) {
    if (!s) return ESP_OK;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;  // workers drain the queue, then exit
    }
    s->cv_work.notify_all();
    for (std::thread &t : s->threads) t.join();
    for (esp_ctx *c : s->ctxs) esp_destroy(c);
    for (auto &kv : s->live) delete kv.second;
    delete s;
    return ESP_OK;
}

int esp_server_submit(esp_server *s, const esp_request *req, const esp_request_out *out,
                      uint64_t *ticket) {
    if (!s || !req || !out || !ticket) return ESP_E_INVAL;
    *ticket = 0;
    if (req->kind != ESP_REQ_EXTEND && req->kind != ESP_REQ_RESCUE) return ESP_E_INVAL;
    const uint32_t n = req->npairs;
    if (n && (!req->reads || !req->read_len || !req->contig)) return ESP_E_INVAL;
    if (req->kind == ESP_REQ_EXTEND && n && (!req->win_start || !req->win_end || !req->hint))
        return ESP_E_INVAL;
    if (req->kind == ESP_REQ_RESCUE && n && !req->mate_pos) return ESP_E_INVAL;
    if (n && (!out->cigar || !out->cig_n || (req->kind == ESP_REQ_EXTEND ? !out->ext : !out->mate)))
        return ESP_E_INVAL;
    Request *r = new Request();
    r->kind = req->kind;
    r->npairs = n;
    r->out = *out;
    r->read_len.assign(req->read_len, req->read_len + n);
    r->contig.assign(req->contig, req->contig + n);
    r->read_off.resize(n);
    uint64_t total = 0;
    for (uint32_t k = 0; k < n; k++) {
        const uint32_t m = r->read_len[k];
        if (m > s->cfg.max_read_len || r->contig[k] >= s->ref->ncontig) {
            delete r;
            return ESP_E_INVAL;
        }
        r->read_off[k] = total;
        total += m;
        r->cig_words += esp_cigar_capacity(m);
    }
    r->reads.assign(req->reads, req->reads + total);
    if (req->kind == ESP_REQ_EXTEND) {
        r->win_start.assign(req->win_start, req->win_start + n);
        r->win_end.assign(req->win_end, req->win_end + n);
        r->hint.assign(req->hint, req->hint + n);
        for (uint32_t k = 0; k < n; k++) {
            const uint32_t a = r->win_start[k], b = r->win_end[k];
            if (a > b || b > s->ref->len[r->contig[k]] || b - a > kMaxWin) {
                delete r;
                return ESP_E_INVAL;
            }
            r->win_bytes += b - a;
        }
    } else {
        r->mate_pos.assign(req->mate_pos, req->mate_pos + n);
    }
    r->bytes = request_bytes(*r);
    if (r->bytes > s->data_budget) {
        delete r;
        return ESP_E_NOMEM;
    }
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->stop) {
            delete r;
            return ESP_E_INVAL;
        }
        if (s->pending.size() >= s->cfg.max_pending) {
            delete r;
            return ESP_E_BUSY;
        }
        r->ticket = s->next_ticket++;
        r->info.submit_ns = now_ns();
        s->live[r->ticket] = r;
        s->st.submitted++;
        if (n == 0) {  // nothing to compute
            r->info.done = 1;
            r->info.done_ns = r->info.submit_ns;
            s->st.completed++;
        } else {
            s->pending.push_back(r);
        }
        *ticket = r->ticket;
    }
    s->cv_work.notify_one();
    return ESP_OK;
}

int esp_server_wait(esp_server *s, uint64_t ticket, int64_t timeout_us, esp_request_info *info) {
    if (!s || !info) return ESP_E_INVAL;
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->live.find(ticket);
    if (it == s->live.end()) return ESP_E_INVAL;
    Request *r = it->second;
    auto ready = [&] { return r->info.done != 0; };
    if (timeout_us < 0) {
        s->cv_done.wait(lk, ready);
    } else {
        s->cv_done.wait_for(lk, std::chrono::microseconds(timeout_us), ready);
    }
    *info = r->info;
    if (r->info.done) {
        s->live.erase(it);
        delete r;
    }
    return ESP_OK;
}

int esp_server_stats_get(esp_server *s, esp_server_stats *out) {
    if (!s || !out) return ESP_E_INVAL;
    std::lock_guard<std::mutex> lk(s->mu);
    *out = s->st;
    out->inflight_bytes = s->inflight_bytes;
    out->pending = (uint32_t)s->pending.size();
    return ESP_OK;
}

}  // extern "C"
