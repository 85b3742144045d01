// esp_window.cpp -- resident reference and the window pipelines behind the C
// ABI: window gather, mate-window rescue and the spliced window DP.
//
// Each pipeline runs on the context stream as device stages around the same
// planner + band/strip Gotoh launches that esp_gotoh_batch uses, with no host
// round trip between stages other than the planner's size summary.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "../../include/esp_align.h"
#include "esp_ctx.hpp"
#include "esp_kernels.hpp"

using namespace esph;


namespace {

void ref_release(esp_ref *r) {
    for (void *p : {r->packed, r->pk_off, r->dlen, r->iv, r->iv_off})
        if (p) (void)hipFree(p);
    r->packed = r->pk_off = r->dlen = r->iv = r->iv_off = nullptr;
}

// Carves aligned sub-arrays out of one scratch buffer.
struct Carve {
    uint8_t *base;
    size_t off = 0;
    template <class T>
    T *take(size_t n) {
        off = align_up(off, 256);
        T *p = (T *)(base + off);
        off += n * sizeof(T);
        return p;
    }
};

size_t mate_desc_bytes(size_t n) {
    // 6 x u64 + 1 x i64 + 4 x u32 + 5 x i32 + esp_ext, plus alignment slack
    return n * (7 * 8 + 4 * 4 + 5 * 4 + sizeof(esp_ext)) + 20 * 256;
}

}  // namespace

// The window EA-Myers batch (esp_window_myers_batch and its allocation-free
// variant): argument checks, then the prep kernel decodes every job's text
// into `tarena` (job k at k * tstride) and writes the Myers descriptors into
// `desc`; the Myers kernel drains the jobs through `counter`, which is
// cleared here by a fill kernel.  Every descriptor is written for every job,
// so the counter is the only scratch that must start at zero.
int esph::window_myers_check(esp_ctx *c, const esp_ref *r, const esp_window_job *jobs, size_t njobs,
                             const uint8_t *reads, uint32_t max_text, int mode, int32_t *dist,
                             uint64_t *pos) {
    if (!c) return ESP_E_INVAL;
    if (!r || (njobs && (!jobs || !reads || !dist || (mode != 2 && !pos))))
        return fail(c, ESP_E_INVAL, "NULL argument");
    if (mode < 0 || mode > 2) return fail(c, ESP_E_INVAL, "mode not in {0,1,2}");
    if (r->device != c->device) return fail(c, ESP_E_INVAL, "reference lives on another device");
    if (njobs > 0xFFFFFF00ull || max_text > kMaxWin) return fail(c, ESP_E_INVAL, "batch too large");
    return ESP_OK;
}

size_t esph::window_myers_desc_bytes(size_t njobs) { return njobs * (2 * 8 + 2 * 4) + 8 * 256; }

size_t esph::window_myers_arena_bytes(size_t njobs, uint32_t max_text) {
    return njobs * align_up((uint64_t)max_text + 16, 16) + 64;
}

int esph::window_myers_core(esp_ctx *c, const esp_ref *r, const esp_window_job *jobs, size_t njobs,
                            const uint8_t *reads, uint32_t max_pattern, uint32_t max_text, int mode,
                            int32_t *dist, uint64_t *pos, int32_t *status, uint8_t *desc, uint8_t *tarena,
                            uint32_t *counter) {
    const size_t n = njobs;
    espk::WinMyersArgs a{};
    a.njobs = (uint32_t)n;
    a.R = r->view();
    a.jobs = jobs;
    a.max_pattern = max_pattern;
    a.max_text = max_text;
    a.tstride = align_up((uint64_t)max_text + 16, 16);
    a.tarena = tarena;
    Carve cv{desc};
    a.p_off = cv.take<uint64_t>(n);
    a.t_off = cv.take<uint64_t>(n);
    a.p_len = cv.take<uint32_t>(n);
    a.t_len = cv.take<uint32_t>(n);
    a.status = status;
    HIP_TRY(c, ev_record(c, c->ev0));
    tim_begin(c);
    HIP_TRY(c, espk::launch_winmyers_prep(a, c->stream));
    tim_mark(c, 5, 0);
    const int mp = (int)std::min<uint32_t>(std::max<uint32_t>(max_pattern, 1), 256);
    uint64_t blocks = std::min<uint64_t>((uint64_t)c->ncu * espk::myers_blocks_per_cu(mp, false), (n + 63) / 64);
    espk::MyersArgs m{};
    m.npairs = (uint32_t)n;
    m.reads = reads;
    m.read_off = a.p_off;
    m.read_len = a.p_len;
    m.texts = a.tarena;
    m.text_off = a.t_off;
    m.text_len = a.t_len;
    m.dist = dist;
    m.end = mode == 0 ? pos : nullptr;
    m.start = mode == 1 ? pos : nullptr;
    m.counter = counter;
    HIP_TRY(c, espk::launch_zero_u32(m.counter, 1, c->stream));
    HIP_TRY(c, espk::launch_myers(m, mp, (int)std::max<uint64_t>(blocks, 1), c->stream));
    tim_mark(c, 3, 0);
    HIP_TRY(c, ev_record(c, c->ev1));
    return ESP_OK;
}

extern "C" {

int esp_ref_upload(esp_ctx *c, size_t n_contigs, const uint32_t *len, const uint8_t *packed,
                   const uint64_t *packed_off, const uint32_t *iv, const uint64_t *iv_off,
                   esp_ref **out) {
    if (!c || !out) return ESP_E_INVAL;
    *out = nullptr;
    if (n_contigs == 0 || n_contigs > 0xFFFFFFFFull || !len || !packed_off || !iv_off)
        return fail(c, ESP_E_INVAL, "reference needs at least one contig and its tables");
    // N intervals must be merged, sorted and inside their contig (fasta.rs:295-304
    // builds them that way); the gather's per-run binary search relies on it.
    for (size_t k = 0; k < n_contigs; k++) {
        if (iv_off[k + 1] < iv_off[k]) return fail(c, ESP_E_INVAL, "iv_off not monotone");
        uint32_t prev_end = 0;
        for (uint64_t j = iv_off[k]; j < iv_off[k + 1]; j++) {
            const uint32_t s = iv[2 * j], e = iv[2 * j + 1];
            if (s >= e || e > len[k] || (j > iv_off[k] && s < prev_end))
                return fail(c, ESP_E_INVAL, "N intervals must be sorted, disjoint and in range");
            prev_end = e;
        }
        if (len[k] && !packed) return fail(c, ESP_E_INVAL, "packed is NULL");
    }
    HIP_TRY(c, hipSetDevice(c->device));
    // Device layout: each contig's packed bytes 16-byte aligned and followed
    // by at least 16 zero bytes.
    std::vector<uint64_t> doff(n_contigs);
    uint64_t total = 0;
    for (size_t k = 0; k < n_contigs; k++) {
        doff[k] = total;
        total += align_up(((uint64_t)len[k] + 3) / 4 + 16, 16);
    }
    std::vector<uint8_t> host(total, 0);
    for (size_t k = 0; k < n_contigs; k++)
        if (len[k]) std::memcpy(host.data() + doff[k], packed + packed_off[k], ((size_t)len[k] + 3) / 4);
    const uint64_t niv = iv_off[n_contigs];
    esp_ref *r = new esp_ref();
    r->device = c->device;
    r->ncontig = (uint32_t)n_contigs;
    r->len.assign(len, len + n_contigs);
    auto up = [&](void **dst, const void *src, size_t bytes) -> hipError_t {
        hipError_t e = hipMalloc(dst, std::max<size_t>(bytes, 16));
        if (e != hipSuccess) return e;
        r->bytes += std::max<size_t>(bytes, 16);
        if (bytes) e = hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice);
        return e;
    };
    hipError_t e = up(&r->packed, host.data(), total);
    if (e == hipSuccess) e = up(&r->pk_off, doff.data(), n_contigs * 8);
    if (e == hipSuccess) e = up(&r->dlen, len, n_contigs * 4);
    if (e == hipSuccess) e = up(&r->iv, iv, niv * 8);
    if (e == hipSuccess) e = up(&r->iv_off, iv_off, (n_contigs + 1) * 8);
    if (e != hipSuccess) {
        ref_release(r);
        delete r;
        return fail(c, e == hipErrorOutOfMemory ? ESP_E_NOMEM : ESP_E_DEVICE,
                    std::string("reference upload: ") + hipGetErrorString(e));
    }
    *out = r;
    return ESP_OK;
}

int esp_ref_free(esp_ctx *c, esp_ref *r) {
    if (!r) return ESP_OK;
    if (c) {
        HIP_TRY(c, hipSetDevice(r->device));
        HIP_TRY(c, hipStreamSynchronize(c->stream));
    }
    ref_release(r);
    delete r;
    return ESP_OK;
}

int esp_ref_device_bytes(const esp_ref *r, size_t *bytes) {
    if (!r || !bytes) return ESP_E_INVAL;
    *bytes = r->bytes;
    return ESP_OK;
}

int esp_gather_batch(esp_ctx *c, const esp_ref *r, const esp_gather_job *jobs, size_t njobs,
                     uint8_t *dst, int32_t *status) {
    if (!c) return ESP_E_INVAL;
    if (!r || (njobs && (!jobs || !dst))) return fail(c, ESP_E_INVAL, "NULL argument");
    if (njobs > 0xFFFFFFFFull) return fail(c, ESP_E_INVAL, "too many jobs");
    if (r->device != c->device) return fail(c, ESP_E_INVAL, "reference lives on another device");
    if (!njobs) return ESP_OK;
    HIP_TRY(c, hipSetDevice(c->device));
    espk::GatherArgs a;
    a.R = r->view();
    a.jobs = jobs;
    a.njobs = (uint32_t)njobs;
    a.dst = dst;
    a.status = status;
    HIP_TRY(c, ev_record(c, c->ev0));
    tim_begin(c);
    HIP_TRY(c, espk::launch_gather(a, c->stream));
    tim_mark(c, 5, 0);
    HIP_TRY(c, ev_record(c, c->ev1));
    return ESP_OK;
}

int esp_mate_rescue_batch(esp_ctx *c, const esp_ref *r, const esp_mate_batch *b,
                          const esp_params *p, esp_mate_rec *out, uint32_t *cigar,
                          const uint64_t *cig_off, uint32_t *cig_n) {
    if (!c) return ESP_E_INVAL;
    int rc = check_params(c, p);
    if (rc) return rc;
    if (!r || !b || !out || !cigar || !cig_off || !cig_n)
        return fail(c, ESP_E_INVAL, "NULL argument");
    if (r->device != c->device) return fail(c, ESP_E_INVAL, "reference lives on another device");
    if (b->npairs == 0) return ESP_OK;
    if (b->npairs > 0xFFFFFF00ull) return fail(c, ESP_E_INVAL, "too many pairs");
    if (!b->reads || !b->read_off || !b->read_len || !b->mate_contig || !b->mate_pos)
        return fail(c, ESP_E_INVAL, "batch arrays are required");
    HIP_TRY(c, hipSetDevice(c->device));
    const size_t n = b->npairs;
    const uint32_t maxm = std::max<uint32_t>(1, std::min<uint32_t>(b->max_read_len, 256));
    espk::MateArgs a{};
    a.npairs = (uint32_t)n;
    a.R = r->view();
    a.reads = b->reads;
    a.read_off = b->read_off;
    a.read_len = b->read_len;
    a.mate_contig = b->mate_contig;
    a.mate_pos = b->mate_pos;
    a.max_read = maxm;
    a.qstride = (uint32_t)(2 * align_up(maxm, 16));
    a.wstride = mate_window_stride(maxm);
    HIP_TRY(c, c->wquery.ensure(n * a.qstride + 64));
    HIP_TRY(c, c->warena.ensure(n * a.wstride + 64));
    HIP_TRY(c, c->wdesc.ensure(mate_desc_bytes(n)));
    a.qarena = (uint8_t *)c->wquery.p;
    a.warena = (uint8_t *)c->warena.p;
    Carve cv{(uint8_t *)c->wdesc.p};
    a.q_off_plus = cv.take<uint64_t>(n);
    a.q_off_minus = cv.take<uint64_t>(n);
    a.w_off = cv.take<uint64_t>(n);
    a.s_plus = cv.take<uint64_t>(n);
    a.s_minus = cv.take<uint64_t>(n);
    a.g_read_off = cv.take<uint64_t>(n);
    a.g_offset = cv.take<int64_t>(n);
    a.q_len = cv.take<uint32_t>(n);
    a.w_len = cv.take<uint32_t>(n);
    a.wstart = cv.take<uint32_t>(n);
    a.g_read_len = cv.take<uint32_t>(n);
    a.state = cv.take<int32_t>(n);
    a.d_plus = cv.take<int32_t>(n);
    a.d_minus = cv.take<int32_t>(n);
    a.sel = cv.take<int32_t>(n);
    a.dist = cv.take<int32_t>(n);
    a.ext = cv.take<esp_ext>(n);
    a.out = out;
    a.cig_n = cig_n;
    HIP_TRY(c, c->counter.ensure(256));

    HIP_TRY(c, ev_record(c, c->ev0));
    tim_begin(c);
    HIP_TRY(c, espk::launch_mate_prep(a, c->stream));
    tim_mark(c, 5, 0);
    // EA-Myers best start of both orientations over the window.
    const int mp = (int)std::min<uint64_t>(std::max<uint64_t>(maxm, 1), 256);
    const int per_cu = espk::myers_blocks_per_cu(mp, false);
    uint64_t blocks = std::min<uint64_t>((uint64_t)c->ncu * per_cu, (n + 63) / 64);
    blocks = std::max<uint64_t>(blocks, 1);
    for (int strand = 0; strand < 2; strand++) {
        espk::MyersArgs m{};
        m.npairs = (uint32_t)n;
        m.reads = a.qarena;
        m.read_off = strand ? a.q_off_minus : a.q_off_plus;
        m.read_len = a.q_len;
        m.texts = a.warena;
        m.text_off = a.w_off;
        m.text_len = a.w_len;
        m.dist = strand ? a.d_minus : a.d_plus;
        m.end = nullptr;
        m.start = strand ? a.s_minus : a.s_plus;
        m.counter = (uint32_t *)c->counter.p + 16 * strand;
        HIP_TRY(c, espk::launch_zero_u32(m.counter, 1, c->stream));
        HIP_TRY(c, espk::launch_myers(m, mp, (int)blocks, c->stream));
        tim_mark(c, 3, 0);
    }
    HIP_TRY(c, espk::launch_mate_select(a, c->stream));
    tim_mark(c, 6, 0);
    // extend_hint at the chosen start.
    esp_batch gb{};
    gb.npairs = n;
    gb.reads = a.qarena;
    gb.read_off = a.g_read_off;
    gb.read_len = a.g_read_len;
    gb.wins = a.warena;
    gb.win_off = a.w_off;
    gb.win_len = a.w_len;
    gb.offset = a.g_offset;
    gb.half = nullptr;
    esp_batch_out go{};
    go.ext = a.ext;
    go.cigar = cigar;
    go.cig_off = cig_off;
    go.cig_n = cig_n;
    rc = gotoh_batch_core(c, &gb, p, ESP_MODE_HINT, &go);
    if (rc) return rc;
    HIP_TRY(c, espk::launch_mate_finish(a, c->stream));
    tim_mark(c, 6, 0);
    HIP_TRY(c, ev_record(c, c->ev1));
    return ESP_OK;
}

int esp_window_myers_batch(esp_ctx *c, const esp_ref *r, const esp_window_job *jobs,
                           size_t njobs, const uint8_t *reads, uint32_t max_pattern,
                           uint32_t max_text, int mode, int32_t *dist, uint64_t *pos,
                           int32_t *status) {
    int rc = window_myers_check(c, r, jobs, njobs, reads, max_text, mode, dist, pos);
    if (rc || njobs == 0) return rc;
    HIP_TRY(c, hipSetDevice(c->device));
    HIP_TRY(c, c->warena.ensure(window_myers_arena_bytes(njobs, max_text)));
    HIP_TRY(c, c->wdesc.ensure(window_myers_desc_bytes(njobs)));
    HIP_TRY(c, c->counter.ensure(256));
    return window_myers_core(c, r, jobs, njobs, reads, max_pattern, max_text, mode, dist, pos, status,
                             (uint8_t *)c->wdesc.p, (uint8_t *)c->warena.p, (uint32_t *)c->counter.p);
}

int esp_unspliced_batch(esp_ctx *c, const esp_ref *r, const esp_chain_batch *b,
                        const esp_params *p, esp_unspliced_rec *out, uint32_t *cigar,
                        const uint64_t *cig_off, uint32_t *cig_n) {
    if (!c) return ESP_E_INVAL;
    int rc = check_params(c, p);
    if (rc) return rc;
    if (!r || !b || !out || !cigar || !cig_off || !cig_n) return fail(c, ESP_E_INVAL, "NULL argument");
    if (r->device != c->device) return fail(c, ESP_E_INVAL, "reference lives on another device");
    if (b->npairs == 0) return ESP_OK;
    if (b->npairs > 0xFFFFFF00ull) return fail(c, ESP_E_INVAL, "too many pairs");
    if (!b->reads || !b->read_off || !b->read_len || !b->contig || !b->minus || !b->ref_start ||
        !b->ref_end || !b->read_start || !b->read_end || !b->allow_dp)
        return fail(c, ESP_E_INVAL, "batch arrays are required");
    if (b->max_read_len > kMaxRead) return fail(c, ESP_E_INVAL, "max_read_len above 16000");
    HIP_TRY(c, hipSetDevice(c->device));
    const size_t n = b->npairs;
    const uint32_t maxm = std::max<uint32_t>(1, b->max_read_len);
    espk::UnspArgs a{};
    a.npairs = (uint32_t)n;
    a.R = r->view();
    a.reads = b->reads;
    a.read_off = b->read_off;
    a.read_len = b->read_len;
    a.contig = b->contig;
    a.minus = b->minus;
    a.rs = b->ref_start;
    a.re = b->ref_end;
    a.qs = b->read_start;
    a.qe = b->read_end;
    a.allow_dp = b->allow_dp;
    a.max_read = maxm;
    a.P = dev_params(p);
    a.qstride = (uint32_t)align_up(maxm, 16);
    a.wstride = align_up(2 * (uint64_t)maxm + 64, 16);  // read + 2 * (m/2 + 30) at most
    HIP_TRY(c, c->wquery.ensure(n * a.qstride + 64));
    HIP_TRY(c, c->warena.ensure(n * a.wstride + 64));
    HIP_TRY(c, c->wdesc.ensure(n * (5 * 8 + 6 * 4 + 3 * 4 + sizeof(esp_ext)) + 16 * 256));
    a.qarena = (uint8_t *)c->wquery.p;
    a.warena = (uint8_t *)c->warena.p;
    Carve cv{(uint8_t *)c->wdesc.p};
    a.q_off = cv.take<uint64_t>(n);
    a.w_off = cv.take<uint64_t>(n);
    a.s_start = cv.take<uint64_t>(n);
    a.g_read_off = cv.take<uint64_t>(n);
    a.g_hint = cv.take<int64_t>(n);
    a.q_len = cv.take<uint32_t>(n);
    a.w_len = cv.take<uint32_t>(n);
    a.wstart = cv.take<uint32_t>(n);
    a.g_read_len = cv.take<uint32_t>(n);
    a.ea_bi = cv.take<uint32_t>(n);
    a.ea_bj = cv.take<uint32_t>(n);
    a.state = cv.take<int32_t>(n);
    a.dist = cv.take<int32_t>(n);
    a.ea_sum = cv.take<int32_t>(n);
    a.ext = cv.take<esp_ext>(n);
    a.out = out;
    a.cigar = cigar;
    a.cig_off = cig_off;
    a.cig_n = cig_n;
    HIP_TRY(c, c->counter.ensure(256));

    HIP_TRY(c, ev_record(c, c->ev0));
    tim_begin(c);
    HIP_TRY(c, espk::launch_unsp_prep(a, c->stream));
    tim_mark(c, 5, 0);
    {  // EA-Myers best start of the read over its window (reads of 1..256 bases)
        const int mp = (int)std::min<uint32_t>(std::max<uint32_t>(maxm, 1), 256);
        uint64_t blocks = std::min<uint64_t>((uint64_t)c->ncu * espk::myers_blocks_per_cu(mp, false), (n + 63) / 64);
        espk::MyersArgs m{};
        m.npairs = (uint32_t)n;
        m.reads = a.qarena;
        m.read_off = a.q_off;
        m.read_len = a.q_len;
        m.texts = a.warena;
        m.text_off = a.w_off;
        m.text_len = a.w_len;
        m.dist = a.dist;
        m.end = nullptr;
        m.start = a.s_start;
        m.counter = (uint32_t *)c->counter.p;
        HIP_TRY(c, espk::launch_zero_u32(m.counter, 1, c->stream));
        HIP_TRY(c, espk::launch_myers(m, mp, (int)std::max<uint64_t>(blocks, 1), c->stream));
        tim_mark(c, 3, 0);
    }
    HIP_TRY(c, espk::launch_unsp_ea(a, c->stream));
    tim_mark(c, 6, 0);
    esp_batch gb{};
    gb.npairs = n;
    gb.reads = a.qarena;
    gb.read_off = a.g_read_off;
    gb.read_len = a.g_read_len;
    gb.wins = a.warena;
    gb.win_off = a.w_off;
    gb.win_len = a.w_len;
    gb.offset = a.g_hint;
    gb.half = nullptr;
    esp_batch_out go{};
    go.ext = a.ext;
    go.cigar = cigar;
    go.cig_off = cig_off;
    go.cig_n = cig_n;
    rc = gotoh_batch_core(c, &gb, p, ESP_MODE_HINT, &go);
    if (rc) return rc;
    HIP_TRY(c, espk::launch_unsp_finish(a, c->stream));
    tim_mark(c, 6, 0);
    HIP_TRY(c, ev_record(c, c->ev1));
    return ESP_OK;
}

int esp_spliced_batch(esp_ctx *c, const esp_ref *r, const esp_splice_batch *b,
                      const esp_params *p, esp_batch_out *o, uint32_t *stitched,
                      const uint64_t *stitch_off, esp_splice_rec *rec) {
    if (!c) return ESP_E_INVAL;
    int rc = check_params(c, p);
    if (rc) return rc;
    if (!r || !b || !o || !o->ext || !o->cigar || !o->cig_off || !o->cig_n || !stitched ||
        !stitch_off || !rec)
        return fail(c, ESP_E_INVAL, "NULL argument");
    if (r->device != c->device) return fail(c, ESP_E_INVAL, "reference lives on another device");
    if (b->npairs == 0) return ESP_OK;
    if (b->npairs > 0xFFFFFF00ull) return fail(c, ESP_E_INVAL, "too many pairs");
    if (!b->reads || !b->read_off || !b->read_len || !b->contig || !b->hint || !b->exons ||
        !b->exon_off || !b->pseudo_off)
        return fail(c, ESP_E_INVAL, "batch arrays are required");
    HIP_TRY(c, hipSetDevice(c->device));
    HIP_TRY(c, ev_record(c, c->ev0));
    tim_begin(c);
    rc = spliced_core(c, r, b, p, o, stitched, stitch_off, rec);
    if (rc) return rc;
    HIP_TRY(c, ev_record(c, c->ev1));
    return ESP_OK;
}

}  // extern "C"

// The spliced window pipeline on the context stream (no timing reset):
// pseudo-reference prep, extend_hint, stitch.  Arguments already checked.
int esph::spliced_core(esp_ctx *c, const esp_ref *r, const esp_splice_batch *b, const esp_params *p,
                       esp_batch_out *o, uint32_t *stitched, const uint64_t *stitch_off, esp_splice_rec *rec) {
    const size_t n = b->npairs;
    espk::SpliceArgs a{};
    a.npairs = (uint32_t)n;
    a.R = r->view();
    a.read_len = b->read_len;
    a.contig = b->contig;
    a.exons = b->exons;
    a.exon_off = b->exon_off;
    a.pseudo_off = b->pseudo_off;
    a.n_exons = b->n_exons;
    a.pseudo_bytes = b->pseudo_bytes;
    HIP_TRY(c, c->warena.ensure(b->pseudo_bytes + 64));
    HIP_TRY(c, c->wdesc.ensure(b->n_exons * 12 + n * 12 + 4 * 256));
    a.parena = (uint8_t *)c->warena.p;
    Carve cv{(uint8_t *)c->wdesc.p};
    a.bounds = cv.take<uint32_t>(3 * b->n_exons);
    a.state = cv.take<int32_t>(n);
    a.g_read_len = cv.take<uint32_t>(n);
    a.p_len = cv.take<uint32_t>(n);
    a.ext = o->ext;
    a.cigar = o->cigar;
    a.cig_off = o->cig_off;
    a.cig_n = o->cig_n;
    a.stitched = stitched;
    a.stitch_off = stitch_off;
    a.rec = rec;

    HIP_TRY(c, espk::launch_splice_prep(a, c->stream));
    tim_mark(c, 5, 0);
    esp_batch gb{};
    gb.npairs = n;
    gb.reads = b->reads;
    gb.read_off = b->read_off;
    gb.read_len = a.g_read_len;
    gb.wins = a.parena;
    gb.win_off = b->pseudo_off;
    gb.win_len = a.p_len;
    gb.offset = b->hint;
    gb.half = nullptr;
    const int rc = gotoh_batch_core(c, &gb, p, ESP_MODE_HINT, o);
    if (rc) return rc;
    HIP_TRY(c, espk::launch_splice_finish(a, c->stream));
    tim_mark(c, 6, 0);
    return ESP_OK;
}
