// esp_junction.cpp -- host side of the junction refinement batch, the
// align_spliced chain batch and the reextend_with batch (esp_junction.hip).
//
// Both chain and placement batches need one size from the device before their
// arena can be laid out (the pseudo-reference bytes of the prepared chains,
// the window bytes of the re-extended placements): one prefix-sum kernel
// writes the offsets and the total, and the host reads back 16 bytes.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>

#include "../../include/esp_align.h"
#include "esp_ctx.hpp"
#include "esp_kernels.hpp"

using namespace esph;

namespace {

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

int junction_view(esp_ctx *c, const esp_junctions *j, espk::JunctionView &v) {
    v.jn = nullptr;
    v.cnt = nullptr;
    v.jmask = 0;
    if (!j || !j->d) return ESP_OK;  // an empty library
    if (j->device != c->device) return fail(c, ESP_E_INVAL, "junctions live on another device");
    if (!j->counts)
        return fail(c, ESP_E_INVAL, "junction library was uploaded without strands (use esp_junctions_upload_stranded)");
    v.jn = (const uint4 *)j->d;
    v.cnt = (const uint2 *)j->counts;
    v.jmask = j->cap - 1;
    return ESP_OK;
}

int splice_cfg(esp_ctx *c, const esp_splice_params *sp, espk::SpliceCfgDev &d) {
    if (!sp) return fail(c, ESP_E_INVAL, "splice params are required");
    d.min_intron = sp->min_intron;
    d.max_intron = sp->max_intron;
    d.gt_ag = sp->gt_ag_bonus;
    d.gc_ag = sp->gc_ag_bonus;
    d.at_ac = sp->at_ac_bonus;
    d.noncanon = sp->noncanonical_penalty;
    d.known = sp->known_bonus;
    return ESP_OK;
}

// Reads back the two totals a scan wrote (entry n of each array).
int read_totals(esp_ctx *c, const uint64_t *a_total, const uint64_t *b_total, uint64_t &a, uint64_t &b) {
    HIP_TRY(c, c->hout.ensure(64));
    uint64_t *h = (uint64_t *)c->hout.p;
    h[0] = h[1] = 0;
    if (a_total) HIP_TRY(c, hipMemcpyAsync(h, a_total, 8, hipMemcpyDeviceToHost, c->stream));
    if (b_total) HIP_TRY(c, hipMemcpyAsync(h + 1, b_total, 8, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    a = h[0];
    b = h[1];
    return ESP_OK;
}

}  // namespace

extern "C" {

int esp_refine_junction_batch(esp_ctx *c, const esp_ref *r, const esp_junctions *j,
                              const esp_junction_query *q, size_t n, const uint8_t *reads,
                              const esp_splice_params *sp, esp_refined_junction *out) {
    if (!c) return ESP_E_INVAL;
    if (!r || (n && (!q || !reads || !out))) return fail(c, ESP_E_INVAL, "NULL argument");
    if (r->device != c->device) return fail(c, ESP_E_INVAL, "reference lives on another device");
    if (n > 0xFFFFFF00ull) return fail(c, ESP_E_INVAL, "too many queries");
    espk::RefineArgs a{};
    int rc = junction_view(c, j, a.J);
    if (rc) return rc;
    rc = splice_cfg(c, sp, a.sp);
    if (rc) return rc;
    if (n == 0) return ESP_OK;
    HIP_TRY(c, hipSetDevice(c->device));
    a.R = r->view();
    a.n = (uint32_t)n;
    a.q = q;
    a.reads = reads;
    a.out = out;
    HIP_TRY(c, ev_record(c, c->ev0));
    tim_begin(c);
    HIP_TRY(c, espk::launch_refine(a, c->stream));
    tim_mark(c, 5, 0);
    HIP_TRY(c, ev_record(c, c->ev1));
    return ESP_OK;
}

int esp_align_spliced_batch(esp_ctx *c, const esp_ref *r, const esp_junctions *j,
                            const esp_spliced_chain_batch *b, uint32_t k, const esp_splice_params *sp,
                            const esp_params *p, esp_batch_out *o, uint32_t *stitched,
                            const uint64_t *stitch_off, esp_spliced_chain_rec *rec,
                            esp_refined_junction *junctions_out, uint32_t *bounds_out) {
    if (!c) return ESP_E_INVAL;
    int rc = check_params(c, p);
    if (rc) return rc;
    if (!r || !b || !o || !o->ext || !o->cigar || !o->cig_off || !o->cig_n || !stitched || !stitch_off || !rec)
        return fail(c, ESP_E_INVAL, "NULL argument");
    if (r->device != c->device) return fail(c, ESP_E_INVAL, "reference lives on another device");
    espk::ChainPrepArgs a{};
    rc = junction_view(c, j, a.J);
    if (rc) return rc;
    rc = splice_cfg(c, sp, a.sp);
    if (rc) return rc;
    const size_t n = b->nchain;
    if (n == 0) return ESP_OK;
    if (n > 0xFFFFFF00ull) return fail(c, ESP_E_INVAL, "too many chains");
    if (!b->reads || !b->read_off || !b->read_len || !b->contig || !b->minus || !b->anchors || !b->anc_off)
        return fail(c, ESP_E_INVAL, "batch arrays are required");
    HIP_TRY(c, hipSetDevice(c->device));
    // The anchor count sizes the per-anchor slots (segments, junctions, bounds).
    uint64_t nanc = 0, unused = 0;
    rc = read_totals(c, b->anc_off + n, nullptr, nanc, unused);
    if (rc) return rc;
    const size_t slots = std::max<uint64_t>(nanc, 1);
    HIP_TRY(c, c->jscratch.ensure(slots * (16 + 12 + sizeof(esp_refined_junction)) + n * 64 + 16 * 256));
    Carve cv{(uint8_t *)c->jscratch.p};
    a.segs = cv.take<uint32_t>(4 * slots);
    a.jrec = junctions_out ? junctions_out : cv.take<esp_refined_junction>(slots);
    a.bounds = bounds_out ? bounds_out : cv.take<uint32_t>(3 * slots);
    a.state = cv.take<int32_t>(n);
    a.refusal = cv.take<int32_t>(n);
    a.nseg = cv.take<uint32_t>(n);
    a.nex = cv.take<uint32_t>(n);
    a.plen = cv.take<uint64_t>(n);
    uint64_t *exon_off = cv.take<uint64_t>(n + 1);
    uint64_t *pseudo_off = cv.take<uint64_t>(n + 1);
    a.hint = cv.take<int64_t>(n);
    esp_splice_rec *srec = cv.take<esp_splice_rec>(n);
    a.R = r->view();
    a.npairs = (uint32_t)n;
    a.k = k;
    a.reads = b->reads;
    a.read_off = b->read_off;
    a.read_len = b->read_len;
    a.contig = b->contig;
    a.minus = b->minus;
    a.anchors = b->anchors;
    a.anc_off = b->anc_off;
    a.exon_off = exon_off;
    a.pseudo_off = pseudo_off;
    a.srec = srec;
    a.rec = rec;

    HIP_TRY(c, ev_record(c, c->ev0));
    tim_begin(c);
    HIP_TRY(c, espk::launch_chain_prep(a, c->stream));
    rc = scan2(c, a.nex, a.plen, n, exon_off, pseudo_off);
    if (rc) return rc;
    tim_mark(c, 5, 0);
    uint64_t nex = 0, pbytes = 0;
    rc = read_totals(c, exon_off + n, pseudo_off + n, nex, pbytes);
    if (rc) return rc;
    HIP_TRY(c, c->jarena.ensure(std::max<uint64_t>(nex, 1) * 8));
    a.exons = (uint32_t *)c->jarena.p;
    HIP_TRY(c, espk::launch_chain_exons(a, c->stream));
    esp_splice_batch sb{};
    sb.npairs = n;
    sb.reads = b->reads;
    sb.read_off = b->read_off;
    sb.read_len = b->read_len;
    sb.contig = b->contig;
    sb.hint = a.hint;
    sb.exons = a.exons;
    sb.exon_off = exon_off;
    sb.pseudo_off = pseudo_off;
    sb.pseudo_bytes = pbytes;
    sb.n_exons = nex;
    rc = spliced_core(c, r, &sb, p, o, stitched, stitch_off, srec);
    if (rc) return rc;
    HIP_TRY(c, espk::launch_chain_rec(a, c->stream));
    HIP_TRY(c, ev_record(c, c->ev1));
    return ESP_OK;
}

int esp_reextend_batch_run(esp_ctx *c, const esp_ref *r, const esp_reextend_batch *b, const esp_params *p,
                           int edit_dir, esp_reextend_rec *rec, uint32_t *cigar_out, const uint64_t *cig_out_off,
                           uint32_t *cig_out_n) {
    if (!c) return ESP_E_INVAL;
    if (!p) return fail(c, ESP_E_INVAL, "params is NULL");
    esp_params ep = *p;  // reextend_with runs with the candidate's EditDir
    ep.edit_dir = edit_dir;
    int rc = check_params(c, &ep);
    if (rc) return rc;
    if (!r || !b || !rec || !cigar_out || !cig_out_off || !cig_out_n) return fail(c, ESP_E_INVAL, "NULL argument");
    if (r->device != c->device) return fail(c, ESP_E_INVAL, "reference lives on another device");
    const size_t n = b->nplace;
    if (n == 0) return ESP_OK;
    if (n > 0xFFFFFF00ull) return fail(c, ESP_E_INVAL, "too many placements");
    if (!b->reads || !b->read_off || !b->read_len || !b->contig || !b->pos || !b->cigar || !b->cig_off || !b->cig_n)
        return fail(c, ESP_E_INVAL, "batch arrays are required");
    HIP_TRY(c, hipSetDevice(c->device));
    HIP_TRY(c, c->jscratch.ensure(n * 64 + 16 * 256));
    Carve cv{(uint8_t *)c->jscratch.p};
    espk::ReextArgs a{};
    a.R = r->view();
    a.n = (uint32_t)n;
    a.read_len = b->read_len;
    a.contig = b->contig;
    a.pos = b->pos;
    a.cigar = b->cigar;
    a.cig_off = b->cig_off;
    a.cig_n = b->cig_n;
    a.state = cv.take<int32_t>(n);
    a.g_read_len = cv.take<uint32_t>(n);
    a.wlen = cv.take<uint32_t>(n);
    a.hint = cv.take<int64_t>(n);
    uint64_t *win_off = cv.take<uint64_t>(n + 1);
    esp_ext *ext = cv.take<esp_ext>(n);
    a.win_off = win_off;
    a.ext = ext;
    a.cig_out = cigar_out;
    a.cig_out_off = cig_out_off;
    a.cig_out_n = cig_out_n;
    a.rec = rec;

    HIP_TRY(c, ev_record(c, c->ev0));
    tim_begin(c);
    HIP_TRY(c, espk::launch_reext_prep(a, c->stream));
    rc = scan2(c, a.wlen, nullptr, n, win_off, nullptr);
    if (rc) return rc;
    uint64_t wbytes = 0, unused = 0;
    rc = read_totals(c, win_off + n, nullptr, wbytes, unused);
    if (rc) return rc;
    HIP_TRY(c, c->jarena.ensure(wbytes + 64));
    a.arena = (uint8_t *)c->jarena.p;
    HIP_TRY(c, espk::launch_reext_gather(a, c->stream));
    tim_mark(c, 5, 0);
    esp_batch gb{};
    gb.npairs = n;
    gb.reads = b->reads;
    gb.read_off = b->read_off;
    gb.read_len = a.g_read_len;
    gb.wins = a.arena;
    gb.win_off = win_off;
    gb.win_len = a.wlen;
    gb.offset = a.hint;
    gb.half = nullptr;
    esp_batch_out go{};
    go.ext = ext;
    go.cigar = cigar_out;
    go.cig_off = cig_out_off;
    go.cig_n = cig_out_n;
    rc = gotoh_batch_core(c, &gb, &ep, ESP_MODE_HINT, &go);
    if (rc) return rc;
    HIP_TRY(c, espk::launch_reext_finish(a, c->stream));
    tim_mark(c, 6, 0);
    HIP_TRY(c, ev_record(c, c->ev1));
    return ESP_OK;
}

int esp_intron_chain_batch_run(esp_ctx *c, const esp_ref *r, const esp_junctions *j,
                               const esp_intron_chain_batch *b, uint32_t k, const esp_intron_params *ip,
                               const esp_params *p, esp_batch_out *o, uint32_t *stitched,
                               const uint64_t *stitch_off, esp_intron_chain_rec *rec) {
    if (!c) return ESP_E_INVAL;
    int rc = check_params(c, p);
    if (rc) return rc;
    if (!r || !b || !ip || !o || !o->ext || !o->cigar || !o->cig_off || !o->cig_n || !stitched || !stitch_off ||
        !rec)
        return fail(c, ESP_E_INVAL, "NULL argument");
    if (r->device != c->device) return fail(c, ESP_E_INVAL, "reference lives on another device");
    espk::IcArgs a{};
    rc = junction_view(c, j, a.J);
    if (rc) return rc;
    const size_t n = b->nread;
    if (n == 0) return ESP_OK;
    if (n > 0x7FFFFF00ull) return fail(c, ESP_E_INVAL, "too many reads");
    if (!b->reads || !b->read_off || !b->read_len || !b->anchors || !b->anc_off)
        return fail(c, ESP_E_INVAL, "batch arrays are required");
    if (b->max_read_len > kMaxRead) return fail(c, ESP_E_INVAL, "max_read_len above 16000");
    HIP_TRY(c, hipSetDevice(c->device));
    uint64_t nanc = 0, unused = 0;
    rc = read_totals(c, b->anc_off + n, nullptr, nanc, unused);
    if (rc) return rc;
    const size_t slots = std::max<uint64_t>(nanc, 1), nj = 2 * n;
    HIP_TRY(c, c->jscratch.ensure(slots * 56 + n * (sizeof(espk::IcState) + 96) + nj * 64 + 32 * 256));
    Carve cv{(uint8_t *)c->jscratch.p};
    a.mems = cv.take<uint32_t>(6 * slots);
    a.frags = cv.take<uint32_t>(8 * slots);
    a.st = cv.take<espk::IcState>(n);
    a.p_off = cv.take<uint64_t>(nj);
    a.p_len = cv.take<uint32_t>(nj);
    a.t_ctg = cv.take<uint32_t>(nj);
    a.t_lo = cv.take<uint32_t>(nj);
    a.t_len = cv.take<uint32_t>(nj);
    uint64_t *t_off = cv.take<uint64_t>(nj + 1);
    int32_t *dist = cv.take<int32_t>(nj);
    uint64_t *endp = cv.take<uint64_t>(nj);
    uint64_t *startp = cv.take<uint64_t>(nj);
    a.exons = cv.take<uint32_t>(4 * n);
    a.nex = cv.take<uint32_t>(n);
    a.ctg_out = cv.take<uint32_t>(n);
    a.exon_off = cv.take<uint64_t>(n + 1);
    a.plen = cv.take<uint64_t>(n);
    uint64_t *pseudo_off = cv.take<uint64_t>(n + 1);
    a.hint = cv.take<int64_t>(n);
    esp_splice_rec *srec = cv.take<esp_splice_rec>(n);
    a.R = r->view();
    a.sp.min_intron = 20;  // SpliceParams::default with max_intron = intron_max (intron_chain.rs:462-473)
    a.sp.max_intron = ip->intron_max;
    a.sp.gt_ag = 4;
    a.sp.gc_ag = 2;
    a.sp.at_ac = 1;
    a.sp.noncanon = 8;
    a.sp.known = 6;
    a.n = (uint32_t)n;
    a.k = k;
    a.max_read = b->max_read_len;
    a.min_frag_cov = ip->min_frag_cov;
    a.refine_window = ip->refine_window;
    a.refine_pattern_len = ip->refine_pattern_len;
    a.intron_max = ip->intron_max;
    a.reads = b->reads;
    a.read_off = b->read_off;
    a.read_len = b->read_len;
    a.anchors = b->anchors;
    a.anc_off = b->anc_off;
    a.t_off = t_off;
    a.dist = dist;
    a.end = endp;
    a.start = startp;
    a.srec = srec;
    a.ext = o->ext;
    a.stitched = stitched;
    a.stitch_off = stitch_off;
    a.rec = rec;
    HIP_TRY(c, c->counter.ensure(256));
    uint32_t *counter = (uint32_t *)c->counter.p;

    // Myers over jobs [first, first + count) of the current job arrays.
    auto myers = [&](size_t first, size_t count, int maxm, bool want_end, bool want_start) -> int {
        espk::MyersArgs m{};
        m.npairs = (uint32_t)count;
        m.reads = b->reads;
        m.read_off = a.p_off + first;
        m.read_len = a.p_len + first;
        m.texts = a.arena;
        m.text_off = t_off + first;
        m.text_len = a.t_len + first;
        m.dist = dist + first;
        m.end = want_end ? endp + first : nullptr;
        m.start = want_start ? startp + first : nullptr;
        m.counter = counter;
        const int mp = std::min(std::max(maxm, 1), 256);
        uint64_t blocks = std::min<uint64_t>((uint64_t)c->ncu * espk::myers_blocks_per_cu(mp, false), (count + 63) / 64);
        HIP_TRY(c, espk::launch_zero_u32(counter, 1, c->stream));
        HIP_TRY(c, espk::launch_myers(m, mp, (int)std::max<uint64_t>(blocks, 1), c->stream));
        return ESP_OK;
    };
    // Texts of the current jobs: offsets, one read-back of the total, decode.
    auto gather = [&]() -> int {
        int r1 = scan2(c, a.t_len, nullptr, nj, t_off, nullptr);
        if (r1) return r1;
        uint64_t tbytes = 0, u = 0;
        int r2 = read_totals(c, t_off + nj, nullptr, tbytes, u);
        if (r2) return r2;
        HIP_TRY(c, c->jarena.ensure(tbytes + 64));
        a.arena = (uint8_t *)c->jarena.p;
        HIP_TRY(c, espk::launch_ic(a, 1, c->stream));
        return ESP_OK;
    };

    HIP_TRY(c, ev_record(c, c->ev0));
    tim_begin(c);
    HIP_TRY(c, espk::launch_ic(a, 0, c->stream));
    rc = gather();
    if (rc) return rc;
    tim_mark(c, 5, 0);
    rc = myers(0, n, (int)ip->refine_pattern_len, true, false);
    if (rc) return rc;
    rc = myers(n, n, (int)ip->refine_pattern_len, false, true);
    if (rc) return rc;
    tim_mark(c, 3, 0);
    HIP_TRY(c, espk::launch_ic(a, 2, c->stream));
    rc = gather();
    if (rc) return rc;
    rc = myers(0, nj, (int)(b->max_read_len - b->max_read_len / 2), false, false);
    if (rc) return rc;
    tim_mark(c, 3, 0);
    HIP_TRY(c, espk::launch_ic(a, 3, c->stream));
    rc = scan2(c, nullptr, a.plen, n, nullptr, pseudo_off);
    if (rc) return rc;
    uint64_t pbytes = 0, u2 = 0;
    rc = read_totals(c, pseudo_off + n, nullptr, pbytes, u2);
    if (rc) return rc;
    tim_mark(c, 6, 0);
    esp_splice_batch sb{};
    sb.npairs = n;
    sb.reads = b->reads;
    sb.read_off = b->read_off;
    sb.read_len = b->read_len;
    sb.contig = a.ctg_out;
    sb.hint = a.hint;
    sb.exons = a.exons;
    sb.exon_off = a.exon_off;
    sb.pseudo_off = pseudo_off;
    sb.pseudo_bytes = pbytes;
    sb.n_exons = 2 * n;
    rc = spliced_core(c, r, &sb, p, o, stitched, stitch_off, srec);
    if (rc) return rc;
    HIP_TRY(c, espk::launch_ic(a, 4, c->stream));
    HIP_TRY(c, ev_record(c, c->ev1));
    return ESP_OK;
}

}  // extern "C"
