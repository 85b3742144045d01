/*
 * esp_align.h -- C ABI of the MI355X alignment core.
 *
 * Drop-in boundary for the reference's crates/map alignment core.  Every
 * entry point is extern "C", takes plain pointers and sizes, never throws and
 * returns an esp_status.  CIGARs are BAM u32 words (len << 4 | op) with
 * op M=0 I=1 D=2 N=3 S=4 (reference crates/bamio/src/encode.rs:169-177).
 *
 * Reference interfaces replaced (file:line in intell-biol-lab/esperanto):
 *   esp_extend             crates/map/src/extend.rs:347   extend()
 *   esp_extend_hint        crates/map/src/extend.rs:392   extend_hint()
 *   esp_band               crates/map/src/extend.rs:293   run_dispatch(.., Some((offset,1,1,half)))
 *   esp_myers_*            crates/map/src/myers_ea.rs:118-185, :249-310
 *   esp_stitch_cigar       crates/map/src/splice.rs:435   stitch_cigar()
 *   esp_ctx                crates/map/src/extend.rs:231   ExtendBuffer (per-thread scratch)
 *   esp_*_batch            crates/map/src/pipeline.rs:393-404 (rayon par_iter over reads)
 *
 * Sequences are ASCII bytes exactly as the reference receives them
 * (uppercase/lowercase ACGT, N and any other byte scoring as the reference's
 * base_code, extend.rs:151-159).
 *
 * Scoring envelope accepted by the device path (ESP_E_INVAL otherwise):
 *   gap_open >= 0, gap_ext >= 0, gap_open + gap_ext <= 16384,
 *   |match_score| <= 1024, |mismatch| <= 16384, read length <= 16000,
 *   window length <= 1 << 24.  Inside it results equal the reference's for
 *   every byte pattern, on both of its accumulator routes (run_dispatch,
 *   extend.rs:293-344): inputs whose bytes are all A, C, G, T or N follow the
 *   i16 route, whose wrapping add (extend.rs:473-501) drops every diagonal
 *   candidate above 32767 -- reachable here when read length times
 *   max(match_score, mismatch) exceeds 32767 -- and all other inputs follow
 *   the i32 route.
 *
 * A batch call that fails after queueing kernels waits for them before it
 * returns, so its arrays may be freed as soon as it returns.
 */
#ifndef ESP_ALIGN_H
#define ESP_ALIGN_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    ESP_OK = 0,
    ESP_E_INVAL = 1,   /* invalid argument / scoring outside the envelope */
    ESP_E_CAP = 2,     /* caller buffer too small; required size reported */
    ESP_E_DEVICE = 3,  /* HIP runtime error */
    ESP_E_NOMEM = 4,   /* device allocation failed / admission refused */
    ESP_E_BUSY = 5     /* server queue full */
} esp_status;

/* ExtendParams, crates/map/src/extend.rs:96-122.  edit_dir: 0 Off, 1 Ag, 2 Tc. */
typedef struct {
    int32_t match_score;
    int32_t mismatch;
    int32_t gap_open;
    int32_t gap_ext;
    int32_t edit_dir;
} esp_params;

/* Extension minus the CIGAR vector, crates/map/src/extend.rs:202-214. */
typedef struct {
    uint32_t read_start;
    uint32_t read_end;
    uint32_t ref_start;
    int32_t score;
} esp_ext;

/* TouchRec, crates/map/src/extend.rs:220-229. */
typedef struct {
    int32_t touched;
    int64_t row;
    int32_t left;
    int32_t clamped;
} esp_touch;

/* Context: one HIP stream plus device scratch; replaces ExtendBuffer.  Not
 * thread-safe; use one context per host thread.  (extend_hint's pruned
 * full-matrix reruns fork onto a second stream of the context and join back
 * before the call's next stage; callers see only the context stream.) */
typedef struct esp_ctx esp_ctx;

int esp_create(int device, esp_ctx **out);
int esp_destroy(esp_ctx *ctx);
const char *esp_last_error(const esp_ctx *ctx);
int esp_sync(esp_ctx *ctx);

/* ---- single-pair ("plain") forms: one kernel launch per call ----------- */

/* extend(): full-matrix local extension.  cap = CIGAR capacity in ops; on
 * ESP_E_CAP *n_ops holds the required count. */
int esp_extend(esp_ctx *ctx, const uint8_t *read, size_t m, const uint8_t *win, size_t n,
               const esp_params *p, esp_ext *out, uint32_t *cigar, size_t cap, size_t *n_ops);

/* extend_hint(): half-band 30 around hint_offset, escalating to the full
 * matrix when a row maximum touches the band edge or score*20 < 17*perfect. */
int esp_extend_hint(esp_ctx *ctx, const uint8_t *read, size_t m, const uint8_t *win, size_t n,
                    const esp_params *p, int64_t hint_offset, esp_ext *out, uint32_t *cigar,
                    size_t cap, size_t *n_ops);

/* run_dispatch(.., Some((offset,1,1,half))) when has_band != 0, else
 * run_dispatch(.., None); also returns the TouchRec. */
int esp_band(esp_ctx *ctx, const uint8_t *read, size_t m, const uint8_t *win, size_t n,
             const esp_params *p, int has_band, int64_t offset, int64_t half, esp_ext *out,
             esp_touch *touch, uint32_t *cigar, size_t cap, size_t *n_ops);

/* myers_ea::long::infix / infix_best_end / infix_best_start (myers_ea.rs:
 * 118-185, :249-310; m == 0 -> (0,0), m > 256 -> (m,0)); the short-module
 * functions are the same for m <= 128. */
int esp_myers_infix(esp_ctx *ctx, const uint8_t *read, size_t m, const uint8_t *text, size_t n,
                    int32_t *dist);
int esp_myers_best_end(esp_ctx *ctx, const uint8_t *read, size_t m, const uint8_t *text,
                       size_t n, int32_t *dist, size_t *end);
int esp_myers_best_start(esp_ctx *ctx, const uint8_t *read, size_t m, const uint8_t *text,
                         size_t n, int32_t *dist, size_t *start);

/* ---- batched forms ------------------------------------------------------ */

/* Pairs as concatenated bytes.  All pointers are DEVICE pointers for the
 * *_batch calls (inputs resident in HBM).  offset/half: per pair (band mode)
 * or the hint offset (extend_hint mode); half may be NULL otherwise. */
typedef struct {
    size_t npairs;
    const uint8_t *reads;
    const uint64_t *read_off;
    const uint32_t *read_len;
    const uint8_t *wins;
    const uint64_t *win_off;
    const uint32_t *win_len;
    const int64_t *offset;
    const int64_t *half;
} esp_batch;

/* Output arena (device pointers).  cig_off has npairs + 1 entries; pair k
 * owns cigar[cig_off[k] .. cig_off[k+1]) and needs at least 2*m + 5 words
 * (esp_cigar_capacity).  touch and aux may be NULL. */
typedef struct {
    esp_ext *ext;
    esp_touch *touch;
    uint32_t *cigar;
    const uint64_t *cig_off;
    uint32_t *cig_n;
    int32_t *aux; /* extend_hint: 0 band kept, 1 escalated, 2 empty, 3 band covers window */
} esp_batch_out;

typedef enum { ESP_MODE_BAND = 0, ESP_MODE_FULL = 1, ESP_MODE_HINT = 2 } esp_mode;

size_t esp_cigar_capacity(size_t m);

/* Asynchronous on the context stream; esp_sync() waits.  The host copies
 * of read_len / win_len / offset / half (host_meta, may be NULL) let the
 * planner bucket pairs without a device round trip. */
typedef struct {
    const uint32_t *read_len;
    const uint32_t *win_len;
    const int64_t *offset;
    const int64_t *half;
    /* Optional (0: unknown): a bound on the batch's read lengths.  With it,
     * esp_myers_batch sizes its kernel from the bound and reads nothing back;
     * reads past the bound still get exact results (a wider kernel takes
     * them, more slowly). */
    uint32_t max_read_len;
} esp_host_meta;

int esp_gotoh_batch(esp_ctx *ctx, const esp_batch *b, const esp_host_meta *meta,
                    const esp_params *p, int mode, esp_batch_out *out);

/* Per pair, text = window (myers_ea.rs:249-310):
 *   end != NULL            long::infix_best_end    -> dist, end
 *   start != NULL          long::infix_best_start  -> dist, start
 *   both                   both (dist is the same minimum)
 *   neither                long::infix             -> dist only
 * With meta->max_read_len set, the call queues its kernels and returns
 * without waiting (no host round trip; it can be captured in a graph).
 * Without it, the call reduces the device length arrays to the longest
 * pattern and text (one 8-byte read-back, which waits for the work already
 * queued on the context stream), refuses texts over 1<<24, and returns with
 * the kernel queued. */
int esp_myers_batch(esp_ctx *ctx, const esp_batch *b, const esp_host_meta *meta,
                    int32_t *dist, uint64_t *end, uint64_t *start);

/* Autotuner of the band kernels' persistent grids: runs the probe batch in
 * band mode (1 + reps times per candidate) with every band half at k blocks
 * (waves) per CU for k = 1 .. its occupancy limit, times each half's kernel
 * with HIP events, keeps each half's fastest k in the context for every
 * later batch and graph capture, and reports the choice.  The probe's outputs
 * are written as by esp_gotoh_batch.  No counterpart in the reference: it
 * picks launch shapes, not results. */
typedef struct {
    int32_t half[5];           /* band halves 16, 30, 32, 48, 64 */
    int32_t occupancy[5];      /* blocks per CU the occupancy limit allows */
    int32_t blocks_per_cu[5];  /* chosen */
    double ms_default[5];      /* kernel ms at the occupancy limit */
    double ms_best[5];         /* kernel ms at the choice */
    int32_t candidates;
} esp_tune_report;
int esp_autotune_band(esp_ctx *ctx, const esp_batch *b, const esp_host_meta *meta, const esp_params *p,
                      esp_batch_out *out, int reps, esp_tune_report *report);

/* Autotuner of the pruned full-matrix passes' tile shape: runs the probe
 * batch in extend_hint mode (1 + reps times per candidate) at each stripe
 * half-width and head pad of the candidate list (the default 32/16 first),
 * times the full-matrix stages (escalation sort, pruned passes, one-wave and
 * unpruned reruns) with HIP events, and keeps the fastest shape in the
 * context (as ESP_OPT_FULL_PRUNE_SHAPE would set it) when it beats the
 * default by more than 2%.  Results do not depend on the shape.  A probe with
 * no escalation keeps the default.  No counterpart in the reference: it picks
 * launch shapes, not results. */
typedef struct {
    int32_t candidates;
    int32_t half[16];  /* stripe half-width of each candidate (rows) */
    int32_t pad[16];   /* head pad of each candidate (rows) */
    double ms[16];     /* the candidate's full-matrix stages, ms per run */
    int32_t chosen_half, chosen_pad;
    double ms_default, ms_best;
} esp_prune_tune_report;
int esp_autotune_prune(esp_ctx *ctx, const esp_batch *b, const esp_host_meta *meta, const esp_params *p,
                       esp_batch_out *out, int reps, esp_prune_tune_report *report);

/* One esp_gotoh_batch call captured as a HIP graph, in any mode (extend_hint's
 * escalations included: the escalation list, its sort and the pruned and
 * unpruned full-matrix passes all run from device counters).  The graph
 * sizes every launch from the host lengths (meta->read_len and meta->win_len
 * are required) instead of reading the planner back, owns its own scratch,
 * and is recorded by running the call once eagerly (which writes the outputs)
 * and once under stream capture.  esp_graph_launch redoes the whole batch on
 * what the batch's device arrays hold at that time (same arrays, same pair
 * count, reads and windows no longer than at capture) and writes the same
 * outputs as esp_gotoh_batch would.  Pairs past those bounds are left alone
 * and counted; esp_graph_check (waits for the stream) returns ESP_E_INVAL
 * when a launch met any.  Launches on one context run in stream order; a
 * graph must not be launched on two contexts at once.  No counterpart in the
 * reference: it is the batched form's replay without per-kernel launches or
 * host round trips. */
typedef struct esp_graph esp_graph;
int esp_gotoh_graph_create(esp_ctx *ctx, const esp_batch *b, const esp_host_meta *meta, const esp_params *p,
                           int mode, esp_batch_out *out, esp_graph **graph);
int esp_graph_launch(esp_ctx *ctx, esp_graph *graph);
int esp_graph_check(esp_ctx *ctx, esp_graph *graph);
int esp_graph_destroy(esp_ctx *ctx, esp_graph *graph);

/* Device time (ms) of the last batch call (planner plus every kernel),
 * measured with HIP events on the context stream. */
int esp_last_kernel_ms(esp_ctx *ctx, double *ms);

/* Per-launch device times of the last batch call: kind (0 planner, 1 band
 * kernel, 2 strip kernel, 3 Myers kernel, 5 window gather / preparation,
 * 6 selection / stitching, 7 extend_hint full-matrix rerun of escalated band
 * attempts (the pruned pass when pruning is on), 8 placement walk kernel,
 * 9 fast lane kernel, 10 a whole graph launch, 11 the unpruned full-matrix
 * kernel, 12 the sort of the escalation list, 13 the one-wavefront-per-pair
 * rerun of what the pruned pass could not prove), half (band kernel) and
 * ms. */
typedef struct {
    int32_t kind;
    int32_t half;
    double ms;
} esp_timing;
int esp_last_timings(esp_ctx *ctx, esp_timing *out, size_t cap, size_t *n);

/* Context options.  ESP_OPT_FULL_KERNEL picks the kernel of extend_hint's
 * full-matrix reruns: 0 by batch size (default), 1 lane per pair; both run
 * the lane-per-pair kernel in this build (2, wave per pair, is refused).
 * ESP_OPT_TRACE_BUDGET caps the traceback scratch of one launch in bytes
 * (the full-matrix kernels get twice that); grids shrink to fit.
 * ESP_OPT_FULL_PRUNE (1, the default, 2 or 0): the full-matrix reruns first
 * take a pruned pass that computes only the cells that can still reach the
 * band attempt's score and proves per pair that no skipped cell could (1:
 * when the longest escalated window has 512 bases or more, pairs with a
 * head of at most 3 row groups in a pass that holds only the head and the
 * stripe in registers, beside the whole-tile pass for the rest; 2: all in
 * the whole-tile pass); the
 * pairs it cannot prove rerun one wavefront per pair over the whole matrix
 * (windows up to 4096 bases; longer ones in the unpruned lane kernel);
 * results are the same either way.  ESP_OPT_FULL_PRUNE_SHAPE sets the
 * pruned pass's stripe half-width and head pad (rows): value = half |
 * pad << 8, each 1..255 (0 restores 32, 16).
 * ESP_OPT_PLAIN_KERNEL picks the kernel of the single-pair calls (esp_band,
 * esp_extend, esp_extend_hint): 0 (default) one wavefront per pair, its 64
 * lanes sweeping 64 rows at once, for windows up to 4096 bases (longer ones
 * take the lane kernel); 1 the lane-per-pair strip kernel.  Results are the
 * same either way; esp_last_kernel_ms reads the call's kernel time.
 * ESP_OPT_FULL_CENSUS (0, the default, or 1): the full-matrix reruns also
 * count the cells they compute (esp_plan_stats.full_cells_*), at the cost of
 * a few atomics per wave; results are unchanged. */
enum { ESP_OPT_FULL_KERNEL = 1, ESP_OPT_TRACE_BUDGET = 2, ESP_OPT_BAND_GRID = 3, ESP_OPT_FULL_PRUNE = 4,
       ESP_OPT_FULL_PRUNE_SHAPE = 5, ESP_OPT_PLAIN_KERNEL = 6, ESP_OPT_FULL_CENSUS = 7 };
/* ESP_OPT_BAND_GRID sets one band half's persistent grid as esp_autotune_band
 * would: value = half << 16 | blocks per CU (1 .. the half's occupancy
 * limit; 0 restores the limit). */
int esp_ctx_set_option(esp_ctx *ctx, int option, int64_t value);

/* Routing of the last Gotoh batch (or pipeline) on this context: pairs per
 * band-kernel half (16, 30, 32, 48, 64), pairs in the general strip kernel,
 * extend_hint band attempts that escalated to the full matrix, and of those
 * the pairs the pruned pass finished and the pairs the unpruned kernel ran
 * (every escalated pair when pruning is off).  Waits for the context
 * stream. */
typedef struct {
    uint32_t band_pairs[5];
    uint32_t strip_pairs;
    uint32_t escalated;
    uint32_t full_pruned;
    uint32_t full_unpruned;
    uint32_t full_second;  /* escalated pairs the pruned pass could not prove
                              (rerun one wavefront per pair) */
    /* With ESP_OPT_FULL_CENSUS on (else 0), over every full-matrix pass of
     * the batch: cells computed inside the pairs' matrices, lane-cells issued
     * (64 lanes x 8-row groups x columns, padding included), and the cells
     * of the escalated pairs' full matrices (sum of m x n). */
    uint64_t full_cells_computed;
    uint64_t full_lane_cells_issued;
    uint64_t full_cells_matrix;
    /* Of the escalations the pruned passes took: pairs with a head of at
     * most 3 row groups and the rest (the escalation sort's classes). */
    uint32_t full_small_head;
    uint32_t full_large_head;
} esp_plan_stats;
int esp_last_plan_stats(esp_ctx *ctx, esp_plan_stats *out);

/* Device memory helpers for callers without an allocator of their own.
 * Copies are synchronous with respect to the context stream. */
int esp_device_alloc(esp_ctx *ctx, size_t bytes, void **out);
int esp_device_free(esp_ctx *ctx, void *p);
int esp_copy_to_device(esp_ctx *ctx, void *dst, const void *src, size_t bytes);
int esp_copy_to_host(esp_ctx *ctx, void *dst, const void *src, size_t bytes);
/* Plain device-to-device copy kernel (16-byte vectors); the reference byte
 * count for calibrating HBM traffic counters. */
int esp_device_copy(esp_ctx *ctx, void *dst, const void *src, size_t bytes);

/* ---- splice / mate-rescue glue (host side, no device work) ------------- */

/* stitch_cigar, splice.rs:435-494.  bounds: 3 u32 per exon (pseudo_start,
 * ref_start, ref_end); skips[k] = junction k length (end - start). */
int esp_stitch_cigar(const uint32_t *cigar, size_t n_ops, uint32_t p_start,
                     const uint32_t *bounds, size_t n_exons, const uint32_t *skips,
                     size_t n_junctions, uint32_t *out, size_t cap, size_t *n_out);

/* ---- resident reference and window pipelines ---------------------------- */

/* 2-bit packed reference in the layout of crates/map/src/fasta.rs:1-135 /
 * :291-310 (Contig{len, packed, n_intervals}).  Contig k has len[k] bases;
 * its ceil(len/4) packed bytes start at packed + packed_off[k] (base at pos
 * in byte pos/4 at shift 6 - 2*(pos%4)); its merged, sorted, half-open N
 * intervals are (iv[2j], iv[2j+1]) for j in [iv_off[k], iv_off[k+1]).
 * Host pointers: the reference is copied to the device once and stays
 * resident until esp_ref_free (the Reference held by the aligner). */
typedef struct esp_ref esp_ref;
int esp_ref_upload(esp_ctx *ctx, size_t n_contigs, const uint32_t *len, const uint8_t *packed,
                   const uint64_t *packed_off, const uint32_t *iv, const uint64_t *iv_off,
                   esp_ref **out);
int esp_ref_free(esp_ctx *ctx, esp_ref *ref);
/* Device bytes the resident reference holds. */
int esp_ref_device_bytes(const esp_ref *ref, size_t *bytes);

/* Window gather.  Job k writes contig[start, end) as ASCII (Contig::decode_into,
 * fasta.rs:77-135: A C G T, N over N intervals), reverse-complemented when
 * flags & 1 (revcomp_into, align.rs:2525-2533), to dst + dst_off.  Jobs and
 * dst are device pointers.  A job outside its contig writes nothing and
 * sets status[k] = ESP_E_INVAL (0 otherwise; status may be NULL). */
typedef struct {
    uint32_t contig;
    uint32_t start;
    uint32_t end;
    uint32_t flags;
    uint64_t dst_off;
} esp_gather_job;
int esp_gather_batch(esp_ctx *ctx, const esp_ref *ref, const esp_gather_job *jobs, size_t njobs,
                     uint8_t *dst, int32_t *status);

/* Mate-window rescue, mate_window_placement (crates/map/src/align.rs:851-921),
 * per pair: window [mate_pos - span, mate_pos + span + m) clipped to the
 * contig with span = 750 + m (MATE_WINDOW_INSERT = 500, align.rs:181); the
 * EA-Myers best start of the read and of its reverse complement over the
 * window, kept when dist <= max(m/3, 10) and start + m <= window length,
 * Plus unless Minus is strictly closer; extend_hint at that start; placed
 * when the score is positive.  Device pointers; max_read_len (host) bounds
 * read_len over the batch and sizes the scratch (at most 256: longer reads
 * never place). */
typedef struct {
    size_t npairs;
    const uint8_t *reads;
    const uint64_t *read_off;
    const uint32_t *read_len;
    const uint32_t *mate_contig;
    const uint32_t *mate_pos;
    uint32_t max_read_len;
} esp_mate_batch;
typedef struct {
    int32_t placed;      /* 1 placed, 0 no placement, -1 invalid input (contig
                            index out of range or read longer than max_read_len) */
    int32_t strand;      /* 0 Plus, 1 Minus */
    uint32_t pos;        /* window start + extension ref_start */
    int32_t score;       /* extension score */
    int32_t dist;        /* EA-Myers distance of the chosen orientation */
    uint32_t wstart;     /* window [wstart, wstart + wlen) */
    uint32_t wlen;
    uint32_t read_start; /* extension read span */
    uint32_t read_end;
} esp_mate_rec;
/* cigar/cig_off/cig_n: the extension CIGAR arena (device; pair k needs
 * esp_cigar_capacity(read_len[k]) words); cig_n = 0 when not placed. */
int esp_mate_rescue_batch(esp_ctx *ctx, const esp_ref *ref, const esp_mate_batch *b,
                          const esp_params *p, esp_mate_rec *out, uint32_t *cigar,
                          const uint64_t *cig_off, uint32_t *cig_n);

/* Spliced window DP: the extension and stitching of align_spliced
 * (splice.rs:625-664, after prepare_spliced) and of the intron-chain fast
 * path (intron_chain.rs:497-536).  Pair k's pseudo-reference concatenates
 * contig[k][rs, re) over its exons (exons[2e], exons[2e+1]) for e in
 * [exon_off[k], exon_off[k+1]); extend_hint(read, pseudo, hint[k]); the CIGAR
 * is stitched with a RefSkip of rs[e+1] - re[e] at each seam; the genomic
 * start is SplicedAlignment::genomic_start.  pseudo_off (npairs + 1) holds
 * the prefix sums of the pairs' exon lengths.  Device pointers except
 * pseudo_bytes and n_exons (host copies of the last prefix entries). */
typedef struct {
    size_t npairs;
    const uint8_t *reads;
    const uint64_t *read_off;
    const uint32_t *read_len;
    const uint32_t *contig;
    const int64_t *hint;
    const uint32_t *exons;
    const uint64_t *exon_off;
    const uint64_t *pseudo_off;
    uint64_t pseudo_bytes;  /* host: pseudo_off[npairs] */
    size_t n_exons;         /* host: exon_off[npairs] */
} esp_splice_batch;
typedef struct {
    int32_t status;         /* 1 aligned, 0 rejected (an exon with re <= rs),
                               -1 invalid (contig or exon out of range, or
                               pseudo_off inconsistent), -2 stitched CIGAR
                               capacity too small (n_stitched = required) */
    uint32_t genomic_start;
    uint32_t pseudo_len;
    uint32_t n_stitched;
} esp_splice_rec;
/* ext_out: extension records and the pseudo-reference CIGAR arena (as for
 * esp_gotoh_batch in ESP_MODE_HINT); stitched/stitch_off: the genomic CIGAR
 * arena (pair k needs esp_cigar_capacity(m) + 2 * n_exons words). */
int esp_spliced_batch(esp_ctx *ctx, const esp_ref *ref, const esp_splice_batch *b,
                      const esp_params *p, esp_batch_out *ext_out, uint32_t *stitched,
                      const uint64_t *stitch_off, esp_splice_rec *rec);

/* The mate-rescue and spliced window pipelines (esp_mate_rescue_batch,
 * esp_spliced_batch) captured as HIP graphs the same way as
 * esp_gotoh_graph_create (launch, check and destroy with esp_graph_*): one
 * graph holds the window preparation, the EA-Myers launches, the extend_hint
 * batch with its escalations and the selection or stitching.  Bounds: the mate batch's
 * max_read_len (and the window stride it implies); for the spliced batch the
 * caller states the longest read and the longest pseudo-reference any replay
 * may hold.  A replay reads the batch arrays again (mate positions, exons,
 * hints, reads may change; pair count, pseudo_bytes and n_exons may not). */
int esp_mate_rescue_graph_create(esp_ctx *ctx, const esp_ref *ref, const esp_mate_batch *b,
                                 const esp_params *p, esp_mate_rec *out, uint32_t *cigar,
                                 const uint64_t *cig_off, uint32_t *cig_n, esp_graph **graph);
int esp_spliced_graph_create(esp_ctx *ctx, const esp_ref *ref, const esp_splice_batch *b,
                             const esp_params *p, esp_batch_out *ext_out, uint32_t *stitched,
                             const uint64_t *stitch_off, esp_splice_rec *rec, uint32_t max_read_len,
                             uint32_t max_pseudo_len, esp_graph **graph);

/* Unspliced chain extension, extend_unspliced_chain after its fast lane
 * (align.rs:1770-1919): per chain, the window continuous_extension_window
 * (align.rs:33-55) of the chain's reference span [ref_start, ref_end) and read
 * span [read_start, read_end); the read in alignment orientation (reverse
 * complement for minus chains); for reads of 1..256 bases the EA-Myers lane
 * (best start within max(m/3, 10), a per-base mismatch recount equal to the
 * distance, no N, Kadane-optimal soft clips kept when they cover half the
 * read); otherwise, when allow_dp[k] != 0 (the caller's DP budget),
 * extend_hint over the window at the chain diagonal.  Device pointers;
 * max_read_len (host) bounds read_len and sizes the scratch. */
typedef struct {
    size_t npairs;
    const uint8_t *reads;
    const uint64_t *read_off;
    const uint32_t *read_len;
    const uint32_t *contig;
    const uint32_t *minus;      /* 0 Plus, 1 Minus */
    const uint32_t *ref_start;  /* Chain::ref_span */
    const uint32_t *ref_end;
    const uint32_t *read_start; /* Chain::read_span */
    const uint32_t *read_end;
    const uint32_t *allow_dp;
    uint32_t max_read_len;
} esp_chain_batch;
typedef struct {
    int32_t status;   /* 1 EA lane placement, 2 banded DP placement, 0 none,
                         -1 invalid (contig out of range, read past max_read_len) */
    int32_t strand;
    uint32_t pos;
    int32_t score;
    int32_t dist;     /* EA-Myers distance of an EA-lane placement */
    uint32_t wstart;  /* the window [wstart, wstart + wlen) */
    uint32_t wlen;
} esp_unspliced_rec;
int esp_unspliced_batch(esp_ctx *ctx, const esp_ref *ref, const esp_chain_batch *b,
                        const esp_params *p, esp_unspliced_rec *out, uint32_t *cigar,
                        const uint64_t *cig_off, uint32_t *cig_n);

/* EA-Myers of read sub-patterns against reference windows: the breakpoint
 * refinements of the intron-chain fast path (refine_donor_end: infix_best_end,
 * refine_acceptor_start: infix_best_start, intron_chain.rs:285-342) and its
 * split-read verification (verify_split_read: long::infix, :344-373).  Job k:
 * pattern reads[read_off + qstart, read_off + qend), text contig[tstart, tend).
 * mode 0: dist + best end, 1: dist + best start, 2: dist only (pos untouched).
 * status[k] = ESP_E_INVAL for a job outside its contig or the host bounds
 * max_pattern / max_text (its dist and pos are then 0).  Device pointers. */
typedef struct {
    uint64_t read_off;
    uint32_t qstart;
    uint32_t qend;
    uint32_t contig;
    uint32_t tstart;
    uint32_t tend;
    uint32_t pad;
} esp_window_job;
int esp_window_myers_batch(esp_ctx *ctx, const esp_ref *ref, const esp_window_job *jobs,
                           size_t njobs, const uint8_t *reads, uint32_t max_pattern,
                           uint32_t max_text, int mode, int32_t *dist, uint64_t *pos,
                           int32_t *status);

/* esp_window_myers_batch without any device allocation.  No counterpart in
 * the reference: it is for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments and results; its scratch (per job: the text, max_text + 16 bytes
 * rounded up to 16, and 24 bytes of descriptors; plus about 2.3 KiB) lives in
 * a 32 MiB device array that every GPU the library runs on holds until the
 * process ends.  A batch needing more returns ESP_E_INVAL before anything
 * runs (split it).  All calls on one GPU share that array, so calls that
 * overlap on one GPU (several contexts or threads) race; calls queued on one
 * context's stream run in order. */
int esp_window_myers_batch_noalloc(esp_ctx *ctx, const esp_ref *ref, const esp_window_job *jobs,
                                   size_t njobs, const uint8_t *reads, uint32_t max_pattern,
                                   uint32_t max_text, int mode, int32_t *dist, uint64_t *pos,
                                   int32_t *status);

/* ---- placement walks ---------------------------------------------------------- */

/* Annotated junctions (JunctionLib, crates/map/src/gtf.rs:120-169): n
 * (contig, start, end) triples in any order, duplicates allowed (host
 * pointer).  The device keeps them as a hash set until esp_junctions_free;
 * the walk asks it what contains_any_strand (gtf.rs:165-169) asks of the
 * sorted library: is (contig, start, end) present. */
typedef struct esp_junctions esp_junctions;
int esp_junctions_upload(esp_ctx *ctx, size_t n, const uint32_t *triples, esp_junctions **out);
int esp_junctions_free(esp_ctx *ctx, esp_junctions *j);

/* JunctionLib::build (gtf.rs:84-114) with strands: n (contig, start, end,
 * minus) quads in any order, identical ones counted (host pointer).  Besides
 * contains_any_strand, the library then answers JunctionLib::support
 * (gtf.rs:142-149, the count of an exact stranded junction), which the
 * junction refinement below needs.  Freed by esp_junctions_free. */
int esp_junctions_upload_stranded(esp_ctx *ctx, size_t n, const uint32_t *quads, esp_junctions **out);

/* SpliceParams (splice.rs:21-53; defaults 20, 1000000, 4, 2, 1, 8, 6, 12). */
typedef struct {
    uint32_t min_intron;
    uint32_t max_intron;
    int32_t gt_ag_bonus;
    int32_t gc_ag_bonus;
    int32_t at_ac_bonus;
    int32_t noncanonical_penalty;
    int32_t known_bonus;
    uint32_t refine_radius;
} esp_splice_params;

/* RefinedJunction (gtf.rs:52-61).  signal: 0 GT-AG, 1 GC-AG, 2 AT-AC,
 * 3 non-canonical. */
typedef struct {
    int32_t found;    /* 1 Some, 0 None */
    uint32_t start;   /* intron [start, end), plus-strand coordinates */
    uint32_t end;
    uint32_t minus;   /* transcript strand */
    int32_t signal;
    uint32_t support; /* JunctionLib::support of the stranded junction */
} esp_refined_junction;

/* One refine_junction_read call (splice.rs:208-317): the read
 * reads[read_off .. + read_len) in alignment orientation, naive breakpoints
 * and the read positions of the two anchors.  With use_fallback set, a None
 * becomes the junction [fallback_start, fallback_end) on the query's strand
 * with its junction_signal and support (intron_chain.rs:475-486); found is
 * then 2. */
typedef struct {
    uint64_t read_off;
    uint32_t read_len;
    uint32_t contig;
    uint32_t naive_start;
    uint32_t naive_end;
    uint32_t q_a;
    uint32_t q_b;
    uint32_t minus;
    uint32_t use_fallback;
    uint32_t fallback_start;
    uint32_t fallback_end;
} esp_junction_query;

/* refine_junction_read per query (device pointers; sp is host).  junctions
 * must come from esp_junctions_upload_stranded (or be NULL: an empty
 * library).  A query on a contig out of range gets found = 0. */
int esp_refine_junction_batch(esp_ctx *ctx, const esp_ref *ref, const esp_junctions *junctions,
                              const esp_junction_query *queries, size_t n, const uint8_t *reads,
                              const esp_splice_params *sp, esp_refined_junction *out);

/* align_spliced (splice.rs:625-664) after prepare_spliced (:545-621) per
 * chain: segment_chain (:364-396) over the chain's anchors (qpos, rpos) at
 * anchors[2 * anc_off[k] .. 2 * anc_off[k + 1]) in (rpos, qpos) order with
 * seed length k; read_aware refinement of every adjacent segment pair, the
 * intron and support guards, the pseudo-reference with 30-base outer flanks,
 * extend_hint at offset 0 and the stitched genomic CIGAR.  Device pointers;
 * the chain's read is in alignment orientation. */
typedef struct {
    size_t nchain;
    const uint8_t *reads;
    const uint64_t *read_off;
    const uint32_t *read_len;
    const uint32_t *contig;
    const uint32_t *minus;     /* chain strand: 0 Plus, 1 Minus */
    const uint32_t *anchors;
    const uint64_t *anc_off;   /* nchain + 1 */
} esp_spliced_chain_batch;
typedef struct {
    int32_t status;        /* 1 aligned, 0 refused (prepare_spliced returned None),
                              -1 invalid (contig out of range), -2 stitched CIGAR
                              capacity too small (n_stitched = required) */
    int32_t refusal;       /* with status 0: 1 fewer than two segments, 2 a
                              refinement found nothing, 3 a refined intron
                              outside [min_intron, max_intron], 4 the splice
                              support guard, 5 an exon piece with end <= start */
    uint32_t n_segments;
    uint32_t genomic_start;
    uint32_t pseudo_len;
    uint32_t n_stitched;
} esp_spliced_chain_rec;
/* Outputs: ext_out as for esp_gotoh_batch in extend_hint mode (pseudo-
 * reference CIGARs; pair k needs esp_cigar_capacity(read_len[k]) words);
 * stitched/stitch_off the genomic CIGARs (pair k needs
 * esp_cigar_capacity(read_len[k]) + 2 * anchors words); junctions and
 * bounds: chain k's refined junctions at junctions[anc_off[k] + j] and its
 * exon bounds (pseudo_start, ref_start, ref_end) at bounds[3 * (anc_off[k] + s)]
 * (either may be NULL). */
int esp_align_spliced_batch(esp_ctx *ctx, const esp_ref *ref, const esp_junctions *junctions,
                            const esp_spliced_chain_batch *b, uint32_t k, const esp_splice_params *sp,
                            const esp_params *p, esp_batch_out *ext_out, uint32_t *stitched,
                            const uint64_t *stitch_off, esp_spliced_chain_rec *rec,
                            esp_refined_junction *junctions_out, uint32_t *bounds);

/* Aligner::reextend_with (align.rs:1170-1225) per genomic placement, with
 * one EditDir for the whole call (the mapper's candidate_dir; callers group
 * placements by it): a placement whose CIGAR holds a RefSkip is left alone;
 * otherwise the window [pos - (lead + 30), min(pos + ref_span + tail + 30,
 * contig length)) around its footprint and soft clips is extended with
 * extend_hint at hint (pos - wstart) - lead, and the result replaces the
 * placement when its score is positive and it aligns at least as many read
 * bases (M + I) as before, but no more than the read has. */
typedef struct {
    size_t nplace;
    const uint8_t *reads;      /* the query in the placement's orientation */
    const uint64_t *read_off;
    const uint32_t *read_len;
    const uint32_t *contig;
    const uint32_t *pos;
    const uint32_t *cigar;
    const uint64_t *cig_off;
    const uint32_t *cig_n;
} esp_reextend_batch;
typedef struct {
    int32_t status;        /* 1 replaced, 0 kept, 2 spliced (left alone), 3 empty
                              window, -1 contig out of range */
    uint32_t pos;          /* the placement's position after the call */
    int32_t score;         /* the new extension's score */
    uint32_t wstart;
    uint32_t wlen;
    uint32_t aligned;      /* M + I before */
    uint32_t new_aligned;  /* M + I of the new extension */
} esp_reextend_rec;
/* cigar_out/cig_out_off/cig_out_n: the new extension's CIGAR for every
 * placement whose DP ran (placement k needs esp_cigar_capacity(read_len[k])
 * words; cig_out_n = 0 otherwise).  Device pointers. */
int esp_reextend_batch_run(esp_ctx *ctx, const esp_ref *ref, const esp_reextend_batch *b, const esp_params *p,
                           int edit_dir, esp_reextend_rec *rec, uint32_t *cigar_out, const uint64_t *cig_out_off,
                           uint32_t *cig_out_n);

/* The intron-chain fast path, try_intron_chain_placement
 * (crates/map/src/intron_chain.rs:392-538), per read: anchors extended into
 * maximal exact matches, clustered into fragments by implied read start,
 * the best fragment pair, EA-Myers refinement of the donor end and acceptor
 * start, split-read verification, the read-aware junction (or the pair's own
 * junction when refinement finds none), the two-exon pseudo-reference
 * extension and its acceptance (positive score, read-conserving stitched
 * CIGAR holding a RefSkip).  Anchors: 4 u32 each (qpos, rpos, contig, strand
 * 0 Plus / 1 Minus) at anchors[4 * anc_off[k] .. 4 * anc_off[k + 1]), the
 * read's whole anchor pool in any order.  Device pointers; max_read_len
 * (host) bounds read_len over the batch. */
typedef struct {
    size_t nread;
    const uint8_t *reads;      /* alignment orientation */
    const uint64_t *read_off;
    const uint32_t *read_len;
    const uint32_t *anchors;
    const uint64_t *anc_off;   /* nread + 1 */
    uint32_t max_read_len;
} esp_intron_chain_batch;
/* IntronParams (intron_chain.rs:20-42; defaults 15, 30, 50, 105000). */
typedef struct {
    uint32_t min_frag_cov;
    uint32_t refine_window;
    uint32_t refine_pattern_len;
    uint32_t intron_max;
} esp_intron_params;
typedef struct {
    int32_t status;          /* 1 placed, 0 None, -1 read longer than max_read_len */
    int32_t reason;          /* with status 0: 1 no anchors or empty read, 2 no fragment
                                of min_frag_cov, 3 a dominant fragment (unspliced read),
                                4 no fragment pair, 5 contig out of range, 6 fewer than
                                two anchors, 7 intron length, 8 split verification,
                                9 junction, 10 exon bounds, 11 extension score, 12 CIGAR
                                not read-conserving, 13 no RefSkip */
    uint32_t genomic_start;
    int32_t score;
    uint32_t left_start;     /* exon bounds [left_start, jstart) ++ [jend, right_end) */
    uint32_t jstart;
    uint32_t jend;
    uint32_t right_end;
    int32_t signal;          /* the junction's signal (0 GT-AG .. 3 non-canonical) */
    uint32_t support;
    uint32_t minus;
    uint32_t n_stitched;
    int64_t donor_end;       /* the EA-Myers refinements (reasons >= 6) */
    int64_t acceptor_start;
    int32_t vd;              /* split verification distances */
    int32_t va;
} esp_intron_chain_rec;
/* ext_out: pseudo-reference extensions (pair k needs esp_cigar_capacity(m)
 * words); stitched/stitch_off: genomic CIGARs (esp_cigar_capacity(m) + 4). */
int esp_intron_chain_batch_run(esp_ctx *ctx, const esp_ref *ref, const esp_junctions *junctions,
                               const esp_intron_chain_batch *b, uint32_t k, const esp_intron_params *ip,
                               const esp_params *p, esp_batch_out *ext_out, uint32_t *stitched,
                               const uint64_t *stitch_off, esp_intron_chain_rec *rec);

/* The SpliceParams fields rescore_with reads (splice.rs:21-52; defaults 6,
 * 4, 2, 1, 8) and the seed k-mer length of exact_kmer_support (SeedParams
 * k, 15 in the reference's configurations). */
typedef struct {
    int32_t known_bonus;
    int32_t gt_ag_bonus;
    int32_t gc_ag_bonus;
    int32_t at_ac_bonus;
    int32_t noncanonical_penalty;
    uint32_t seed_k;
} esp_walk_params;

/* StrictScore (align.rs:188-211). */
typedef struct {
    int32_t score;
    uint32_t aligned, matches, mm, ea_ag, ea_tc, ins, del, clipped, n_splice;
    int32_t splice_score;
} esp_strict_score;

/* recount_mm_ea_in's counts (align.rs:2473-2516) and exact_kmer_support's
 * k-mer count (align.rs:1231-1288). */
typedef struct {
    uint32_t mm, ea_ag, ea_tc, kmers;
} esp_walk_counts;

/* Placements to walk: placement k is contig[k] at pos[k] with the BAM CIGAR
 * cigar[cig_off[k] .. cig_off[k] + cig_n[k]) over the query bytes
 * reads[read_off[k] .. + read_len[k]) in the alignment's frame (reverse
 * complemented for minus placements, as the callers pass them).  dir[k] is
 * the EditDir exempted (0 Off, 1 Ag, 2 Tc): rescore_with's `dir`, and the
 * placement's ea_dir for the k-mer scan.  Device pointers. */
typedef struct {
    size_t nplace;
    const uint8_t *reads;
    const uint64_t *read_off;
    const uint32_t *read_len;
    const uint32_t *contig;
    const uint32_t *pos;
    const uint32_t *cigar;
    const uint64_t *cig_off;
    const uint32_t *cig_n;
    const uint8_t *dir;
} esp_walk_batch;

/* Per placement: rescore_with (align.rs:1060-1163) into strict, and
 * recount_mm_ea_in + exact_kmer_support into counts (either may be NULL).
 * Reference bytes past a contig end and query bytes past the read read as
 * N, as in the reference.  junctions may be NULL (an empty library).  A
 * placement on a contig out of range gets all-zero records and
 * status[k] = ESP_E_INVAL (status may be NULL).  Device pointers. */
int esp_walk_placements(esp_ctx *ctx, const esp_ref *ref, const esp_junctions *junctions,
                        const esp_walk_batch *b, const esp_params *p, const esp_walk_params *wp,
                        esp_strict_score *strict, esp_walk_counts *counts, int32_t *status);

/* The transcriptome engine's EA-Myers (crates/engine/src/myers.rs): edit
 * distance with identity and the editing pairs (pattern G / text A, pattern
 * C / text T) free, N matching nothing, case folded, indels 1.  Per pair,
 * pattern = read, text = window of the batch:
 *   mode 0  myers::infix / long::infix (text overhang free on both sides)
 *   mode 1  myers::global / long::global (both ends forced)
 * Pattern lengths (meta->read_len, host) must be 1..256, as the reference
 * asserts; otherwise ESP_E_INVAL and nothing runs. */
int esp_engine_myers_batch(esp_ctx *ctx, const esp_batch *b, const esp_host_meta *meta, int mode,
                           int32_t *dist);

/* verify_dist (crates/engine/src/lib.rs:851-876) per candidate: the read in
 * the candidate's orientation, the candidate's transcript sequence at
 * tx[tx_off[k] .. + tx_len[k]), anchor = ext.tx_lo - ext.read_lo; the window
 * is [max(anchor - flank, 0), min(max(anchor + m + flank, 0), tx_len)).
 * status[k] 1: dist[k] is myers infix distance over the window; 0: empty
 * window (the reference's None).  Device pointers; host_read_len as above. */
typedef struct {
    size_t n;
    const uint8_t *reads;
    const uint64_t *read_off;
    const uint32_t *read_len;
    const uint8_t *tx;
    const uint64_t *tx_off;
    const uint32_t *tx_len;
    const int64_t *anchor;
} esp_engine_verify;
int esp_engine_verify_batch(esp_ctx *ctx, const esp_engine_verify *v, const uint32_t *host_read_len,
                            uint32_t flank, int32_t *dist, int32_t *status);

/* The Kadane fast lane, Aligner::try_fast_lane (align.rs:2274-2401), over a
 * batch of chains: chain k is contig[k] with its anchors (qpos, rpos) at
 * anchors[2 * anc_off[k] .. 2 * (anc_off[k] + anc_n[k])) in the chain's
 * (rpos, qpos) order, and its read reads[read_off[k] .. + read_len[k]) in the
 * chain's orientation (reverse complemented for minus chains).  Device
 * pointers. */
typedef struct {
    size_t nchain;
    const uint8_t *reads;
    const uint64_t *read_off;
    const uint32_t *read_len;
    const uint32_t *contig;
    const uint32_t *anchors;
    const uint64_t *anc_off;
    const uint32_t *anc_n;
} esp_fast_batch;
/* status 1: the lane accepts with Extension (score, ref_start relative to
 * wstart, read_start, read_end) and CIGAR cigar[3k .. 3k + n_ops); pos =
 * wstart + ref_start is the placement's genomic start.  status 0: the lane
 * refuses (the caller goes on to the EA lane / DP).  status -1: contig out of
 * range. */
typedef struct {
    int32_t status;
    int32_t score;
    uint32_t wstart, ref_start, read_start, read_end, pos, n_ops;
} esp_fast_rec;
/* k, w: SeedParams (15, 5 in the reference's configurations); p supplies
 * match_score, mismatch and edit_dir.  cigar: 3 words per chain. */
int esp_fast_lane_batch(esp_ctx *ctx, const esp_ref *ref, const esp_fast_batch *b, const esp_params *p,
                        uint32_t k, uint32_t w, esp_fast_rec *out, uint32_t *cigar);

/* intron_length_penalty (splice.rs:119-121) as the walk kernel evaluates
 * it: from a threshold table built on the host with the reference's f64
 * formula, so that device and host agree exactly. */
int32_t esp_intron_length_penalty(uint32_t len);

/* ---- serving loop ---------------------------------------------------------- */

/* A request loop over the resident reference with continuous batching: the
 * replacement for the per-read rayon pipeline (pipeline.rs:393-404) when the
 * callers are many independent clients.  Requests queue in arrival order;
 * each of `workers` worker threads owns a context (stream + scratch) and
 * takes the next batch of queued requests that fits its pair cap and the
 * device-memory budget, so batches overlap on the device and a request that
 * arrives while others run joins the next free slot.  Results are identical
 * to running each request alone (every pair is computed independently). */
typedef struct esp_server esp_server;
typedef struct {
    esp_params params;          /* ExtendParams of every request */
    uint64_t mem_budget;        /* device bytes for in-flight request data */
    uint32_t workers;           /* batches in flight (contexts), 1..8 */
    uint32_t max_batch_pairs;   /* pairs per batch; a larger request runs alone */
    uint32_t max_pending;       /* queued requests before submit returns ESP_E_BUSY */
    uint32_t max_read_len;      /* longest read accepted (<= 16000; rescue reads <= 256) */
} esp_server_config;

/* kind ESP_REQ_EXTEND: per pair, extend_hint(read, contig[win_start, win_end),
 *   hint) (the window decoded from the resident reference, extend.rs:392);
 * kind ESP_REQ_RESCUE: per pair, mate_window_placement(read, contig, mate_pos)
 *   (align.rs:851).
 * Inputs are host memory and are copied before submit returns. */
enum { ESP_REQ_EXTEND = 0, ESP_REQ_RESCUE = 1 };
typedef struct {
    uint32_t kind;
    uint32_t npairs;
    const uint8_t *reads;       /* concatenated */
    const uint32_t *read_len;
    const uint32_t *contig;
    const uint32_t *win_start;  /* EXTEND */
    const uint32_t *win_end;    /* EXTEND */
    const int64_t *hint;        /* EXTEND */
    const uint32_t *mate_pos;   /* RESCUE */
} esp_request;

/* Host output buffers, written before the request completes; they must stay
 * valid until esp_server_wait reports it done.  cigar holds, per pair k,
 * esp_cigar_capacity(read_len[k]) words in order (pair k starts at the sum
 * of the capacities before it). */
typedef struct {
    esp_ext *ext;           /* EXTEND: npairs records */
    esp_mate_rec *mate;     /* RESCUE: npairs records */
    uint32_t *cigar;
    uint32_t *cig_n;
} esp_request_out;

typedef struct {
    int32_t status;         /* ESP_OK, or the error the batch hit */
    int32_t done;
    uint64_t submit_ns;     /* steady-clock times */
    uint64_t done_ns;
    uint64_t batch;         /* batch sequence number that carried it */
} esp_request_info;

typedef struct {
    uint64_t submitted, completed, batches, pairs;
    uint64_t inflight_bytes, peak_inflight_bytes;
    uint32_t pending, max_batch_requests;
    /* summed over batches (ns): host staging, extension phase (launch to
     * results on the host), rescue phase, scattering results to callers */
    uint64_t stage_ns, extend_ns, rescue_ns, scatter_ns;
} esp_server_stats;

int esp_server_create(int device, const esp_ref *ref, const esp_server_config *cfg,
                      esp_server **out);
int esp_server_destroy(esp_server *s);
/* ESP_E_BUSY: queue full; ESP_E_NOMEM: the request alone exceeds the budget;
 * ESP_E_INVAL: malformed request (lengths, contig, kind). */
int esp_server_submit(esp_server *s, const esp_request *req, const esp_request_out *out,
                      uint64_t *ticket);
/* Waits up to timeout_us (< 0: forever) for the ticket; info->done tells.
 * A completed ticket is released by this call. */
int esp_server_wait(esp_server *s, uint64_t ticket, int64_t timeout_us, esp_request_info *info);
int esp_server_stats_get(esp_server *s, esp_server_stats *out);
/* Steady-clock now, the clock of submit_ns / done_ns. */
uint64_t esp_now_ns(void);

#ifdef __cplusplus
}
#endif

#endif /* ESP_ALIGN_H */
