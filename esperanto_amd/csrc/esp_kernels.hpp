// esp_kernels.hpp -- launch interface between the host runtime (esp_host.cpp)
// and the device kernels (esp_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/esp_align.h"
#include "esp_device.hpp"

namespace espk {

// Per-launch description of a Gotoh batch.  Sequences are ASCII bytes.
struct GotohArgs {
    uint32_t npairs;
    const uint32_t *order;  // visiting order (length-bucketed), or nullptr for identity
    const uint32_t *range;  // device {start, count} into order for this launch, or nullptr
    const uint8_t *reads;
    const uint64_t *read_off;
    const uint32_t *read_len;
    const uint8_t *wins;
    const uint64_t *win_off;
    const uint32_t *win_len;
    const int64_t *offset;  // band offset or extend_hint hint offset
    const int64_t *half;    // band half-width (band mode)
    int32_t mode;           // esp_mode
    espd::GParams P;
    esp_ext *ext;
    esp_touch *touch;
    uint32_t *cigar;
    const uint64_t *cig_off;
    uint32_t *cig_n;
    int32_t *aux;
    // scratch
    uint32_t *counter;  // work-queue head, zeroed by the launcher
    int32_t *rows;      // 5 arrays of maxm x NL int32, lane-interleaved
    uint32_t *trace;    // maxw x NL trace words, lane-interleaved
    uint32_t maxm;
    uint32_t maxw;
    uint32_t nl;        // lanes in the persistent grid
    // Band kernel substitution rows: for read code 0..4, 5-bit signed scores
    // of window codes 0..5 packed at bit 5*code.
    uint32_t lut[5];
    // The same rows with -open_ext folded in (score - oe, still 5-bit
    // signed fields); lut_folded says they fit.  The packed wide-band
    // kernels then form D = U + (score - oe) from the packed U in one SDWA add.
    uint32_t lutf[5];
    int32_t lut_folded;
    // Full-matrix kernel substitution columns: for window code 0..5, 5-bit
    // signed scores of read codes 0..4 at bit 5*code, and -16 at bit 25 (the
    // code of rows past the read).
    uint32_t flut[6];
    // flut with -open_ext folded in (as lutf for the band kernel).
    uint32_t flutf[6];
    int32_t flut_folded;
    // extend_hint escalation (band kernel in ESP_MODE_HINT): pairs whose band
    // attempt must be redone on the full matrix are appended to esc_list
    // through the counter esc_range[1].
    uint32_t *esc_range;
    uint32_t *esc_list;
    // With esc_info set, an escalating pair p also records its band attempt's
    // best score and the diagonal bj - bi of its best cell at esc_info[2p],
    // esc_info[2p + 1]: the pruned full-matrix pass reads them.
    int32_t *esc_info;
    // Pruned full-matrix pass (k_gotoh_full with pruning): prune_s is the
    // largest substitution score; a pair whose check fails is appended to
    // fb_list through the counter fb_range[1] instead of being traced back.
    int32_t prune_s;
    int32_t prune_half;  // stripe rows each side of a pair's diagonal
    int32_t prune_pad;   // head rows past the last row a 0 start can reach L from
    uint32_t *fb_list;
    uint32_t *fb_range;
    // Pairs per work unit of the full-matrix kernel: 64 (0 means 64) or 1
    // (one pair a wave, for short lists whose pairs share no stripe).
    uint32_t unit;
    // One-wave-per-pair list kernel: entries from list_cap on are passed to
    // fb_list instead (a long list runs faster in the lane kernel); its LDS
    // holds windows up to wave_lds_n bases and wave_lds_tr trace words.
    uint32_t list_cap;
    uint32_t wave_lds_n, wave_lds_tr;
    // Cell census of the full-matrix kernel (nullptr: off).  census[0] +=
    // cells computed inside the pairs' matrices (rows 1..m, columns 1..n),
    // census[1] += lane-cells issued (64 lanes x 8-row groups x columns,
    // padding included), census[2] += m x n of each pair when census_first
    // is set (the unpruned kernel run alone; with pruning the escalation sort
    // counts the matrices).
    unsigned long long *census;
    int32_t census_first;
    // Always 0 (value-initialised).  Read from the argument block, the band
    // kernel's floor at 0 is an operand the compiler cannot see through, so
    // F's floor and H's max stay two v_max3.
    int32_t zero;
};

struct MyersArgs {
    uint32_t npairs;
    const uint32_t *order;
    const uint32_t *range;  // device {start, count} into order, or nullptr
    const uint8_t *reads;
    const uint64_t *read_off;
    const uint32_t *read_len;
    const uint8_t *texts;
    const uint64_t *text_off;
    const uint32_t *text_len;
    int32_t *dist;
    uint64_t *end;    // nullptr: no forward scan unless start is nullptr too (infix)
    uint64_t *start;  // nullptr: no reversed scan
    uint32_t *counter;
    int32_t global;   // 1: global distance (both ends forced; engine/src/myers.rs:149-168)
    // With ovf_list set, the 32-bit kernel appends the pairs whose pattern is
    // longer than its words hold (up to 256 bases) to ovf_list through the
    // counter ovf_range[1], for an 8-word launch over that list.
    uint32_t *ovf_list;
    uint32_t *ovf_range;
};

// verify_dist windows (engine/src/lib.rs:851-876): per candidate the text is
// its transcript's [wlo, whi), wlo = max(anchor - flank, 0), whi =
// min(max(anchor + m + flank, 0), tx_len); an empty window is status 0.
struct EnginePrepArgs {
    uint32_t n;
    const uint32_t *read_len;
    const uint64_t *tx_off;
    const uint32_t *tx_len;
    const int64_t *anchor;
    uint32_t flank;
    uint64_t *text_off;
    uint32_t *text_len;
    int32_t *status;
};
hipError_t launch_engine_prep(const EnginePrepArgs &a, hipStream_t s);

// Device-side batch planner: pairs are classified (one class per band
// kernel specialisation, plus the general strip class) and counting-sorted
// by (class, read length descending) into `order`.
constexpr int PLAN_NCLASS = 6;  // halves 16, 30, 32, 48, 64, then the strip class
constexpr int PLAN_STRIP = 5;
constexpr int PLAN_NBIN = 512;
struct PlanSummary {
    uint32_t range[PLAN_NCLASS][2];  // {start, count} per class
    uint32_t strip_maxm;
    uint32_t bad;                   // pairs outside the length limits
    unsigned long long strip_maxw;  // largest strip-layout trace of the strip class (words)
    uint32_t esc_maxm;              // extend_hint band pairs: largest read,
    uint32_t esc_maxn;              // largest window and
    unsigned long long esc_maxw;    // largest full-matrix strip trace (words)
    uint32_t strip_work;            // strip-class pairs with a read and a window
    uint32_t strip_maxn;            // their largest window
};
struct PlanDev {
    PlanSummary sum;  // first: the host reads back only this part
    uint32_t esc_range[2];  // {start in order (= npairs), escalated count}
    // strip class run one wave per pair: {0, pairs passed on to the lane
    // kernel}, and that kernel's queue head
    uint32_t wave_ovf[2], wave_ovf_counter;
    uint32_t counters[PLAN_NCLASS + 2];
    uint32_t hist[PLAN_NCLASS * PLAN_NBIN];
    uint32_t cursor[PLAN_NCLASS * PLAN_NBIN];
};
struct PlanArgs {
    uint32_t npairs;
    const uint32_t *read_len;
    const uint32_t *win_len;
    const int64_t *offset;
    const int64_t *half;
    int32_t band_ok;    // 1 when mode is band and the scoring fits the band kernel
    int32_t hint_mode;  // 1 when mode is extend_hint and the scoring fits the band kernel
    // 1 when bands clamped on the left (0 <= offset < half) may take the band
    // kernel too: the mismatch score is <= 0, so the 'N' columns that stand
    // for the columns left of the window keep H = 0 (see the band kernel)
    int32_t left_clamp_ok;
    uint32_t max_read;
    uint32_t max_win;
    // strip-class pairs whose strip trace exceeds this many words are counted
    // as bad (a captured graph's strip scratch); ~0 when the call sizes its own
    unsigned long long max_strip_w;
    PlanDev *plan;
    uint32_t *order;
};
hipError_t launch_plan(const PlanArgs &a, int ncu, hipStream_t s);
int plan_class_half(int c);

// Strip width of the general kernel (columns per register tile).
constexpr int STRIP_T = 32;
constexpr int GOTOH_BLOCK = 64;
constexpr int MYERS_BLOCK = 64;

// Trace words one pair needs in the strip layout.
__host__ __device__ inline uint64_t strip_trace_words(uint64_t m, uint64_t n) {
    return ((n + STRIP_T - 1) / STRIP_T) * m * (STRIP_T / 8);
}

// One pair per wave (esp_wave.hip): the longest window it takes, and the
// trace words it keeps in LDS (more go to the wave's global scratch).
constexpr int WAVE_NMAX = 4096;
constexpr uint32_t WAVE_TR_LDS = 4096;
// Trace words of one pair in the wave kernel: rows times the widest row
// window in words; FULL and HINT size for the full matrix (an extend_hint
// band attempt may escalate), BAND for its clamped half-width.
__host__ __device__ inline uint64_t wave_trace_words(uint64_t m, uint64_t n, int mode, int64_t half) {
    uint64_t w = n;
    if (mode == ESP_MODE_BAND) {
        const int64_t h = half < 0 ? 0 : (half > (int64_t)n ? (int64_t)n : half);
        w = (uint64_t)(2 * h + 1) < n ? (uint64_t)(2 * h + 1) : n;
    }
    return m * ((w + 7) / 8);
}
hipError_t launch_gotoh_wave(const GotohArgs &a, int blocks, hipStream_t s);
// The same kernel over a device list of full-matrix reruns (order, range),
// one wave per pair; maxw trace words per block; windows over WAVE_NMAX and
// traces over maxw go to fb_list / fb_range.
// LDS of the list kernel: windows up to ln bases (a multiple of 16, at most
// WAVE_NMAX) and tr trace words kept in LDS (more go to the block's global
// trace); blocks per CU at that size.
hipError_t launch_gotoh_wave_list(const GotohArgs &a, int blocks, hipStream_t s);
size_t wave_list_lds(uint32_t ln, uint32_t tr);
int gotoh_wave_blocks_per_cu(uint32_t ln, uint32_t tr);

hipError_t launch_gotoh_strip(const GotohArgs &a, int blocks, hipStream_t s);

// Diagonal-frame band kernel for interior bands (every row window inside
// [1, n]); half must be one of band_halves().
constexpr int BAND_BLOCK = 64;
// Longest read the band kernel takes at each half (longer ones run in the
// strip kernel): 287 at half 16, 259 at 30, 255 at 32, 48 and 64.
__host__ __device__ constexpr int band_max_read(int half) {
    return (half <= 32 ? 320 : (half <= 48 ? 352 : 384)) - (2 * half + 1);
}
// LDS ring of window-code rows per wave (64 bytes a row, plus 2*half mirror
// rows and the LUT row): sized so that LDS allows as many waves per CU as
// the register arrays do (16 at half 16, 12 at 30-48, 8 at 64), whatever
// the read length.
__host__ __device__ constexpr int band_ring(int half) {
    return half <= 16 ? 127 : (half <= 32 ? 148 : (half <= 48 ? 116 : 191));
}
bool band_half_supported(int half);
hipError_t launch_gotoh_band(const GotohArgs &a, int half, int blocks, hipStream_t s);
int gotoh_band_blocks_per_cu(int half);
__host__ __device__ inline uint64_t band_trace_words(uint64_t m, uint64_t half) { return m * ((2 * half + 1 + 7) / 8); }
// Full-matrix kernel for extend_hint's second attempt (reads up to 264
// bases): rows = full_rows(max read length) picks the register tile.  With
// prune set it computes only the head rows and a diagonal stripe of each
// wave's pairs (esc_info, prune_s, fb_list / fb_range must be set).
int full_rows(uint32_t maxm);
constexpr int FULL_PRUNE_MAX_ROWS = 192;  // largest row tile with a pruned pass
hipError_t launch_gotoh_full(const GotohArgs &a, int rows, bool prune, int blocks, hipStream_t s);
// The pruned pass over a register window (head and stripe groups only) for
// the escalation sort's first head class; rows 64..192.
hipError_t launch_gotoh_pruned(const GotohArgs &a, int rows, int blocks, hipStream_t s);
// The register-window pass runs when the batch's longest escalated window
// has at least this many bases.
constexpr uint64_t PRUNE_WINDOW_MIN_N = 512;
int gotoh_pruned_blocks_per_cu(int rows);
int gotoh_full_blocks_per_cu(int rows);

// Ordering of the escalation list for the pruned pass: a counting sort by
// (head-row class, diagonal of the band attempt's best cell), so that the 64
// pairs of a wave share one stripe; pairs the pruned pass cannot help go to
// the fallback list at once.  Every count lives in EscDev (zeroed before).
// The pruned pass takes the sorted list with (ESC_HALF_A, ESC_PAD_A): a
// stripe row costs the read's span of columns, a head row the whole window.
// Its failures go to the one-pair-per-wave kernel (launch_gotoh_wave_list).
constexpr int ESC_HALF_A = 32, ESC_PAD_A = 16;
constexpr int ESC_NDIAG = 1024;    // diagonal buckets per class (four diagonals each)
constexpr int ESC_NCLASS = 3;      // head groups <= 3, <= 6, <= 12
constexpr int ESC_NB = ESC_NCLASS * ESC_NDIAG;
struct EscDev {
    uint32_t prange[2];   // {0, pairs in the sorted list}
    uint32_t crange0[2];  // {0, pairs of the first head class}: the register-window pass
    uint32_t crange1[2];  // {first class's count, the other classes' pairs}: the whole-tile pass
    uint32_t c0counter;   // work-queue head of the register-window pass
    uint32_t fbrange[2];  // {0, pairs the sort or the pruned pass left for the one-wave rerun}
    uint32_t fb2range[2]; // {0, pairs for the unpruned kernel}
    uint32_t pcounter, fbcounter, fb2counter;  // work-queue heads (fbcounter unused)
    uint32_t hist[ESC_NB];
    uint32_t cursor[ESC_NB];
};
struct EscSortArgs {
    const uint32_t *list;   // escalated pairs
    const uint32_t *count;  // device count of list
    const uint32_t *read_len;
    const int32_t *esc_info;
    int32_t s;              // largest substitution score
    int32_t pad;            // head pad of the first pruned pass
    EscDev *esc;
    uint32_t *plist;
    uint32_t *fblist;   // pairs the pruned pass cannot help go straight to its failure list
    // census (nullptr: off): census[2] += m x n of every listed pair
    const uint32_t *win_len;
    unsigned long long *census;
};
hipError_t launch_esc_sort(const EscSortArgs &a, uint32_t bound, hipStream_t s);
// Head groups (8 rows each) a pruned pass computes for a pair: every row
// from which a path starting at 0 can still reach L, plus pad rows.
__host__ __device__ inline int32_t esc_head_groups(int32_t m, int32_t L, int32_t s, int32_t pad) {
    const int32_t r = m - (L + s - 1) / s;
    const int32_t need = r > 0 ? r : 0;
    return (need + pad + 7) >> 3;
}
// maxm: the longest pattern of the launch (1..256 after clamping).
hipError_t launch_myers(const MyersArgs &a, int maxm, int blocks, hipStream_t s);
hipError_t launch_zero_u32(uint32_t *p, size_t n, hipStream_t s);
hipError_t launch_len_max(const uint32_t *rl, const uint32_t *wl, size_t n, uint32_t *out, int ncu,
                          hipStream_t s);
hipError_t launch_copy(void *dst, const void *src, size_t bytes, hipStream_t s);
int gotoh_strip_blocks_per_cu(int edit);
int myers_blocks_per_cu(int maxm, bool global);

// ---- resident reference and window pipelines (esp_window.hip) ----------

// Device view of the resident 2-bit reference.  Contig k's packed bytes
// start at packed + pk_off[k] (16-byte aligned) and are followed by at least
// 16 zero bytes, so a 12-byte window read at any in-range base stays inside
// the contig's region.
struct RefView {
    const uint8_t *packed;
    const uint64_t *pk_off;
    const uint32_t *len;
    const uint32_t *iv;      // 2 u32 per N interval
    const uint64_t *iv_off;  // ncontig + 1
    uint32_t ncontig;
};

// Placement walks (rescore_with, recount_mm_ea_in, exact_kmer_support).
// Slot hash of a junction triple in the library's open-addressing set.
__host__ __device__ inline uint32_t junction_hash(uint32_t c, uint32_t s, uint32_t e) {
    uint32_t h = c * 0x9E3779B1u ^ s * 0x85EBCA77u ^ e * 0xC2B2AE3Du;
    h ^= h >> 15;
    h *= 0x2C1B3C6Du;
    h ^= h >> 12;
    return h;
}
constexpr int PEN_UP = 64, PEN_DOWN = 32;
struct WalkArgs {
    RefView R;
    const uint4 *jn;     // junction hash set slots (x contig, y start, z end, w used); NULL if none
    uint32_t jmask;      // slots - 1
    uint32_t n;
    const uint8_t *reads;
    const uint64_t *read_off;
    const uint32_t *read_len;
    const uint32_t *contig;
    const uint32_t *pos;
    const uint32_t *cigar;
    const uint64_t *cig_off;
    const uint32_t *cig_n;
    const uint8_t *dir;
    int32_t match, mismatch, gap_open, gap_ext;
    int32_t known, gtag, gcag, atac, noncanon;
    uint32_t k;
    // intron_length_penalty thresholds: for len >= 1000 the penalty is the
    // number of up[v] <= len; for len < 1000 the number of down[v] >= len.
    uint32_t up[PEN_UP];
    uint32_t down[PEN_DOWN];
    int32_t nup, ndown;
    esp_strict_score *strict;
    esp_walk_counts *counts;
    int32_t *status;
};
hipError_t launch_walk(const WalkArgs &a, int blocks, hipStream_t s);
int walk_blocks_per_cu();
int fast_lane_blocks_per_cu();

// try_fast_lane over chains (anchors as (qpos, rpos) pairs).
struct FastArgs {
    RefView R;
    uint32_t n;
    const uint8_t *reads;
    const uint64_t *read_off;
    const uint32_t *read_len;
    const uint32_t *contig;
    const uint32_t *anchors;
    const uint64_t *anc_off;
    const uint32_t *anc_n;
    uint32_t k, w;
    int32_t match, mismatch, edit_dir;
    esp_fast_rec *out;
    uint32_t *cigar;  // 3 words per chain
};
hipError_t launch_fast_lane(const FastArgs &a, int blocks, hipStream_t s);

struct GatherArgs {
    RefView R;
    const esp_gather_job *jobs;
    uint32_t njobs;
    uint8_t *dst;
    int32_t *status;
};

struct MateArgs {
    uint32_t npairs;
    RefView R;
    const uint8_t *reads;
    const uint64_t *read_off;
    const uint32_t *read_len;
    const uint32_t *mate_contig;
    const uint32_t *mate_pos;
    uint32_t max_read;
    uint32_t qstride;  // query slot per pair: read at 0, reverse complement at qstride / 2
    uint64_t wstride;  // window slot per pair
    uint8_t *qarena;
    uint8_t *warena;
    // per-pair scratch
    uint64_t *q_off_plus, *q_off_minus, *w_off, *s_plus, *s_minus, *g_read_off;
    int64_t *g_offset;
    uint32_t *q_len, *w_len, *wstart, *g_read_len;
    int32_t *state, *d_plus, *d_minus, *sel, *dist;
    esp_ext *ext;
    // outputs
    esp_mate_rec *out;
    uint32_t *cig_n;
};

struct SpliceArgs {
    uint32_t npairs;
    RefView R;
    const uint32_t *read_len;
    const uint32_t *contig;
    const uint32_t *exons;
    const uint64_t *exon_off;
    const uint64_t *pseudo_off;
    uint64_t n_exons;       // exon_off[npairs]: exon indices must stay below it
    uint64_t pseudo_bytes;  // pseudo_off[npairs]: the arena size
    uint8_t *parena;
    uint32_t *bounds;  // 3 per exon: pseudo_start, ref_start, ref_end
    int32_t *state;
    uint32_t *g_read_len;
    uint32_t *p_len;
    const esp_ext *ext;
    const uint32_t *cigar;
    const uint64_t *cig_off;
    const uint32_t *cig_n;
    uint32_t *stitched;
    const uint64_t *stitch_off;
    esp_splice_rec *rec;
};

struct UnspArgs {
    uint32_t npairs;
    RefView R;
    const uint8_t *reads;
    const uint64_t *read_off;
    const uint32_t *read_len;
    const uint32_t *contig, *minus, *rs, *re, *qs, *qe, *allow_dp;
    uint32_t max_read;
    uint32_t qstride;
    uint64_t wstride;
    uint8_t *qarena;
    uint8_t *warena;
    espd::GParams P;
    // per-pair scratch
    uint64_t *q_off, *w_off, *s_start, *g_read_off;
    int64_t *g_hint;
    uint32_t *q_len, *w_len, *wstart, *g_read_len, *ea_bi, *ea_bj;
    int32_t *state, *dist, *ea_sum;
    esp_ext *ext;
    // outputs
    esp_unspliced_rec *out;
    uint32_t *cigar;
    const uint64_t *cig_off;
    uint32_t *cig_n;
};

struct WinMyersArgs {
    uint32_t njobs;
    RefView R;
    const esp_window_job *jobs;
    uint32_t max_pattern;
    uint32_t max_text;
    uint64_t tstride;
    uint8_t *tarena;
    uint64_t *p_off, *t_off;
    uint32_t *p_len, *t_len;
    int32_t *status;
};

// ---- junction refinement, chain preparation, re-extension (esp_junction.hip)

// The stranded junction library: hash slots (x contig, y start, z end, w
// used) and per slot the plus / minus support counts; jn == nullptr is the
// empty library.
struct JunctionView {
    const uint4 *jn;
    const uint2 *cnt;
    uint32_t jmask;
};
struct SpliceCfgDev {
    uint32_t min_intron, max_intron;
    int32_t gt_ag, gc_ag, at_ac, noncanon, known;
};
struct RefineArgs {
    RefView R;
    JunctionView J;
    SpliceCfgDev sp;
    uint32_t n;
    const esp_junction_query *q;
    const uint8_t *reads;
    esp_refined_junction *out;
};
hipError_t launch_refine(const RefineArgs &a, hipStream_t s);

struct ChainPrepArgs {
    RefView R;
    JunctionView J;
    SpliceCfgDev sp;
    uint32_t npairs, k;
    const uint8_t *reads;
    const uint64_t *read_off;
    const uint32_t *read_len;
    const uint32_t *contig, *minus, *anchors;
    const uint64_t *anc_off;
    uint32_t *segs;                // 4 per anchor slot: q_start, q_end, r_start, r_end
    esp_refined_junction *jrec;    // per anchor slot
    uint32_t *bounds;              // 3 per anchor slot: pseudo_start, ref_start, ref_end
    int32_t *state;                // 1 prepared, 0 refused, -1 invalid
    int32_t *refusal;
    uint32_t *nseg, *nex;          // segments; exons handed on (0 unless prepared)
    uint64_t *plen;                // pseudo-reference length (0 unless prepared)
    // after the scans
    const uint64_t *exon_off, *pseudo_off;
    uint32_t *exons;               // 2 per exon: ref_start, ref_end
    int64_t *hint;
    const esp_splice_rec *srec;    // the spliced pipeline's records
    esp_spliced_chain_rec *rec;
};
hipError_t launch_chain_prep(const ChainPrepArgs &a, hipStream_t s);
hipError_t launch_chain_exons(const ChainPrepArgs &a, hipStream_t s);
hipError_t launch_chain_rec(const ChainPrepArgs &a, hipStream_t s);

struct ReextArgs {
    RefView R;
    uint32_t n;
    const uint32_t *read_len, *contig, *pos, *cigar, *cig_n;
    const uint64_t *cig_off;
    int32_t *state;       // 1 runs the DP, else the final status
    uint32_t *g_read_len, *wlen;
    int64_t *hint;
    const uint64_t *win_off;
    uint8_t *arena;
    const esp_ext *ext;
    const uint32_t *cig_out;
    const uint64_t *cig_out_off;
    uint32_t *cig_out_n;
    esp_reextend_rec *rec;
};
hipError_t launch_reext_prep(const ReextArgs &a, hipStream_t s);
hipError_t launch_reext_gather(const ReextArgs &a, hipStream_t s);
hipError_t launch_reext_finish(const ReextArgs &a, hipStream_t s);

// Intron-chain fast path: per-read state between the stages.
struct IcState {
    int32_t reason;
    uint32_t ctg, minus, junction_off;
    uint32_t l_q_lo, l_q_hi, l_r_lo, l_r_hi, l_nm;
    uint32_t r_q_lo, r_q_hi, r_r_lo, r_r_hi, r_nm;
    uint32_t istart, iend;
    uint32_t d_run, a_run, d_lo, a_lo, v_run;
    int64_t donor_end, acceptor_start;
};
struct IcArgs {
    RefView R;
    JunctionView J;
    SpliceCfgDev sp;
    uint32_t n, k, max_read;
    uint32_t min_frag_cov, refine_window, refine_pattern_len, intron_max;
    const uint8_t *reads;
    const uint64_t *read_off;
    const uint32_t *read_len;
    const uint32_t *anchors;
    const uint64_t *anc_off;
    uint32_t *mems;   // 6 per anchor slot
    uint32_t *frags;  // 8 per anchor slot
    IcState *st;
    // EA-Myers jobs, two per read (job k and n + k)
    uint64_t *p_off;
    uint32_t *p_len, *t_ctg, *t_lo, *t_len;
    const uint64_t *t_off;
    uint8_t *arena;
    const int32_t *dist;
    const uint64_t *end, *start;
    // the spliced pipeline's inputs and records
    uint32_t *exons, *nex, *ctg_out;
    uint64_t *exon_off;
    uint64_t *plen;
    int64_t *hint;
    const esp_splice_rec *srec;
    const esp_ext *ext;
    const uint32_t *stitched;
    const uint64_t *stitch_off;
    esp_intron_chain_rec *rec;
};
hipError_t launch_ic(const IcArgs &a, int stage, hipStream_t s);

// Exclusive prefix sums (totals at [n]) of u32 counts and u64 lengths, either
// may be null; tmp holds scan2_tmp_bytes(n) of device scratch.
size_t scan2_tmp_bytes(size_t n);
hipError_t launch_scan2(const uint32_t *cnt, const uint64_t *len, size_t n, uint64_t *cnt_off, uint64_t *len_off,
                        uint64_t *tmp, hipStream_t s);

hipError_t launch_gather(const GatherArgs &a, hipStream_t s);
hipError_t launch_winmyers_prep(const WinMyersArgs &a, hipStream_t s);
hipError_t launch_unsp_prep(const UnspArgs &a, hipStream_t s);
hipError_t launch_unsp_ea(const UnspArgs &a, hipStream_t s);
hipError_t launch_unsp_finish(const UnspArgs &a, hipStream_t s);
hipError_t launch_mate_prep(const MateArgs &a, hipStream_t s);
hipError_t launch_mate_select(const MateArgs &a, hipStream_t s);
hipError_t launch_mate_finish(const MateArgs &a, hipStream_t s);
hipError_t launch_splice_prep(const SpliceArgs &a, hipStream_t s);
hipError_t launch_splice_finish(const SpliceArgs &a, hipStream_t s);

}  // namespace espk
