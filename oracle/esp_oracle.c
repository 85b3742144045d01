/*
 * esp_oracle.c -- CPU restatement of the reference alignment core, used ONLY
 * as test infrastructure: the checker in tests/, the smoke check and the
 * timed CPU baseline leg of bench.py.  Nothing in the product library
 * (esperanto_amd/) links or calls this file.
 *
 * It restates, function by function, the reference's
 *   crates/map/src/extend.rs    (scalar run_band<i16|i32>, finish_extension,
 *                                run_dispatch, extend, extend_hint, push_op)
 *   crates/map/src/myers_ea.rs  (build_peq, calculate_block, scan, infix*,
 *                                long::*)
 *   crates/map/src/splice.rs    (stitch_cigar, align_spliced after
 *                                prepare_spliced, SplicedAlignment::genomic_start)
 *   crates/map/src/fasta.rs     (Contig::decode_into / base / is_n)
 *   crates/map/src/align.rs     (mate_window_placement, revcomp_into,
 *                                continuous_extension_window, the EA-Myers
 *                                lane of extend_unspliced_chain, the
 *                                placement walks rescore_with,
 *                                recount_mm_ea_in, exact_kmer_support)
 *   crates/map/src/splice.rs    (splice_signal, intron_length_penalty)
 *   crates/map/src/gtf.rs       (JunctionLib::contains_any_strand)
 * plus two independent textbook checkers (local affine-gap DP max score over
 * a row-window set, semi-global EA edit distance) used to pin the
 * restatement where the reference has no literal expected outputs.
 *
 * The reference's AVX2 kernels (extend.rs:1094-2026) are specified to be
 * cell-for-cell identical to run_band::<i16> (their own tests,
 * extend.rs:2578-2913), so run_dispatch here routes every pure input to the
 * scalar i16 restatement; the results are the same by that contract.
 *
 * Result layout: CIGAR ops as BAM u32 (len << 4 | op), op M=0 I=1 D=2 N=3
 * S=4 (bamio/src/encode.rs:169-177).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <pthread.h>

#define OP_M 0u
#define OP_I 1u
#define OP_D 2u
#define OP_N 3u
#define OP_S 4u

typedef struct {
    int32_t match_score; /* extend.rs:98-122 */
    int32_t mismatch;
    int32_t gap_open;
    int32_t gap_ext;
    int32_t edit_dir; /* 0 Off, 1 Ag, 2 Tc (extend.rs:43-50) */
} espo_params;

typedef struct {
    uint32_t read_start, read_end, ref_start;
    int32_t score;
} espo_ext;

typedef struct {
    int32_t touched;
    int64_t row;
    int32_t left;
    int32_t clamped;
} espo_touch;

/* ---------------------------------------------------------------- codes */

/* extend.rs:151-159 */
static inline int base_code(uint8_t b) {
    switch (b) {
    case 'A': case 'a': return 0;
    case 'C': case 'c': return 1;
    case 'G': case 'g': return 2;
    case 'T': case 't': return 3;
    default: return 4;
    }
}

/* extend.rs:82-93: Some(true) = ref A/read G, Some(false) = ref T/read C. */
static inline int exempt_dir(uint8_t ref_b, uint8_t read_b) {
    if (ref_b == 'A' && read_b == 'G') return 1;
    if (ref_b == 'T' && read_b == 'C') return 0;
    return -1;
}

/* extend.rs:68-73 */
static inline int edit_exempt(int dir, uint8_t ref_b, uint8_t read_b) {
    int e = exempt_dir(ref_b, read_b);
    return (dir == 1 && e == 1) || (dir == 2 && e == 0);
}

typedef struct {
    int16_t lut[25]; /* [read_code*5 + ref_code], extend.rs:161-186 */
    int32_t open_ext, ext; /* extend.rs:450-459 */
} core_params;

static void core_new(const espo_params *p, core_params *c) {
    int16_t m = (int16_t)p->match_score, mm = (int16_t)p->mismatch;
    for (int k = 0; k < 25; k++) c->lut[k] = mm;
    for (int k = 0; k < 4; k++) c->lut[k * 5 + k] = m;
    if (edit_exempt(p->edit_dir, 'A', 'G')) c->lut[10] = 0;
    if (edit_exempt(p->edit_dir, 'T', 'C')) c->lut[8] = 0;
    c->open_ext = -(p->gap_open + p->gap_ext);
    c->ext = -p->gap_ext;
}

/* ---------------------------------------------------------------- CIGAR */

typedef struct {
    uint32_t *ops;
    size_t n, cap;
    int overflow;
} opvec;

static inline uint32_t op_kind(uint32_t op) { return op & 0xfu; }
static inline uint32_t op_len(uint32_t op) { return op >> 4; }

/* push_op: extend.rs:254-282 (merge with the previous op of the same kind) */
static void push_op(opvec *v, uint32_t kind, uint32_t len) {
    if (v->n > 0 && op_kind(v->ops[v->n - 1]) == kind) {
        v->ops[v->n - 1] = ((op_len(v->ops[v->n - 1]) + len) << 4) | kind;
        return;
    }
    if (v->n >= v->cap) {
        v->overflow = 1;
        v->n++;
        return;
    }
    v->ops[v->n++] = (len << 4) | kind;
}

/* ---------------------------------------------------------------- Acc */

#define NEG16 (-30000)
#define NEG32 (-1000000000)

/* i16 Acc (extend.rs:473-501): exact-equality sentinel, wrapping add, clamp.
 * i32 Acc (extend.rs:503-528).  Values are carried in int32. */
typedef struct {
    int is16;
    int32_t NEG;
} acc_t;

static inline int32_t acc_from(const acc_t *a, int32_t v) {
    return a->is16 ? (int32_t)(int16_t)v : v;
}
static inline int32_t acc_add(const acc_t *a, int32_t x, int32_t y) {
    if (x == a->NEG || y == a->NEG) return a->NEG;
    if (a->is16) {
        int32_t s = (int32_t)(int16_t)(uint16_t)((uint32_t)x + (uint32_t)y);
        return s > a->NEG ? s : a->NEG;
    }
    return (int32_t)((uint32_t)x + (uint32_t)y);
}

/* ---------------------------------------------------------------- run_band */

#define H_DIAG 0
#define H_FROM_E 1
#define H_FROM_F 2
#define H_START 3

typedef struct {
    int64_t lo;
    size_t off, len;
} row_trace;

typedef struct {
    uint8_t op;   /* OP_M / OP_I / OP_D */
    int32_t delta;
} step_t;

/* finish_extension: extend.rs:864-1068. */
static void finish_extension(int64_t m, const core_params *core, const row_trace *rows,
                             const uint8_t *rows_flat, const int *read_codes,
                             const int *ref_codes, int64_t n_ref, int64_t best_i,
                             int64_t best_j, espo_ext *out, opvec *cig) {
    size_t steps_cap = (size_t)(best_i + best_j + 4);
    step_t *steps = (step_t *)malloc(steps_cap * sizeof(step_t));
    size_t ns = 0;
    int state = 0; /* 0 H, 1 E, 2 F */
    int64_t i = best_i, j = best_j;
    for (;;) {
        if (i <= 0 || j <= 0) break;
        const row_trace *row = &rows[i];
        if (j < row->lo || j >= row->lo + (int64_t)row->len) break;
        uint8_t cell = rows_flat[row->off + (size_t)(j - row->lo)];
        if (state == 0) {
            int hd = cell & 3;
            if (hd == H_DIAG) {
                int32_t s = core->lut[read_codes[i - 1] * 5 + ref_codes[j - 1]];
                steps[ns].op = OP_M; steps[ns].delta = s; ns++;
                i--; j--;
            } else if (hd == H_FROM_E) {
                steps[ns].op = OP_I; steps[ns].delta = core->open_ext; ns++;
                state = ((cell >> 2) & 1) ? 1 : 0;
                i--;
            } else if (hd == H_FROM_F) {
                steps[ns].op = OP_D; steps[ns].delta = core->open_ext; ns++;
                state = ((cell >> 4) & 1) ? 2 : 0;
                j--;
            } else {
                break;
            }
        } else if (state == 1) {
            steps[ns].op = OP_I; steps[ns].delta = core->ext; ns++;
            if (((cell >> 2) & 1) == 0) state = 0;
            i--;
        } else {
            steps[ns].op = OP_D; steps[ns].delta = core->ext; ns++;
            if (((cell >> 4) & 1) == 0) state = 0;
            j--;
        }
    }
    /* steps.reverse() */
    for (size_t a = 0, b = ns ? ns - 1 : 0; a < b; a++, b--) {
        step_t t = steps[a]; steps[a] = steps[b]; steps[b] = t;
    }
    /* End trimming (extend.rs:932-960). */
    size_t lead = 0, trail = 0;
    {
        int32_t acc = 0;
        for (size_t k = 0; k < ns; k++) {
            acc += steps[k].delta;
            if (acc <= 0) lead = k + 1;
        }
    }
    {
        int32_t acc = 0;
        for (size_t kk = ns; kk > 0; kk--) {
            size_t k = kk - 1;
            acc += steps[k].delta;
            if (acc <= 0) trail = ns - k;
        }
    }
    if (ns > 0 && lead + trail >= ns) {
        lead = ns / 2;
        trail = ns - lead - 1;
    }
    uint32_t read_start = (uint32_t)(i > 0 ? i : 0);
    uint32_t ref_start = (uint32_t)(j > 0 ? j : 0);
    for (size_t k = 0; k < lead; k++) {
        if (steps[k].op == OP_M || steps[k].op == OP_I) read_start++;
        if (steps[k].op == OP_M || steps[k].op == OP_D) ref_start++;
    }
    /* Rebuild (extend.rs:974-990). */
    int32_t score = 0;
    uint32_t *tmp_ops = (uint32_t *)malloc((ns + 4) * sizeof(uint32_t));
    opvec ops = {tmp_ops, 0, ns + 4, 0};
    uint32_t rm = 0, used_ref = 0;
    for (size_t k = lead; k < ns - trail; k++) {
        score += steps[k].delta;
        push_op(&ops, steps[k].op, 1);
        if (steps[k].op == OP_M || steps[k].op == OP_I) rm++;
        if (steps[k].op == OP_M || steps[k].op == OP_D) used_ref++;
    }
    uint32_t read_end = read_start + rm;
    uint32_t neutral_left = 0, neutral_right = 0;
    if (score > 0) {
        /* Neutral flank scan (extend.rs:991-1049). */
        static const uint8_t ACGT[4] = {'A', 'C', 'G', 'T'};
#define NEUTRAL(q, r) ((q) < 4 && (r) < 4 && ((q) == (r) || exempt_dir(ACGT[r], ACGT[q]) >= 0))
        int32_t delta = 0;
        uint32_t lim = read_start < ref_start ? read_start : ref_start;
        for (uint32_t nn = 1; nn <= lim; nn++) {
            int q = read_codes[read_start - nn];
            int r = ref_codes[ref_start - nn];
            if (!NEUTRAL(q, r)) break;
            delta += core->lut[q * 5 + r];
            if (delta == 0) neutral_left = nn;
        }
        uint32_t ref_end = ref_start + used_ref;
        delta = 0;
        uint32_t av1 = (uint32_t)m - read_end, av2 = (uint32_t)n_ref - ref_end;
        uint32_t available = av1 < av2 ? av1 : av2;
        for (uint32_t nn = 0; nn < available; nn++) {
            int q = read_codes[read_end + nn];
            int r = ref_codes[ref_end + nn];
            if (!NEUTRAL(q, r)) break;
            delta += core->lut[q * 5 + r];
            if (delta == 0) neutral_right = nn + 1;
        }
#undef NEUTRAL
        read_start -= neutral_left;
        ref_start -= neutral_left;
        read_end += neutral_right;
    }
    /* CIGAR assembly (extend.rs:1036-1059). */
    if (read_start > 0) push_op(cig, OP_S, read_start);
    if (neutral_left > 0) push_op(cig, OP_M, neutral_left);
    for (size_t k = 0; k < ops.n; k++) push_op(cig, op_kind(ops.ops[k]), op_len(ops.ops[k]));
    if (neutral_right > 0) push_op(cig, OP_M, neutral_right);
    if (read_end < (uint32_t)m) push_op(cig, OP_S, (uint32_t)m - read_end);
    out->read_start = read_start;
    out->read_end = read_end;
    out->ref_start = ref_start;
    out->score = score;
    free(tmp_ops);
    free(steps);
}

static void empty_extension(uint32_t m, espo_ext *out, opvec *cig) {
    out->read_start = out->read_end = out->ref_start = 0;
    out->score = 0;
    if (m > 0) push_op(cig, OP_S, m);
}

/* Planted mistakes for the tests that must catch them (0 = none): 1 tests
 * the right band edge before the left one, 2 keeps the last row maximum
 * (>=) instead of the first, 3 lets E win a tie with the diagonal, 4 takes a
 * gap extension on a tie with the open.  Only tests/ set it. */
static int g_mutation = 0;
void espo_set_mutation(int k) { g_mutation = k; }

/* run_band<T>: extend.rs:606-851.  is16 selects the i16 Acc.  best_cell, when
 * not NULL, receives (best_i, best_j). */
static void run_band(const uint8_t *read, int64_t m, const uint8_t *rf, int64_t n,
                     const core_params *core, int is16, int has_band, int64_t b_offset,
                     int64_t b_half, espo_ext *out, opvec *cig, espo_touch *touch,
                     int64_t *best_cell) {
    memset(touch, 0, sizeof(*touch));
    if (best_cell) best_cell[0] = best_cell[1] = 0;
    if (m == 0 || n == 0) {
        empty_extension((uint32_t)m, out, cig);
        return;
    }
    acc_t A = {is16, is16 ? NEG16 : NEG32};
    int64_t offset, half;
    if (has_band) {
        offset = b_offset;
        half = b_half > 0 ? b_half : 0;
    } else {
        offset = 0;
        half = n;
    }
    size_t W = (size_t)n + 2; /* every stored row is at most n + 1 wide */
    int32_t *h_prev = (int32_t *)malloc(W * sizeof(int32_t));
    int32_t *h_cur = (int32_t *)malloc(W * sizeof(int32_t));
    int32_t *e_prev = (int32_t *)malloc(W * sizeof(int32_t));
    int32_t *e_cur = (int32_t *)malloc(W * sizeof(int32_t));
    int32_t *f_cur = (int32_t *)malloc(W * sizeof(int32_t));
    for (size_t k = 0; k < W; k++) h_prev[k] = h_cur[k] = e_prev[k] = e_cur[k] = f_cur[k] = A.NEG;
    row_trace *rows = (row_trace *)malloc((size_t)(m + 1) * sizeof(row_trace));
    size_t flat_cap = 1024, flat_n = 0;
    uint8_t *flat = (uint8_t *)malloc(flat_cap);
    int *rc_codes = (int *)malloc((size_t)m * sizeof(int));
    int *rf_codes = (int *)malloc((size_t)n * sizeof(int));
    for (int64_t k = 0; k < m; k++) rc_codes[k] = base_code(read[k]);
    for (int64_t k = 0; k < n; k++) rf_codes[k] = base_code(rf[k]);

#define FLAT_GROW(extra)                                              \
    do {                                                              \
        while (flat_n + (extra) > flat_cap) flat_cap *= 2;            \
        flat = (uint8_t *)realloc(flat, flat_cap);                    \
    } while (0)

    int64_t center0 = offset;
    int64_t lo0 = center0 - half > 0 ? center0 - half : 0;
    int64_t hi0 = center0 + half < n ? center0 + half : n;
    if (lo0 <= hi0) {
        size_t len = (size_t)(hi0 - lo0 + 1);
        FLAT_GROW(len);
        memset(flat + flat_n, H_START, len);
        for (size_t c = 0; c < len; c++) h_prev[c] = acc_from(&A, 0);
        rows[0].lo = lo0; rows[0].off = flat_n; rows[0].len = len;
        flat_n += len;
    } else {
        rows[0].lo = lo0; rows[0].off = flat_n; rows[0].len = 0;
    }
    int32_t open_ext = acc_from(&A, core->open_ext);
    int32_t ext_p = acc_from(&A, core->ext);
    int32_t zero = 0;
    int32_t best_score = zero;
    int64_t best_i = 0, best_j = 0;
    int64_t plo = lo0, phi = hi0;
    for (int64_t i = 1; i <= m; i++) {
        int64_t center = offset + i;
        int64_t c = center < 1 ? 1 : (center > n ? n : center);
        int64_t lo = c - half > 1 ? c - half : 1;
        int64_t hi = c + half < n ? c + half : n;
        if (lo > hi) {
            rows[i].lo = lo; rows[i].off = flat_n; rows[i].len = 0;
            plo = lo; phi = hi;
            continue;
        }
        size_t len = (size_t)(hi - lo + 1);
        FLAT_GROW(len);
        size_t off = flat_n;
        memset(flat + off, 0, len);
        flat_n += len;
        int rc = rc_codes[i - 1];
        int32_t row_max = A.NEG;
        int64_t row_max_cell = 0;
        for (int64_t jj = lo; jj <= hi; jj++) {
            size_t cidx = (size_t)(jj - lo);
            int in_prev = jj >= plo && jj <= phi;
            int32_t h_up = in_prev ? h_prev[jj - plo] : zero;
            int32_t e_up = in_prev ? e_prev[jj - plo] : A.NEG;
            int32_t e_open = acc_add(&A, h_up, open_ext);
            int32_t e_ext = acc_add(&A, e_up, ext_p);
            int32_t e_val; int edir;
            int e_better = g_mutation == 4 ? e_ext >= e_open : e_ext > e_open;
            if (e_better) { e_val = e_ext; edir = 1; } else { e_val = e_open; edir = 0; }
            int32_t h_left = jj > lo ? h_cur[cidx - 1] : zero;
            int32_t f_left = jj > lo ? f_cur[cidx - 1] : A.NEG;
            int32_t f_open = acc_add(&A, h_left, open_ext);
            int32_t f_ext = acc_add(&A, f_left, ext_p);
            int32_t f_val; int fdir;
            int f_better = g_mutation == 4 ? f_ext >= f_open : f_ext > f_open;
            if (f_better) { f_val = f_ext; fdir = 1; } else { f_val = f_open; fdir = 0; }
            int64_t hp = jj - 1;
            int32_t h_diag = (hp >= plo && hp <= phi) ? h_prev[hp - plo] : zero;
            int32_t s = core->lut[rc * 5 + rf_codes[jj - 1]];
            int32_t diag = acc_add(&A, h_diag, acc_from(&A, s));
            int32_t h_val = zero;
            int hdir = H_START;
            if (f_val != A.NEG && !(h_val > f_val)) { h_val = f_val; hdir = H_FROM_F; }
            if (g_mutation == 3) {
                if (diag != A.NEG && !(h_val > diag)) { h_val = diag; hdir = H_DIAG; }
                if (e_val != A.NEG && !(h_val > e_val)) { h_val = e_val; hdir = H_FROM_E; }
            } else {
                if (e_val != A.NEG && !(h_val > e_val)) { h_val = e_val; hdir = H_FROM_E; }
                if (diag != A.NEG && !(h_val > diag)) { h_val = diag; hdir = H_DIAG; }
            }
            if (h_val == zero) hdir = H_START;
            h_cur[cidx] = h_val;
            e_cur[cidx] = e_val;
            f_cur[cidx] = f_val;
            flat[off + cidx] = (uint8_t)(hdir | (edir << 2) | (fdir << 4));
            if (g_mutation == 2 ? h_val >= row_max : h_val > row_max) {
                row_max = h_val;
                row_max_cell = (int64_t)cidx;
            }
        }
        int left_strict = lo > 0;
        int right_strict = hi < n;
        if (!touch->touched && g_mutation == 1) {
            if (right_strict && h_cur[len - 1] == row_max) {
                touch->touched = 1; touch->row = i; touch->left = 0; touch->clamped = hi >= n;
            } else if (left_strict && h_cur[0] == row_max) {
                touch->touched = 1; touch->row = i; touch->left = 1; touch->clamped = lo <= 1;
            }
        } else if (!touch->touched) {
            if (left_strict && h_cur[0] == row_max) {
                touch->touched = 1; touch->row = i; touch->left = 1; touch->clamped = lo <= 1;
            } else if (right_strict && h_cur[len - 1] == row_max) {
                touch->touched = 1; touch->row = i; touch->left = 0; touch->clamped = hi >= n;
            }
        }
        if (row_max > best_score) {
            best_score = row_max;
            best_i = i;
            best_j = lo + row_max_cell;
        }
        rows[i].lo = lo; rows[i].off = off; rows[i].len = len;
        int32_t *t;
        t = h_prev; h_prev = h_cur; h_cur = t;
        t = e_prev; e_prev = e_cur; e_cur = t;
        plo = lo; phi = hi;
    }
#undef FLAT_GROW
    if (best_cell) { best_cell[0] = best_i; best_cell[1] = best_j; }
    finish_extension(m, core, rows, flat, rc_codes, rf_codes, n, best_i, best_j, out, cig);
    free(h_prev); free(h_cur); free(e_prev); free(e_cur); free(f_cur);
    free(rows); free(flat); free(rc_codes); free(rf_codes);
}

/* ------------------------------------------------------------ AVX2 kernels */

#include <immintrin.h>

/* Route of run_dispatch for pure inputs: 0 the scalar run_band<i16> (the
 * checker's default: the reference requires its AVX2 kernels to equal it),
 * 1 the AVX2 kernels exactly where the reference takes them (pure, m and n
 * >= 8, AVX2 present).  Set by tests and the CPU baseline. */
static int g_route = 0;
void espo_set_route(int r) { g_route = r; }
int espo_avx2_available(void) { return __builtin_cpu_supports("avx2"); }

/* Per-thread scratch of the AVX2 kernels, grown on demand (ExtendBuffer's
 * packed buffers, extend.rs:231). */
typedef struct {
    int16_t *h[3], *e[2], *f[2], *rmax, *hl, *hr, *rd, *rr;
    int32_t *rarg;
    int64_t *lo, *hi, *alo, *ahi, *base;
    size_t cap_m, cap_n;
} avx_scratch;
static __thread avx_scratch g_avx;

static void avx_reserve(size_t m, size_t n) {
    avx_scratch *S = &g_avx;
    if (m > S->cap_m) {
        size_t c = m + 64;
        for (int k = 0; k < 3; k++) S->h[k] = (int16_t *)realloc(S->h[k], (c + 33) * 2);
        for (int k = 0; k < 2; k++) S->e[k] = (int16_t *)realloc(S->e[k], (c + 33) * 2);
        for (int k = 0; k < 2; k++) S->f[k] = (int16_t *)realloc(S->f[k], (c + 33) * 2);
        S->rmax = (int16_t *)realloc(S->rmax, (c + 1) * 2);
        S->hl = (int16_t *)realloc(S->hl, (c + 1) * 2);
        S->hr = (int16_t *)realloc(S->hr, (c + 1) * 2);
        S->rd = (int16_t *)realloc(S->rd, (c + 16) * 2);
        S->rarg = (int32_t *)realloc(S->rarg, (c + 1) * 4);
        S->lo = (int64_t *)realloc(S->lo, (c + 2) * 8);
        S->hi = (int64_t *)realloc(S->hi, (c + 2) * 8);
        S->alo = (int64_t *)realloc(S->alo, (c + 2) * 8);
        S->ahi = (int64_t *)realloc(S->ahi, (c + 2) * 8);
        S->base = (int64_t *)realloc(S->base, (c + 2) * 8);
        S->cap_m = c;
    }
    if (n > S->cap_n) {
        size_t c = n + 64;
        S->rr = (int16_t *)realloc(S->rr, (c + 16) * 2);
        S->cap_n = c;
    }
}

/* The vector body shared by avx2_full / avx2_band (extend.rs:1352-1430,
 * :1855-1933): one 16-lane block of anti-diagonal d starting at row v. */
typedef struct {
    __m256i neg, zero, ones, one, two, three, four, open, ext, sdiag, smm, sga, stc, cg, ca, cc, ct;
} avx_consts;

static inline __m256i avx_addc(const avx_consts *K, __m256i x, __m256i c) {
    const __m256i isn = _mm256_cmpeq_epi16(x, K->neg);
    const __m256i r = _mm256_max_epi16(_mm256_add_epi16(x, c), K->neg);
    return _mm256_blendv_epi8(r, K->neg, isn);
}

__attribute__((target("avx2"))) static inline void avx_block(const avx_consts *K, const int16_t *rd,
                                                             const int16_t *rr, const int16_t *h2p,
                                                             const int16_t *h1p, int16_t *h0p, const int16_t *e1p,
                                                             int16_t *e0p, const int16_t *f1p, int16_t *f0p,
                                                             int64_t v, int64_t n, int64_t d, int16_t hb[16],
                                                             int16_t db[16]) {
    const __m256i rc = _mm256_loadu_si256((const __m256i *)(rd + (v - 1)));
    const __m256i rfc = _mm256_loadu_si256((const __m256i *)(rr + (v + (n - d))));
    const __m256i eq = _mm256_cmpeq_epi16(rc, rfc);
    const __m256i is_match = _mm256_and_si256(eq, _mm256_cmpgt_epi16(K->four, rc));
    const __m256i ga = _mm256_and_si256(_mm256_cmpeq_epi16(rc, K->cg), _mm256_cmpeq_epi16(rfc, K->ca));
    const __m256i tc = _mm256_and_si256(_mm256_cmpeq_epi16(rc, K->cc), _mm256_cmpeq_epi16(rfc, K->ct));
    __m256i sc = _mm256_blendv_epi8(K->smm, K->sga, ga);
    sc = _mm256_blendv_epi8(sc, K->stc, tc);
    sc = _mm256_blendv_epi8(sc, K->sdiag, is_match);
    const __m256i h_up = _mm256_loadu_si256((const __m256i *)(h1p + (v - 1)));
    const __m256i e_up = _mm256_loadu_si256((const __m256i *)(e1p + (v - 1)));
    const __m256i e_open = avx_addc(K, h_up, K->open), e_ext = avx_addc(K, e_up, K->ext);
    const __m256i egt = _mm256_cmpgt_epi16(e_ext, e_open);
    const __m256i e_val = _mm256_blendv_epi8(e_open, e_ext, egt);
    const __m256i edir = _mm256_and_si256(egt, K->one);
    const __m256i h_left = _mm256_loadu_si256((const __m256i *)(h1p + v));
    const __m256i f_left = _mm256_loadu_si256((const __m256i *)(f1p + v));
    const __m256i f_open = avx_addc(K, h_left, K->open), f_ext = avx_addc(K, f_left, K->ext);
    const __m256i fgt = _mm256_cmpgt_epi16(f_ext, f_open);
    const __m256i f_val = _mm256_blendv_epi8(f_open, f_ext, fgt);
    const __m256i fdir = _mm256_and_si256(fgt, K->one);
    const __m256i h_diag = _mm256_loadu_si256((const __m256i *)(h2p + (v - 1)));
    const __m256i diag = avx_addc(K, h_diag, sc);
    __m256i hv = K->zero, hd = K->three;
    const __m256i fmask = _mm256_andnot_si256(
        _mm256_or_si256(_mm256_cmpeq_epi16(f_val, K->neg), _mm256_cmpgt_epi16(hv, f_val)), K->ones);
    hv = _mm256_blendv_epi8(hv, f_val, fmask);
    hd = _mm256_blendv_epi8(hd, K->two, fmask);
    const __m256i emask = _mm256_andnot_si256(
        _mm256_or_si256(_mm256_cmpeq_epi16(e_val, K->neg), _mm256_cmpgt_epi16(hv, e_val)), K->ones);
    hv = _mm256_blendv_epi8(hv, e_val, emask);
    hd = _mm256_blendv_epi8(hd, K->one, emask);
    const __m256i dmask = _mm256_andnot_si256(
        _mm256_or_si256(_mm256_cmpeq_epi16(diag, K->neg), _mm256_cmpgt_epi16(hv, diag)), K->ones);
    hv = _mm256_blendv_epi8(hv, diag, dmask);
    hd = _mm256_blendv_epi8(hd, K->zero, dmask);
    hd = _mm256_blendv_epi8(hd, K->three, _mm256_cmpeq_epi16(hv, K->zero));
    const __m256i dir16 =
        _mm256_or_si256(hd, _mm256_or_si256(_mm256_slli_epi16(edir, 2), _mm256_slli_epi16(fdir, 4)));
    _mm256_storeu_si256((__m256i *)(h0p + v), hv);
    _mm256_storeu_si256((__m256i *)(e0p + v), e_val);
    _mm256_storeu_si256((__m256i *)(f0p + v), f_val);
    _mm256_storeu_si256((__m256i *)hb, hv);
    _mm256_storeu_si256((__m256i *)db, dir16);
}

__attribute__((target("avx2"))) static void avx_consts_init(avx_consts *K, int16_t oe, int16_t ex, int16_t sd,
                                                            int16_t sm, int16_t sg, int16_t st) {
    K->neg = _mm256_set1_epi16(NEG16);
    K->zero = _mm256_set1_epi16(0);
    K->ones = _mm256_set1_epi16(-1);
    K->one = _mm256_set1_epi16(1);
    K->two = _mm256_set1_epi16(2);
    K->three = _mm256_set1_epi16(3);
    K->four = _mm256_set1_epi16(4);
    K->open = _mm256_set1_epi16(oe);
    K->ext = _mm256_set1_epi16(ex);
    K->sdiag = _mm256_set1_epi16(sd);
    K->smm = _mm256_set1_epi16(sm);
    K->sga = _mm256_set1_epi16(sg);
    K->stc = _mm256_set1_epi16(st);
    K->cg = _mm256_set1_epi16(2);
    K->ca = _mm256_set1_epi16(0);
    K->cc = _mm256_set1_epi16(1);
    K->ct = _mm256_set1_epi16(3);
}

/* LUT constants of both kernels and their scalar fallback (extend.rs:1158-1168):
 * returns 0 when one of them is the i16 sentinel. */
static int avx_lut(const core_params *core, int16_t *sd, int16_t *sm, int16_t *sg, int16_t *st, int16_t *oe,
                   int16_t *ex) {
    *sd = core->lut[0];
    *sm = core->lut[1];
    *sg = core->lut[10];
    *st = core->lut[8];
    *oe = (int16_t)core->open_ext;
    *ex = (int16_t)core->ext;
    return !(*sd == NEG16 || *sm == NEG16 || *sg == NEG16 || *st == NEG16 || *oe == NEG16 || *ex == NEG16);
}

/* avx2_full::run + dp_diagonals: extend.rs:1118-1488. */
__attribute__((target("avx2"))) static void avx2_full_run(const uint8_t *read, int64_t m, const uint8_t *rf,
                                                          int64_t n, const core_params *core, espo_ext *out,
                                                          opvec *cig, espo_touch *touch, int64_t *best_cell) {
    int16_t sd, sm, sg, st, oe, ex;
    if (!avx_lut(core, &sd, &sm, &sg, &st, &oe, &ex)) {
        run_band(read, m, rf, n, core, 1, 0, 0, 0, out, cig, touch, best_cell);
        return;
    }
    memset(touch, 0, sizeof(*touch));
    avx_reserve((size_t)m, (size_t)n);
    avx_scratch *S = &g_avx;
    int *rc_codes = (int *)malloc((size_t)m * sizeof(int)), *rf_codes = (int *)malloc((size_t)n * sizeof(int));
    for (int64_t k = 0; k < m; k++) rc_codes[k] = base_code(read[k]);
    for (int64_t k = 0; k < n; k++) rf_codes[k] = base_code(rf[k]);
    const size_t flat_n = (size_t)(n + 1 + m * n);
    uint8_t *flat = (uint8_t *)calloc(flat_n, 1);
    memset(flat, H_START, (size_t)n + 1);
    row_trace *rows = (row_trace *)malloc((size_t)(m + 1) * sizeof(row_trace));
    rows[0].lo = 0; rows[0].off = 0; rows[0].len = (size_t)n + 1;
    for (int64_t i = 1; i <= m; i++) {
        rows[i].lo = 1;
        rows[i].off = (size_t)((n + 1) + (i - 1) * n);
        rows[i].len = (size_t)n;
    }
    const size_t cap = (size_t)m + 33;
    for (int k = 0; k < 3; k++) memset(S->h[k], 0, cap * 2);
    for (int k = 0; k < 2; k++) { memset(S->e[k], 0, cap * 2); memset(S->f[k], 0, cap * 2); }
    for (int64_t i = 0; i <= m; i++) { S->rmax[i] = NEG16; S->rarg[i] = 1; S->hl[i] = NEG16; }
    for (int64_t i = 0; i < m + 16; i++) S->rd[i] = i < m ? (int16_t)rc_codes[i] : 0;
    for (int64_t q = 0; q < n + 16; q++) S->rr[q] = q < n ? (int16_t)rf_codes[n - 1 - q] : 0;
    avx_consts K;
    avx_consts_init(&K, oe, ex, sd, sm, sg, st);
    int h2 = 0, h1 = 1, h0 = 2, e1 = 0, e0 = 1;
    int16_t hb[16], db[16];
    for (int64_t d = 0; d <= m + n; d++) {
        const int64_t ilo = d - n > 1 ? d - n : 1, ihi = m < d - 1 ? m : d - 1;
        for (int64_t v = ilo; v <= ihi; v += 16) {
            avx_block(&K, S->rd, S->rr, S->h[h2], S->h[h1], S->h[h0], S->e[e1], S->e[e0], S->f[e1], S->f[e0], v,
                      n, d, hb, db);
            for (int k = 0; k < 16; k++) {
                const int64_t i = v + k;
                if (i > ihi) break;
                flat[(size_t)(i * (n - 1) + d)] = (uint8_t)db[k];
                if (hb[k] > S->rmax[i]) { S->rmax[i] = hb[k]; S->rarg[i] = (int32_t)(d - i); }
                if (d - i == 1) S->hl[i] = hb[k];
            }
        }
        if (d <= n) { S->h[h0][0] = 0; S->e[e0][0] = NEG16; S->f[e0][0] = NEG16; }
        if (d <= m) { S->h[h0][d] = 0; S->e[e0][d] = NEG16; S->f[e0][d] = NEG16; }
        const int t = h2; h2 = h1; h1 = h0; h0 = t;
        const int t2 = e1; e1 = e0; e0 = t2;
    }
    int16_t best = 0;
    int64_t bi = 0, bj = 0;
    for (int64_t i = 1; i <= m; i++) {
        const int16_t rm = S->rmax[i];
        if (!touch->touched && S->hl[i] == rm) { touch->touched = 1; touch->row = i; touch->left = 1; touch->clamped = 1; }
        if (rm > best) { best = rm; bi = i; bj = S->rarg[i]; }
    }
    if (best_cell) { best_cell[0] = bi; best_cell[1] = bj; }
    finish_extension(m, core, rows, flat, rc_codes, rf_codes, n, bi, bj, out, cig);
    free(rows); free(flat); free(rc_codes); free(rf_codes);
}

/* avx2_band::run + dp_diagonals: extend.rs:1538-2026. */
__attribute__((target("avx2"))) static void avx2_band_run(const uint8_t *read, int64_t m, const uint8_t *rf,
                                                          int64_t n, const core_params *core, int64_t offset,
                                                          int64_t half, espo_ext *out, opvec *cig,
                                                          espo_touch *touch, int64_t *best_cell) {
    if (half < 0) half = 0;
    int16_t sd, sm, sg, st, oe, ex;
    if (!avx_lut(core, &sd, &sm, &sg, &st, &oe, &ex)) {
        run_band(read, m, rf, n, core, 1, 1, offset, half, out, cig, touch, best_cell);
        return;
    }
    memset(touch, 0, sizeof(*touch));
    avx_reserve((size_t)m, (size_t)n);
    avx_scratch *S = &g_avx;
    int *rc_codes = (int *)malloc((size_t)m * sizeof(int)), *rf_codes = (int *)malloc((size_t)n * sizeof(int));
    for (int64_t k = 0; k < m; k++) rc_codes[k] = base_code(read[k]);
    for (int64_t k = 0; k < n; k++) rf_codes[k] = base_code(rf[k]);
    row_trace *rows = (row_trace *)malloc((size_t)(m + 1) * sizeof(row_trace));
    const int64_t lo0 = offset - half > 0 ? offset - half : 0, hi0 = offset + half < n ? offset + half : n;
    const size_t len0 = lo0 <= hi0 ? (size_t)(hi0 - lo0 + 1) : 0;
    rows[0].lo = lo0; rows[0].off = 0; rows[0].len = len0;
    size_t off_acc = len0;
    for (int64_t i = 1; i <= m; i++) {
        int64_t c = offset + i;
        c = c < 1 ? 1 : (c > n ? n : c);
        const int64_t lo = c - half > 1 ? c - half : 1, hi = c + half < n ? c + half : n;
        const size_t len = (size_t)(hi - lo + 1);
        S->lo[i] = lo; S->hi[i] = hi; S->alo[i] = lo + i; S->ahi[i] = hi + i;
        S->base[i] = (int64_t)off_acc - lo;
        rows[i].lo = lo; rows[i].off = off_acc; rows[i].len = len;
        off_acc += len;
    }
    uint8_t *flat = (uint8_t *)calloc(off_acc ? off_acc : 1, 1);
    memset(flat, H_START, len0);
    const int64_t dmin = S->alo[1], dmax = S->ahi[m];
    const size_t cap = (size_t)m + 33;
    for (int k = 0; k < 3; k++) memset(S->h[k], 0, cap * 2);
    for (int k = 0; k < 2; k++)
        for (size_t q = 0; q < cap; q++) { S->e[k][q] = NEG16; S->f[k][q] = NEG16; }
    for (int64_t i = 0; i <= m; i++) { S->rmax[i] = NEG16; S->rarg[i] = 1; S->hl[i] = NEG16; S->hr[i] = NEG16; }
    for (int64_t i = 0; i < m + 16; i++) S->rd[i] = i < m ? (int16_t)rc_codes[i] : 0;
    for (int64_t q = 0; q < n + 16; q++) S->rr[q] = q < n ? (int16_t)rf_codes[n - 1 - q] : 0;
    avx_consts K;
    avx_consts_init(&K, oe, ex, sd, sm, sg, st);
    int h2 = 0, h1 = 1, h0 = 2, e1 = 0, e0 = 1;
    int64_t blo_c = 1, bhi_c = 1, me_r = 1, mf_r = 1, md_r = 1;
    int16_t hb[16], db[16];
    for (int64_t d = dmin; d <= dmax; d++) {
        while (blo_c <= m && S->ahi[blo_c] < d) blo_c++;
        while (bhi_c < m && S->alo[bhi_c + 1] <= d) bhi_c++;
        const int64_t blo = blo_c, bhi = bhi_c;
        for (int64_t v = blo; v <= bhi; v += 16) {
            avx_block(&K, S->rd, S->rr, S->h[h2], S->h[h1], S->h[h0], S->e[e1], S->e[e0], S->f[e1], S->f[e0], v,
                      n, d, hb, db);
            for (int k = 0; k < 16; k++) {
                const int64_t i = v + k;
                if (i > bhi) break;
                const int64_t j = d - i;
                flat[(size_t)(S->base[i] + j)] = (uint8_t)db[k];
                if (hb[k] > S->rmax[i]) { S->rmax[i] = hb[k]; S->rarg[i] = (int32_t)j; }
                if (j == S->lo[i]) S->hl[i] = hb[k];
                if (j == S->hi[i]) S->hr[i] = hb[k];
            }
        }
        while (me_r <= m && S->ahi[me_r] < d + 1) me_r++;
        if (me_r <= m && S->ahi[me_r] == d + 1 && me_r >= 2 && S->hi[me_r] == S->hi[me_r - 1] + 1) {
            S->h[h0][me_r - 1] = 0; S->e[e0][me_r - 1] = NEG16; S->f[e0][me_r - 1] = NEG16;
        }
        while (mf_r <= m && S->alo[mf_r] < d + 1) mf_r++;
        if (mf_r <= m && S->alo[mf_r] == d + 1) {
            S->h[h0][mf_r] = 0; S->e[e0][mf_r] = NEG16; S->f[e0][mf_r] = NEG16;
        }
        while (md_r <= m && S->alo[md_r] < d + 2) md_r++;
        if (md_r <= m && S->alo[md_r] == d + 2 && md_r >= 2 && S->lo[md_r] == S->lo[md_r - 1]) {
            S->h[h0][md_r - 1] = 0; S->e[e0][md_r - 1] = NEG16; S->f[e0][md_r - 1] = NEG16;
        }
        const int t = h2; h2 = h1; h1 = h0; h0 = t;
        const int t2 = e1; e1 = e0; e0 = t2;
    }
    int16_t best = 0;
    int64_t bi = 0, bj = 0;
    for (int64_t i = 1; i <= m; i++) {
        const int16_t rm = S->rmax[i];
        if (!touch->touched) {
            if (S->lo[i] > 0 && S->hl[i] == rm) {
                touch->touched = 1; touch->row = i; touch->left = 1; touch->clamped = S->lo[i] <= 1;
            } else if (S->hi[i] < n && S->hr[i] == rm) {
                touch->touched = 1; touch->row = i; touch->left = 0; touch->clamped = S->hi[i] >= n;
            }
        }
        if (rm > best) { best = rm; bi = i; bj = S->rarg[i]; }
    }
    if (best_cell) { best_cell[0] = bi; best_cell[1] = bj; }
    finish_extension(m, core, rows, flat, rc_codes, rf_codes, n, bi, bj, out, cig);
    free(rows); free(flat); free(rc_codes); free(rf_codes);
}

/* ---------------------------------------------------------------- API */

static int pure_byte(uint8_t b) { return b == 'A' || b == 'C' || b == 'G' || b == 'T' || b == 'N'; }

/* run_dispatch: extend.rs:293-344.  The AVX2 branches are cell-for-cell the
 * i16 scalar kernel by the reference's own contract, so pure inputs run the
 * i16 restatement and everything else the i32 one; with espo_set_route(1)
 * pure inputs of m, n >= 8 take the AVX2 restatements instead, as the
 * reference does on an AVX2 host. */
static void run_dispatch(const uint8_t *read, int64_t m, const uint8_t *rf, int64_t n,
                         const espo_params *p, int has_band, int64_t offset, int64_t half,
                         espo_ext *out, opvec *cig, espo_touch *touch) {
    core_params core;
    core_new(p, &core);
    int pure = 1;
    for (int64_t k = 0; k < m && pure; k++) pure = pure_byte(read[k]);
    for (int64_t k = 0; k < n && pure; k++) pure = pure_byte(rf[k]);
    if (g_route == 1 && pure && m >= 8 && n >= 8 && espo_avx2_available()) {
        if (has_band) avx2_band_run(read, m, rf, n, &core, offset, half, out, cig, touch, NULL);
        else avx2_full_run(read, m, rf, n, &core, out, cig, touch, NULL);
        return;
    }
    run_band(read, m, rf, n, &core, pure, has_band, offset, half, out, cig, touch, NULL);
}

#define INITIAL_BAND 30 /* extend.rs:126 */

static int finish_call(opvec *cig, size_t *n_ops) {
    *n_ops = cig->n;
    return cig->overflow ? 2 : 0;
}

/* Internal run_band with explicit accumulator width (tests compare both). */
int espo_run_band(const uint8_t *read, size_t m, const uint8_t *rf, size_t n,
                  const espo_params *p, int is16, int has_band, int64_t offset, int64_t half,
                  espo_ext *out, espo_touch *touch, uint32_t *cigar, size_t cap, size_t *n_ops) {
    core_params core;
    core_new(p, &core);
    opvec cig = {cigar, 0, cap, 0};
    run_band(read, (int64_t)m, rf, (int64_t)n, &core, is16, has_band, offset, half, out, &cig, touch,
             NULL);
    return finish_call(&cig, n_ops);
}

/* espo_run_band plus the best cell (best_i, best_j) the traceback starts from. */
int espo_run_band_cell(const uint8_t *read, size_t m, const uint8_t *rf, size_t n,
                       const espo_params *p, int is16, int has_band, int64_t offset, int64_t half,
                       espo_ext *out, espo_touch *touch, uint32_t *cigar, size_t cap, size_t *n_ops,
                       int64_t *best_cell) {
    core_params core;
    core_new(p, &core);
    opvec cig = {cigar, 0, cap, 0};
    run_band(read, (int64_t)m, rf, (int64_t)n, &core, is16, has_band, offset, half, out, &cig, touch,
             best_cell);
    return finish_call(&cig, n_ops);
}

/* run_dispatch(.., Some((offset,1,1,half))) or None when has_band == 0. */
int espo_band(const uint8_t *read, size_t m, const uint8_t *rf, size_t n, const espo_params *p,
              int has_band, int64_t offset, int64_t half, espo_ext *out, espo_touch *touch,
              uint32_t *cigar, size_t cap, size_t *n_ops) {
    opvec cig = {cigar, 0, cap, 0};
    run_dispatch(read, (int64_t)m, rf, (int64_t)n, p, has_band, offset, half, out, &cig, touch);
    return finish_call(&cig, n_ops);
}

/* extend: extend.rs:347-354 */
int espo_extend(const uint8_t *read, size_t m, const uint8_t *rf, size_t n, const espo_params *p,
                espo_ext *out, uint32_t *cigar, size_t cap, size_t *n_ops) {
    espo_touch t;
    return espo_band(read, m, rf, n, p, 0, 0, 0, out, &t, cigar, cap, n_ops);
}

/* extend_hint: extend.rs:392-428.  *escalated reports which attempt won
 * (0 band accepted, 1 full matrix, 2 empty, 3 band covers window). */
int espo_extend_hint(const uint8_t *read, size_t m, const uint8_t *rf, size_t n,
                     const espo_params *p, int64_t hint_offset, espo_ext *out, uint32_t *cigar,
                     size_t cap, size_t *n_ops, int32_t *escalated) {
    opvec cig = {cigar, 0, cap, 0};
    espo_touch t;
    if (m == 0 || n == 0) {
        empty_extension((uint32_t)m, out, &cig);
        *escalated = 2;
        return finish_call(&cig, n_ops);
    }
    if ((int64_t)INITIAL_BAND >= (int64_t)n) {
        run_dispatch(read, (int64_t)m, rf, (int64_t)n, p, 0, 0, 0, out, &cig, &t);
        *escalated = 3;
        return finish_call(&cig, n_ops);
    }
    run_dispatch(read, (int64_t)m, rf, (int64_t)n, p, 1, hint_offset, INITIAL_BAND, out, &cig, &t);
    /* extend.rs:423-424 in release i32 arithmetic, which wraps on overflow
     * (the scaled scores of the envelope can take score * 20 past 2^31). */
    int32_t perfect = (int32_t)((uint32_t)m * (uint32_t)p->match_score);
    int32_t lhs = (int32_t)((uint32_t)out->score * 20u), rhs = (int32_t)((uint32_t)perfect * 17u);
    if (t.touched || lhs < rhs) {
        cig.n = 0;
        cig.overflow = 0;
        run_dispatch(read, (int64_t)m, rf, (int64_t)n, p, 0, 0, 0, out, &cig, &t);
        *escalated = 1;
        return finish_call(&cig, n_ops);
    }
    *escalated = 0;
    return finish_call(&cig, n_ops);
}

/* ---------------------------------------------------------------- Myers */

typedef unsigned __int128 u128;

/* myers_ea.rs:27-38 */
static int ea_matches(uint8_t read, uint8_t reference) {
    uint8_t r = (read >= 'a' && read <= 'z') ? (uint8_t)(read - 32) : read;
    uint8_t f = (reference >= 'a' && reference <= 'z') ? (uint8_t)(reference - 32) : reference;
    switch (r) {
    case 'A': return f == 'A';
    case 'C': return f == 'C' || f == 'T';
    case 'G': return f == 'G' || f == 'A';
    case 'T': return f == 'T';
    default: return 0;
    }
}

/* myers_ea.rs:18-24 */
static u128 mask_for(size_t m) { return m >= 128 ? ~(u128)0 : (((u128)1 << m) - 1); }

/* myers_ea.rs:42-56 (block 0) and :194-213 (2-block) */
static void build_peq_blocks(const uint8_t *read, size_t m, size_t nblk, u128 *peq /* nblk*256 */) {
    memset(peq, 0, nblk * 256 * sizeof(u128));
    static const uint8_t ACGT[4] = {'A', 'C', 'G', 'T'};
    size_t lim = nblk * 128 < m ? nblk * 128 : m;
    for (size_t i = 0; i < lim; i++)
        for (int f = 0; f < 4; f++)
            if (ea_matches(read[i], ACGT[f])) peq[(i / 128) * 256 + ACGT[f]] |= (u128)1 << (i % 128);
    static const uint8_t lo[4] = {'a', 'c', 'g', 't'};
    for (size_t b = 0; b < nblk; b++)
        for (int f = 0; f < 4; f++) peq[b * 256 + lo[f]] = peq[b * 256 + ACGT[f]];
}

/* calculate_block: myers_ea.rs:64-91 */
static void calculate_block(u128 pv, u128 mv, u128 eq_in, int hin, u128 mask, u128 high_bit,
                            u128 *pv_out, u128 *mv_out, int *hout) {
    u128 hin_is_neg = (u128)(hin < 0);
    u128 ph_bit0 = (u128)(hin > 0);
    u128 xv = eq_in | mv;
    u128 eq = eq_in | hin_is_neg;
    u128 xh = (((eq & pv) + pv) ^ pv) | eq;
    u128 ph = mv | ~(xh | pv);
    u128 mh = pv & xh;
    *hout = (ph & high_bit) ? 1 : ((mh & high_bit) ? -1 : 0);
    u128 ph_sh = (ph << 1) | ph_bit0;
    u128 mh_sh = (mh << 1) | hin_is_neg;
    *pv_out = (mh_sh | ~(xv | ph_sh)) & mask;
    *mv_out = (ph_sh & xv) & mask;
}

/* scan (myers_ea.rs:93-114) / scan2 (:215-246), with the best/best_at update
 * of infix_best_end (strict <, earliest end). */
static void myers_scan(const uint8_t *read, size_t m, const uint8_t *text, size_t n, int32_t *dist,
                       size_t *end) {
    if (m <= 128) {
        u128 peq[256];
        build_peq_blocks(read, m, 1, peq);
        u128 mask = mask_for(m);
        u128 high = (u128)1 << (m - 1);
        u128 vp = mask, vn = 0;
        int32_t score = (int32_t)m, best = score;
        size_t best_at = 0;
        for (size_t j = 0; j < n; j++) {
            int h;
            calculate_block(vp, vn, peq[text[j]], 0, mask, high, &vp, &vn, &h);
            score += h;
            if (score < best) { best = score; best_at = j + 1; }
        }
        *dist = best;
        *end = best_at;
        return;
    }
    u128 *peq = (u128 *)malloc(2 * 256 * sizeof(u128));
    build_peq_blocks(read, m, 2, peq);
    size_t m1 = m - 128;
    u128 mask1 = mask_for(m1);
    u128 high0 = (u128)1 << 127;
    u128 high1 = (u128)1 << (m1 - 1);
    u128 vp0 = ~(u128)0, vn0 = 0, vp1 = mask1, vn1 = 0;
    int32_t score = (int32_t)m, best = score;
    size_t best_at = 0;
    for (size_t j = 0; j < n; j++) {
        int h0, h1;
        u128 a, b, c, d;
        calculate_block(vp0, vn0, peq[text[j]], 0, ~(u128)0, high0, &a, &b, &h0);
        calculate_block(vp1, vn1, peq[256 + text[j]], h0, mask1, high1, &c, &d, &h1);
        vp0 = a; vn0 = b; vp1 = c; vn1 = d;
        score += h1;
        if (score < best) { best = score; best_at = j + 1; }
    }
    free(peq);
    *dist = best;
    *end = best_at;
}

/* long::infix_best_end (myers_ea.rs:271-286): m == 0 -> (0,0), m > 256 -> (m,0). */
void espo_myers_best_end(const uint8_t *read, size_t m, const uint8_t *text, size_t n,
                         int32_t *dist, size_t *end) {
    if (m == 0) { *dist = 0; *end = 0; return; }
    if (m > 256) { *dist = (int32_t)m; *end = 0; return; }
    myers_scan(read, m, text, n, dist, end);
}

/* long::infix_best_start_into (myers_ea.rs:294-310) */
void espo_myers_best_start(const uint8_t *read, size_t m, const uint8_t *text, size_t n,
                           int32_t *dist, size_t *start) {
    if (m == 0) { *dist = 0; *start = 0; return; }
    if (m > 256) { *dist = (int32_t)m; *start = 0; return; }
    uint8_t *rr = (uint8_t *)malloc(m ? m : 1), *rt = (uint8_t *)malloc(n ? n : 1);
    for (size_t k = 0; k < m; k++) rr[k] = read[m - 1 - k];
    for (size_t k = 0; k < n; k++) rt[k] = text[n - 1 - k];
    size_t rev_end;
    myers_scan(rr, m, rt, n, dist, &rev_end);
    *start = n - rev_end;
    free(rr);
    free(rt);
}

/* ---------------------------------------------------------------- splice */

/* stitch_cigar: splice.rs:435-494.  bounds: 3 u32 per exon (pseudo_start,
 * ref_start, ref_end); skips[k] = junctions[k].end - junctions[k].start. */
int espo_stitch_cigar(const uint32_t *cigar, size_t n_ops, uint32_t p_start,
                      const uint32_t *bounds, size_t n_exons, const uint32_t *skips,
                      size_t n_junctions, uint32_t *out, size_t cap, size_t *n_out) {
    opvec o = {out, 0, cap, 0};
    if (n_exons == 0) {
        for (size_t k = 0; k < n_ops; k++) {
            if (o.n < cap) out[o.n] = cigar[k];
            else o.overflow = 1;
            o.n++;
        }
        *n_out = o.n;
        return o.overflow ? 2 : 0;
    }
#define EXON_END(ei) ((uint64_t)bounds[3 * (ei)] + (uint64_t)(bounds[3 * (ei) + 2] - bounds[3 * (ei) + 1]))
    uint64_t p = p_start;
    size_t ei = 0;
    while (ei + 1 < n_exons && p >= EXON_END(ei)) ei++;
    for (size_t k = 0; k < n_ops; k++) {
        uint32_t kind = op_kind(cigar[k]), len = op_len(cigar[k]);
        if (kind == OP_I || kind == OP_S || kind == OP_N) {
            push_op(&o, kind, len);
            continue;
        }
        uint64_t remaining = len;
        while (remaining > 0) {
            if (ei >= n_exons) {
                push_op(&o, kind, (uint32_t)remaining);
                p += remaining;
                break;
            }
            if (p < bounds[3 * ei]) p = bounds[3 * ei];
            uint64_t ee = EXON_END(ei);
            uint64_t rie = ee > p ? ee - p : 0;
            uint64_t take = remaining < rie ? remaining : rie;
            if (take > 0) push_op(&o, kind, (uint32_t)take);
            p += take;
            remaining -= take;
            if (remaining > 0) {
                if (ei + 1 >= n_exons) {
                    push_op(&o, kind, (uint32_t)remaining);
                    p += remaining;
                    break;
                }
                if (ei < n_junctions) push_op(&o, OP_N, skips[ei]);
                ei++;
            }
        }
    }
#undef EXON_END
    *n_out = o.n;
    return o.overflow ? 2 : 0;
}

/* ------------------------------------------------------ reference windows */

/* Resident reference in the layout of fasta.rs:62-71: contig k has len[k]
 * bases, packed bytes at packed + packed_off[k] (base at pos in byte pos/4
 * at shift 6 - 2*(pos%4)), N intervals (iv[2j], iv[2j+1]) for
 * j in [iv_off[k], iv_off[k+1]). */
typedef struct {
    size_t ncontig;
    const uint32_t *len;
    const uint8_t *packed;
    const uint64_t *packed_off;
    const uint32_t *iv;
    const uint64_t *iv_off;
} espo_ref;

/* Contig::is_n: fasta.rs:119-135 (binary search per base). */
static int contig_is_n(const espo_ref *R, size_t k, uint32_t pos) {
    size_t lo = 0, hi = (size_t)(R->iv_off[k + 1] - R->iv_off[k]);
    const uint32_t *iv = R->iv + 2 * R->iv_off[k];
    while (lo < hi) {
        size_t mid = (lo + hi) / 2;
        if (pos < iv[2 * mid]) hi = mid;
        else if (pos >= iv[2 * mid + 1]) lo = mid + 1;
        else return 1;
    }
    return 0;
}

/* Contig::base + Base::to_ascii: fasta.rs:105-117, :138-146. */
static uint8_t contig_base(const espo_ref *R, size_t k, uint32_t pos) {
    if (contig_is_n(R, k, pos)) return 'N';
    uint8_t byte = R->packed[R->packed_off[k] + (pos >> 2)];
    unsigned shift = 6 - 2 * (pos & 3);
    static const uint8_t ascii[4] = {'A', 'C', 'G', 'T'};
    return ascii[(byte >> shift) & 3];
}

/* Base::from_ascii(b).complement().to_ascii(): fasta.rs:30-39, :147-156. */
static uint8_t comp_ascii(uint8_t b) {
    switch (b) {
    case 'A': case 'a': return 'T';
    case 'C': case 'c': return 'G';
    case 'G': case 'g': return 'C';
    case 'T': case 't': return 'A';
    default: return 'N';
    }
}

/* Contig::decode_into (fasta.rs:77-82), or revcomp_into of it
 * (align.rs:2525-2533) when rc != 0.  Returns 1 when out of range. */
int espo_decode(const espo_ref *R, uint32_t contig, uint32_t start, uint32_t end, int rc,
                uint8_t *out) {
    if (contig >= R->ncontig || start > end || end > R->len[contig]) return 1;
    uint32_t L = end - start;
    for (uint32_t k = 0; k < L; k++)
        out[k] = rc ? comp_ascii(contig_base(R, contig, end - 1 - k)) : contig_base(R, contig, start + k);
    return 0;
}

/* revcomp_into: align.rs:2525-2533. */
void espo_revcomp(const uint8_t *seq, size_t m, uint8_t *out) {
    for (size_t k = 0; k < m; k++) out[k] = comp_ascii(seq[m - 1 - k]);
}

typedef struct {
    int32_t placed, strand;
    uint32_t pos;
    int32_t score, dist;
    uint32_t wstart, wlen, read_start, read_end;
} espo_mate_rec;

/* mate_window_placement: align.rs:851-921.  placed = -1 for a contig index
 * out of range (the reference indexes its contig table and panics).  The
 * m <= 128 route (short::infix_best_start_into) equals the long one. */
int espo_mate_window(const espo_ref *R, const uint8_t *read, size_t m, uint32_t contig,
                     uint32_t mate_pos, const espo_params *p, espo_mate_rec *rec, uint32_t *cigar,
                     size_t cap, size_t *n_ops) {
    memset(rec, 0, sizeof(*rec));
    *n_ops = 0;
    if (m == 0 || m > 256) return 0;
    if (contig >= R->ncontig) {
        rec->placed = -1;
        return 0;
    }
    uint32_t read_len = (uint32_t)m;
    uint32_t clen = R->len[contig];
    uint32_t span = 500u + 500u / 2u + read_len; /* MATE_WINDOW_INSERT = 500 (align.rs:181) */
    uint32_t wstart = mate_pos > span ? mate_pos - span : 0u;
    uint64_t we = (uint64_t)mate_pos + span + read_len;
    if (we > 0xFFFFFFFFull) we = 0xFFFFFFFFull;
    uint32_t wend = (uint32_t)(we < clen ? we : clen);
    if (wend <= wstart || (size_t)(wend - wstart) < m) return 0;
    size_t n = wend - wstart;
    uint8_t *win = (uint8_t *)malloc(n), *rc = (uint8_t *)malloc(m);
    espo_decode(R, contig, wstart, wend, 0, win);
    espo_revcomp(read, m, rc);
    rec->wstart = wstart;
    rec->wlen = (uint32_t)n;
    int32_t threshold = (int32_t)read_len / 3;
    if (threshold < 10) threshold = 10;
    int best = 0; /* 0 none, 1 Plus, 2 Minus */
    int32_t best_d = 0;
    size_t best_start = 0;
    for (int strand = 1; strand <= 2; strand++) {
        const uint8_t *q = strand == 1 ? read : rc;
        int32_t d;
        size_t start;
        espo_myers_best_start(q, m, win, n, &d, &start);
        if (d > threshold || start + m > n) continue;
        if (best == 0 || d < best_d) {
            best = strand;
            best_d = d;
            best_start = start;
        }
    }
    int rc_code = 0;
    if (best) {
        espo_ext e;
        int32_t esc;
        size_t nops = 0;
        rc_code = espo_extend_hint(best == 1 ? read : rc, m, win, n, p, (int64_t)best_start, &e,
                                   cigar, cap, &nops, &esc);
        rec->strand = best == 2;
        rec->dist = best_d;
        rec->pos = wstart + e.ref_start;
        rec->score = e.score;
        rec->read_start = e.read_start;
        rec->read_end = e.read_end;
        rec->placed = e.score > 0;
        *n_ops = rec->placed ? nops : 0;
    }
    free(win);
    free(rc);
    return rc_code;
}

typedef struct {
    int32_t status;
    uint32_t genomic_start, pseudo_len, n_stitched;
} espo_splice_rec;

/* SplicedAlignment::genomic_start: splice.rs:414-425. */
static uint32_t genomic_start_of(uint32_t p, const uint32_t *bounds, size_t n_exons) {
    for (size_t e = 0; e < n_exons; e++) {
        uint32_t ps = bounds[3 * e], rs = bounds[3 * e + 1], re = bounds[3 * e + 2];
        if (p >= ps && p < ps + (re - rs)) return rs + (p - ps);
    }
    uint32_t ps = bounds[3 * (n_exons - 1)], rs = bounds[3 * (n_exons - 1) + 1];
    return rs + (p > ps ? p - ps : 0u);
}

/* Spliced window DP: the pseudo-reference concatenation of prepare_spliced
 * (splice.rs:589-608) / intron_chain.rs:497-505 over the given exons, then
 * align_spliced's extend_hint + stitch_cigar (splice.rs:625-664) with the
 * junction between adjacent exons.  status: 1 aligned, 0 an exon with
 * re <= rs (prepare_spliced returns None), -1 no exons or an exon outside
 * the contig, -2 stitched CIGAR capacity too small. */
int espo_spliced(const espo_ref *R, const uint8_t *read, size_t m, uint32_t contig,
                 const uint32_t *exons, size_t n_exons, int64_t hint, const espo_params *p,
                 espo_ext *ext, uint32_t *cigar, size_t cap, size_t *n_ops, uint32_t *stitched,
                 size_t scap, espo_splice_rec *rec) {
    memset(rec, 0, sizeof(*rec));
    memset(ext, 0, sizeof(*ext));
    *n_ops = 0;
    if (contig >= R->ncontig || n_exons == 0) {
        rec->status = -1;
        return 0;
    }
    uint64_t total = 0;
    int rejected = 0;
    for (size_t e = 0; e < n_exons; e++) {
        uint32_t rs = exons[2 * e], re = exons[2 * e + 1];
        if (re <= rs) rejected = 1;
        else if (re > R->len[contig]) { rec->status = -1; return 0; }
        else total += re - rs;
    }
    if (rejected) return 0;
    uint8_t *pseudo = (uint8_t *)malloc(total ? total : 1);
    uint32_t *bounds = (uint32_t *)malloc(3 * n_exons * sizeof(uint32_t));
    uint32_t *skips = (uint32_t *)malloc(n_exons * sizeof(uint32_t));
    uint32_t ps = 0;
    for (size_t e = 0; e < n_exons; e++) {
        uint32_t rs = exons[2 * e], re = exons[2 * e + 1];
        bounds[3 * e] = ps;
        bounds[3 * e + 1] = rs;
        bounds[3 * e + 2] = re;
        espo_decode(R, contig, rs, re, 0, pseudo + ps);
        ps += re - rs;
        if (e + 1 < n_exons) skips[e] = exons[2 * (e + 1)] - re;
    }
    int32_t esc;
    int rc = espo_extend_hint(read, m, pseudo, total, p, hint, ext, cigar, cap, n_ops, &esc);
    size_t ns = 0;
    int src = espo_stitch_cigar(cigar, *n_ops, ext->ref_start, bounds, n_exons, skips,
                                n_exons - 1, stitched, scap, &ns);
    rec->status = src ? -2 : 1;
    rec->n_stitched = (uint32_t)ns;
    rec->pseudo_len = (uint32_t)total;
    rec->genomic_start = genomic_start_of(ext->ref_start, bounds, n_exons);
    free(pseudo);
    free(bounds);
    free(skips);
    return rc;
}

typedef struct {
    const espo_ref *R;
    const espo_params *p;
    const uint8_t *reads;
    const uint64_t *read_off;
    const uint32_t *read_len;
    const uint32_t *contig;
    const uint32_t *mate_pos;
    const int64_t *hint;
    const uint32_t *exons;
    const uint64_t *exon_off;
    espo_mate_rec *mrec;
    espo_splice_rec *srec;
    espo_ext *ext;
    uint32_t *cigar;
    const uint64_t *cig_off;
    uint32_t *cig_n;
    uint32_t *stitched;
    const uint64_t *stitch_off;
    size_t lo, hi;
} window_job;

static void *mate_worker(void *arg) {
    window_job *J = (window_job *)arg;
    for (size_t k = J->lo; k < J->hi; k++) {
        size_t nops = 0;
        espo_mate_window(J->R, J->reads + J->read_off[k], J->read_len[k], J->contig[k],
                         J->mate_pos[k], J->p, &J->mrec[k], J->cigar + J->cig_off[k],
                         (size_t)(J->cig_off[k + 1] - J->cig_off[k]), &nops);
        J->cig_n[k] = (uint32_t)nops;
    }
    return NULL;
}

static void *splice_worker(void *arg) {
    window_job *J = (window_job *)arg;
    for (size_t k = J->lo; k < J->hi; k++) {
        size_t nops = 0;
        espo_spliced(J->R, J->reads + J->read_off[k], J->read_len[k], J->contig[k],
                     J->exons + 2 * J->exon_off[k], (size_t)(J->exon_off[k + 1] - J->exon_off[k]),
                     J->hint[k], J->p, &J->ext[k], J->cigar + J->cig_off[k],
                     (size_t)(J->cig_off[k + 1] - J->cig_off[k]), &nops,
                     J->stitched + J->stitch_off[k],
                     (size_t)(J->stitch_off[k + 1] - J->stitch_off[k]), &J->srec[k]);
        J->cig_n[k] = (uint32_t)nops;
    }
    return NULL;
}

static void run_window_jobs(window_job *tmpl, size_t npairs, int nthreads, void *(*fn)(void *)) {
    if (nthreads < 1) nthreads = 1;
    window_job *jobs = (window_job *)calloc((size_t)nthreads, sizeof(window_job));
    pthread_t *th = (pthread_t *)calloc((size_t)nthreads, sizeof(pthread_t));
    size_t chunk = (npairs + (size_t)nthreads - 1) / (size_t)nthreads;
    for (int t = 0; t < nthreads; t++) {
        jobs[t] = *tmpl;
        jobs[t].lo = (size_t)t * chunk < npairs ? (size_t)t * chunk : npairs;
        jobs[t].hi = jobs[t].lo + chunk < npairs ? jobs[t].lo + chunk : npairs;
        pthread_create(&th[t], NULL, fn, &jobs[t]);
    }
    for (int t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
    free(jobs);
    free(th);
}

/* mate_window_placement over a batch (nthreads POSIX threads). */
void espo_mate_batch(const espo_ref *R, size_t npairs, const uint8_t *reads,
                     const uint64_t *read_off, const uint32_t *read_len, const uint32_t *contig,
                     const uint32_t *mate_pos, const espo_params *p, espo_mate_rec *rec,
                     uint32_t *cigar, const uint64_t *cig_off, uint32_t *cig_n, int nthreads) {
    window_job J;
    memset(&J, 0, sizeof(J));
    J.R = R; J.p = p; J.reads = reads; J.read_off = read_off; J.read_len = read_len;
    J.contig = contig; J.mate_pos = mate_pos; J.mrec = rec; J.cigar = cigar; J.cig_off = cig_off;
    J.cig_n = cig_n;
    run_window_jobs(&J, npairs, nthreads, mate_worker);
}

/* Spliced window DP over a batch (nthreads POSIX threads). */
void espo_spliced_batch(const espo_ref *R, size_t npairs, const uint8_t *reads,
                        const uint64_t *read_off, const uint32_t *read_len, const uint32_t *contig,
                        const int64_t *hint, const uint32_t *exons, const uint64_t *exon_off,
                        const espo_params *p, espo_ext *ext, uint32_t *cigar,
                        const uint64_t *cig_off, uint32_t *cig_n, uint32_t *stitched,
                        const uint64_t *stitch_off, espo_splice_rec *rec, int nthreads) {
    window_job J;
    memset(&J, 0, sizeof(J));
    J.R = R; J.p = p; J.reads = reads; J.read_off = read_off; J.read_len = read_len;
    J.contig = contig; J.hint = hint; J.exons = exons; J.exon_off = exon_off; J.ext = ext;
    J.cigar = cigar; J.cig_off = cig_off; J.cig_n = cig_n; J.stitched = stitched;
    J.stitch_off = stitch_off; J.srec = rec;
    run_window_jobs(&J, npairs, nthreads, splice_worker);
}

/* ---------------------------------------------- unspliced chain extension */

static uint32_t sat_add32(uint32_t a, uint32_t b) { uint64_t s = (uint64_t)a + b; return s > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)s; }
static uint32_t sat_sub32(uint32_t a, uint32_t b) { return a > b ? a - b : 0u; }
static uint32_t min32(uint32_t a, uint32_t b) { return a < b ? a : b; }
static uint32_t max32(uint32_t a, uint32_t b) { return a > b ? a : b; }

/* continuous_extension_window: align.rs:33-55 (rs, re = Chain::ref_span,
 * qs, qe = Chain::read_span). */
void espo_ext_window(uint32_t rs, uint32_t re, uint32_t qs, uint32_t qe, uint32_t read_len,
                     uint32_t contig_len, uint32_t *start, uint32_t *end) {
    uint32_t flank = read_len / 2 + 30;
    uint32_t head = max32(flank, qs);
    uint32_t tail = max32(flank, sat_sub32(read_len, qe));
    uint32_t st = min32(sat_sub32(rs, head), contig_len);
    uint32_t en = min32(sat_add32(re, tail), contig_len);
    en = min32(en, sat_add32(sat_add32(rs, read_len), flank));
    uint32_t f2 = (uint32_t)((uint64_t)flank * 2 > 0xFFFFFFFFull ? 0xFFFFFFFFu : flank * 2);
    en = min32(en, sat_add32(sat_add32(st, read_len), f2));
    *start = st;
    *end = max32(en, st);
}

typedef struct {
    int32_t status; /* 1 EA lane, 2 banded DP, 0 none, -1 contig out of range */
    int32_t strand;
    uint32_t pos;
    int32_t score;
    int32_t dist;   /* EA-Myers distance (EA lane) */
    uint32_t wstart, wlen;
} espo_unspliced_rec;

/* EditDir::exempt: extend.rs:71-93 (uppercase bytes only). */
static int dir_exempt(int dir, uint8_t ref_b, uint8_t read_b) {
    int ed = exempt_dir(ref_b, read_b); /* 1 plus (A/G), 0 minus (T/C), -1 none */
    return (dir == 1 && ed == 1) || (dir == 2 && ed == 0);
}

/* extend_unspliced_chain after the fast lane (align.rs:1770-1919): the
 * EA-Myers window lane for reads of 1..256 bases (best start within
 * max(m/3, 10), per-base mismatch recount equal to the distance, no N, then
 * the Kadane-optimal soft clips when the kept segment covers half the read),
 * else, when allow_dp, extend_hint over the same window at the chain's
 * diagonal. */
int espo_unspliced(const espo_ref *R, const uint8_t *read, size_t m, uint32_t contig, int minus,
                   uint32_t rs, uint32_t re, uint32_t qs, uint32_t qe, int allow_dp,
                   const espo_params *p, espo_unspliced_rec *rec, uint32_t *cigar, size_t cap,
                   size_t *n_ops) {
    memset(rec, 0, sizeof(*rec));
    *n_ops = 0;
    rec->strand = minus ? 1 : 0;
    if (contig >= R->ncontig) {
        rec->status = -1;
        return 0;
    }
    uint32_t read_len = (uint32_t)m;
    uint32_t ws, we;
    espo_ext_window(rs, re, qs, qe, read_len, R->len[contig], &ws, &we);
    size_t n = we - ws;
    rec->wstart = ws;
    rec->wlen = (uint32_t)n;
    uint8_t *win = (uint8_t *)malloc(n ? n : 1), *q = (uint8_t *)malloc(m ? m : 1);
    espo_decode(R, contig, ws, we, 0, win);
    if (minus) espo_revcomp(read, m, q);
    else if (m) memcpy(q, read, m);
    int rc = 0;
    if (read_len > 0 && read_len <= 256 && n >= m) {
        int32_t threshold = (int32_t)read_len / 3;
        if (threshold < 10) threshold = 10;
        int32_t d;
        size_t start;
        espo_myers_best_start(q, m, win, n, &d, &start);
        int usable = d <= threshold && start + m <= n;
        if (usable) {
            int32_t mm = 0;
            int has_n = 0;
            for (size_t k = 0; k < m; k++) {
                uint8_t qb = q[k], rb = win[start + k];
                if (qb == 'N' || rb == 'N') { has_n = 1; break; }
                if (qb != rb && exempt_dir(rb, qb) < 0) mm++;
            }
            if (has_n || mm != d) usable = 0;
        }
        if (usable) {
            int32_t best_sum = 0, cur_sum = 0;
            size_t bi = 0, bj = 0, cur_i = 0;
            for (size_t k = 0; k < m; k++) {
                uint8_t qb = q[k], rb = win[start + k];
                int32_t sc = qb == rb ? p->match_score : (dir_exempt(p->edit_dir, rb, qb) ? 0 : p->mismatch);
                if (cur_sum <= 0) { cur_sum = sc; cur_i = k; }
                else cur_sum += sc;
                if (cur_sum > best_sum) { best_sum = cur_sum; bi = cur_i; bj = k + 1; }
            }
            uint32_t seg = (uint32_t)(bj - bi);
            if (best_sum > 0 && seg * 2 >= read_len) {
                size_t k = 0;
                if (bi > 0 && k < cap) cigar[k++] = ((uint32_t)bi << 4) | OP_S;
                if (k < cap) cigar[k++] = (seg << 4) | OP_M;
                if (m - bj > 0 && k < cap) cigar[k++] = ((uint32_t)(m - bj) << 4) | OP_S;
                *n_ops = k;
                rec->status = 1;
                rec->pos = ws + (uint32_t)(start + bi);
                rec->score = best_sum;
                rec->dist = d;
                free(win);
                free(q);
                return 0;
            }
        }
    }
    if (allow_dp) {
        espo_ext e;
        int32_t esc;
        size_t nops = 0;
        int64_t hint = (int64_t)(rs - ws) - (int64_t)qs;
        rc = espo_extend_hint(q, m, win, n, p, hint, &e, cigar, cap, &nops, &esc);
        if (e.score > 0) {
            rec->status = 2;
            rec->pos = ws + e.ref_start;
            rec->score = e.score;
            *n_ops = nops;
        }
    }
    free(win);
    free(q);
    return rc;
}

/* ------------------------------------------------- independent checkers */

/* Textbook local affine-gap DP (Gotoh 1982) over the cell set the band
 * defines: row i >= 1 holds columns [lo_i, hi_i] with c = clamp(offset+i,
 * 1, n), lo_i = max(c-half,1), hi_i = min(c+half,n).  A local alignment
 * may begin at any cell, so H outside the cell set is 0 while the gap states
 * E/F there are -inf.  Gap of length k costs open + k*ext.  Returns the
 * maximum H.  Written from the recurrence, not from run_band. */
int32_t espo_bf_local_affine(const uint8_t *read, size_t m, const uint8_t *rf, size_t n,
                             const espo_params *p, int has_band, int64_t offset, int64_t half) {
    if (m == 0 || n == 0) return 0;
    const int64_t NINF = -((int64_t)1 << 40);
    size_t W = n + 1;
    int64_t *H = (int64_t *)malloc((m + 1) * W * sizeof(int64_t));
    int64_t *E = (int64_t *)malloc((m + 1) * W * sizeof(int64_t));
    int64_t *F = (int64_t *)malloc((m + 1) * W * sizeof(int64_t));
    for (size_t k = 0; k < (m + 1) * W; k++) { H[k] = 0; E[k] = NINF; F[k] = NINF; }
    int64_t go = p->gap_open, ge = p->gap_ext, best = 0;
    if (!has_band) { offset = 0; half = (int64_t)n; }
    if (half < 0) half = 0;
    for (size_t i = 1; i <= m; i++) {
        int64_t c = offset + (int64_t)i;
        if (c < 1) c = 1;
        if (c > (int64_t)n) c = (int64_t)n;
        int64_t lo = c - half > 1 ? c - half : 1, hi = c + half < (int64_t)n ? c + half : (int64_t)n;
        int rcode = base_code(read[i - 1]);
        for (int64_t j = lo; j <= hi; j++) {
            int fcode = base_code(rf[j - 1]);
            int64_t s;
            if (rcode < 4 && rcode == fcode) s = p->match_score;
            else if (rcode < 4 && fcode < 4 && edit_exempt(p->edit_dir, "ACGT"[fcode], "ACGT"[rcode])) s = 0;
            else s = p->mismatch;
            size_t ij = i * W + (size_t)j, up = (i - 1) * W + (size_t)j, lf = ij - 1, dg = up - 1;
            int64_t e = NINF, f = NINF;
            e = H[up] - go - ge;
            if (E[up] > NINF && E[up] - ge > e) e = E[up] - ge;
            f = H[lf] - go - ge;
            if (F[lf] > NINF && F[lf] - ge > f) f = F[lf] - ge;
            int64_t h = 0;
            if (H[dg] + s > h) h = H[dg] + s;
            if (e > h) h = e;
            if (f > h) h = f;
            H[ij] = h; E[ij] = e; F[ij] = f;
            if (h > best) best = h;
        }
    }
    free(H); free(E); free(F);
    return (int32_t)best;
}

/* The same textbook DP held as whole matrices, read back for what the
 * band's row scan reports, from the values alone:
 *   best      the first maximum in row-major order (i ascending, then j),
 *             (0, 0) when no cell is positive;
 *   touch     the first row i whose H at the left window edge equals the
 *             row's maximum (left = 1, clamped = lo_i <= 1), or else whose
 *             H at a right edge below n does (left = 0, clamped = 0);
 * with diag_cap >= 0 a diagonal candidate above diag_cap is not taken
 * (the reference's i16 route, whose wrapped sum is negative there).
 * Written from the definitions, not from run_band. */
typedef struct {
    int32_t best;
    int64_t bi, bj;
    int32_t touched;
    int64_t row;
    int32_t left, clamped;
} espo_bf_rows_out;

void espo_bf_band_rows(const uint8_t *read, size_t m, const uint8_t *rf, size_t n,
                       const espo_params *p, int has_band, int64_t offset, int64_t half,
                       int64_t diag_cap, espo_bf_rows_out *out) {
    memset(out, 0, sizeof(*out));
    if (m == 0 || n == 0) return;
    const int64_t NINF = -((int64_t)1 << 40);
    size_t W = n + 1;
    int64_t *H = (int64_t *)calloc((m + 1) * W, sizeof(int64_t));
    int64_t *E = (int64_t *)malloc((m + 1) * W * sizeof(int64_t));
    int64_t *F = (int64_t *)malloc((m + 1) * W * sizeof(int64_t));
    int64_t *LO = (int64_t *)malloc((m + 1) * sizeof(int64_t));
    int64_t *HI = (int64_t *)malloc((m + 1) * sizeof(int64_t));
    for (size_t k = 0; k < (m + 1) * W; k++) { E[k] = NINF; F[k] = NINF; }
    int64_t go = p->gap_open, ge = p->gap_ext;
    if (!has_band) { offset = 0; half = (int64_t)n; }
    if (half < 0) half = 0;
    for (size_t i = 1; i <= m; i++) {
        int64_t c = offset + (int64_t)i;
        c = c < 1 ? 1 : (c > (int64_t)n ? (int64_t)n : c);
        LO[i] = c - half > 1 ? c - half : 1;
        HI[i] = c + half < (int64_t)n ? c + half : (int64_t)n;
        int rcode = base_code(read[i - 1]);
        for (int64_t j = LO[i]; j <= HI[i]; j++) {
            int fcode = base_code(rf[j - 1]);
            int64_t s;
            if (rcode < 4 && rcode == fcode) s = p->match_score;
            else if (rcode < 4 && fcode < 4 && edit_exempt(p->edit_dir, "ACGT"[fcode], "ACGT"[rcode])) s = 0;
            else s = p->mismatch;
            size_t ij = i * W + (size_t)j, up = (i - 1) * W + (size_t)j;
            int64_t e = H[up] - go - ge, f = H[ij - 1] - go - ge;
            if (E[up] > NINF && E[up] - ge > e) e = E[up] - ge;
            if (F[ij - 1] > NINF && F[ij - 1] - ge > f) f = F[ij - 1] - ge;
            int64_t d = H[up - 1] + s, h = 0;
            if (!(diag_cap >= 0 && d > diag_cap) && d > h) h = d;
            if (e > h) h = e;
            if (f > h) h = f;
            H[ij] = h; E[ij] = e; F[ij] = f;
        }
    }
    int64_t best = 0;
    for (size_t i = 1; i <= m; i++)
        for (int64_t j = LO[i]; j <= HI[i]; j++)
            if (H[i * W + (size_t)j] > best) { best = H[i * W + (size_t)j]; out->bi = (int64_t)i; out->bj = j; }
    out->best = (int32_t)best;
    for (size_t i = 1; i <= m && !out->touched; i++) {
        if (LO[i] > HI[i]) continue;
        int64_t rmax = NINF;
        for (int64_t j = LO[i]; j <= HI[i]; j++)
            if (H[i * W + (size_t)j] > rmax) rmax = H[i * W + (size_t)j];
        if (H[i * W + (size_t)LO[i]] == rmax) {
            out->touched = 1; out->row = (int64_t)i; out->left = 1; out->clamped = LO[i] <= 1;
        } else if (HI[i] < (int64_t)n && H[i * W + (size_t)HI[i]] == rmax) {
            out->touched = 1; out->row = (int64_t)i; out->left = 0; out->clamped = 0;
        }
    }
    free(H); free(E); free(F); free(LO); free(HI);
}

/* Textbook semi-global unit-cost edit distance under EA equality
 * (free text prefix and suffix): D[0][j] = 0, D[i][0] = i.  Returns
 * (min_j D[m][j], earliest j attaining it, counting j = 0 with D = m). */
void espo_bf_ea_edit(const uint8_t *read, size_t m, const uint8_t *text, size_t n, int32_t *dist,
                     size_t *end) {
    int32_t *prev = (int32_t *)malloc((m + 1) * sizeof(int32_t));
    int32_t *cur = (int32_t *)malloc((m + 1) * sizeof(int32_t));
    for (size_t i = 0; i <= m; i++) prev[i] = (int32_t)i; /* column j = 0 */
    int32_t best = (int32_t)m;
    size_t best_at = 0;
    for (size_t j = 1; j <= n; j++) {
        cur[0] = 0;
        for (size_t i = 1; i <= m; i++) {
            int32_t d = prev[i - 1] + (ea_matches(read[i - 1], text[j - 1]) ? 0 : 1);
            int32_t a = prev[i] + 1, b = cur[i - 1] + 1;
            if (a < d) d = a;
            if (b < d) d = b;
            cur[i] = d;
        }
        if (cur[m] < best) { best = cur[m]; best_at = j; }
        int32_t *t = prev; prev = cur; cur = t;
    }
    free(prev); free(cur);
    *dist = best;
    *end = best_at;
}

/* ---------------------------------------------------------------- batches */

/* Packed batch description shared by the batch helpers: sequences are
 * concatenated bytes addressed by 64-bit offsets. */
typedef struct {
    const uint8_t *reads, *wins;
    const uint64_t *read_off, *win_off;
    const uint32_t *read_len, *win_len;
    const int64_t *offset, *half; /* band mode */
    const espo_params *p;
    int mode; /* 0 band, 1 full, 2 extend_hint (offset = hint) */
    size_t lo, hi;
    espo_ext *out;
    espo_touch *touch;
    uint32_t *cigar;
    const uint64_t *cig_off;
    uint32_t *cig_n;
    int32_t *aux;
    int rc;
} band_job;

static void *band_worker(void *arg) {
    band_job *J = (band_job *)arg;
    for (size_t k = J->lo; k < J->hi; k++) {
        const uint8_t *r = J->reads + J->read_off[k], *w = J->wins + J->win_off[k];
        size_t cap = (size_t)(J->cig_off[k + 1] - J->cig_off[k]), nops = 0;
        uint32_t *cg = J->cigar + J->cig_off[k];
        espo_touch t = {0, 0, 0, 0};
        int rc;
        if (J->mode == 2) {
            int32_t esc = 0;
            rc = espo_extend_hint(r, J->read_len[k], w, J->win_len[k], J->p, J->offset[k],
                                  &J->out[k], cg, cap, &nops, &esc);
            if (J->aux) J->aux[k] = esc;
        } else {
            rc = espo_band(r, J->read_len[k], w, J->win_len[k], J->p, J->mode == 0,
                           J->mode == 0 ? J->offset[k] : 0, J->mode == 0 ? J->half[k] : 0,
                           &J->out[k], &t, cg, cap, &nops);
        }
        if (J->touch) J->touch[k] = t;
        J->cig_n[k] = (uint32_t)nops;
        if (rc) J->rc = rc;
    }
    return NULL;
}

/* Runs many pairs over nthreads POSIX threads (static chunks, one scratch
 * set per call, mirroring the reference's per-worker ExtendBuffer). */
int espo_band_batch(size_t npairs, const uint8_t *reads, const uint64_t *read_off,
                    const uint32_t *read_len, const uint8_t *wins, const uint64_t *win_off,
                    const uint32_t *win_len, const int64_t *offset, const int64_t *half,
                    const espo_params *p, int mode, espo_ext *out, espo_touch *touch,
                    uint32_t *cigar, const uint64_t *cig_off, uint32_t *cig_n, int32_t *aux,
                    int nthreads) {
    if (nthreads < 1) nthreads = 1;
    band_job *jobs = (band_job *)calloc((size_t)nthreads, sizeof(band_job));
    pthread_t *th = (pthread_t *)calloc((size_t)nthreads, sizeof(pthread_t));
    size_t chunk = (npairs + (size_t)nthreads - 1) / (size_t)nthreads;
    for (int t = 0; t < nthreads; t++) {
        band_job *J = &jobs[t];
        J->reads = reads; J->wins = wins; J->read_off = read_off; J->win_off = win_off;
        J->read_len = read_len; J->win_len = win_len; J->offset = offset; J->half = half;
        J->p = p; J->mode = mode; J->out = out; J->touch = touch; J->cigar = cigar;
        J->cig_off = cig_off; J->cig_n = cig_n; J->aux = aux;
        J->lo = (size_t)t * chunk < npairs ? (size_t)t * chunk : npairs;
        J->hi = J->lo + chunk < npairs ? J->lo + chunk : npairs;
        pthread_create(&th[t], NULL, band_worker, J);
    }
    int rc = 0;
    for (int t = 0; t < nthreads; t++) {
        pthread_join(th[t], NULL);
        if (jobs[t].rc) rc = jobs[t].rc;
    }
    free(jobs);
    free(th);
    return rc;
}

typedef struct {
    const uint8_t *reads, *texts;
    const uint64_t *read_off, *text_off;
    const uint32_t *read_len, *text_len;
    size_t lo, hi;
    int32_t *dist;
    uint64_t *end, *start;
} myers_job;

static void *myers_worker(void *arg) {
    myers_job *J = (myers_job *)arg;
    for (size_t k = J->lo; k < J->hi; k++) {
        const uint8_t *r = J->reads + J->read_off[k], *t = J->texts + J->text_off[k];
        size_t e = 0, s = 0;
        int32_t d = 0, d2 = 0;
        espo_myers_best_end(r, J->read_len[k], t, J->text_len[k], &d, &e);
        J->dist[k] = d;
        J->end[k] = e;
        if (J->start) {
            espo_myers_best_start(r, J->read_len[k], t, J->text_len[k], &d2, &s);
            J->start[k] = s;
        }
    }
    return NULL;
}

/* long::infix_best_end (+ infix_best_start when start != NULL) over a batch. */
int espo_myers_batch(size_t npairs, const uint8_t *reads, const uint64_t *read_off,
                     const uint32_t *read_len, const uint8_t *texts, const uint64_t *text_off,
                     const uint32_t *text_len, int32_t *dist, uint64_t *end, uint64_t *start,
                     int nthreads) {
    if (nthreads < 1) nthreads = 1;
    myers_job *jobs = (myers_job *)calloc((size_t)nthreads, sizeof(myers_job));
    pthread_t *th = (pthread_t *)calloc((size_t)nthreads, sizeof(pthread_t));
    size_t chunk = (npairs + (size_t)nthreads - 1) / (size_t)nthreads;
    for (int t = 0; t < nthreads; t++) {
        myers_job *J = &jobs[t];
        J->reads = reads; J->texts = texts; J->read_off = read_off; J->text_off = text_off;
        J->read_len = read_len; J->text_len = text_len; J->dist = dist; J->end = end;
        J->start = start;
        J->lo = (size_t)t * chunk < npairs ? (size_t)t * chunk : npairs;
        J->hi = J->lo + chunk < npairs ? J->lo + chunk : npairs;
        pthread_create(&th[t], NULL, myers_worker, J);
    }
    for (int t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
    free(jobs);
    free(th);
    return 0;
}

/* ====================================================== placement walks */

/* SpliceParams fields used by rescore_with (splice.rs:21-52); defaults
 * known 6, GT-AG 4, GC-AG 2, AT-AC 1, non-canonical 8. */
typedef struct {
    int32_t known_bonus;
    int32_t gt_ag_bonus;
    int32_t gc_ag_bonus;
    int32_t at_ac_bonus;
    int32_t noncanonical_penalty;
} espo_splice_params;

/* StrictScore (align.rs:188-211). */
typedef struct {
    int32_t score;
    uint32_t aligned, matches, mm, ea_ag, ea_tc, ins, del, clipped, n_splice;
    int32_t splice_score;
} espo_strict;

/* intron_length_penalty: splice.rs:119-121, 8 x |log10(max(len,1)) - 3| as
 * i32 (f64 arithmetic, truncation toward zero). */
int32_t espo_intron_length_penalty(uint32_t len) {
    double v = 8.0 * fabs(log10((double)(len > 1 ? len : 1)) - 3.0);
    return (int32_t)v;
}

static uint8_t upper(uint8_t b) { return (b >= 'a' && b <= 'z') ? (uint8_t)(b - 32) : b; }

/* splice_signal(donor, acceptor, minus).score(sp): splice.rs:61-103. */
int32_t espo_splice_signal_score(uint8_t d0, uint8_t d1, uint8_t a0, uint8_t a1, int minus,
                                 const espo_splice_params *sp) {
    d0 = upper(d0); d1 = upper(d1); a0 = upper(a0); a1 = upper(a1);
    int sig; /* 0 GT-AG, 1 GC-AG, 2 AT-AC, 3 non-canonical */
    if (!minus) {
        if (d0 == 'G' && d1 == 'T' && a0 == 'A' && a1 == 'G') sig = 0;
        else if (d0 == 'G' && d1 == 'C' && a0 == 'A' && a1 == 'G') sig = 1;
        else if (d0 == 'A' && d1 == 'T' && a0 == 'A' && a1 == 'C') sig = 2;
        else sig = 3;
    } else if (d0 == 'C' && d1 == 'T' && a0 == 'A' && a1 == 'C') sig = 0;
    else if (d0 == 'C' && d1 == 'T' && a0 == 'G' && a1 == 'C') sig = 1;
    else if (d0 == 'G' && d1 == 'T' && a0 == 'A' && a1 == 'T') sig = 2;
    else sig = 3;
    switch (sig) {
    case 0: return sp->gt_ag_bonus;
    case 1: return sp->gc_ag_bonus;
    case 2: return sp->at_ac_bonus;
    default: return -sp->noncanonical_penalty;
    }
}

/* JunctionLib::contains_any_strand (gtf.rs:165-169): binary search of
 * (contig, start, end) in the sorted triples jn[3j .. 3j+2]. */
static int jlib_contains(const uint32_t *jn, size_t nj, uint32_t contig, uint32_t start, uint32_t end) {
    size_t lo = 0, hi = nj;
    while (lo < hi) {
        size_t mid = (lo + hi) / 2;
        const uint32_t *t = jn + 3 * mid;
        int c = t[0] != contig ? (t[0] < contig ? -1 : 1)
              : t[1] != start  ? (t[1] < start ? -1 : 1)
              : t[2] != end    ? (t[2] < end ? -1 : 1) : 0;
        if (c == 0) return 1;
        if (c < 0) lo = mid + 1;
        else hi = mid;
    }
    return 0;
}

/* Reference byte at pos (N past the contig end): the decode_append reads of
 * the walks, clamped at contig.len (align.rs:1082-1086, :2487-2492). */
static uint8_t ref_byte(const espo_ref *R, uint32_t contig, uint32_t pos) {
    return pos < R->len[contig] ? contig_base(R, contig, pos) : (uint8_t)'N';
}

/* Aligner::rescore_with (align.rs:1060-1163).  query is in the alignment's
 * frame; bytes past its end read as N.  dir is the EditDir exempted. */
int espo_rescore(const espo_ref *R, const uint8_t *query, size_t qlen, uint32_t contig,
                 uint32_t pos, const uint32_t *cigar, size_t n_ops, int dir, const espo_params *ep,
                 const espo_splice_params *sp, const uint32_t *jn, size_t nj, espo_strict *out) {
    if (contig >= R->ncontig) return 1;
    espo_strict s;
    memset(&s, 0, sizeof s);
    int32_t score = 0;
    uint32_t ro = 0, rf = pos;
    for (size_t o = 0; o < n_ops; o++) {
        const uint32_t op = cigar[o] & 15u, n = cigar[o] >> 4;
        if (op == OP_M) {
            for (uint32_t k = 0; k < n; k++) {
                const uint8_t qb = (size_t)ro + k < qlen ? query[ro + k] : (uint8_t)'N';
                const uint8_t rb = ref_byte(R, contig, rf + k);
                if (qb == 'N' || rb == 'N') continue; /* 0 score, uncounted */
                s.aligned++;
                if (qb == rb) {
                    score += ep->match_score;
                    s.matches++;
                } else {
                    if (!edit_exempt(dir, rb, qb)) score += ep->mismatch;
                    const int e = exempt_dir(rb, qb);
                    if (e == 1) s.ea_ag++;
                    else if (e == 0) s.ea_tc++;
                    else s.mm++;
                }
            }
            ro += n;
            rf += n;
        } else if (op == OP_I) {
            score -= ep->gap_open + ep->gap_ext * (int32_t)n;
            s.aligned += n;
            s.ins++;
            ro += n;
        } else if (op == OP_D) {
            score -= ep->gap_open + ep->gap_ext * (int32_t)n;
            s.aligned += n;
            s.del++;
            rf += n;
        } else if (op == OP_S) {
            s.clipped += n;
            ro += n;
        } else if (op == OP_N) {
            const int32_t before = score;
            if (jlib_contains(jn, nj, contig, rf, rf + n)) {
                score += sp->known_bonus;
            } else {
                const uint32_t rf1 = rf == 0xFFFFFFFFu ? rf : rf + 1; /* saturating_add */
                const uint8_t d0 = ref_byte(R, contig, rf), d1 = ref_byte(R, contig, rf1);
                const uint8_t a0 = ref_byte(R, contig, rf + (n >= 2 ? n - 2 : 0));
                const uint8_t a1 = ref_byte(R, contig, rf + (n >= 1 ? n - 1 : 0));
                const int32_t sp_plus = espo_splice_signal_score(d0, d1, a0, a1, 0, sp);
                const int32_t sp_minus = espo_splice_signal_score(d0, d1, a0, a1, 1, sp);
                score += (sp_plus > sp_minus ? sp_plus : sp_minus) - espo_intron_length_penalty(n);
            }
            s.n_splice++;
            s.splice_score += score - before;
            rf += n;
        }
    }
    s.score = score;
    *out = s;
    return 0;
}

/* recount_mm_ea_in (align.rs:2473-2516): (mm, ea_ag, ea_tc) over the M
 * columns where neither byte is N and they differ. */
int espo_recount(const espo_ref *R, const uint8_t *query, size_t qlen, uint32_t contig, uint32_t pos,
                 const uint32_t *cigar, size_t n_ops, uint32_t out[3]) {
    if (contig >= R->ncontig) return 1;
    uint32_t ro = 0, rf = pos, mm = 0, ag = 0, tc = 0;
    for (size_t o = 0; o < n_ops; o++) {
        const uint32_t op = cigar[o] & 15u, n = cigar[o] >> 4;
        if (op == OP_M) {
            for (uint32_t k = 0; k < n; k++) {
                const uint8_t qb = (size_t)ro + k < qlen ? query[ro + k] : (uint8_t)'N';
                const uint8_t rb = ref_byte(R, contig, rf + k);
                if (qb != 'N' && rb != 'N' && qb != rb) {
                    const int e = exempt_dir(rb, qb);
                    if (e == 1) ag++;
                    else if (e == 0) tc++;
                    else mm++;
                }
            }
            ro += n;
            rf += n;
        } else if (op == OP_I || op == OP_S) {
            ro += n;
        } else if (op == OP_D || op == OP_N) {
            rf += n;
        }
    }
    out[0] = mm;
    out[1] = ag;
    out[2] = tc;
    return 0;
}

/* Aligner::exact_kmer_support (align.rs:1231-1288): non-overlapping exact
 * k-mers along the M blocks, scanned from each block's start (a k-mer with
 * a bad column moves the scan one base on), exact in the reduced alphabet of
 * dir (the placement's ea_dir). */
uint32_t espo_kmer_support(const espo_ref *R, const uint8_t *query, size_t qlen, uint32_t contig,
                           uint32_t pos, const uint32_t *cigar, size_t n_ops, int dir, uint32_t k) {
    if (k == 0 || qlen == 0 || contig >= R->ncontig) return 0;
    uint32_t count = 0, ro = 0, rf = pos;
    for (size_t o = 0; o < n_ops; o++) {
        const uint32_t op = cigar[o] & 15u, n = cigar[o] >> 4;
        if (op == OP_M) {
            uint32_t i = 0;
            while ((uint64_t)i + k <= n) {
                int ok = 1;
                for (uint32_t t = 0; t < k; t++) {
                    const uint8_t qb = (size_t)ro + i + t < qlen ? query[ro + i + t] : (uint8_t)'N';
                    const uint8_t rb = ref_byte(R, contig, rf + i + t);
                    if (qb != rb && !edit_exempt(dir, rb, qb)) {
                        ok = 0;
                        break;
                    }
                }
                if (ok) {
                    count++;
                    i += k;
                } else {
                    i++;
                }
            }
            ro += n;
            rf += n;
        } else if (op == OP_I || op == OP_S) {
            ro += n;
        } else if (op == OP_D || op == OP_N) {
            rf += n;
        }
    }
    return count;
}

/* The three walks over a batch of placements (nthreads POSIX threads): the
 * CPU form of the device walk, used as the checker and the CPU baseline.
 * counts[4k..4k+3] = (mm, ea_ag, ea_tc, k-mers). */
typedef struct {
    const espo_ref *R;
    const uint8_t *reads;
    const uint64_t *read_off, *cig_off;
    const uint32_t *read_len, *contig, *pos, *cigar, *cig_n;
    const uint8_t *dir;
    const espo_params *ep;
    const espo_splice_params *sp;
    const uint32_t *jn;
    size_t nj, lo, hi;
    uint32_t k;
    espo_strict *strict;
    uint32_t *counts;
    int32_t *status;
} walk_job;

static void *walk_worker(void *arg) {
    walk_job *J = (walk_job *)arg;
    for (size_t i = J->lo; i < J->hi; i++) {
        const uint8_t *q = J->reads + J->read_off[i];
        const uint32_t *c = J->cigar + J->cig_off[i];
        const int d = J->dir ? J->dir[i] : 0;
        espo_strict s;
        memset(&s, 0, sizeof s);
        int bad = espo_rescore(J->R, q, J->read_len[i], J->contig[i], J->pos[i], c, J->cig_n[i], d, J->ep,
                               J->sp, J->jn, J->nj, &s);
        J->strict[i] = s;
        J->status[i] = bad;
        uint32_t *o = J->counts + 4 * i;
        o[0] = s.mm;
        o[1] = s.ea_ag;
        o[2] = s.ea_tc;
        o[3] = bad ? 0 : espo_kmer_support(J->R, q, J->read_len[i], J->contig[i], J->pos[i], c, J->cig_n[i],
                                           d, J->k);
    }
    return NULL;
}

void espo_walk_batch(const espo_ref *R, size_t n, const uint8_t *reads, const uint64_t *read_off,
                     const uint32_t *read_len, const uint32_t *contig, const uint32_t *pos,
                     const uint32_t *cigar, const uint64_t *cig_off, const uint32_t *cig_n,
                     const uint8_t *dir, const espo_params *ep, const espo_splice_params *sp,
                     const uint32_t *jn, size_t nj, uint32_t k, espo_strict *strict, uint32_t *counts,
                     int32_t *status, int nthreads) {
    if (nthreads < 1) nthreads = 1;
    walk_job *jobs = (walk_job *)calloc((size_t)nthreads, sizeof(walk_job));
    pthread_t *th = (pthread_t *)calloc((size_t)nthreads, sizeof(pthread_t));
    const size_t chunk = (n + (size_t)nthreads - 1) / (size_t)nthreads;
    for (int t = 0; t < nthreads; t++) {
        walk_job *J = &jobs[t];
        J->R = R; J->reads = reads; J->read_off = read_off; J->read_len = read_len;
        J->contig = contig; J->pos = pos; J->cigar = cigar; J->cig_off = cig_off; J->cig_n = cig_n;
        J->dir = dir; J->ep = ep; J->sp = sp; J->jn = jn; J->nj = nj; J->k = k;
        J->strict = strict; J->counts = counts; J->status = status;
        J->lo = (size_t)t * chunk < n ? (size_t)t * chunk : n;
        J->hi = J->lo + chunk < n ? J->lo + chunk : n;
        pthread_create(&th[t], NULL, walk_worker, J);
    }
    for (int t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
    free(jobs);
    free(th);
}

/* ==================================================== Kadane fast lane */

/* try_fast_lane's result (align.rs:2274-2401): status 1 with the Extension
 * (score, ref_start, read_start, read_end, CIGAR of at most three ops) and
 * the window start wstart, 0 when the lane refuses. */
typedef struct {
    int32_t status;
    int32_t score;
    uint32_t wstart, ref_start, read_start, read_end;
    uint32_t n_ops;
    uint32_t cigar[3];
} espo_fast;

/* Aligner::try_fast_lane (align.rs:2274-2401).  anchors holds the chain's
 * (qpos, rpos) pairs in its (rpos, qpos) order; query is the read in the
 * chain's orientation; k and w are SeedParams k and w. */
int espo_fast_lane(const espo_ref *R, const uint8_t *query, uint32_t read_len, uint32_t contig,
                   const uint32_t *anchors, size_t n_anchors, uint32_t k, uint32_t w, const espo_params *ep,
                   espo_fast *out) {
    memset(out, 0, sizeof *out);
    if (contig >= R->ncontig) return 1;
    if (n_anchors == 0) return 0; /* anchors.first()? */
    const int64_t diag = (int64_t)anchors[1] - (int64_t)anchors[0];
    for (size_t a = 0; a < n_anchors; a++)
        if ((int64_t)anchors[2 * a + 1] - (int64_t)anchors[2 * a] != diag) return 0;
    const uint32_t max_gap = k + w + 5;
    for (size_t a = 1; a < n_anchors; a++)
        if (anchors[2 * a] - anchors[2 * (a - 1)] > max_gap) return 0; /* u32 difference */
    const uint32_t qstart = anchors[0], qend = anchors[2 * (n_anchors - 1)] + k;
    if (qstart > 2 || read_len - qend > 2) return 0;
    const uint32_t rs = anchors[1];
    const uint32_t flank = read_len / 2 + 30;
    const uint32_t wstart = rs > flank ? rs - flank : 0;
    const uint32_t clen = R->len[contig];
    int32_t best_score = 0, cur_score = 0;
    size_t best_start = 0, best_end = 0, cur_start = 0;
    uint32_t cur_mm = 0, best_mm = 0;
    for (size_t i = 0; i < read_len; i++) {
        const int64_t rpos = (int64_t)i + diag;
        int32_t s = 0;
        int is_mm = 0;
        if (rpos >= 0 && rpos < (int64_t)clen) {
            const uint8_t qb = query[i], rb = contig_base(R, contig, (uint32_t)rpos);
            if (qb == 'N' || rb == 'N') s = 0;
            else if (qb == rb) s = ep->match_score;
            else if (edit_exempt(ep->edit_dir, rb, qb)) s = 0;
            else {
                s = ep->mismatch;
                is_mm = 1;
            }
        }
        if (cur_score <= 0) {
            cur_score = s;
            cur_start = i;
            cur_mm = (uint32_t)is_mm;
        } else {
            cur_score += s;
            cur_mm += (uint32_t)is_mm;
        }
        if (cur_score > best_score) {
            best_score = cur_score;
            best_start = cur_start;
            best_end = i + 1;
            best_mm = cur_mm;
        }
    }
    if (best_score <= 0 || best_mm > 0) return 0;
    const uint32_t read_start = (uint32_t)best_start, read_end = (uint32_t)best_end;
    uint32_t n = 0;
    if (read_start > 0) out->cigar[n++] = read_start << 4 | OP_S;
    out->cigar[n++] = (read_end - read_start) << 4 | OP_M;
    if (read_len - read_end > 0) out->cigar[n++] = (read_len - read_end) << 4 | OP_S;
    out->status = 1;
    out->score = best_score;
    out->wstart = wstart;
    out->ref_start = (uint32_t)((int64_t)read_start + diag - (int64_t)wstart);
    out->read_start = read_start;
    out->read_end = read_end;
    out->n_ops = n;
    return 0;
}

/* ===================================== engine EA-Myers (128-bit blocks) */

typedef unsigned __int128 u128;

/* code5 (engine/src/myers.rs:19-28). */
static int eng_code5(uint8_t b) {
    switch (b) {
    case 'A': case 'a': return 0;
    case 'C': case 'c': return 1;
    case 'G': case 'g': return 2;
    case 'T': case 't': return 3;
    default: return 4;
    }
}

typedef struct {
    size_t m, blocks;
    u128 peq[2][5], init_pv[2], top[2];
} eng_myers;

/* MyersEa::new (myers.rs:77-104). */
static void eng_new(eng_myers *s, const uint8_t *pat, size_t m) {
    memset(s, 0, sizeof *s);
    s->m = m;
    s->blocks = (m + 127) / 128;
    for (size_t i = 0; i < m; i++) s->peq[i / 128][eng_code5(pat[i])] |= (u128)1 << (i % 128);
    const size_t b0 = m < 128 ? m : 128;
    s->init_pv[0] = b0 == 128 ? ~(u128)0 : (((u128)1 << b0) - 1);
    s->top[0] = (u128)1 << (b0 - 1);
    if (s->blocks == 2) {
        const size_t b1 = m - 128;
        s->init_pv[1] = b1 == 128 ? ~(u128)0 : (((u128)1 << b1) - 1);
        s->top[1] = (u128)1 << (b1 - 1);
    }
}

/* eq_vec (myers.rs:109-120): text A also pairs with pattern G, text T with
 * pattern C; text N never matches. */
static u128 eng_eq(const eng_myers *s, size_t blk, int tcode) {
    if (tcode > 3) return 0;
    u128 eq = s->peq[blk][tcode];
    if (tcode == 0) eq |= s->peq[blk][2];
    else if (tcode == 3) eq |= s->peq[blk][1];
    return eq;
}

/* calculate_block (myers.rs:180-208): returns hout, writes the adder carry. */
static int eng_block(u128 eq, u128 *pv, u128 *mv, int hin, int cin, u128 top, int *cout) {
    const u128 xv = eq | *mv;
    const u128 a = eq & *pv;
    u128 sum = a + *pv;
    int c = sum < a;
    const u128 sum2 = sum + (u128)(cin ? 1 : 0);
    c |= sum2 < sum;
    sum = sum2;
    *cout = c;
    const u128 xh = (sum ^ *pv) | xv;
    u128 ph = *mv | ~(xh | *pv);
    u128 mh = *pv & xh;
    int hout = 0;
    if (ph & top) hout = 1;
    else if (mh & top) hout = -1;
    ph <<= 1;
    mh <<= 1;
    if (hin == 1) ph |= 1;
    else if (hin == -1) mh |= 1;
    *pv = mh | ~(xv | ph);
    *mv = ph & xv;
    return hout;
}

/* myers::infix / long::infix (infix_min, myers.rs:123-146) when global == 0,
 * myers::global / long::global (myers.rs:149-168) when global == 1.
 * Returns -1 for a pattern length outside 1..=256 (the reference asserts). */
int64_t espo_engine_myers(const uint8_t *pat, size_t m, const uint8_t *text, size_t n, int global) {
    if (m == 0 || m > 256) return -1;
    eng_myers s;
    eng_new(&s, pat, m);
    u128 pv[2] = {s.init_pv[0], s.init_pv[1]}, mv[2] = {0, 0};
    int32_t score = (int32_t)m;
    uint32_t best = (uint32_t)m;
    for (size_t j = 0; j < n; j++) {
        const int tcode = eng_code5(text[j]);
        int hin = global ? 1 : 0, cin = 0;
        for (size_t blk = 0; blk < s.blocks; blk++) {
            int cout = 0;
            const int h = eng_block(eng_eq(&s, blk, tcode), &pv[blk], &mv[blk], hin, cin, s.top[blk], &cout);
            hin = h;
            cin = cout;
        }
        score += hin;
        if (!global && score >= 0 && (uint32_t)score < best) best = (uint32_t)score;
    }
    if (global) return score > 0 ? score : 0;
    return best;
}

/* ================================ alignment-path checker (test invariants) */

/* The affine score of the alignment path an Extension reports, scored as the
 * extension DP scores it: SubstLut over the M columns (extend.rs:161-186),
 * gap_open + gap_ext x len per I or D run (extend.rs:450-459).  bad[k] is 1
 * when the path does not fit the pair: soft clips other than read_start
 * at the front and m - read_end at the back, M/I consuming other than
 * [read_start, read_end), or M/D leaving the window from ref_start. */
typedef struct {
    const uint8_t *reads, *wins;
    const uint64_t *read_off, *win_off, *cig_off;
    const uint32_t *read_len, *win_len, *cigar, *cig_n;
    const espo_ext *ext;
    const espo_params *p;
    size_t lo, hi;
    int64_t *score;
    int32_t *bad;
} path_job;

static void *path_worker(void *arg) {
    path_job *J = (path_job *)arg;
    core_params core;
    core_new(J->p, &core);
    for (size_t k = J->lo; k < J->hi; k++) {
        const uint8_t *r = J->reads + J->read_off[k], *w = J->wins + J->win_off[k];
        const uint32_t m = J->read_len[k], n = J->win_len[k];
        const uint32_t *cg = J->cigar + J->cig_off[k];
        const uint32_t nops = J->cig_n[k];
        const espo_ext e = J->ext[k];
        int64_t s = 0;
        int bad = 0;
        uint64_t q = 0, t = e.ref_start;
        for (uint32_t o = 0; o < nops; o++) {
            const uint32_t op = cg[o] & 15u, len = cg[o] >> 4;
            if (op == OP_S) {
                if (o == 0) {
                    if (len != e.read_start) bad = 1;
                    q += len;
                } else if (o == nops - 1) {
                    if (q != e.read_end || len != m - e.read_end) bad = 1;
                    q += len;
                } else {
                    bad = 1;
                }
            } else if (op == OP_M) {
                if (o == 0 && e.read_start != 0) bad = 1;
                for (uint32_t x = 0; x < len && !bad; x++) {
                    if (q + x >= m || t + x >= n) {
                        bad = 1;
                        break;
                    }
                    s += core.lut[base_code(r[q + x]) * 5 + base_code(w[t + x])];
                }
                q += len;
                t += len;
            } else if (op == OP_I || op == OP_D) {
                s -= (int64_t)J->p->gap_open + (int64_t)J->p->gap_ext * len;
                if (op == OP_I) q += len;
                else t += len;
            } else {
                bad = 1;
            }
        }
        if (nops && (cg[nops - 1] & 15u) != OP_S && q != e.read_end) bad = 1;
        if (q != (nops ? m : 0) && !(nops == 0 && e.read_start == e.read_end)) bad = 1;
        if (t > n) bad = 1;
        J->score[k] = s;
        J->bad[k] = bad;
    }
    return NULL;
}

void espo_path_scores(size_t npairs, const uint8_t *reads, const uint64_t *read_off, const uint32_t *read_len,
                      const uint8_t *wins, const uint64_t *win_off, const uint32_t *win_len, const espo_ext *ext,
                      const uint32_t *cigar, const uint64_t *cig_off, const uint32_t *cig_n,
                      const espo_params *p, int64_t *score, int32_t *bad, int nthreads) {
    if (nthreads < 1) nthreads = 1;
    path_job *jobs = (path_job *)calloc((size_t)nthreads, sizeof(path_job));
    pthread_t *th = (pthread_t *)calloc((size_t)nthreads, sizeof(pthread_t));
    const size_t chunk = (npairs + (size_t)nthreads - 1) / (size_t)nthreads;
    for (int t = 0; t < nthreads; t++) {
        path_job *J = &jobs[t];
        J->reads = reads; J->wins = wins; J->read_off = read_off; J->win_off = win_off; J->cig_off = cig_off;
        J->read_len = read_len; J->win_len = win_len; J->cigar = cigar; J->cig_n = cig_n; J->ext = ext;
        J->p = p; J->score = score; J->bad = bad;
        J->lo = (size_t)t * chunk < npairs ? (size_t)t * chunk : npairs;
        J->hi = J->lo + chunk < npairs ? J->lo + chunk : npairs;
        pthread_create(&th[t], NULL, path_worker, J);
    }
    for (int t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
    free(jobs);
    free(th);
}

/* ------------------------------------------- junction refinement / re-extension */

/* SpliceParams (splice.rs:21-53). */
typedef struct {
    uint32_t min_intron, max_intron;
    int32_t gt_ag_bonus, gc_ag_bonus, at_ac_bonus, noncanonical_penalty, known_bonus;
    uint32_t refine_radius;
} espo_splice_cfg;

/* JunctionLib::build + support (gtf.rs:84-149): junctions sorted by
 * (contig, start, end, minus) with identical ones merged into counts. */
typedef struct {
    uint32_t *q; /* 4 per junction */
    uint32_t *count;
    size_t n;
} espo_jlib;

static int quad_cmp(const void *a, const void *b) {
    const uint32_t *x = (const uint32_t *)a, *y = (const uint32_t *)b;
    for (int k = 0; k < 4; k++)
        if (x[k] != y[k]) return x[k] < y[k] ? -1 : 1;
    return 0;
}

espo_jlib *espo_jlib_build(const uint32_t *quads, size_t n) {
    espo_jlib *L = (espo_jlib *)calloc(1, sizeof(espo_jlib));
    uint32_t *tmp = (uint32_t *)malloc((n ? n : 1) * 4 * sizeof(uint32_t));
    for (size_t k = 0; k < n; k++) {
        memcpy(tmp + 4 * k, quads + 4 * k, 4 * sizeof(uint32_t));
        tmp[4 * k + 3] = tmp[4 * k + 3] ? 1u : 0u;
    }
    qsort(tmp, n, 4 * sizeof(uint32_t), quad_cmp);
    L->q = (uint32_t *)malloc((n ? n : 1) * 4 * sizeof(uint32_t));
    L->count = (uint32_t *)malloc((n ? n : 1) * sizeof(uint32_t));
    for (size_t k = 0; k < n; k++) {
        if (L->n && quad_cmp(L->q + 4 * (L->n - 1), tmp + 4 * k) == 0) {
            L->count[L->n - 1]++;
        } else {
            memcpy(L->q + 4 * L->n, tmp + 4 * k, 4 * sizeof(uint32_t));
            L->count[L->n++] = 1;
        }
    }
    free(tmp);
    return L;
}

void espo_jlib_free(espo_jlib *L) {
    if (!L) return;
    free(L->q);
    free(L->count);
    free(L);
}

uint32_t espo_jlib_support(const espo_jlib *L, uint32_t contig, uint32_t start, uint32_t end, uint32_t minus) {
    if (!L) return 0;
    const uint32_t key[4] = {contig, start, end, minus ? 1u : 0u};
    size_t lo = 0, hi = L->n;
    while (lo < hi) {
        size_t mid = (lo + hi) / 2;
        int c = quad_cmp(L->q + 4 * mid, key);
        if (c == 0) return L->count[mid];
        if (c < 0) lo = mid + 1;
        else hi = mid;
    }
    return 0;
}

/* splice_signal (splice.rs:75-104): 0 GT-AG, 1 GC-AG, 2 AT-AC, 3 non-canonical. */
static int signal_of(const uint8_t d[2], const uint8_t a[2], int minus) {
    uint8_t d0 = upper(d[0]), d1 = upper(d[1]), a0 = upper(a[0]), a1 = upper(a[1]);
    if (!minus) {
        if (d0 == 'G' && d1 == 'T' && a0 == 'A' && a1 == 'G') return 0;
        if (d0 == 'G' && d1 == 'C' && a0 == 'A' && a1 == 'G') return 1;
        if (d0 == 'A' && d1 == 'T' && a0 == 'A' && a1 == 'C') return 2;
        return 3;
    }
    if (d0 == 'C' && d1 == 'T' && a0 == 'A' && a1 == 'C') return 0;
    if (d0 == 'C' && d1 == 'T' && a0 == 'G' && a1 == 'C') return 1;
    if (d0 == 'G' && d1 == 'T' && a0 == 'A' && a1 == 'T') return 2;
    return 3;
}

/* SpliceSignal::score (splice.rs:62-71). */
static int32_t signal_bonus(int sig, const espo_splice_cfg *sp) {
    switch (sig) {
    case 0: return sp->gt_ag_bonus;
    case 1: return sp->gc_ag_bonus;
    case 2: return sp->at_ac_bonus;
    default: return -sp->noncanonical_penalty;
    }
}

/* dinuc (splice.rs:330-347): donor [p, p+2), acceptor [p-2, p), N-padded
 * past the contig end. */
static void dinuc_at(const espo_ref *R, uint32_t contig, uint32_t p, int donor, uint8_t out[2]) {
    out[0] = out[1] = 'N';
    if (!donor && p < 2) return;
    uint32_t a = donor ? p : p - 2, b = donor ? p + 2 : p;
    if (a >= R->len[contig]) return;
    if (b > R->len[contig]) b = R->len[contig];
    for (uint32_t x = a; x < b; x++) out[x - a] = contig_base(R, contig, x);
}

/* Base::from_ascii (fasta.rs:30-38) as a code, N = 4. */
static int base_of(uint8_t b) {
    switch (b) {
    case 'A': case 'a': return 0;
    case 'C': case 'c': return 1;
    case 'G': case 'g': return 2;
    case 'T': case 't': return 3;
    default: return 4;
    }
}

/* edit_equiv (splice.rs:320-325): equal Bases (N equals N), or A/G, T/C either way. */
static int edit_equiv(int a, int b) {
    return a == b || (a == 0 && b == 2) || (a == 2 && b == 0) || (a == 3 && b == 1) || (a == 1 && b == 3);
}

typedef struct {
    int32_t found; /* 1: Some(RefinedJunction), 0: None */
    uint32_t start, end, minus;
    int32_t signal; /* 0 GT-AG, 1 GC-AG, 2 AT-AC, 3 non-canonical */
    uint32_t support;
} espo_refined;

/* refine_junction_read: splice.rs:208-317. */
int espo_refine_junction_read(const espo_ref *R, const uint8_t *read, size_t read_len_sz, uint32_t contig,
                              uint32_t naive_start, uint32_t naive_end, uint32_t q_a, uint32_t q_b, int minus,
                              const espo_jlib *L, const espo_splice_cfg *sp, espo_refined *out) {
    memset(out, 0, sizeof(*out));
    if (contig >= R->ncontig) return 0;
    const uint32_t len = R->len[contig];
    const uint32_t read_len = (uint32_t)read_len_sz;
    const int64_t ref_gap = (int64_t)naive_end - (int64_t)naive_start;
    const int64_t read_gap = (int64_t)q_b - (int64_t)q_a;
    if (read_gap > 20) return 0;
    const int64_t intron_len = ref_gap - read_gap;
    if (intron_len < (int64_t)sp->min_intron || intron_len > (int64_t)sp->max_intron) return 0;
    const int64_t lo = (read_gap < 0 ? read_gap : 0) - 2, hi = (read_gap > 0 ? read_gap : 0) + 2;
    int have = 0;
    int32_t best = 0;
    for (int64_t delta = lo; delta <= hi; delta++) {
        const int64_t s = (int64_t)naive_start + delta;
        if (s < 0) continue;
        const int64_t e = s + intron_len;
        if (e < 2 || (uint64_t)e > (uint64_t)len) continue;
        const int64_t qb = (int64_t)q_a + delta;
        if (qb < 1 || (uint32_t)qb >= read_len) continue;
        int32_t mt = 0;
        for (int64_t i = 1; i <= 4; i++) {
            const int64_t q1 = qb - i, g1 = s - i;
            /* (u64)q < read_len: the reference would panic past the read */
            if (q1 >= 0 && g1 >= 0 && (uint64_t)q1 < read_len &&
                edit_equiv(base_of(read[q1]), base_of(contig_base(R, contig, (uint32_t)g1))))
                mt++;
            const int64_t q2 = qb + i - 1, g2 = e + i - 1;
            if ((uint32_t)q2 < read_len && (uint64_t)q2 < read_len && (uint64_t)g2 < (uint64_t)len &&
                edit_equiv(base_of(read[q2]), base_of(contig_base(R, contig, (uint32_t)g2))))
                mt++;
        }
        uint8_t dn[2], ac[2];
        dinuc_at(R, contig, (uint32_t)s, 1, dn);
        dinuc_at(R, contig, (uint32_t)e, 0, ac);
        for (int t = 0; t < 2; t++) {
            const int tm = t == 0 ? minus : !minus;
            const int sig = signal_of(dn, ac, tm);
            const uint32_t support = espo_jlib_support(L, contig, (uint32_t)s, (uint32_t)e, (uint32_t)tm);
            const int64_t ad = delta < 0 ? -delta : delta;
            int32_t score = mt * 50 + signal_bonus(sig, sp) * 5 - (int32_t)ad * 5;
            if (support > 0) score += sp->known_bonus;
            if (!have || score > best) {
                have = 1;
                best = score;
                out->found = 1;
                out->start = (uint32_t)s;
                out->end = (uint32_t)e;
                out->minus = (uint32_t)tm;
                out->signal = sig;
                out->support = support;
            }
        }
    }
    return have;
}

/* splice_support_pass: splice.rs:107-115 (MIN_INTRON_WITHOUT_SUPPORT = 50). */
static int support_pass(uint32_t intron_len, int sig, uint32_t support) {
    if (intron_len < 50 && support == 0) return 0;
    if (sig == 3 && support == 0) return 0;
    return 1;
}

/* segment_chain: splice.rs:364-396.  anchors as (qpos, rpos) pairs in the
 * chain's (rpos, qpos) order; segs receives 4 u32 (q_start, q_end, r_start,
 * r_end) per segment (capacity n_anchors).  Returns the segment count. */
size_t espo_segment_chain(const uint32_t *anchors, size_t n, uint32_t k, uint32_t min_intron, uint32_t *segs) {
    if (n == 0) return 0;
    size_t ns = 0;
    uint32_t cur[4] = {anchors[0], anchors[0] + k, anchors[1], anchors[1] + k};
    for (size_t a = 1; a < n; a++) {
        const uint32_t q1 = anchors[2 * (a - 1)], r1 = anchors[2 * (a - 1) + 1];
        const uint32_t q2 = anchors[2 * a], r2 = anchors[2 * a + 1];
        const uint32_t ref_gap = r2 > r1 + k ? r2 - (r1 + k) : 0u;
        const uint32_t read_gap = q2 > q1 + k ? q2 - (q1 + k) : 0u;
        if (ref_gap > min_intron && read_gap <= 10) {
            memcpy(segs + 4 * ns++, cur, sizeof(cur));
            cur[0] = q2; cur[1] = q2 + k; cur[2] = r2; cur[3] = r2 + k;
        } else {
            cur[1] = q2 + k;
            cur[3] = r2 + k;
        }
    }
    memcpy(segs + 4 * ns++, cur, sizeof(cur));
    return ns;
}

/* Refusal codes of prepare_spliced (splice.rs:545-621), each a `None`:
 * 1 fewer than two segments, 2 a refinement found nothing, 3 a refined intron
 * outside [min_intron, max_intron], 4 the support guard, 5 an exon piece with
 * re <= rs; 6 a contig out of range (`contigs.get(..)?`). */
typedef struct {
    int32_t status;   /* 1 aligned, 0 refused (refusal says why), -2 stitched capacity too small */
    int32_t refusal;
    uint32_t n_segments, genomic_start, pseudo_len, n_stitched;
} espo_align_splice_rec;

/* align_spliced: splice.rs:625-664 after prepare_spliced (:545-621).
 * junctions receives n_segments - 1 records, bounds 3 u32 per segment. */
int espo_align_spliced(const espo_ref *R, const espo_jlib *L, const uint8_t *read, size_t m, uint32_t contig,
                       int minus, const uint32_t *anchors, size_t n_anchors, uint32_t k,
                       const espo_splice_cfg *sp, const espo_params *p, espo_ext *ext, uint32_t *cigar,
                       size_t cap, size_t *n_ops, uint32_t *stitched, size_t scap,
                       espo_align_splice_rec *rec, espo_refined *junctions, uint32_t *bounds) {
    memset(rec, 0, sizeof(*rec));
    memset(ext, 0, sizeof(*ext));
    *n_ops = 0;
    uint32_t *segs = (uint32_t *)malloc((n_anchors ? n_anchors : 1) * 4 * sizeof(uint32_t));
    const size_t ns = espo_segment_chain(anchors, n_anchors, k, sp->min_intron, segs);
    rec->n_segments = (uint32_t)ns;
    int refusal = 0;
    if (ns < 2) refusal = 1;
    else if (contig >= R->ncontig) refusal = 6;
    for (size_t w = 0; !refusal && w + 1 < ns; w++) {
        espo_refined *j = &junctions[w];
        if (!espo_refine_junction_read(R, read, m, contig, segs[4 * w + 3], segs[4 * (w + 1) + 2], segs[4 * w + 1],
                                       segs[4 * (w + 1)], minus, L, sp, j)) {
            refusal = 2;
            break;
        }
        const uint32_t il = j->end - j->start;
        if (il < sp->min_intron || il > sp->max_intron) refusal = 3;
        else if (!support_pass(il, j->signal, j->support)) refusal = 4;
    }
    uint64_t total = 0;
    for (size_t i = 0; !refusal && i < ns; i++) {
        const uint32_t rs = i == 0 ? (segs[2] > 30 ? segs[2] - 30 : 0u) : junctions[i - 1].end;
        uint32_t re;
        if (i + 1 == ns) {
            const uint64_t x = (uint64_t)segs[4 * i + 3] + 30u;
            re = (uint32_t)(x < R->len[contig] ? x : R->len[contig]);
        } else {
            re = junctions[i].start;
        }
        if (re <= rs) {
            refusal = 5;
            break;
        }
        bounds[3 * i] = (uint32_t)total;
        bounds[3 * i + 1] = rs;
        bounds[3 * i + 2] = re;
        total += re - rs;
    }
    free(segs);
    if (refusal) {
        rec->refusal = refusal;
        return 0;
    }
    uint8_t *pseudo = (uint8_t *)malloc(total ? total : 1);
    uint32_t *skips = (uint32_t *)malloc(ns * sizeof(uint32_t));
    for (size_t i = 0; i < ns; i++) {
        espo_decode(R, contig, bounds[3 * i + 1], bounds[3 * i + 2], 0, pseudo + bounds[3 * i]);
        if (i + 1 < ns) skips[i] = junctions[i].end - junctions[i].start;
    }
    int32_t esc;
    int rc = espo_extend_hint(read, m, pseudo, total, p, 0, ext, cigar, cap, n_ops, &esc);
    size_t nst = 0;
    int src = espo_stitch_cigar(cigar, *n_ops, ext->ref_start, bounds, ns, skips, ns - 1, stitched, scap, &nst);
    rec->status = src ? -2 : 1;
    rec->n_stitched = (uint32_t)nst;
    rec->pseudo_len = (uint32_t)total;
    rec->genomic_start = genomic_start_of(ext->ref_start, bounds, ns);
    free(pseudo);
    free(skips);
    return rc;
}

typedef struct {
    int32_t status; /* 1 replaced, 0 kept, 2 spliced (left alone), 3 empty window, -1 contig out of range */
    uint32_t pos;   /* the placement's position after the call */
    int32_t score;  /* the new extension's score (0 when none ran) */
    uint32_t wstart, wlen, aligned, new_aligned;
} espo_reext_rec;

/* Aligner::reextend_with: align.rs:1170-1225.  dir replaces p->edit_dir.
 * cigar_out holds the new extension's CIGAR whenever the DP ran. */
int espo_reextend_with(const espo_ref *R, const uint8_t *query, size_t m, uint32_t contig, uint32_t pos,
                       const uint32_t *cigar, size_t n_ops, const espo_params *p, int dir, espo_reext_rec *rec,
                       uint32_t *cigar_out, size_t cap, size_t *n_out) {
    memset(rec, 0, sizeof(*rec));
    rec->pos = pos;
    *n_out = 0;
    for (size_t k = 0; k < n_ops; k++)
        if (op_kind(cigar[k]) == OP_N) {
            rec->status = 2;
            return 0;
        }
    if (contig >= R->ncontig) {
        rec->status = -1;
        return 0;
    }
    const uint32_t read_len = (uint32_t)m;
    uint32_t lead = 0, tail = 0, ref_span = 0, aligned = 0;
    int seen_m = 0;
    for (size_t k = 0; k < n_ops; k++) {
        const uint32_t kind = op_kind(cigar[k]), n = op_len(cigar[k]);
        if (kind == OP_S) {
            if (seen_m) tail += n;
            else lead += n;
        } else if (kind == OP_M) {
            seen_m = 1;
            ref_span += n;
            aligned += n;
        } else if (kind == OP_I) {
            aligned += n;
        } else if (kind == OP_D) {
            ref_span += n;
        }
    }
    rec->aligned = aligned;
    const uint32_t back = lead + 30u;
    const uint32_t wstart = pos > back ? pos - back : 0u;
    uint32_t wend = pos + ref_span + tail + 30u; /* u32 arithmetic, as the reference's */
    if (wend > R->len[contig]) wend = R->len[contig];
    if (wend <= wstart) {
        rec->status = 3;
        return 0;
    }
    rec->wstart = wstart;
    rec->wlen = wend - wstart;
    uint8_t *win = (uint8_t *)malloc(wend - wstart);
    espo_decode(R, contig, wstart, wend, 0, win);
    espo_params ep = *p;
    ep.edit_dir = dir;
    espo_ext e;
    int32_t esc;
    const int64_t hint = (int64_t)(pos - wstart) - (int64_t)lead;
    int rc = espo_extend_hint(query, m, win, wend - wstart, &ep, hint, &e, cigar_out, cap, n_out, &esc);
    free(win);
    uint32_t new_aligned = 0;
    for (size_t k = 0; k < *n_out; k++)
        if (op_kind(cigar_out[k]) == OP_M || op_kind(cigar_out[k]) == OP_I) new_aligned += op_len(cigar_out[k]);
    rec->score = e.score;
    rec->new_aligned = new_aligned;
    if (e.score > 0 && new_aligned >= aligned && new_aligned <= read_len) {
        rec->status = 1;
        rec->pos = wstart + e.ref_start;
    }
    return rc;
}

/* ------------------------------------------------------ intron-chain fast path */

/* IntronParams (intron_chain.rs:20-42). */
typedef struct {
    uint32_t min_frag_cov, refine_window, refine_pattern_len, intron_max;
} espo_intron_cfg;

typedef struct {
    uint32_t contig, strand, q_lo, q_hi, r_lo, r_hi;
} ic_mem;

typedef struct {
    uint32_t contig, strand, q_lo, q_hi, r_lo, r_hi;
    size_t m0, nm; /* member MEMs mems[m0 .. m0 + nm), sorted by q_lo */
} ic_frag;

/* base_eq: intron_chain.rs:113-117. */
static int ic_base_eq(uint8_t a, uint8_t b) {
    a &= 0xDF;
    b &= 0xDF;
    return a >= 'A' && a <= 'T' && a == b && (a == 'A' || a == 'C' || a == 'G' || a == 'T');
}

static int64_t mem_diag(const ic_mem *m) { return (int64_t)m->q_lo - (int64_t)m->r_lo; }

static int mem_key_cmp(const ic_mem *x, const ic_mem *y) {
    if (x->contig != y->contig) return x->contig < y->contig ? -1 : 1;
    if (x->strand != y->strand) return x->strand < y->strand ? -1 : 1;
    if (mem_diag(x) != mem_diag(y)) return mem_diag(x) < mem_diag(y) ? -1 : 1;
    if (x->q_lo != y->q_lo) return x->q_lo < y->q_lo ? -1 : 1;
    if (x->r_lo != y->r_lo) return x->r_lo < y->r_lo ? -1 : 1;
    return 0;
}

/* Frag::cov (intron_chain.rs:91-102) over q_lo-sorted members. */
static uint32_t frag_cov(const ic_frag *f, const ic_mem *mems) {
    uint32_t total = 0, end = 0;
    for (size_t i = 0; i < f->nm; i++) {
        const ic_mem *m = &mems[f->m0 + i];
        if (m->q_hi > end) {
            total += m->q_hi - (m->q_lo > end ? m->q_lo : end);
            end = m->q_hi;
        }
    }
    return total;
}

static int64_t frag_off_lo(const ic_frag *f) { return (int64_t)f->r_lo - (int64_t)f->q_lo; }

typedef struct {
    int32_t status;   /* 1 placed, 0 None (reason says which stage) */
    int32_t reason;   /* 1 no anchors/read, 2 no fragment, 3 dominant fragment, 4 no pair, 5 contig,
                         6 anchors, 7 intron length, 8 split verification, 9 junction, 10 exon bounds,
                         11 extension score, 12 CIGAR not conserving, 13 no RefSkip */
    uint32_t genomic_start;
    int32_t score;
    uint32_t left_start, jstart, jend, right_end;
    int32_t signal;
    uint32_t support, minus, n_stitched;
    int64_t donor_end, acceptor_start;
    int32_t vd, va;
} espo_ic_rec;

/* try_intron_chain_placement: intron_chain.rs:392-538 with extend_mems
 * (:122-166), cluster_frags (:171-214), best_frag_pair (:233-277),
 * refine_donor_end (:285-309), refine_acceptor_start (:318-342) and
 * verify_split_read (:346-373).  anchors: 4 u32 each (qpos, rpos, contig,
 * strand 0 Plus / 1 Minus).  stitched receives the genomic CIGAR. */
int espo_intron_chain(const espo_ref *R, const espo_jlib *L, const uint8_t *read, size_t m,
                      const uint32_t *anchors, size_t na, uint32_t k, const espo_intron_cfg *ip,
                      const espo_params *p, espo_ext *ext, uint32_t *cigar, size_t cap, size_t *n_ops,
                      uint32_t *stitched, size_t scap, espo_ic_rec *rec) {
    memset(rec, 0, sizeof(*rec));
    memset(ext, 0, sizeof(*ext));
    *n_ops = 0;
    if (na == 0 || m == 0) { rec->reason = 1; return 0; }
    const uint32_t rl = (uint32_t)m;
    ic_mem *mems = (ic_mem *)malloc(na * sizeof(ic_mem));
    size_t nm = 0;
    for (size_t a = 0; a < na; a++) {
        const uint32_t qpos = anchors[4 * a], rpos = anchors[4 * a + 1], ctg = anchors[4 * a + 2];
        if (ctg >= R->ncontig) continue;
        const uint32_t clen = R->len[ctg];
        const uint32_t win_lo = rpos > rl ? rpos - rl : 0u;
        uint64_t wh = (uint64_t)rpos + k + rl;
        const uint32_t win_hi = (uint32_t)(wh < clen ? wh : clen);
        if (win_hi <= win_lo) continue;
        uint32_t q_lo = qpos, r_lo = rpos;
        /* q_lo <= rl, r_lo <= win_hi: an anchor past the read or the contig (the
         * reference would panic) does not extend */
        while (q_lo > 0 && r_lo > win_lo && q_lo <= rl && r_lo <= win_hi &&
               ic_base_eq(read[q_lo - 1], contig_base(R, ctg, r_lo - 1))) {
            q_lo--;
            r_lo--;
        }
        uint32_t q_hi = qpos + k, r_hi = rpos + k;
        while (q_hi < rl && r_hi < win_hi && ic_base_eq(read[q_hi], contig_base(R, ctg, r_hi))) { q_hi++; r_hi++; }
        if (q_hi > q_lo) {
            ic_mem x = {ctg, anchors[4 * a + 3] ? 1u : 0u, q_lo, q_hi, r_lo, r_hi};
            mems[nm++] = x;
        }
    }
    /* stable insertion sort by (contig, strand, diag, q_lo, r_lo) */
    for (size_t i = 1; i < nm; i++) {
        ic_mem x = mems[i];
        size_t j = i;
        while (j > 0 && mem_key_cmp(&mems[j - 1], &x) > 0) { mems[j] = mems[j - 1]; j--; }
        mems[j] = x;
    }
    ic_frag *frags = (ic_frag *)malloc((nm ? nm : 1) * sizeof(ic_frag));
    size_t nf = 0;
    for (size_t i = 0; i < nm; i++) {
        const ic_mem *x = &mems[i];
        int start_new = 1;
        if (nf > 0 && frags[nf - 1].contig == x->contig && frags[nf - 1].strand == x->strand)
            start_new = mem_diag(x) - ((int64_t)frags[nf - 1].q_lo - (int64_t)frags[nf - 1].r_lo) > 5;
        if (start_new) {
            ic_frag f = {x->contig, x->strand, x->q_lo, x->q_hi, x->r_lo, x->r_hi, i, 1};
            frags[nf++] = f;
        } else {
            ic_frag *f = &frags[nf - 1];
            if (x->q_lo < f->q_lo) f->q_lo = x->q_lo;
            if (x->q_hi > f->q_hi) f->q_hi = x->q_hi;
            if (x->r_lo < f->r_lo) f->r_lo = x->r_lo;
            if (x->r_hi > f->r_hi) f->r_hi = x->r_hi;
            f->nm++;
        }
    }
    for (size_t f = 0; f < nf; f++) /* members by q_lo, stable */
        for (size_t i = frags[f].m0 + 1; i < frags[f].m0 + frags[f].nm; i++) {
            ic_mem x = mems[i];
            size_t j = i;
            while (j > frags[f].m0 && mems[j - 1].q_lo > x.q_lo) { mems[j] = mems[j - 1]; j--; }
            mems[j] = x;
        }
    /* keep fragments with cov >= min_frag_cov */
    size_t nk = 0;
    uint32_t dom = 0;
    for (size_t f = 0; f < nf; f++) {
        const uint32_t cv = frag_cov(&frags[f], mems);
        if (cv >= ip->min_frag_cov) {
            frags[nk++] = frags[f];
            if (cv > dom) dom = cv;
        }
    }
    int reason = 0;
    if (nk == 0) reason = 2;
    else if ((uint64_t)dom * 20 >= (uint64_t)m * 11) reason = 3;
    int have = 0;
    size_t bi = 0, bj = 0;
    int32_t bscore = 0;
    uint32_t split = 0, istart = 0, iend = 0;
    for (size_t i = 0; !reason && i < nk; i++)
        for (size_t j = 0; j < nk; j++) {
            if (i == j) continue;
            const ic_frag *a = &frags[i], *b = &frags[j];
            if (a->contig != b->contig || a->strand != b->strand || a->q_hi > b->q_lo + 3) continue;
            const uint32_t sp = a->q_hi < b->q_lo ? a->q_hi : b->q_lo;
            const int64_t off_a = (int64_t)a->r_hi - (int64_t)a->q_hi, off_b = frag_off_lo(b);
            const int64_t len0 = off_b - off_a;
            if (len0 < 20 || len0 > (int64_t)ip->intron_max) continue;
            const uint32_t ca = frag_cov(a, mems), cb = frag_cov(b, mems);
            const int64_t bal = (int64_t)ca - (int64_t)cb;
            const int32_t score = (int32_t)((ca + cb) * 4) + (int32_t)((a->nm + b->nm) * 3) -
                                  (int32_t)(bal < 0 ? -bal : bal);
            if (!have || score > bscore) {
                have = 1;
                bscore = score;
                bi = i;
                bj = j;
                split = sp;
                istart = (uint32_t)((int64_t)sp + off_a);
                iend = (uint32_t)((int64_t)sp + off_b);
            }
        }
    if (!reason && !have) reason = 4;
    ic_frag left = {0}, right = {0};
    if (!reason) {
        left = frags[bi];
        right = frags[bj];
        if (left.contig >= R->ncontig) reason = 5;
    }
    (void)split;
    if (reason) {
        rec->reason = reason;
        free(mems);
        free(frags);
        return 0;
    }
    const uint32_t ctg = left.contig, clen = R->len[ctg];
    const int minus = left.strand == 1;
    const uint32_t j_est = (uint32_t)(m / 2);
    const uint32_t junction_off =
        left.q_hi >= right.q_lo ? j_est : (j_est < left.q_hi ? left.q_hi : (j_est > right.q_lo ? right.q_lo : j_est));
    const int64_t w = ip->refine_window;
    const size_t plen = ip->refine_pattern_len;
    /* refine_donor_end */
    int64_t donor_end;
    {
        const size_t off = junction_off < m ? junction_off : m;
        const int64_t lo64 = (int64_t)left.r_hi - w, hi64 = (int64_t)left.r_hi + w;
        const uint32_t lo = (uint32_t)(lo64 > 0 ? lo64 : 0), hi = (uint32_t)(hi64 < (int64_t)clen ? hi64 : clen);
        const uint32_t dflt = left.r_hi > 0 ? left.r_hi - 1 : 0u;
        donor_end = dflt;
        if (!(hi <= lo || off == 0)) {
            const int64_t tl = (int64_t)lo - frag_off_lo(&left) + 2;
            const size_t text_lo_in_read = (size_t)(tl > 0 ? tl : 0);
            size_t pat_start = off > plen ? off - plen : 0;
            if (text_lo_in_read > pat_start) pat_start = text_lo_in_read;
            if (pat_start < off) {
                uint8_t *text = (uint8_t *)malloc(hi - lo);
                espo_decode(R, ctg, lo, hi, 0, text);
                int32_t d;
                size_t end;
                espo_myers_best_end(read + pat_start, off - pat_start, text, hi - lo, &d, &end);
                donor_end = (int64_t)lo + (int64_t)end;
                free(text);
            }
        }
    }
    /* refine_acceptor_start */
    int64_t acceptor_start;
    {
        const size_t off = junction_off < m ? junction_off : m;
        const int64_t lo64 = (int64_t)right.r_lo - w, hi64 = (int64_t)right.r_lo + w;
        const uint32_t lo = (uint32_t)(lo64 > 0 ? lo64 : 0), hi = (uint32_t)(hi64 < (int64_t)clen ? hi64 : clen);
        acceptor_start = right.r_lo;
        if (hi > lo) {
            const int64_t th = (int64_t)hi - frag_off_lo(&right) - 2;
            const size_t text_hi_in_read = (size_t)(th > 0 ? th : 0);
            size_t pat_end = off + plen;
            if (pat_end > m) pat_end = m;
            if (pat_end > text_hi_in_read) pat_end = text_hi_in_read;
            if (pat_end > off) {
                uint8_t *text = (uint8_t *)malloc(hi - lo);
                espo_decode(R, ctg, lo, hi, 0, text);
                int32_t d;
                size_t start;
                espo_myers_best_start(read + off, pat_end - off, text, hi - lo, &d, &start);
                acceptor_start = (int64_t)lo + (int64_t)start;
                free(text);
            }
        }
    }
    rec->donor_end = donor_end;
    rec->acceptor_start = acceptor_start;
    const int64_t intron_len = acceptor_start - donor_end - 1;
    /* verify_split_read */
    int32_t vd = INT32_MAX, va = INT32_MAX;
    {
        const size_t mid = m / 2;
        const int64_t ds = frag_off_lo(&left) - 30;
        const uint32_t d_start = (uint32_t)(ds > 0 ? ds : 0);
        int64_t de = donor_end + 1 + 30;
        if (de < 1) de = 1;
        if (de > (int64_t)clen) de = clen;
        const uint32_t d_end = (uint32_t)de;
        int64_t as = acceptor_start - 30;
        if (as < 0) as = 0;
        if (as > (int64_t)clen) as = clen;
        const uint32_t a_start = (uint32_t)as;
        int64_t ae = (int64_t)right.r_hi + 30;
        if (ae < (int64_t)a_start + 1) ae = (int64_t)a_start + 1;
        if (ae > (int64_t)clen) ae = clen;
        const uint32_t a_end = (uint32_t)ae;
        if (!(d_start >= d_end || a_start >= a_end || mid == 0 || mid >= m)) {
            uint8_t *t1 = (uint8_t *)malloc(d_end - d_start), *t2 = (uint8_t *)malloc(a_end - a_start);
            espo_decode(R, ctg, d_start, d_end, 0, t1);
            espo_decode(R, ctg, a_start, a_end, 0, t2);
            size_t e;
            espo_myers_best_end(read, mid, t1, d_end - d_start, &vd, &e);
            espo_myers_best_end(read + mid, m - mid, t2, a_end - a_start, &va, &e);
            free(t1);
            free(t2);
        }
    }
    rec->vd = vd;
    rec->va = va;
    const size_t total_anchors = left.nm + right.nm;
    const int32_t max_v = vd > va ? vd : va;
    const int32_t sum_v = (int32_t)((uint32_t)vd + (uint32_t)va); /* release i32 wraps */
    const int32_t max_thr = (int32_t)((double)m * 0.17), sum_thr = (int32_t)((double)m * 0.24);
    if (total_anchors < 2) reason = 6;
    else if (intron_len < 20 || intron_len > (int64_t)ip->intron_max) reason = 7;
    else if (max_v > max_thr || sum_v > sum_thr) reason = 8;
    espo_refined rj;
    memset(&rj, 0, sizeof(rj));
    uint32_t left_start = 0, right_end = 0;
    if (!reason) {
        espo_splice_cfg sp = {20, ip->intron_max, 4, 2, 1, 8, 6, 12};
        if (!espo_refine_junction_read(R, read, m, ctg, left.r_hi, right.r_lo, left.q_hi, right.q_lo, minus, L, &sp,
                                       &rj)) {
            uint8_t dn[2], ac[2];
            dinuc_at(R, ctg, istart, 1, dn);
            dinuc_at(R, ctg, iend, 0, ac);
            rj.found = 2;
            rj.start = istart;
            rj.end = iend;
            rj.minus = (uint32_t)minus;
            rj.signal = signal_of(dn, ac, minus);
            rj.support = espo_jlib_support(L, ctg, istart, iend, (uint32_t)minus);
        }
        if (rj.end <= rj.start || rj.end - rj.start < 20 || rj.end - rj.start > ip->intron_max) reason = 9;
    }
    if (!reason) {
        left_start = left.r_lo > rl ? left.r_lo - rl : 0u;
        const uint64_t re = (uint64_t)right.r_hi + rl;
        right_end = (uint32_t)(re < clen ? re : clen);
        if (left_start >= rj.start || rj.end >= right_end) reason = 10;
    }
    free(mems);
    free(frags);
    rec->jstart = rj.start;
    rec->jend = rj.end;
    rec->signal = rj.signal;
    rec->support = rj.support;
    rec->minus = (uint32_t)minus;
    if (reason) {
        rec->reason = reason;
        return 0;
    }
    rec->left_start = left_start;
    rec->right_end = right_end;
    const uint32_t exons[4] = {left_start, rj.start, rj.end, right_end};
    const int64_t hint = frag_off_lo(&left) - (int64_t)left_start;
    espo_splice_rec srec;
    int rc = espo_spliced(R, read, m, ctg, exons, 2, hint, p, ext, cigar, cap, n_ops, stitched, scap, &srec);
    rec->genomic_start = srec.genomic_start;
    rec->score = ext->score;
    rec->n_stitched = srec.n_stitched;
    uint64_t q = 0;
    int has_n = 0;
    for (size_t i = 0; i < srec.n_stitched; i++) {
        const uint32_t kind = op_kind(stitched[i]);
        if (kind == OP_M || kind == OP_I || kind == OP_S) q += op_len(stitched[i]);
        if (kind == OP_N) has_n = 1;
    }
    if (srec.status != 1) reason = 12;
    else if (ext->score <= 0) reason = 11;
    else if (q != m) reason = 12;
    else if (!has_n) reason = 13;
    rec->reason = reason;
    rec->status = reason ? 0 : 1;
    return rc;
}
