/*
 * group_agg_node.c -- drop-in GPU GroupAggNode (src/group_agg.h:14-31).
 *
 * Same constructor, struct prefix, output schema (keys, then one double
 * column per aggregate; src/group_agg.c:641-663) and single-result-batch
 * protocol as the reference.  Ordering follows the reference's choice of
 * path (src/group_agg.c:624-638): with temp_dir set and at least one key
 * the groups come out ascending by key with NA last, as the SortNode-fed
 * path produces them; otherwise in first-occurrence order, as the hash
 * path does.  The SortNode itself is not built -- the GPU table yields the
 * key order directly -- so explain() no longer lists it.
 *
 * Every child batch is copied to the device and fed to one vgpu_groupby;
 * the finished table comes back as malloc'd columns.
 */
#include "devscan.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* The reference sorts the rows of a key-ordered group-by with a top-down
 * merge sort whose comparator finds NaN equal to every number
 * (src/sort.c:32-47, 97-116), so where NaN-key groups land -- and whether
 * the groups of one number are split around them -- follows the merge
 * pattern over all input rows.  The GPU does not reproduce that order and
 * says so instead of returning another one. */
#define NAN_KEY_ORDER_ERROR \
    "GPU group_by: NaN values in a double group key with key order (the reference's merge sort places " \
    "NaN keys by its merge pattern over all rows; not reproduced on the GPU)"

#define MK_MAX_KEYS 8

/* a batch's string key column, kept on the host until the groups' strings
 * are read from it; base = physical position of its first row */
typedef struct {
    VecArray a;
    int64_t  base;
} KeptStr;

typedef struct {
    vgpu_groupby *live_gb;            /* the group-by being fed (released on errors by free) */
    KeptStr *kept[MK_MAX_KEYS];       /* string keys: batches' string columns kept */
    int64_t n_kept;
    DevCol fj_probe[VGPU_MAX_COLS];   /* fused join: probe columns of one batch */
    DevCol fj_cast;                   /* fused join: probe key cast to the build key type */
    DevCol key;
    DevCol in[VGPU_MAX_INPUTS];
    DevMem sel;
    DevMem codes;      /* multi-column keys: one tuple code per row */
    VxDevScan ds;      /* open while a .vtr child is fed straight to the device */
    int ds_open;
    /* string key columns: the batch's offsets, bytes and validity, and the
     * two hashes and positions computed from them */
    DevMem s_val[VGPU_MAX_COLS], s_code[VGPU_MAX_COLS], s_pos[VGPU_MAX_COLS];
    vgpu_strdict *sdict[VGPU_MAX_COLS];   /* per string key: its exact dictionary (this run) */
    int fused;         /* the last run took the fused Project <- Join path */
} GroupGpu;

/* the group-by a run holds: kept in the node state so that free_node
 * releases it when an error unwinds the run */
static vgpu_groupby *gb_hold(GroupAggNode *ga, vgpu_groupby *gb) {
    ((GroupGpu *)ga->gpu)->live_gb = gb;
    return gb;
}

static void gb_drop(GroupAggNode *ga, vgpu_groupby *gb) {
    GroupGpu *g = (GroupGpu *)ga->gpu;
    if (g->live_gb == gb) g->live_gb = NULL;
    vgpu_groupby_destroy(gb);
}

/* file column of a scan's k-th output column (its columns in file order) */
static int scan_file_col(const ScanNode *sn, int k) {
    const int n = sn->file->header.schema.n_cols;
    for (int c = 0; c < n; c++)
        if (sn->col_mask[c] && k-- == 0) return c;
    return -1;
}

/* the index of file column c in fcols[0..*n), appended when absent */
static int fed_slot(int *fcols, int *n, int c) {
    for (int j = 0; j < *n; j++)
        if (fcols[j] == c) return j;
    if (*n >= VX_DS_COLS) vectra_error("GPU scan feed: more than %d columns", VX_DS_COLS);
    fcols[*n] = c;
    return (*n)++;
}

/* the .vtr scan under a child that is one, or a filter over one */
static ScanNode *scan_below(VecNode *c) {
    if (c->kind && strcmp(c->kind, "FilterNode") == 0) c = ((FilterNode *)c)->child;
    return c->kind && strcmp(c->kind, "ScanNode") == 0 ? (ScanNode *)c : NULL;
}

/* A bound on the groups of an int64 / bool key read from a .vtr scan: its
 * value range over every row group's statistics, plus the NA group (0 when
 * the file has no statistics or the range is no tighter than the rows).
 * With it the group table is sized once and trusted, as for a caller's
 * group count, instead of reading its occupancy back before every batch. */
static int64_t scan_group_bound(const ScanNode *sn, int key_out_col) {
    const Vtr1File *f = sn->file;
    const VecSchema *fs = &f->header.schema;
    int c = -1;
    for (int i = 0, k = key_out_col; i < fs->n_cols; i++)
        if (sn->col_mask[i] && k-- == 0) c = i;
    if (c < 0 || f->header.version < 3) return 0;
    if (fs->col_types[c] == VEC_BOOL) return 3;
    if (fs->col_types[c] != VEC_INT64) return 0;
    int64_t rows = 0, lo = 0, hi = 0;
    int any = 0;
    for (uint32_t g = 0; g < f->header.n_rowgroups; g++) {
        const Vtr1RowGroup *rg = &f->rowgroups[g];
        rows += rg->n_rows;
        if (!rg->col_stats || !rg->col_stats[c].has_stats) continue;
        const int64_t a = rg->col_stats[c].i64.min, b = rg->col_stats[c].i64.max;
        if (!any || a < lo) lo = a;
        if (!any || b > hi) hi = b;
        any = 1;
    }
    if (!any) return rows > 0 ? 1 : 0;
    const uint64_t span = (uint64_t)hi - (uint64_t)lo;
    if (span >= (uint64_t)rows) return 0;
    return (int64_t)span + 2;
}

/* A child that is a .vtr scan, or a filter over one, is fed from the file
 * straight to the device (knob "scan.device_feed"): row groups pruned by
 * the scan's pushed-down predicate as the scan prunes them, the filter's
 * predicate evaluated inside the group-by feed instead of as a selection
 * vector that comes back to the host.  Same groups, order and values as
 * pulling the child's batches: positions count the rows of the groups fed,
 * which orders first occurrences as the filtered batches would. */
static int feed_from_scan(GroupAggNode *ga, vgpu_groupby *gb, int key_idx, int n_in, const int *in_col) {
    VecNode *c = ga->child;
    FilterNode *fn = NULL;
    if (c->kind && strcmp(c->kind, "FilterNode") == 0) {
        fn = (FilterNode *)c;
        c = fn->child;
    }
    if (!c->kind || strcmp(c->kind, "ScanNode") != 0 || vgpu_tune_get("scan.device_feed") <= 1) return 0;
    ScanNode *sn = (ScanNode *)c;
    GroupGpu *g = (GroupGpu *)ga->gpu;
    int fcols[VX_DS_COLS], nf = 0, kslot = -1, islot[VGPU_MAX_INPUTS], pslot[VGPU_MAX_COLS];
    if (key_idx >= 0) kslot = fed_slot(fcols, &nf, scan_file_col(sn, key_idx));
    for (int s = 0; s < n_in; s++) islot[s] = fed_slot(fcols, &nf, scan_file_col(sn, in_col[s]));
    CompiledExpr cx;
    memset(&cx, 0, sizeof cx);
    if (fn) {
        /* compiled against the scan's schema (names and types only) */
        const VecSchema *os = &sn->base.output_schema;
        VecBatch *shape = vec_batch_alloc(os->n_cols, 0);
        for (int k = 0; k < os->n_cols; k++) {
            shape->columns[k] = vec_array_alloc(os->col_types[k], 0);
            shape->col_names[k] = vx_strdup(os->col_names[k]);
        }
        const int rc = vx_try_compile_expr(fn->predicate, shape, &cx);
        vec_batch_free(shape);
        if (rc) vectra_error("%s", vx_last_error());
        if (cx.result_type != VEC_BOOL) vectra_error("filter predicate must evaluate to boolean");
        if (cx.n_scmp > 0) return 0;   /* string comparisons: the FilterNode's batches take them */
        for (int p = 0; p < cx.n_cols; p++) pslot[p] = fed_slot(fcols, &nf, scan_file_col(sn, cx.col_idx[p]));
    }
    vx_devscan_open(&g->ds, sn, fcols, nf);
    g->ds_open = 1;
    vgpu_col cols[VX_DS_COLS], ins[VGPU_MAX_INPUTS], pcols[VGPU_MAX_COLS];
    int64_t n, pos0;
    while (vx_devscan_next(&g->ds, cols, &n, &pos0)) {
        for (int s = 0; s < n_in; s++) ins[s] = cols[islot[s]];
        for (int p = 0; p < cx.n_cols; p++) pcols[p] = cols[pslot[p]];
        if (vgpu_groupby_feed(gb, kslot >= 0 ? &cols[kslot] : NULL, ins, n, NULL, 0, pos0, fn ? &cx.prog : NULL,
                              fn ? pcols : NULL, fn ? cx.n_cols : 0) < 0) {
            gb_drop(ga, gb);
            vectra_error("GPU: %s", vgpu_last_error());
        }
    }
    vx_devscan_close(&g->ds);
    g->ds_open = 0;
    return 1;
}

/* ---- several grouping columns ----------------------------------------
 * The groups of a key tuple (src/group_agg.c:155-281 with n_keys > 1;
 * equality per column as src/hash.c:48-78: NA equals NA, -0 equals +0, NaN
 * equals nothing) are the groups of a 64-bit code of the tuple
 * (vgpu_key_codes), grouped in first-occurrence order like one int64 key.
 * Every key column also feeds hidden aggregates -- min and max (exact
 * int64 bits) or, for bool, the sum, and the valid count -- which give the
 * group's key values back and prove that no two different tuples shared a
 * code: within a group each column must be all NA or all valid with
 * min == max (sum 0 or all TRUE). A string column enters the code as its
 * exact dictionary code (vgpu_strdict_append: equal bytes, equal code;
 * src/hash.c:58-66 compares bytes), which also feeds min / max (one string
 * per group, as exact as for int64), and the row position to min: the
 * batches' string columns stay on the host until the end, and each group's
 * string is read from that row. In key order
 * (temp_dir set) the groups are then sorted stably by the tuple, each
 * column ascending with NA last (NaN after the numbers, in
 * first-occurrence order; strings by bytes, a prefix first, src/sort.c:58-66). */

typedef struct {
    int       m;
    VecType   type[MK_MAX_KEYS];
    const double  *v[MK_MAX_KEYS];     /* key values (int64 as raw bits) */
    const uint8_t *ok[MK_MAX_KEYS];    /* 1 = valid */
    const char *const *sp[MK_MAX_KEYS];   /* string keys: bytes and length */
    const int64_t *sl[MK_MAX_KEYS];
} TupleCmp;

/* rank within a column: number, NaN, NA */
static int tuple_rank(const TupleCmp *t, int j, int64_t g) {
    if (!t->ok[j][g]) return 2;
    return t->type[j] == VEC_DOUBLE && t->v[j][g] != t->v[j][g] ? 1 : 0;
}

static int tuple_cmp(const TupleCmp *t, int64_t a, int64_t b) {
    for (int j = 0; j < t->m; j++) {
        const int ra = tuple_rank(t, j, a), rb = tuple_rank(t, j, b);
        if (ra != rb) return ra < rb ? -1 : 1;
        if (ra) continue;
        if (t->type[j] == VEC_STRING) {
            const int64_t la = t->sl[j][a], lb = t->sl[j][b];
            const int c = memcmp(t->sp[j][a], t->sp[j][b], (size_t)(la < lb ? la : lb));
            if (c) return c < 0 ? -1 : 1;
            if (la != lb) return la < lb ? -1 : 1;
        } else if (t->type[j] == VEC_INT64) {
            int64_t x, y;
            memcpy(&x, &t->v[j][a], 8);
            memcpy(&y, &t->v[j][b], 8);
            if (x != y) return x < y ? -1 : 1;
        } else if (t->v[j][a] != t->v[j][b]) {
            return t->v[j][a] < t->v[j][b] ? -1 : 1;
        }
    }
    return 0;
}

/* stable merge sort of idx[0..n) by tuple */
static void tuple_sort(const TupleCmp *t, int64_t *idx, int64_t *tmp, int64_t n) {
    for (int64_t w = 1; w < n; w *= 2) {
        for (int64_t lo = 0; lo < n; lo += 2 * w) {
            const int64_t mid = lo + w < n ? lo + w : n, hi = lo + 2 * w < n ? lo + 2 * w : n;
            int64_t i = lo, j = mid, k = lo;
            while (i < mid && j < hi) tmp[k++] = tuple_cmp(t, idx[j], idx[i]) < 0 ? idx[j++] : idx[i++];
            while (i < mid) tmp[k++] = idx[i++];
            while (j < hi) tmp[k++] = idx[j++];
        }
        memcpy(idx, tmp, (size_t)n * sizeof *idx);
    }
}

/* The same order on the device: a stable LSD radix sort of the group order
 * (first-seen to begin with), column by column from the last.  A number
 * column takes one pass by its value -- int64 with the sign bit flipped,
 * double by its IEEE bits mapped to an unsigned order (-0 as +0), bool 0 /
 * 1; a string column takes passes from its least significant key up: the
 * length, then its bytes in 8-byte big-endian chunks, zero-padded, from the
 * last chunk to the first (zero-padded order with the shorter string first
 * on a tie is memcmp order with a prefix first, src/sort.c:58-66); then,
 * where the column has NaN or NA, a pass by its class (number, NaN, NA).
 * A pass whose keys are all equal is skipped.  Groups equal on a column
 * keep their relative order, as in tuple_sort.  Returns 0 (nothing done)
 * for fewer than two groups or when device memory runs out. */
static uint64_t order_bits(const TupleCmp *t, int j, int64_t g) {
    if (t->type[j] == VEC_BOOL) return t->v[j][g] > 0;
    uint64_t u;
    if (t->type[j] == VEC_INT64) {
        memcpy(&u, &t->v[j][g], 8);
        return u ^ 0x8000000000000000ull;
    }
    const double x = t->v[j][g] == 0.0 ? 0.0 : t->v[j][g];
    memcpy(&u, &x, 8);
    return (u >> 63) ? ~u : u | 0x8000000000000000ull;
}

/* bytes [8c, 8c + 8) of group g's string, zero-padded, big-endian */
static uint64_t chunk_bits(const TupleCmp *t, int j, int64_t g, int64_t c) {
    const int64_t len = t->sl[j][g];
    const unsigned char *p = (const unsigned char *)t->sp[j][g];
    uint64_t u = 0;
    for (int b = 0; b < 8; b++) {
        const int64_t k = 8 * c + b;
        u = (u << 8) | (k < len ? p[k] : 0u);
    }
    return u;
}

/* the keys of one pass: pass -1 = the class; for strings 0 = the length
 * and c + 1 = chunk c; for numbers 0 = the value */
static int pass_keys(const TupleCmp *t, int j, int64_t pass, const uint32_t *hp, uint64_t *hk, int64_t n) {
    int differ = 0;
    for (int64_t i = 0; i < n; i++) {
        const int64_t g = hp[i];
        const int r = tuple_rank(t, j, g);
        uint64_t k;
        if (pass < 0) k = (uint64_t)r;
        else if (r) k = 0;
        else if (t->type[j] != VEC_STRING) k = order_bits(t, j, g);
        else k = pass == 0 ? (uint64_t)t->sl[j][g] : chunk_bits(t, j, g, pass - 1);
        hk[i] = k;
        differ |= i > 0 && k != hk[0];
    }
    return differ;
}

static int device_tuple_order(const TupleCmp *t, int64_t *order, int64_t n) {
    if (n < 2 || n >= ((int64_t)1 << 32)) return 0;
    uint64_t *hk = (uint64_t *)malloc((size_t)n * 8);
    uint32_t *hp = (uint32_t *)malloc((size_t)n * 4);
    uint64_t *dk = (uint64_t *)vgpu_malloc((size_t)n * 8);
    uint32_t *dp = (uint32_t *)vgpu_malloc((size_t)n * 4);
    if (!hk || !hp || !dk || !dp) {
        free(hk);
        free(hp);
        if (dk) vgpu_free(dk);
        if (dp) vgpu_free(dp);
        return 0;   /* the host sort takes it */
    }
    for (int64_t i = 0; i < n; i++) hp[i] = (uint32_t)order[i];
    int rc = 0;
    for (int j = t->m - 1; j >= 0 && rc == 0; j--) {
        int64_t chunks = 0;
        if (t->type[j] == VEC_STRING)
            for (int64_t g = 0; g < n; g++)
                if (t->ok[j][g] && (t->sl[j][g] + 7) / 8 > chunks) chunks = (t->sl[j][g] + 7) / 8;
        /* value passes, least significant first, then the class */
        const int64_t n_pass = t->type[j] == VEC_STRING ? 1 + chunks : 1;
        for (int64_t q = 0; q <= n_pass && rc == 0; q++) {
            const int64_t pass = q == n_pass ? -1 : t->type[j] == VEC_STRING ? (q == 0 ? 0 : n_pass - q) : 0;
            if (!pass_keys(t, j, pass, hp, hk, n)) continue;
            const int end_bit = pass < 0 ? 2 : t->type[j] == VEC_BOOL ? 1 : 64;
            if (vgpu_memcpy_h2d(dk, hk, (size_t)n * 8) < 0 || vgpu_memcpy_h2d(dp, hp, (size_t)n * 4) < 0 ||
                vgpu_sort_pairs_u64(dk, dp, n, 0, end_bit) < 0 || vgpu_memcpy_d2h(hp, dp, (size_t)n * 4) < 0)
                rc = -1;
        }
    }
    vgpu_free(dk);
    vgpu_free(dp);
    free(hk);
    if (rc == 0)
        for (int64_t i = 0; i < n; i++) order[i] = hp[i];
    free(hp);
    if (rc) vectra_error("GPU: %s", vgpu_last_error());
    return 1;
}

static int slot_for(const VecSchema *cs, int ci, int *in_col, int32_t *in_type, int *n_in) {
    for (int s = 0; s < *n_in; s++)
        if (in_col[s] == ci) return s;
    if (*n_in >= VGPU_MAX_INPUTS) vectra_error("GPU summarise: more than %d input columns", VGPU_MAX_INPUTS);
    in_col[*n_in] = ci;
    in_type[*n_in] = (int32_t)cs->col_types[ci];
    return (*n_in)++;
}


static int synth_slot(int *in_col, int32_t *in_type, int *n_in, int code) {
    if (*n_in >= VGPU_MAX_INPUTS) vectra_error("GPU summarise: more than %d input columns", VGPU_MAX_INPUTS);
    in_col[*n_in] = code;
    in_type[*n_in] = VGPU_INT64;
    return (*n_in)++;
}

static void kept_free(KeptStr **kept, int m, int64_t n) {
    for (int j = 0; j < m; j++) {
        if (!kept[j]) continue;
        for (int64_t i = 0; i < n; i++) vec_array_free(&kept[j][i].a);
        free(kept[j]);
        kept[j] = NULL;
    }
}

static VecBatch *gpu_group_run_multi(GroupAggNode *ga) {
    GroupGpu *g = (GroupGpu *)ga->gpu;
    const VecSchema *cs = &ga->child->output_schema;
    const int m = ga->n_keys;
    if (m > MK_MAX_KEYS) vectra_error("GPU group_by: %d grouping columns exceed %d", m, MK_MAX_KEYS);
    int kcol[MK_MAX_KEYS], kslot[MK_MAX_KEYS];
    VecType ktype[MK_MAX_KEYS];
    for (int j = 0; j < m; j++) {
        kcol[j] = vec_schema_find_col(cs, ga->key_names[j]);
        if (kcol[j] < 0) vectra_error("group_by: column not found: %s", ga->key_names[j]);
        ktype[j] = cs->col_types[kcol[j]];
    }
    int n_str = 0;
    for (int j = 0; j < m; j++) n_str += ktype[j] == VEC_STRING;
    int n_in = 0, in_col[VGPU_MAX_INPUTS];
    int32_t in_type[VGPU_MAX_INPUTS];
    vgpu_agg aggs[VGPU_MAX_AGGS];
    const int hidden = 1 + 3 * m + n_str;
    if (ga->n_aggs + hidden > VGPU_MAX_AGGS)
        vectra_error("GPU summarise: %d aggregates over %d grouping columns exceed the GPU path's %d (each "
                     "grouping column takes 3, a string column 4, plus one)", ga->n_aggs, m, VGPU_MAX_AGGS);
    for (int a = 0; a < ga->n_aggs; a++) {
        const AggSpec *sp = &ga->agg_specs[a];
        aggs[a].kind = (int32_t)sp->kind;
        aggs[a].na_rm = sp->na_rm;
        aggs[a].reserved = 0;
        aggs[a].input = -1;
        if (sp->kind == AGG_COUNT_STAR) continue;
        const int ci = vec_schema_find_col(cs, sp->input_col);
        if (ci < 0) vectra_error("summarise: column not found: %s", sp->input_col);
        if (cs->col_types[ci] == VEC_STRING)
            vectra_error("GPU summarise: string column '%s' is not aggregated on the GPU", sp->input_col);
        aggs[a].input = slot_for(cs, ci, in_col, in_type, &n_in);
    }
    /* a string key's slots are its dictionary code (min / max) and the row position */
    int pslot[MK_MAX_KEYS];
    for (int j = 0; j < m; j++) {
        if (ktype[j] == VEC_STRING) {
            kslot[j] = synth_slot(in_col, in_type, &n_in, -2 - 2 * j);
            pslot[j] = synth_slot(in_col, in_type, &n_in, -3 - 2 * j);
        } else {
            kslot[j] = slot_for(cs, kcol[j], in_col, in_type, &n_in);
            pslot[j] = -1;
        }
    }
    /* hidden aggregates: n, then per key column (min | sum), (max | -),
     * count, and for strings the smallest row position */
    int na = ga->n_aggs;
    const int a_n = na;
    aggs[na++] = (vgpu_agg){VGPU_AGG_COUNT_STAR, -1, 0, 0};
    int a_lo[MK_MAX_KEYS], a_hi[MK_MAX_KEYS], a_cnt[MK_MAX_KEYS], a_pos[MK_MAX_KEYS];
    for (int j = 0; j < m; j++) {
        a_pos[j] = -1;
        if (ktype[j] == VEC_STRING) {
            a_pos[j] = na;
            aggs[na++] = (vgpu_agg){VGPU_AGG_MIN, pslot[j], 0, VGPU_AGG_RAW};
        }
        if (ktype[j] == VEC_BOOL) {
            a_lo[j] = na;
            aggs[na++] = (vgpu_agg){VGPU_AGG_SUM, kslot[j], 0, 0};
            a_hi[j] = -1;
        } else {
            a_lo[j] = na;
            aggs[na++] = (vgpu_agg){VGPU_AGG_MIN, kslot[j], 0, VGPU_AGG_RAW};
            a_hi[j] = na;
            aggs[na++] = (vgpu_agg){VGPU_AGG_MAX, kslot[j], 0, VGPU_AGG_RAW};
        }
        a_cnt[j] = na;
        aggs[na++] = (vgpu_agg){VGPU_AGG_COUNT, kslot[j], 0, 0};
    }
    vgpu_groupby *gb = gb_hold(ga, vgpu_groupby_create(VGPU_INT64, n_in, in_type, na, aggs, VGPU_ORDER_FIRST_SEEN, 0));
    if (!gb) vectra_error("GPU: %s", vgpu_last_error());

    int64_t pos = 0, phys = 0;
    KeptStr **kept = g->kept;
    int64_t n_kept = 0, cap_kept = 0;
    VecBatch *b;
    while ((b = ga->child->next_batch(ga->child)) != NULL) {
        const int64_t nl = vec_batch_logical_rows(b);
        if (nl > 0) {
            vgpu_col ins[VGPU_MAX_INPUTS], keys[MK_MAX_KEYS];
            for (int s2 = 0; s2 < n_in; s2++) {
                if (in_col[s2] < 0) continue;   /* a string key's codes or positions: below */
                if (b->columns[in_col[s2]].type != (VecType)in_type[s2]) {
                    gb_drop(ga, gb);
                    vec_batch_free(b);
                    vectra_error("summarise: column type changed between batches");
                }
                devcol_upload(&g->in[s2], &b->columns[in_col[s2]], b->n_rows);
                ins[s2] = g->in[s2].col;
            }
            if (n_str > 0 && n_kept == cap_kept) {
                cap_kept = cap_kept ? 2 * cap_kept : 16;
                for (int j = 0; j < m; j++)
                    if (ktype[j] == VEC_STRING) {
                        kept[j] = (KeptStr *)realloc(kept[j], (size_t)cap_kept * sizeof(KeptStr));
                        if (!kept[j]) vectra_error("alloc failed");
                    }
            }
            for (int j = 0; j < m; j++) {
                if (ktype[j] != VEC_STRING) {
                    keys[j] = ins[kslot[j]];
                    continue;
                }
                VecArray *sa = &b->columns[kcol[j]];
                if (sa->type != VEC_STRING) {
                    gb_drop(ga, gb);
                    vec_batch_free(b);
                    kept_free(kept, m, n_kept);
                    vectra_error("group_by: column type changed between batches");
                }
                const int64_t nb = b->n_rows, bytes = sa->buf.str.offsets[nb];
                (void)bytes;
                uint8_t *d_val = NULL;
                if (!vx_all_valid(sa->validity, nb)) {
                    const size_t vb = (size_t)vec_validity_bytes(nb);
                    d_val = (uint8_t *)devmem_reserve(&g->s_val[j], vb + 8);
                    VX_GPU(vgpu_memcpy_h2d(d_val, sa->validity, vb));
                }
                /* the string's exact dictionary code stands for it in the
                 * tuple code and in the hidden min / max (equal bytes, equal
                 * code: one string per group is then exact, as for int64) */
                if (!g->sdict[j] && !(g->sdict[j] = vgpu_strdict_create())) vectra_error("GPU: %s", vgpu_last_error());
                int64_t *code = (int64_t *)devmem_reserve(&g->s_code[j], (size_t)(nb > 0 ? nb : 1) * 8);
                int64_t *ps = (int64_t *)devmem_reserve(&g->s_pos[j], (size_t)(nb > 0 ? nb : 1) * 8);
                VX_GPU(vgpu_strdict_append(g->sdict[j], sa->buf.str.offsets, (const uint8_t *)sa->buf.str.data, d_val,
                                           nb, phys, code, ps));
                keys[j] = (vgpu_col){VGPU_INT64, 0, code, d_val};
                ins[kslot[j]] = (vgpu_col){VGPU_INT64, 0, code, d_val};
                ins[pslot[j]] = (vgpu_col){VGPU_INT64, 0, ps, NULL};
                /* the column moves out of the batch (its bytes are copied
                 * when the batch only borrows them) */
                KeptStr *k = &kept[j][n_kept];
                k->a = *sa;
                k->base = phys;
                if (!sa->owns_data) {
                    char *d = (char *)malloc((size_t)(bytes > 0 ? bytes : 1));
                    if (!d) vectra_error("alloc failed");
                    if (bytes > 0) memcpy(d, sa->buf.str.data, (size_t)bytes);
                    k->a.buf.str.data = d;
                    k->a.owns_data = 1;
                }
                memset(sa, 0, sizeof *sa);
            }
            if (n_str > 0) g->n_kept = ++n_kept;
            int64_t *codes = (int64_t *)devmem_reserve(&g->codes, (size_t)b->n_rows * 8);
            vgpu_col key = {VGPU_INT64, 0, codes, NULL};
            const int32_t *d_sel = NULL;
            if (b->sel) {
                d_sel = (const int32_t *)devmem_reserve(&g->sel, (size_t)b->sel_n * 4);
                VX_GPU(vgpu_memcpy_h2d((void *)d_sel, b->sel, (size_t)b->sel_n * 4));
            }
            /* codes of NaN tuples come from physical positions, unique over batches */
            if (vgpu_key_codes(keys, m, b->n_rows, phys, codes, NULL) < 0 ||
                vgpu_groupby_feed(gb, &key, ins, b->n_rows, d_sel, b->sel ? b->sel_n : 0, pos, NULL, NULL, 0) < 0) {
                gb_drop(ga, gb);
                vec_batch_free(b);
                vectra_error("GPU: %s", vgpu_last_error());
            }
            VX_GPU(vgpu_synchronize());
        }
        pos += nl;
        phys += b->n_rows;
        vec_batch_free(b);
    }
    const int64_t ng = vgpu_groupby_finish(gb);
    if (ng < 0) {
        gb_drop(ga, gb);
        kept_free(kept, m, n_kept);
        vectra_error("GPU: %s", vgpu_last_error());
    }
    for (int j = 0; j < m; j++) {   /* fed: the groups' strings come from the kept batches */
        vgpu_strdict_free(g->sdict[j]);
        g->sdict[j] = NULL;
    }
    const size_t cap = (size_t)(ng > 0 ? ng : 1);
    double **vals = (double **)calloc((size_t)na, sizeof(double *));
    uint8_t **valid = (uint8_t **)calloc((size_t)na, sizeof(uint8_t *));
    uint64_t *codes_h = (uint64_t *)malloc(cap * 8);
    if (!vals || !valid || !codes_h) vectra_error("alloc failed");
    for (int a = 0; a < na; a++) {
        vals[a] = (double *)malloc(cap * 8);
        valid[a] = (uint8_t *)calloc((cap + 7) / 8, 1);
        if (!vals[a] || !valid[a]) vectra_error("alloc failed");
    }
    const int64_t warns = vgpu_groupby_fetch(gb, codes_h, NULL, vals, valid);
    gb_drop(ga, gb);
    free(codes_h);
    if (warns < 0) {
        kept_free(kept, m, n_kept);
        vectra_error("GPU: %s", vgpu_last_error());
    }
    /* the key values, and the check that every group holds one tuple */
    TupleCmp t;
    memset(&t, 0, sizeof t);
    t.m = m;
    uint8_t *kok[MK_MAX_KEYS];
    double *kv[MK_MAX_KEYS];
    const char **ksp[MK_MAX_KEYS] = {0};
    int64_t *ksl[MK_MAX_KEYS] = {0};
    int collided = 0;
    for (int j = 0; j < m; j++) {
        t.type[j] = ktype[j];
        kv[j] = (double *)malloc(cap * 8);
        kok[j] = (uint8_t *)malloc(cap);
        if (!kv[j] || !kok[j]) vectra_error("alloc failed");
        if (ktype[j] == VEC_STRING) {
            ksp[j] = (const char **)malloc(cap * sizeof(char *));
            ksl[j] = (int64_t *)malloc(cap * 8);
            if (!ksp[j] || !ksl[j]) vectra_error("alloc failed");
        }
        for (int64_t r = 0; r < ng; r++) {
            const double n = vals[a_n][r], c = vals[a_cnt[j]][r];
            kok[j][r] = c > 0;
            if (c != 0 && c != n) collided = 1;
            if (ksp[j]) {
                ksp[j][r] = "";
                ksl[j][r] = 0;
            }
            if (c == 0) {
                kv[j][r] = 0;
                continue;
            }
            const double lo = vals[a_lo[j]][r];
            if (ktype[j] == VEC_STRING) {
                if (memcmp(&lo, &vals[a_hi[j]][r], 8) != 0) collided = 1;
                /* the group's string, at its smallest row position */
                int64_t p;
                memcpy(&p, &vals[a_pos[j]][r], 8);
                int64_t lo_i = 0, hi_i = n_kept - 1;
                while (lo_i < hi_i) {
                    const int64_t mid = (lo_i + hi_i + 1) / 2;
                    if (kept[j][mid].base <= p) lo_i = mid;
                    else hi_i = mid - 1;
                }
                const VecArray *sa = n_kept > 0 ? &kept[j][lo_i].a : NULL;
                const int64_t row = n_kept > 0 ? p - kept[j][lo_i].base : -1;
                if (!sa || row < 0 || row >= sa->length || !vec_array_is_valid(sa, row)) {
                    collided = 1;
                } else {
                    ksp[j][r] = sa->buf.str.data + sa->buf.str.offsets[row];
                    ksl[j][r] = sa->buf.str.offsets[row + 1] - sa->buf.str.offsets[row];
                }
                kv[j][r] = 0;
            } else if (ktype[j] == VEC_BOOL) {
                if (lo != 0 && lo != c) collided = 1;
                kv[j][r] = lo > 0 ? 1.0 : 0.0;
            } else if (ktype[j] == VEC_INT64) {
                if (memcmp(&lo, &vals[a_hi[j]][r], 8) != 0) collided = 1;
                kv[j][r] = lo;
            } else {
                if (n > 1 && !(lo == vals[a_hi[j]][r])) collided = 1;
                kv[j][r] = lo;
            }
        }
        t.v[j] = kv[j];
        t.ok[j] = kok[j];
        t.sp[j] = ksp[j];
        t.sl[j] = ksl[j];
    }
    int nan_sorted = 0;
    for (int j = 0; j < m && ga->use_sorted; j++)
        for (int64_t r = 0; r < ng && ktype[j] == VEC_DOUBLE && !nan_sorted; r++)
            nan_sorted = kok[j][r] && isnan(kv[j][r]);
    if (collided || nan_sorted) {
        for (int a = 0; a < na; a++) { free(vals[a]); free(valid[a]); }
        for (int j = 0; j < m; j++) { free(kv[j]); free(kok[j]); free(ksp[j]); free(ksl[j]); }
        free(vals);
        free(valid);
        kept_free(kept, m, n_kept);
        if (nan_sorted) vectra_error("%s", NAN_KEY_ORDER_ERROR);
        vectra_error("GPU group_by: two different key tuples share a 64-bit code (probability about 2^-64 "
                     "per pair of groups)");
    }
    int64_t *order = (int64_t *)malloc(cap * sizeof(int64_t));
    int64_t *tmp = (int64_t *)malloc(cap * sizeof(int64_t));
    if (!order || !tmp) vectra_error("alloc failed");
    for (int64_t r = 0; r < ng; r++) order[r] = r;
    if (ga->use_sorted && !device_tuple_order(&t, order, ng)) tuple_sort(&t, order, tmp, ng);
    free(tmp);
    VecBatch *res = vec_batch_alloc(m + ga->n_aggs, ng);
    for (int j = 0; j < m; j++) {
        VecArray a = vec_array_alloc(ktype[j], ng);
        if (ktype[j] == VEC_STRING) {
            int64_t total = 0;
            for (int64_t r = 0; r < ng; r++)
                if (kok[j][order[r]]) total += ksl[j][order[r]];
            a.buf.str.data = (char *)malloc((size_t)(total > 0 ? total : 1));
            if (!a.buf.str.data) vectra_error("alloc failed");
            a.buf.str.data_len = total;
            int64_t at = 0;
            for (int64_t r = 0; r < ng; r++) {
                const int64_t s2 = order[r];
                a.buf.str.offsets[r] = at;
                if (!kok[j][s2]) continue;
                vec_array_set_valid(&a, r);
                memcpy(a.buf.str.data + at, ksp[j][s2], (size_t)ksl[j][s2]);
                at += ksl[j][s2];
            }
            a.buf.str.offsets[ng] = at;
        }
        for (int64_t r = 0; r < ng && ktype[j] != VEC_STRING; r++) {
            const int64_t s2 = order[r];
            if (!kok[j][s2]) continue;
            vec_array_set_valid(&a, r);
            if (ktype[j] == VEC_BOOL) a.buf.bln[r] = kv[j][s2] > 0;
            else if (ktype[j] == VEC_INT64) memcpy(&a.buf.i64[r], &kv[j][s2], 8);
            else a.buf.dbl[r] = kv[j][s2];
        }
        res->columns[j] = a;
        res->col_names[j] = vx_strdup(ga->key_names[j]);
    }
    for (int a = 0; a < ga->n_aggs; a++) {
        VecArray o = vec_array_alloc(VEC_DOUBLE, ng);
        for (int64_t r = 0; r < ng; r++) {
            const int64_t s2 = order[r];
            o.buf.dbl[r] = vals[a][s2];
            if ((valid[a][s2 >> 3] >> (s2 & 7)) & 1) vec_array_set_valid(&o, r);
        }
        res->columns[m + a] = o;
        res->col_names[m + a] = vx_strdup(ga->agg_specs[a].output_name);
    }
    free(order);
    for (int a = 0; a < na; a++) { free(vals[a]); free(valid[a]); }
    for (int j = 0; j < m; j++) { free(kv[j]); free(kok[j]); free(ksp[j]); free(ksl[j]); }
    free(vals);
    free(valid);
    kept_free(kept, m, n_kept);
    for (int64_t w = 0; w < warns; w++)
        vectra_warning("no non-missing arguments to min/max; returning Inf/-Inf");
    return res;
}

/* finishes gb and builds the result batch: the key column (n_keys == 1),
 * then one double column per aggregate; destroys gb */
static VecBatch *group_result(GroupAggNode *ga, vgpu_groupby *gb, VecType key_type) {
    int64_t ng = vgpu_groupby_finish(gb);
    if (ng < 0) {
        gb_drop(ga, gb);
        vectra_error("GPU: %s", vgpu_last_error());
    }
    int n_out = ga->n_keys + ga->n_aggs;
    VecBatch *res = vec_batch_alloc(n_out, ng);
    double **vals = (double **)calloc((size_t)(ga->n_aggs + 1), sizeof(double *));
    uint8_t **valid = (uint8_t **)calloc((size_t)(ga->n_aggs + 1), sizeof(uint8_t *));
    for (int a = 0; a < ga->n_aggs; a++) {
        res->columns[ga->n_keys + a] = vec_array_alloc(VEC_DOUBLE, ng);
        vals[a] = res->columns[ga->n_keys + a].buf.dbl;
        valid[a] = res->columns[ga->n_keys + a].validity;
        res->col_names[ga->n_keys + a] = vx_strdup(ga->agg_specs[a].output_name);
    }
    uint64_t *kbuf = NULL;
    if (ga->n_keys == 1) {
        res->columns[0] = vec_array_alloc(key_type, ng);
        res->col_names[0] = vx_strdup(ga->key_names[0]);
        kbuf = key_type == VEC_BOOL ? (uint64_t *)malloc((size_t)(ng > 0 ? ng : 1) * 8)
                                    : (uint64_t *)res->columns[0].buf.i64;
    }
    int64_t warns = vgpu_groupby_fetch(gb, kbuf, ga->n_keys ? res->columns[0].validity : NULL, vals, valid);
    gb_drop(ga, gb);
    if (warns < 0) {
        free(vals); free(valid);
        vec_batch_free(res);
        vectra_error("GPU: %s", vgpu_last_error());
    }
    if (ga->n_keys == 1 && key_type == VEC_BOOL) {
        for (int64_t i = 0; i < ng; i++) res->columns[0].buf.bln[i] = (uint8_t)kbuf[i];
        free(kbuf);
    }
    if (ga->use_sorted && ga->n_keys == 1 && key_type == VEC_DOUBLE) {
        for (int64_t i = 0; i < ng; i++)
            if (vec_array_is_valid(&res->columns[0], i) && isnan(res->columns[0].buf.dbl[i])) {
                free(vals);
                free(valid);
                vec_batch_free(res);
                vectra_error("%s", NAN_KEY_ORDER_ERROR);
            }
    }
    free(vals);
    free(valid);
    /* one warning per min/max group without a value (src/agg_ops.c:218-239) */
    for (int64_t w = 0; w < warns; w++)
        vectra_warning("no non-missing arguments to min/max; returning Inf/-Inf");
    return res;
}

/* ---- GroupAgg <- Project(tmp = a * b) <- Join, fused -----------------
 * summarise(s = sum(value * weight)) over a join is, in R, a GroupAggNode
 * over the hidden ProjectNode that computes .vectra_tmp_s = value * weight
 * over the JoinNode (R/verbs.R:266-284, src/project.c:51-252,
 * src/join.c:262-480).  When the Project's entries the aggregation reads
 * are join columns or products of two, and the join is a GPU join of one
 * key, this node has the join build its table and then streams the join's
 * left child through vgpu_groupby_feed_join: probe, product and
 * aggregation in one device pass, no join rows materialized.  Group order
 * is the join output's: feed k's pairs take positions k * 2^44 + (probe
 * row, match) with each batch's unmatched left-join rows after its matches
 * (src/join.c:439-448). */

typedef struct {
    int op;              /* 0 = one join column, 1 = product of two */
    int side_a, col_a;   /* side 0 = left (probe), 1 = right (build) */
    int side_b, col_b;
} FjSrc;

static int fj_col(const VecNode *join, const char *name, int *side, int *col) {
    int oc = vec_schema_find_col(&join->output_schema, name);
    return oc < 0 ? -1 : vx_gpu_join_output_source(join, oc, side, col);
}

static VecType fj_type(const JoinNode *jn, int side, int col) {
    const VecSchema *sc = side == 0 ? &jn->left->output_schema : &jn->right->output_schema;
    return sc->col_types[col];
}

/* the source of Project output column `name`, or -1 when it is not a join
 * column (pass-through or reference) or, with allow_product, a product of
 * two numeric join columns */
static int fj_resolve(const ProjectNode *pn, const char *name, int allow_product, FjSrc *src) {
    const JoinNode *jn = (const JoinNode *)pn->child;
    memset(src, 0, sizeof *src);
    for (int i = 0; i < pn->n_entries; i++) {
        const ProjEntry *pe = &pn->entries[i];
        if (strcmp(pe->output_name, name) != 0) continue;
        const VecExpr *e = pe->expr;
        if (!e) return fj_col(pn->child, name, &src->side_a, &src->col_a);
        if (e->kind == EXPR_COL_REF) return fj_col(pn->child, e->col_name, &src->side_a, &src->col_a);
        if (!allow_product || e->kind != EXPR_ARITH || e->op != '*' || !e->left || !e->right ||
            e->left->kind != EXPR_COL_REF || e->right->kind != EXPR_COL_REF)
            return -1;
        if (fj_col(pn->child, e->left->col_name, &src->side_a, &src->col_a) ||
            fj_col(pn->child, e->right->col_name, &src->side_b, &src->col_b))
            return -1;
        VecType ta = fj_type(jn, src->side_a, src->col_a), tb = fj_type(jn, src->side_b, src->col_b);
        if ((ta != VEC_INT64 && ta != VEC_DOUBLE) || (tb != VEC_INT64 && tb != VEC_DOUBLE)) return -1;
        src->op = 1;
        return 0;
    }
    return -1;
}

/* slot of col in a side's compact column list, appended if absent; the
 * probe side's slot 0 holds the (possibly cast) probe key alone, so probe
 * columns search from slot 1 */
static int fj_slot(int *cols, int *n, int col, int first) {
    for (int i = first; i < *n; i++)
        if (cols[i] == col) return i;
    if (*n >= VGPU_MAX_COLS) return -1;
    cols[*n] = col;
    return (*n)++;
}

/* Returns the result batch, or NULL when the tree below is not the fused
 * shape (then nothing has been pulled; the join may have built its table). */
static VecBatch *gpu_group_run_fused(GroupAggNode *ga) {
    GroupGpu *g = (GroupGpu *)ga->gpu;
    /* knob "gb.fuse_join" = 1 runs the tree node by node */
    if (vgpu_tune_get("gb.fuse_join") < 2) return NULL;
    if (!ga->child->kind || strcmp(ga->child->kind, "ProjectNode") != 0) return NULL;
    const ProjectNode *pn = (const ProjectNode *)ga->child;
    if (!vx_gpu_join_fusable(pn->child) || ga->n_keys > 1 || ga->n_aggs > VGPU_MAX_AGGS) return NULL;
    const JoinNode *jn = (const JoinNode *)pn->child;
    /* probe columns: slot 0 is the probe key; build columns by reference */
    int pcols[VGPU_MAX_COLS], npc = 1, bcols[VGPU_MAX_COLS], nbc = 0;
    pcols[0] = jn->keys[0].left_col;
    int key_side = 0, key_slot = 0;
    VecType key_type = VEC_INT64;
    if (ga->n_keys == 1) {
        FjSrc k;
        if (fj_resolve(pn, ga->key_names[0], 0, &k)) return NULL;
        key_type = fj_type(jn, k.side_a, k.col_a);
        if (key_type == VEC_STRING) return NULL;
        key_side = k.side_a;
        key_slot = k.side_a == 0 ? fj_slot(pcols, &npc, k.col_a, 1) : fj_slot(bcols, &nbc, k.col_a, 0);
        if (key_slot < 0) return NULL;
    }
    vgpu_agg aggs[VGPU_MAX_AGGS];
    vgpu_input ins[VGPU_MAX_INPUTS];
    int32_t in_type[VGPU_MAX_INPUTS];
    FjSrc in_src[VGPU_MAX_INPUTS];
    int n_in = 0;
    for (int a = 0; a < ga->n_aggs; a++) {
        const AggSpec *sp = &ga->agg_specs[a];
        aggs[a].kind = (int32_t)sp->kind;
        aggs[a].na_rm = sp->na_rm;
        aggs[a].reserved = 0;
        aggs[a].input = -1;
        if (sp->kind == AGG_COUNT_STAR) continue;
        FjSrc src;
        if (!sp->input_col || fj_resolve(pn, sp->input_col, 1, &src)) return NULL;
        if (src.op == 0 && fj_type(jn, src.side_a, src.col_a) == VEC_STRING) return NULL;
        int slot = -1;
        for (int i = 0; i < n_in; i++)
            if (!memcmp(&in_src[i], &src, sizeof src)) slot = i;
        if (slot < 0) {
            if (n_in >= VGPU_MAX_INPUTS) return NULL;
            slot = n_in++;
            in_src[slot] = src;
            VecType ta = fj_type(jn, src.side_a, src.col_a);
            if (src.op == 1) {
                VecType tb = fj_type(jn, src.side_b, src.col_b);
                in_type[slot] = (ta == VEC_DOUBLE || tb == VEC_DOUBLE) ? VGPU_DOUBLE : VGPU_INT64;
            } else {
                in_type[slot] = (int32_t)ta;
            }
            vgpu_input *in = &ins[slot];
            memset(in, 0, sizeof *in);
            in->op = src.op;
            in->side_a = src.side_a;
            in->col_a = src.side_a == 0 ? fj_slot(pcols, &npc, src.col_a, 1) : fj_slot(bcols, &nbc, src.col_a, 0);
            if (src.op == 1) {
                in->side_b = src.side_b;
                in->col_b = src.side_b == 0 ? fj_slot(pcols, &npc, src.col_b, 1) : fj_slot(bcols, &nbc, src.col_b, 0);
            }
            if (in->col_a < 0 || (src.op == 1 && in->col_b < 0)) return NULL;
        }
        aggs[a].input = slot;
    }
    for (int i = 1; i < npc; i++)
        if (jn->left->output_schema.col_types[pcols[i]] == VEC_STRING) return NULL;

    /* the shape fits: build the join table with the build columns embedded */
    VxJoinFusion fu;
    vx_gpu_join_prepare_fusion(pn->child, bcols, nbc, &fu);
    /* The fused pass orders at most 4,096 matches per probe row, and a
     * coerced left join continues its chains on uncoerced key bytes
     * (src/join.c:398-411): those joins run unfused, through the Project. */
    if (fu.max_dup > 4096 || (fu.coerced && fu.kind == JOIN_LEFT && fu.max_dup > 1)) return NULL;

    int order = ga->use_sorted ? VGPU_ORDER_KEY : VGPU_ORDER_FIRST_SEEN;
    vgpu_groupby *gb = gb_hold(ga, vgpu_groupby_create(ga->n_keys ? (int)key_type : VGPU_KEY_NONE, n_in, in_type, ga->n_aggs,
                                           aggs, order, 0));
    if (!gb) vectra_error("GPU: %s", vgpu_last_error());
    vgpu_col bdev[VGPU_MAX_COLS];
    for (int i = 0; i < nbc; i++) bdev[i] = fu.r_dev[bcols[i]].col;
    VecNode *left = fu.left;
    const VecSchema *ls = &left->output_schema;
    int64_t feed = 0;
    VecBatch *b;
    while ((b = left->next_batch(left)) != NULL) {
        int64_t nl = vec_batch_logical_rows(b);
        if (nl > 0) {
            vgpu_col pdev[VGPU_MAX_COLS];
            for (int i = 0; i < npc; i++) {
                if (b->columns[pcols[i]].type != ls->col_types[pcols[i]]) {
                    gb_drop(ga, gb);
                    vec_batch_free(b);
                    vectra_error("join: column type changed between batches");
                }
                devcol_upload(&g->fj_probe[i], &b->columns[pcols[i]], b->n_rows);
                pdev[i] = g->fj_probe[i].col;
            }
            if (fu.coerced) {
                devcol_reserve(&g->fj_cast, (size_t)(b->n_rows > 0 ? b->n_rows : 1) * 8, 0);
                if (vgpu_cast(&pdev[0], b->n_rows, (int)fu.build_key_type, g->fj_cast.d_data) < 0) {
                    gb_drop(ga, gb);
                    vec_batch_free(b);
                    vectra_error("GPU: %s", vgpu_last_error());
                }
                pdev[0].data = g->fj_cast.d_data;
                pdev[0].type = (int32_t)fu.build_key_type;
            }
            const int32_t *d_sel = NULL;
            if (b->sel) {
                d_sel = (const int32_t *)devmem_reserve(&g->sel, (size_t)(b->sel_n > 0 ? b->sel_n : 1) * 4);
                VX_GPU(vgpu_memcpy_h2d((void *)d_sel, b->sel, (size_t)b->sel_n * 4));
            }
            vgpu_feed_spec sp;
            memset(&sp, 0, sizeof sp);
            sp.n_rows = b->n_rows;
            sp.sel = d_sel;
            sp.n_sel = b->sel ? b->sel_n : 0;
            sp.pos_base = feed << 44;
            sp.join = fu.jt;
            sp.join_kind = fu.kind == JOIN_LEFT ? VGPU_JOIN_LEFT : VGPU_JOIN_INNER;
            sp.key_side = key_side;
            sp.key_col = key_slot;
            sp.probe_key_col = 0;
            sp.n_probe_cols = npc;
            sp.n_build_cols = nbc;
            sp.probe_cols = pdev;
            sp.build_cols = bdev;
            sp.n_inputs = n_in;
            sp.inputs = ins;
            if (vgpu_groupby_feed_join(gb, &sp) < 0) {
                gb_drop(ga, gb);
                vec_batch_free(b);
                vectra_error("GPU: %s", vgpu_last_error());
            }
            /* the next batch's uploads reuse the probe buffers */
            VX_GPU(vgpu_synchronize());
        }
        feed++;
        vec_batch_free(b);
    }
    g->fused = 1;
    return group_result(ga, gb, key_type);
}

static VecBatch *gpu_group_run(GroupAggNode *ga) {
    GroupGpu *g = (GroupGpu *)ga->gpu;
    const VecSchema *cs = &ga->child->output_schema;
    VecBatch *fused = gpu_group_run_fused(ga);
    if (fused) return fused;
    if (ga->n_keys > 1) return gpu_group_run_multi(ga);
    if (ga->n_keys == 1) {
        const int ki = vec_schema_find_col(cs, ga->key_names[0]);
        if (ki >= 0 && cs->col_types[ki] == VEC_STRING) return gpu_group_run_multi(ga);
    }
    int key_idx = -1;
    VecType key_type = VEC_INT64;
    if (ga->n_keys == 1) {
        key_idx = vec_schema_find_col(cs, ga->key_names[0]);
        if (key_idx < 0) vectra_error("group_by: column not found: %s", ga->key_names[0]);
        key_type = cs->col_types[key_idx];
        if (key_type == VEC_STRING)
            vectra_error("GPU group_by: string key '%s' is not grouped on the GPU", ga->key_names[0]);
    }
    if (ga->n_aggs > VGPU_MAX_AGGS)
        vectra_error("GPU summarise: %d aggregates exceed %d", ga->n_aggs, VGPU_MAX_AGGS);
    /* input slots: one per distinct input column */
    int n_in = 0, in_col[VGPU_MAX_INPUTS];
    int32_t in_type[VGPU_MAX_INPUTS];
    vgpu_agg aggs[VGPU_MAX_AGGS];
    for (int a = 0; a < ga->n_aggs; a++) {
        const AggSpec *sp = &ga->agg_specs[a];
        aggs[a].kind = (int32_t)sp->kind;
        aggs[a].na_rm = sp->na_rm;
        aggs[a].reserved = 0;
        aggs[a].input = -1;
        if (sp->kind == AGG_COUNT_STAR) continue;
        int ci = vec_schema_find_col(cs, sp->input_col);
        if (ci < 0) vectra_error("summarise: column not found: %s", sp->input_col);
        if (cs->col_types[ci] == VEC_STRING)
            vectra_error("GPU summarise: string column '%s' is not aggregated on the GPU", sp->input_col);
        int slot = -1;
        for (int s = 0; s < n_in; s++)
            if (in_col[s] == ci) { slot = s; break; }
        if (slot < 0) {
            if (n_in >= VGPU_MAX_INPUTS)
                vectra_error("GPU summarise: more than %d input columns", VGPU_MAX_INPUTS);
            slot = n_in++;
            in_col[slot] = ci;
            in_type[slot] = (int32_t)cs->col_types[ci];
        }
        aggs[a].input = slot;
    }
    int order = ga->use_sorted ? VGPU_ORDER_KEY : VGPU_ORDER_FIRST_SEEN;
    const ScanNode *below = scan_below(ga->child);
    const int64_t hint = below && key_idx >= 0 ? scan_group_bound(below, key_idx) : 0;
    vgpu_groupby *gb = gb_hold(ga, vgpu_groupby_create(ga->n_keys ? (int)key_type : VGPU_KEY_NONE, n_in, in_type,
                                           ga->n_aggs, aggs, order, hint));
    if (!gb) vectra_error("GPU: %s", vgpu_last_error());

    int64_t pos = 0;
    VecBatch *b;
    const int from_scan = feed_from_scan(ga, gb, key_idx, n_in, in_col);
    while (!from_scan && (b = ga->child->next_batch(ga->child)) != NULL) {
        int64_t nl = vec_batch_logical_rows(b);
        if (nl > 0) {
            vgpu_col key, ins[VGPU_MAX_INPUTS];
            if (key_idx >= 0) {
                devcol_upload(&g->key, &b->columns[key_idx], b->n_rows);
                key = g->key.col;
            }
            for (int s = 0; s < n_in; s++) {
                if (b->columns[in_col[s]].type != (VecType)in_type[s]) {
                    gb_drop(ga, gb);
                    vectra_error("summarise: column type changed between batches");
                }
                devcol_upload(&g->in[s], &b->columns[in_col[s]], b->n_rows);
                ins[s] = g->in[s].col;
            }
            const int32_t *d_sel = NULL;
            if (b->sel) {
                d_sel = (const int32_t *)devmem_reserve(&g->sel, (size_t)b->sel_n * 4);
                VX_GPU(vgpu_memcpy_h2d((void *)d_sel, b->sel, (size_t)b->sel_n * 4));
            }
            if (vgpu_groupby_feed(gb, key_idx >= 0 ? &key : NULL, ins, b->n_rows, d_sel, b->sel ? b->sel_n : 0,
                                  pos, NULL, NULL, 0) < 0) {
                gb_drop(ga, gb);
                vec_batch_free(b);
                vectra_error("GPU: %s", vgpu_last_error());
            }
            /* the uploads above are synchronous, so the buffers may be
             * reused by the next batch once the feed kernel has run */
            VX_GPU(vgpu_synchronize());
        }
        pos += nl;
        vec_batch_free(b);
    }

    return group_result(ga, gb, key_type);
}

static VecBatch *gpu_group_next(VecNode *self) {
    GroupAggNode *ga = (GroupAggNode *)self;
    if (ga->done) return NULL;
    ga->done = 1;
    return gpu_group_run(ga);
}

static void gpu_group_free(VecNode *self) {
    GroupAggNode *ga = (GroupAggNode *)self;
    GroupGpu *g = (GroupGpu *)ga->gpu;
    ga->child->free_node(ga->child);
    for (int k = 0; k < ga->n_keys; k++) free(ga->key_names[k]);
    free(ga->key_names);
    for (int a = 0; a < ga->n_aggs; a++) {
        free(ga->agg_specs[a].output_name);
        free(ga->agg_specs[a].input_col);
    }
    free(ga->agg_specs);
    if (g) {
        if (g->live_gb) vgpu_groupby_destroy(g->live_gb);
        kept_free(g->kept, MK_MAX_KEYS, g->n_kept);
        devcol_free(&g->key);
        for (int i = 0; i < VGPU_MAX_INPUTS; i++) devcol_free(&g->in[i]);
        for (int i = 0; i < VGPU_MAX_COLS; i++) devcol_free(&g->fj_probe[i]);
        devcol_free(&g->fj_cast);
        devmem_free(&g->sel);
        devmem_free(&g->codes);
        for (int j = 0; j < VGPU_MAX_COLS; j++) {
            devmem_free(&g->s_val[j]);
            devmem_free(&g->s_code[j]);
            devmem_free(&g->s_pos[j]);
            vgpu_strdict_free(g->sdict[j]);
            g->sdict[j] = NULL;
        }
        if (g->ds_open) vx_devscan_close(&g->ds);
        free(g);
    }
    vec_schema_free(&ga->base.output_schema);
    free(ga);
}

GroupAggNode *group_agg_node_create(VecNode *child, int n_keys, char **key_names, int n_aggs,
                                    AggSpec *agg_specs, const char *temp_dir) {
    vx_gpu_init();
    GroupAggNode *ga = (GroupAggNode *)calloc(1, sizeof *ga);
    if (!ga) vectra_error("alloc failed for GroupAggNode");
    ga->gpu = calloc(1, sizeof(GroupGpu));
    if (!ga->gpu) vectra_error("alloc failed for GroupAggNode");
    const VecSchema *cs = &child->output_schema;
    if (temp_dir && n_keys > 0) {
        /* the reference resolves the sort keys here and fails early */
        for (int k = 0; k < n_keys; k++)
            if (vec_schema_find_col(cs, key_names[k]) < 0)
                vectra_error("group_by: column not found: %s", key_names[k]);
        ga->use_sorted = 1;
    }
    ga->child = child;
    ga->n_keys = n_keys;
    ga->key_names = key_names;
    ga->n_aggs = n_aggs;
    ga->agg_specs = agg_specs;
    ga->done = 0;
    int n_out = n_keys + n_aggs;
    char **names = (char **)malloc(sizeof(char *) * (size_t)(n_out > 0 ? n_out : 1));
    VecType *types = (VecType *)malloc(sizeof(VecType) * (size_t)(n_out > 0 ? n_out : 1));
    for (int k = 0; k < n_keys; k++) {
        int idx = vec_schema_find_col(cs, key_names[k]);
        names[k] = key_names[k];
        types[k] = idx >= 0 ? cs->col_types[idx] : VEC_DOUBLE;
    }
    for (int a = 0; a < n_aggs; a++) {
        names[n_keys + a] = agg_specs[a].output_name;
        types[n_keys + a] = VEC_DOUBLE;
    }
    ga->base.output_schema = vec_schema_create(n_out, names, types);
    free(names);
    free(types);
    ga->base.next_batch = gpu_group_next;
    ga->base.free_node = gpu_group_free;
    ga->base.kind = "GroupAggNode";
    return ga;
}

int vx_gpu_group_note(const VecNode *n, char *buf, int size) {
    if (n->next_batch != gpu_group_next) return -1;
    const GroupAggNode *ga = (const GroupAggNode *)n;
    const VecSchema *cs = &ga->child->output_schema;
    int strs = 0;
    for (int k = 0; k < ga->n_keys; k++) {
        const int i = vec_schema_find_col(cs, ga->key_names[k]);
        strs += i >= 0 && cs->col_types[i] == VEC_STRING;
    }
    const int tuple = ga->n_keys > 1 || strs > 0;
    const VecNode *c = ga->child;
    const int filtered = c->kind && strcmp(c->kind, "FilterNode") == 0;
    if (filtered) c = ((const FilterNode *)c)->child;
    const int feed = !tuple && c->kind && strcmp(c->kind, "ScanNode") == 0 && vgpu_tune_get("scan.device_feed") > 1;
    /* over Project <- GPU Join: fused at run time when the shape holds */
    const GroupGpu *g = (const GroupGpu *)ga->gpu;
    const int over_join = ga->child->kind && strcmp(ga->child->kind, "ProjectNode") == 0 &&
                          vx_gpu_join_note(((const ProjectNode *)ga->child)->child, NULL, 0) >= 0;
    const char *fuse = g && g->fused ? ", fused with the Project and Join below (probe, product and aggregate in one "
                                       "device pass)"
                       : over_join && vgpu_tune_get("gb.fuse_join") >= 2
                           ? ", fuses the Project and Join below when their shape allows" : "";
    return snprintf(buf, (size_t)size, "gpu: %s group-by%s, %s%s%s", tuple ? "tuple-code" : "hash",
                    strs ? " (string keys by exact device dictionary codes)" : "", ga->use_sorted ? "key order" : "first-seen order",
                    feed ? (filtered ? ", device scan feed with the filter in the feed" : ", device scan feed") : "",
                    fuse);
}
