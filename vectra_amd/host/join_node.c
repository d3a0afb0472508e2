/*
 * join_node.c -- drop-in GPU JoinNode (src/join.h:45-78).
 *
 * Same constructor checks and messages, struct prefix, output schema (all
 * left columns, then the right non-key columns, suffixed on name
 * collisions; src/join.c:631-791) and batch protocol as the reference:
 * BUILD pulls the whole right side once, PROBE emits one output batch per
 * left batch (skipping empty ones), FULL then emits the never-matched
 * build rows in batches of at most 65,536 (src/join.c:487-599).
 *
 * The build side lives on the device for the node's lifetime; each probe
 * batch is copied up, probed (count / scan / ordered write), and the output
 * columns are gathered on the device and copied back as malloc'd arrays.
 * String payload columns are gathered on the host from the pair list.
 */
#include "gpu_common.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define FINALIZE_ROWS 65536   /* src/join.c:487 */
#define JN_MAX_KEYS 8

typedef struct {
    int       built;
    int       state;          /* 0 probe, 1 finalize, 2 done */
    int       r_ncols;
    VecArray *r_host;         /* materialized right side (host) */
    int64_t   n_build;
    DevCol   *r_dev;          /* right columns on the device (key coerced) */
    VecType   build_key_type;
    vgpu_jtable *jt;
    int      *r_nonkey;
    int       n_r_nonkey;
    int       first_only;     /* see join_build_side */
    int       pay_cols[4];    /* right columns embedded in the table entries */
    int       n_pay;          /* (set by a fused GroupAgg before the build) */
    DevCol    l_dev[64];
    DevCol    pkey_cast;
    DevMem    sel;
    DevMem    out_data, out_valid;
    int64_t  *h_build;        /* FULL finalize: unmatched build rows */
    int64_t   n_unmatched, fin_cursor;
    /* several key columns: the table is keyed by tuple codes */
    int       multi;
    VecType   common[JN_MAX_KEYS];
    DevCol    pcast[JN_MAX_KEYS];
    DevMem    bcode, bvalid, pcode, pvalid;
    /* string key columns: the build side's strings in an exact dictionary
     * (vgpu_strdict), the build rows' codes kept for the pair check, and
     * each probe batch's codes (looked up, -1 for strings not in the build) */
    vgpu_strdict *sdict[JN_MAX_KEYS];
    DevMem    s_bcode[JN_MAX_KEYS], s_bval[JN_MAX_KEYS], s_pcode[JN_MAX_KEYS], s_pval[JN_MAX_KEYS];
    vgpu_col  s_bcol[JN_MAX_KEYS];
} JoinGpu;

/* the rows [0, n) of host string column a as a device int64 column of
 * exact dictionary codes with the column's validity (equal bytes, equal
 * code; src/join.c:96-108 compares bytes): the build side appends its
 * strings, a probe batch looks them up */
static vgpu_col string_key_col(JoinGpu *g, int k, const VecArray *a, int64_t n, int build) {
    DevMem *cm = build ? &g->s_bcode[k] : &g->s_pcode[k], *vm = build ? &g->s_bval[k] : &g->s_pval[k];
    int64_t *codes = (int64_t *)devmem_reserve(cm, (size_t)(n > 0 ? n : 1) * 8);
    uint8_t *d_val = NULL;
    if (build && !g->sdict[k] && !(g->sdict[k] = vgpu_strdict_create())) vectra_error("GPU: %s", vgpu_last_error());
    if (n > 0) {
        if (!vx_all_valid(a->validity, n)) {
            const size_t vb = (size_t)vec_validity_bytes(n);
            d_val = (uint8_t *)devmem_reserve(vm, vb + 8);
            VX_GPU(vgpu_memcpy_h2d(d_val, a->validity, vb));
        }
        const uint8_t *chars = (const uint8_t *)a->buf.str.data;
        if (build) VX_GPU(vgpu_strdict_append(g->sdict[k], a->buf.str.offsets, chars, d_val, n, 0, codes, NULL));
        else VX_GPU(vgpu_strdict_lookup(g->sdict[k], a->buf.str.offsets, chars, d_val, n, codes));
    }
    const vgpu_col c = {VGPU_INT64, 0, codes, d_val};
    if (build) g->s_bcol[k] = c;
    return c;
}

/* Several key columns (src/join.c:67-113 hashes and compares the tuple):
 * the table is built on a 64-bit code of each build row's tuple, every
 * column coerced to the pair's common type, rows with an NA or NaN
 * component left out (they never match); probes use the probe tuples'
 * codes, and every matched pair's tuples are compared afterwards, so a code
 * shared by two different tuples (about 2^-64 per pair) is an error, never
 * a wrong match.  A string column (one key or several) enters the codes as
 * its exact dictionary code, so equal codes are equal bytes
 * (src/join.c:96-108): one string key cannot collide at all (a tuple code
 * of one column is a bijection of its value), and the pair check compares
 * the codes on the device like numbers. */
static void multi_build_side(JoinNode *jn) {
    JoinGpu *g = (JoinGpu *)jn->gpu;
    const int m = jn->n_keys;
    const VecSchema *rs = &jn->right->output_schema;
    const VecSchema *ls = &jn->left->output_schema;
    if (m > JN_MAX_KEYS) vectra_error("GPU join: %d key columns exceed %d", m, JN_MAX_KEYS);
    g->multi = 1;
    g->r_ncols = rs->n_cols;
    g->n_build = vx_concat_batches(jn->right, &g->r_host);
    g->first_only = 0;
    for (int k = 0; k < m; k++) {
        const VecType lt = ls->col_types[jn->keys[k].left_col], rt = rs->col_types[jn->keys[k].right_col];
        if (lt == VEC_STRING || rt == VEC_STRING) {
            if (lt != rt) vectra_error("GPU join: string key joined to a %s key", vx_type_name(lt == VEC_STRING ? rt : lt));
            g->common[k] = VEC_STRING;
            continue;
        }
        g->common[k] = vx_common_type(lt, rt);
        /* A left / full join continues its chain walk comparing every probe
         * key in its original type with the coerced build key's bytes
         * (src/join.c:398-400, 413-415): with several keys which later
         * matches survive depends on every column's bytes, which this path
         * does not reproduce, so it does not take the join. */
        if ((jn->kind == JOIN_LEFT || jn->kind == JOIN_FULL) && lt != g->common[k])
            vectra_error("GPU join: left / full join on %d key columns with the probe key '%s' coerced to %s is "
                         "not evaluated on the GPU", m, ls->col_names[jn->keys[k].left_col],
                         vx_type_name(g->common[k]));
    }
    g->r_dev = (DevCol *)calloc((size_t)(g->r_ncols > 0 ? g->r_ncols : 1), sizeof(DevCol));
    for (int c = 0; c < g->r_ncols; c++) {
        if (g->r_host[c].type == VEC_STRING) continue;
        devcol_upload(&g->r_dev[c], &g->r_host[c], g->n_build);
    }
    vgpu_col keys[JN_MAX_KEYS];
    const size_t nb = (size_t)(g->n_build > 0 ? g->n_build : 1);
    for (int k = 0; k < m; k++) {
        const int rk = jn->keys[k].right_col;
        if (g->common[k] == VEC_STRING) {
            keys[k] = string_key_col(g, k, &g->r_host[rk], g->n_build, 1);
            continue;
        }
        DevCol *dc = &g->r_dev[rk];
        if ((VecType)dc->col.type != g->common[k]) {
            void *dst = vgpu_malloc(nb * 8);
            if (!dst) vectra_error("GPU: %s", vgpu_last_error());
            VX_GPU(vgpu_cast(&dc->col, g->n_build, (int)g->common[k], dst));
            vgpu_free(dc->d_data);
            dc->d_data = dst;
            dc->cap_data = nb * 8;
            dc->col.data = dst;
            dc->col.type = (int32_t)g->common[k];
        }
        keys[k] = dc->col;
    }
    int64_t *codes = (int64_t *)devmem_reserve(&g->bcode, nb * 8);
    uint8_t *valid = (uint8_t *)devmem_reserve(&g->bvalid, (nb + 63) / 64 * 8);
    VX_GPU(vgpu_key_codes(keys, m, g->n_build, 0, codes, valid));
    vgpu_col bkey = {VGPU_INT64, 0, codes, valid};
    g->jt = vgpu_join_build(&bkey, g->n_build);
    if (!g->jt) vectra_error("GPU: %s", vgpu_last_error());
}

static void join_build_side(JoinNode *jn) {
    JoinGpu *g = (JoinGpu *)jn->gpu;
    if (jn->n_keys > 1 || jn->left->output_schema.col_types[jn->keys[0].left_col] == VEC_STRING ||
        jn->right->output_schema.col_types[jn->keys[0].right_col] == VEC_STRING) {
        multi_build_side(jn);
        return;
    }
    const VecSchema *rs = &jn->right->output_schema;
    const VecSchema *ls = &jn->left->output_schema;
    g->r_ncols = rs->n_cols;
    g->n_build = vx_concat_batches(jn->right, &g->r_host);
    int rk = jn->keys[0].right_col;
    VecType lt = ls->col_types[jn->keys[0].left_col];
    VecType rt = rs->col_types[rk];
    if (lt == VEC_STRING || rt == VEC_STRING)
        vectra_error("GPU join: string keys are not joined on the GPU");
    VecType common = vx_common_type(lt, rt);
    g->build_key_type = common;
    /* Left/full joins continue a chain walk by comparing the probe key in
     * its ORIGINAL type against the build key's bytes (src/join.c:398-400,
     * 413-415); when the probe key had to be coerced the probe passes the
     * uncoerced column so that the device applies the same rule. */
    g->first_only = (jn->kind == JOIN_LEFT || jn->kind == JOIN_FULL) && lt != common;
    g->r_dev = (DevCol *)calloc((size_t)(g->r_ncols > 0 ? g->r_ncols : 1), sizeof(DevCol));
    for (int c = 0; c < g->r_ncols; c++) {
        if (g->r_host[c].type == VEC_STRING) continue;   /* gathered on the host */
        devcol_upload(&g->r_dev[c], &g->r_host[c], g->n_build);
    }
    vgpu_col bkey = g->r_dev[rk].col;
    if (rt != common) {
        /* coerce the build key on the device (src/join.c:190-205) */
        DevCol *dc = &g->r_dev[rk];
        void *dst = vgpu_malloc((size_t)(g->n_build > 0 ? g->n_build : 1) * 8);
        if (!dst) vectra_error("GPU: %s", vgpu_last_error());
        VX_GPU(vgpu_cast(&dc->col, g->n_build, (int)common, dst));
        vgpu_free(dc->d_data);
        dc->d_data = dst;
        dc->cap_data = (size_t)(g->n_build > 0 ? g->n_build : 1) * 8;
        dc->col.data = dst;
        dc->col.type = (int32_t)common;
        bkey = dc->col;
    }
    if (g->n_pay > 0) {
        vgpu_col pay[4];
        for (int i = 0; i < g->n_pay; i++) pay[i] = g->r_dev[g->pay_cols[i]].col;
        g->jt = vgpu_join_build_payload(&bkey, g->n_build, pay, g->n_pay);
    } else {
        g->jt = vgpu_join_build(&bkey, g->n_build);
    }
    if (!g->jt) vectra_error("GPU: %s", vgpu_last_error());
}

/* copies one gathered device column back into a fresh VecArray */
static VecArray fetch_gathered(JoinGpu *g, VecType t, int64_t n) {
    VecArray a = vec_array_alloc(t, n);
    if (n > 0) {
        size_t w = t == VEC_BOOL ? 1 : 8;
        VX_GPU(vgpu_memcpy_d2h(a.buf.i64, g->out_data.ptr, (size_t)n * w));
        VX_GPU(vgpu_memcpy_d2h(a.validity, g->out_valid.ptr, (size_t)vec_validity_bytes(n)));
    }
    return a;
}

static VecArray gather_device(JoinGpu *g, const vgpu_col *src, VecType t, const void *idx, int idx64, int64_t n) {
    devmem_reserve(&g->out_data, (size_t)(n > 0 ? n : 1) * 8);
    devmem_reserve(&g->out_valid, (size_t)((n + 63) / 64 + 1) * 8);
    if (n > 0) VX_GPU(vgpu_gather(src, idx, idx64, n, g->out_data.ptr, (uint8_t *)g->out_valid.ptr));
    return fetch_gathered(g, t, n);
}

/* host gather of a string column through a row list (row < 0 -> NA) */
static VecArray gather_strings(const VecArray *src, const int64_t *rows, int64_t n) {
    VecArray d = vec_array_alloc(VEC_STRING, n);
    int64_t total = 0;
    for (int64_t i = 0; i < n; i++) {
        int64_t r = rows[i];
        if (r >= 0 && vec_array_is_valid(src, r)) total += src->buf.str.offsets[r + 1] - src->buf.str.offsets[r];
    }
    d.buf.str.data = (char *)malloc((size_t)(total > 0 ? total : 1));
    d.buf.str.data_len = total;
    int64_t off = 0;
    for (int64_t i = 0; i < n; i++) {
        int64_t r = rows[i];
        d.buf.str.offsets[i] = off;
        if (r >= 0 && vec_array_is_valid(src, r)) {
            vec_array_set_valid(&d, i);
            int64_t s = src->buf.str.offsets[r], l = src->buf.str.offsets[r + 1] - s;
            memcpy(d.buf.str.data + off, src->buf.str.data + s, (size_t)l);
            off += l;
        }
    }
    d.buf.str.offsets[n] = off;
    return d;
}

static VecBatch *make_output(JoinNode *jn, int64_t n) {
    VecBatch *res = vec_batch_alloc(jn->base.output_schema.n_cols, n);
    for (int c = 0; c < res->n_cols; c++) res->col_names[c] = vx_strdup(jn->base.output_schema.col_names[c]);
    return res;
}

static VecBatch *probe_batch_multi(JoinNode *jn, VecBatch *pb);

/* every matched pair (build row >= 0) joins equal tuples, compared on the
 * device (pkeys: the probe's key columns in the common types, string keys
 * as dictionary codes) */
static void verify_pairs(JoinNode *jn, const vgpu_col *pkeys, const int32_t *d_prow, const int64_t *d_brow,
                         int64_t n) {
    JoinGpu *g = (JoinGpu *)jn->gpu;
    vgpu_col bkeys[JN_MAX_KEYS];
    for (int k = 0; k < jn->n_keys; k++)
        bkeys[k] = g->common[k] == VEC_STRING ? g->s_bcol[k] : g->r_dev[jn->keys[k].right_col].col;
    int64_t bad = 0;
    VX_GPU(vgpu_pairs_unequal(pkeys, bkeys, jn->n_keys, d_prow, d_brow, n, &bad));
    if (bad) vectra_error("GPU join: two different key tuples share a 64-bit code (probability about 2^-64 per pair)");
}

static VecBatch *probe_batch(JoinNode *jn, VecBatch *pb) {
    JoinGpu *g = (JoinGpu *)jn->gpu;
    int lnc = pb->n_cols;
    if (lnc > 64) vectra_error("GPU join: more than 64 left columns");
    int lk = jn->keys[0].left_col;
    int64_t nl = vec_batch_logical_rows(pb);
    if (nl == 0) return NULL;   /* empty probe batch: no output (src/join.c:451-459) */
    if (g->multi) return probe_batch_multi(jn, pb);
    /* probe key, coerced to the build key type (src/join.c:268-289) */
    devcol_upload(&g->l_dev[lk], &pb->columns[lk], pb->n_rows);
    vgpu_col pkey = g->l_dev[lk].col;
    if ((VecType)pkey.type != g->build_key_type) {
        devcol_reserve(&g->pkey_cast, (size_t)(pb->n_rows > 0 ? pb->n_rows : 1) * 8, 8);
        VX_GPU(vgpu_cast(&pkey, pb->n_rows, (int)g->build_key_type, g->pkey_cast.d_data));
        pkey.data = g->pkey_cast.d_data;
        pkey.type = (int32_t)g->build_key_type;
    }
    const int32_t *d_sel = NULL;
    if (pb->sel) {
        d_sel = (const int32_t *)devmem_reserve(&g->sel, (size_t)(pb->sel_n > 0 ? pb->sel_n : 1) * 4);
        VX_GPU(vgpu_memcpy_h2d((void *)d_sel, pb->sel, (size_t)pb->sel_n * 4));
    }
    int64_t n_out = 0;
    VX_GPU(vgpu_join_probe(g->jt, (int)jn->kind, &pkey, pb->n_rows, d_sel, pb->sel ? pb->sel_n : 0,
                           g->first_only ? &g->l_dev[lk].col : NULL, &n_out));
    if (n_out == 0) return NULL;
    const int32_t *d_prow;
    const int64_t *d_brow;
    vgpu_join_pairs(g->jt, &d_prow, &d_brow);
    /* host copies of the pairs, needed only for string columns */
    int64_t *h_p = NULL, *h_b = NULL;
    int need_host = 0;
    for (int c = 0; c < lnc; c++) need_host |= pb->columns[c].type == VEC_STRING;
    if (jn->kind != JOIN_SEMI && jn->kind != JOIN_ANTI)
        for (int j = 0; j < g->n_r_nonkey; j++) need_host |= g->r_host[g->r_nonkey[j]].type == VEC_STRING;
    if (need_host) {
        int32_t *tmp = (int32_t *)malloc((size_t)n_out * 4);
        h_p = (int64_t *)malloc((size_t)n_out * 8);
        h_b = (int64_t *)malloc((size_t)n_out * 8);
        VX_GPU(vgpu_memcpy_d2h(tmp, d_prow, (size_t)n_out * 4));
        VX_GPU(vgpu_memcpy_d2h(h_b, d_brow, (size_t)n_out * 8));
        for (int64_t i = 0; i < n_out; i++) h_p[i] = tmp[i];
        free(tmp);
    }
    VecBatch *res = make_output(jn, n_out);
    for (int c = 0; c < lnc; c++) {
        const VecArray *src = &pb->columns[c];
        if (src->type == VEC_STRING) {
            res->columns[c] = gather_strings(src, h_p, n_out);
            continue;
        }
        if (c != lk) devcol_upload(&g->l_dev[c], src, pb->n_rows);
        vgpu_col col = g->l_dev[c].col;
        if (c == lk) col = g->l_dev[lk].col;   /* original (uncoerced) key values */
        res->columns[c] = gather_device(g, &col, src->type, d_prow, 0, n_out);
    }
    if (jn->kind != JOIN_SEMI && jn->kind != JOIN_ANTI) {
        for (int j = 0; j < g->n_r_nonkey; j++) {
            int rc = g->r_nonkey[j];
            if (g->r_host[rc].type == VEC_STRING) {
                res->columns[lnc + j] = gather_strings(&g->r_host[rc], h_b, n_out);
                continue;
            }
            res->columns[lnc + j] = gather_device(g, &g->r_dev[rc].col, g->r_host[rc].type, d_brow, 1, n_out);
        }
    }
    free(h_p);
    free(h_b);
    return res;
}

/* gathers the left columns through d_prow and, unless semi / anti, the
 * right non-key columns through d_brow (n output rows) */
static VecBatch *gather_output(JoinNode *jn, VecBatch *pb, const int32_t *d_prow, const int64_t *d_brow, int64_t n) {
    JoinGpu *g = (JoinGpu *)jn->gpu;
    const int lnc = pb->n_cols;
    int64_t *h_p = NULL, *h_b = NULL;
    int need_host = 0;
    for (int c = 0; c < lnc; c++) need_host |= pb->columns[c].type == VEC_STRING;
    if (jn->kind != JOIN_SEMI && jn->kind != JOIN_ANTI)
        for (int j = 0; j < g->n_r_nonkey; j++) need_host |= g->r_host[g->r_nonkey[j]].type == VEC_STRING;
    if (need_host) {
        int32_t *tmp = (int32_t *)malloc((size_t)n * 4);
        h_p = (int64_t *)malloc((size_t)n * 8);
        h_b = (int64_t *)malloc((size_t)n * 8);
        VX_GPU(vgpu_memcpy_d2h(tmp, d_prow, (size_t)n * 4));
        if (d_brow) VX_GPU(vgpu_memcpy_d2h(h_b, d_brow, (size_t)n * 8));
        for (int64_t i = 0; i < n; i++) h_p[i] = tmp[i];
        free(tmp);
    }
    VecBatch *res = make_output(jn, n);
    for (int c = 0; c < lnc; c++) {
        const VecArray *src = &pb->columns[c];
        if (src->type == VEC_STRING) {
            res->columns[c] = gather_strings(src, h_p, n);
            continue;
        }
        devcol_upload(&g->l_dev[c], src, pb->n_rows);
        res->columns[c] = gather_device(g, &g->l_dev[c].col, src->type, d_prow, 0, n);
    }
    if (jn->kind != JOIN_SEMI && jn->kind != JOIN_ANTI) {
        for (int j = 0; j < g->n_r_nonkey; j++) {
            const int rc = g->r_nonkey[j];
            if (g->r_host[rc].type == VEC_STRING) {
                res->columns[lnc + j] = gather_strings(&g->r_host[rc], h_b, n);
                continue;
            }
            res->columns[lnc + j] = gather_device(g, &g->r_dev[rc].col, g->r_host[rc].type, d_brow, 1, n);
        }
    }
    free(h_p);
    free(h_b);
    return res;
}

/* Several key columns: the probe tuples' codes against the build codes.
 * Semi and anti joins are derived from an inner probe, whose pairs can be
 * checked: semi keeps each probe row with a pair once, anti the logical
 * rows without one, both in probe order (src/join.c:371-384). */
static VecBatch *probe_batch_multi(JoinNode *jn, VecBatch *pb) {
    JoinGpu *g = (JoinGpu *)jn->gpu;
    const int m = jn->n_keys;
    vgpu_col keys[JN_MAX_KEYS];
    for (int k = 0; k < m; k++) {
        const int lk = jn->keys[k].left_col;
        if (g->common[k] == VEC_STRING) {
            keys[k] = string_key_col(g, k, &pb->columns[lk], pb->n_rows, 0);
            continue;
        }
        devcol_upload(&g->l_dev[lk], &pb->columns[lk], pb->n_rows);
        keys[k] = g->l_dev[lk].col;
        if ((VecType)keys[k].type != g->common[k]) {
            devcol_reserve(&g->pcast[k], (size_t)(pb->n_rows > 0 ? pb->n_rows : 1) * 8, 8);
            VX_GPU(vgpu_cast(&keys[k], pb->n_rows, (int)g->common[k], g->pcast[k].d_data));
            keys[k].data = g->pcast[k].d_data;
            keys[k].type = (int32_t)g->common[k];
        }
    }
    const size_t np = (size_t)(pb->n_rows > 0 ? pb->n_rows : 1);
    int64_t *codes = (int64_t *)devmem_reserve(&g->pcode, np * 8);
    uint8_t *valid = (uint8_t *)devmem_reserve(&g->pvalid, (np + 63) / 64 * 8);
    VX_GPU(vgpu_key_codes(keys, m, pb->n_rows, 0, codes, valid));
    vgpu_col pkey = {VGPU_INT64, 0, codes, valid};
    const int32_t *d_sel = NULL;
    if (pb->sel) {
        d_sel = (const int32_t *)devmem_reserve(&g->sel, (size_t)(pb->sel_n > 0 ? pb->sel_n : 1) * 4);
        VX_GPU(vgpu_memcpy_h2d((void *)d_sel, pb->sel, (size_t)pb->sel_n * 4));
    }
    const int semi_anti = jn->kind == JOIN_SEMI || jn->kind == JOIN_ANTI;
    int64_t n_out = 0;
    VX_GPU(vgpu_join_probe(g->jt, semi_anti ? (int)JOIN_INNER : (int)jn->kind, &pkey, pb->n_rows, d_sel,
                           pb->sel ? pb->sel_n : 0, NULL, &n_out));
    const int32_t *d_prow = NULL;
    const int64_t *d_brow = NULL;
    if (n_out > 0) {
        vgpu_join_pairs(g->jt, &d_prow, &d_brow);
        verify_pairs(jn, keys, d_prow, d_brow, n_out);
    }
    if (!semi_anti) return n_out > 0 ? gather_output(jn, pb, d_prow, d_brow, n_out) : NULL;
    /* probe rows with a pair, in probe order */
    const int64_t nl = vec_batch_logical_rows(pb);
    int32_t *pr = (int32_t *)malloc((size_t)(n_out > 0 ? n_out : 1) * 4);
    int32_t *rows = (int32_t *)malloc((size_t)(nl > 0 ? nl : 1) * 4);
    uint8_t *hit = (uint8_t *)calloc((size_t)(pb->n_rows > 0 ? pb->n_rows : 1), 1);
    if (!pr || !rows || !hit) vectra_error("alloc failed");
    if (n_out > 0) VX_GPU(vgpu_memcpy_d2h(pr, d_prow, (size_t)n_out * 4));
    for (int64_t i = 0; i < n_out; i++) hit[pr[i]] = 1;
    int64_t nr = 0;
    for (int64_t l = 0; l < nl; l++) {
        const int32_t r = pb->sel ? pb->sel[l] : (int32_t)l;
        if (hit[r] == (jn->kind == JOIN_SEMI)) rows[nr++] = r;
    }
    free(pr);
    free(hit);
    VecBatch *res = NULL;
    if (nr > 0) {
        int32_t *d_rows = (int32_t *)devmem_reserve(&g->sel, (size_t)nr * 4);
        VX_GPU(vgpu_memcpy_h2d(d_rows, rows, (size_t)nr * 4));
        res = gather_output(jn, pb, d_rows, NULL, nr);
    }
    free(rows);
    return res;
}

/* FULL: one batch of never-matched build rows (src/join.c:487-551) */
static VecBatch *finalize_batch(JoinNode *jn) {
    JoinGpu *g = (JoinGpu *)jn->gpu;
    if (!g->h_build) {
        int64_t n = 0;
        VX_GPU(vgpu_join_unmatched_build(g->jt, &n));
        g->n_unmatched = n;
        g->h_build = (int64_t *)malloc((size_t)(n > 0 ? n : 1) * 8);
        if (n > 0) {
            const int64_t *d_b;
            vgpu_join_pairs(g->jt, NULL, &d_b);
            VX_GPU(vgpu_memcpy_d2h(g->h_build, d_b, (size_t)n * 8));
        }
        g->fin_cursor = 0;
    }
    int64_t n = g->n_unmatched - g->fin_cursor;
    if (n <= 0) return NULL;
    if (n > FINALIZE_ROWS) n = FINALIZE_ROWS;
    const int64_t *rows = g->h_build + g->fin_cursor;
    g->fin_cursor += n;
    DevMem idx = {0, 0};
    devmem_reserve(&idx, (size_t)n * 8);
    VX_GPU(vgpu_memcpy_h2d(idx.ptr, rows, (size_t)n * 8));
    const VecSchema *ls = &jn->left->output_schema;
    int lnc = ls->n_cols;
    VecBatch *res = make_output(jn, n);
    for (int c = 0; c < lnc; c++) {
        int rsrc = -1;
        for (int k = 0; k < jn->n_keys; k++)
            if (c == jn->keys[k].left_col) rsrc = jn->keys[k].right_col;
        VecType t = ls->col_types[c];
        if (rsrc < 0) {
            res->columns[c] = vec_array_alloc(t, n);   /* all NA */
        } else if (t == VEC_STRING || g->r_host[rsrc].type == VEC_STRING) {
            res->columns[c] = gather_strings(&g->r_host[rsrc], rows, n);
        } else {
            /* the left-typed output takes the build key's bytes as they are
             * (src/builder.c:110-138 writes through the builder's type) */
            vgpu_col col = g->r_dev[rsrc].col;
            if (t == VEC_BOOL && col.type != VGPU_BOOL) col.type = VGPU_BOOL;
            else if (t != VEC_BOOL && col.type == VGPU_BOOL) col.type = VGPU_BOOL;
            res->columns[c] = gather_device(g, &col, col.type == VGPU_BOOL ? VEC_BOOL : t, idx.ptr, 1, n);
            res->columns[c].type = t;
        }
    }
    for (int j = 0; j < g->n_r_nonkey; j++) {
        int rc = g->r_nonkey[j];
        if (g->r_host[rc].type == VEC_STRING)
            res->columns[lnc + j] = gather_strings(&g->r_host[rc], rows, n);
        else
            res->columns[lnc + j] = gather_device(g, &g->r_dev[rc].col, g->r_host[rc].type, idx.ptr, 1, n);
    }
    devmem_free(&idx);
    return res;
}

static VecBatch *gpu_join_next(VecNode *self) {
    JoinNode *jn = (JoinNode *)self;
    JoinGpu *g = (JoinGpu *)jn->gpu;
    if (!g->built) {
        join_build_side(jn);
        g->built = 1;
    }
    while (g->state == 0) {
        VecBatch *pb = jn->left->next_batch(jn->left);
        if (!pb) {
            g->state = jn->kind == JOIN_FULL ? 1 : 2;
            break;
        }
        VecBatch *res = probe_batch(jn, pb);
        vec_batch_free(pb);
        if (res) return res;
    }
    if (g->state == 1) {
        VecBatch *res = finalize_batch(jn);
        if (res) return res;
        g->state = 2;
    }
    return NULL;
}

static void gpu_join_free(VecNode *self) {
    JoinNode *jn = (JoinNode *)self;
    JoinGpu *g = (JoinGpu *)jn->gpu;
    jn->left->free_node(jn->left);
    jn->right->free_node(jn->right);
    if (g) {
        if (g->r_host) {
            for (int c = 0; c < g->r_ncols; c++) vec_array_free(&g->r_host[c]);
            free(g->r_host);
        }
        if (g->r_dev) {
            for (int c = 0; c < g->r_ncols; c++) devcol_free(&g->r_dev[c]);
            free(g->r_dev);
        }
        for (int c = 0; c < 64; c++) devcol_free(&g->l_dev[c]);
        devcol_free(&g->pkey_cast);
        for (int k = 0; k < JN_MAX_KEYS; k++) {
            devcol_free(&g->pcast[k]);
            devmem_free(&g->s_bcode[k]);
            devmem_free(&g->s_bval[k]);
            devmem_free(&g->s_pcode[k]);
            devmem_free(&g->s_pval[k]);
            vgpu_strdict_free(g->sdict[k]);
            g->sdict[k] = NULL;
        }
        devmem_free(&g->bcode);
        devmem_free(&g->bvalid);
        devmem_free(&g->pcode);
        devmem_free(&g->pvalid);
        devmem_free(&g->sel);
        devmem_free(&g->out_data);
        devmem_free(&g->out_valid);
        if (g->jt) vgpu_join_destroy(g->jt);
        free(g->r_nonkey);
        free(g->h_build);
        free(g);
    }
    free(jn->keys);
    free(jn->suffix_x);
    free(jn->suffix_y);
    vec_schema_free(&jn->base.output_schema);
    free(jn);
}

JoinNode *join_node_create(VecNode *left, VecNode *right, JoinKind kind, int n_keys, JoinKey *keys,
                           const char *suffix_x, const char *suffix_y) {
    static const char *kind_names[] = {"inner_join", "left_join", "full_join", "semi_join", "anti_join"};
    vx_gpu_init();
    const VecSchema *ls = &left->output_schema, *rs = &right->output_schema;
    for (int k = 0; k < n_keys; k++) {
        VecType lt = ls->col_types[keys[k].left_col], rt = rs->col_types[keys[k].right_col];
        if (lt != rt && (lt == VEC_STRING || rt == VEC_STRING))
            vectra_error("%s key type mismatch: x.%s (%s) vs y.%s (%s)", kind_names[kind],
                         ls->col_names[keys[k].left_col], vx_type_name(lt), rs->col_names[keys[k].right_col],
                         vx_type_name(rt));
    }
    JoinNode *jn = (JoinNode *)calloc(1, sizeof *jn);
    JoinGpu *g = (JoinGpu *)calloc(1, sizeof *g);
    if (!jn || !g) vectra_error("alloc failed for JoinNode");
    jn->gpu = g;
    jn->left = left;
    jn->right = right;
    jn->kind = kind;
    jn->n_keys = n_keys;
    jn->keys = keys;
    jn->suffix_x = vx_strdup(suffix_x);
    jn->suffix_y = vx_strdup(suffix_y);
    int *r_is_key = (int *)calloc((size_t)(rs->n_cols + 1), sizeof(int));
    for (int k = 0; k < n_keys; k++) r_is_key[keys[k].right_col] = 1;
    g->r_nonkey = (int *)malloc(sizeof(int) * (size_t)(rs->n_cols + 1));
    for (int c = 0; c < rs->n_cols; c++)
        if (!r_is_key[c]) g->r_nonkey[g->n_r_nonkey++] = c;
    int semi = kind == JOIN_SEMI || kind == JOIN_ANTI;
    int out_n = semi ? ls->n_cols : ls->n_cols + g->n_r_nonkey;
    char **names = (char **)calloc((size_t)(out_n + 1), sizeof(char *));
    VecType *types = (VecType *)malloc(sizeof(VecType) * (size_t)(out_n + 1));
    for (int i = 0; i < ls->n_cols; i++) {
        names[i] = vx_strdup(ls->col_names[i]);
        types[i] = ls->col_types[i];
    }
    if (!semi) {
        for (int j = 0; j < g->n_r_nonkey; j++) {
            int c = g->r_nonkey[j];
            const char *rn = rs->col_names[c];
            int hit = -1;
            for (int i = 0; i < ls->n_cols; i++)
                if (strcmp(ls->col_names[i], rn) == 0) { hit = i; break; }
            if (hit >= 0) {
                size_t L = strlen(rn) + strlen(suffix_y) + 1;
                names[ls->n_cols + j] = (char *)malloc(L);
                snprintf(names[ls->n_cols + j], L, "%s%s", rn, suffix_y);
                size_t L2 = strlen(ls->col_names[hit]) + strlen(suffix_x) + 1;
                free(names[hit]);
                names[hit] = (char *)malloc(L2);
                snprintf(names[hit], L2, "%s%s", ls->col_names[hit], suffix_x);
            } else {
                names[ls->n_cols + j] = vx_strdup(rn);
            }
            types[ls->n_cols + j] = rs->col_types[c];
        }
    }
    jn->base.output_schema = vec_schema_create(out_n, names, types);
    for (int i = 0; i < out_n; i++) free(names[i]);
    free(names);
    free(types);
    free(r_is_key);
    jn->base.next_batch = gpu_join_next;
    jn->base.free_node = gpu_join_free;
    jn->base.kind = "JoinNode";
    return jn;
}

int vx_gpu_join_note(const VecNode *n, char *buf, int size) {
    if (n->next_batch != gpu_join_next) return -1;
    const JoinNode *jn = (const JoinNode *)n;
    int strs = 0;
    for (int k = 0; k < jn->n_keys; k++)
        strs += jn->left->output_schema.col_types[jn->keys[k].left_col] == VEC_STRING;
    const int tuple = jn->n_keys > 1 || strs > 0;
    return snprintf(buf, (size_t)size, "gpu: %s table built on the device%s, probe per batch",
                    tuple ? "tuple-code" : "hash", strs ? " (string keys by exact device dictionary codes)" : "");
}

/* ---- fusion with a GroupAggNode above a ProjectNode ------------------ */

int vx_gpu_join_fusable(const VecNode *n) {
    if (!n || n->next_batch != gpu_join_next) return 0;
    const JoinNode *jn = (const JoinNode *)n;
    const JoinGpu *g = (const JoinGpu *)jn->gpu;
    if (g->built || g->state != 0) return 0;
    if (jn->kind != JOIN_INNER && jn->kind != JOIN_LEFT) return 0;
    if (jn->n_keys != 1) return 0;
    VecType lt = jn->left->output_schema.col_types[jn->keys[0].left_col];
    VecType rt = jn->right->output_schema.col_types[jn->keys[0].right_col];
    return lt != VEC_STRING && rt != VEC_STRING;
}

int vx_gpu_join_output_source(const VecNode *n, int out_col, int *side, int *col) {
    const JoinNode *jn = (const JoinNode *)n;
    const JoinGpu *g = (const JoinGpu *)jn->gpu;
    const int nl = jn->left->output_schema.n_cols;
    if (out_col < 0) return -1;
    if (out_col < nl) {
        *side = 0;
        *col = out_col;
        return 0;
    }
    if (out_col - nl >= g->n_r_nonkey || jn->kind == JOIN_SEMI || jn->kind == JOIN_ANTI) return -1;
    *side = 1;
    *col = g->r_nonkey[out_col - nl];
    return 0;
}

void vx_gpu_join_prepare_fusion(VecNode *n, const int *payload_cols, int n_payload, VxJoinFusion *out) {
    JoinNode *jn = (JoinNode *)n;
    JoinGpu *g = (JoinGpu *)jn->gpu;
    g->n_pay = 0;
    for (int i = 0; i < n_payload && g->n_pay < 4; i++) {
        VecType t = jn->right->output_schema.col_types[payload_cols[i]];
        if (t != VEC_STRING && payload_cols[i] != jn->keys[0].right_col) g->pay_cols[g->n_pay++] = payload_cols[i];
    }
    join_build_side(jn);
    g->built = 1;
    memset(out, 0, sizeof *out);
    out->jt = g->jt;
    out->left = jn->left;
    out->kind = jn->kind;
    out->left_key = jn->keys[0].left_col;
    out->right_key = jn->keys[0].right_col;
    out->build_key_type = g->build_key_type;
    out->r_dev = g->r_dev;
    out->r_ncols = g->r_ncols;
    out->max_dup = vgpu_join_max_dup(g->jt);
    out->coerced = jn->left->output_schema.col_types[jn->keys[0].left_col] != g->build_key_type;
}
