/*
 * serve.c -- the request loop of include/vx_serve.h: plain (one query at a
 * time) and continuous batching (shared scans of the fact columns), over
 * device-resident columns and a resident join table.
 */
#define _POSIX_C_SOURCE 199309L
#include "vx_serve.h"

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#define SV_MAX_BATCH 256

static double sv_now(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

static void sv_sleep_until(double t0, double at) {
    double dt = at - (sv_now() - t0);
    if (dt <= 0) return;
    struct timespec ts;
    ts.tv_sec = (time_t)dt;
    ts.tv_nsec = (long)((dt - (double)ts.tv_sec) * 1e9);
    nanosleep(&ts, NULL);
}

/* value > t over the column in slot `col` */
static vgpu_prog sv_pred(int col, double t) {
    vgpu_prog p;
    memset(&p, 0, sizeof p);
    p.n_insn = 3;
    p.insn[0].op = VGPU_OP_COL;
    p.insn[0].type = VGPU_DOUBLE;
    p.insn[0].col = col;
    p.insn[1].op = VGPU_OP_CONST;
    p.insn[1].type = VGPU_DOUBLE;
    p.insn[1].imm.d = t;
    p.insn[2].op = VGPU_OP_GT;
    p.insn[2].type = VGPU_DOUBLE;
    p.insn[2].a = 0;
    p.insn[2].b = 1;
    return p;
}

static const vgpu_agg SV_AGGS[2] = {{VGPU_AGG_SUM, 0, 0, 0}, {VGPU_AGG_COUNT_STAR, -1, 0, 0}};
static const int32_t SV_IN_TYPES[1] = {VGPU_DOUBLE};

/* host copy of one result (n groups) into r */
static int sv_keep(vx_result *r, int64_t n, int keep) {
    r->n_groups = n;
    if (!keep) return 0;
    size_t m = (size_t)(n > 0 ? n : 1), vb = (size_t)((n + 7) / 8 + 1);
    r->keys = (int64_t *)malloc(m * 8);
    r->key_valid = (uint8_t *)calloc(vb, 1);
    r->sum = (double *)malloc(m * 8);
    r->sum_valid = (uint8_t *)calloc(vb, 1);
    r->count = (double *)malloc(m * 8);
    return (r->keys && r->key_valid && r->sum && r->sum_valid && r->count) ? 0 : -1;
}

/* ---- plain: one group-by per request ---------------------------------- */

static int sv_plain_one(const vx_serve_cfg *c, vgpu_jtable *jt, const vx_request *q, vx_result *r) {
    vgpu_groupby *g = vgpu_groupby_create(VGPU_INT64, 1, SV_IN_TYPES, 2, SV_AGGS, VGPU_ORDER_KEY, 0);
    if (!g) return -1;
    int rc;
    if (q->cls == 0) {
        vgpu_prog p = sv_pred(0, q->threshold);
        rc = vgpu_groupby_feed(g, &c->fact_group, &c->fact_value, c->n_rows, NULL, 0, 0, &p, &c->fact_value, 1);
    } else {
        vgpu_prog p = sv_pred(1, q->threshold);   /* probe columns: key, value */
        vgpu_col probe[2] = {c->fact_key, c->fact_value};
        vgpu_col build[2] = {c->dim_key, c->dim_attr};
        vgpu_input in = {0, 0, 1, 0, 0, 0};
        vgpu_feed_spec sp;
        memset(&sp, 0, sizeof sp);
        sp.n_rows = c->n_rows;
        sp.pred = &p;
        sp.join = jt;
        sp.join_kind = VGPU_JOIN_INNER;
        sp.key_side = 1;
        sp.key_col = 1;
        sp.probe_key_col = 0;
        sp.n_probe_cols = 2;
        sp.n_build_cols = 2;
        sp.probe_cols = probe;
        sp.build_cols = build;
        sp.n_inputs = 1;
        sp.inputs = &in;
        rc = vgpu_groupby_feed_join(g, &sp);
    }
    int64_t n = rc < 0 ? -1 : vgpu_groupby_finish(g);
    if (n < 0 || sv_keep(r, n, c->keep_results)) {
        vgpu_groupby_destroy(g);
        return -1;
    }
    double *vals[2] = {NULL, NULL};
    uint8_t *valid[2] = {NULL, NULL};
    int64_t *keys = NULL;
    uint8_t *kv = NULL;
    if (r->keys) {
        keys = r->keys;
        kv = r->key_valid;
        vals[0] = r->sum;
        vals[1] = r->count;
        valid[0] = r->sum_valid;
    }
    rc = vgpu_groupby_fetch(g, keys, kv, vals, valid) < 0 ? -1 : 0;
    vgpu_groupby_destroy(g);
    return rc;
}

/* ---- batched: pass templates ------------------------------------------ */

/* a batch of max_batch queries over the fact columns, planned once; a pass
 * reruns its first k queries with the pass's thresholds */
static vgpu_batch *sv_template(const vx_serve_cfg *c, vgpu_jtable *jt, int cls, int kmax) {
    vgpu_prog p = sv_pred(0, 0.0);
    vgpu_batch *b = vgpu_batch_create(VGPU_INT64, 1, SV_IN_TYPES, 2, SV_AGGS, VGPU_ORDER_KEY, &p, 1, 1);
    if (!b) return NULL;
    if (cls == 1 && vgpu_batch_join(b, jt, 0) < 0) {
        vgpu_batch_destroy(b);
        return NULL;
    }
    vgpu_query *qs = (vgpu_query *)calloc((size_t)kmax, sizeof(vgpu_query));
    if (!qs) {
        vgpu_batch_destroy(b);
        return NULL;
    }
    double inf = INFINITY;
    for (int k = 0; k < kmax; k++) {
        if (cls == 0) qs[k].key = c->fact_group;
        else qs[k].probe_key = c->fact_key;
        qs[k].inputs = &c->fact_value;
        qs[k].pred_cols = &c->fact_value;
        qs[k].n_rows = c->n_rows;
        memcpy(&qs[k].param, &inf, 8);   /* the planning run selects nothing */
    }
    int rc = vgpu_batch_plan(b, qs, kmax);
    free(qs);
    if (rc < 0) {
        vgpu_batch_destroy(b);
        return NULL;
    }
    return b;
}

static int sv_pass(const vx_serve_cfg *c, vgpu_batch *b, const vx_request *reqs, const int *take, int k,
                   vx_result *out, double t0, int pass) {
    uint64_t params[SV_MAX_BATCH];
    for (int j = 0; j < k; j++) memcpy(&params[j], &reqs[take[j]].threshold, 8);
    double start = sv_now() - t0;
    if (vgpu_batch_rerun(b, params, k) < 0) return -1;
    int64_t counts[SV_MAX_BATCH];
    if (vgpu_batch_counts(b, counts, NULL) < 0) return -1;
    for (int j = 0; j < k; j++) {
        vx_result *r = &out[take[j]];
        if (sv_keep(r, counts[j], c->keep_results)) return -1;
        if (c->keep_results) {
            double *vals[2] = {r->sum, r->count};
            uint8_t *valid[2] = {r->sum_valid, NULL};
            if (vgpu_batch_fetch(b, j, r->keys, r->key_valid, vals, valid) < 0) return -1;
        }
    }
    double end = sv_now() - t0;
    for (int j = 0; j < k; j++) {
        out[take[j]].start_s = start;
        out[take[j]].end_s = end;
        out[take[j]].pass = pass;
        out[take[j]].batch = k;
    }
    return 0;
}

int vx_serve_run(const vx_serve_cfg *c, const vx_request *reqs, int n, vx_result *out, vx_serve_stats *st) {
    memset(st, 0, sizeof *st);
    memset(out, 0, (size_t)n * sizeof *out);
    int kmax = c->max_batch < 1 ? 1 : (c->max_batch > SV_MAX_BATCH ? SV_MAX_BATCH : c->max_batch);
    double s0 = sv_now();
    /* the dimension's join table stays resident for every class-1 request */
    vgpu_jtable *jt = vgpu_join_build_payload(&c->dim_key, c->n_dim, &c->dim_attr, 1);
    if (!jt) return -1;
    vgpu_batch *tpl[2] = {NULL, NULL};
    int rc = 0;
    if (c->batched) {
        for (int cls = 0; cls < 2 && rc == 0; cls++) {
            tpl[cls] = sv_template(c, jt, cls, kmax);
            if (!tpl[cls]) rc = -1;
        }
        if (rc == 0) {
            /* admission by the query-state budget: a request holds its share
             * of the larger pass template */
            int64_t per = 0;
            for (int cls = 0; cls < 2; cls++) {
                int64_t bytes = vgpu_batch_device_bytes(tpl[cls]);
                st->resident_bytes += bytes;
                if (bytes / kmax > per) per = bytes / kmax;
            }
            st->state_bytes = per;
            if (c->state_budget > 0 && per > 0 && c->state_budget / per < kmax)
                kmax = (int)(c->state_budget / per) > 0 ? (int)(c->state_budget / per) : 1;
        }
    }
    st->setup_s = sv_now() - s0;
    int *pending = (int *)malloc((size_t)(n > 0 ? n : 1) * sizeof(int));
    int *take = (int *)malloc((size_t)kmax * sizeof(int));
    if (!pending || !take) rc = -1;
    double t0 = sv_now();
    int next = 0, n_pend = 0, done = 0;
    while (rc == 0 && done < n) {
        double now = sv_now() - t0;
        while (next < n && reqs[next].arrival_s <= now) pending[n_pend++] = next++;
        if (n_pend == 0) {
            sv_sleep_until(t0, reqs[next].arrival_s);
            continue;
        }
        if (!c->batched) {
            int i = pending[0];
            memmove(pending, pending + 1, (size_t)(n_pend - 1) * sizeof(int));
            n_pend--;
            out[i].start_s = sv_now() - t0;
            if (sv_plain_one(c, jt, &reqs[i], &out[i]) < 0) {
                rc = -1;
                break;
            }
            out[i].end_s = sv_now() - t0;
            out[i].pass = -1;
            out[i].batch = 1;
            done++;
            continue;
        }
        /* the next pass: the oldest request's class, every waiting request
         * of that class in arrival order, up to kmax */
        int cls = reqs[pending[0]].cls, k = 0, w = 0;
        for (int j = 0; j < n_pend; j++) {
            int i = pending[j];
            if (reqs[i].cls == cls && k < kmax) take[k++] = i;
            else pending[w++] = i;
        }
        n_pend = w;
        if (sv_pass(c, tpl[cls], reqs, take, k, out, t0, st->n_passes) < 0) {
            rc = -1;
            break;
        }
        st->n_passes++;
        if (k > st->max_pass) st->max_pass = k;
        done += k;
    }
    st->n_done = done;
    st->wall_s = sv_now() - t0 - (n > 0 ? reqs[0].arrival_s : 0);
    free(pending);
    free(take);
    for (int cls = 0; cls < 2; cls++)
        if (tpl[cls]) vgpu_batch_destroy(tpl[cls]);
    vgpu_join_destroy(jt);
    return rc;
}

void vx_result_free(vx_result *r) {
    free(r->keys);
    free(r->key_valid);
    free(r->sum);
    free(r->sum_valid);
    free(r->count);
    r->keys = NULL;
    r->key_valid = NULL;
    r->sum = NULL;
    r->sum_valid = NULL;
    r->count = NULL;
}
