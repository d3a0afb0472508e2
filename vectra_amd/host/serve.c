/*
 * serve.c -- the request loop of include/vx_serve.h: plain (one query at a
 * time) and continuous batching (shared scans of the fact columns, several
 * passes in flight on streams of their own), over device-resident columns
 * and a resident join table.
 */
#define _POSIX_C_SOURCE 200809L
#include "vx_serve.h"

#include <pthread.h>

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

/* ---- batched: pass slots ---------------------------------------------- */

#define SV_MAX_SLOTS 8

struct SvSched;

/* one pass slot: a host thread driving its own stream, with a pass template
 * of each class planned on that stream */
typedef struct {
    struct SvSched *S;
    pthread_t th;
    int       started;
    void     *stream;
    vgpu_batch *tpl[2];
    int       kcap[2];     /* requests a pass of each class takes (its template's size) */
    int64_t   tpl_bytes;
    int       ready;       /* templates planned (1) or failed (-1) */
    int       state;       /* 0 idle, 1 pass assigned, 2 stop */
    int       cls, k, pass;
    int       take[SV_MAX_BATCH];
    char      err[512];
} SvSlot;

typedef struct SvSched {
    pthread_mutex_t mu;
    pthread_cond_t  cv_slot;   /* slots wait for work */
    pthread_cond_t  cv_main;   /* the scheduler waits for a slot or an arrival */
    const vx_serve_cfg *c;
    vgpu_jtable *jt;
    const vx_request *reqs;
    vx_result *out;
    int       kmax;
    int       kcap[2];     /* requests per pass of each class (at most kmax) */
    double    t0;
    int       done, in_flight, failed;
    char      err[512];
} SvSched;

static void *sv_slot_main(void *arg) {
    SvSlot *w = (SvSlot *)arg;
    SvSched *S = w->S;
    int ok = vgpu_stream_use(w->stream) == 0;
    for (int cls = 0; cls < 2 && ok; cls++) {
        w->tpl[cls] = sv_template(S->c, S->jt, cls, S->kmax);
        ok = w->tpl[cls] != NULL;
        w->kcap[cls] = S->kmax;
        /* A template too large for the bucketed scan (its per-block table
         * of groups x requests outgrows LDS) would run every request of a
         * pass as a scan of its own: take the largest pass that is still
         * bucketed instead (the table grows with the requests, so a binary
         * search finds it); none: keep kmax. */
        if (ok && S->kmax > 1 && !vgpu_batch_bucketed(w->tpl[cls])) {
            int lo = 0, hi = S->kmax - 1;   /* largest bucketed size in [1, kmax) is in (lo, hi] or none */
            vgpu_batch *best = NULL;
            while (lo < hi && ok) {
                const int mid = (lo + hi + 1) / 2;
                vgpu_batch *t = sv_template(S->c, S->jt, cls, mid);
                ok = t != NULL;
                if (!ok) break;
                if (vgpu_batch_bucketed(t)) {
                    if (best) vgpu_batch_destroy(best);
                    best = t;
                    lo = mid;
                } else {
                    vgpu_batch_destroy(t);
                    hi = mid - 1;
                }
            }
            if (!ok) {
                if (best) vgpu_batch_destroy(best);
            } else if (best) {
                vgpu_batch_destroy(w->tpl[cls]);
                w->tpl[cls] = best;
                w->kcap[cls] = lo;
            }
        }
    }
    if (ok) {
        w->tpl_bytes = vgpu_batch_device_bytes(w->tpl[0]) + vgpu_batch_device_bytes(w->tpl[1]);
        ok = vgpu_synchronize() == 0;
    }
    pthread_mutex_lock(&S->mu);
    if (!ok) snprintf(w->err, sizeof w->err, "%s", vgpu_last_error());
    w->ready = ok ? 1 : -1;
    pthread_cond_broadcast(&S->cv_main);
    for (;;) {
        while (w->state == 0) pthread_cond_wait(&S->cv_slot, &S->mu);
        if (w->state == 2) break;
        int take[SV_MAX_BATCH];
        const int cls = w->cls, k = w->k, pass = w->pass;
        memcpy(take, w->take, (size_t)k * sizeof(int));
        pthread_mutex_unlock(&S->mu);
        /* the pass runs on this slot's stream while other slots' passes run */
        int rc = sv_pass(S->c, w->tpl[cls], S->reqs, take, k, S->out, S->t0, pass);
        pthread_mutex_lock(&S->mu);
        if (rc < 0 && !S->failed) {
            S->failed = 1;
            snprintf(S->err, sizeof S->err, "%s", vgpu_last_error());
        }
        S->done += k;
        S->in_flight--;
        w->state = 0;
        pthread_cond_broadcast(&S->cv_main);
    }
    pthread_mutex_unlock(&S->mu);
    for (int cls = 0; cls < 2; cls++)
        if (w->tpl[cls]) vgpu_batch_destroy(w->tpl[cls]);
    vgpu_stream_use(NULL);
    return NULL;
}

/* starts slot w and waits until its templates are planned; 0 or -1 */
static int sv_slot_open(SvSched *S, SvSlot *w) {
    memset(w, 0, sizeof *w);
    w->S = S;
    w->stream = vgpu_stream_create();
    if (!w->stream) return -1;
    if (pthread_create(&w->th, NULL, sv_slot_main, w) != 0) {
        vgpu_stream_destroy(w->stream);
        w->stream = NULL;
        return vgpu_error("serve: cannot start a pass slot thread");
    }
    w->started = 1;
    pthread_mutex_lock(&S->mu);
    while (w->ready == 0) pthread_cond_wait(&S->cv_main, &S->mu);
    const int ok = w->ready > 0;
    pthread_mutex_unlock(&S->mu);
    return ok ? 0 : -1;
}

static void sv_slot_close(SvSched *S, SvSlot *w) {
    if (!w->started) return;
    pthread_mutex_lock(&S->mu);
    w->state = 2;
    pthread_cond_broadcast(&S->cv_slot);
    pthread_mutex_unlock(&S->mu);
    pthread_join(w->th, NULL);
    vgpu_stream_destroy(w->stream);
    w->started = 0;
}

static void sv_abs_deadline(double t0, double at, struct timespec *ts) {
    /* cv_main runs on CLOCK_MONOTONIC, like sv_now */
    double t = t0 + at;
    ts->tv_sec = (time_t)t;
    ts->tv_nsec = (long)((t - (double)ts->tv_sec) * 1e9);
    if (ts->tv_nsec >= 1000000000L) {
        ts->tv_sec++;
        ts->tv_nsec -= 1000000000L;
    }
}

static int sv_run_batched(const vx_serve_cfg *c, vgpu_jtable *jt, const vx_request *reqs, int n, vx_result *out,
                          vx_serve_stats *st, int64_t resident) {
    int kmax = c->max_batch < 1 ? 1 : (c->max_batch > SV_MAX_BATCH ? SV_MAX_BATCH : c->max_batch);
    int want = c->in_flight <= 0 ? 2 : (c->in_flight > SV_MAX_SLOTS ? SV_MAX_SLOTS : c->in_flight);
    SvSched S;
    memset(&S, 0, sizeof S);
    pthread_condattr_t ca;
    pthread_condattr_init(&ca);
    pthread_condattr_setclock(&ca, CLOCK_MONOTONIC);
    pthread_mutex_init(&S.mu, NULL);
    pthread_cond_init(&S.cv_slot, NULL);
    pthread_cond_init(&S.cv_main, &ca);
    pthread_condattr_destroy(&ca);
    S.c = c;
    S.jt = jt;
    S.reqs = reqs;
    S.out = out;
    S.kmax = kmax;
    SvSlot slots[SV_MAX_SLOTS];
    memset(slots, 0, sizeof slots);
    int n_slots = 0, rc = 0;
    /* Slots open while their pass state fits the budget (at least one):
     * each holds a template of each class planned for kmax requests. */
    while (n_slots < want && rc == 0) {
        if (sv_slot_open(&S, &slots[n_slots]) < 0) {
            if (slots[n_slots].err[0]) vgpu_error(slots[n_slots].err);
            rc = -1;
            n_slots++;
            break;
        }
        const int64_t tb = slots[n_slots].tpl_bytes;
        if (n_slots > 0 && resident + tb > st->hbm_budget) {
            sv_slot_close(&S, &slots[n_slots]);
            break;
        }
        resident += tb;
        n_slots++;
    }
    st->slots = n_slots;
    st->resident_bytes = resident;
    S.kcap[0] = S.kcap[1] = kmax;
    if (rc == 0 && n_slots > 0) {
        /* every slot planned the same templates */
        for (int cls = 0; cls < 2; cls++) S.kcap[cls] = slots[0].kcap[cls];
        /* a request holds its share of the larger template */
        int64_t per = 0;
        for (int cls = 0; cls < 2; cls++) {
            int64_t b = vgpu_batch_device_bytes(slots[0].tpl[cls]) / S.kcap[cls];
            if (b > per) per = b;
        }
        st->state_bytes = per;
        if (c->state_budget > 0 && per > 0 && c->state_budget / per < kmax) {
            kmax = (int)(c->state_budget / per) > 0 ? (int)(c->state_budget / per) : 1;
            for (int cls = 0; cls < 2; cls++)
                if (S.kcap[cls] > kmax) S.kcap[cls] = kmax;
        }
    }
    int *pending = (int *)malloc((size_t)(n > 0 ? n : 1) * sizeof(int));
    if (!pending && rc == 0) rc = vgpu_error("serve: alloc failed");
    S.t0 = sv_now();
    int next = 0, n_pend = 0;
    pthread_mutex_lock(&S.mu);
    while (rc == 0 && !S.failed && S.done < n) {
        double now = sv_now() - S.t0;
        while (next < n && reqs[next].arrival_s <= now) pending[n_pend++] = next++;
        /* fill every free slot: the oldest waiting request's class, every
         * waiting request of that class in arrival order, up to kmax */
        for (int s = 0; s < n_slots && n_pend > 0; s++) {
            SvSlot *w = &slots[s];
            if (w->state != 0) continue;
            int cls = reqs[pending[0]].cls, k = 0, wr = 0;
            for (int j = 0; j < n_pend; j++) {
                int i = pending[j];
                if (reqs[i].cls == cls && k < S.kcap[cls]) w->take[k++] = i;
                else pending[wr++] = i;
            }
            n_pend = wr;
            w->cls = cls;
            w->k = k;
            w->pass = st->n_passes++;
            w->state = 1;
            S.in_flight++;
            if (S.in_flight > st->max_in_flight) st->max_in_flight = S.in_flight;
            if (k > st->max_pass) st->max_pass = k;
            pthread_cond_broadcast(&S.cv_slot);
        }
        if (S.done >= n || S.failed) break;
        /* wait for a slot to free or for the next arrival */
        if (next < n && (n_pend == 0 || S.in_flight < n_slots)) {
            struct timespec ts;
            sv_abs_deadline(S.t0, reqs[next].arrival_s, &ts);
            pthread_cond_timedwait(&S.cv_main, &S.mu, &ts);
        } else {
            pthread_cond_wait(&S.cv_main, &S.mu);
        }
    }
    /* let passes in flight finish */
    while (S.in_flight > 0) pthread_cond_wait(&S.cv_main, &S.mu);
    pthread_mutex_unlock(&S.mu);
    st->n_done = S.done;
    st->wall_s = sv_now() - S.t0 - (n > 0 ? reqs[0].arrival_s : 0);
    for (int s = 0; s < n_slots; s++) sv_slot_close(&S, &slots[s]);
    free(pending);
    if (rc == 0 && S.failed) rc = vgpu_error(S.err);
    pthread_cond_destroy(&S.cv_slot);
    pthread_cond_destroy(&S.cv_main);
    pthread_mutex_destroy(&S.mu);
    return rc;
}

static int64_t sv_col_bytes(const vgpu_col *col, int64_t n) {
    if (!col->data) return 0;
    return n * (col->type == VGPU_BOOL ? 1 : 8) + (col->validity ? (n + 7) / 8 : 0);
}

int vx_serve_run(const vx_serve_cfg *c, const vx_request *reqs, int n, vx_result *out, vx_serve_stats *st) {
    memset(st, 0, sizeof *st);
    memset(out, 0, (size_t)n * sizeof *out);
    double s0 = sv_now();
    int64_t total = 0;
    if (vgpu_mem_info(NULL, &total) < 0) return -1;
    st->hbm_budget = c->hbm_budget > 0 ? c->hbm_budget : total;
    /* the dimension's join table stays resident for every class-1 request */
    vgpu_jtable *jt = vgpu_join_build_payload(&c->dim_key, c->n_dim, &c->dim_attr, 1);
    if (!jt) return -1;
    int64_t resident = sv_col_bytes(&c->fact_key, c->n_rows) + sv_col_bytes(&c->fact_value, c->n_rows) +
                       sv_col_bytes(&c->fact_group, c->n_rows) + sv_col_bytes(&c->dim_key, c->n_dim) +
                       sv_col_bytes(&c->dim_attr, c->n_dim) + vgpu_join_device_bytes(jt);
    int rc = 0;
    if (resident > st->hbm_budget) {
        vgpu_join_destroy(jt);
        char msg[160];
        snprintf(msg, sizeof msg, "serve: resident columns and join table (%lld bytes) exceed the budget of %lld",
                 (long long)resident, (long long)st->hbm_budget);
        return vgpu_error(msg);
    }
    /* the pass slots plan on their own streams: the table must be built first */
    if (vgpu_synchronize() < 0) rc = -1;
    if (rc == 0 && c->batched) {
        st->setup_s = sv_now() - s0;
        rc = sv_run_batched(c, jt, reqs, n, out, st, resident);
        vgpu_join_destroy(jt);
        return rc;
    }
    st->resident_bytes = resident;
    st->setup_s = sv_now() - s0;
    double t0 = sv_now();
    int done = 0;
    for (int i = 0; rc == 0 && i < n; i++) {
        sv_sleep_until(t0, reqs[i].arrival_s);
        out[i].start_s = sv_now() - t0;
        if (sv_plain_one(c, jt, &reqs[i], &out[i]) < 0) {
            rc = -1;
            break;
        }
        out[i].end_s = sv_now() - t0;
        out[i].pass = -1;
        out[i].batch = 1;
        done++;
    }
    st->n_done = done;
    st->max_in_flight = done > 0 ? 1 : 0;
    st->slots = 1;
    st->wall_s = sv_now() - t0 - (n > 0 ? reqs[0].arrival_s : 0);
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
