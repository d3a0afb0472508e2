/*
 * oracle.c -- CPU restatement of the reference hot path.  TEST
 * INFRASTRUCTURE ONLY: the checker for tests/, smoke() and bench.py's
 * cpu_baseline.  The product (vectra_amd/) never links or calls this file.
 *
 * Every function names the reference lines it follows (gcol33/vectra 0.2.2).
 * The arithmetic is restated, not copied; where the reference has a quirk
 * (NaN compare, left-join chain walk on uncoerced keys, bool min/max) the
 * quirk is reproduced, since parity means the reference's results.
 */
#define _GNU_SOURCE
#include "oracle.h"
#include "vx_core.h"

#include <math.h>
#include <setjmp.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* ==================================================================== */
/* coercion (src/coerce.c:7-71)                                          */
/* ==================================================================== */

static VecType common_type(VecType a, VecType b) {
    if (a == b) return a;
    if (a == VEC_STRING || b == VEC_STRING)
        vectra_error("cannot coerce string in arithmetic/comparison");
    if (a == VEC_DOUBLE || b == VEC_DOUBLE) return VEC_DOUBLE;
    if (a == VEC_INT64 || b == VEC_INT64) return VEC_INT64;
    return VEC_BOOL;
}

static VecArray *new_array(VecType t, int64_t n) {
    VecArray *a = (VecArray *)malloc(sizeof *a);
    *a = vec_array_alloc(t, n);
    return a;
}

static void drop_array(VecArray *a) {
    if (!a) return;
    vec_array_free(a);
    free(a);
}

static void copy_validity(VecArray *dst, const VecArray *src) {
    for (int64_t i = 0; i < src->length; i++)
        if (vec_array_is_valid(src, i)) vec_array_set_valid(dst, i);
}

/* vec_coerce: same type = deep copy; bool->int64, bool->double,
 * int64->double widen; anything else is an error. */
static VecArray *coerce(const VecArray *a, VecType to) {
    int64_t n = a->length;
    VecArray *o = new_array(to, n);
    copy_validity(o, a);
    if (a->type == to) {
        switch (to) {
        case VEC_INT64:  memcpy(o->buf.i64, a->buf.i64, (size_t)n * 8); break;
        case VEC_DOUBLE: memcpy(o->buf.dbl, a->buf.dbl, (size_t)n * 8); break;
        case VEC_BOOL:   memcpy(o->buf.bln, a->buf.bln, (size_t)n); break;
        case VEC_STRING:
            memcpy(o->buf.str.offsets, a->buf.str.offsets, (size_t)(n + 1) * 8);
            o->buf.str.data_len = a->buf.str.data_len;
            o->buf.str.data = (char *)malloc((size_t)(a->buf.str.data_len > 0 ? a->buf.str.data_len : 1));
            if (a->buf.str.data_len > 0)
                memcpy(o->buf.str.data, a->buf.str.data, (size_t)a->buf.str.data_len);
            break;
        }
        return o;
    }
    if (a->type == VEC_BOOL && to == VEC_INT64) {
        for (int64_t i = 0; i < n; i++) o->buf.i64[i] = a->buf.bln[i];
    } else if (a->type == VEC_BOOL && to == VEC_DOUBLE) {
        for (int64_t i = 0; i < n; i++) o->buf.dbl[i] = a->buf.bln[i];
    } else if (a->type == VEC_INT64 && to == VEC_DOUBLE) {
        for (int64_t i = 0; i < n; i++) o->buf.dbl[i] = (double)a->buf.i64[i];
    } else {
        drop_array(o);
        vectra_error("unsupported coercion: %d -> %d", a->type, to);
    }
    return o;
}

/* ==================================================================== */
/* scalar kernels (src/scalar_ops.c)                                     */
/* ==================================================================== */

/* vec_arith (src/scalar_ops.c:11-66): NA if either side NA; integer
 * division or modulo by zero gives NA; doubles use IEEE + - * / and fmod.
 * Signed overflow is undefined in the reference; it wraps here. */
static VecArray *arith(const VecArray *l, const VecArray *r, char op) {
    if (l->length != r->length) vectra_error("arith: mismatched lengths");
    VecType t = common_type(l->type, r->type);
    if (t == VEC_BOOL) t = VEC_INT64;
    VecArray *cl = coerce(l, t), *cr = coerce(r, t);
    int64_t n = l->length;
    VecArray *o = new_array(t, n);
    for (int64_t i = 0; i < n; i++) {
        if (!vec_array_is_valid(cl, i) || !vec_array_is_valid(cr, i)) continue;
        vec_array_set_valid(o, i);
        if (t == VEC_INT64) {
            uint64_t a = (uint64_t)cl->buf.i64[i], b = (uint64_t)cr->buf.i64[i];
            int64_t sa = cl->buf.i64[i], sb = cr->buf.i64[i];
            switch (op) {
            case '+': o->buf.i64[i] = (int64_t)(a + b); break;
            case '-': o->buf.i64[i] = (int64_t)(a - b); break;
            case '*': o->buf.i64[i] = (int64_t)(a * b); break;
            /* INT64_MIN / -1 overflows: undefined in the reference (x86
             * traps); it wraps here, and INT64_MIN % -1 is 0 */
            case '/':
                if (sb == 0) vec_array_set_null(o, i);
                else o->buf.i64[i] = sb == -1 ? (int64_t)(0 - a) : sa / sb;
                break;
            case '%':
                if (sb == 0) vec_array_set_null(o, i);
                else o->buf.i64[i] = sb == -1 ? 0 : sa % sb;
                break;
            default: vectra_error("unknown arith op: %c", op);
            }
        } else {
            double a = cl->buf.dbl[i], b = cr->buf.dbl[i];
            switch (op) {
            case '+': o->buf.dbl[i] = a + b; break;
            case '-': o->buf.dbl[i] = a - b; break;
            case '*': o->buf.dbl[i] = a * b; break;
            case '/': o->buf.dbl[i] = a / b; break;
            case '%': o->buf.dbl[i] = fmod(a, b); break;
            default: vectra_error("unknown arith op: %c", op);
            }
        }
    }
    drop_array(cl);
    drop_array(cr);
    return o;
}

/* The comparison outcome is decided on the three-way sign, exactly as
 * src/scalar_ops.c:125-141 does: cmp = (a>b)-(a<b), so NaN gives 0 and
 * satisfies ==, <= and >=. */
static int cmp_apply(int cmp, char op, char op2) {
    if (op == '=' && op2 == '=') return cmp == 0;
    if (op == '!' && op2 == '=') return cmp != 0;
    if (op == '<' && op2 == ' ') return cmp < 0;
    if (op == '<' && op2 == '=') return cmp <= 0;
    if (op == '>' && op2 == ' ') return cmp > 0;
    if (op == '>' && op2 == '=') return cmp >= 0;
    return 0;
}

/* vec_cmp (src/scalar_ops.c:70-147) */
static VecArray *compare(const VecArray *l, const VecArray *r, char op, char op2) {
    if (l->length != r->length) vectra_error("cmp: mismatched lengths");
    int64_t n = l->length;
    VecArray *o = new_array(VEC_BOOL, n);
    if (l->type == VEC_STRING && r->type == VEC_STRING) {
        for (int64_t i = 0; i < n; i++) {
            if (!vec_array_is_valid(l, i) || !vec_array_is_valid(r, i)) continue;
            vec_array_set_valid(o, i);
            int64_t ls = l->buf.str.offsets[i], ll = l->buf.str.offsets[i + 1] - ls;
            int64_t rs = r->buf.str.offsets[i], rl = r->buf.str.offsets[i + 1] - rs;
            int64_t m = ll < rl ? ll : rl;
            int c = memcmp(l->buf.str.data + ls, r->buf.str.data + rs, (size_t)m);
            if (c == 0) c = (ll > rl) - (ll < rl);
            o->buf.bln[i] = (uint8_t)cmp_apply(c, op, op2);
        }
        return o;
    }
    VecType t = common_type(l->type, r->type);
    if (t == VEC_BOOL) t = VEC_INT64;
    VecArray *cl = coerce(l, t), *cr = coerce(r, t);
    for (int64_t i = 0; i < n; i++) {
        if (!vec_array_is_valid(cl, i) || !vec_array_is_valid(cr, i)) continue;
        vec_array_set_valid(o, i);
        int c;
        if (t == VEC_INT64) {
            int64_t a = cl->buf.i64[i], b = cr->buf.i64[i];
            c = (a > b) - (a < b);
        } else {
            double a = cl->buf.dbl[i], b = cr->buf.dbl[i];
            c = (a > b) - (a < b);
        }
        o->buf.bln[i] = (uint8_t)cmp_apply(c, op, op2);
    }
    drop_array(cl);
    drop_array(cr);
    return o;
}

/* vec_bool_binary (src/scalar_ops.c:151-185): R three-valued & and |. */
static VecArray *bool_binary(const VecArray *l, const VecArray *r, char op) {
    if (l->length != r->length) vectra_error("bool: mismatched lengths");
    if (l->type != VEC_BOOL || r->type != VEC_BOOL)
        vectra_error("bool: operands must be boolean");
    int64_t n = l->length;
    VecArray *o = new_array(VEC_BOOL, n);
    for (int64_t i = 0; i < n; i++) {
        int lv = vec_array_is_valid(l, i), rv = vec_array_is_valid(r, i);
        int lb = lv ? l->buf.bln[i] : 0, rb = rv ? r->buf.bln[i] : 0;
        int known, val;
        if (op == '&') {
            if ((lv && !lb) || (rv && !rb)) { known = 1; val = 0; }
            else if (lv && rv)              { known = 1; val = 1; }
            else                            { known = 0; val = 0; }
        } else {
            if ((lv && lb) || (rv && rb)) { known = 1; val = 1; }
            else if (lv && rv)            { known = 1; val = 0; }
            else                          { known = 0; val = 0; }
        }
        if (known) { vec_array_set_valid(o, i); o->buf.bln[i] = (uint8_t)val; }
    }
    return o;
}

/* vec_bool_not (src/scalar_ops.c:187-203) */
static VecArray *bool_not(const VecArray *a) {
    if (a->type != VEC_BOOL) vectra_error("bool not: operand must be boolean");
    VecArray *o = new_array(VEC_BOOL, a->length);
    copy_validity(o, a);
    for (int64_t i = 0; i < a->length; i++)
        if (vec_array_is_valid(a, i)) o->buf.bln[i] = (uint8_t)!a->buf.bln[i];
    return o;
}

/* vec_negate (src/scalar_ops.c:205-228) */
static VecArray *negate(const VecArray *a) {
    if (a->type != VEC_INT64 && a->type != VEC_DOUBLE)
        vectra_error("negate: unsupported type %d", a->type);
    VecArray *o = new_array(a->type, a->length);
    copy_validity(o, a);
    for (int64_t i = 0; i < a->length; i++) {
        if (a->type == VEC_INT64) o->buf.i64[i] = (int64_t)(0 - (uint64_t)a->buf.i64[i]);
        else o->buf.dbl[i] = -a->buf.dbl[i];
    }
    return o;
}

/* ==================================================================== */
/* expression evaluation (src/expr.c:91-146)                             */
/* ==================================================================== */

static const VecArray *find_col(const VecBatch *b, const char *name) {
    for (int i = 0; i < b->n_cols; i++)
        if (strcmp(b->col_names[i], name) == 0) return &b->columns[i];
    vectra_error("column not found: %s", name);
}

VecArray *oracle_expr_eval(const VecExpr *e, const VecBatch *b) {
    int64_t n = b->n_rows;
    switch (e->kind) {
    case EXPR_COL_REF: {
        const VecArray *c = find_col(b, e->col_name);
        return coerce(c, c->type);
    }
    case EXPR_LIT_INT64: {
        VecArray *o = new_array(VEC_INT64, n);
        vec_array_set_all_valid(o);
        for (int64_t i = 0; i < n; i++) o->buf.i64[i] = e->lit_i64;
        return o;
    }
    case EXPR_LIT_DOUBLE: {
        VecArray *o = new_array(VEC_DOUBLE, n);
        vec_array_set_all_valid(o);
        for (int64_t i = 0; i < n; i++) o->buf.dbl[i] = e->lit_dbl;
        return o;
    }
    case EXPR_LIT_BOOL: {
        VecArray *o = new_array(VEC_BOOL, n);
        vec_array_set_all_valid(o);
        for (int64_t i = 0; i < n; i++) o->buf.bln[i] = e->lit_bln;
        return o;
    }
    case EXPR_LIT_STRING: {
        VecArray *o = new_array(VEC_STRING, n);
        vec_array_set_all_valid(o);
        int64_t L = (int64_t)strlen(e->lit_str);
        o->buf.str.data = (char *)malloc((size_t)(L * n > 0 ? L * n : 1));
        o->buf.str.data_len = L * n;
        for (int64_t i = 0; i < n; i++) {
            o->buf.str.offsets[i] = i * L;
            memcpy(o->buf.str.data + i * L, e->lit_str, (size_t)L);
        }
        o->buf.str.offsets[n] = L * n;
        return o;
    }
    case EXPR_LIT_NA:
        return new_array(e->result_type, n);
    case EXPR_ARITH: {
        VecArray *l = oracle_expr_eval(e->left, b), *r = oracle_expr_eval(e->right, b);
        VecArray *o = arith(l, r, e->op);
        drop_array(l); drop_array(r);
        return o;
    }
    case EXPR_CMP: {
        VecArray *l = oracle_expr_eval(e->left, b), *r = oracle_expr_eval(e->right, b);
        VecArray *o = compare(l, r, e->op, e->op2);
        drop_array(l); drop_array(r);
        return o;
    }
    case EXPR_BOOL: {
        if (e->op == '!') {
            VecArray *a = oracle_expr_eval(e->operand, b);
            VecArray *o = bool_not(a);
            drop_array(a);
            return o;
        }
        VecArray *l = oracle_expr_eval(e->left, b), *r = oracle_expr_eval(e->right, b);
        VecArray *o = bool_binary(l, r, e->op);
        drop_array(l); drop_array(r);
        return o;
    }
    case EXPR_IS_NA: {
        VecArray *a = oracle_expr_eval(e->operand, b);
        VecArray *o = new_array(VEC_BOOL, a->length);
        vec_array_set_all_valid(o);
        for (int64_t i = 0; i < a->length; i++) o->buf.bln[i] = (uint8_t)!vec_array_is_valid(a, i);
        drop_array(a);
        return o;
    }
    case EXPR_NEGATE: {
        VecArray *a = oracle_expr_eval(e->operand, b);
        VecArray *o = negate(a);
        drop_array(a);
        return o;
    }
    case EXPR_IF_ELSE:
    case EXPR_CAST:
        vectra_error("if_else/cast not yet implemented");
    default:
        vectra_error("oracle: expression kind %d not on the hot path", (int)e->kind);
    }
}

int oracle_try_expr_eval(const VecExpr *e, const VecBatch *b, VecArray **out) {
    vx_frame f;
    *out = NULL;
    vx_frame_push(&f);
    if (setjmp(f.jb) == 0) {
        *out = oracle_expr_eval(e, b);
        vx_frame_pop(&f);
        return 0;
    }
    return -1;
}

void oracle_array_free(VecArray *a) { drop_array(a); }

/* ==================================================================== */
/* filter (src/filter.c:13-84)                                           */
/* ==================================================================== */

typedef struct {
    VecNode  base;
    VecNode *child;
    VecExpr *pred;
} OFilter;

static VecBatch *ofilter_next(VecNode *self) {
    OFilter *f = (OFilter *)self;
    VecBatch *b;
    while ((b = f->child->next_batch(f->child)) != NULL) {
        VecArray *m = oracle_expr_eval(f->pred, b);
        if (m->type != VEC_BOOL) {
            drop_array(m);
            vec_batch_free(b);
            vectra_error("filter predicate must evaluate to boolean");
        }
        int64_t nl = vec_batch_logical_rows(b);
        int32_t *sel = (int32_t *)malloc((size_t)(nl > 0 ? nl : 1) * 4);
        int32_t k = 0;
        for (int64_t li = 0; li < nl; li++) {
            int64_t p = vec_batch_physical_row(b, li);
            if (vec_array_is_valid(m, p) && m->buf.bln[p]) sel[k++] = (int32_t)p;
        }
        drop_array(m);
        free(b->sel);
        b->sel = sel;
        b->sel_n = k;
        if (k > 0) return b;
        vec_batch_free(b);
    }
    return NULL;
}

static void ofilter_free(VecNode *self) {
    OFilter *f = (OFilter *)self;
    f->child->free_node(f->child);
    vec_expr_free(f->pred);
    vec_schema_free(&f->base.output_schema);
    free(f);
}

VecNode *oracle_filter_node_create(VecNode *child, VecExpr *pred) {
    OFilter *f = (OFilter *)calloc(1, sizeof *f);
    f->child = child;
    f->pred = pred;
    f->base.output_schema = vec_schema_copy(&child->output_schema);
    f->base.next_batch = ofilter_next;
    f->base.free_node = ofilter_free;
    f->base.kind = "FilterNode";
    return &f->base;
}

/* ==================================================================== */
/* hashing and key equality (src/hash.c:12-78, src/hash.h:11-14)         */
/* ==================================================================== */

#define FNV_BASIS 0xcbf29ce484222325ULL
#define FNV_MUL   0x100000001b3ULL

static uint64_t fnv_bytes(const uint8_t *p, int64_t n) {
    uint64_t h = FNV_BASIS;
    for (int64_t k = 0; k < n; k++) { h ^= p[k]; h *= FNV_MUL; }
    return h;
}

uint64_t oracle_hash_value(const VecArray *a, int64_t row) {
    if (!vec_array_is_valid(a, row)) return FNV_BASIS ^ 0xFF;
    switch (a->type) {
    case VEC_INT64: {
        int64_t v = a->buf.i64[row];
        return fnv_bytes((const uint8_t *)&v, 8);
    }
    case VEC_DOUBLE: {
        double v = a->buf.dbl[row];
        if (v == 0.0) v = 0.0;   /* -0 hashes as +0 */
        return fnv_bytes((const uint8_t *)&v, 8);
    }
    case VEC_BOOL:
        return fnv_bytes(&a->buf.bln[row], 1);
    case VEC_STRING: {
        int64_t s = a->buf.str.offsets[row], e = a->buf.str.offsets[row + 1];
        return fnv_bytes((const uint8_t *)a->buf.str.data + s, e - s);
    }
    }
    return 0;
}

static uint64_t hash_combine(uint64_t h1, uint64_t h2) {
    return ((h1 << 5) | (h1 >> 59)) ^ h2;
}

/* val_equal (src/hash.c:48-70): NA == NA for grouping. */
static int group_val_equal(const VecArray *a, int64_t ra, const VecArray *b, int64_t rb) {
    int va = vec_array_is_valid(a, ra), vb = vec_array_is_valid(b, rb);
    if (!va && !vb) return 1;
    if (!va || !vb) return 0;
    switch (a->type) {
    case VEC_INT64:  return a->buf.i64[ra] == b->buf.i64[rb];
    case VEC_DOUBLE: return a->buf.dbl[ra] == b->buf.dbl[rb];
    case VEC_BOOL:   return a->buf.bln[ra] == b->buf.bln[rb];
    case VEC_STRING: {
        int64_t as = a->buf.str.offsets[ra], al = a->buf.str.offsets[ra + 1] - as;
        int64_t bs = b->buf.str.offsets[rb], bl = b->buf.str.offsets[rb + 1] - bs;
        return al == bl && memcmp(a->buf.str.data + as, b->buf.str.data + bs, (size_t)al) == 0;
    }
    }
    return 0;
}

/* ==================================================================== */
/* accumulators (src/agg_ops.c:25-255)                                   */
/* ==================================================================== */

typedef struct {
    AggKind  kind;
    VecType  in_type;
    int      na_rm;
    int64_t  n, cap;
    int64_t *count;
    int64_t *sum_i;
    double  *sum_d;
    int64_t *mm_i;
    double  *mm_d;
    uint8_t *has_value;
    uint8_t *has_na;
} Acc;

static void acc_grow(Acc *a, int64_t n) {
    if (n <= a->cap) { if (n > a->n) a->n = n; return; }
    int64_t c = a->cap ? a->cap : 64;
    while (c < n) c *= 2;
#define GROW(p, T) do { p = (T *)realloc(p, (size_t)c * sizeof(T)); \
                        memset(p + a->cap, 0, (size_t)(c - a->cap) * sizeof(T)); } while (0)
    GROW(a->count, int64_t);
    GROW(a->sum_i, int64_t);
    GROW(a->sum_d, double);
    GROW(a->mm_i, int64_t);
    GROW(a->mm_d, double);
    GROW(a->has_value, uint8_t);
    GROW(a->has_na, uint8_t);
#undef GROW
    a->cap = c;
    a->n = n;
}

static void acc_free(Acc *a) {
    free(a->count); free(a->sum_i); free(a->sum_d); free(a->mm_i);
    free(a->mm_d); free(a->has_value); free(a->has_na);
    memset(a, 0, sizeof *a);
}

/* agg_accum_feed (src/agg_ops.c:93-158) */
static void acc_feed(Acc *a, int64_t g, const VecArray *col, int64_t row) {
    int valid = col ? vec_array_is_valid(col, row) : 0;
    switch (a->kind) {
    case AGG_COUNT_STAR: a->count[g]++; return;
    case AGG_COUNT: if (valid) a->count[g]++; return;
    case AGG_SUM:
    case AGG_MEAN:
        if (!valid) { if (!a->na_rm) a->has_na[g] = 1; return; }
        a->count[g]++;
        if (a->in_type == VEC_DOUBLE) a->sum_d[g] += col->buf.dbl[row];
        else if (a->in_type == VEC_INT64)
            a->sum_i[g] = (int64_t)((uint64_t)a->sum_i[g] + (uint64_t)col->buf.i64[row]);
        else if (a->in_type == VEC_BOOL) a->sum_d[g] += (double)col->buf.bln[row];
        return;
    case AGG_MIN:
    case AGG_MAX: {
        if (!valid) { if (!a->na_rm) a->has_na[g] = 1; return; }
        int is_min = a->kind == AGG_MIN;
        if (a->in_type == VEC_DOUBLE) {
            double v = col->buf.dbl[row];
            if (!a->has_value[g] || (is_min ? v < a->mm_d[g] : v > a->mm_d[g])) {
                a->mm_d[g] = v;
                a->has_value[g] = 1;
            }
        } else if (a->in_type == VEC_INT64) {
            int64_t v = col->buf.i64[row];
            if (!a->has_value[g] || (is_min ? v < a->mm_i[g] : v > a->mm_i[g])) {
                a->mm_i[g] = v;
                a->has_value[g] = 1;
            }
        }
        /* bool input: the reference updates nothing (src/agg_ops.c:118-135) */
        return;
    }
    }
}

/* agg_accum_finish (src/agg_ops.c:160-255): every result is double. */
static VecArray acc_finish(const Acc *a) {
    int64_t n = a->n;
    VecArray o = vec_array_alloc(VEC_DOUBLE, n);
    for (int64_t g = 0; g < n; g++) {
        switch (a->kind) {
        case AGG_COUNT:
        case AGG_COUNT_STAR:
            vec_array_set_valid(&o, g);
            o.buf.dbl[g] = (double)a->count[g];
            break;
        case AGG_SUM:
            if (a->has_na[g]) break;
            vec_array_set_valid(&o, g);
            o.buf.dbl[g] = a->in_type == VEC_INT64 ? (double)a->sum_i[g] : a->sum_d[g];
            break;
        case AGG_MEAN:
            if (a->has_na[g]) break;
            vec_array_set_valid(&o, g);
            if (a->count[g] == 0) o.buf.dbl[g] = NAN;
            else o.buf.dbl[g] = (a->in_type == VEC_INT64 ? (double)a->sum_i[g] : a->sum_d[g])
                                / (double)a->count[g];
            break;
        case AGG_MIN:
        case AGG_MAX:
            if (a->has_na[g]) break;
            vec_array_set_valid(&o, g);
            if (!a->has_value[g]) {
                o.buf.dbl[g] = a->kind == AGG_MIN ? INFINITY : -INFINITY;
                vectra_warning(a->kind == AGG_MIN
                    ? "no non-missing arguments to min; returning Inf"
                    : "no non-missing arguments to max; returning -Inf");
            } else {
                o.buf.dbl[g] = a->in_type == VEC_DOUBLE ? a->mm_d[g] : (double)a->mm_i[g];
            }
            break;
        }
    }
    return o;
}

/* ==================================================================== */
/* group-by (src/group_agg.c)                                            */
/* ==================================================================== */

typedef struct {
    VecNode  base;
    VecNode *child;
    int      n_keys;
    char   **key_names;
    int      n_aggs;
    AggSpec *specs;
    int      sorted;
    int      done;
} OGroup;

/* growing dense copy of whole input rows (keys + agg inputs), used as the
 * key arena for the hash path and as the materialized input of the sort;
 * string values are appended to the column's data (builder.c's layout:
 * offsets[n + 1] into one byte buffer) */
typedef struct {
    int       nc;
    VecType  *types;
    VecArray *cols;
    int64_t  *dcap;     /* per string column: bytes allocated for data */
    int64_t   n, cap;
} RowStore;

static void rs_init(RowStore *s, int nc, const VecType *types) {
    s->nc = nc;
    s->types = (VecType *)malloc(sizeof(VecType) * (size_t)(nc ? nc : 1));
    memcpy(s->types, types, sizeof(VecType) * (size_t)nc);
    s->cols = (VecArray *)calloc((size_t)(nc ? nc : 1), sizeof(VecArray));
    s->dcap = (int64_t *)calloc((size_t)(nc ? nc : 1), sizeof(int64_t));
    s->n = 0;
    s->cap = 64;
    for (int c = 0; c < nc; c++) {
        s->cols[c] = vec_array_alloc(types[c], s->cap);
        if (types[c] == VEC_STRING) {
            s->dcap[c] = 256;
            s->cols[c].buf.str.data = (char *)malloc(256);
        }
    }
}

static void rs_append(RowStore *s, const VecArray *const *src, int64_t row) {
    if (s->n == s->cap) {
        int64_t nc2 = s->cap * 2;
        for (int c = 0; c < s->nc; c++) {
            VecArray *a = &s->cols[c];
            if (a->type == VEC_STRING) {
                a->buf.str.offsets = (int64_t *)realloc(a->buf.str.offsets, (size_t)(nc2 + 1) * 8);
            } else {
                size_t w = a->type == VEC_BOOL ? 1 : 8;
                a->buf.i64 = (int64_t *)realloc(a->buf.i64, (size_t)nc2 * w);
            }
            a->validity = (uint8_t *)realloc(a->validity, (size_t)vec_validity_bytes(nc2));
            memset(a->validity + vec_validity_bytes(s->cap), 0,
                   (size_t)(vec_validity_bytes(nc2) - vec_validity_bytes(s->cap)));
            a->length = nc2;
        }
        s->cap = nc2;
    }
    for (int c = 0; c < s->nc; c++) {
        VecArray *a = &s->cols[c];
        const VecArray *x = src[c];
        if (x->type != a->type) vectra_error("oracle: column type changed between batches");
        const int ok = vec_array_is_valid(x, row);
        if (ok) vec_array_set_valid(a, s->n);
        if (a->type == VEC_STRING) {
            const int64_t at = s->n ? a->buf.str.offsets[s->n] : 0;
            int64_t len = 0;
            if (ok) {
                const int64_t b0 = x->buf.str.offsets[row];
                len = x->buf.str.offsets[row + 1] - b0;
                if (at + len > s->dcap[c]) {
                    while (at + len > s->dcap[c]) s->dcap[c] *= 2;
                    a->buf.str.data = (char *)realloc(a->buf.str.data, (size_t)s->dcap[c]);
                }
                memcpy(a->buf.str.data + at, x->buf.str.data + b0, (size_t)len);
            }
            a->buf.str.offsets[s->n] = at;
            a->buf.str.offsets[s->n + 1] = at + len;
            a->buf.str.data_len = at + len;
        } else if (ok) {
            if (a->type == VEC_BOOL) a->buf.bln[s->n] = x->buf.bln[row];
            else a->buf.i64[s->n] = x->buf.i64[row];   /* 8 raw bytes */
        }
    }
    s->n++;
}

static void rs_free(RowStore *s) {
    for (int c = 0; c < s->nc; c++) vec_array_free(&s->cols[c]);
    free(s->cols);
    free(s->types);
    free(s->dcap);
}

/* out[0..n) = the strings s[rows[i]] (NA for rows[i] < 0 or an NA value),
 * written as one offsets + data array (src/builder.c:110-138) */
static void str_gather(VecArray *d, const VecArray *s, const int64_t *rows, int64_t n) {
    int64_t total = 0;
    for (int64_t i = 0; i < n; i++)
        if (rows[i] >= 0 && vec_array_is_valid(s, rows[i]))
            total += s->buf.str.offsets[rows[i] + 1] - s->buf.str.offsets[rows[i]];
    free(d->buf.str.data);
    d->buf.str.data = (char *)malloc((size_t)(total > 0 ? total : 1));
    d->buf.str.data_len = total;
    int64_t at = 0;
    for (int64_t i = 0; i < n; i++) {
        d->buf.str.offsets[i] = at;
        const int64_t r = rows[i];
        if (r < 0 || !vec_array_is_valid(s, r)) continue;
        vec_array_set_valid(d, i);
        const int64_t b0 = s->buf.str.offsets[r], len = s->buf.str.offsets[r + 1] - b0;
        memcpy(d->buf.str.data + at, s->buf.str.data + b0, (size_t)len);
        at += len;
    }
    d->buf.str.offsets[n] = at;
}

static int resolve(const VecSchema *s, const char *name, const char *what) {
    int i = vec_schema_find_col(s, name);
    if (i < 0) vectra_error("%s: column not found: %s", what, name);
    return i;
}

/* hash path: src/group_agg.c:155-281 with src/hash.c:80-165 */
static VecBatch *ogroup_hash(OGroup *g) {
    const VecSchema *cs = &g->child->output_schema;
    int nk = g->n_keys, na = g->n_aggs;
    int *kidx = (int *)malloc(sizeof(int) * (size_t)(nk + 1));
    VecType *ktypes = (VecType *)malloc(sizeof(VecType) * (size_t)(nk + 1));
    for (int k = 0; k < nk; k++) {
        kidx[k] = resolve(cs, g->key_names[k], "group_by");
        ktypes[k] = cs->col_types[kidx[k]];
    }
    int *aidx = (int *)malloc(sizeof(int) * (size_t)(na + 1));
    Acc *acc = (Acc *)calloc((size_t)(na + 1), sizeof(Acc));
    for (int a = 0; a < na; a++) {
        if (g->specs[a].kind == AGG_COUNT_STAR) { aidx[a] = -1; acc[a].in_type = VEC_INT64; }
        else {
            aidx[a] = resolve(cs, g->specs[a].input_col, "summarise");
            acc[a].in_type = cs->col_types[aidx[a]];
        }
        acc[a].kind = g->specs[a].kind;
        acc[a].na_rm = g->specs[a].na_rm;
    }
    RowStore arena;
    rs_init(&arena, nk, ktypes);
    int64_t nslots = 64, ngroups = 0;
    int64_t *slot_gid = (int64_t *)malloc((size_t)nslots * 8);
    uint64_t *slot_h = (uint64_t *)malloc((size_t)nslots * 8);
    memset(slot_gid, -1, (size_t)nslots * 8);
    const VecArray **krow = (const VecArray **)malloc(sizeof(void *) * (size_t)(nk + 1));

    VecBatch *b;
    while ((b = g->child->next_batch(g->child)) != NULL) {
        for (int k = 0; k < nk; k++) krow[k] = &b->columns[kidx[k]];
        int64_t nl = vec_batch_logical_rows(b);
        for (int64_t li = 0; li < nl; li++) {
            int64_t r = vec_batch_physical_row(b, li);
            uint64_t h = 0;
            for (int k = 0; k < nk; k++) {
                uint64_t kh = oracle_hash_value(krow[k], r);
                h = k == 0 ? kh : hash_combine(h, kh);
            }
            /* grow before probing once load passes 70% (src/hash.c:109) */
            if (ngroups * 10 > nslots * 7) {
                int64_t n2 = nslots * 2;
                int64_t *g2 = (int64_t *)malloc((size_t)n2 * 8);
                uint64_t *h2 = (uint64_t *)malloc((size_t)n2 * 8);
                memset(g2, -1, (size_t)n2 * 8);
                for (int64_t s = 0; s < nslots; s++) {
                    if (slot_gid[s] < 0) continue;
                    int64_t t = (int64_t)(slot_h[s] & (uint64_t)(n2 - 1));
                    while (g2[t] >= 0) t = (t + 1) & (n2 - 1);
                    g2[t] = slot_gid[s];
                    h2[t] = slot_h[s];
                }
                free(slot_gid); free(slot_h);
                slot_gid = g2; slot_h = h2; nslots = n2;
            }
            int64_t s = (int64_t)(h & (uint64_t)(nslots - 1)), gid;
            for (;;) {
                if (slot_gid[s] < 0) {
                    gid = ngroups++;
                    slot_gid[s] = gid;
                    slot_h[s] = h;
                    rs_append(&arena, krow, r);
                    for (int a = 0; a < na; a++) acc_grow(&acc[a], ngroups);
                    break;
                }
                if (slot_h[s] == h) {
                    int64_t cand = slot_gid[s];
                    int eq = 1;
                    for (int k = 0; k < nk && eq; k++)
                        eq = group_val_equal(&arena.cols[k], cand, krow[k], r);
                    if (eq) { gid = cand; break; }
                }
                s = (s + 1) & (nslots - 1);
            }
            for (int a = 0; a < na; a++)
                acc_feed(&acc[a], gid, aidx[a] >= 0 ? &b->columns[aidx[a]] : NULL, r);
        }
        vec_batch_free(b);
    }

    VecBatch *out = vec_batch_alloc(nk + na, ngroups);
    for (int k = 0; k < nk; k++) {
        VecArray d = vec_array_alloc(ktypes[k], ngroups);
        if (d.type == VEC_STRING) {
            int64_t *ix = (int64_t *)malloc((size_t)(ngroups > 0 ? ngroups : 1) * 8);
            for (int64_t i = 0; i < ngroups; i++) ix[i] = i;
            str_gather(&d, &arena.cols[k], ix, ngroups);
            free(ix);
        }
        for (int64_t i = 0; i < ngroups && d.type != VEC_STRING; i++) {
            if (!vec_array_is_valid(&arena.cols[k], i)) continue;
            vec_array_set_valid(&d, i);
            if (d.type == VEC_BOOL) d.buf.bln[i] = arena.cols[k].buf.bln[i];
            else d.buf.i64[i] = arena.cols[k].buf.i64[i];
        }
        out->columns[k] = d;
        out->col_names[k] = vx_strdup(g->key_names[k]);
    }
    for (int a = 0; a < na; a++) {
        acc[a].n = ngroups;
        if (ngroups > acc[a].cap) acc_grow(&acc[a], ngroups);
        out->columns[nk + a] = acc_finish(&acc[a]);
        out->col_names[nk + a] = vx_strdup(g->specs[a].output_name);
        acc_free(&acc[a]);
    }
    rs_free(&arena);
    free(slot_gid); free(slot_h); free(krow);
    free(kidx); free(ktypes); free(aidx); free(acc);
    return out;
}

/* compare_value (src/sort.c:32-71), ascending: NA last; doubles via < and >
 * so NaN compares equal to everything. */
static int sort_cmp_value(const VecArray *a, int64_t ra, int64_t rb) {
    int va = vec_array_is_valid(a, ra), vb = vec_array_is_valid(a, rb);
    if (!va && !vb) return 0;
    if (!va) return 1;
    if (!vb) return -1;
    switch (a->type) {
    case VEC_DOUBLE: {
        double x = a->buf.dbl[ra], y = a->buf.dbl[rb];
        return x < y ? -1 : x > y ? 1 : 0;
    }
    case VEC_INT64: {
        int64_t x = a->buf.i64[ra], y = a->buf.i64[rb];
        return x < y ? -1 : x > y ? 1 : 0;
    }
    case VEC_BOOL:
        return (int)a->buf.bln[ra] - (int)a->buf.bln[rb];
    case VEC_STRING: {
        /* bytes over the shorter length, then the shorter first (src/sort.c:58-66) */
        const int64_t sa = a->buf.str.offsets[ra], la = a->buf.str.offsets[ra + 1] - sa;
        const int64_t sb = a->buf.str.offsets[rb], lb = a->buf.str.offsets[rb + 1] - sb;
        const int c = memcmp(a->buf.str.data + sa, a->buf.str.data + sb, (size_t)(la < lb ? la : lb));
        if (c) return c < 0 ? -1 : 1;
        return la < lb ? -1 : la > lb ? 1 : 0;
    }
    }
    return 0;
}

typedef struct { const VecArray *cols; int nk; } SortCtx;

static int sort_cmp_rows(const SortCtx *c, int64_t a, int64_t b) {
    for (int k = 0; k < c->nk; k++) {
        int r = sort_cmp_value(&c->cols[k], a, b);
        if (r) return r;
    }
    return 0;
}

/* merge_sort_impl (src/sort.c:97-116): top-down, split at n/2, ties and
 * NaN-induced "equal" results keep the left element. */
static void merge_sort(int64_t *ix, int64_t *tmp, int64_t n, const SortCtx *c) {
    if (n <= 1) return;
    int64_t mid = n / 2;
    merge_sort(ix, tmp, mid, c);
    merge_sort(ix + mid, tmp, n - mid, c);
    int64_t i = 0, j = mid, k = 0;
    while (i < mid && j < n)
        tmp[k++] = sort_cmp_rows(c, ix[i], ix[j]) <= 0 ? ix[i++] : ix[j++];
    while (i < mid) tmp[k++] = ix[i++];
    while (j < n) tmp[k++] = ix[j++];
    memcpy(ix, tmp, (size_t)n * 8);
}

/* group boundary test of snap_matches (src/group_agg.c:326-359): validity
 * must agree and values compare with != (NaN never matches). */
static int same_group(const VecArray *cols, int nk, int64_t a, int64_t b) {
    for (int k = 0; k < nk; k++) {
        const VecArray *x = &cols[k];
        int va = vec_array_is_valid(x, a), vb = vec_array_is_valid(x, b);
        if (va != vb) return 0;
        if (!va) continue;
        switch (x->type) {
        case VEC_INT64:  if (x->buf.i64[a] != x->buf.i64[b]) return 0; break;
        case VEC_DOUBLE: if (x->buf.dbl[a] != x->buf.dbl[b]) return 0; break;
        case VEC_BOOL:   if (x->buf.bln[a] != x->buf.bln[b]) return 0; break;
        case VEC_STRING: if (!group_val_equal(x, a, x, b)) return 0; break;
        }
    }
    return 1;
}

/* sort path: SortNode in-memory (src/sort.c:488-529, 566-636) feeding
 * sorted_agg_next_batch (src/group_agg.c:466-585). */
static VecBatch *ogroup_sorted(OGroup *g) {
    const VecSchema *cs = &g->child->output_schema;
    int nk = g->n_keys, na = g->n_aggs, nc = nk + na;
    int *cidx = (int *)malloc(sizeof(int) * (size_t)(nc + 1));
    VecType *ctypes = (VecType *)malloc(sizeof(VecType) * (size_t)(nc + 1));
    for (int k = 0; k < nk; k++) cidx[k] = resolve(cs, g->key_names[k], "group_by");
    for (int a = 0; a < na; a++)
        cidx[nk + a] = g->specs[a].kind == AGG_COUNT_STAR ? -1
                     : resolve(cs, g->specs[a].input_col, "summarise");
    for (int c = 0; c < nc; c++) ctypes[c] = cidx[c] >= 0 ? cs->col_types[cidx[c]] : VEC_INT64;

    /* materialize every logical row of the child in pull order; count-star
     * slots reuse the first key column so every stored column has a source */
    for (int c = nk; c < nc; c++) if (cidx[c] < 0) { cidx[c] = cidx[0]; ctypes[c] = ctypes[0]; }
    RowStore st;
    rs_init(&st, nc, ctypes);
    const VecArray **src = (const VecArray **)malloc(sizeof(void *) * (size_t)(nc + 1));
    VecBatch *b;
    while ((b = g->child->next_batch(g->child)) != NULL) {
        for (int c = 0; c < nc; c++) src[c] = &b->columns[cidx[c]];
        int64_t nl = vec_batch_logical_rows(b);
        for (int64_t li = 0; li < nl; li++)
            rs_append(&st, src, vec_batch_physical_row(b, li));
        vec_batch_free(b);
    }
    for (int a = 0; a < na; a++)
        if (g->specs[a].kind != AGG_COUNT_STAR) ctypes[nk + a] = st.cols[nk + a].type;

    int64_t n = st.n;
    int64_t *ix = (int64_t *)malloc((size_t)(n > 0 ? n : 1) * 8);
    int64_t *tmp = (int64_t *)malloc((size_t)(n > 0 ? n : 1) * 8);
    for (int64_t i = 0; i < n; i++) ix[i] = i;
    SortCtx ctx = { st.cols, nk };
    merge_sort(ix, tmp, n, &ctx);

    /* linear scan: one accumulator slot per group, emitted at each boundary */
    Acc *acc = (Acc *)calloc((size_t)(na + 1), sizeof(Acc));
    for (int a = 0; a < na; a++) {
        acc[a].kind = g->specs[a].kind;
        acc[a].in_type = ctypes[nk + a];
        acc[a].na_rm = g->specs[a].na_rm;
    }
    int64_t ngroups = 0;
    for (int64_t i = 0; i < n; i++)
        if (i == 0 || !same_group(st.cols, nk, ix[i - 1], ix[i])) ngroups++;
    for (int a = 0; a < na; a++) acc_grow(&acc[a], ngroups > 0 ? ngroups : 1);
    int64_t *first_row = (int64_t *)malloc((size_t)(ngroups > 0 ? ngroups : 1) * 8);
    int64_t gi = -1;
    for (int64_t i = 0; i < n; i++) {
        int64_t r = ix[i];
        if (i == 0 || !same_group(st.cols, nk, ix[i - 1], r)) first_row[++gi] = r;
        for (int a = 0; a < na; a++)
            acc_feed(&acc[a], gi, g->specs[a].kind == AGG_COUNT_STAR ? NULL : &st.cols[nk + a], r);
    }

    VecBatch *out = vec_batch_alloc(nk + na, ngroups);
    for (int k = 0; k < nk; k++) {
        VecArray d = vec_array_alloc(ctypes[k], ngroups);
        if (d.type == VEC_STRING) str_gather(&d, &st.cols[k], first_row, ngroups);
        for (int64_t q = 0; q < ngroups && d.type != VEC_STRING; q++) {
            int64_t r = first_row[q];
            if (!vec_array_is_valid(&st.cols[k], r)) continue;
            vec_array_set_valid(&d, q);
            if (d.type == VEC_BOOL) d.buf.bln[q] = st.cols[k].buf.bln[r];
            else d.buf.i64[q] = st.cols[k].buf.i64[r];
        }
        out->columns[k] = d;
        out->col_names[k] = vx_strdup(g->key_names[k]);
    }
    for (int a = 0; a < na; a++) {
        acc[a].n = ngroups;
        out->columns[nk + a] = acc_finish(&acc[a]);
        out->col_names[nk + a] = vx_strdup(g->specs[a].output_name);
        acc_free(&acc[a]);
    }
    free(acc); free(first_row); free(ix); free(tmp); free(src);
    rs_free(&st);
    free(cidx); free(ctypes);
    return out;
}

static VecBatch *ogroup_next(VecNode *self) {
    OGroup *g = (OGroup *)self;
    if (g->done) return NULL;
    g->done = 1;
    return g->sorted ? ogroup_sorted(g) : ogroup_hash(g);
}

static void ogroup_free(VecNode *self) {
    OGroup *g = (OGroup *)self;
    g->child->free_node(g->child);
    for (int k = 0; k < g->n_keys; k++) free(g->key_names[k]);
    free(g->key_names);
    for (int a = 0; a < g->n_aggs; a++) {
        free(g->specs[a].output_name);
        free(g->specs[a].input_col);
    }
    free(g->specs);
    vec_schema_free(&g->base.output_schema);
    free(g);
}

VecNode *oracle_group_agg_node_create(VecNode *child, int n_keys,
                                      char **key_names, int n_aggs,
                                      AggSpec *specs, const char *temp_dir) {
    OGroup *g = (OGroup *)calloc(1, sizeof *g);
    g->child = child;
    g->n_keys = n_keys;
    g->key_names = key_names;
    g->n_aggs = n_aggs;
    g->specs = specs;
    g->sorted = temp_dir != NULL && n_keys > 0;   /* src/group_agg.c:624-638 */
    int no = n_keys + n_aggs;
    char **names = (char **)malloc(sizeof(char *) * (size_t)(no + 1));
    VecType *types = (VecType *)malloc(sizeof(VecType) * (size_t)(no + 1));
    for (int k = 0; k < n_keys; k++) {
        int i = vec_schema_find_col(&child->output_schema, key_names[k]);
        names[k] = key_names[k];
        types[k] = i >= 0 ? child->output_schema.col_types[i] : VEC_DOUBLE;
    }
    for (int a = 0; a < n_aggs; a++) {
        names[n_keys + a] = specs[a].output_name;
        types[n_keys + a] = VEC_DOUBLE;
    }
    g->base.output_schema = vec_schema_create(no, names, types);
    free(names); free(types);
    g->base.next_batch = ogroup_next;
    g->base.free_node = ogroup_free;
    g->base.kind = "GroupAggNode";
    return &g->base;
}

/* ==================================================================== */
/* join (src/join.c)                                                     */
/* ==================================================================== */

typedef struct {
    VecNode  base;
    VecNode *left, *right;
    JoinKind kind;
    int      nk;
    JoinKey *keys;
    int      built;
    int      finalizing, finished;
    int      r_ncols;
    VecArray *r_cols;
    int64_t  n_build, n_slots;
    int64_t *head;
    uint64_t *slot_h;
    int64_t *next;
    uint8_t *build_matched;
    int64_t  fin_cursor;
    int     *r_nonkey;
    int      n_r_nonkey;
} OJoin;

static uint64_t join_hash(const VecArray *cols, const int *kidx, int nk, int64_t row) {
    uint64_t h = 0;
    for (int k = 0; k < nk; k++) {
        uint64_t kh = oracle_hash_value(&cols[kidx[k]], row);
        h = k == 0 ? kh : hash_combine(h, kh);
    }
    return h;
}

/* join_keys_equal (src/join.c:77-113): NA never matches; the switch is on
 * the PROBE column's type, the build column is read through that view. */
static int join_eq(const VecArray *pc, const int *pk, const VecArray *bc,
                   const int *bk, int nk, int64_t pr, int64_t br) {
    for (int k = 0; k < nk; k++) {
        const VecArray *p = &pc[pk[k]], *b = &bc[bk[k]];
        if (!vec_array_is_valid(p, pr) || !vec_array_is_valid(b, br)) return 0;
        switch (p->type) {
        case VEC_INT64:  if (p->buf.i64[pr] != b->buf.i64[br]) return 0; break;
        case VEC_DOUBLE: if (p->buf.dbl[pr] != b->buf.dbl[br]) return 0; break;
        case VEC_BOOL:   if (p->buf.bln[pr] != b->buf.bln[br]) return 0; break;
        case VEC_STRING: if (!group_val_equal(p, pr, b, br)) return 0; break;
        }
    }
    return 1;
}

static int *key_idx(const OJoin *j, int right) {
    int *ix = (int *)malloc(sizeof(int) * (size_t)(j->nk + 1));
    for (int k = 0; k < j->nk; k++) ix[k] = right ? j->keys[k].right_col : j->keys[k].left_col;
    return ix;
}

/* join_build (src/join.c:156-220) with jht_create/jht_insert (:23-61) */
static void ojoin_build(OJoin *j) {
    const VecSchema *rs = &j->right->output_schema;
    j->r_ncols = rs->n_cols;
    RowStore st;
    rs_init(&st, rs->n_cols, rs->col_types);
    const VecArray **src = (const VecArray **)malloc(sizeof(void *) * (size_t)(rs->n_cols + 1));
    VecBatch *b;
    while ((b = j->right->next_batch(j->right)) != NULL) {
        for (int c = 0; c < rs->n_cols; c++) src[c] = &b->columns[c];
        int64_t nl = vec_batch_logical_rows(b);
        for (int64_t li = 0; li < nl; li++) rs_append(&st, src, vec_batch_physical_row(b, li));
        vec_batch_free(b);
    }
    free(src);
    j->n_build = st.n;
    j->r_cols = st.cols;
    for (int c = 0; c < st.nc; c++) j->r_cols[c].length = st.n;
    free(st.types);
    /* coerce build keys to the common type */
    const VecSchema *ls = &j->left->output_schema;
    for (int k = 0; k < j->nk; k++) {
        VecType lt = ls->col_types[j->keys[k].left_col];
        VecArray *rc = &j->r_cols[j->keys[k].right_col];
        if (lt != rc->type) {
            VecType t = common_type(lt, rc->type);
            if (t != rc->type) {
                VecArray *cv = coerce(rc, t);
                vec_array_free(rc);
                *rc = *cv;
                free(cv);
            }
        }
    }
    int64_t nb = j->n_build > 0 ? j->n_build : 1;
    int64_t ns = 64;
    while (ns < nb * 2) ns *= 2;
    j->n_slots = ns;
    j->head = (int64_t *)malloc((size_t)ns * 8);
    j->slot_h = (uint64_t *)malloc((size_t)ns * 8);
    j->next = (int64_t *)malloc((size_t)nb * 8);
    memset(j->head, -1, (size_t)ns * 8);
    memset(j->next, -1, (size_t)nb * 8);
    int *rk = key_idx(j, 1);
    for (int64_t r = 0; r < j->n_build; r++) {
        uint64_t h = join_hash(j->r_cols, rk, j->nk, r);
        int64_t s = (int64_t)(h & (uint64_t)(ns - 1));
        for (;;) {
            if (j->head[s] < 0) { j->head[s] = r; j->slot_h[s] = h; break; }
            if (j->slot_h[s] == h) { j->next[r] = j->head[s]; j->head[s] = r; break; }
            s = (s + 1) & (ns - 1);
        }
    }
    free(rk);
    if (j->kind == JOIN_FULL)
        j->build_matched = (uint8_t *)calloc((size_t)((j->n_build + 7) / 8 + 1), 1);
}

static int64_t chain_from(const OJoin *j, int64_t br, const VecArray *pc, const int *pk,
                          const int *rk, int64_t pr) {
    while (br >= 0) {
        if (join_eq(pc, pk, j->r_cols, rk, j->nk, pr, br)) return br;
        br = j->next[br];
    }
    return -1;
}

/* jht_probe (src/join.c:115-135) */
static int64_t probe_first(const OJoin *j, uint64_t h, const VecArray *pc, const int *pk,
                           const int *rk, int64_t pr) {
    int64_t s = (int64_t)(h & (uint64_t)(j->n_slots - 1));
    for (;;) {
        if (j->head[s] < 0) return -1;
        if (j->slot_h[s] == h) {
            int64_t br = chain_from(j, j->head[s], pc, pk, rk, pr);
            if (br >= 0) return br;
        }
        s = (s + 1) & (j->n_slots - 1);
    }
}

/* output row list for one probe batch: pairs (probe phys row, build row or -1) */
typedef struct { int64_t *p, *b; int64_t n, cap; } PairList;

static void pl_push(PairList *l, int64_t p, int64_t b) {
    if (l->n == l->cap) {
        l->cap = l->cap ? l->cap * 2 : 1024;
        l->p = (int64_t *)realloc(l->p, (size_t)l->cap * 8);
        l->b = (int64_t *)realloc(l->b, (size_t)l->cap * 8);
    }
    l->p[l->n] = p;
    l->b[l->n] = b;
    l->n++;
}

static void put_value(VecArray *d, int64_t i, const VecArray *s, int64_t r) {
    if (r < 0 || !vec_array_is_valid(s, r)) return;
    vec_array_set_valid(d, i);
    /* the builder writes through its own type's view (src/builder.c:110-138) */
    if (d->type == VEC_BOOL) d->buf.bln[i] = s->type == VEC_BOOL ? s->buf.bln[r] : ((const uint8_t *)s->buf.i64)[r];
    else if (s->type == VEC_BOOL) d->buf.i64[i] = s->buf.bln[r];
    else d->buf.i64[i] = s->buf.i64[r];
}

/* join_probe_one (src/join.c:262-480) */
static VecBatch *ojoin_probe(OJoin *j, VecBatch *pb) {
    const VecSchema *ls = &j->left->output_schema;
    int lnc = ls->n_cols, onc = j->base.output_schema.n_cols;
    int64_t nl = vec_batch_logical_rows(pb);
    int *pk = key_idx(j, 0), *rk = key_idx(j, 1);

    /* probe keys coerced to the build key type for hashing/compare */
    VecArray *hash_cols = (VecArray *)malloc(sizeof(VecArray) * (size_t)(lnc + 1));
    memcpy(hash_cols, pb->columns, sizeof(VecArray) * (size_t)lnc);
    VecArray **owned = (VecArray **)calloc((size_t)(j->nk + 1), sizeof(VecArray *));
    for (int k = 0; k < j->nk; k++) {
        VecType pt = pb->columns[pk[k]].type, bt = j->r_cols[rk[k]].type;
        if (pt != bt) {
            owned[k] = coerce(&pb->columns[pk[k]], bt);
            hash_cols[pk[k]] = *owned[k];
        }
    }
    PairList out = {0};
    uint8_t *matched = (uint8_t *)calloc((size_t)(nl / 8 + 1), 1);
    for (int64_t li = 0; li < nl; li++) {
        int64_t pr = vec_batch_physical_row(pb, li);
        uint64_t h = join_hash(hash_cols, pk, j->nk, pr);
        int64_t br = probe_first(j, h, hash_cols, pk, rk, pr);
        switch (j->kind) {
        case JOIN_SEMI: if (br >= 0) pl_push(&out, pr, -1); break;
        case JOIN_ANTI: if (br < 0) pl_push(&out, pr, -1); break;
        case JOIN_INNER:
            while (br >= 0) {
                pl_push(&out, pr, br);
                br = chain_from(j, j->next[br], hash_cols, pk, rk, pr);
            }
            break;
        case JOIN_LEFT:
        case JOIN_FULL:
            if (br >= 0) matched[li >> 3] |= (uint8_t)(1u << (li & 7));
            while (br >= 0) {
                if (j->kind == JOIN_FULL) j->build_matched[br >> 3] |= (uint8_t)(1u << (br & 7));
                pl_push(&out, pr, br);
                /* the chain walk here compares against the UNCOERCED probe
                 * columns (src/join.c:398-400, 413-415) */
                br = chain_from(j, j->next[br], pb->columns, pk, rk, pr);
            }
            break;
        }
    }
    if (j->kind == JOIN_LEFT || j->kind == JOIN_FULL)
        for (int64_t li = 0; li < nl; li++)
            if (!(matched[li >> 3] & (1u << (li & 7))))
                pl_push(&out, vec_batch_physical_row(pb, li), -1);
    free(matched);
    for (int k = 0; k < j->nk; k++) drop_array(owned[k]);
    free(owned);
    free(hash_cols);
    free(pk); free(rk);

    if (out.n == 0) { free(out.p); free(out.b); return NULL; }
    VecBatch *res = vec_batch_alloc(onc, out.n);
    for (int c = 0; c < onc; c++) {
        VecArray d = vec_array_alloc(j->base.output_schema.col_types[c], out.n);
        const VecArray *s = c < lnc ? &pb->columns[c] : &j->r_cols[j->r_nonkey[c - lnc]];
        const int64_t *rows = c < lnc ? out.p : out.b;
        if (d.type == VEC_STRING) str_gather(&d, s, rows, out.n);
        else for (int64_t i = 0; i < out.n; i++) put_value(&d, i, s, rows[i]);
        res->columns[c] = d;
        res->col_names[c] = vx_strdup(j->base.output_schema.col_names[c]);
    }
    free(out.p); free(out.b);
    return res;
}

/* join_finalize (src/join.c:487-551): unmatched build rows, 65,536 per batch */
static VecBatch *ojoin_finalize(OJoin *j) {
    const VecSchema *ls = &j->left->output_schema;
    int lnc = ls->n_cols, onc = j->base.output_schema.n_cols;
    PairList out = {0};
    int64_t br = j->fin_cursor;
    for (; br < j->n_build && out.n < 65536; br++)
        if (!(j->build_matched[br >> 3] & (1u << (br & 7)))) pl_push(&out, -1, br);
    j->fin_cursor = br;
    if (out.n == 0) { free(out.p); free(out.b); return NULL; }
    VecBatch *res = vec_batch_alloc(onc, out.n);
    for (int c = 0; c < onc; c++) {
        VecArray d = vec_array_alloc(j->base.output_schema.col_types[c], out.n);
        const VecArray *s = NULL;
        if (c < lnc) {
            for (int k = 0; k < j->nk; k++)
                if (j->keys[k].left_col == c) { s = &j->r_cols[j->keys[k].right_col]; break; }
        } else {
            s = &j->r_cols[j->r_nonkey[c - lnc]];
        }
        if (s && d.type == VEC_STRING) str_gather(&d, s, out.b, out.n);
        else if (s) for (int64_t i = 0; i < out.n; i++) put_value(&d, i, s, out.b[i]);
        res->columns[c] = d;
        res->col_names[c] = vx_strdup(j->base.output_schema.col_names[c]);
    }
    free(out.p); free(out.b);
    return res;
}

static VecBatch *ojoin_next(VecNode *self) {
    OJoin *j = (OJoin *)self;
    if (!j->built) { ojoin_build(j); j->built = 1; }
    while (!j->finalizing && !j->finished) {
        VecBatch *pb = j->left->next_batch(j->left);
        if (!pb) {
            if (j->kind == JOIN_FULL) j->finalizing = 1;
            else j->finished = 1;
            break;
        }
        VecBatch *r = ojoin_probe(j, pb);
        vec_batch_free(pb);
        if (r) return r;
    }
    if (j->finalizing) {
        VecBatch *r = ojoin_finalize(j);
        if (r) return r;
        j->finalizing = 0;
        j->finished = 1;
    }
    return NULL;
}

static void ojoin_free(VecNode *self) {
    OJoin *j = (OJoin *)self;
    j->left->free_node(j->left);
    j->right->free_node(j->right);
    if (j->r_cols) {
        for (int c = 0; c < j->r_ncols; c++) vec_array_free(&j->r_cols[c]);
        free(j->r_cols);
    }
    free(j->head); free(j->slot_h); free(j->next); free(j->build_matched);
    free(j->keys); free(j->r_nonkey);
    vec_schema_free(&j->base.output_schema);
    free(j);
}

/* join_node_create (src/join.c:631-791): key type check, output schema
 * with suffixes on name collisions. */
VecNode *oracle_join_node_create(VecNode *left, VecNode *right, JoinKind kind,
                                 int n_keys, JoinKey *keys,
                                 const char *sx, const char *sy) {
    static const char *kn[] = {"inner_join", "left_join", "full_join", "semi_join", "anti_join"};
    const VecSchema *ls = &left->output_schema, *rs = &right->output_schema;
    for (int k = 0; k < n_keys; k++) {
        VecType lt = ls->col_types[keys[k].left_col], rt = rs->col_types[keys[k].right_col];
        if (lt != rt && (lt == VEC_STRING || rt == VEC_STRING))
            vectra_error("%s key type mismatch: x.%s (%s) vs y.%s (%s)", kn[kind],
                         ls->col_names[keys[k].left_col],
                         lt == VEC_INT64 ? "int64" : lt == VEC_DOUBLE ? "double" : lt == VEC_BOOL ? "bool" : "string",
                         rs->col_names[keys[k].right_col],
                         rt == VEC_INT64 ? "int64" : rt == VEC_DOUBLE ? "double" : rt == VEC_BOOL ? "bool" : "string");
    }
    OJoin *j = (OJoin *)calloc(1, sizeof *j);
    j->left = left; j->right = right; j->kind = kind; j->nk = n_keys; j->keys = keys;
    int *rkey = (int *)calloc((size_t)(rs->n_cols + 1), sizeof(int));
    for (int k = 0; k < n_keys; k++) rkey[keys[k].right_col] = 1;
    j->r_nonkey = (int *)malloc(sizeof(int) * (size_t)(rs->n_cols + 1));
    for (int c = 0; c < rs->n_cols; c++) if (!rkey[c]) j->r_nonkey[j->n_r_nonkey++] = c;
    int on = (kind == JOIN_SEMI || kind == JOIN_ANTI) ? ls->n_cols : ls->n_cols + j->n_r_nonkey;
    char **names = (char **)calloc((size_t)(on + 1), sizeof(char *));
    VecType *types = (VecType *)malloc(sizeof(VecType) * (size_t)(on + 1));
    for (int i = 0; i < ls->n_cols; i++) { names[i] = vx_strdup(ls->col_names[i]); types[i] = ls->col_types[i]; }
    if (on > ls->n_cols) {
        for (int q = 0; q < j->n_r_nonkey; q++) {
            int c = j->r_nonkey[q];
            const char *rn = rs->col_names[c];
            int hit = -1;
            for (int i = 0; i < ls->n_cols; i++) if (strcmp(ls->col_names[i], rn) == 0) { hit = i; break; }
            char *nm;
            if (hit >= 0) {
                size_t L = strlen(rn) + strlen(sy) + 1;
                nm = (char *)malloc(L);
                snprintf(nm, L, "%s%s", rn, sy);
                size_t L2 = strlen(ls->col_names[hit]) + strlen(sx) + 1;
                char *lx = (char *)malloc(L2);
                snprintf(lx, L2, "%s%s", ls->col_names[hit], sx);
                free(names[hit]);
                names[hit] = lx;
            } else {
                nm = vx_strdup(rn);
            }
            names[ls->n_cols + q] = nm;
            types[ls->n_cols + q] = rs->col_types[c];
        }
    }
    j->base.output_schema = vec_schema_create(on, names, types);
    for (int i = 0; i < on; i++) free(names[i]);
    free(names); free(types); free(rkey);
    j->base.next_batch = ojoin_next;
    j->base.free_node = ojoin_free;
    j->base.kind = "JoinNode";
    return &j->base;
}

/* ==================================================================== */
/* hidden mutate (src/project.c:51-252, src/scalar_ops.c:11-66)          */
/* ==================================================================== */

typedef struct {
    VecNode  base;
    VecNode *child;
    char    *a, *b;
} OProject;

static VecBatch *oproject_next(VecNode *self) {
    OProject *p = (OProject *)self;
    VecBatch *in = p->child->next_batch(p->child);
    if (!in) return NULL;
    /* gather the logical rows, then evaluate a * b over them */
    int64_t n = vec_batch_logical_rows(in);
    int nc = in->n_cols;
    VecBatch *out = vec_batch_alloc(nc + 1, n);
    for (int c = 0; c < nc; c++) {
        const VecArray *s = &in->columns[c];
        VecArray d = vec_array_alloc(s->type, n);
        for (int64_t i = 0; i < n; i++) put_value(&d, i, s, vec_batch_physical_row(in, i));
        out->columns[c] = d;
        out->col_names[c] = vx_strdup(in->col_names[c]);
    }
    out->n_cols = nc;
    VecArray *prod = arith(find_col(out, p->a), find_col(out, p->b), '*');
    out->n_cols = nc + 1;
    out->columns[nc] = *prod;
    free(prod);
    out->col_names[nc] = vx_strdup(p->base.output_schema.col_names[nc]);
    vec_batch_free(in);
    return out;
}

static void oproject_free(VecNode *self) {
    OProject *p = (OProject *)self;
    p->child->free_node(p->child);
    free(p->a);
    free(p->b);
    vec_schema_free(&p->base.output_schema);
    free(p);
}

VecNode *oracle_project_mul_node_create(VecNode *child, const char *out_name,
                                        const char *a_name, const char *b_name) {
    OProject *p = (OProject *)calloc(1, sizeof *p);
    const VecSchema *cs = &child->output_schema;
    int ia = resolve(cs, a_name, "mutate"), ib = resolve(cs, b_name, "mutate");
    int nc = cs->n_cols;
    char **names = (char **)malloc(sizeof(char *) * (size_t)(nc + 1));
    VecType *types = (VecType *)malloc(sizeof(VecType) * (size_t)(nc + 1));
    for (int c = 0; c < nc; c++) { names[c] = cs->col_names[c]; types[c] = cs->col_types[c]; }
    names[nc] = (char *)out_name;
    VecType t = common_type(cs->col_types[ia], cs->col_types[ib]);
    types[nc] = t == VEC_BOOL ? VEC_INT64 : t;
    p->base.output_schema = vec_schema_create(nc + 1, names, types);
    free(names);
    free(types);
    p->child = child;
    p->a = vx_strdup(a_name);
    p->b = vx_strdup(b_name);
    p->base.next_batch = oproject_next;
    p->base.free_node = oproject_free;
    p->base.kind = "ProjectNode";
    return &p->base;
}

/* ==================================================================== */
/* exact summation                                                        */
/* ==================================================================== */

/* A fixed-point accumulator wide enough for every finite double: bit k of
 * the 2240-bit two's-complement integer weighs 2^(k - 1126), so subnormals
 * (2^-1074) and the largest finite value (< 2^1024) both fit with 60 bits of
 * carry headroom.  The final value is rounded once, to nearest-even. */
#define XS_WORDS 35
typedef struct { uint64_t w[XS_WORDS]; } XSum;

static void xs_add_shifted(XSum *s, uint64_t mant, int shift, int neg) {
    /* add or subtract mant << shift */
    int wi = shift / 64, bi = shift % 64;
    uint64_t lo = mant << bi;
    uint64_t hi = bi ? mant >> (64 - bi) : 0;
    if (!neg) {
        unsigned __int128 c = 0;
        for (int i = wi; i < XS_WORDS; i++) {
            uint64_t add = i == wi ? lo : i == wi + 1 ? hi : 0;
            c += (unsigned __int128)s->w[i] + add;
            s->w[i] = (uint64_t)c;
            c >>= 64;
            if (!c && i > wi) break;
        }
    } else {
        uint64_t borrow = 0;
        for (int i = wi; i < XS_WORDS; i++) {
            uint64_t sub = i == wi ? lo : i == wi + 1 ? hi : 0;
            uint64_t x = s->w[i];
            uint64_t r = x - sub - borrow;
            borrow = (x < sub) || (x - sub < borrow);
            s->w[i] = r;
            if (!borrow && i > wi) break;
        }
    }
}

double oracle_exact_sum(const double *x, int64_t n) {
    XSum s;
    memset(&s, 0, sizeof s);
    double special = 0.0;
    int has_special = 0;
    for (int64_t i = 0; i < n; i++) {
        double v = x[i];
        if (!isfinite(v)) { special += v; has_special = 1; continue; }
        if (v == 0.0) continue;
        uint64_t bits;
        memcpy(&bits, &v, 8);
        int e = (int)((bits >> 52) & 0x7FF);
        uint64_t m = bits & ((1ULL << 52) - 1);
        if (e) m |= 1ULL << 52; else e = 1;
        /* value = m * 2^(e - 1075); bit position = e - 1075 + 1126 = e + 51 */
        xs_add_shifted(&s, m, e + 51, (bits >> 63) != 0);
    }
    if (has_special) return special;
    int neg = (s.w[XS_WORDS - 1] >> 63) != 0;
    if (neg) {   /* magnitude = two's complement negation */
        uint64_t carry = 1;
        for (int i = 0; i < XS_WORDS; i++) {
            uint64_t v = ~s.w[i] + carry;
            carry = carry && v == 0;
            s.w[i] = v;
        }
    }
    int top = -1;
    for (int i = XS_WORDS - 1; i >= 0; i--)
        if (s.w[i]) { top = i * 64 + 63 - __builtin_clzll(s.w[i]); break; }
    if (top < 0) return 0.0;
    /* take 53 bits below and including top, round to nearest even */
    int lsb = top - 52;
    if (lsb < 0) lsb = 0;
    uint64_t mant = 0;
    for (int b = top; b >= lsb; b--)
        mant = (mant << 1) | ((s.w[b / 64] >> (b % 64)) & 1);
    int round_up = 0;
    if (lsb > 0) {
        int half = (int)((s.w[(lsb - 1) / 64] >> ((lsb - 1) % 64)) & 1);
        int sticky = 0;
        for (int b = lsb - 2; b >= 0 && !sticky; b--)
            if ((s.w[b / 64] >> (b % 64)) & 1) sticky = 1;
        round_up = half && (sticky || (mant & 1));
    }
    double r = ldexp((double)mant, lsb - 1126);
    if (round_up) r = ldexp((double)(mant + 1), lsb - 1126);
    return neg ? -r : r;
}
