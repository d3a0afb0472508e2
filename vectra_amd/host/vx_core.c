/*
 * vx_core.c -- standalone engine plumbing (see vx_core.h).
 *
 * Semantics follow the reference helpers they stand in for:
 *   vec_array_alloc/free ........ src/array.c:6-57 (validity calloc'd = all NA)
 *   vec_batch_alloc/free ........ src/batch.c:7-29
 *   vec_schema_* ................ src/schema.c:6-58
 *   vec_expr_alloc/free ......... src/expr.c:9-26
 *   vectra_error ................ src/error.c:5-12
 *   collect pull loop ........... src/collect.c:170-206
 */
#define _GNU_SOURCE
#include "vx_core.h"

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ */
/* errors                                                             */
/* ------------------------------------------------------------------ */

static __thread vx_frame *g_top;
static __thread char g_err[1024];
static __thread char g_warn[1024];
static __thread int64_t g_nwarn;

void vx_frame_push(vx_frame *f) { f->prev = g_top; g_top = f; }
void vx_frame_pop(vx_frame *f)  { g_top = f->prev; }

void vectra_error(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    vx_frame *f = g_top;
    if (!f) {
        fprintf(stderr, "vectra error (no handler): %s\n", g_err);
        abort();
    }
    g_top = f->prev;
    longjmp(f->jb, 1);
}

void vectra_warning(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_warn, sizeof g_warn, fmt, ap);
    va_end(ap);
    g_nwarn++;
}

const char *vx_last_error(void)   { return g_err; }
const char *vx_last_warning(void) { return g_warn; }
int64_t     vx_warning_count(void) { return g_nwarn; }
void        vx_clear_warnings(void) { g_nwarn = 0; g_warn[0] = 0; }

char *vx_strdup(const char *s) {
    if (!s) return NULL;
    size_t n = strlen(s) + 1;
    char *d = (char *)malloc(n);
    if (!d) vectra_error("alloc failed for string");
    memcpy(d, s, n);
    return d;
}

/* ------------------------------------------------------------------ */
/* arrays                                                             */
/* ------------------------------------------------------------------ */

VecArray vec_array_alloc(VecType type, int64_t length) {
    VecArray a;
    memset(&a, 0, sizeof a);
    a.type = type;
    a.length = length;
    a.owns_data = 1;
    size_t vb = (size_t)vec_validity_bytes(length);
    a.validity = (uint8_t *)calloc(vb ? vb : 1, 1);
    if (!a.validity) vectra_error("failed to allocate validity bitmap");
    size_t n = (size_t)(length > 0 ? length : 1);
    switch (type) {
    case VEC_INT64:  a.buf.i64 = (int64_t *)calloc(n, 8); break;
    case VEC_DOUBLE: a.buf.dbl = (double *)calloc(n, 8); break;
    case VEC_BOOL:   a.buf.bln = (uint8_t *)calloc(n, 1); break;
    case VEC_STRING:
        a.buf.str.offsets = (int64_t *)calloc(n + 1, 8);
        break;
    }
    if (!a.buf.i64) vectra_error("failed to allocate %s array", "column");
    return a;
}

void vec_array_free(VecArray *a) {
    if (!a) return;
    free(a->validity);
    a->validity = NULL;
    if (a->type == VEC_STRING) {
        free(a->buf.str.offsets);
        if (a->owns_data) free(a->buf.str.data);
        a->buf.str.offsets = NULL;
        a->buf.str.data = NULL;
    } else {
        free(a->buf.i64);
        a->buf.i64 = NULL;
    }
    a->length = 0;
}

void vec_array_set_all_valid(VecArray *a) {
    if (a->validity)
        memset(a->validity, 0xFF, (size_t)vec_validity_bytes(a->length));
}

/* ------------------------------------------------------------------ */
/* batches                                                            */
/* ------------------------------------------------------------------ */

VecBatch *vec_batch_alloc(int n_cols, int64_t n_rows) {
    VecBatch *b = (VecBatch *)calloc(1, sizeof *b);
    if (!b) vectra_error("failed to allocate VecBatch");
    b->n_cols = n_cols;
    b->n_rows = n_rows;
    b->columns = (VecArray *)calloc((size_t)(n_cols > 0 ? n_cols : 1), sizeof(VecArray));
    b->col_names = (char **)calloc((size_t)(n_cols > 0 ? n_cols : 1), sizeof(char *));
    if (!b->columns || !b->col_names) vectra_error("failed to allocate batch columns");
    return b;
}

void vec_batch_free(VecBatch *b) {
    if (!b) return;
    for (int i = 0; i < b->n_cols; i++) {
        vec_array_free(&b->columns[i]);
        free(b->col_names[i]);
    }
    free(b->columns);
    free(b->col_names);
    free(b->sel);
    free(b);
}

/* ------------------------------------------------------------------ */
/* schemas                                                            */
/* ------------------------------------------------------------------ */

VecSchema vec_schema_create(int n_cols, char **col_names, VecType *col_types) {
    VecSchema s;
    size_t n = (size_t)(n_cols > 0 ? n_cols : 1);
    s.n_cols = n_cols;
    s.col_names = (char **)calloc(n, sizeof(char *));
    s.col_types = (VecType *)calloc(n, sizeof(VecType));
    s.col_annotations = (char **)calloc(n, sizeof(char *));
    if (!s.col_names || !s.col_types || !s.col_annotations)
        vectra_error("failed to allocate schema");
    for (int i = 0; i < n_cols; i++) {
        s.col_names[i] = vx_strdup(col_names[i]);
        s.col_types[i] = col_types[i];
    }
    return s;
}

void vec_schema_free(VecSchema *s) {
    if (!s) return;
    for (int i = 0; i < s->n_cols; i++) {
        free(s->col_names[i]);
        if (s->col_annotations) free(s->col_annotations[i]);
    }
    free(s->col_names);
    free(s->col_types);
    free(s->col_annotations);
    memset(s, 0, sizeof *s);
}

VecSchema vec_schema_copy(const VecSchema *src) {
    VecSchema s = vec_schema_create(src->n_cols, src->col_names, src->col_types);
    for (int i = 0; i < src->n_cols; i++)
        if (src->col_annotations && src->col_annotations[i])
            s.col_annotations[i] = vx_strdup(src->col_annotations[i]);
    return s;
}

int vec_schema_find_col(const VecSchema *s, const char *name) {
    for (int i = 0; i < s->n_cols; i++)
        if (strcmp(s->col_names[i], name) == 0) return i;
    return -1;
}

/* ------------------------------------------------------------------ */
/* expressions                                                        */
/* ------------------------------------------------------------------ */

VecExpr *vec_expr_alloc(VecExprKind kind) {
    VecExpr *e = (VecExpr *)calloc(1, sizeof *e);
    if (!e) vectra_error("alloc failed for VecExpr");
    e->kind = kind;
    return e;
}

void vec_expr_free(VecExpr *e) {
    if (!e) return;
    free(e->col_name);
    free(e->lit_str);
    vec_expr_free(e->left);
    vec_expr_free(e->right);
    vec_expr_free(e->operand);
    vec_expr_free(e->cond);
    vec_expr_free(e->then_expr);
    vec_expr_free(e->else_expr);
    free(e);
}

/* ------------------------------------------------------------------ */
/* dense slice copy (used by the source node and by collect)          */
/* ------------------------------------------------------------------ */

/* Copies rows [phys(0) .. phys(n-1)] of src into a new dense array,
 * where phys(i) = sel ? sel[i] : start + i. */
static VecArray copy_rows(const VecArray *src, int64_t start,
                          const int32_t *sel, int64_t n) {
    VecArray d = vec_array_alloc(src->type, n);
    int64_t str_total = 0;
    if (src->type == VEC_STRING) {
        for (int64_t i = 0; i < n; i++) {
            int64_t p = sel ? sel[i] : start + i;
            if (vec_array_is_valid(src, p))
                str_total += src->buf.str.offsets[p + 1] - src->buf.str.offsets[p];
        }
        d.buf.str.data = (char *)malloc((size_t)(str_total > 0 ? str_total : 1));
        d.buf.str.data_len = str_total;
    }
    int64_t off = 0;
    for (int64_t i = 0; i < n; i++) {
        int64_t p = sel ? sel[i] : start + i;
        int v = vec_array_is_valid(src, p);
        if (v) vec_array_set_valid(&d, i);
        switch (src->type) {
        case VEC_INT64:  if (v) d.buf.i64[i] = src->buf.i64[p]; break;
        case VEC_DOUBLE: if (v) d.buf.dbl[i] = src->buf.dbl[p]; break;
        case VEC_BOOL:   if (v) d.buf.bln[i] = src->buf.bln[p]; break;
        case VEC_STRING: {
            d.buf.str.offsets[i] = off;
            if (v) {
                int64_t s = src->buf.str.offsets[p];
                int64_t l = src->buf.str.offsets[p + 1] - s;
                memcpy(d.buf.str.data + off, src->buf.str.data + s, (size_t)l);
                off += l;
            }
            break;
        }
        }
    }
    if (src->type == VEC_STRING) d.buf.str.offsets[n] = off;
    return d;
}

/* ------------------------------------------------------------------ */
/* row-group source node                                              */
/* ------------------------------------------------------------------ */

typedef struct {
    VecNode         base;
    int             n_cols;
    const VecArray *cols;
    int64_t         n_rows;
    int64_t         rg_rows;
    int64_t         cursor;
    int             emitted;
} MemScanNode;

static VecBatch *mem_scan_next(VecNode *self) {
    MemScanNode *m = (MemScanNode *)self;
    if (m->cursor >= m->n_rows && (m->emitted || m->n_rows > 0)) return NULL;
    int64_t n = m->n_rows - m->cursor;
    if (m->rg_rows > 0 && n > m->rg_rows) n = m->rg_rows;
    VecBatch *b = vec_batch_alloc(m->n_cols, n);
    for (int c = 0; c < m->n_cols; c++) {
        b->columns[c] = copy_rows(&m->cols[c], m->cursor, NULL, n);
        b->col_names[c] = vx_strdup(m->base.output_schema.col_names[c]);
    }
    m->cursor += n;
    m->emitted = 1;
    if (n == 0) { vec_batch_free(b); return NULL; }
    return b;
}

static void mem_scan_free(VecNode *self) {
    MemScanNode *m = (MemScanNode *)self;
    vec_schema_free(&m->base.output_schema);
    free(m);
}

VecNode *vx_mem_scan_create(int n_cols, const char *const *names,
                            const VecArray *cols, int64_t n_rows,
                            int64_t rg_rows) {
    MemScanNode *m = (MemScanNode *)calloc(1, sizeof *m);
    if (!m) vectra_error("alloc failed for MemScanNode");
    VecType *types = (VecType *)malloc(sizeof(VecType) * (size_t)(n_cols > 0 ? n_cols : 1));
    for (int c = 0; c < n_cols; c++) types[c] = cols[c].type;
    m->base.output_schema = vec_schema_create(n_cols, (char **)names, types);
    free(types);
    m->n_cols = n_cols;
    m->cols = cols;
    m->n_rows = n_rows;
    m->rg_rows = rg_rows;
    m->base.next_batch = mem_scan_next;
    m->base.free_node = mem_scan_free;
    /* not "ScanNode": code that meets that kind reads the .vtr scan's
     * struct (file handle, pushed-down predicate) */
    m->base.kind = "MemScanNode";
    return &m->base;
}

/* ------------------------------------------------------------------ */
/* collect                                                            */
/* ------------------------------------------------------------------ */

VecBatch *vx_collect(VecNode *root) {
    const VecSchema *s = &root->output_schema;
    int nc = s->n_cols;
    int64_t cap = 0, total = 0, nb = 0;
    VecBatch **parts = NULL;
    VecBatch *b;
    while ((b = root->next_batch(root)) != NULL) {
        if (nb == cap) {
            cap = cap ? cap * 2 : 16;
            parts = (VecBatch **)realloc(parts, (size_t)cap * sizeof *parts);
        }
        parts[nb++] = b;
        total += vec_batch_logical_rows(b);
    }
    VecBatch *out = vec_batch_alloc(nc, total);
    for (int c = 0; c < nc; c++) {
        out->col_names[c] = vx_strdup(s->col_names[c]);
        VecType t = s->col_types[c];
        VecArray d = vec_array_alloc(t, total);
        int64_t row = 0, str_total = 0;
        if (t == VEC_STRING) {
            for (int64_t i = 0; i < nb; i++) {
                VecArray *src = &parts[i]->columns[c];
                int64_t n = vec_batch_logical_rows(parts[i]);
                for (int64_t li = 0; li < n; li++) {
                    int64_t p = vec_batch_physical_row(parts[i], li);
                    if (vec_array_is_valid(src, p))
                        str_total += src->buf.str.offsets[p + 1] - src->buf.str.offsets[p];
                }
            }
            d.buf.str.data = (char *)malloc((size_t)(str_total > 0 ? str_total : 1));
            d.buf.str.data_len = str_total;
        }
        int64_t off = 0;
        for (int64_t i = 0; i < nb; i++) {
            VecBatch *pb = parts[i];
            VecArray *src = &pb->columns[c];
            int64_t n = vec_batch_logical_rows(pb);
            for (int64_t li = 0; li < n; li++, row++) {
                int64_t p = vec_batch_physical_row(pb, li);
                int v = vec_array_is_valid(src, p);
                if (v) vec_array_set_valid(&d, row);
                switch (t) {
                case VEC_INT64:  if (v) d.buf.i64[row] = src->buf.i64[p]; break;
                case VEC_DOUBLE: if (v) d.buf.dbl[row] = src->buf.dbl[p]; break;
                case VEC_BOOL:   if (v) d.buf.bln[row] = src->buf.bln[p]; break;
                case VEC_STRING:
                    d.buf.str.offsets[row] = off;
                    if (v) {
                        int64_t st = src->buf.str.offsets[p];
                        int64_t l = src->buf.str.offsets[p + 1] - st;
                        memcpy(d.buf.str.data + off, src->buf.str.data + st, (size_t)l);
                        off += l;
                    }
                    break;
                }
            }
        }
        if (t == VEC_STRING) d.buf.str.offsets[total] = off;
        out->columns[c] = d;
    }
    for (int64_t i = 0; i < nb; i++) vec_batch_free(parts[i]);
    free(parts);
    return out;
}

int vx_try_collect(VecNode *root, VecBatch **out) {
    vx_frame f;
    *out = NULL;
    vx_frame_push(&f);
    if (setjmp(f.jb) == 0) {
        *out = vx_collect(root);
        vx_frame_pop(&f);
        return 0;
    }
    return -1;
}

int vx_try_next(VecNode *node, VecBatch **out) {
    vx_frame f;
    *out = NULL;
    vx_frame_push(&f);
    if (setjmp(f.jb) == 0) {
        *out = node->next_batch(node);
        vx_frame_pop(&f);
        return 0;
    }
    return -1;
}

void vx_node_free(VecNode *node) {
    if (node) node->free_node(node);
}

int vx_try_filter(vx_filter_ctor fn, VecNode *child, VecExpr *pred, VecNode **out) {
    vx_frame f;
    *out = NULL;
    vx_frame_push(&f);
    if (setjmp(f.jb) == 0) {
        *out = fn(child, pred);
        vx_frame_pop(&f);
        return 0;
    }
    return -1;
}

int vx_try_group(vx_group_ctor fn, VecNode *child, int n_keys, char **key_names, int n_aggs,
                 AggSpec *specs, const char *temp_dir, VecNode **out) {
    vx_frame f;
    *out = NULL;
    vx_frame_push(&f);
    if (setjmp(f.jb) == 0) {
        *out = fn(child, n_keys, key_names, n_aggs, specs, temp_dir);
        vx_frame_pop(&f);
        return 0;
    }
    return -1;
}

int vx_try_join(vx_join_ctor fn, VecNode *left, VecNode *right, JoinKind kind, int n_keys,
                JoinKey *keys, const char *sx, const char *sy, VecNode **out) {
    vx_frame f;
    *out = NULL;
    vx_frame_push(&f);
    if (setjmp(f.jb) == 0) {
        *out = fn(left, right, kind, n_keys, keys, sx, sy);
        vx_frame_pop(&f);
        return 0;
    }
    return -1;
}

/* ==================================================================== */
/* ProjectNode (the reference's src/project.c, for the harness)          */
/* ==================================================================== */

/* Stands in for the reference's own ProjectNode where R is absent, with the
 * reference's struct layout (src/project.h:7-18) and output schema (entry
 * names; an expression entry takes its expression's type, which the
 * reference's parser stores in expr->result_type, a pass-through entry its
 * child column's type; src/project.c:287-320).  Each batch comes out
 * dense (the child's selection gathered), pass-through columns copied and
 * expression entries evaluated over the gathered rows.  The expressions
 * evaluated are the ones summarise() inserts for nested aggregate inputs:
 * column references, and + - * / % over columns and literals with vectra's
 * arithmetic (src/scalar_ops.c:11-66: NA if either side is NA, bool
 * promoted to int64, int64 wrapping, integer / and % by zero NA, double
 * IEEE and fmod); anything else is an error. */

typedef struct {
    VecType  type;
    int64_t  n;
    const VecArray *col;    /* column operand (physical rows), or NULL */
    const int32_t  *sel;    /* gather map for col, or NULL */
    int64_t  li;            /* literal (int64 / bool) */
    double   ld;            /* literal (double) */
    int      lit_na;
} PjOperand;

static const VecArray *pj_find(const VecBatch *b, const char *name) {
    for (int c = 0; c < b->n_cols; c++)
        if (strcmp(b->col_names[c], name) == 0) return &b->columns[c];
    vectra_error("column not found: %s", name);
}

static VecArray pj_gather(const VecArray *a, const VecBatch *b) {
    int64_t n = vec_batch_logical_rows(b);
    VecArray o = vec_array_alloc(a->type, n);
    for (int64_t i = 0; i < n; i++) {
        int64_t r = vec_batch_physical_row(b, i);
        if (!vec_array_is_valid(a, r)) continue;
        vec_array_set_valid(&o, i);
        switch (a->type) {
        case VEC_INT64: o.buf.i64[i] = a->buf.i64[r]; break;
        case VEC_DOUBLE: o.buf.dbl[i] = a->buf.dbl[r]; break;
        case VEC_BOOL: o.buf.bln[i] = a->buf.bln[r]; break;
        case VEC_STRING: break;
        }
    }
    if (a->type == VEC_STRING) {
        int64_t total = 0;
        for (int64_t i = 0; i < n; i++) {
            int64_t r = vec_batch_physical_row(b, i);
            total += a->buf.str.offsets[r + 1] - a->buf.str.offsets[r];
        }
        free(o.buf.str.data);
        o.buf.str.data = (char *)malloc((size_t)(total > 0 ? total : 1));
        o.buf.str.data_len = total;
        int64_t at = 0;
        for (int64_t i = 0; i < n; i++) {
            int64_t r = vec_batch_physical_row(b, i);
            int64_t len = a->buf.str.offsets[r + 1] - a->buf.str.offsets[r];
            o.buf.str.offsets[i] = at;
            memcpy(o.buf.str.data + at, a->buf.str.data + a->buf.str.offsets[r], (size_t)len);
            at += len;
        }
        o.buf.str.offsets[n] = at;
    }
    return o;
}

static PjOperand pj_operand(const VecExpr *e, const VecBatch *b) {
    PjOperand o;
    memset(&o, 0, sizeof o);
    o.n = vec_batch_logical_rows(b);
    switch (e->kind) {
    case EXPR_COL_REF:
        o.col = pj_find(b, e->col_name);
        o.type = o.col->type;
        if (o.type == VEC_STRING) vectra_error("cannot coerce string in arithmetic/comparison");
        break;
    case EXPR_LIT_INT64: o.type = VEC_INT64; o.li = e->lit_i64; break;
    case EXPR_LIT_DOUBLE: o.type = VEC_DOUBLE; o.ld = e->lit_dbl; break;
    case EXPR_LIT_BOOL: o.type = VEC_BOOL; o.li = e->lit_bln; break;
    case EXPR_LIT_NA: o.type = VEC_DOUBLE; o.lit_na = 1; break;
    default:
        vectra_error("project: the harness evaluates column references and arithmetic over columns and literals "
                     "(expression kind %d)", (int)e->kind);
    }
    return o;
}

static int pj_valid(const PjOperand *o, const VecBatch *b, int64_t i) {
    if (o->col) return vec_array_is_valid(o->col, vec_batch_physical_row(b, i));
    return !o->lit_na;
}

static int64_t pj_i64(const PjOperand *o, const VecBatch *b, int64_t i) {
    if (!o->col) return o->li;
    int64_t r = vec_batch_physical_row(b, i);
    return o->type == VEC_BOOL ? (int64_t)o->col->buf.bln[r] : o->col->buf.i64[r];
}

static double pj_dbl(const PjOperand *o, const VecBatch *b, int64_t i) {
    if (!o->col) return o->type == VEC_DOUBLE ? o->ld : (double)o->li;
    int64_t r = vec_batch_physical_row(b, i);
    switch (o->type) {
    case VEC_DOUBLE: return o->col->buf.dbl[r];
    case VEC_INT64: return (double)o->col->buf.i64[r];
    default: return (double)o->col->buf.bln[r];
    }
}

static VecArray pj_eval(const VecExpr *e, const VecBatch *b) {
    if (e->kind == EXPR_COL_REF) return pj_gather(pj_find(b, e->col_name), b);
    if (e->kind != EXPR_ARITH)
        vectra_error("project: the harness evaluates column references and arithmetic over columns and literals "
                     "(expression kind %d)", (int)e->kind);
    PjOperand l = pj_operand(e->left, b), r = pj_operand(e->right, b);
    VecType t = (l.type == VEC_DOUBLE || r.type == VEC_DOUBLE) ? VEC_DOUBLE : VEC_INT64;
    int64_t n = vec_batch_logical_rows(b);
    VecArray o = vec_array_alloc(t, n);
    for (int64_t i = 0; i < n; i++) {
        if (!pj_valid(&l, b, i) || !pj_valid(&r, b, i)) continue;
        vec_array_set_valid(&o, i);
        if (t == VEC_INT64) {
            uint64_t x = (uint64_t)pj_i64(&l, b, i), y = (uint64_t)pj_i64(&r, b, i);
            int64_t sx = (int64_t)x, sy = (int64_t)y;
            switch (e->op) {
            case '+': o.buf.i64[i] = (int64_t)(x + y); break;
            case '-': o.buf.i64[i] = (int64_t)(x - y); break;
            case '*': o.buf.i64[i] = (int64_t)(x * y); break;
            case '/':
                if (sy == 0) vec_array_set_null(&o, i);
                else o.buf.i64[i] = sy == -1 ? (int64_t)(0 - x) : sx / sy;
                break;
            case '%':
                if (sy == 0) vec_array_set_null(&o, i);
                else o.buf.i64[i] = sy == -1 ? 0 : sx % sy;
                break;
            default: vec_array_free(&o); vectra_error("unknown arith op: %c", e->op);
            }
        } else {
            double x = pj_dbl(&l, b, i), y = pj_dbl(&r, b, i);
            switch (e->op) {
            case '+': o.buf.dbl[i] = x + y; break;
            case '-': o.buf.dbl[i] = x - y; break;
            case '*': o.buf.dbl[i] = x * y; break;
            case '/': o.buf.dbl[i] = x / y; break;
            case '%': o.buf.dbl[i] = fmod(x, y); break;
            default: vec_array_free(&o); vectra_error("unknown arith op: %c", e->op);
            }
        }
    }
    return o;
}

static VecBatch *pj_next(VecNode *self) {
    ProjectNode *pn = (ProjectNode *)self;
    VecBatch *in = pn->child->next_batch(pn->child);
    if (!in) return NULL;
    int64_t n = vec_batch_logical_rows(in);
    VecBatch *out = vec_batch_alloc(pn->n_entries, n);
    for (int i = 0; i < pn->n_entries; i++) {
        const ProjEntry *pe = &pn->entries[i];
        out->columns[i] = pe->expr ? pj_eval(pe->expr, in) : pj_gather(pj_find(in, pe->output_name), in);
        out->col_names[i] = vx_strdup(pe->output_name);
    }
    vec_batch_free(in);
    return out;
}

static void pj_free(VecNode *self) {
    ProjectNode *pn = (ProjectNode *)self;
    pn->child->free_node(pn->child);
    for (int i = 0; i < pn->n_entries; i++) {
        free(pn->entries[i].output_name);
        if (pn->entries[i].expr) vec_expr_free(pn->entries[i].expr);
    }
    free(pn->entries);
    vec_schema_free(&pn->base.output_schema);
    free(pn);
}

/* the type an entry's expression evaluates to over the child schema (the
 * reference's parser records it in expr->result_type, src/r_bridge.c:918-935) */
static VecType pj_type(const VecExpr *e, const VecSchema *cs) {
    switch (e->kind) {
    case EXPR_COL_REF: {
        int i = vec_schema_find_col(cs, e->col_name);
        if (i < 0) vectra_error("column not found: %s", e->col_name);
        return cs->col_types[i];
    }
    case EXPR_LIT_INT64: return VEC_INT64;
    case EXPR_LIT_BOOL: return VEC_BOOL;
    case EXPR_ARITH: {
        VecType l = pj_type(e->left, cs), r = pj_type(e->right, cs);
        return (l == VEC_DOUBLE || r == VEC_DOUBLE) ? VEC_DOUBLE : VEC_INT64;
    }
    default: return VEC_DOUBLE;
    }
}

ProjectNode *project_node_create(VecNode *child, int n_entries, ProjEntry *entries) {
    ProjectNode *pn = (ProjectNode *)calloc(1, sizeof *pn);
    if (!pn) vectra_error("alloc failed");
    pn->child = child;
    pn->n_entries = n_entries;
    pn->entries = entries;
    char **names = (char **)malloc(sizeof(char *) * (size_t)(n_entries > 0 ? n_entries : 1));
    VecType *types = (VecType *)malloc(sizeof(VecType) * (size_t)(n_entries > 0 ? n_entries : 1));
    for (int i = 0; i < n_entries; i++) {
        names[i] = entries[i].output_name;
        if (entries[i].expr) {
            types[i] = pj_type(entries[i].expr, &child->output_schema);
            entries[i].expr->result_type = types[i];
        } else {
            int idx = vec_schema_find_col(&child->output_schema, entries[i].output_name);
            if (idx < 0) {
                free(names);
                free(types);
                vectra_error("select: column not found: %s", entries[i].output_name);
            }
            types[i] = child->output_schema.col_types[idx];
        }
    }
    pn->base.output_schema = vec_schema_create(n_entries, names, types);
    free(names);
    free(types);
    pn->base.next_batch = pj_next;
    pn->base.free_node = pj_free;
    pn->base.kind = "ProjectNode";
    return pn;
}

int vx_try_project(VecNode *child, int n_entries, ProjEntry *entries, VecNode **out) {
    vx_frame f;
    *out = NULL;
    vx_frame_push(&f);
    if (setjmp(f.jb) == 0) {
        *out = (VecNode *)project_node_create(child, n_entries, entries);
        vx_frame_pop(&f);
        return 0;
    }
    return -1;
}
