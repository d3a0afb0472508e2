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
