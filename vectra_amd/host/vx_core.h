/*
 * vx_core.h -- engine plumbing the GPU nodes need around them when they run
 * outside the R package: array/batch/schema helpers, the error channel, a
 * row-group source node and the collect() pull loop.
 *
 * Inside vectra these symbols come from the engine itself
 * (src/array.c, src/batch.c, src/schema.c, src/expr.c:9-26, src/error.c,
 * src/collect.c:170-206); this file supplies the same names and semantics
 * so the nodes link and behave identically in a standalone harness.
 */
#ifndef VX_CORE_H
#define VX_CORE_H

#include <setjmp.h>
#include "vectra_abi.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- errors --------------------------------------------------------
 * vectra_error() formats into 1024 bytes and never returns, like the
 * reference (src/error.c:5-12, which longjmps through Rf_error).  Here it
 * longjmps to the innermost vx_try frame; with no frame it aborts. */
typedef struct vx_frame {
    jmp_buf           jb;
    struct vx_frame  *prev;
} vx_frame;

void        vx_frame_push(vx_frame *f);
void        vx_frame_pop(vx_frame *f);
/* usage: vx_frame_push(&f); if (setjmp(f.jb) == 0) { ...; vx_frame_pop(&f); } */

void        vectra_error(const char *fmt, ...) __attribute__((noreturn, format(printf, 1, 2)));
void        vectra_warning(const char *fmt, ...) __attribute__((format(printf, 1, 2)));
const char *vx_last_error(void);
const char *vx_last_warning(void);
int64_t     vx_warning_count(void);
void        vx_clear_warnings(void);

/* ---- arrays / batches / schemas (src/array.c, batch.c, schema.c) ---- */
VecArray  vec_array_alloc(VecType type, int64_t length);
void      vec_array_free(VecArray *arr);
void      vec_array_set_all_valid(VecArray *arr);
static inline int64_t vec_validity_bytes(int64_t n) { return (n + 7) / 8; }
static inline void vec_array_set_valid(VecArray *a, int64_t i) {
    if (a->validity) a->validity[i >> 3] |= (uint8_t)(1u << (i & 7));
}
static inline void vec_array_set_null(VecArray *a, int64_t i) {
    if (a->validity) a->validity[i >> 3] &= (uint8_t)~(1u << (i & 7));
}

VecBatch *vec_batch_alloc(int n_cols, int64_t n_rows);
void      vec_batch_free(VecBatch *batch);

VecSchema vec_schema_create(int n_cols, char **col_names, VecType *col_types);
void      vec_schema_free(VecSchema *schema);
VecSchema vec_schema_copy(const VecSchema *src);
int       vec_schema_find_col(const VecSchema *schema, const char *name);

VecExpr  *vec_expr_alloc(VecExprKind kind);
void      vec_expr_free(VecExpr *expr);

char     *vx_strdup(const char *s);

/* ---- row-group source ----------------------------------------------
 * Emits the borrowed columns as consecutive batches of rg_rows rows (the
 * last one ragged), each a deep copy the consumer frees -- the shape the
 * .vtr scan produces (src/scan.c:107-127).  rg_rows <= 0 means one batch. */
VecNode *vx_mem_scan_create(int n_cols, const char *const *names,
                            const VecArray *cols, int64_t n_rows,
                            int64_t rg_rows);

/* ---- collect --------------------------------------------------------
 * Pulls root to exhaustion and concatenates the logical rows of every
 * batch into one dense batch (no sel), as vec_collect does before it
 * converts to R (src/collect.c:170-206).  Frees nothing of root. */
VecBatch *vx_collect(VecNode *root);

/* ctypes-friendly wrappers that turn vectra_error into a return code. */
int  vx_try_collect(VecNode *root, VecBatch **out);
int  vx_try_next(VecNode *node, VecBatch **out);
void vx_node_free(VecNode *node);

/* the same for the three constructors, called through a function pointer
 * with the constructor's own signature */
typedef VecNode *(*vx_filter_ctor)(VecNode *, VecExpr *);
typedef VecNode *(*vx_group_ctor)(VecNode *, int, char **, int, AggSpec *, const char *);
typedef VecNode *(*vx_join_ctor)(VecNode *, VecNode *, JoinKind, int, JoinKey *, const char *, const char *);
int vx_try_filter(vx_filter_ctor fn, VecNode *child, VecExpr *pred, VecNode **out);
int vx_try_group(vx_group_ctor fn, VecNode *child, int n_keys, char **key_names, int n_aggs,
                 AggSpec *specs, const char *temp_dir, VecNode **out);
int vx_try_join(vx_join_ctor fn, VecNode *left, VecNode *right, JoinKind kind, int n_keys,
                JoinKey *keys, const char *sx, const char *sy, VecNode **out);

/* ---- ProjectNode ----------------------------------------------------
 * The reference's project_node_create (src/project.h:20-23), for trees the
 * harness builds where R is absent: summarise(s = sum(a * b)) inserts
 * Project(child columns..., .vectra_tmp_s = a * b) under the GroupAggNode
 * (R/verbs.R:266-284).  Takes ownership of entries, their names and
 * expressions; evaluates column references and + - * / % over columns and
 * literals (other expressions raise). */
ProjectNode *project_node_create(VecNode *child, int n_entries, ProjEntry *entries);
int vx_try_project(VecNode *child, int n_entries, ProjEntry *entries, VecNode **out);

/* ---- .vtr harness entry points -------------------------------------- */
/* the optimizer's pushdown of a filter's predicate onto the .vtr scan
 * below it (src/optimize.c:353-369), along a chain of filters */
void vx_pushdown_predicates(VecNode *node);
/* scan_node_create / a multi-row-group .vtr writer, returning -1 with the
 * message in vx_last_error() instead of raising */
int vx_try_scan(const char *path, int *col_indices, int n_selected, VecNode **out);
int vx_try_vtr_write(const char *path, int n_cols, const char *const *names, const VecArray *cols, int64_t n_rows,
                     int64_t rg_rows);

#ifdef __cplusplus
}
#endif
#endif
