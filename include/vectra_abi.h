/*
 * vectra_abi.h -- the engine types that the GPU operator nodes exchange with
 * the rest of vectra, restated so that every struct is byte-for-byte
 * layout-compatible with the reference engine.
 *
 * A node built here can be handed to vectra's optimizer, explain() and
 * collect() unchanged: those walk nodes through the VecNode vtable and read
 * the struct prefixes of FilterNode / GroupAggNode / JoinNode after a strcmp
 * on `kind`.  Every field that code touches sits at the same offset here.
 *
 * Reference layouts restated (file:line of gcol33/vectra 0.2.2):
 *   VecType, VecArray, VecBatch, VecSchema, VecNode ... src/types.h:7-104
 *   validity bitmap (LSB-first, 1 = valid) ........... src/array.h:13-33
 *   VecExprKind, VecExpr ............................. src/expr.h:6-62
 *   AggKind ........................................... src/agg_ops.h:7-14
 *   AggSpec, GroupAggNode ............................. src/group_agg.h:7-23
 *   JoinKind, JoinKey, JoinNode prefix ................ src/join.h:6-61
 *   FilterNode ........................................ src/filter.h:7-11
 *   ProjEntry, ProjectNode ............................ src/project.h:7-18
 *   Vtr1Header, Vtr1ColStat, Vtr1RowGroup, Vtr1File ... src/vtr1.h:7-37
 *   ScanNode .......................................... src/scan.h:9-16
 *
 * Ownership rules are the reference's: every buffer reachable from a
 * VecBatch returned by next_batch() is malloc/calloc memory that the
 * consumer releases with free() (src/batch.c:19-29).  No pinned or device
 * allocation ever escapes a node.
 */
#ifndef VECTRA_ABI_H
#define VECTRA_ABI_H

#include <stdint.h>
#include <stddef.h>
#include <stdio.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- column types (src/types.h:7-12) -------------------------------- */
typedef enum {
    VEC_INT64  = 0,
    VEC_DOUBLE = 1,
    VEC_BOOL   = 2,
    VEC_STRING = 3
} VecType;

/* ---- one column (src/types.h:44-62) ---------------------------------
 * validity: bit i of byte i/8 set = row i valid; NULL = all rows valid. */
typedef struct {
    VecType   type;
    int64_t   length;
    uint8_t  *validity;
    uint8_t   owns_data;
    union {
        int64_t  *i64;
        double   *dbl;
        uint8_t  *bln;
        struct {
            int64_t *offsets;
            char    *data;
            int64_t  data_len;
        } str;
    } buf;
} VecArray;

/* ---- one batch (src/types.h:64-75) ----------------------------------
 * sel, when non-NULL, lists sel_n physical row indices in logical order. */
typedef struct {
    int64_t    n_rows;
    int        n_cols;
    char     **col_names;
    VecArray  *columns;
    int32_t   *sel;
    int32_t    sel_n;
} VecBatch;

static inline int64_t vec_batch_logical_rows(const VecBatch *b) {
    return b->sel ? (int64_t)b->sel_n : b->n_rows;
}

static inline int64_t vec_batch_physical_row(const VecBatch *b, int64_t li) {
    return b->sel ? (int64_t)b->sel[li] : li;
}

static inline int vec_array_is_valid(const VecArray *a, int64_t i) {
    if (!a->validity) return 1;
    return (a->validity[i >> 3] >> (i & 7)) & 1;
}

/* ---- schema (src/types.h:87-92) ------------------------------------- */
typedef struct {
    int       n_cols;
    char    **col_names;
    VecType  *col_types;
    char    **col_annotations;
} VecSchema;

/* ---- pull-based node vtable (src/types.h:95-104) -------------------- */
typedef struct VecNode VecNode;
typedef VecBatch *(*NextBatchFn)(VecNode *self);
typedef void      (*FreeFn)(VecNode *self);

struct VecNode {
    NextBatchFn   next_batch;
    FreeFn        free_node;
    VecSchema     output_schema;
    const char   *kind;
};

/* ---- expression tree (src/expr.h:6-62) ------------------------------ */
typedef enum {
    EXPR_COL_REF,
    EXPR_LIT_INT64,
    EXPR_LIT_DOUBLE,
    EXPR_LIT_BOOL,
    EXPR_LIT_STRING,
    EXPR_LIT_NA,
    EXPR_ARITH,
    EXPR_CMP,
    EXPR_BOOL,
    EXPR_IS_NA,
    EXPR_IF_ELSE,
    EXPR_CAST,
    EXPR_NEGATE,
    EXPR_NCHAR,
    EXPR_SUBSTR,
    EXPR_GREPL
} VecExprKind;

typedef struct VecExpr VecExpr;
struct VecExpr {
    VecExprKind kind;
    VecType     result_type;
    char       *col_name;      /* COL_REF */
    int64_t     lit_i64;       /* LIT_* */
    double      lit_dbl;
    uint8_t     lit_bln;
    char       *lit_str;
    char        op;            /* ARITH: + - * / %   CMP: < > = !   BOOL: & | ! */
    char        op2;           /* CMP second char: '=' or ' ' */
    VecExpr    *left;
    VecExpr    *right;
    VecExpr    *operand;       /* BOOL '!', IS_NA, NEGATE */
    VecExpr    *cond;          /* IF_ELSE (unimplemented in the engine) */
    VecExpr    *then_expr;
    VecExpr    *else_expr;
    VecType     cast_to;       /* CAST (unimplemented in the engine) */
};

/* ---- aggregation (src/agg_ops.h:7-14, src/group_agg.h:7-12) --------- */
typedef enum {
    AGG_COUNT,
    AGG_COUNT_STAR,
    AGG_SUM,
    AGG_MEAN,
    AGG_MIN,
    AGG_MAX
} AggKind;

typedef struct {
    char    *output_name;
    AggKind  kind;
    char    *input_col;    /* NULL for AGG_COUNT_STAR */
    int      na_rm;
} AggSpec;

/* ---- join (src/join.h:6-17) ----------------------------------------- */
typedef enum {
    JOIN_INNER,
    JOIN_LEFT,
    JOIN_FULL,
    JOIN_SEMI,
    JOIN_ANTI
} JoinKind;

typedef struct {
    int left_col;
    int right_col;
} JoinKey;

/* ---- the three hot-path nodes ---------------------------------------
 * Public prefixes match the reference; GPU state hangs off the trailing
 * opaque pointer, which nothing outside the node reads. */

/* src/filter.h:7-11 -- optimizer reads child/predicate
 * (src/optimize.c:102-125, 356-372), explain reads child
 * (src/r_bridge.c:532-535). */
typedef struct {
    VecNode   base;
    VecNode  *child;
    VecExpr  *predicate;
    void     *gpu;         /* GPU-private state */
} FilterNode;

/* src/group_agg.h:14-23 -- optimizer reads child/n_keys/key_names/
 * n_aggs/agg_specs (src/optimize.c:243-259), explain reads n_keys. */
typedef struct {
    VecNode     base;
    VecNode    *child;
    int         n_keys;
    char      **key_names;
    int         n_aggs;
    AggSpec    *agg_specs;
    int         done;
    int         use_sorted;   /* 1 = key-sorted output (R path, temp_dir set) */
    void       *gpu;          /* GPU-private state */
} GroupAggNode;

/* src/join.h:45-53 prefix -- optimizer and explain read left/right/kind/
 * n_keys (src/optimize.c:313-333, 396-400; src/r_bridge.c:552-556,
 * 638-650). */
typedef struct {
    VecNode   base;
    VecNode  *left;
    VecNode  *right;
    JoinKind  kind;
    int       n_keys;
    JoinKey  *keys;
    char     *suffix_x;
    char     *suffix_y;
    void     *gpu;           /* GPU-private state */
} JoinNode;

/* src/project.h:7-18 -- the reference's ProjectNode, which summarise()
 * inserts under a GroupAggNode for an aggregate of an expression such as
 * sum(value * weight) (R/verbs.R:266-284).  The GPU GroupAggNode reads
 * these fields, as the optimizer does (src/optimize.c:128-150), to fuse
 * GroupAgg <- Project(tmp = a * b) <- Join into one device pass. */
typedef struct {
    char    *output_name;
    VecExpr *expr;        /* NULL: pass the child column output_name through */
} ProjEntry;

typedef struct {
    VecNode    base;
    VecNode   *child;
    int        n_entries;
    ProjEntry *entries;
} ProjectNode;

/* ---- drop-in constructors (same signatures and ownership) ----------- */

/* src/filter.h:14.  Takes ownership of child and predicate. */
FilterNode *filter_node_create(VecNode *child, VecExpr *predicate);

/* src/group_agg.h:28-31.  Takes ownership of child, key_names (and its
 * strings) and agg_specs (and their strings); temp_dir is only tested for
 * NULL: non-NULL with n_keys > 0 selects key-sorted output order. */
GroupAggNode *group_agg_node_create(VecNode *child,
                                    int n_keys, char **key_names,
                                    int n_aggs, AggSpec *agg_specs,
                                    const char *temp_dir);

/* src/join.h:76-78.  Takes ownership of left, right and keys; copies the
 * suffixes. */
JoinNode *join_node_create(VecNode *left, VecNode *right,
                           JoinKind kind, int n_keys, JoinKey *keys,
                           const char *suffix_x, const char *suffix_y);

/* The plan text of a node tree, as explain() prints it (the node lines of
 * src/r_bridge.c:575-698, C_node_plan without the optimizer pass): one line
 * per node in pre-order, "<2 spaces per level><kind> [<annotation>]\n",
 * the reference's annotation for ScanNode / FilterNode / GroupAggNode /
 * JoinNode, and for GPU nodes "; gpu: ..." saying what the device does.
 * Writes at most size bytes (NUL-terminated when size > 0) and returns the
 * length of the whole text, as snprintf does (-1 for a NULL root). */
int64_t vx_explain(const VecNode *root, char *buf, int64_t size);

/* ---- .vtr files and the scan node ------------------------------------
 * The GPU nodes read a scan child's open file directly (row-group offsets,
 * schema, per-column statistics, the pushed-down predicate), so these
 * layouts are the reference's. */

/* src/vtr1.h:7-12 */
typedef struct {
    uint16_t  version;
    VecSchema schema;
    uint32_t  n_rowgroups;
} Vtr1Header;

/* src/vtr1.h:14-22 -- per column and row group (version 3 files) */
typedef struct {
    uint8_t has_stats;    /* 0: string column, or no valid value */
    union {
        struct { int64_t min, max; } i64;
        struct { double min, max; } dbl;
        struct { uint8_t min, max; } bln;   /* some valid FALSE / some valid TRUE */
    };
} Vtr1ColStat;

/* src/vtr1.h:24-29 */
typedef struct {
    int64_t      file_offset;   /* of the row group's n_rows field */
    int64_t      n_rows;
    Vtr1ColStat *col_stats;     /* n_cols entries; NULL before version 3 */
} Vtr1RowGroup;

/* src/vtr1.h:31-36 */
typedef struct {
    FILE         *fp;
    Vtr1Header    header;
    Vtr1RowGroup *rowgroups;
} Vtr1File;

/* src/scan.h:9-16 -- the optimizer sets predicate (borrowed from a
 * FilterNode above, pred_borrowed = 1) for row-group pruning */
typedef struct {
    VecNode    base;
    Vtr1File  *file;
    int       *col_mask;      /* file columns the scan reads */
    uint32_t   next_rg;
    VecExpr   *predicate;
    int        pred_borrowed;
} ScanNode;

/* src/vtr1.h:39-55, src/scan.h:18-21 */
Vtr1File *vtr1_open(const char *path);
VecBatch *vtr1_read_rowgroup(Vtr1File *file, uint32_t rg_idx, const int *col_mask);
void      vtr1_close(Vtr1File *file);
void      vtr1_write(const char *path, const VecBatch *batch);
void      vtr1_write_header(FILE *fp, const VecSchema *schema, uint32_t n_rowgroups);
void      vtr1_write_rowgroup(FILE *fp, const VecBatch *batch);
ScanNode *scan_node_create(const char *path, int *col_indices, int n_selected);

#ifdef __cplusplus
}
#endif

#endif /* VECTRA_ABI_H */
