/*
 * gpu_common.h -- helpers shared by the three GPU operator nodes: device
 * copies of host columns, the VecExpr -> vgpu_prog compiler, and the
 * mapping of libvgpu failures onto vectra_error().
 */
#ifndef VX_GPU_COMMON_H
#define VX_GPU_COMMON_H

#include "vectra_abi.h"
#include "vgpu.h"
#include "vx_core.h"

#ifdef __cplusplus
extern "C" {
#endif

/* raises vectra_error with libvgpu's message when status < 0 */
#define VX_GPU(call)                                                   \
    do {                                                               \
        if ((call) < 0) vectra_error("GPU: %s", vgpu_last_error());   \
    } while (0)

void vx_gpu_init(void);

/* A device copy of one host column; buffers grow and are reused. */
typedef struct {
    vgpu_col col;
    void    *d_data;
    size_t   cap_data;
    uint8_t *d_valid;
    size_t   cap_valid;
} DevCol;

/* uploads a[0..n) (int64/double/bool); validity only if some row is NA */
void devcol_upload(DevCol *dc, const VecArray *a, int64_t n);
/* device buffer of at least `bytes` inside dc (for kernel outputs) */
void devcol_reserve(DevCol *dc, size_t data_bytes, size_t valid_bytes);
void devcol_free(DevCol *dc);

/* generic growable device buffer */
typedef struct { void *ptr; size_t cap; } DevMem;
void *devmem_reserve(DevMem *m, size_t bytes);
void  devmem_free(DevMem *m);

int vx_all_valid(const uint8_t *validity, int64_t n);
VecType vx_common_type(VecType a, VecType b);   /* src/coerce.c:7-19 */
const char *vx_type_name(VecType t);            /* src/types.h:14-22 */

/* VecExpr -> straight-line program over the batch's columns.  Column slots
 * index col_idx[] (batch column indices); *result_type gets the
 * expression's type.  Raises vectra_error with the reference's messages
 * for type errors and with an explicit message for shapes the GPU path
 * does not evaluate (string operands). */
/* a comparison of two strings (src/scalar_ops.c:77-106), evaluated before
 * the program into an int64 column of -1 / 0 / 1 that the program compares
 * against 0; each side is a batch column (col >= 0), a literal (lit) or
 * the NA string (na) */
#define VX_MAX_SCMP 4
typedef struct {
    int         col[2];
    const char *lit[2];
    int         na[2];
} VxStrCmp;

typedef struct {
    vgpu_prog prog;
    int       n_cols;
    int       col_idx[VGPU_MAX_COLS];   /* batch column, or -1 - k: string comparison k */
    VecType   result_type;
    int       n_scmp;
    VxStrCmp  scmp[VX_MAX_SCMP];
} CompiledExpr;

void vx_compile_expr(const VecExpr *e, const VecBatch *batch, CompiledExpr *out);

/* the column of string comparison k of cx over batch b on the device
 * (n_rows entries; buffers in dc, each operand's offsets, bytes and
 * validity staged in stage[0..2] and stage[3..5]); ordered on the library
 * stream before the work that reads it */
vgpu_col vx_string_cmp_col(const CompiledExpr *cx, int k, const VecBatch *b, DevCol *dc, DevMem stage[6]);
/* the same on the host: out (b->n_rows) and a validity bitmap */
void vx_string_cmp_host(const CompiledExpr *cx, int k, const VecBatch *b, int64_t *out, uint8_t *valid);
/* the same, returning -1 (message in vx_last_error) instead of raising */
int vx_try_compile_expr(const VecExpr *e, const VecBatch *batch, CompiledExpr *out);

/* concatenates the logical rows of batches (consumed) into dense columns
 * of the given schema; returns the row count */
int64_t vx_concat_batches(VecNode *child, VecArray **cols_out);

/* ---- a GPU JoinNode read by the GPU GroupAggNode above it ------------
 * GroupAgg <- Project(tmp = a * b) <- Join (what summarise(sum(a * b))
 * builds over a join, R/verbs.R:266-284) runs as one fused probe ->
 * aggregate pass: the GroupAgg node has the JoinNode build its table (with
 * the build columns it aggregates embedded) and then feeds the join's left
 * child itself.  If it finds the built table unsuited (see group_agg_node.c)
 * it pulls batches from the Project as usual and the JoinNode probes its
 * already built table. */
typedef struct {
    vgpu_jtable  *jt;
    VecNode      *left;            /* the join's probe child, not yet pulled */
    int           kind;            /* JOIN_INNER or JOIN_LEFT */
    int           left_key, right_key;
    VecType       build_key_type;
    const DevCol *r_dev;           /* right columns on the device, key coerced */
    int           r_ncols;
    int64_t       max_dup;         /* most build rows of one key */
    int           coerced;         /* the probe key is cast to build_key_type */
} VxJoinFusion;

/* explain() notes of the GPU nodes: -1 when n is not such a node */
int vx_gpu_join_note(const VecNode *n, char *buf, int size);

/* 1 for a GPU JoinNode that has not started, one non-string key, inner or left */
int vx_gpu_join_fusable(const VecNode *n);
/* join output column -> (0, left column) or (1, right column); -1 if none */
int vx_gpu_join_output_source(const VecNode *n, int out_col, int *side, int *col);
/* builds the table (right columns payload_cols embedded, at most four) */
void vx_gpu_join_prepare_fusion(VecNode *n, const int *payload_cols, int n_payload, VxJoinFusion *out);

#ifdef __cplusplus
}
#endif
#endif
