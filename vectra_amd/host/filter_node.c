/*
 * filter_node.c -- drop-in GPU FilterNode (src/filter.h:7-14).
 *
 * Same constructor, struct prefix, vtable and output as the reference
 * node (src/filter.c:41-84): each child batch keeps its columns and gets a
 * malloc'd int32 selection vector holding, in logical order, the physical
 * rows whose predicate is valid and TRUE (composing with an existing sel);
 * batches left empty are freed and skipped.  The predicate is compiled per
 * batch to a device program and evaluated by vgpu_filter; only the columns
 * it reads are copied to the device.
 */
#include "gpu_common.h"

#include <stdlib.h>
#include <string.h>

typedef struct {
    DevCol cols[VGPU_MAX_COLS];
    DevMem sel_in;
    DevMem sel_out;
    DevMem sstage[6];   /* string comparisons: operands staged on the device */
} FilterGpu;

static VecBatch *gpu_filter_next(VecNode *self) {
    FilterNode *fn = (FilterNode *)self;
    FilterGpu *g = (FilterGpu *)fn->gpu;
    VecBatch *b;
    while ((b = fn->child->next_batch(fn->child)) != NULL) {
        CompiledExpr cx;
        vx_compile_expr(fn->predicate, b, &cx);
        if (cx.result_type != VEC_BOOL) {
            vec_batch_free(b);
            vectra_error("filter predicate must evaluate to boolean");
        }
        int64_t n_logical = vec_batch_logical_rows(b);
        int32_t *sel = (int32_t *)malloc((size_t)(n_logical > 0 ? n_logical : 1) * sizeof(int32_t));
        if (!sel) vectra_error("alloc failed for filter sel");
        int64_t n_out = 0;
        if (n_logical > 0) {
            vgpu_col cols[VGPU_MAX_COLS];
            for (int s = 0; s < cx.n_cols; s++) {
                if (cx.col_idx[s] < 0) {   /* a string comparison's -1 / 0 / 1 column */
                    cols[s] = vx_string_cmp_col(&cx, -1 - cx.col_idx[s], b, &g->cols[s], g->sstage);
                    continue;
                }
                devcol_upload(&g->cols[s], &b->columns[cx.col_idx[s]], b->n_rows);
                cols[s] = g->cols[s].col;
            }
            const int32_t *d_sel = NULL;
            if (b->sel) {
                d_sel = (const int32_t *)devmem_reserve(&g->sel_in, (size_t)b->sel_n * 4);
                VX_GPU(vgpu_memcpy_h2d((void *)d_sel, b->sel, (size_t)b->sel_n * 4));
            }
            int32_t *d_out = (int32_t *)devmem_reserve(&g->sel_out, (size_t)n_logical * 4);
            VX_GPU(vgpu_filter(&cx.prog, cols, cx.n_cols, b->n_rows, d_sel, b->sel ? b->sel_n : 0, d_out, &n_out));
            if (n_out > 0) VX_GPU(vgpu_memcpy_d2h(sel, d_out, (size_t)n_out * 4));
        }
        free(b->sel);
        b->sel = sel;
        b->sel_n = (int32_t)n_out;
        if (n_out > 0) return b;
        vec_batch_free(b);
    }
    return NULL;
}

static void gpu_filter_free(VecNode *self) {
    FilterNode *fn = (FilterNode *)self;
    FilterGpu *g = (FilterGpu *)fn->gpu;
    fn->child->free_node(fn->child);
    vec_expr_free(fn->predicate);
    if (g) {
        for (int i = 0; i < VGPU_MAX_COLS; i++) devcol_free(&g->cols[i]);
        devmem_free(&g->sel_in);
        devmem_free(&g->sel_out);
        for (int i = 0; i < 6; i++) devmem_free(&g->sstage[i]);
        free(g);
    }
    vec_schema_free(&fn->base.output_schema);
    free(fn);
}

FilterNode *filter_node_create(VecNode *child, VecExpr *predicate) {
    vx_gpu_init();
    FilterNode *fn = (FilterNode *)calloc(1, sizeof *fn);
    if (!fn) vectra_error("alloc failed for FilterNode");
    fn->gpu = calloc(1, sizeof(FilterGpu));
    if (!fn->gpu) vectra_error("alloc failed for FilterNode");
    fn->child = child;
    fn->predicate = predicate;
    fn->base.output_schema = vec_schema_copy(&child->output_schema);
    fn->base.next_batch = gpu_filter_next;
    fn->base.free_node = gpu_filter_free;
    fn->base.kind = "FilterNode";
    return fn;
}

int vx_gpu_filter_note(const VecNode *n, char *buf, int size) {
    if (n->next_batch != gpu_filter_next) return -1;
    return snprintf(buf, (size_t)size, "gpu: predicate on the device, selection vector");
}
