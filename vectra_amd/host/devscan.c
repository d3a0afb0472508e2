/*
 * devscan.c -- device scan feed over a .vtr scan node (see devscan.h).
 *
 * Per row group (layout: src/vtr1.c:39-135): the fed columns' validity
 * bitmaps and values are read with fread straight into one of two pinned
 * staging slots, walking the group's columns in file order and seeking
 * past the others; the values, and the bitmap of a column that has NAs in
 * the group, are copied to the slot's device buffers on the copy stream.
 * The library stream waits for those copies before the caller's kernels;
 * a slot's staging is refilled only after its copies are done and its
 * device buffers only after the kernels that read them.
 */
#define _GNU_SOURCE
#include "devscan.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static int64_t g_groups_fed;

int64_t vx_scan_feed_groups(void) { return g_groups_fed; }

static size_t type_width(VecType t) { return t == VEC_BOOL ? 1 : 8; }

static uint8_t *slot_validity(VxDevScan *d, int s, int j) {
    return d->pinned[s] + (size_t)j * (d->vb_cap + (size_t)d->max_rows * 8);
}

void vx_devscan_open(VxDevScan *d, ScanNode *scan, const int *fcols, int n) {
    memset(d, 0, sizeof *d);
    d->scan = scan;
    d->cur = -1;
    const Vtr1File *f = scan->file;
    const VecSchema *fs = &f->header.schema;
    if (n > VX_DS_COLS) vectra_error("GPU scan feed: %d columns exceed %d", n, VX_DS_COLS);
    d->n = n;
    for (int j = 0; j < n; j++) {
        const int c = fcols[j];
        if (c < 0 || c >= fs->n_cols) vectra_error("GPU scan feed: column index %d out of range", c);
        for (int k = 0; k < j; k++)
            if (fcols[k] == c) vectra_error("GPU scan feed: column %d named twice", c);
        if (fs->col_types[c] == VEC_STRING)
            vectra_error("GPU: string columns are not evaluated on the GPU");
        d->fcol[j] = c;
        d->type[j] = fs->col_types[c];
    }
    for (uint32_t g = 0; g < f->header.n_rowgroups; g++)
        if (f->rowgroups[g].n_rows > d->max_rows) d->max_rows = f->rowgroups[g].n_rows;
    if (d->max_rows == 0 || n == 0) return;
    d->vb_cap = ((size_t)vec_validity_bytes(d->max_rows) + 7) & ~(size_t)7;
    const size_t per_col = d->vb_cap + (size_t)d->max_rows * 8;
    for (int s = 0; s < 2; s++) {
        d->pinned[s] = (uint8_t *)vgpu_host_alloc(per_col * (size_t)n);
        if (!d->pinned[s]) vectra_error("GPU: %s", vgpu_last_error());
        for (int j = 0; j < n; j++) {
            d->d_data[s][j] = vgpu_malloc((size_t)d->max_rows * type_width(d->type[j]));
            d->d_valid[s][j] = (uint8_t *)vgpu_malloc(d->vb_cap + 8);
            if (!d->d_data[s][j] || !d->d_valid[s][j]) vectra_error("GPU: %s", vgpu_last_error());
        }
    }
}

/* reads row group g's fed columns into staging slot s */
static void read_group(VxDevScan *d, int s, uint32_t g) {
    Vtr1File *f = d->scan->file;
    const VecSchema *fs = &f->header.schema;
    const int64_t n = f->rowgroups[g].n_rows;
    const int64_t vb = vec_validity_bytes(n);
    int64_t off = f->rowgroups[g].file_offset + 8 + (f->header.version >= 3 ? 17 * (int64_t)fs->n_cols : 0);
    if (fseeko(f->fp, (off_t)off, SEEK_SET) != 0) vectra_error("unexpected end of file");
    for (int c = 0; c < fs->n_cols; c++) {
        int j = -1;
        for (int k = 0; k < d->n; k++)
            if (d->fcol[k] == c) j = k;
        const VecType t = fs->col_types[c];
        if (j < 0) {
            /* a column the consumer does not read */
            if (t == VEC_STRING) {
                if (fseeko(f->fp, (off_t)(vb + 8 * (n + 1)), SEEK_CUR) != 0) vectra_error("unexpected end of file");
                uint8_t lb[8];
                if (fread(lb, 1, 8, f->fp) != 8) vectra_error("unexpected end of file");
                uint64_t len = 0;
                for (int i = 7; i >= 0; i--) len = (len << 8) | lb[i];
                if (fseeko(f->fp, (off_t)len, SEEK_CUR) != 0) vectra_error("unexpected end of file");
            } else {
                const int64_t w = (int64_t)type_width(t);
                if (fseeko(f->fp, (off_t)(vb + w * n), SEEK_CUR) != 0) vectra_error("unexpected end of file");
            }
            continue;
        }
        uint8_t *v = slot_validity(d, s, j);
        if (vb && fread(v, 1, (size_t)vb, f->fp) != (size_t)vb)
            vectra_error("unexpected end of file reading validity bitmap");
        const size_t bytes = (size_t)n * type_width(t);
        if (bytes && fread(v + d->vb_cap, 1, bytes, f->fp) != bytes)
            vectra_error("unexpected end of file reading %s data", vx_type_name(t));
    }
}

int vx_devscan_next(VxDevScan *d, vgpu_col *cols, int64_t *n_rows, int64_t *pos_base) {
    Vtr1File *f = d->scan->file;
    /* the caller has enqueued its work on the previous group by now */
    if (d->cur >= 0) VX_GPU(vgpu_upload_release(d->cur));
    uint32_t g = d->next_rg;
    for (; g < f->header.n_rowgroups; g++) {
        const Vtr1RowGroup *rg = &f->rowgroups[g];
        if (d->scan->predicate && f->header.version >= 3 &&
            !vx_vtr_might_match(d->scan->predicate, rg->col_stats, &f->header.schema)) {
            d->groups_pruned++;
            continue;
        }
        if (rg->n_rows > 0) break;
    }
    d->next_rg = g + 1;
    if (g >= f->header.n_rowgroups) return 0;
    const int64_t n = f->rowgroups[g].n_rows;
    const int s = d->cur < 0 ? 0 : 1 - d->cur;
    VX_GPU(vgpu_upload_host_wait(s));   /* slot s's staging is no longer being copied */
    read_group(d, s, g);
    for (int j = 0; j < d->n; j++) {
        const uint8_t *v = slot_validity(d, s, j);
        cols[j].type = (int32_t)d->type[j];
        cols[j].reserved = 0;
        cols[j].data = d->d_data[s][j];
        cols[j].validity = NULL;
        VX_GPU(vgpu_upload_async(s, d->d_data[s][j], v + d->vb_cap, (size_t)n * type_width(d->type[j])));
        if (!vx_all_valid(v, n)) {
            VX_GPU(vgpu_upload_async(s, d->d_valid[s][j], v, (size_t)vec_validity_bytes(n)));
            cols[j].validity = d->d_valid[s][j];
        }
    }
    VX_GPU(vgpu_upload_acquire(s));
    d->cur = s;
    *n_rows = n;
    *pos_base = d->rows_fed;
    d->rows_fed += n;
    d->groups_fed++;
    g_groups_fed++;
    return 1;
}

void vx_devscan_close(VxDevScan *d) {
    /* nothing may still copy into or read these buffers */
    vgpu_synchronize();
    for (int s = 0; s < 2; s++) {
        vgpu_upload_host_wait(s);
        for (int j = 0; j < d->n; j++) {
            if (d->d_data[s][j]) vgpu_free(d->d_data[s][j]);
            if (d->d_valid[s][j]) vgpu_free(d->d_valid[s][j]);
        }
        if (d->pinned[s]) vgpu_host_free(d->pinned[s]);
    }
    memset(d, 0, sizeof *d);
    d->cur = -1;
}
