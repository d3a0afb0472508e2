/*
 * devscan.c -- device scan feed over a .vtr scan node (see devscan.h).
 *
 * Per row group (layout: src/vtr1.c:39-135): the fed columns' validity
 * bitmaps and values are read with pread, by several threads, straight
 * into one of two pinned staging slots, the other columns skipped by their
 * sizes; the values, and the bitmap of a column that has NAs in the group,
 * are copied to the slot's device buffers on the copy stream.
 * The library stream waits for those copies before the caller's kernels;
 * a slot's staging is refilled only after its copies are done and its
 * device buffers only after the kernels that read them.
 */
#define _GNU_SOURCE
#include "devscan.h"
#include "vx_zonemap.h"

#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

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

/* One contiguous piece of the file for the staging buffer. */
typedef struct {
    uint8_t *dst;
    int64_t  off;
    size_t   len;
} Span;

typedef struct {
    int         fd;
    const Span *spans;
    int         n_spans, first, stride;
    int         failed;
} ReadJob;

static void *read_spans(void *arg) {
    ReadJob *j = (ReadJob *)arg;
    for (int i = j->first; i < j->n_spans; i += j->stride) {
        const Span *sp = &j->spans[i];
        size_t done = 0;
        while (done < sp->len) {
            const ssize_t got = pread(j->fd, sp->dst + done, sp->len - done, (off_t)(sp->off + (int64_t)done));
            if (got <= 0) {
                j->failed = 1;
                return NULL;
            }
            done += (size_t)got;
        }
    }
    return NULL;
}

static uint64_t le64(const uint8_t *b) {
    uint64_t v = 0;
    for (int i = 7; i >= 0; i--) v = (v << 8) | b[i];
    return v;
}

/* Reads row group g's fed columns into staging slot s: every fed column's
 * validity bitmap and its values in pieces of up to DS_PIECE bytes, read
 * with pread by up to DS_THREADS threads: one thread copying from the
 * page cache (about 14 GB/s) is slower than the copy stream. */
#define DS_THREADS 8
#define DS_PIECE ((size_t)4 << 20)
#define DS_MAX_SPANS 2048

static void read_group(VxDevScan *d, int s, uint32_t g) {
    Vtr1File *f = d->scan->file;
    const VecSchema *fs = &f->header.schema;
    const int fd = fileno(f->fp);
    const int64_t n = f->rowgroups[g].n_rows;
    const int64_t vb = vec_validity_bytes(n);
    int64_t off = f->rowgroups[g].file_offset + 8 + (f->header.version >= 3 ? 17 * (int64_t)fs->n_cols : 0);
    Span spans[DS_MAX_SPANS];
    int ns = 0;
    for (int c = 0; c < fs->n_cols; c++) {
        const VecType t = fs->col_types[c];
        int j = -1;
        for (int k = 0; k < d->n; k++)
            if (d->fcol[k] == c) j = k;
        if (t == VEC_STRING) {
            /* never fed: skip it by its data length */
            uint8_t lb[8];
            if (pread(fd, lb, 8, (off_t)(off + vb + 8 * (n + 1))) != 8) vectra_error("unexpected end of file");
            off += vb + 8 * (n + 1) + 8 + (int64_t)le64(lb);
            continue;
        }
        const int64_t w = (int64_t)type_width(t);
        if (j >= 0) {
            uint8_t *v = slot_validity(d, s, j);
            if (vb) spans[ns++] = (Span){v, off, (size_t)vb};
            for (size_t p = 0; p < (size_t)(w * n); p += DS_PIECE) {
                if (ns >= DS_MAX_SPANS) vectra_error("GPU scan feed: row group %u too large", g);
                const size_t len = (size_t)(w * n) - p < DS_PIECE ? (size_t)(w * n) - p : DS_PIECE;
                spans[ns++] = (Span){v + d->vb_cap + p, off + vb + (int64_t)p, len};
            }
        }
        off += vb + w * n;
    }
    const int nt = ns < DS_THREADS ? ns : DS_THREADS;
    ReadJob jobs[DS_THREADS];
    pthread_t th[DS_THREADS];
    int started = 0;
    for (int t = 0; t < nt; t++) {
        jobs[t] = (ReadJob){fd, spans, ns, t, nt, 0};
        if (t > 0 && pthread_create(&th[t], NULL, read_spans, &jobs[t]) == 0) {
            started |= 1 << t;
        } else if (t > 0) {
            read_spans(&jobs[t]);   /* no thread: read its pieces here */
        }
    }
    if (nt > 0) read_spans(&jobs[0]);
    int failed = 0;
    for (int t = 0; t < nt; t++) {
        if (started & (1 << t)) pthread_join(th[t], NULL);
        failed |= jobs[t].failed;
    }
    if (failed) vectra_error("unexpected end of file reading row group %u", g);
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
