/*
 * vx_vtr.c -- the .vtr columnar file format and the scan node over it, for
 * the standalone harness (inside vectra they come from src/vtr1.c and
 * src/scan.c; the GPU nodes only rely on the layouts declared in
 * vectra_abi.h, which both implementations share).
 *
 * File layout (all integers little-endian; src/vtr1.c:39-135 writes it,
 * :164-300 reads it):
 *
 *   "VTR1"  u16 version (1..3)  u16 n_cols
 *   n_cols x { u16 name_len, name bytes, u8 type,
 *              [v2+] u16 annotation_len, annotation bytes }
 *   u32 n_rowgroups
 *   n_rowgroups x row group:
 *     u64 n_rows
 *     [v3] n_cols x { u8 has_stats, u64 min, u64 max }   (17 bytes each)
 *     n_cols x column: validity bitmap (ceil(n_rows / 8) bytes), then
 *       int64 / double: 8 x n_rows bytes; bool: n_rows bytes;
 *       string: (n_rows + 1) x i64 offsets, u64 data_len, data_len bytes
 *
 * Statistics (v3, src/vtr1.c:66-113): int64 / double columns keep the
 * smallest and largest valid value as the scan meets them with `<` / `>`
 * (so a double column whose first valid value is NaN records NaN for both),
 * has_stats = 0 and zero placeholders when no value is valid, when the row
 * group is empty and for strings; bool columns always have has_stats = 1
 * with min = "some valid FALSE", max = "some valid TRUE".  The writer always
 * writes version 3.
 *
 * Zone-map pruning (src/scan.c:10-117): a row group is skipped when its
 * statistics show that a `col op literal` comparison (either side; AND
 * needs both sides, OR either) cannot hold for any row, comparing as
 * doubles; bool and string columns and other expressions never prune.
 */
#define _GNU_SOURCE
#include "vx_core.h"
#include "vx_zonemap.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* ---- little-endian field I/O ------------------------------------------ */

static void put_le(FILE *fp, uint64_t v, int bytes) {
    uint8_t b[8];
    for (int i = 0; i < bytes; i++) b[i] = (uint8_t)(v >> (8 * i));
    fwrite(b, 1, (size_t)bytes, fp);
}

static uint64_t get_le(FILE *fp, int bytes) {
    uint8_t b[8];
    if (fread(b, 1, (size_t)bytes, fp) != (size_t)bytes) vectra_error("unexpected end of file");
    uint64_t v = 0;
    for (int i = bytes - 1; i >= 0; i--) v = (v << 8) | b[i];
    return v;
}

static void read_exact(FILE *fp, void *dst, size_t n, const char *what) {
    if (n && fread(dst, 1, n, fp) != n) vectra_error("unexpected end of file reading %s", what);
}

static void skip_bytes(FILE *fp, int64_t n) {
    if (n && fseeko(fp, (off_t)n, SEEK_CUR) != 0) vectra_error("unexpected end of file");
}

/* bytes of one fixed-width column's values */
static int64_t value_bytes(VecType t, int64_t n) { return t == VEC_BOOL ? n : 8 * n; }

/* ---- writing ------------------------------------------------------------ */

void vtr1_write_header(FILE *fp, const VecSchema *schema, uint32_t n_rowgroups) {
    fwrite("VTR1", 1, 4, fp);
    put_le(fp, 3, 2);
    put_le(fp, (uint64_t)schema->n_cols, 2);
    for (int c = 0; c < schema->n_cols; c++) {
        const char *name = schema->col_names[c];
        const char *ann = schema->col_annotations ? schema->col_annotations[c] : NULL;
        size_t nl = strlen(name), al = ann ? strlen(ann) : 0;
        put_le(fp, nl, 2);
        fwrite(name, 1, nl, fp);
        put_le(fp, (uint64_t)schema->col_types[c], 1);
        put_le(fp, al, 2);
        if (al) fwrite(ann, 1, al, fp);
    }
    put_le(fp, n_rowgroups, 4);
}

/* statistics of one column of a row group: has_stats, min bits, max bits */
static void column_stats(const VecArray *a, int64_t n, uint8_t *has, uint64_t *lo, uint64_t *hi) {
    *has = 0;
    *lo = *hi = 0;
    if (a->type == VEC_STRING || n == 0) return;
    if (a->type == VEC_BOOL) {
        uint64_t seen_f = 0, seen_t = 0;
        for (int64_t i = 0; i < n; i++) {
            if (!vec_array_is_valid(a, i)) continue;
            if (a->buf.bln[i]) seen_t = 1;
            else seen_f = 1;
        }
        *has = 1;
        *lo = seen_f;
        *hi = seen_t;
        return;
    }
    int64_t first = -1;
    for (int64_t i = 0; i < n && first < 0; i++)
        if (vec_array_is_valid(a, i)) first = i;
    if (first < 0) return;
    *has = 1;
    if (a->type == VEC_INT64) {
        int64_t mn = a->buf.i64[first], mx = mn;
        for (int64_t i = first + 1; i < n; i++) {
            if (!vec_array_is_valid(a, i)) continue;
            const int64_t v = a->buf.i64[i];
            mn = v < mn ? v : mn;
            mx = v > mx ? v : mx;
        }
        *lo = (uint64_t)mn;
        *hi = (uint64_t)mx;
    } else {
        /* `v < mn` / `v > mx` as the scan meets the values: a NaN never
         * replaces a bound, and a NaN first value is never replaced */
        double mn = a->buf.dbl[first], mx = mn;
        for (int64_t i = first + 1; i < n; i++) {
            if (!vec_array_is_valid(a, i)) continue;
            const double v = a->buf.dbl[i];
            if (v < mn) mn = v;
            if (v > mx) mx = v;
        }
        memcpy(lo, &mn, 8);
        memcpy(hi, &mx, 8);
    }
}

void vtr1_write_rowgroup(FILE *fp, const VecBatch *batch) {
    const int64_t n = batch->n_rows;
    put_le(fp, (uint64_t)n, 8);
    for (int c = 0; c < batch->n_cols; c++) {
        uint8_t has;
        uint64_t lo, hi;
        column_stats(&batch->columns[c], n, &has, &lo, &hi);
        put_le(fp, has, 1);
        put_le(fp, lo, 8);
        put_le(fp, hi, 8);
    }
    const int64_t vb = vec_validity_bytes(n);
    uint8_t *ones = NULL;
    for (int c = 0; c < batch->n_cols; c++) {
        const VecArray *a = &batch->columns[c];
        if (a->validity) {
            fwrite(a->validity, 1, (size_t)vb, fp);
        } else {
            /* an all-valid array without a bitmap: bits 0..n-1 set */
            if (!ones) {
                ones = (uint8_t *)calloc((size_t)(vb ? vb : 1), 1);
                if (!ones) vectra_error("alloc failed");
                for (int64_t i = 0; i < n; i++) ones[i >> 3] |= (uint8_t)(1u << (i & 7));
            }
            fwrite(ones, 1, (size_t)vb, fp);
        }
        switch (a->type) {
        case VEC_INT64:
        case VEC_DOUBLE:
        case VEC_BOOL:
            fwrite(a->buf.i64, 1, (size_t)value_bytes(a->type, n), fp);
            break;
        case VEC_STRING:
            fwrite(a->buf.str.offsets, 8, (size_t)(n + 1), fp);
            put_le(fp, (uint64_t)a->buf.str.data_len, 8);
            fwrite(a->buf.str.data, 1, (size_t)a->buf.str.data_len, fp);
            break;
        }
    }
    free(ones);
}

void vtr1_write(const char *path, const VecBatch *batch) {
    FILE *fp = fopen(path, "wb");
    if (!fp) vectra_error("cannot open file for writing: %s", path);
    VecType *types = (VecType *)malloc((size_t)(batch->n_cols > 0 ? batch->n_cols : 1) * sizeof(VecType));
    if (!types) {
        fclose(fp);
        vectra_error("alloc failed");
    }
    for (int c = 0; c < batch->n_cols; c++) types[c] = batch->columns[c].type;
    VecSchema s;
    memset(&s, 0, sizeof s);
    s.n_cols = batch->n_cols;
    s.col_names = batch->col_names;
    s.col_types = types;
    vtr1_write_header(fp, &s, 1);
    vtr1_write_rowgroup(fp, batch);
    free(types);
    fclose(fp);
}

/* ---- reading ------------------------------------------------------------ */

/* skips one column of an n-row row group (the file positioned at it) */
static void skip_column(FILE *fp, VecType t, int64_t n) {
    skip_bytes(fp, vec_validity_bytes(n));
    if (t == VEC_STRING) {
        skip_bytes(fp, 8 * (n + 1));
        skip_bytes(fp, (int64_t)get_le(fp, 8));
    } else {
        skip_bytes(fp, value_bytes(t, n));
    }
}

static char *read_str16(FILE *fp, const char *what) {
    const size_t len = (size_t)get_le(fp, 2);
    char *s = (char *)malloc(len + 1);
    if (!s) vectra_error("alloc failed");
    read_exact(fp, s, len, what);
    s[len] = '\0';
    return s;
}

Vtr1File *vtr1_open(const char *path) {
    FILE *fp = fopen(path, "rb");
    if (!fp) vectra_error("cannot open file: %s", path);
    char magic[4];
    if (fread(magic, 1, 4, fp) != 4 || memcmp(magic, "VTR1", 4) != 0) {
        fclose(fp);
        vectra_error("not a valid .vtr file (bad magic): %s", path);
    }
    Vtr1File *f = (Vtr1File *)calloc(1, sizeof(Vtr1File));
    if (!f) {
        fclose(fp);
        vectra_error("alloc failed");
    }
    f->fp = fp;
    const unsigned version = (unsigned)get_le(fp, 2);
    if (version < 1 || version > 3) {
        fclose(fp);
        free(f);
        vectra_error("unsupported .vtr version: %u", version);
    }
    f->header.version = (uint16_t)version;
    const int n_cols = (int)get_le(fp, 2);
    VecSchema *s = &f->header.schema;
    const size_t nc = (size_t)(n_cols > 0 ? n_cols : 1);
    s->n_cols = n_cols;
    s->col_names = (char **)calloc(nc, sizeof(char *));
    s->col_types = (VecType *)calloc(nc, sizeof(VecType));
    s->col_annotations = (char **)calloc(nc, sizeof(char *));
    if (!s->col_names || !s->col_types || !s->col_annotations) vectra_error("alloc failed reading schema");
    for (int c = 0; c < n_cols; c++) {
        s->col_names[c] = read_str16(fp, "column name");
        s->col_types[c] = (VecType)get_le(fp, 1);
        if (version >= 2) {
            char *ann = read_str16(fp, "annotation");
            if (*ann) s->col_annotations[c] = ann;
            else free(ann);
        }
    }
    f->header.n_rowgroups = (uint32_t)get_le(fp, 4);
    const uint32_t n_rg = f->header.n_rowgroups;
    f->rowgroups = (Vtr1RowGroup *)calloc(n_rg ? n_rg : 1, sizeof(Vtr1RowGroup));
    if (!f->rowgroups) vectra_error("alloc failed for rowgroup index");
    /* the row-group index: each group's offset, size and statistics */
    for (uint32_t g = 0; g < n_rg; g++) {
        Vtr1RowGroup *rg = &f->rowgroups[g];
        rg->file_offset = (int64_t)ftello(fp);
        rg->n_rows = (int64_t)get_le(fp, 8);
        if (version >= 3) {
            rg->col_stats = (Vtr1ColStat *)calloc(nc, sizeof(Vtr1ColStat));
            if (!rg->col_stats) vectra_error("alloc failed");
            for (int c = 0; c < n_cols; c++) {
                Vtr1ColStat *st = &rg->col_stats[c];
                st->has_stats = (uint8_t)get_le(fp, 1);
                const uint64_t lo = get_le(fp, 8), hi = get_le(fp, 8);
                if (!st->has_stats) continue;
                switch (s->col_types[c]) {
                case VEC_INT64:
                    st->i64.min = (int64_t)lo;
                    st->i64.max = (int64_t)hi;
                    break;
                case VEC_DOUBLE:
                    memcpy(&st->dbl.min, &lo, 8);
                    memcpy(&st->dbl.max, &hi, 8);
                    break;
                case VEC_BOOL:
                    st->bln.min = (uint8_t)lo;
                    st->bln.max = (uint8_t)hi;
                    break;
                default:
                    break;
                }
            }
        }
        for (int c = 0; c < n_cols; c++) skip_column(fp, s->col_types[c], rg->n_rows);
    }
    return f;
}

/* file offset of row group g's first column */
static int64_t rowgroup_data_offset(const Vtr1File *f, uint32_t g) {
    return f->rowgroups[g].file_offset + 8 + (f->header.version >= 3 ? 17 * (int64_t)f->header.schema.n_cols : 0);
}

VecBatch *vtr1_read_rowgroup(Vtr1File *f, uint32_t g, const int *col_mask) {
    if (g >= f->header.n_rowgroups)
        vectra_error("row group index out of range: %u >= %u", g, f->header.n_rowgroups);
    const VecSchema *s = &f->header.schema;
    const int64_t n = f->rowgroups[g].n_rows;
    int n_out = 0;
    for (int c = 0; c < s->n_cols; c++) n_out += col_mask[c] != 0;
    VecBatch *b = vec_batch_alloc(n_out, n);
    if (fseeko(f->fp, (off_t)rowgroup_data_offset(f, g), SEEK_SET) != 0) vectra_error("unexpected end of file");
    int o = 0;
    for (int c = 0; c < s->n_cols; c++) {
        const VecType t = s->col_types[c];
        if (!col_mask[c]) {
            skip_column(f->fp, t, n);
            continue;
        }
        VecArray a = vec_array_alloc(t, n);
        read_exact(f->fp, a.validity, (size_t)vec_validity_bytes(n), "validity bitmap");
        if (t == VEC_STRING) {
            read_exact(f->fp, a.buf.str.offsets, (size_t)(8 * (n + 1)), "string offsets");
            const uint64_t len = get_le(f->fp, 8);
            a.buf.str.data_len = (int64_t)len;
            a.buf.str.data = (char *)malloc(len ? (size_t)len : 1);
            if (!a.buf.str.data) vectra_error("alloc failed for string data");
            read_exact(f->fp, a.buf.str.data, (size_t)len, "string data");
        } else {
            read_exact(f->fp, a.buf.i64, (size_t)value_bytes(t, n),
                       t == VEC_INT64 ? "int64 data" : t == VEC_DOUBLE ? "double data" : "bool data");
        }
        b->columns[o] = a;
        b->col_names[o] = vx_strdup(s->col_names[c]);
        o++;
    }
    return b;
}

void vtr1_close(Vtr1File *f) {
    if (!f) return;
    if (f->fp) fclose(f->fp);
    if (f->rowgroups) {
        for (uint32_t g = 0; g < f->header.n_rowgroups; g++) free(f->rowgroups[g].col_stats);
        free(f->rowgroups);
    }
    vec_schema_free(&f->header.schema);
    free(f);
}

/* ---- the scan node (src/scan.h:9-16, src/scan.c:119-184) -------------- */

static VecBatch *scan_next(VecNode *self) {
    ScanNode *sn = (ScanNode *)self;
    Vtr1File *f = sn->file;
    while (sn->next_rg < f->header.n_rowgroups) {
        const uint32_t g = sn->next_rg++;
        if (sn->predicate && f->header.version >= 3 &&
            !vx_vtr_might_match(sn->predicate, f->rowgroups[g].col_stats, &f->header.schema))
            continue;
        return vtr1_read_rowgroup(f, g, sn->col_mask);
    }
    return NULL;
}

static void scan_free(VecNode *self) {
    ScanNode *sn = (ScanNode *)self;
    if (sn->predicate && !sn->pred_borrowed) vec_expr_free(sn->predicate);
    vtr1_close(sn->file);
    free(sn->col_mask);
    vec_schema_free(&sn->base.output_schema);
    free(sn);
}

ScanNode *scan_node_create(const char *path, int *col_indices, int n_selected) {
    ScanNode *sn = (ScanNode *)calloc(1, sizeof(ScanNode));
    if (!sn) vectra_error("alloc failed for ScanNode");
    sn->file = vtr1_open(path);
    const VecSchema *fs = &sn->file->header.schema;
    sn->col_mask = (int *)calloc((size_t)(fs->n_cols > 0 ? fs->n_cols : 1), sizeof(int));
    if (!sn->col_mask) vectra_error("alloc failed for col_mask");
    if (!col_indices) {
        for (int c = 0; c < fs->n_cols; c++) sn->col_mask[c] = 1;
        sn->base.output_schema = vec_schema_copy(fs);
    } else {
        for (int i = 0; i < n_selected; i++) {
            if (col_indices[i] < 0 || col_indices[i] >= fs->n_cols)
                vectra_error("column index out of range: %d", col_indices[i]);
            sn->col_mask[col_indices[i]] = 1;
        }
        /* the selected columns in file order */
        char **names = (char **)malloc((size_t)(n_selected > 0 ? n_selected : 1) * sizeof(char *));
        VecType *types = (VecType *)malloc((size_t)(n_selected > 0 ? n_selected : 1) * sizeof(VecType));
        if (!names || !types) vectra_error("alloc failed");
        int k = 0;
        for (int c = 0; c < fs->n_cols; c++)
            if (sn->col_mask[c]) {
                names[k] = fs->col_names[c];
                types[k] = fs->col_types[c];
                k++;
            }
        sn->base.output_schema = vec_schema_create(k, names, types);
        free(names);
        free(types);
    }
    sn->base.next_batch = scan_next;
    sn->base.free_node = scan_free;
    sn->base.kind = "ScanNode";
    return sn;
}

/* ---- harness entry points (no reference counterpart) ------------------ */

/* The optimizer's predicate pushdown onto a scan below a filter
 * (src/optimize.c:353-369): the scan borrows the filter's predicate for
 * row-group pruning when its file has statistics.  Applied to every
 * FilterNode directly above a ScanNode in the chain starting at node. */
void vx_pushdown_predicates(VecNode *node) {
    while (node && node->kind && strcmp(node->kind, "FilterNode") == 0) {
        FilterNode *fn = (FilterNode *)node;
        if (fn->child->kind && strcmp(fn->child->kind, "ScanNode") == 0) {
            ScanNode *sn = (ScanNode *)fn->child;
            if (sn->file->header.version >= 3 && !sn->predicate) {
                sn->predicate = fn->predicate;
                sn->pred_borrowed = 1;
            }
        }
        node = fn->child;
    }
}

int vx_try_scan(const char *path, int *col_indices, int n_selected, VecNode **out) {
    vx_frame f;
    *out = NULL;
    vx_frame_push(&f);
    if (setjmp(f.jb) == 0) {
        *out = (VecNode *)scan_node_create(path, col_indices, n_selected);
        vx_frame_pop(&f);
        return 0;
    }
    return -1;
}

/* writes columns cols[0..n_cols) of n_rows rows as a .vtr file of row
 * groups of rg_rows rows (one group when rg_rows <= 0) */
int vx_try_vtr_write(const char *path, int n_cols, const char *const *names, const VecArray *cols, int64_t n_rows,
                     int64_t rg_rows) {
    vx_frame f;
    FILE *volatile fp = NULL;
    VecBatch *volatile b = NULL;
    vx_frame_push(&f);
    if (setjmp(f.jb) == 0) {
        fp = fopen(path, "wb");
        if (!fp) vectra_error("cannot open file for writing: %s", path);
        const int64_t rg = rg_rows > 0 ? rg_rows : (n_rows > 0 ? n_rows : 1);
        const uint32_t n_rg = (uint32_t)((n_rows + rg - 1) / rg);
        VecType types[64];
        if (n_cols > 64) vectra_error("too many columns");
        for (int c = 0; c < n_cols; c++) types[c] = cols[c].type;
        VecSchema s;
        memset(&s, 0, sizeof s);
        s.n_cols = n_cols;
        s.col_names = (char **)names;
        s.col_types = types;
        vtr1_write_header(fp, &s, n_rg);
        for (uint32_t g = 0; g < n_rg; g++) {
            const int64_t r0 = (int64_t)g * rg;
            const int64_t n = n_rows - r0 < rg ? n_rows - r0 : rg;
            b = vec_batch_alloc(n_cols, n);
            for (int c = 0; c < n_cols; c++) {
                const VecArray *a = &cols[c];
                VecArray o = vec_array_alloc(a->type, n);
                for (int64_t i = 0; i < n; i++)
                    if (vec_array_is_valid(a, r0 + i)) vec_array_set_valid(&o, i);
                if (a->type == VEC_STRING) {
                    const int64_t d0 = a->buf.str.offsets[r0], d1 = a->buf.str.offsets[r0 + n];
                    for (int64_t i = 0; i <= n; i++) o.buf.str.offsets[i] = a->buf.str.offsets[r0 + i] - d0;
                    o.buf.str.data_len = d1 - d0;
                    o.buf.str.data = (char *)malloc((size_t)(d1 - d0 > 0 ? d1 - d0 : 1));
                    if (!o.buf.str.data) vectra_error("alloc failed");
                    memcpy(o.buf.str.data, a->buf.str.data + d0, (size_t)(d1 - d0));
                } else {
                    const size_t w = a->type == VEC_BOOL ? 1 : 8;
                    memcpy(o.buf.i64, (const uint8_t *)a->buf.i64 + (size_t)r0 * w, (size_t)n * w);
                }
                b->columns[c] = o;
                b->col_names[c] = vx_strdup(names[c]);
            }
            vtr1_write_rowgroup(fp, b);
            vec_batch_free(b);
            b = NULL;
        }
        fclose(fp);
        vx_frame_pop(&f);
        return 0;
    }
    if (b) vec_batch_free(b);
    if (fp) fclose(fp);
    return -1;
}
