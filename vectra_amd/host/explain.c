/*
 * explain.c -- the plan text of a node tree, as explain() prints it
 * (src/r_bridge.c:672-698): one line per node in pre-order, two spaces of
 * indent per level, "<kind> [<annotation>]", with the reference's
 * annotation for the node kinds of the path (src/r_bridge.c:575-670) and,
 * for the GPU nodes, what the device does after "; ".
 */
#include "gpu_common.h"

#include <stdarg.h>
#include <stdio.h>
#include <string.h>

int vx_gpu_filter_note(const VecNode *n, char *buf, int size);
int vx_gpu_group_note(const VecNode *n, char *buf, int size);
int vx_gpu_join_note(const VecNode *n, char *buf, int size);

#define EX_MAX_LINES 64   /* src/r_bridge.c:706 */

/* appends to buf[0..size) at *pos, never past its end; *pos counts what
 * would have been written */
static void put(char *buf, size_t size, size_t *pos, const char *fmt, ...) __attribute__((format(printf, 4, 5)));
static void put(char *buf, size_t size, size_t *pos, const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    const size_t room = *pos < size ? size - *pos : 0;
    const int n = vsnprintf(room ? buf + *pos : NULL, room, fmt, ap);
    va_end(ap);
    if (n > 0) *pos += (size_t)n;
}

static void annotation(const VecNode *node, char *buf, size_t size) {
    const char *kind = node->kind ? node->kind : "Unknown";
    size_t pos = 0;
    buf[0] = '\0';
    if (strcmp(kind, "ScanNode") == 0) {
        const ScanNode *sn = (const ScanNode *)node;
        const int file_cols = sn->file->header.schema.n_cols, read_cols = sn->base.output_schema.n_cols;
        if (read_cols < file_cols) put(buf, size, &pos, "streaming, %d/%d cols (pruned)", read_cols, file_cols);
        else put(buf, size, &pos, "streaming, %d cols", read_cols);
        if (sn->predicate) put(buf, size, &pos, ", predicate pushdown");
        if (sn->file->header.version >= 3) put(buf, size, &pos, ", v3 stats");
    } else if (strcmp(kind, "MemScanNode") == 0) {
        put(buf, size, &pos, "in memory, %d cols", node->output_schema.n_cols);
    } else if (strcmp(kind, "FilterNode") == 0) {
        put(buf, size, &pos, "streaming");
    } else if (strcmp(kind, "ProjectNode") == 0) {
        const ProjectNode *pn = (const ProjectNode *)node;
        int tmp = 0;
        for (int i = 0; i < pn->n_entries; i++) tmp |= strncmp(pn->entries[i].output_name, ".vectra_tmp_", 12) == 0;
        put(buf, size, &pos, tmp ? "streaming, hidden mutate" : "streaming");
    } else if (strcmp(kind, "GroupAggNode") == 0) {
        put(buf, size, &pos, "materializes, %d keys", ((const GroupAggNode *)node)->n_keys);
    } else if (strcmp(kind, "JoinNode") == 0) {
        static const char *const names[] = {"inner", "left", "full", "semi", "anti"};
        const JoinNode *jn = (const JoinNode *)node;
        const char *jk = (int)jn->kind >= 0 && (int)jn->kind < 5 ? names[jn->kind] : "unknown";
        put(buf, size, &pos, "build right + stream left, %s, %d keys", jk, jn->n_keys);
    }
    char note[320];
    if (vx_gpu_filter_note(node, note, (int)sizeof note) >= 0 || vx_gpu_group_note(node, note, (int)sizeof note) >= 0 ||
        vx_gpu_join_note(node, note, (int)sizeof note) >= 0)
        put(buf, size, &pos, "%s%s", pos ? "; " : "", note);
}

static int children(const VecNode *n, const VecNode **out) {
    const char *kind = n->kind ? n->kind : "";
    if (strcmp(kind, "FilterNode") == 0) {
        out[0] = ((const FilterNode *)n)->child;
        return 1;
    }
    if (strcmp(kind, "GroupAggNode") == 0) {
        out[0] = ((const GroupAggNode *)n)->child;
        return 1;
    }
    if (strcmp(kind, "ProjectNode") == 0) {
        out[0] = ((const ProjectNode *)n)->child;
        return 1;
    }
    if (strcmp(kind, "JoinNode") == 0) {
        out[0] = ((const JoinNode *)n)->left;
        out[1] = ((const JoinNode *)n)->right;
        return 2;
    }
    return 0;
}

static void lines(const VecNode *n, int depth, char *buf, size_t size, size_t *pos, int *count) {
    if (*count >= EX_MAX_LINES) return;
    char ann[512];
    annotation(n, ann, sizeof ann);
    put(buf, size, pos, "%*s%s [%s]\n", 2 * depth, "", n->kind ? n->kind : "Unknown", ann);
    (*count)++;
    const VecNode *c[2];
    const int nc = children(n, c);
    for (int i = 0; i < nc; i++) lines(c[i], depth + 1, buf, size, pos, count);
}

int64_t vx_explain(const VecNode *root, char *buf, int64_t size) {
    if (!root || size < 0) return -1;
    size_t pos = 0;
    int count = 0;
    if (buf && size > 0) buf[0] = '\0';
    lines(root, 0, buf, (size_t)size, &pos, &count);
    return (int64_t)pos;
}
