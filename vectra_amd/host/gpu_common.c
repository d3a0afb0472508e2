/*
 * gpu_common.c -- device column staging and the expression compiler used by
 * the GPU operator nodes.
 *
 * The compiler applies vectra's typing rules at compile time so the device
 * never coerces implicitly:
 *   common type bool < int64 < double, string never coerces . src/coerce.c:7-19
 *   arithmetic / comparison promote bool to int64 .......... src/scalar_ops.c:16-17, 109-110
 *   & | ! need boolean operands ............................ src/scalar_ops.c:151-203
 *   unary minus on int64 / double only ..................... src/scalar_ops.c:205-228
 *   if_else / cast unimplemented ........................... src/expr.c:275-278
 */
#define _GNU_SOURCE
#include "gpu_common.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static int g_gpu_ready;

void vx_gpu_init(void) {
    if (g_gpu_ready) return;
    const char *dev = getenv("VECTRA_GPU_DEVICE");
    VX_GPU(vgpu_init(dev ? atoi(dev) : 0));
    g_gpu_ready = 1;
}

int vx_all_valid(const uint8_t *v, int64_t n) {
    if (!v) return 1;
    int64_t full = n / 8;
    for (int64_t i = 0; i < full; i++)
        if (v[i] != 0xFF) return 0;
    int rem = (int)(n % 8);
    if (rem && (v[full] & ((1u << rem) - 1)) != ((1u << rem) - 1)) return 0;
    return 1;
}

void *devmem_reserve(DevMem *m, size_t bytes) {
    if (bytes <= m->cap && m->ptr) return m->ptr;
    size_t n = m->cap ? m->cap : 4096;
    while (n < bytes) n *= 2;
    if (m->ptr) vgpu_free(m->ptr);
    m->ptr = vgpu_malloc(n);
    if (!m->ptr) {
        m->cap = 0;
        vectra_error("GPU: %s", vgpu_last_error());
    }
    m->cap = n;
    return m->ptr;
}

void devmem_free(DevMem *m) {
    if (m->ptr) vgpu_free(m->ptr);
    m->ptr = NULL;
    m->cap = 0;
}

void devcol_reserve(DevCol *dc, size_t data_bytes, size_t valid_bytes) {
    DevMem d = { dc->d_data, dc->cap_data }, v = { dc->d_valid, dc->cap_valid };
    devmem_reserve(&d, data_bytes ? data_bytes : 8);
    devmem_reserve(&v, valid_bytes ? valid_bytes : 8);
    dc->d_data = d.ptr; dc->cap_data = d.cap;
    dc->d_valid = (uint8_t *)v.ptr; dc->cap_valid = v.cap;
}

void devcol_upload(DevCol *dc, const VecArray *a, int64_t n) {
    if (a->type == VEC_STRING)
        vectra_error("GPU: string columns are not evaluated on the GPU");
    size_t w = a->type == VEC_BOOL ? 1 : 8;
    size_t vb = (size_t)vec_validity_bytes(n);
    devcol_reserve(dc, (size_t)n * w, vb + 8);
    VX_GPU(vgpu_memcpy_h2d(dc->d_data, a->buf.i64, (size_t)n * w));
    dc->col.type = (int32_t)a->type;
    dc->col.reserved = 0;
    dc->col.data = dc->d_data;
    if (vx_all_valid(a->validity, n)) {
        dc->col.validity = NULL;
    } else {
        VX_GPU(vgpu_memcpy_h2d(dc->d_valid, a->validity, vb));
        dc->col.validity = dc->d_valid;
    }
}

void devcol_free(DevCol *dc) {
    if (dc->d_data) vgpu_free(dc->d_data);
    if (dc->d_valid) vgpu_free(dc->d_valid);
    memset(dc, 0, sizeof *dc);
}

VecType vx_common_type(VecType a, VecType b) {
    if (a == b) return a;
    if (a == VEC_STRING || b == VEC_STRING)
        vectra_error("cannot coerce string in arithmetic/comparison");
    if (a == VEC_DOUBLE || b == VEC_DOUBLE) return VEC_DOUBLE;
    if (a == VEC_INT64 || b == VEC_INT64) return VEC_INT64;
    return VEC_BOOL;
}

const char *vx_type_name(VecType t) {
    switch (t) {
    case VEC_INT64:  return "int64";
    case VEC_DOUBLE: return "double";
    case VEC_BOOL:   return "bool";
    case VEC_STRING: return "string";
    }
    return "unknown";
}

/* ------------------------------------------------------------------ */
/* expression compiler                                                */
/* ------------------------------------------------------------------ */

typedef struct {
    CompiledExpr  *out;
    const VecBatch *batch;
    VecType        rtype[VGPU_MAX_INSN];
} Cx;

static int cx_push(Cx *c, uint8_t op, VecType type, int a, int b) {
    vgpu_prog *p = &c->out->prog;
    if (p->n_insn >= VGPU_MAX_INSN)
        vectra_error("GPU: expression longer than %d instructions", VGPU_MAX_INSN);
    vgpu_insn *in = &p->insn[p->n_insn];
    memset(in, 0, sizeof *in);
    in->op = op;
    in->type = (uint8_t)type;
    in->a = (uint8_t)(a < 0 ? 0 : a);
    in->b = (uint8_t)(b < 0 ? 0 : b);
    c->rtype[p->n_insn] = type;
    return p->n_insn++;
}

/* coerce register r to type t (folding constants) */
static int cx_cast(Cx *c, int r, VecType t) {
    if (c->rtype[r] == t) return r;
    vgpu_insn *src = &c->out->prog.insn[r];
    if (src->op == VGPU_OP_CONST && t == VEC_DOUBLE) {
        double d = (double)src->imm.i;   /* int64 or bool 0/1 */
        int k = cx_push(c, VGPU_OP_CONST, VEC_DOUBLE, -1, -1);
        c->out->prog.insn[k].imm.d = d;
        return k;
    }
    if (src->op == VGPU_OP_CONST && t == VEC_INT64) {
        int k = cx_push(c, VGPU_OP_CONST, VEC_INT64, -1, -1);
        c->out->prog.insn[k].imm.i = src->imm.i;
        return k;
    }
    if (src->op == VGPU_OP_NA) return cx_push(c, VGPU_OP_NA, t, -1, -1);
    return cx_push(c, VGPU_OP_CAST, t, r, -1);
}

static int cx_col(Cx *c, const char *name) {
    const VecBatch *b = c->batch;
    int ci = -1;
    for (int i = 0; i < b->n_cols; i++)
        if (strcmp(b->col_names[i], name) == 0) { ci = i; break; }
    if (ci < 0) vectra_error("column not found: %s", name);
    VecType t = b->columns[ci].type;
    if (t == VEC_STRING)
        vectra_error("GPU: string column '%s' in an expression is not evaluated on the GPU", name);
    CompiledExpr *o = c->out;
    int slot = -1;
    for (int s = 0; s < o->n_cols; s++)
        if (o->col_idx[s] == ci) { slot = s; break; }
    if (slot < 0) {
        if (o->n_cols >= VGPU_MAX_COLS)
            vectra_error("GPU: expression reads more than %d columns", VGPU_MAX_COLS);
        slot = o->n_cols++;
        o->col_idx[slot] = ci;
    }
    int r = cx_push(c, VGPU_OP_COL, t, -1, -1);
    o->prog.insn[r].col = slot;
    return r;
}

/* the batch column a string operand names, -1 for a literal / NA (then
 * *lit / *na are set), -2 when e is no string operand */
static int str_operand(Cx *c, const VecExpr *e, const char **lit, int *na) {
    *lit = NULL;
    *na = 0;
    if (e->kind == EXPR_LIT_STRING) {
        *lit = e->lit_str ? e->lit_str : "";
        return -1;
    }
    if (e->kind == EXPR_LIT_NA && e->result_type == VEC_STRING) {
        *na = 1;
        return -1;
    }
    if (e->kind != EXPR_COL_REF) return -2;
    const VecBatch *b = c->batch;
    for (int i = 0; i < b->n_cols; i++)
        if (strcmp(b->col_names[i], e->col_name) == 0) return b->columns[i].type == VEC_STRING ? i : -2;
    return -2;
}

static uint8_t cmp_opcode(char op, char op2);

/* `l op r` on two strings: a synthetic int64 column cmp(l, r) in a new
 * column slot, compared against 0; -1 when an operand is no string */
static int cx_string_cmp(Cx *c, const VecExpr *e) {
    const char *ll, *rl;
    int lna, rna;
    const int lc = str_operand(c, e->left, &ll, &lna), rc = str_operand(c, e->right, &rl, &rna);
    const int ls = lc != -2, rs = rc != -2;
    if (!ls && !rs) return -1;
    if (ls != rs) vectra_error("cannot coerce string in arithmetic/comparison");
    CompiledExpr *o = c->out;
    if (o->n_scmp >= VX_MAX_SCMP) vectra_error("GPU: more than %d string comparisons in an expression", VX_MAX_SCMP);
    if (o->n_cols >= VGPU_MAX_COLS) vectra_error("GPU: expression reads more than %d columns", VGPU_MAX_COLS);
    const int k = o->n_scmp++;
    o->scmp[k] = (VxStrCmp){{lc, rc}, {ll, rl}, {lna, rna}};
    const int slot = o->n_cols++;
    o->col_idx[slot] = -1 - k;
    const int col = cx_push(c, VGPU_OP_COL, VEC_INT64, -1, -1);
    o->prog.insn[col].col = slot;
    const int zero = cx_push(c, VGPU_OP_CONST, VEC_INT64, -1, -1);
    o->prog.insn[zero].imm.i = 0;
    const uint8_t op = cmp_opcode(e->op, e->op2);
    const int r = op ? cx_push(c, op, VEC_INT64, col, zero) : cx_push(c, VGPU_OP_NE, VEC_INT64, col, col);
    c->rtype[r] = VEC_BOOL;
    return r;
}

static uint8_t cmp_opcode(char op, char op2) {
    if (op == '=' && op2 == '=') return VGPU_OP_EQ;
    if (op == '!' && op2 == '=') return VGPU_OP_NE;
    if (op == '<' && op2 == ' ') return VGPU_OP_LT;
    if (op == '<' && op2 == '=') return VGPU_OP_LE;
    if (op == '>' && op2 == ' ') return VGPU_OP_GT;
    if (op == '>' && op2 == '=') return VGPU_OP_GE;
    /* the reference yields FALSE for an unknown operator pair
     * (src/scalar_ops.c:135-141): compare against an impossible outcome */
    return 0;
}

static int cx_emit(Cx *c, const VecExpr *e) {
    switch (e->kind) {
    case EXPR_COL_REF:
        return cx_col(c, e->col_name);
    case EXPR_LIT_INT64: {
        int r = cx_push(c, VGPU_OP_CONST, VEC_INT64, -1, -1);
        c->out->prog.insn[r].imm.i = e->lit_i64;
        return r;
    }
    case EXPR_LIT_DOUBLE: {
        int r = cx_push(c, VGPU_OP_CONST, VEC_DOUBLE, -1, -1);
        c->out->prog.insn[r].imm.d = e->lit_dbl;
        return r;
    }
    case EXPR_LIT_BOOL: {
        int r = cx_push(c, VGPU_OP_CONST, VEC_BOOL, -1, -1);
        c->out->prog.insn[r].imm.i = e->lit_bln ? 1 : 0;
        return r;
    }
    case EXPR_LIT_NA:
        if (e->result_type == VEC_STRING)
            vectra_error("GPU: string NA literal is not evaluated on the GPU");
        return cx_push(c, VGPU_OP_NA, e->result_type, -1, -1);
    case EXPR_LIT_STRING:
        vectra_error("GPU: string literal is not evaluated on the GPU");
    case EXPR_ARITH: {
        int l = cx_emit(c, e->left), r = cx_emit(c, e->right);
        VecType t = vx_common_type(c->rtype[l], c->rtype[r]);
        if (t == VEC_BOOL) t = VEC_INT64;
        l = cx_cast(c, l, t);
        r = cx_cast(c, r, t);
        uint8_t op;
        switch (e->op) {
        case '+': op = VGPU_OP_ADD; break;
        case '-': op = VGPU_OP_SUB; break;
        case '*': op = VGPU_OP_MUL; break;
        case '/': op = VGPU_OP_DIV; break;
        case '%': op = VGPU_OP_MOD; break;
        default: vectra_error("unknown arith op: %c", e->op);
        }
        return cx_push(c, op, t, l, r);
    }
    case EXPR_CMP: {
        const int sc = cx_string_cmp(c, e);
        if (sc >= 0) return sc;
        int l = cx_emit(c, e->left), r = cx_emit(c, e->right);
        VecType t = vx_common_type(c->rtype[l], c->rtype[r]);
        if (t == VEC_BOOL) t = VEC_INT64;
        l = cx_cast(c, l, t);
        r = cx_cast(c, r, t);
        uint8_t op = cmp_opcode(e->op, e->op2);
        /* an unknown operator pair is FALSE wherever both sides are valid:
         * l != l, decided on (a>b)-(a<b) = 0, is never true */
        int k = op ? cx_push(c, op, t, l, r) : cx_push(c, VGPU_OP_NE, t, l, l);
        c->rtype[k] = VEC_BOOL;   /* the instruction's type field is the operand type */
        return k;
    }
    case EXPR_BOOL: {
        if (e->op == '!') {
            int a = cx_emit(c, e->operand);
            if (c->rtype[a] != VEC_BOOL) vectra_error("bool not: operand must be boolean");
            return cx_push(c, VGPU_OP_NOT, VEC_BOOL, a, -1);
        }
        int l = cx_emit(c, e->left), r = cx_emit(c, e->right);
        if (c->rtype[l] != VEC_BOOL || c->rtype[r] != VEC_BOOL)
            vectra_error("bool: operands must be boolean");
        return cx_push(c, e->op == '&' ? VGPU_OP_AND : VGPU_OP_OR, VEC_BOOL, l, r);
    }
    case EXPR_IS_NA: {
        int a = cx_emit(c, e->operand);
        return cx_push(c, VGPU_OP_ISNA, VEC_BOOL, a, -1);
    }
    case EXPR_NEGATE: {
        int a = cx_emit(c, e->operand);
        VecType t = c->rtype[a];
        if (t != VEC_INT64 && t != VEC_DOUBLE) vectra_error("negate: unsupported type %d", (int)t);
        return cx_push(c, VGPU_OP_NEG, t, a, -1);
    }
    case EXPR_IF_ELSE:
    case EXPR_CAST:
        vectra_error("if_else/cast not yet implemented");
    case EXPR_NCHAR:
    case EXPR_SUBSTR:
    case EXPR_GREPL:
        vectra_error("GPU: string functions are not evaluated on the GPU");
    }
    vectra_error("unknown expr kind: %d", (int)e->kind);
}

void vx_compile_expr(const VecExpr *e, const VecBatch *batch, CompiledExpr *out) {
    memset(out, 0, sizeof *out);
    Cx c;
    memset(&c, 0, sizeof c);
    c.out = out;
    c.batch = batch;
    int r = cx_emit(&c, e);
    /* the result must be the last instruction */
    if (r != out->prog.n_insn - 1) {
        int k = cx_push(&c, VGPU_OP_CAST, c.rtype[r], r, -1);
        (void)k;
    }
    out->result_type = c.rtype[out->prog.n_insn - 1];
}

int vx_try_compile_expr(const VecExpr *e, const VecBatch *batch, CompiledExpr *out) {
    vx_frame f;
    vx_frame_push(&f);
    if (setjmp(f.jb) == 0) {
        vx_compile_expr(e, batch, out);
        vx_frame_pop(&f);
        return 0;
    }
    return -1;
}

/* one side of a string comparison on the device: the batch column's
 * offsets / bytes / validity uploaded, or the literal as a one-row column */
static vgpu_strcol scmp_side(const VxStrCmp *s, int side, const VecBatch *b, DevMem *off, DevMem *chr, DevMem *val) {
    vgpu_strcol sc = {NULL, NULL, NULL, 0, 0};
    if (s->na[side]) {
        sc.null_const = 1;
        return sc;
    }
    if (s->col[side] < 0) {
        const int64_t len = (int64_t)strlen(s->lit[side]);
        const int64_t o[2] = {0, len};
        sc.offsets = (const int64_t *)devmem_reserve(off, 16);
        sc.chars = (const uint8_t *)devmem_reserve(chr, (size_t)(len > 0 ? len : 1));
        VX_GPU(vgpu_memcpy_h2d((void *)sc.offsets, o, 16));
        if (len > 0) VX_GPU(vgpu_memcpy_h2d((void *)sc.chars, s->lit[side], (size_t)len));
        sc.constant = 1;
        return sc;
    }
    const VecArray *a = &b->columns[s->col[side]];
    const int64_t n = b->n_rows, bytes = a->buf.str.offsets[n];   /* the bytes the offsets index */
    int64_t *d_off = (int64_t *)devmem_reserve(off, (size_t)(n + 1) * 8);
    uint8_t *d_chr = (uint8_t *)devmem_reserve(chr, (size_t)(bytes > 0 ? bytes : 1));
    VX_GPU(vgpu_memcpy_h2d(d_off, a->buf.str.offsets, (size_t)(n + 1) * 8));
    if (bytes > 0) VX_GPU(vgpu_memcpy_h2d(d_chr, a->buf.str.data, (size_t)bytes));
    sc.offsets = d_off;
    sc.chars = d_chr;
    if (!vx_all_valid(a->validity, n)) {
        const size_t vb = (size_t)vec_validity_bytes(n);
        sc.validity = (const uint8_t *)devmem_reserve(val, vb + 8);
        VX_GPU(vgpu_memcpy_h2d((void *)sc.validity, a->validity, vb));
    }
    return sc;
}

vgpu_col vx_string_cmp_col(const CompiledExpr *cx, int k, const VecBatch *b, DevCol *dc, DevMem stage[6]) {
    const int64_t n = b->n_rows;
    devcol_reserve(dc, (size_t)(n > 0 ? n : 1) * 8, (size_t)((n + 63) / 64 + 1) * 8);
    const VxStrCmp *s = &cx->scmp[k];
    const vgpu_strcol l = scmp_side(s, 0, b, &stage[0], &stage[1], &stage[2]);
    const vgpu_strcol r = scmp_side(s, 1, b, &stage[3], &stage[4], &stage[5]);
    VX_GPU(vgpu_string_compare(&l, &r, n, (int64_t *)dc->d_data, dc->d_valid));
    dc->col = (vgpu_col){VGPU_INT64, 0, dc->d_data, dc->d_valid};
    return dc->col;
}

static int scmp_host_get(const VxStrCmp *s, int side, const VecBatch *b, int64_t r, const char **p, int64_t *len) {
    if (s->na[side]) return 0;
    if (s->col[side] < 0) {
        *p = s->lit[side];
        *len = (int64_t)strlen(s->lit[side]);
        return 1;
    }
    const VecArray *a = &b->columns[s->col[side]];
    if (!vec_array_is_valid(a, r)) return 0;
    *p = a->buf.str.data + a->buf.str.offsets[r];
    *len = a->buf.str.offsets[r + 1] - a->buf.str.offsets[r];
    return 1;
}

void vx_string_cmp_host(const CompiledExpr *cx, int k, const VecBatch *b, int64_t *out, uint8_t *valid) {
    const VxStrCmp *s = &cx->scmp[k];
    memset(valid, 0, (size_t)vec_validity_bytes(b->n_rows));
    for (int64_t r = 0; r < b->n_rows; r++) {
        const char *pa = NULL, *pb = NULL;
        int64_t la = 0, lb = 0;
        out[r] = 0;
        if (!scmp_host_get(s, 0, b, r, &pa, &la) || !scmp_host_get(s, 1, b, r, &pb, &lb)) continue;
        const int64_t m = la < lb ? la : lb;
        int c = m > 0 ? memcmp(pa, pb, (size_t)m) : 0;
        c = c < 0 ? -1 : c > 0 ? 1 : (la > lb) - (la < lb);
        out[r] = c;
        valid[r >> 3] |= (uint8_t)(1u << (r & 7));
    }
}

/* ------------------------------------------------------------------ */
/* build-side materialization                                         */
/* ------------------------------------------------------------------ */

int64_t vx_concat_batches(VecNode *child, VecArray **cols_out) {
    VecBatch *solo = NULL;
    /* vx_collect already implements the logical-row concatenation */
    solo = vx_collect(child);
    int64_t n = solo->n_rows;
    *cols_out = solo->columns;
    solo->columns = NULL;
    for (int i = 0; i < solo->n_cols; i++) free(solo->col_names[i]);
    free(solo->col_names);
    free(solo);
    return n;
}
