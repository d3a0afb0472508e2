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
