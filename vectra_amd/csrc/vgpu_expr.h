// vgpu_expr.h -- device evaluation of compiled expression programs.
//
// Semantics are vectra's (src/scalar_ops.c:11-185, src/expr.c:137-146):
// NA propagates through arithmetic and comparison; integer / and % by zero
// give NA; comparisons decide on cmp = (a>b)-(a<b) so a NaN operand
// satisfies ==, <= and >=; & and | are R's three-valued operators.
//
// The general path is a register VM whose registers live in LDS, one
// column of LDS per thread, so a uniform instruction stream indexes them
// without scratch memory.  Programs of the shape  COL op CONST  (the common
// filter) take a fused fast path with no LDS traffic.
#pragma once

#include "vgpu_internal.h"

struct ColSet {
    vgpu_col c[VGPU_MAX_COLS];
};

// host-derived summary of a program: fast path when it is one comparison
// of a column against a constant
struct PredInfo {
    int32_t present;     // 0 = no predicate (every row passes)
    int32_t fast;        // 1 = fast path fields below are used
    int32_t col;         // column slot
    int32_t cmp_op;      // VGPU_OP_EQ..GE, oriented as  col op k
    int32_t cmp_type;    // VGPU_INT64 or VGPU_DOUBLE
    int32_t col_type;    // type of the loaded column
    uint64_t k;          // constant bits in cmp_type
};

PredInfo vg_pred_info(const vgpu_prog *p);

__host__ __device__ __forceinline__ int vg_cmp_apply(int c, int op) {
    switch (op) {
    case VGPU_OP_EQ: return c == 0;
    case VGPU_OP_NE: return c != 0;
    case VGPU_OP_LT: return c < 0;
    case VGPU_OP_LE: return c <= 0;
    case VGPU_OP_GT: return c > 0;
    default:         return c >= 0;
    }
}

__host__ __device__ __forceinline__ int vg_cmp3_i64(int64_t a, int64_t b) { return (a > b) - (a < b); }
__host__ __device__ __forceinline__ int vg_cmp3_f64(double a, double b) { return (a > b) - (a < b); }

// type of the value an instruction produces (a comparison's type field
// names its operands; its result, like the boolean operators', is bool)
__host__ __device__ __forceinline__ int vg_result_type(const vgpu_insn &in) {
    return (in.op >= VGPU_OP_EQ && in.op <= VGPU_OP_ISNA) ? VGPU_BOOL : in.type;
}

// value of a loaded column in the comparison type
// `value op k` for the loaded bits of a valid value of the fast path's column
__host__ __device__ __forceinline__ bool vg_fast_cmp(const PredInfo &pi, uint64_t bits, uint64_t k) {
    int cmp;
    if (pi.cmp_type == VGPU_DOUBLE) {
        double v = pi.col_type == VGPU_DOUBLE ? vg_bits_to_f64(bits) : (double)(int64_t)bits;
        cmp = vg_cmp3_f64(v, vg_bits_to_f64(k));
    } else {
        cmp = vg_cmp3_i64((int64_t)bits, (int64_t)k);
    }
    return vg_cmp_apply(cmp, pi.cmp_op);
}

__host__ __device__ __forceinline__ bool vg_fast_pred(const PredInfo &pi, const ColSet &cs, int64_t row) {
    const vgpu_col &c = cs.c[pi.col];
    if (!vg_valid(c.validity, row)) return false;
    return vg_fast_cmp(pi, vg_load_bits(c, row), pi.k);
}

// General VM.  regs/regv: LDS arrays [n_insn][stride], this thread's column
// at index tid.  Returns the result register's (value, valid).  Also built
// for the host (vgpu_eval_host) so the compiler and VM are testable on CPU.
__host__ __device__ inline void vg_vm_eval(const vgpu_prog &p, const ColSet &cs, int64_t row,
                                           uint64_t *regs, uint8_t *regv, int stride, int tid,
                                           uint64_t &out_val, bool &out_valid) {
    for (int i = 0; i < p.n_insn; i++) {
        const vgpu_insn in = p.insn[i];
        uint64_t v = 0;
        bool ok = true;
        uint64_t a = 0, b = 0;
        bool av = true, bv = true;
        if (in.op >= VGPU_OP_CAST) {
            a = regs[in.a * stride + tid];
            av = regv[in.a * stride + tid];
        }
        if (in.op >= VGPU_OP_ADD && in.op <= VGPU_OP_OR) {
            b = regs[in.b * stride + tid];
            bv = regv[in.b * stride + tid];
        }
        switch (in.op) {
        case VGPU_OP_COL: {
            const vgpu_col &c = cs.c[in.col];
            ok = vg_valid(c.validity, row);
            v = ok ? vg_load_bits(c, row) : 0;
            break;
        }
        case VGPU_OP_CONST: v = (uint64_t)in.imm.i; break;
        case VGPU_OP_NA: ok = false; break;
        case VGPU_OP_CAST: {
            ok = av;
            int from = vg_result_type(p.insn[in.a]);
            if (in.type == VGPU_DOUBLE && from != VGPU_DOUBLE)
                v = vg_f64_to_bits((double)(int64_t)a);
            else
                v = a;   // bool -> int64 keeps the 0/1 integer
            break;
        }
        case VGPU_OP_ADD: case VGPU_OP_SUB: case VGPU_OP_MUL:
        case VGPU_OP_DIV: case VGPU_OP_MOD: {
            ok = av && bv;
            if (!ok) break;
            if (in.type == VGPU_DOUBLE) {
                double x = vg_bits_to_f64(a), y = vg_bits_to_f64(b), r;
                switch (in.op) {
                case VGPU_OP_ADD: r = x + y; break;
                case VGPU_OP_SUB: r = x - y; break;
                case VGPU_OP_MUL: r = x * y; break;
                case VGPU_OP_DIV: r = x / y; break;
                default:          r = fmod(x, y); break;
                }
                v = vg_f64_to_bits(r);
            } else {
                int64_t x = (int64_t)a, y = (int64_t)b;
                switch (in.op) {
                case VGPU_OP_ADD: v = a + b; break;
                case VGPU_OP_SUB: v = a - b; break;
                case VGPU_OP_MUL: v = a * b; break;
                case VGPU_OP_DIV:
                    if (y == 0) ok = false;
                    else v = (y == -1) ? (uint64_t)0 - a : (uint64_t)(x / y);
                    break;
                default:
                    if (y == 0) ok = false;
                    else v = (y == -1) ? 0 : (uint64_t)(x % y);
                    break;
                }
            }
            break;
        }
        case VGPU_OP_EQ: case VGPU_OP_NE: case VGPU_OP_LT:
        case VGPU_OP_LE: case VGPU_OP_GT: case VGPU_OP_GE: {
            ok = av && bv;
            if (!ok) break;
            int c = in.type == VGPU_DOUBLE ? vg_cmp3_f64(vg_bits_to_f64(a), vg_bits_to_f64(b))
                                           : vg_cmp3_i64((int64_t)a, (int64_t)b);
            v = (uint64_t)vg_cmp_apply(c, in.op);
            break;
        }
        case VGPU_OP_AND: {
            if ((av && !a) || (bv && !b)) { ok = true; v = 0; }
            else if (av && bv)            { ok = true; v = 1; }
            else                          { ok = false; }
            break;
        }
        case VGPU_OP_OR: {
            if ((av && a) || (bv && b)) { ok = true; v = 1; }
            else if (av && bv)          { ok = true; v = 0; }
            else                        { ok = false; }
            break;
        }
        case VGPU_OP_NOT: ok = av; v = av ? (uint64_t)(a == 0) : 0; break;
        case VGPU_OP_ISNA: ok = true; v = (uint64_t)!av; break;
        case VGPU_OP_NEG:
            ok = av;
            if (in.type == VGPU_DOUBLE) v = a ^ 0x8000000000000000ull;
            else v = (uint64_t)0 - a;
            break;
        default: ok = false; break;
        }
        regs[i * stride + tid] = v;
        regv[i * stride + tid] = ok;
    }
    int r = p.n_insn - 1;
    out_val = regs[r * stride + tid];
    out_valid = regv[r * stride + tid];
}

// predicate test of one row: valid && true
__host__ __device__ __forceinline__ bool vg_pred_row(const PredInfo &pi, const vgpu_prog &p, const ColSet &cs,
                                            int64_t row, uint64_t *regs, uint8_t *regv, int stride, int tid) {
    if (!pi.present) return true;
    if (pi.fast) return vg_fast_pred(pi, cs, row);
    uint64_t v;
    bool ok;
    vg_vm_eval(p, cs, row, regs, regv, stride, tid, v, ok);
    return ok && v != 0;
}

// LDS bytes the VM needs for a block of `threads`
static inline size_t vg_vm_lds_bytes(const vgpu_prog *p, int threads) {
    if (!p) return 0;
    return (size_t)p->n_insn * threads * 9;
}
