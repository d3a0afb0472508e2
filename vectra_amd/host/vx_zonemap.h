/*
 * vx_zonemap.h -- row-group pruning with .vtr statistics (src/scan.c:10-117),
 * shared by the harness scan node and the GPU nodes' device scan feed (the
 * reference keeps its version static inside its scan, so the GPU library
 * carries its own).
 *
 * 0 when a row group whose statistics are `stats` cannot hold a row that
 * satisfies e: a `col op literal` comparison with the literal on either
 * side (int64 or double, compared as doubles against the column's min and
 * max), or an AND (both sides) / OR (either side) of such; 1 otherwise,
 * including bool and string columns and every other expression.
 */
#ifndef VX_ZONEMAP_H
#define VX_ZONEMAP_H

#include "vectra_abi.h"
#include "vx_core.h"

static inline int vx_vtr_might_match(const VecExpr *e, const Vtr1ColStat *stats, const VecSchema *schema) {
    if (!e || !stats) return 1;
    if (e->kind == EXPR_BOOL && (e->op == '&' || e->op == '|')) {
        const int l = vx_vtr_might_match(e->left, stats, schema);
        const int r = vx_vtr_might_match(e->right, stats, schema);
        return e->op == '&' ? (l && r) : (l || r);
    }
    if (e->kind != EXPR_CMP) return 1;
    const VecExpr *col, *lit;
    int flip;
    const int lit_r = e->right->kind == EXPR_LIT_INT64 || e->right->kind == EXPR_LIT_DOUBLE;
    const int lit_l = e->left->kind == EXPR_LIT_INT64 || e->left->kind == EXPR_LIT_DOUBLE;
    if (e->left->kind == EXPR_COL_REF && lit_r) {
        col = e->left;
        lit = e->right;
        flip = 0;
    } else if (e->right->kind == EXPR_COL_REF && lit_l) {
        col = e->right;
        lit = e->left;
        flip = 1;
    } else {
        return 1;
    }
    const int ci = vec_schema_find_col(schema, col->col_name);
    if (ci < 0 || !stats[ci].has_stats) return 1;
    double lo, hi;
    if (schema->col_types[ci] == VEC_INT64) {
        lo = (double)stats[ci].i64.min;
        hi = (double)stats[ci].i64.max;
    } else if (schema->col_types[ci] == VEC_DOUBLE) {
        lo = stats[ci].dbl.min;
        hi = stats[ci].dbl.max;
    } else {
        return 1;
    }
    const double v = lit->kind == EXPR_LIT_DOUBLE ? lit->lit_dbl : (double)lit->lit_i64;
    char op = e->op;
    if (flip) op = op == '<' ? '>' : op == '>' ? '<' : op;   /* literal op col == col op' literal */
    const int eq = e->op2 == '=';
    switch (op) {
    case '>': return eq ? hi >= v : hi > v;
    case '<': return eq ? lo <= v : lo < v;
    case '=': return eq ? (lo <= v && hi >= v) : 1;
    case '!': return eq ? !(lo == v && hi == v) : 1;
    default: return 1;
    }
}

#endif
