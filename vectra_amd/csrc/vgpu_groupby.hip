// vgpu_groupby.hip -- hash group-by aggregation (n, count, sum, mean, min,
// max over int64 / double / bool inputs) with vectra's NA, NaN and
// ordering rules.
//
// Reference semantics (gcol33/vectra 0.2.2):
//   grouping, NA == NA, NaN != NaN, -0 == +0 ..... src/hash.c:12-78
//   first-occurrence group order (hash path) ..... src/hash.c:117-122
//   key-sorted order, NA last (sorted path) ...... src/sort.c:32-85
//   accumulators and their finish ................ src/agg_ops.c:93-255
//
// Design.  Groups live in one open-addressing table in HBM whose records
// are arrays of 64-bit words ("hot": key, n, first position, per input
// sum / min / max; "cold": per input NA count and the NaN / signed-zero
// positions that only NA, NaN and zero values touch).  Every update is an
// atomic on a word, so blocks never coordinate.  When the expected group
// count fits, each block first aggregates into a private copy of the table
// in LDS and merges its occupied records into HBM once at the end (the
// low-cardinality path); otherwise rows update the HBM table directly.
//
// Exactness.  Counts and int64 sum/min/max are exact (wrapping int64 sum,
// order-preserving unsigned maps for min/max).  Double min/max reproduce
// the reference's order-dependent cases exactly by position: the result is
// NaN iff the group's first valid value is NaN, and a zero result takes the
// sign of the group's first zero.  Double sums use FP64 atomic adds, so
// they differ from the reference's sequential sum only in rounding order.
#include "vgpu_expr.h"
#include "vgpu_join_table.h"

#include <algorithm>
#include <cmath>
#include <vector>

#define GB_THREADS 256
#define GB_MAXW 64
#define GB_NOPOS 0xFFFFFFFFFFFFFFFFull
#define GB_LDS_BUDGET (64 * 1024)

struct GbLayout {
    int32_t RH, RC;
    int32_t w_fpos;
    int32_t n_inputs;
    int32_t key_is_double;
    int32_t key_none;
    int32_t in_type[VGPU_MAX_INPUTS];
    int32_t w_sum[VGPU_MAX_INPUTS], w_min[VGPU_MAX_INPUTS], w_max[VGPU_MAX_INPUTS];
    int32_t w_fvpos[VGPU_MAX_INPUTS];
    int32_t c_na[VGPU_MAX_INPUTS], c_nan[VGPU_MAX_INPUTS], c_zneg[VGPU_MAX_INPUTS], c_zpos[VGPU_MAX_INPUTS];
    // initial value class per word: 0 = zero, 1 = all-ones
    uint64_t hot_init_ones;   // bit w set -> hot word w starts at all-ones
    uint64_t cold_init_ones;
};

struct GbAggs {
    int32_t n;
    vgpu_agg a[VGPU_MAX_AGGS];
};

struct vgpu_groupby {
    GbLayout L;
    GbAggs aggs;
    int32_t order;
    int64_t cap = 0;          // hashed slots (pow2); +2 reserved slots follow
    int64_t n_groups = 0;     // after finish
    int64_t hint = 0;
    int64_t est_groups = 0;   // groups expected when no hint was given (sketch or groups seen); 0 = unknown
    int32_t finished = 0;
    int32_t any_fed = 0;
    DevBuf hot, cold;
    DevBuf sketch;            // cardinality sketch registers (GB_SK_REGS u32) + NaN-key row count (u64)
    DevBuf ctl;               // [0] occupied count, [1] error flag, [2] zero-key -0 pos,
                              // [3] zero-key +0 pos, [4] warn count, [5] tmp counter,
                              // [6] or-reduce, [7] and-reduce
    // finish buffers
    DevBuf sort_keys, sort_vals, sort_tmp;
    DevBuf out_keys, out_kvalid;
    std::vector<DevBuf> out_vals, out_valid;
    int64_t warn_count = 0;
};

// ---------------------------------------------------------------------------
// record updates (generic pointers: the same code runs on LDS and HBM)
// ---------------------------------------------------------------------------

__device__ __forceinline__ void gb_atomic_min(uint64_t *p, uint64_t v) { atomicMin((unsigned long long *)p, (unsigned long long)v); }
__device__ __forceinline__ void gb_atomic_max(uint64_t *p, uint64_t v) { atomicMax((unsigned long long *)p, (unsigned long long)v); }
__device__ __forceinline__ void gb_atomic_add(uint64_t *p, uint64_t v) { atomicAdd((unsigned long long *)p, (unsigned long long)v); }
__device__ __forceinline__ void gb_atomic_addf(uint64_t *p, double v) { atomicAdd(reinterpret_cast<double *>(p), v); }

// one input value of the row being aggregated
struct InVal {
    uint64_t bits;
    bool valid;
};

// the per-row words: n and the first position
__device__ __forceinline__ void gb_update_row(const GbLayout &L, uint64_t *hot, uint64_t pos) {
    gb_atomic_add(&hot[1], 1);
    if (L.w_fpos >= 0) gb_atomic_min(&hot[L.w_fpos], pos);
}

// the words of input i (no local arrays: everything stays in registers)
__device__ __forceinline__ void gb_update_input(const GbLayout &L, uint64_t *hot, uint64_t *cold, int i,
                                                uint64_t pos, InVal v) {
    if (!v.valid) {
        gb_atomic_add(&cold[L.c_na[i]], 1);
        return;
    }
    uint64_t b = v.bits;
    int t = L.in_type[i];
    if (L.w_fvpos[i] >= 0) gb_atomic_min(&hot[L.w_fvpos[i]], pos);
    if (L.w_sum[i] >= 0) {
        if (t == VGPU_DOUBLE) gb_atomic_addf(&hot[L.w_sum[i]], vg_bits_to_f64(b));
        else gb_atomic_add(&hot[L.w_sum[i]], b);   // int64 wraps; bool adds 0/1
    }
    if (t == VGPU_BOOL) return;   // the reference never updates min/max for bool
    if (L.w_min[i] < 0 && L.w_max[i] < 0) return;
    uint64_t o;
    if (t == VGPU_DOUBLE) {
        if ((b & 0x7FFFFFFFFFFFFFFFull) > 0x7FF0000000000000ull) {
            gb_atomic_min(&cold[L.c_nan[i]], pos);
            return;
        }
        if ((b & 0x7FFFFFFFFFFFFFFFull) == 0) {
            gb_atomic_min(&cold[(b >> 63) ? L.c_zneg[i] : L.c_zpos[i]], pos);
            b = 0;
        }
        o = vg_f64_ordered(b);
    } else {
        o = vg_i64_ordered(b);
    }
    if (L.w_min[i] >= 0) gb_atomic_min(&hot[L.w_min[i]], o);
    if (L.w_max[i] >= 0) gb_atomic_max(&hot[L.w_max[i]], o);
}

// merge a whole record (LDS partial -> HBM)
__device__ __forceinline__ void gb_merge(const GbLayout &L, uint64_t *dh, uint64_t *dc,
                                         const uint64_t *sh, const uint64_t *sc) {
    gb_atomic_add(&dh[1], sh[1]);
    if (L.w_fpos >= 0) gb_atomic_min(&dh[L.w_fpos], sh[L.w_fpos]);
    for (int i = 0; i < L.n_inputs; i++) {
        if (sc[L.c_na[i]]) gb_atomic_add(&dc[L.c_na[i]], sc[L.c_na[i]]);
        if (L.w_fvpos[i] >= 0) gb_atomic_min(&dh[L.w_fvpos[i]], sh[L.w_fvpos[i]]);
        if (L.w_sum[i] >= 0) {
            if (L.in_type[i] == VGPU_DOUBLE) gb_atomic_addf(&dh[L.w_sum[i]], vg_bits_to_f64(sh[L.w_sum[i]]));
            else gb_atomic_add(&dh[L.w_sum[i]], sh[L.w_sum[i]]);
        }
        if (L.w_min[i] >= 0) gb_atomic_min(&dh[L.w_min[i]], sh[L.w_min[i]]);
        if (L.w_max[i] >= 0) gb_atomic_max(&dh[L.w_max[i]], sh[L.w_max[i]]);
        if (L.c_nan[i] >= 0) {
            if (sc[L.c_nan[i]] != GB_NOPOS) gb_atomic_min(&dc[L.c_nan[i]], sc[L.c_nan[i]]);
            if (sc[L.c_zneg[i]] != GB_NOPOS) gb_atomic_min(&dc[L.c_zneg[i]], sc[L.c_zneg[i]]);
            if (sc[L.c_zpos[i]] != GB_NOPOS) gb_atomic_min(&dc[L.c_zpos[i]], sc[L.c_zpos[i]]);
        }
    }
}

// Find or claim the slot of key kn among cap hashed slots of stride RH,
// giving up after max_probes.  A claim bumps *claimed when given.  Returns
// -1 when no slot was found (the HBM table never gets there: the host
// keeps its load at or below one half).
__device__ __forceinline__ int64_t gb_slot(uint64_t *hot, int64_t cap, int RH, uint64_t kn, uint64_t empty,
                                           int64_t max_probes, uint64_t *claimed) {
    uint64_t mask = (uint64_t)cap - 1;
    uint64_t s = vg_key_hash(kn) & mask;
    for (int64_t probes = 0; probes < max_probes; probes++) {
        uint64_t *kp = &hot[s * RH];
        uint64_t k = vg_atomic_load(kp);
        if (k == kn) return (int64_t)s;
        if (k == empty) {
            uint64_t old = atomicCAS((unsigned long long *)kp, (unsigned long long)empty, (unsigned long long)kn);
            if (old == empty) {
                if (claimed) atomicAdd((unsigned long long *)claimed, 1ull);
                return (int64_t)s;
            }
            if (old == kn) return (int64_t)s;
        }
        s = (s + 1) & mask;
    }
    return -1;
}

#define GB_LDS_PROBES 32

// map a row's key to a slot of the HBM table (reserved slots cap / cap+1
// for NA and the EMPTY-valued int64); NaN keys become unique synthetic
// keys so every NaN row forms its own group
__device__ __forceinline__ int64_t gb_global_slot(const GbLayout &L, uint64_t *hot, int64_t cap, int cls,
                                                  uint64_t kn, uint64_t pos, uint64_t *ctl) {
    if (cls == VG_KEY_NA) return cap;
    if (cls == VG_KEY_SPECIAL) return cap + 1;
    uint64_t empty = L.key_is_double ? VG_EMPTY_F64 : VG_EMPTY_I64;
    if (cls == VG_KEY_NAN) kn = 0x7FF8000000000000ull | (pos & 0x0007FFFFFFFFFFFFull);
    int64_t s = gb_slot(hot, cap, L.RH, kn, empty, cap, &ctl[0]);
    if (s < 0) atomicOr((unsigned long long *)&ctl[1], 1ull);
    return s;
}

__device__ __forceinline__ void gb_note_zero_key(uint64_t *ctl, uint64_t raw_bits, uint64_t pos) {
    gb_atomic_min(&ctl[(raw_bits >> 63) ? 2 : 3], pos);
}

// ---------------------------------------------------------------------------
// table init
// ---------------------------------------------------------------------------

__global__ void k_gb_init(uint64_t *hot, uint64_t *cold, int64_t n_slots, int RH, int RC,
                          uint64_t hot_ones, uint64_t cold_ones, uint64_t empty) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < n_slots; i += stride) {
        uint64_t *h = hot + i * RH;
        h[0] = empty;
        for (int w = 1; w < RH; w++) h[w] = ((hot_ones >> w) & 1) ? ~0ull : 0ull;
        uint64_t *c = cold + i * RC;
        for (int w = 0; w < RC; w++) c[w] = ((cold_ones >> w) & 1) ? ~0ull : 0ull;
    }
}

// ---------------------------------------------------------------------------
// feed kernels
// ---------------------------------------------------------------------------

struct InSet {
    vgpu_col c[VGPU_MAX_INPUTS];
};

struct FeedArgs {
    int64_t n_logical;
    const int32_t *sel;
    int64_t pos_base;
    vgpu_col key;
    uint64_t *hot;
    uint64_t *cold;
    uint64_t *ctl;
    int64_t cap;
    int32_t lds_slots;     // 0 = no LDS privatization
    int32_t vm_lds;        // bytes of VM registers at the start of dyn LDS
};

// The LDS-private table holds S hashed records plus two reserved ones (NA
// key, EMPTY-valued int64 key), like the HBM table's slots cap / cap+1, so
// that no row class funnels every block into one HBM record.
__device__ __forceinline__ void gb_lds_init(const GbLayout &L, uint64_t *lhot, uint64_t *lcold, int S,
                                            uint64_t empty) {
    for (int s = threadIdx.x; s < S + 2; s += GB_THREADS) {
        uint64_t *h = lhot + (size_t)s * L.RH;
        h[0] = empty;
        for (int w = 1; w < L.RH; w++) h[w] = ((L.hot_init_ones >> w) & 1) ? ~0ull : 0ull;
        uint64_t *c = lcold + (size_t)s * L.RC;
        for (int w = 0; w < L.RC; w++) c[w] = ((L.cold_init_ones >> w) & 1) ? ~0ull : 0ull;
    }
    __syncthreads();
}

// LDS record of a key: reserved records for NA / EMPTY-valued keys; -1 for
// NaN keys (each its own group, kept in HBM) or an exhausted probe window
__device__ __forceinline__ int64_t gb_lds_slot(uint64_t *lhot, int S, int RH, int cls, uint64_t kn, uint64_t empty) {
    if (cls == VG_KEY_NA) return S;
    if (cls == VG_KEY_SPECIAL) return S + 1;
    if (cls == VG_KEY_NAN) return -1;
    return gb_slot(lhot, S, RH, kn, empty, GB_LDS_PROBES, nullptr);
}

__device__ __forceinline__ void gb_lds_merge(const GbLayout &L, const uint64_t *lhot, const uint64_t *lcold, int S,
                                             uint64_t empty, uint64_t *hot, uint64_t *cold, int64_t cap,
                                             uint64_t *ctl) {
    __syncthreads();
    for (int s = threadIdx.x; s < S + 2; s += GB_THREADS) {
        const uint64_t *h = lhot + (size_t)s * L.RH;
        int64_t gs;
        if (s < S) {
            if (h[0] == empty) continue;
            gs = gb_global_slot(L, hot, cap, VG_KEY_REGULAR, h[0], 0, ctl);
        } else {
            if (h[1] == 0) continue;
            gs = cap + (s - S);
        }
        if (gs >= 0) gb_merge(L, hot + (size_t)gs * L.RH, cold + (size_t)gs * L.RC, h, lcold + (size_t)s * L.RC);
    }
}

// all words of one row into one record (inlined separately for the LDS and
// the HBM record so each call site keeps its address space)
__device__ __forceinline__ void gb_apply(const GbLayout &L, const InSet &ins, int64_t row, uint64_t pos,
                                         uint64_t *h, uint64_t *c) {
    gb_update_row(L, h, pos);
    for (int i = 0; i < L.n_inputs; i++) {
        InVal v;
        v.valid = vg_valid(ins.c[i].validity, row);
        v.bits = v.valid ? vg_load_bits(ins.c[i], row) : 0;
        gb_update_input(L, h, c, i, pos, v);
    }
}

__global__ void __launch_bounds__(GB_THREADS)
k_gb_feed(GbLayout L, FeedArgs A, InSet ins, PredInfo pi, vgpu_prog prog, ColSet pcs) {
    extern __shared__ uint8_t dyn[];
    const int tid = threadIdx.x;
    uint64_t *regs = reinterpret_cast<uint64_t *>(dyn);
    uint8_t *regv = dyn + (size_t)prog.n_insn * GB_THREADS * 8;
    uint64_t *lhot = reinterpret_cast<uint64_t *>(dyn + A.vm_lds);
    const int S = A.lds_slots;
    uint64_t *lcold = lhot + (size_t)(S + 2) * L.RH;
    const uint64_t empty = L.key_is_double ? VG_EMPTY_F64 : VG_EMPTY_I64;
    if (S) gb_lds_init(L, lhot, lcold, S, empty);
    const int64_t stride = (int64_t)gridDim.x * GB_THREADS;
    for (int64_t li = (int64_t)blockIdx.x * GB_THREADS + tid; li < A.n_logical; li += stride) {
        int64_t row = A.sel ? (int64_t)A.sel[li] : li;
        if (!vg_pred_row(pi, prog, pcs, row, regs, regv, GB_THREADS, tid)) continue;
        uint64_t pos = (uint64_t)(A.pos_base + li);
        uint64_t kn = 0;
        int cls = VG_KEY_REGULAR;
        if (!L.key_none) {
            cls = vg_key_class(A.key, row, kn);
            if (L.key_is_double && cls == VG_KEY_REGULAR && kn == 0)
                gb_note_zero_key(A.ctl, vg_load_bits(A.key, row), pos);
        }
        int64_t ls = S ? gb_lds_slot(lhot, S, L.RH, cls, kn, empty) : -1;
        if (ls >= 0) {
            gb_apply(L, ins, row, pos, lhot + (size_t)ls * L.RH, lcold + (size_t)ls * L.RC);
        } else {
            int64_t gs = gb_global_slot(L, A.hot, A.cap, cls, kn, pos, A.ctl);
            if (gs >= 0) gb_apply(L, ins, row, pos, A.hot + (size_t)gs * L.RH, A.cold + (size_t)gs * L.RC);
        }
    }
    if (S) gb_lds_merge(L, lhot, lcold, S, empty, A.hot, A.cold, A.cap, A.ctl);
}

// fused probe -> aggregate ----------------------------------------------------

struct JoinFeedArgs {
    int64_t n_logical;
    const int32_t *sel;
    int64_t pos_base;
    int32_t has_join;
    int32_t join_kind;
    int32_t key_side;
    int32_t key_col;
    int32_t probe_key_col;
    int32_t n_inputs;
    int32_t need_rows;                 // some build column is read through rows[]
    int8_t bpay[VGPU_MAX_COLS];        // payload slot of a build column, -1 = not embedded
    vgpu_input in[VGPU_MAX_INPUTS];
    uint64_t *hot;
    uint64_t *cold;
    uint64_t *ctl;
    int64_t cap;
    int32_t lds_slots;
    int32_t vm_lds;
};

struct ColPair {
    ColSet probe;
    ColSet build;
};

// the build side of one output pair: table entry (nullptr = unmatched),
// segment offset, match index and build row (-1 when not needed)
struct BRef {
    const uint64_t *e;
    uint64_t segw;
    uint32_t off;
    uint32_t j;
    int64_t br;
};

// value of (side, col) at the current pair; build side is NA when unmatched
__device__ __forceinline__ InVal jf_col(const JoinFeedArgs &A, const ColPair &cp, const JView &jv, int side,
                                        int col, int64_t pr, const BRef &b) {
    InVal v;
    if (side == 0) {
        const vgpu_col &c = cp.probe.c[col];
        v.valid = vg_valid(c.validity, pr);
        v.bits = v.valid ? vg_load_bits(c, pr) : 0;
        return v;
    }
    if (!b.e) {
        v.valid = false;
        v.bits = 0;
        return v;
    }
    int p = A.bpay[col];
    if (p >= 0) {
        if (b.j == 0) {
            v.valid = (b.segw >> (24 + p)) & 1;
            v.bits = b.e[2 + p];
        } else {
            v.valid = jv.cval[p][b.off + b.j];
            v.bits = jv.cpay[p][b.off + b.j];
        }
        return v;
    }
    const vgpu_col &c = cp.build.c[col];
    v.valid = vg_valid(c.validity, b.br);
    v.bits = v.valid ? vg_load_bits(c, b.br) : 0;
    return v;
}

__device__ __forceinline__ int jf_type(const ColPair &cp, int side, int col) {
    return side == 0 ? cp.probe.c[col].type : cp.build.c[col].type;
}

// product in vectra's arithmetic: common type (bool -> int64), NA if either
// side is NA (src/scalar_ops.c:11-66)
__device__ __forceinline__ InVal jf_mul(InVal a, int ta, InVal b, int tb) {
    InVal r;
    r.valid = a.valid && b.valid;
    r.bits = 0;
    if (!r.valid) return r;
    if (ta == VGPU_DOUBLE || tb == VGPU_DOUBLE) {
        double x = ta == VGPU_DOUBLE ? vg_bits_to_f64(a.bits) : (double)(int64_t)a.bits;
        double y = tb == VGPU_DOUBLE ? vg_bits_to_f64(b.bits) : (double)(int64_t)b.bits;
        r.bits = vg_f64_to_bits(x * y);
    } else {
        r.bits = a.bits * b.bits;
    }
    return r;
}

__device__ __forceinline__ void jf_apply(const GbLayout &L, const JoinFeedArgs &A, const ColPair &cp,
                                         const JView &jv, int64_t pr, const BRef &b, uint64_t pos, uint64_t *h,
                                         uint64_t *c) {
    gb_update_row(L, h, pos);
    for (int i = 0; i < A.n_inputs; i++) {
        const vgpu_input &in = A.in[i];
        InVal v = jf_col(A, cp, jv, in.side_a, in.col_a, pr, b);
        if (in.op == 1) {
            InVal w = jf_col(A, cp, jv, in.side_b, in.col_b, pr, b);
            v = jf_mul(v, jf_type(cp, in.side_a, in.col_a), w, jf_type(cp, in.side_b, in.col_b));
        }
        gb_update_input(L, h, c, i, pos, v);
    }
}

__device__ __forceinline__ void jf_emit(const GbLayout &L, const JoinFeedArgs &A, const ColPair &cp,
                                        const JView &jv, int64_t pr, const BRef &b, uint64_t pos,
                                        uint64_t *lhot, uint64_t *lcold, int S, uint64_t empty) {
    uint64_t kn = 0;
    int cls = VG_KEY_REGULAR;
    if (!L.key_none) {
        InVal kv = jf_col(A, cp, jv, A.key_side, A.key_col, pr, b);
        cls = vg_key_class_bits(jf_type(cp, A.key_side, A.key_col), kv.valid, kv.bits, kn);
        if (L.key_is_double && cls == VG_KEY_REGULAR && kn == 0) gb_note_zero_key(A.ctl, kv.bits, pos);
    }
    int64_t ls = S ? gb_lds_slot(lhot, S, L.RH, cls, kn, empty) : -1;
    if (ls >= 0) {
        jf_apply(L, A, cp, jv, pr, b, pos, lhot + (size_t)ls * L.RH, lcold + (size_t)ls * L.RC);
    } else {
        int64_t gs = gb_global_slot(L, A.hot, A.cap, cls, kn, pos, A.ctl);
        if (gs >= 0) jf_apply(L, A, cp, jv, pr, b, pos, A.hot + (size_t)gs * L.RH, A.cold + (size_t)gs * L.RC);
    }
}

// Output position of a join pair, order-preserving within one feed call:
// matched pairs (probe row li, j-th match) precede the batch's unmatched
// tail, as in the reference's left join output (src/join.c:439-448).
__device__ __forceinline__ uint64_t jf_pos(uint64_t base, int64_t li, uint32_t j, bool tail) {
    uint32_t jj = j < 0xFFFu ? j : 0xFFFu;
    return base + ((tail ? (1ull << 43) : 0ull) | ((uint64_t)li << 12) | jj);
}

__global__ void __launch_bounds__(GB_THREADS)
k_gb_feed_join(GbLayout L, JoinFeedArgs A, ColPair cp, JView jv, PredInfo pi, vgpu_prog prog) {
    extern __shared__ uint8_t dyn[];
    const int tid = threadIdx.x;
    uint64_t *regs = reinterpret_cast<uint64_t *>(dyn);
    uint8_t *regv = dyn + (size_t)prog.n_insn * GB_THREADS * 8;
    uint64_t *lhot = reinterpret_cast<uint64_t *>(dyn + A.vm_lds);
    const int S = A.lds_slots;
    uint64_t *lcold = lhot + (size_t)(S + 2) * L.RH;
    const uint64_t empty = L.key_is_double ? VG_EMPTY_F64 : VG_EMPTY_I64;
    if (S) gb_lds_init(L, lhot, lcold, S, empty);
    const int64_t stride = (int64_t)gridDim.x * GB_THREADS;
    for (int64_t li = (int64_t)blockIdx.x * GB_THREADS + tid; li < A.n_logical; li += stride) {
        int64_t pr = A.sel ? (int64_t)A.sel[li] : li;
        if (!vg_pred_row(pi, prog, cp.probe, pr, regs, regv, GB_THREADS, tid)) continue;
        BRef b;
        b.e = nullptr;
        b.segw = 0;
        b.off = 0;
        b.j = 0;
        b.br = -1;
        if (!A.has_join) {
            jf_emit(L, A, cp, jv, pr, b, (uint64_t)(A.pos_base + li), lhot, lcold, S, empty);
            continue;
        }
        const uint64_t *e = jt_find(jv, cp.probe.c[A.probe_key_col], pr);
        uint32_t cnt = 0;
        if (e) {
            b.segw = e[1];
            cnt = jt_count(jv, b.segw);
            b.off = (uint32_t)(b.segw >> 32);
        }
        for (uint32_t j = 0; j < cnt; j++) {
            b.e = e;
            b.j = j;
            b.br = A.need_rows ? (int64_t)jv.rows[b.off + j] : -1;
            jf_emit(L, A, cp, jv, pr, b, jf_pos((uint64_t)A.pos_base, li, j, false), lhot, lcold, S, empty);
        }
        if (cnt == 0 && A.join_kind == VGPU_JOIN_LEFT) {
            b.e = nullptr;
            jf_emit(L, A, cp, jv, pr, b, jf_pos((uint64_t)A.pos_base, li, 0, true), lhot, lcold, S, empty);
        }
    }
    if (S) gb_lds_merge(L, lhot, lcold, S, empty, A.hot, A.cold, A.cap, A.ctl);
}

// ---------------------------------------------------------------------------
// pipelined probe -> aggregate for the star shape
// ---------------------------------------------------------------------------
//
// The common join-then-aggregate query reads, per probe row, one probe key,
// at most two probe-side values and build-side values that ride in the
// join table's entries (vgpu_join_build_payload).  The generic kernel walks
// one dependent chain per row (key load -> entry load -> LDS update), which
// leaves the wave waiting on memory; this one gives every thread ITEMS rows
// at a time and issues all of their key, value and entry loads before it
// consumes any, so ITEMS table probes are in flight per thread.

struct StarArgs {
    int64_t n_logical;
    const int32_t *sel;
    int64_t pos_base;
    vgpu_col pkey;                 // probe key column
    vgpu_col pval[2];              // probe-side value columns (n_pval of them)
    int32_t n_pval;
    int32_t join_kind;
    int32_t key_src;               // 0 = probe column pval-free key (kcol), 1 = payload slot kslot, 2 = none
    int32_t kslot;
    vgpu_col kcol;                 // probe-side group key (key_src 0)
    int32_t n_inputs;
    // input i: a = probe value pa[i] (-1 none), b = payload slot pb[i] (-1 none); product when both
    int32_t pa[VGPU_MAX_INPUTS];
    int32_t pb[VGPU_MAX_INPUTS];
    int32_t type[VGPU_MAX_INPUTS]; // result type of input i
    int32_t ta[VGPU_MAX_INPUTS];   // type of the probe operand
    int32_t tb[VGPU_MAX_INPUTS];   // type of the payload operand
    int32_t ktype;                 // group key type
    uint64_t *hot;
    uint64_t *cold;
    uint64_t *ctl;
    int64_t cap;
    int32_t lds_slots;
    int32_t vm_lds;
};

template <int ITEMS>
__global__ void __launch_bounds__(GB_THREADS)
k_gb_star(GbLayout L, StarArgs A, JView jv, PredInfo pi, ColSet pcs) {
    extern __shared__ uint8_t dyn[];
    const int tid = threadIdx.x;
    uint64_t *lhot = reinterpret_cast<uint64_t *>(dyn + A.vm_lds);
    const int S = A.lds_slots;
    uint64_t *lcold = lhot + (size_t)(S + 2) * L.RH;
    const uint64_t empty_g = L.key_is_double ? VG_EMPTY_F64 : VG_EMPTY_I64;
    const uint64_t empty_j = jv.key_is_double ? VG_EMPTY_F64 : VG_EMPTY_I64;
    if (S) gb_lds_init(L, lhot, lcold, S, empty_g);
    const uint64_t jmask = (uint64_t)jv.cap - 1;
    const int64_t step = (int64_t)gridDim.x * GB_THREADS * ITEMS;
    for (int64_t base = (int64_t)blockIdx.x * GB_THREADS * ITEMS; base < A.n_logical; base += step) {
        int64_t li[ITEMS], pr[ITEMS];
        bool live[ITEMS];
        uint64_t kn[ITEMS], pv0[ITEMS], pv1[ITEMS];
        bool pv0ok[ITEMS], pv1ok[ITEMS], kfound[ITEMS];
        uint64_t e0[ITEMS], e1[ITEMS], e2[ITEMS], e3[ITEMS], slot[ITEMS];
        uint64_t gk[ITEMS];
        bool gkok[ITEMS];
        // phase 1: probe-side loads
#pragma unroll
        for (int i = 0; i < ITEMS; i++) {
            li[i] = base + (int64_t)i * GB_THREADS + tid;
            live[i] = li[i] < A.n_logical;
            pr[i] = live[i] ? (A.sel ? (int64_t)A.sel[li[i]] : li[i]) : 0;
            if (live[i] && pi.present) live[i] = vg_fast_pred(pi, pcs, pr[i]);
            int cls = VG_KEY_NA;
            kn[i] = 0;
            if (live[i]) cls = vg_key_class(A.pkey, pr[i], kn[i]);
            kfound[i] = live[i] && (cls == VG_KEY_REGULAR || cls == VG_KEY_SPECIAL);
            slot[i] = cls == VG_KEY_SPECIAL ? (uint64_t)jv.cap : (vg_key_hash(kn[i]) & jmask);
            pv0ok[i] = pv1ok[i] = false;
            pv0[i] = pv1[i] = 0;
            if (live[i] && A.n_pval > 0) {
                pv0ok[i] = vg_valid(A.pval[0].validity, pr[i]);
                pv0[i] = pv0ok[i] ? vg_load_bits(A.pval[0], pr[i]) : 0;
            }
            if (live[i] && A.n_pval > 1) {
                pv1ok[i] = vg_valid(A.pval[1].validity, pr[i]);
                pv1[i] = pv1ok[i] ? vg_load_bits(A.pval[1], pr[i]) : 0;
            }
            gkok[i] = false;
            gk[i] = 0;
            if (live[i] && A.key_src == 0) {
                gkok[i] = vg_valid(A.kcol.validity, pr[i]);
                gk[i] = gkok[i] ? vg_load_bits(A.kcol, pr[i]) : 0;
            }
        }
        // phase 2: first table probe of every row, all loads in flight
#pragma unroll
        for (int i = 0; i < ITEMS; i++) {
            e0[i] = e1[i] = e2[i] = e3[i] = 0;
            if (kfound[i]) {
                const uint64_t *e = jv.ent + slot[i] * jv.E;
                e0[i] = e[0];
                e1[i] = e[1];
                e2[i] = e[2];
                e3[i] = jv.E > 3 ? e[3] : 0;
            }
        }
        // phase 3: finish collisions (rare at load <= 1/2), then aggregate
#pragma unroll
        for (int i = 0; i < ITEMS; i++) {
            if (!live[i]) continue;
            uint32_t cnt = 0;
            if (kfound[i]) {
                if (slot[i] != (uint64_t)jv.cap) {
                    int64_t probes = 0;
                    while (e0[i] != kn[i] && e0[i] != empty_j && probes < jv.cap) {
                        slot[i] = (slot[i] + 1) & jmask;
                        const uint64_t *e = jv.ent + slot[i] * jv.E;
                        e0[i] = e[0];
                        e1[i] = e[1];
                        e2[i] = e[2];
                        e3[i] = jv.E > 3 ? e[3] : 0;
                        probes++;
                    }
                    if (e0[i] == kn[i]) cnt = jt_count(jv, e1[i]);
                } else {
                    cnt = jt_count(jv, e1[i]);
                }
            }
            uint32_t off = (uint32_t)(e1[i] >> 32);
            for (uint32_t j = 0; j < cnt || (j == 0 && cnt == 0 && A.join_kind == VGPU_JOIN_LEFT); j++) {
                bool matched = cnt > 0;
                // payload values of this pair
                uint64_t pay0 = 0, pay1 = 0;
                bool pok0 = false, pok1 = false;
                if (matched) {
                    if (j == 0) {
                        pay0 = e2[i]; pok0 = (e1[i] >> 24) & 1;
                        pay1 = e3[i]; pok1 = (e1[i] >> 25) & 1;
                    } else {
                        pay0 = jv.cpay[0][off + j]; pok0 = jv.cval[0][off + j];
                        if (jv.n_payload > 1) { pay1 = jv.cpay[1][off + j]; pok1 = jv.cval[1][off + j]; }
                    }
                }
                uint64_t pos = jf_pos((uint64_t)A.pos_base, li[i], j, !matched);
                // group key
                uint64_t gkn = 0;
                int gcls = VG_KEY_REGULAR;
                if (A.key_src == 0) {
                    gcls = vg_key_class_bits(A.ktype, gkok[i], gk[i], gkn);
                } else if (A.key_src == 1) {
                    bool ok = A.kslot == 0 ? pok0 : pok1;
                    gcls = vg_key_class_bits(A.ktype, ok, A.kslot == 0 ? pay0 : pay1, gkn);
                }
                if (L.key_is_double && gcls == VG_KEY_REGULAR && gkn == 0) {
                    uint64_t raw = A.key_src == 0 ? gk[i] : (A.kslot == 0 ? pay0 : pay1);
                    gb_note_zero_key(A.ctl, raw, pos);
                }
                int64_t ls = S ? gb_lds_slot(lhot, S, L.RH, gcls, gkn, empty_g) : -1;
                // inputs
#define STAR_APPLY(HP, CP)                                                                          \
                do {                                                                                \
                    gb_update_row(L, HP, pos);                                                      \
                    for (int q = 0; q < A.n_inputs; q++) {                                          \
                        InVal a, b, v;                                                              \
                        a.valid = A.pa[q] == 0 ? pv0ok[i] : pv1ok[i];                               \
                        a.bits = A.pa[q] == 0 ? pv0[i] : pv1[i];                                    \
                        b.valid = A.pb[q] == 0 ? pok0 : pok1;                                       \
                        b.bits = A.pb[q] == 0 ? pay0 : pay1;                                        \
                        if (A.pb[q] < 0) v = a;                                                     \
                        else if (A.pa[q] < 0) v = b;                                                \
                        else v = jf_mul(a, A.ta[q], b, A.tb[q]);                                    \
                        if (A.pb[q] >= 0 && !matched) v.valid = false;                              \
                        gb_update_input(L, HP, CP, q, pos, v);                                      \
                    }                                                                               \
                } while (0)
                // one expansion per address space: LDS record or HBM record
                if (ls >= 0) {
                    uint64_t *h = lhot + (size_t)ls * L.RH, *c = lcold + (size_t)ls * L.RC;
                    STAR_APPLY(h, c);
                } else {
                    int64_t gs = gb_global_slot(L, A.hot, A.cap, gcls, gkn, pos, A.ctl);
                    if (gs >= 0) {
                        uint64_t *h = A.hot + (size_t)gs * L.RH, *c = A.cold + (size_t)gs * L.RC;
                        STAR_APPLY(h, c);
                    }
                }
#undef STAR_APPLY
                if (!matched) break;
            }
        }
    }
    if (S) gb_lds_merge(L, lhot, lcold, S, empty_g, A.hot, A.cold, A.cap, A.ctl);
}

// ---------------------------------------------------------------------------
// finish: compact, order, emit
// ---------------------------------------------------------------------------

// collects occupied slots with their sort keys; the NA group (slot cap) is
// kept out of the ORDER_KEY sort and appended last by the host
__global__ void k_gb_compact(GbLayout L, const uint64_t *hot, int64_t cap, int order,
                             uint64_t *ctl, uint64_t *skeys, uint32_t *svals) {
    int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t empty = L.key_is_double ? VG_EMPTY_F64 : VG_EMPTY_I64;
    bool take = false;
    uint64_t sk = 0;
    if (s < cap + 2) {
        const uint64_t *h = hot + s * L.RH;
        if (s < cap) take = h[0] != empty;
        else take = h[1] != 0;
        if (s == cap) ctl[8] = h[1];                                      // rows of the NA group
        if (take && order == VGPU_ORDER_KEY && s == cap) take = false;   // NA: appended last
        if (take) {
            if (order == VGPU_ORDER_FIRST_SEEN) sk = h[L.w_fpos];
            else if (s == cap + 1) sk = 0;                                // INT64_MIN: smallest
            else if (L.key_is_double) sk = vg_f64_ordered(h[0]);          // NaN groups sort above +Inf by position
            else sk = vg_i64_ordered(h[0]);
        }
    }
    uint64_t m = vg_ballot(take);
    if (!m) return;
    uint64_t base = 0;
    int leader = __ffsll((unsigned long long)m) - 1;
    if (vg_lane() == leader) base = atomicAdd((unsigned long long *)&ctl[5], (unsigned long long)__popcll(m));
    base = __shfl(base, leader, 64);
    if (take) {
        uint64_t idx = base + vg_lane_prefix(m);
        skeys[idx] = sk;
        svals[idx] = (uint32_t)s;
        atomicOr((unsigned long long *)&ctl[6], (unsigned long long)sk);
        atomicAnd((unsigned long long *)&ctl[7], (unsigned long long)sk);
    }
}

struct EmitOut {
    void *keys;
    uint64_t *kvalid;            // bitmap words
    double *vals[VGPU_MAX_AGGS];
    uint64_t *valid[VGPU_MAX_AGGS];
};

__device__ __forceinline__ void gb_finish_agg(const GbLayout &L, const vgpu_agg &ag, const uint64_t *h,
                                              const uint64_t *c, double &out, bool &ok, bool &warn) {
    warn = false;
    uint64_t n = h[1];
    if (ag.kind == VGPU_AGG_COUNT_STAR) { ok = true; out = (double)n; return; }
    int i = ag.input;
    uint64_t na = c[L.c_na[i]];
    uint64_t cnt = n - na;
    if (ag.kind == VGPU_AGG_COUNT) { ok = true; out = (double)cnt; return; }
    if (na && !ag.na_rm) { ok = false; out = 0; return; }   // NA poisons (src/agg_ops.c:170-175)
    int t = L.in_type[i];
    ok = true;
    if (ag.kind == VGPU_AGG_SUM || ag.kind == VGPU_AGG_MEAN) {
        uint64_t sw = h[L.w_sum[i]];
        double sum = t == VGPU_DOUBLE ? vg_bits_to_f64(sw) : t == VGPU_INT64 ? (double)(int64_t)sw : (double)sw;
        if (ag.kind == VGPU_AGG_SUM) { out = sum; return; }
        if (cnt == 0) { out = __builtin_nan(""); return; }
        out = sum / (double)cnt;
        return;
    }
    bool is_min = ag.kind == VGPU_AGG_MIN;
    if (t == VGPU_BOOL || cnt == 0) {
        out = is_min ? __builtin_inf() : -__builtin_inf();
        warn = true;
        return;
    }
    uint64_t o = h[is_min ? L.w_min[i] : L.w_max[i]];
    if (t == VGPU_DOUBLE) {
        uint64_t first_valid = L.w_fvpos[i] >= 0 ? h[L.w_fvpos[i]] : h[L.w_fpos];
        uint64_t nanpos = c[L.c_nan[i]];
        if (nanpos != GB_NOPOS && nanpos == first_valid) { out = __builtin_nan(""); return; }
        if ((is_min && o == GB_NOPOS) || (!is_min && o == 0)) { out = __builtin_nan(""); return; }
        uint64_t bits = vg_f64_unordered(o);
        if (bits == 0 && c[L.c_zneg[i]] < c[L.c_zpos[i]]) bits = 0x8000000000000000ull;
        out = vg_bits_to_f64(bits);
    } else {
        out = (double)(int64_t)(o ^ 0x8000000000000000ull);
    }
}

__global__ void k_gb_emit(GbLayout L, GbAggs AG, const uint64_t *hot, const uint64_t *cold, int64_t cap,
                          const uint32_t *order_slots, int64_t n_sorted, int na_last, int64_t n_out,
                          int key_width, EmitOut O, uint64_t *ctl) {
    int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool in = g < n_out;
    int64_t s = 0;
    if (in) s = (na_last && g == n_sorted) ? cap : (int64_t)order_slots[g];
    const uint64_t *h = hot + s * L.RH;
    const uint64_t *c = cold + s * L.RC;
    bool kvalid = in && s != cap;
    if (in && !L.key_none) {
        uint64_t kb = 0;
        if (s == cap + 1) kb = VG_EMPTY_I64;
        else if (s < cap) {
            kb = h[0];
            if (L.key_is_double) {
                if ((kb & 0x7FFFFFFFFFFFFFFFull) > 0x7FF0000000000000ull) kb = 0x7FF8000000000000ull;
                else if (kb == 0 && ctl[2] < ctl[3]) kb = 0x8000000000000000ull;
            }
        }
        if (key_width == 1) static_cast<uint8_t *>(O.keys)[g] = (uint8_t)kb;
        else static_cast<uint64_t *>(O.keys)[g] = kb;
    }
    uint64_t kw = vg_ballot(kvalid);
    if (vg_lane() == 0 && O.kvalid && (g >> 6) * 64 < n_out) O.kvalid[g >> 6] = kw;
    int warns = 0;
    for (int a = 0; a < AG.n; a++) {
        double v = 0;
        bool ok = false, warn = false;
        if (in) gb_finish_agg(L, AG.a[a], h, c, v, ok, warn);
        if (in) O.vals[a][g] = ok ? v : 0.0;
        warns += warn && ok;
        uint64_t w = vg_ballot(in && ok);
        if (vg_lane() == 0 && (g >> 6) * 64 < n_out) O.valid[a][g >> 6] = w;
    }
    if (warns) atomicAdd((unsigned long long *)&ctl[4], (unsigned long long)warns);
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------

static int gb_layout(vgpu_groupby *g, int key_type, int n_inputs, const int32_t *types, int n_aggs,
                     const vgpu_agg *aggs) {
    GbLayout &L = g->L;
    memset(&L, 0, sizeof L);
    L.n_inputs = n_inputs;
    L.key_is_double = key_type == VGPU_DOUBLE;
    L.key_none = key_type == VGPU_KEY_NONE;
    bool need_sum[VGPU_MAX_INPUTS] = {}, need_min[VGPU_MAX_INPUTS] = {}, need_max[VGPU_MAX_INPUTS] = {};
    bool rm_mm[VGPU_MAX_INPUTS] = {};
    bool any_dbl_mm = false;
    for (int a = 0; a < n_aggs; a++) {
        const vgpu_agg &ag = aggs[a];
        if (ag.kind < VGPU_AGG_COUNT || ag.kind > VGPU_AGG_MAX) {
            vg_set_error("groupby: unknown aggregate kind %d", ag.kind);
            return -1;
        }
        if (ag.kind == VGPU_AGG_COUNT_STAR) continue;
        if (ag.input < 0 || ag.input >= n_inputs) {
            vg_set_error("groupby: aggregate %d reads input %d of %d", a, ag.input, n_inputs);
            return -1;
        }
        int i = ag.input;
        if (ag.kind == VGPU_AGG_SUM || ag.kind == VGPU_AGG_MEAN) need_sum[i] = true;
        if (ag.kind == VGPU_AGG_MIN) need_min[i] = true;
        if (ag.kind == VGPU_AGG_MAX) need_max[i] = true;
        if ((ag.kind == VGPU_AGG_MIN || ag.kind == VGPU_AGG_MAX) && types[i] == VGPU_DOUBLE) {
            any_dbl_mm = true;
            if (ag.na_rm) rm_mm[i] = true;
        }
    }
    int h = 2, c = 0;
    uint64_t hot_ones = 0, cold_ones = 0;
    L.w_fpos = -1;
    if (g->order == VGPU_ORDER_FIRST_SEEN || any_dbl_mm) { L.w_fpos = h; hot_ones |= 1ull << h; h++; }
    for (int i = 0; i < n_inputs; i++) {
        if (types[i] != VGPU_INT64 && types[i] != VGPU_DOUBLE && types[i] != VGPU_BOOL) {
            vg_set_error("groupby: input %d has unsupported type %d", i, types[i]);
            return -1;
        }
        L.in_type[i] = types[i];
        L.w_sum[i] = need_sum[i] ? h++ : -1;
        L.w_min[i] = -1;
        L.w_max[i] = -1;
        if (types[i] != VGPU_BOOL) {
            if (need_min[i]) { L.w_min[i] = h; hot_ones |= 1ull << h; h++; }
            if (need_max[i]) L.w_max[i] = h++;
        }
        L.w_fvpos[i] = -1;
        if (rm_mm[i]) { L.w_fvpos[i] = h; hot_ones |= 1ull << h; h++; }
        L.c_na[i] = c++;
        L.c_nan[i] = L.c_zneg[i] = L.c_zpos[i] = -1;
        if (types[i] == VGPU_DOUBLE && (need_min[i] || need_max[i])) {
            L.c_nan[i] = c; cold_ones |= 1ull << c; c++;
            L.c_zneg[i] = c; cold_ones |= 1ull << c; c++;
            L.c_zpos[i] = c; cold_ones |= 1ull << c; c++;
        }
    }
    if (c == 0) c = 1;
    if (h > GB_MAXW || c > GB_MAXW) {
        vg_set_error("groupby: record too wide (%d hot, %d cold words)", h, c);
        return -1;
    }
    L.RH = h;
    L.RC = c;
    L.hot_init_ones = hot_ones;
    L.cold_init_ones = cold_ones;
    return 0;
}

static int64_t pow2_at_least(int64_t x) {
    int64_t p = 64;
    while (p < x) p <<= 1;
    return p;
}

static int gb_alloc_table(vgpu_groupby *g, int64_t cap) {
    int64_t slots = cap + 2;
    if (g->hot.reserve((size_t)slots * g->L.RH * 8) || g->cold.reserve((size_t)slots * g->L.RC * 8)) return -1;
    uint64_t empty = g->L.key_is_double ? VG_EMPTY_F64 : VG_EMPTY_I64;
    k_gb_init<<<vg_blocks(slots, 256, 8192), 256, 0, vg_stream()>>>(g->hot.as<uint64_t>(), g->cold.as<uint64_t>(), slots,
                                                                    g->L.RH, g->L.RC, g->L.hot_init_ones,
                                                                    g->L.cold_init_ones, empty);
    VG_LAUNCH_CHECK();
    g->cap = cap;
    return 0;
}

__global__ void k_gb_rehash(GbLayout L, const uint64_t *ohot, const uint64_t *ocold, int64_t ocap,
                            uint64_t *nhot, uint64_t *ncold, int64_t ncap, uint64_t *ctl) {
    int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= ocap + 2) return;
    const uint64_t *h = ohot + s * L.RH;
    const uint64_t *c = ocold + s * L.RC;
    const uint64_t empty = L.key_is_double ? VG_EMPTY_F64 : VG_EMPTY_I64;
    int64_t d;
    if (s < ocap) {
        if (h[0] == empty) return;
        d = gb_slot(nhot, ncap, L.RH, h[0], empty, ncap, nullptr);
        if (d < 0) { atomicOr((unsigned long long *)&ctl[1], 1ull); return; }
    } else {
        d = ncap + (s - ocap);
    }
    uint64_t *dh = nhot + d * L.RH;
    uint64_t *dc = ncold + d * L.RC;
    for (int w = 1; w < L.RH; w++) dh[w] = h[w];
    for (int w = 0; w < L.RC; w++) dc[w] = c[w];
}

#define GB_CTL_WORDS 16

static int gb_read_ctl(vgpu_groupby *g, uint64_t *ctl_host) {
    VG_TRY(hipMemcpyAsync(ctl_host, g->ctl.ptr, GB_CTL_WORDS * 8, hipMemcpyDeviceToHost, vg_stream()));
    VG_TRY(hipStreamSynchronize(vg_stream()));
    return 0;
}

// control block: [0] claimed slots, [1] overflow flag, [2]/[3] first
// position of a -0 / +0 double key, [4] min/max warnings, [5] compaction
// counter, [6]/[7] OR / AND of the sort keys, [8] rows of the NA group
__global__ void k_gb_ctl_init(uint64_t *ctl, int finish_only) {
    int i = threadIdx.x;
    if (i >= GB_CTL_WORDS) return;
    if (finish_only) {
        if (i == 5 || i == 6 || i == 8) ctl[i] = 0;
        if (i == 7) ctl[i] = ~0ull;
        return;
    }
    ctl[i] = (i == 2 || i == 3 || i == 7) ? ~0ull : 0ull;
}

// grows the table so that `incoming` more groups keep the load <= 1/2
// Cardinality sketch of one feed's keys (HyperLogLog, Flajolet et al. 2007,
// with the linear-counting correction for small counts): sizes the table
// and picks the LDS-private path for a large feed when the caller gave no
// group count.  4096 registers give a relative error of about 1.6%; the
// table is sized at twice a padded estimate, and a table that still fills
// up is reported by finish as an overflow, never as a wrong result.
#define GB_SK_BITS 12
#define GB_SK_REGS (1 << GB_SK_BITS)
#define GB_SK_MIN_ROWS (1 << 20)

__global__ void __launch_bounds__(GB_THREADS)
k_gb_sketch(vgpu_col key, const int32_t *sel, int64_t n_logical, uint32_t *regs, unsigned long long *nan_rows) {
    __shared__ uint32_t r[GB_SK_REGS];
    __shared__ unsigned long long nn;
    for (int i = threadIdx.x; i < GB_SK_REGS; i += GB_THREADS) r[i] = 0;
    if (threadIdx.x == 0) nn = 0;
    __syncthreads();
    unsigned long long my_nan = 0;
    const int64_t stride = (int64_t)gridDim.x * GB_THREADS;
    for (int64_t li = (int64_t)blockIdx.x * GB_THREADS + threadIdx.x; li < n_logical; li += stride) {
        int64_t row = sel ? (int64_t)sel[li] : li;
        uint64_t kn;
        int cls = vg_key_class(key, row, kn);
        if (cls == VG_KEY_NAN) { my_nan++; continue; }
        if (cls != VG_KEY_REGULAR) continue;
        uint64_t h = vg_mix64(kn);
        uint32_t idx = (uint32_t)(h >> (64 - GB_SK_BITS));
        uint64_t rest = (h << GB_SK_BITS) | (1ull << (GB_SK_BITS - 1));
        uint32_t rank = (uint32_t)__builtin_clzll(rest) + 1;
        // repeated keys find their register already set: no atomic
        if (r[idx] < rank) atomicMax(&r[idx], rank);
    }
    if (my_nan) atomicAdd(&nn, my_nan);
    __syncthreads();
    for (int i = threadIdx.x; i < GB_SK_REGS; i += GB_THREADS)
        if (r[i]) atomicMax(&regs[i], r[i]);
    if (threadIdx.x == 0 && nn) atomicAdd(nan_rows, nn);
}

__global__ void k_gb_sketch_clear(uint32_t *regs, unsigned long long *nan_rows) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < GB_SK_REGS) regs[i] = 0;
    if (i == 0) *nan_rows = 0;
}

// estimated groups of a feed: distinct regular keys + one per NaN row + the
// two reserved groups (NA key, EMPTY-valued key)
static int gb_sketch(vgpu_groupby *g, const vgpu_col *key, const int32_t *sel, int64_t n_logical, int64_t *est) {
    size_t bytes = GB_SK_REGS * 4 + 8;
    if (g->sketch.reserve(bytes)) return -1;
    uint32_t *regs = g->sketch.as<uint32_t>();
    auto *nan_rows = reinterpret_cast<unsigned long long *>(regs + GB_SK_REGS);
    k_gb_sketch_clear<<<GB_SK_REGS / 256, 256, 0, vg_stream()>>>(regs, nan_rows);
    VG_LAUNCH_CHECK();
    k_gb_sketch<<<vg_blocks(n_logical, 65536, 2048), GB_THREADS, 0, vg_stream()>>>(*key, sel, n_logical, regs,
                                                                                     nan_rows);
    VG_LAUNCH_CHECK();
    std::vector<uint32_t> h(GB_SK_REGS + 2);
    VG_TRY(hipMemcpyAsync(h.data(), regs, bytes, hipMemcpyDeviceToHost, vg_stream()));
    VG_TRY(hipStreamSynchronize(vg_stream()));
    double m = GB_SK_REGS, z = 0;
    int zeros = 0;
    for (int i = 0; i < GB_SK_REGS; i++) {
        z += std::ldexp(1.0, -(int)h[i]);
        zeros += h[i] == 0;
    }
    double e = 0.7213 / (1.0 + 1.079 / m) * m * m / z;
    if (e <= 2.5 * m && zeros > 0) e = m * std::log(m / zeros);
    uint64_t nan;
    memcpy(&nan, &h[GB_SK_REGS], 8);
    *est = (int64_t)std::ceil(e) + (int64_t)nan + 2;
    return 0;
}

static int gb_ensure_capacity(vgpu_groupby *g, int64_t incoming) {
    uint64_t ctl[GB_CTL_WORDS];
    if (g->cap == 0) {
        int64_t want = pow2_at_least(2 * std::max<int64_t>(incoming, 32));
        if (g->hint > 0) want = std::max(want, pow2_at_least(2 * g->hint));
        return gb_alloc_table(g, want);
    }
    // a table sized for the caller's group count is trusted until finish,
    // which reports an overflow; no per-feed synchronization
    if (g->hint > 0 && g->cap >= pow2_at_least(2 * g->hint)) return 0;
    if (gb_read_ctl(g, ctl)) return -1;
    if (ctl[1]) {
        vg_set_error("groupby: hash table overflow");
        return -1;
    }
    // new groups of this feed are bounded by its rows (or its sketch)
    int64_t occupied = (int64_t)ctl[0];
    if (g->hint <= 0) g->est_groups = std::max(g->est_groups, occupied);
    int64_t need = pow2_at_least(2 * (occupied + incoming));
    if (need <= g->cap) return 0;
    DevBuf nhot, ncold;
    int64_t slots = need + 2;
    if (nhot.reserve((size_t)slots * g->L.RH * 8) || ncold.reserve((size_t)slots * g->L.RC * 8)) return -1;
    uint64_t empty = g->L.key_is_double ? VG_EMPTY_F64 : VG_EMPTY_I64;
    k_gb_init<<<vg_blocks(slots, 256, 8192), 256, 0, vg_stream()>>>(nhot.as<uint64_t>(), ncold.as<uint64_t>(), slots,
                                                                    g->L.RH, g->L.RC, g->L.hot_init_ones,
                                                                    g->L.cold_init_ones, empty);
    VG_LAUNCH_CHECK();
    k_gb_rehash<<<vg_blocks(g->cap + 2, 256), 256, 0, vg_stream()>>>(g->L, g->hot.as<uint64_t>(), g->cold.as<uint64_t>(),
                                                                      g->cap, nhot.as<uint64_t>(), ncold.as<uint64_t>(),
                                                                      need, g->ctl.as<uint64_t>());
    VG_LAUNCH_CHECK();
    VG_TRY(hipStreamSynchronize(vg_stream()));
    g->hot.release();
    g->cold.release();
    g->hot = nhot;
    g->cold = ncold;
    g->cap = need;
    return 0;
}

extern "C" vgpu_groupby *vgpu_groupby_create(int key_type, int n_inputs, const int32_t *input_types, int n_aggs,
                                             const vgpu_agg *aggs, int order, int64_t expected_groups) {
    if (vg_ensure_init()) return nullptr;
    if (key_type != VGPU_INT64 && key_type != VGPU_DOUBLE && key_type != VGPU_BOOL && key_type != VGPU_KEY_NONE) {
        vg_set_error("groupby: unsupported key type %d", key_type);
        return nullptr;
    }
    if (n_inputs < 0 || n_inputs > VGPU_MAX_INPUTS || n_aggs < 0 || n_aggs > VGPU_MAX_AGGS) {
        vg_set_error("groupby: %d inputs / %d aggregates exceed the limits (%d / %d)", n_inputs, n_aggs,
                     VGPU_MAX_INPUTS, VGPU_MAX_AGGS);
        return nullptr;
    }
    if (order != VGPU_ORDER_FIRST_SEEN && order != VGPU_ORDER_KEY) {
        vg_set_error("groupby: unknown order %d", order);
        return nullptr;
    }
    vgpu_groupby *g = new vgpu_groupby();
    g->order = key_type == VGPU_KEY_NONE ? VGPU_ORDER_FIRST_SEEN : order;
    g->hint = expected_groups;
    g->aggs.n = n_aggs;
    for (int a = 0; a < n_aggs; a++) g->aggs.a[a] = aggs[a];
    if (gb_layout(g, key_type, n_inputs, input_types, n_aggs, aggs) || g->ctl.reserve(GB_CTL_WORDS * 8)) {
        delete g;
        return nullptr;
    }
    k_gb_ctl_init<<<1, 64, 0, vg_stream()>>>(g->ctl.as<uint64_t>(), 0);
    if (hipGetLastError() != hipSuccess) {
        vg_set_error("groupby: control block init failed");
        g->ctl.release();
        delete g;
        return nullptr;
    }
    g->out_vals.resize(n_aggs);
    g->out_valid.resize(n_aggs);
    return g;
}

// LDS slots for privatization: the table must hold the expected groups at
// load <= 1/2 within the LDS budget
static int gb_lds_slots(const vgpu_groupby *g, size_t vm_bytes) {
    int64_t groups = g->L.key_none ? 1 : g->hint > 0 ? g->hint : g->est_groups;
    if (groups <= 0) return 0;
    size_t rec = (size_t)(g->L.RH + g->L.RC) * 8;
    size_t lim = (size_t)vg_tune(VG_TUNE_GB_LDS_BUDGET);
    size_t budget = lim > vm_bytes ? lim - vm_bytes : 0;
    int64_t want = pow2_at_least(2 * groups);
    if (want > vg_tune(VG_TUNE_GB_LDS_MAX_SLOTS) || (size_t)(want + 2) * rec > budget) return 0;
    return (int)want;
}

static unsigned gb_grid(int64_t n_logical, int lds_slots) {
    // privatized blocks amortize their merge over many rows
    if (lds_slots)
        return vg_blocks(n_logical, vg_tune(VG_TUNE_GB_LDS_ROWS), vg_tune(VG_TUNE_GB_LDS_MAX_BLOCKS));
    return vg_blocks(n_logical, vg_tune(VG_TUNE_GB_GLB_ROWS), vg_tune(VG_TUNE_GB_GLB_MAX_BLOCKS));
}

extern "C" int vgpu_groupby_feed(vgpu_groupby *g, const vgpu_col *key, const vgpu_col *inputs, int64_t n_rows,
                                 const int32_t *sel, int64_t n_sel, int64_t pos_base, const vgpu_prog *pred,
                                 const vgpu_col *pred_cols, int n_pred_cols) {
    if (g->finished) {
        vg_set_error("groupby: feed after finish");
        return -1;
    }
    int64_t n_logical = sel ? n_sel : n_rows;
    if (n_logical <= 0) return 0;
    if (!g->L.key_none && !key) {
        vg_set_error("groupby: key column missing");
        return -1;
    }
    if (pred && (pred->n_insn <= 0 || pred->n_insn > VGPU_MAX_INSN || n_pred_cols > VGPU_MAX_COLS)) {
        vg_set_error("groupby: invalid predicate program");
        return -1;
    }
    if (pred) {
        for (int i = 0; i < pred->n_insn; i++)
            if (pred->insn[i].op == VGPU_OP_COL && (pred->insn[i].col < 0 || pred->insn[i].col >= n_pred_cols)) {
                vg_set_error("groupby: predicate reads column slot %d of %d", pred->insn[i].col, n_pred_cols);
                return -1;
            }
    }
    for (int i = 0; i < g->L.n_inputs; i++)
        if (inputs[i].type != g->L.in_type[i]) {
            vg_set_error("groupby: input %d type %d differs from declared %d", i, inputs[i].type, g->L.in_type[i]);
            return -1;
        }
    // groups created by this feed are bounded by its rows, by the hint, or
    // (a large feed without a hint) by its sketched key count
    int64_t incoming = n_logical;
    if (g->hint > 0 && incoming > g->hint) incoming = g->hint;
    if (g->hint <= 0 && !g->L.key_none && n_logical >= GB_SK_MIN_ROWS) {
        int64_t est;
        if (gb_sketch(g, key, sel, n_logical, &est)) return -1;
        g->est_groups = std::max(g->est_groups, est);
        incoming = std::min(incoming, est + est / 4 + 4096);
    }
    if (gb_ensure_capacity(g, incoming)) return -1;

    FeedArgs A;
    memset(&A, 0, sizeof A);
    A.n_logical = n_logical;
    A.sel = sel;
    A.pos_base = pos_base;
    if (key) A.key = *key;
    A.hot = g->hot.as<uint64_t>();
    A.cold = g->cold.as<uint64_t>();
    A.ctl = g->ctl.as<uint64_t>();
    A.cap = g->cap;
    InSet ins;
    memset(&ins, 0, sizeof ins);
    for (int i = 0; i < g->L.n_inputs; i++) ins.c[i] = inputs[i];
    PredInfo pi = vg_pred_info(pred);
    vgpu_prog prog;
    memset(&prog, 0, sizeof prog);
    if (pred) prog = *pred;
    ColSet pcs;
    memset(&pcs, 0, sizeof pcs);
    for (int i = 0; i < n_pred_cols && pred_cols; i++) pcs.c[i] = pred_cols[i];
    size_t vm = (pi.present && !pi.fast) ? vg_vm_lds_bytes(pred, GB_THREADS) : 0;
    vm = (vm + 15) & ~(size_t)15;
    A.vm_lds = (int32_t)vm;
    A.lds_slots = gb_lds_slots(g, vm);
    size_t lds = vm + (A.lds_slots ? (size_t)(A.lds_slots + 2) * (g->L.RH + g->L.RC) * 8 : 0);
    vg_kt_begin();
    k_gb_feed<<<gb_grid(n_logical, A.lds_slots), GB_THREADS, lds, vg_stream()>>>(g->L, A, ins, pi, prog, pcs);
    VG_LAUNCH_CHECK();
    vg_kt_end();
    g->any_fed = 1;
    return 0;
}

// Is this feed the star shape k_gb_star handles?  (join with one or two
// embedded payload columns, no rows[] reads, a fast or absent predicate,
// inputs that are a probe column, a payload column or the product of one
// of each.)  Fills st when it is.
static bool gb_star_shape(const vgpu_groupby *g, const vgpu_feed_spec *sp, const JoinFeedArgs &A, const PredInfo &pi,
                          StarArgs *st) {
    if (!sp->join || (sp->join_kind != VGPU_JOIN_INNER && sp->join_kind != VGPU_JOIN_LEFT)) return false;
    if (sp->join->n_payload < 1 || sp->join->n_payload > 2 || sp->join->E != 4 || A.need_rows) return false;
    if (pi.present && !pi.fast) return false;
    memset(st, 0, sizeof *st);
    st->n_logical = A.n_logical;
    st->sel = A.sel;
    st->pos_base = A.pos_base;
    st->pkey = sp->probe_cols[sp->probe_key_col];
    st->join_kind = sp->join_kind;
    int pcol[2] = {-1, -1};
    auto pval_slot = [&](int col) -> int {
        for (int k = 0; k < st->n_pval; k++)
            if (pcol[k] == col) return k;
        if (st->n_pval == 2) return -1;
        pcol[st->n_pval] = col;
        st->pval[st->n_pval] = sp->probe_cols[col];
        return st->n_pval++;
    };
    if (g->L.key_none) {
        st->key_src = 2;
    } else if (sp->key_side == 0) {
        st->key_src = 0;
        st->kcol = sp->probe_cols[sp->key_col];
        st->ktype = st->kcol.type;
    } else {
        st->key_src = 1;
        st->kslot = A.bpay[sp->key_col];
        st->ktype = sp->build_cols[sp->key_col].type;
    }
    st->n_inputs = sp->n_inputs;
    for (int i = 0; i < sp->n_inputs; i++) {
        const vgpu_input &in = sp->inputs[i];
        int sa = in.side_a, ca = in.col_a, sb = in.side_b, cb = in.col_b;
        if (in.op == 1 && sa == 1 && sb == 0) {   // a * b == b * a: probe operand first
            sa = 0; ca = in.col_b; sb = 1; cb = in.col_a;
        }
        st->pa[i] = st->pb[i] = -1;
        if (in.op == 0) {
            if (sa == 0) {
                st->pa[i] = pval_slot(ca);
                if (st->pa[i] < 0) return false;
                st->ta[i] = sp->probe_cols[ca].type;
                st->type[i] = st->ta[i];
            } else {
                st->pb[i] = A.bpay[ca];
                st->tb[i] = sp->build_cols[ca].type;
                st->type[i] = st->tb[i];
            }
            continue;
        }
        if (sa != 0 || sb != 1) return false;
        st->pa[i] = pval_slot(ca);
        if (st->pa[i] < 0) return false;
        st->pb[i] = A.bpay[cb];
        st->ta[i] = sp->probe_cols[ca].type;
        st->tb[i] = sp->build_cols[cb].type;
        st->type[i] = (st->ta[i] == VGPU_DOUBLE || st->tb[i] == VGPU_DOUBLE) ? VGPU_DOUBLE : VGPU_INT64;
    }
    st->hot = A.hot;
    st->cold = A.cold;
    st->ctl = A.ctl;
    st->cap = A.cap;
    return true;
}

extern "C" int vgpu_groupby_feed_join(vgpu_groupby *g, const vgpu_feed_spec *sp) {
    if (g->finished) {
        vg_set_error("groupby: feed after finish");
        return -1;
    }
    int64_t n_logical = sp->sel ? sp->n_sel : sp->n_rows;
    if (n_logical <= 0) return 0;
    if (sp->n_probe_cols < 0 || sp->n_probe_cols > VGPU_MAX_COLS || sp->n_build_cols < 0 ||
        sp->n_build_cols > VGPU_MAX_COLS || sp->n_inputs != g->L.n_inputs) {
        vg_set_error("groupby_feed_join: column / input counts out of range");
        return -1;
    }
    auto check_ref = [&](int side, int col) -> bool {
        if (side == 0) return col >= 0 && col < sp->n_probe_cols;
        return sp->join && side == 1 && col >= 0 && col < sp->n_build_cols;
    };
    if (!g->L.key_none && !check_ref(sp->key_side, sp->key_col)) {
        vg_set_error("groupby_feed_join: key column reference out of range");
        return -1;
    }
    if (sp->join) {
        if (sp->join_kind != VGPU_JOIN_INNER && sp->join_kind != VGPU_JOIN_LEFT) {
            vg_set_error("groupby_feed_join: join kind %d not fused (inner and left only)", sp->join_kind);
            return -1;
        }
        if (!check_ref(0, sp->probe_key_col)) {
            vg_set_error("groupby_feed_join: probe key column out of range");
            return -1;
        }
        if (sp->probe_cols[sp->probe_key_col].type != sp->join->key_type) {
            vg_set_error("groupby_feed_join: probe key type %d differs from build key type %d",
                         sp->probe_cols[sp->probe_key_col].type, sp->join->key_type);
            return -1;
        }
    }
    for (int i = 0; i < sp->n_inputs; i++) {
        const vgpu_input &in = sp->inputs[i];
        if (!check_ref(in.side_a, in.col_a) || (in.op == 1 && !check_ref(in.side_b, in.col_b)) ||
            (in.op != 0 && in.op != 1)) {
            vg_set_error("groupby_feed_join: input %d references an unknown column", i);
            return -1;
        }
    }
    if (sp->pred) {
        for (int i = 0; i < sp->pred->n_insn; i++)
            if (sp->pred->insn[i].op == VGPU_OP_COL &&
                (sp->pred->insn[i].col < 0 || sp->pred->insn[i].col >= sp->n_probe_cols)) {
                vg_set_error("groupby_feed_join: predicate reads column slot %d", sp->pred->insn[i].col);
                return -1;
            }
    }
    int64_t incoming = n_logical * (sp->join ? std::max<int64_t>(1, sp->join->max_dup) : 1);
    if (g->hint > 0 && incoming > g->hint) incoming = g->hint;
    if (gb_ensure_capacity(g, incoming)) return -1;

    JoinFeedArgs A;
    memset(&A, 0, sizeof A);
    A.n_logical = n_logical;
    A.sel = sp->sel;
    A.pos_base = sp->pos_base;
    A.has_join = sp->join != nullptr;
    A.join_kind = sp->join_kind;
    A.key_side = sp->key_side;
    A.key_col = sp->key_col;
    A.probe_key_col = sp->probe_key_col;
    A.n_inputs = sp->n_inputs;
    for (int i = 0; i < sp->n_inputs; i++) A.in[i] = sp->inputs[i];
    // build columns embedded in the table are read from the probed entry
    for (int c = 0; c < VGPU_MAX_COLS; c++) A.bpay[c] = -1;
    if (sp->join) {
        for (int c = 0; c < sp->n_build_cols; c++)
            for (int p = 0; p < sp->join->n_payload; p++) {
                const vgpu_col &q = sp->join->payload[p];
                if (q.data == sp->build_cols[c].data && q.validity == sp->build_cols[c].validity &&
                    q.type == sp->build_cols[c].type) {
                    A.bpay[c] = (int8_t)p;
                    break;
                }
            }
        bool need = !g->L.key_none && sp->key_side == 1 && A.bpay[sp->key_col] < 0;
        for (int i = 0; i < sp->n_inputs; i++) {
            const vgpu_input &in = sp->inputs[i];
            need |= in.side_a == 1 && A.bpay[in.col_a] < 0;
            need |= in.op == 1 && in.side_b == 1 && A.bpay[in.col_b] < 0;
        }
        A.need_rows = need;
    }
    A.hot = g->hot.as<uint64_t>();
    A.cold = g->cold.as<uint64_t>();
    A.ctl = g->ctl.as<uint64_t>();
    A.cap = g->cap;
    ColPair cp;
    memset(&cp, 0, sizeof cp);
    for (int i = 0; i < sp->n_probe_cols; i++) cp.probe.c[i] = sp->probe_cols[i];
    for (int i = 0; i < sp->n_build_cols; i++) cp.build.c[i] = sp->build_cols[i];
    for (int i = 0; i < sp->n_inputs; i++) {
        const vgpu_input &in = sp->inputs[i];
        int ta = in.side_a == 0 ? cp.probe.c[in.col_a].type : cp.build.c[in.col_a].type;
        int t = ta;
        if (in.op == 1) {
            int tb = in.side_b == 0 ? cp.probe.c[in.col_b].type : cp.build.c[in.col_b].type;
            t = (ta == VGPU_DOUBLE || tb == VGPU_DOUBLE) ? VGPU_DOUBLE : VGPU_INT64;
        }
        if (t != g->L.in_type[i]) {
            vg_set_error("groupby_feed_join: input %d evaluates to type %d, declared %d", i, t, g->L.in_type[i]);
            return -1;
        }
    }
    JView jv;
    memset(&jv, 0, sizeof jv);
    if (sp->join) jv = sp->join->view();
    PredInfo pi = vg_pred_info(sp->pred);
    vgpu_prog prog;
    memset(&prog, 0, sizeof prog);
    if (sp->pred) prog = *sp->pred;
    size_t vm = (pi.present && !pi.fast) ? vg_vm_lds_bytes(sp->pred, GB_THREADS) : 0;
    vm = (vm + 15) & ~(size_t)15;
    A.vm_lds = (int32_t)vm;
    A.lds_slots = gb_lds_slots(g, vm);
    size_t lds = vm + (A.lds_slots ? (size_t)(A.lds_slots + 2) * (g->L.RH + g->L.RC) * 8 : 0);
    StarArgs st;
    if (gb_star_shape(g, sp, A, pi, &st)) {
        st.lds_slots = A.lds_slots;
        st.vm_lds = 0;
        size_t slds = A.lds_slots ? (size_t)(A.lds_slots + 2) * (g->L.RH + g->L.RC) * 8 : 0;
        int items = (int)vg_tune(VG_TUNE_GB_STAR_ITEMS);
        unsigned grid = vg_blocks(n_logical, vg_tune(A.lds_slots ? VG_TUNE_GB_LDS_ROWS : VG_TUNE_GB_GLB_ROWS),
                                  vg_tune(A.lds_slots ? VG_TUNE_GB_LDS_MAX_BLOCKS : VG_TUNE_GB_GLB_MAX_BLOCKS));
        vg_kt_begin();
        switch (items) {
        case 1: k_gb_star<1><<<grid, GB_THREADS, slds, vg_stream()>>>(g->L, st, jv, pi, cp.probe); break;
        case 2: k_gb_star<2><<<grid, GB_THREADS, slds, vg_stream()>>>(g->L, st, jv, pi, cp.probe); break;
        case 8: k_gb_star<8><<<grid, GB_THREADS, slds, vg_stream()>>>(g->L, st, jv, pi, cp.probe); break;
        default: k_gb_star<4><<<grid, GB_THREADS, slds, vg_stream()>>>(g->L, st, jv, pi, cp.probe); break;
        }
        VG_LAUNCH_CHECK();
        vg_kt_end();
        g->any_fed = 1;
        return 0;
    }
    vg_kt_begin();
    k_gb_feed_join<<<gb_grid(n_logical, A.lds_slots), GB_THREADS, lds, vg_stream()>>>(g->L, A, cp, jv, pi, prog);
    VG_LAUNCH_CHECK();
    vg_kt_end();
    g->any_fed = 1;
    return 0;
}

static int highest_bit(uint64_t x) { return x ? 63 - __builtin_clzll(x) : -1; }
static int lowest_bit(uint64_t x) { return x ? __builtin_ctzll(x) : 64; }

extern "C" int64_t vgpu_groupby_finish(vgpu_groupby *g) {
    if (g->finished) return g->n_groups;
    hipStream_t s = vg_stream();
    if (g->cap == 0 && gb_alloc_table(g, 64)) return -1;
    int64_t slots = g->cap + 2;
    if (g->sort_keys.reserve((size_t)slots * 8) || g->sort_vals.reserve((size_t)slots * 4)) return -1;
    k_gb_ctl_init<<<1, 64, 0, s>>>(g->ctl.as<uint64_t>(), 1);
    k_gb_compact<<<vg_blocks(slots, 256), 256, 0, s>>>(g->L, g->hot.as<uint64_t>(), g->cap, g->order,
                                                        g->ctl.as<uint64_t>(), g->sort_keys.as<uint64_t>(),
                                                        g->sort_vals.as<uint32_t>());
    VG_LAUNCH_CHECK();
    // the one synchronization of finish: group count, sort-key bit range,
    // overflow flag and the NA group's size
    uint64_t ctl[GB_CTL_WORDS];
    if (gb_read_ctl(g, ctl)) return -1;
    if (ctl[1]) {
        vg_set_error("groupby: hash table overflow (more groups than the table was sized for)");
        return -1;
    }
    int64_t n_sorted = (int64_t)ctl[5];
    // NA group of the key-sorted order: appended after the sort
    int na_last = g->order == VGPU_ORDER_KEY && ctl[8] != 0;
    uint64_t diff = ctl[6] ^ ctl[7];
    if (n_sorted > 1 && diff) {
        int lo = lowest_bit(diff), hi = highest_bit(diff) + 1;
        if (vg_sort_pairs(g->sort_keys.as<uint64_t>(), g->sort_vals.as<uint32_t>(), n_sorted, lo, hi, g->sort_tmp))
            return -1;
    }
    int64_t n_out = n_sorted + na_last;
    size_t vwords = (size_t)((n_out + 63) / 64 + 1);
    if (g->out_keys.reserve((size_t)(n_out + 1) * 8) || g->out_kvalid.reserve(vwords * 8)) return -1;
    EmitOut O;
    memset(&O, 0, sizeof O);
    O.keys = g->out_keys.ptr;
    O.kvalid = g->out_kvalid.as<uint64_t>();
    for (int a = 0; a < g->aggs.n; a++) {
        if (g->out_vals[a].reserve((size_t)(n_out + 1) * 8) || g->out_valid[a].reserve(vwords * 8)) return -1;
        O.vals[a] = g->out_vals[a].as<double>();
        O.valid[a] = g->out_valid[a].as<uint64_t>();
    }
    if (n_out > 0) {
        k_gb_emit<<<vg_blocks(n_out, 256), 256, 0, s>>>(g->L, g->aggs, g->hot.as<uint64_t>(), g->cold.as<uint64_t>(),
                                                         g->cap, g->sort_vals.as<uint32_t>(), n_sorted, na_last, n_out,
                                                         8, O, g->ctl.as<uint64_t>());
        VG_LAUNCH_CHECK();
    }
    g->n_groups = n_out;
    g->finished = 1;
    return n_out;
}

extern "C" int64_t vgpu_groupby_fetch(vgpu_groupby *g, void *keys, uint8_t *key_validity, double *const *agg_values,
                                      uint8_t *const *agg_validity) {
    if (!g->finished) {
        vg_set_error("groupby: fetch before finish");
        return -1;
    }
    hipStream_t s = vg_stream();
    int64_t n = g->n_groups;
    size_t vb = (size_t)(n + 7) / 8;
    uint64_t ctl[GB_CTL_WORDS];
    if (n > 0) {
        // bool keys are emitted as 8-byte 0/1 words; callers narrow them
        if (keys && !g->L.key_none)
            VG_TRY(hipMemcpyAsync(keys, g->out_keys.ptr, (size_t)n * 8, hipMemcpyDeviceToHost, s));
        if (key_validity) VG_TRY(hipMemcpyAsync(key_validity, g->out_kvalid.ptr, vb, hipMemcpyDeviceToHost, s));
        for (int a = 0; a < g->aggs.n; a++) {
            if (agg_values && agg_values[a])
                VG_TRY(hipMemcpyAsync(agg_values[a], g->out_vals[a].ptr, (size_t)n * 8, hipMemcpyDeviceToHost, s));
            if (agg_validity && agg_validity[a])
                VG_TRY(hipMemcpyAsync(agg_validity[a], g->out_valid[a].ptr, vb, hipMemcpyDeviceToHost, s));
        }
    }
    VG_TRY(hipMemcpyAsync(ctl, g->ctl.ptr, sizeof ctl, hipMemcpyDeviceToHost, s));
    VG_TRY(hipStreamSynchronize(s));
    g->warn_count = (int64_t)ctl[4];
    return g->warn_count;
}

extern "C" int vgpu_groupby_result_device(vgpu_groupby *g, const void **keys, const uint8_t **key_validity, int agg,
                                          const double **values, const uint8_t **validity) {
    if (!g->finished) {
        vg_set_error("groupby: result before finish");
        return -1;
    }
    if (keys) *keys = g->out_keys.ptr;
    if (key_validity) *key_validity = g->out_kvalid.as<uint8_t>();
    if (agg >= 0 && agg < g->aggs.n) {
        if (values) *values = g->out_vals[agg].as<double>();
        if (validity) *validity = g->out_valid[agg].as<uint8_t>();
    }
    return 0;
}

extern "C" void vgpu_groupby_destroy(vgpu_groupby *g) {
    if (!g) return;
    // buffers go back to the stream-ordered pool: no synchronization needed
    g->hot.release();
    g->cold.release();
    g->sketch.release();
    g->ctl.release();
    g->sort_keys.release();
    g->sort_vals.release();
    g->sort_tmp.release();
    g->out_keys.release();
    g->out_kvalid.release();
    for (auto &b : g->out_vals) b.release();
    for (auto &b : g->out_valid) b.release();
    delete g;
}
