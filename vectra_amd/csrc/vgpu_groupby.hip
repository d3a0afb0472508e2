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
#include "vgpu_groupby_core.h"
#include "vgpu_join_table.h"

#include <algorithm>
#include <cmath>
#include <vector>

// ---------------------------------------------------------------------------
// partitioned feed (one large feed of many groups)
//
// Per-row HBM atomics do not scale to tens of millions of groups: every
// update is a device-scope atomic on a random record.  A large feed whose
// sketch shows more groups than LDS tables can hold is instead
// hash-partitioned in two passes (2^B1 x 2^B2 partitions of a few hundred
// groups each) into staged rows, and every partition is then aggregated in
// an LDS table by one block, which writes its finished records once,
// contiguously.  Partitions with more than GBP_CH rows (hot keys) are split
// into several items whose partial records one more block combines.  The
// records stay as the group-by's "run" until finish, which sorts them
// directly; any later feed first folds the run into the hash table.
//
// Staged row (W = n_inputs + 2 words): key bits (NaN keys already replaced
// by their per-row synthetic key), input bits, and a meta word: bits 0-31
// the feed-local row index, 32-33 the key class, 34.. input validity.
// ---------------------------------------------------------------------------

#define GBP_G1 2048                // first-pass blocks (tiles of the feed)
#define GBP_T2 32768               // second-pass rows per block
#define GBP_CH 4096                // rows per aggregation item
#define GBP_MAX_BITS 10            // fan-out bits per pass

struct GbpArgs {
    int32_t W, RW, B1, B2, S, G1;
    // range partitioning (key order over int64 keys of a known range): the
    // partition of a regular key grows with (ordered key - omin) >> rshift,
    // so the partitions, and the items in them, follow key order; every
    // item writes its records sorted by key, and finish needs no sort
    int32_t range, rshift;         // f = min((ordered key - omin) >> rshift, fmax)
    // staged rows without the meta word (W = inputs + 1): a plain feed (no
    // predicate, selection or NAs; int64 key) in key order whose double
    // min / max inputs hold no NaN or zero, so no position is ever read: a
    // row's key class follows from its key, every input is valid, and its
    // position reads as 0 (the first pass sets chk[2] when it meets a NaN
    // or a zero, and the feed restarts with meta words)
    int32_t nometa, pad_nm;
    uint64_t omin, fmax, fmul;     // partition (f * fmul) >> 32, fmul = (P - 1) * 2^32 / (fmax + 1)
    int64_t n_rows;                // logical rows of the feed (bounds every buffer)
    int64_t T1;                    // first-pass rows per block
    int64_t P;                     // partitions
    int64_t pos_base;
    uint64_t *S1, *S2;             // staged rows
    uint32_t *H1;                  // first-pass histogram [2^B1][G1]
    uint64_t *O1;                  // its exclusive scan (+ total at [2^B1 * G1])
    uint64_t *blkbase2;            // second-pass blocks per bucket, scanned (2^B1 + 1)
    uint32_t *H2;                  // second-pass histogram (bucket, sub-bucket, block)
    uint64_t *O2;
    int64_t NB2max;
    uint64_t *pstart;              // partition row starts (P + 1)
    uint32_t *nitems;              // items per partition (P)
    uint64_t *itembase;            // scanned items (P + 1)
    uint32_t *ipart;               // partition of every item (NImax)
    uint4 *idesc;                  // every item: first row, end row, partition, flags (1: the
                                   // partition has several items, 2: the item is its first)
    int64_t NImax;
    uint32_t *icount;              // regular records per item
    uint64_t *irec;                // first record index of an item
    uint32_t *rflags;              // reserved partials of an item: 1 NA, 2 EMPTY-valued key
    uint32_t *mlist;               // partitions with more than one item
    uint64_t *chk;                 // balance of range partitions: [0] rows of the largest
                                   // first-pass bucket, [1] items of the largest partition
    uint64_t *rec;                 // records: (n_rows + 2) x RW words; n_rows / n_rows+1 = NA / EMPTY key
    uint64_t *hot, *cold, *ctl;    // hash table (fallback for records that do not fit an LDS table)
    int64_t cap;
};

struct vgpu_groupby {
    GbLayout L;
    GbAggs aggs;
    int32_t order;
    int64_t cap = 0;          // hashed slots (pow2); +2 reserved slots follow
    int64_t n_groups = 0;     // after finish
    int64_t hint = 0;
    int64_t est_groups = 0;   // groups expected when no hint was given (sketch or groups seen); 0 = unknown
    int32_t sk_ranged = 0;    // the last sketch saw regular int64 / bool keys in [sk_lo, sk_hi]
    int32_t path = 0;         // the last feed: 0 hash table, 1 hash partitions, 2 range partitions
    int64_t sk_lo = 0, sk_hi = 0;
    int32_t finished = 0;
    int32_t any_fed = 0;
    DevBuf hot, cold;
    DevBuf sketch;            // cardinality sketch registers (GB_SK_REGS u32) + NaN-key row count (u64)
    DevBuf ctl;               // [0] occupied count, [1] error flag, [2] zero-key -0 pos,
                              // [3] zero-key +0 pos, [4] warn count, [5] tmp counter,
                              // [6] or-reduce, [7] and-reduce
    // finish buffers
    DevBuf sort_keys, sort_vals, sort_tmp, cmp;
    // partitioned feed: staged rows, bookkeeping and the run's records
    DevBuf gbp_s1, gbp_s2, gbp_aux, run_rec;
    // per-block partials of dense LDS tables
    DevBuf part;
    GbpArgs gbp;
    int32_t run = 0;          // the records of a partitioned feed await finish
    // the finished result in one buffer: keys, key validity, then per
    // aggregate its values and validity (offsets in bytes), so a small
    // result comes back in one copy
    DevBuf out;
    size_t off_kvalid = 0, out_bytes = 0;
    std::vector<size_t> off_vals, off_valid;
    int64_t warn_count = 0;
};

#define GB_U 4

// PF > 0: feeds of at most PF inputs load every row's key and input words
// (and validity) of a tile before the first row's slot is looked up, so a
// wave has 4 x (1 + inputs) loads in flight instead of one row's at a time
// behind its LDS probe
// DENSE: int64 / bool keys own record k - A.kmin of the LDS table (the
// range the feed's sketch saw), found without hashing or probing.
template <int PF, bool DENSE>
__global__ void __launch_bounds__(GB_THREADS)
k_gb_feed(GbLayout L, FeedArgs A, InSet ins, PredInfo pi, vgpu_prog prog, ColSet pcs) {
    extern __shared__ uint8_t dyn[];
    const int tid = threadIdx.x;
    uint64_t *regs = reinterpret_cast<uint64_t *>(dyn);
    uint8_t *regv = dyn + (size_t)prog.n_insn * GB_THREADS * 8;
    uint64_t *lhot0 = reinterpret_cast<uint64_t *>(dyn + A.vm_lds);
    const int S = A.lds_slots;
    // Few groups: every block keeps lds_copies copies of its table and lane
    // l updates copy l % lds_copies, so the lanes of one atomic instruction
    // collide on a record lds_copies times less often.
    const int NC = A.lds_copies > 1 ? A.lds_copies : 1;
    uint64_t *lcold0 = A.lds_cold ? lhot0 + (size_t)NC * (S + 2) * L.RH : nullptr;
    const int copy = (tid & (VG_WAVE - 1)) % NC;
    uint64_t *lhot = lhot0 + (size_t)copy * (S + 2) * L.RH;
    uint64_t *lcold = lcold0 ? lcold0 + (size_t)copy * (S + 2) * L.RC : nullptr;
    const uint64_t empty = L.key_is_double ? VG_EMPTY_F64 : VG_EMPTY_I64;
    if (S) {
        if (DENSE) {
            // every copy's records keyed kmin, kmin + 1, ... (the copies'
            // reserved records as well: never keyed, never merged by key)
            for (int c = 0; c < NC; c++)
                gb_lds_init(L, lhot0 + (size_t)c * (S + 2) * L.RH, lcold0 ? lcold0 + (size_t)c * (S + 2) * L.RC : nullptr,
                            S, empty, true, A.kmin);
        } else {
            gb_lds_init(L, lhot0, lcold0, NC * (S + 2) - 2, empty);
        }
    }
    // Blocks walk tiles of GB_U * GB_THREADS contiguous rows; each thread's
    // GB_U rows of a tile have their selection, predicate and key loads
    // issued together, so each wave keeps several loads in flight within
    // one DRAM neighbourhood
    const int64_t tile = (int64_t)GB_THREADS * GB_U;
    for (int64_t t0 = (int64_t)blockIdx.x * tile; t0 < A.n_logical; t0 += (int64_t)gridDim.x * tile) {
        const int64_t base = t0 + tid;
        const int64_t stride = GB_THREADS;
        int64_t row[GB_U];
        bool ok[GB_U];
        int cls[GB_U];
        uint64_t kn[GB_U];
#pragma unroll
        for (int u = 0; u < GB_U; u++) {
            int64_t li = base + u * stride;
            ok[u] = li < A.n_logical;
            row[u] = ok[u] ? (A.sel ? (int64_t)A.sel[li] : li) : 0;
        }
#pragma unroll
        for (int u = 0; u < GB_U; u++) ok[u] = ok[u] && vg_pred_row(pi, prog, pcs, row[u], regs, regv, GB_THREADS, tid);
        // PF: the words of every row first (rows past the end or dropped by
        // the predicate read row 0 and are not applied), then the classes
        uint64_t kraw[GB_U];
        bool kval[GB_U];
        uint64_t ib[GB_U][PF > 0 ? PF : 1];
        bool iv[GB_U][PF > 0 ? PF : 1];
        if (PF > 0) {
#pragma unroll
            for (int u = 0; u < GB_U; u++) {
                const int64_t r = ok[u] ? row[u] : 0;
                kval[u] = !L.key_none && vg_valid(A.key.validity, r);
                kraw[u] = L.key_none ? 0 : vg_load_bits(A.key, r);
#pragma unroll
                for (int i = 0; i < (PF > 0 ? PF : 1); i++) {
                    if (i < L.n_inputs) {
                        iv[u][i] = vg_valid(ins.c[i].validity, r);
                        ib[u][i] = vg_load_bits(ins.c[i], r);
                    }
                }
            }
        }
#pragma unroll
        for (int u = 0; u < GB_U; u++) {
            kn[u] = 0;
            cls[u] = VG_KEY_REGULAR;
            if (PF > 0) {
                if (ok[u] && !L.key_none) cls[u] = vg_key_class_bits(A.key.type, kval[u], kval[u] ? kraw[u] : 0, kn[u]);
            } else {
                if (ok[u] && !L.key_none) cls[u] = vg_key_class(A.key, row[u], kn[u]);
            }
        }
#pragma unroll
        for (int u = 0; u < GB_U; u++) {
            if (!ok[u]) continue;
            uint64_t pos = (uint64_t)(A.pos_base + base + u * stride);
            if (L.key_is_double && cls[u] == VG_KEY_REGULAR && kn[u] == 0)
                gb_note_zero_key(A.ctl, vg_load_bits(A.key, row[u]), pos);
            int64_t ls = !S ? -1 : DENSE ? gb_lds_dense_slot(S, cls[u], kn[u], A.kmin)
                                         : gb_lds_slot(lhot, S, L.RH, cls[u], kn[u], empty);
            auto apply = [&](uint64_t *h, uint64_t *c) {
                if (PF > 0) {
                    gb_update_row(L, h, pos);
#pragma unroll
                    for (int i = 0; i < (PF > 0 ? PF : 1); i++) {
                        if (i < L.n_inputs) {
                            InVal v;
                            v.valid = iv[u][i];
                            v.bits = iv[u][i] ? ib[u][i] : 0;
                            gb_update_input(L, h, c, i, pos, v);
                        }
                    }
                } else {
                    gb_apply(L, ins, row[u], pos, h, c);
                }
            };
            // one expansion per address space: ds_* atomics on the LDS
            // record, global atomics on the HBM one (a pointer that may be
            // either would make every update a flat atomic)
            if (ls >= 0) {
                apply(lhot + (size_t)ls * L.RH, lcold ? lcold + (size_t)ls * L.RC : nullptr);
            } else {
                int64_t gs = gb_global_slot(L, A.hot, A.cap, cls[u], kn[u], pos, A.ctl);
                if (gs < 0) continue;
                apply(A.hot + (size_t)gs * L.RH, A.cold + (size_t)gs * L.RC);
            }
        }
    }
    if (S) {
        // the copies fold into copy 0 in LDS; one merge into HBM
        for (int c = 1; c < NC; c++) {
            __syncthreads();
            uint64_t *ch = lhot0 + (size_t)c * (S + 2) * L.RH;
            uint64_t *cc = lcold0 ? lcold0 + (size_t)c * (S + 2) * L.RC : nullptr;
            for (int s = tid; s < S + 2; s += GB_THREADS) {
                const uint64_t *h = ch + (size_t)s * L.RH;
                if (h[1] == 0) continue;
                const int64_t d = s < S && !DENSE ? gb_lds_slot(lhot0, S, L.RH, VG_KEY_REGULAR, h[0], empty) : s;
                if (d >= 0)
                    gb_merge(L, lhot0 + (size_t)d * L.RH, lcold0 ? lcold0 + (size_t)d * L.RC : nullptr, h,
                             cc ? cc + (size_t)s * L.RC : nullptr);
                else   // copy 0 has no room for the key: straight to HBM
                    gb_merge_global(L, h, cc ? cc + (size_t)s * L.RC : nullptr, A.hot, A.cold, A.cap, A.ctl);
            }
        }
        gb_lds_finish(L, lhot0, lcold0, S, empty, A.hot, A.cold, A.cap, A.ctl, A.part);
    }
}

// The narrow shape of most group-bys: n / count / sum / mean of at most two
// inputs, a fast predicate or none.  The record words it touches come as
// plain integers, so the row loop keeps few values in scalar registers.
// DENSE: int64 / bool keys own record k - kmin of the LDS table (the range
// the feed's sketch saw), found without hashing.
struct NarrowArgs {
    uint64_t *part;           // block partials of the LDS table (nullptr: merge into HBM directly)
    int64_t n_logical, pos_base;
    const int32_t *sel;
    vgpu_col key;
    vgpu_col in[2];
    vgpu_col pcol;            // the fast predicate's column (pi.present)
    PredInfo pi;
    uint64_t *hot, *cold, *ctl;
    int64_t cap, kmin;
    int32_t S, RH, RC, lds_cold;
    int32_t w_fpos, key_double, key_none, pad;
    int32_t w_sum[2], c_na[2], in_f64[2];   // per input: sum word (-1: none), NA count word, double sum
};

template <int NI>
__device__ __forceinline__ void narrow_update(const NarrowArgs &A, uint64_t *h, uint64_t *c, uint64_t pos,
                                              const bool (&iv)[2], const uint64_t (&ib)[2]) {
    gb_atomic_add(&h[1], 1);
    if (A.w_fpos >= 0) gb_atomic_min(&h[A.w_fpos], pos);
#pragma unroll
    for (int i = 0; i < NI; i++) {
        if (!iv[i]) {
            gb_atomic_add(&c[A.c_na[i]], 1);
        } else if (A.w_sum[i] >= 0) {
            if (A.in_f64[i]) gb_atomic_addf(&h[A.w_sum[i]], vg_bits_to_f64(ib[i]));
            else gb_atomic_add(&h[A.w_sum[i]], ib[i]);
        }
    }
}

template <int NI, bool DENSE>
__global__ void __launch_bounds__(GB_THREADS) k_gb_feed_narrow(GbLayout L, NarrowArgs A) {
    extern __shared__ uint64_t dyn64[];
    const int tid = threadIdx.x;
    const int S = A.S;
    uint64_t *lhot = dyn64;
    uint64_t *lcold = A.lds_cold ? lhot + (size_t)(S + 2) * A.RH : nullptr;
    const uint64_t empty = A.key_double ? VG_EMPTY_F64 : VG_EMPTY_I64;
    if (S) gb_lds_init(L, lhot, lcold, S, empty, DENSE, A.kmin);
    const int64_t tile = (int64_t)GB_THREADS * GB_U;
    for (int64_t t0 = (int64_t)blockIdx.x * tile; t0 < A.n_logical; t0 += (int64_t)gridDim.x * tile) {
        int64_t row[GB_U];
        bool ok[GB_U];
        int cls[GB_U];
        uint64_t kn[GB_U];
        bool iv[GB_U][2];
        uint64_t ib[GB_U][2];
#pragma unroll
        for (int u = 0; u < GB_U; u++) {
            const int64_t li = t0 + tid + (int64_t)u * GB_THREADS;
            ok[u] = li < A.n_logical;
            row[u] = ok[u] ? (A.sel ? (int64_t)A.sel[li] : li) : 0;
        }
#pragma unroll
        for (int u = 0; u < GB_U; u++)
            if (ok[u] && A.pi.present)
                ok[u] = vg_valid(A.pcol.validity, row[u]) && vg_fast_cmp(A.pi, vg_load_bits(A.pcol, row[u]), A.pi.k);
#pragma unroll
        for (int u = 0; u < GB_U; u++) {
            kn[u] = 0;
            cls[u] = VG_KEY_REGULAR;
            if (ok[u] && !A.key_none) cls[u] = vg_key_class(A.key, row[u], kn[u]);
#pragma unroll
            for (int i = 0; i < 2; i++) {
                iv[u][i] = false;
                ib[u][i] = 0;
                if (i < NI && ok[u]) {
                    iv[u][i] = vg_valid(A.in[i].validity, row[u]);
                    ib[u][i] = iv[u][i] ? vg_load_bits(A.in[i], row[u]) : 0;
                }
            }
        }
#pragma unroll
        for (int u = 0; u < GB_U; u++) {
            if (!ok[u]) continue;
            const uint64_t pos = (uint64_t)(A.pos_base + t0 + tid + (int64_t)u * GB_THREADS);
            if (A.key_double && cls[u] == VG_KEY_REGULAR && kn[u] == 0)
                gb_note_zero_key(A.ctl, vg_load_bits(A.key, row[u]), pos);
            int64_t ls = -1;
            if (S) ls = DENSE ? gb_lds_dense_slot(S, cls[u], kn[u], A.kmin) : gb_lds_slot(lhot, S, A.RH, cls[u], kn[u], empty);
            // one expansion per address space: ds_* atomics for the LDS record
            if (ls >= 0) {
                narrow_update<NI>(A, lhot + (size_t)ls * A.RH, lcold ? lcold + (size_t)ls * A.RC : nullptr, pos, iv[u],
                                  ib[u]);
            } else {
                const int64_t gs = gb_global_slot(L, A.hot, A.cap, cls[u], kn[u], pos, A.ctl);
                if (gs >= 0)
                    narrow_update<NI>(A, A.hot + (size_t)gs * A.RH, A.cold + (size_t)gs * A.RC, pos, iv[u], ib[u]);
            }
        }
    }
    if (S) gb_lds_finish(L, lhot, lcold, S, empty, A.hot, A.cold, A.cap, A.ctl, A.part);
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
// A feed holds fewer than 2^31 probe rows and build keys of at most
// JF_MAX_MATCHES rows (both checked on the host), so positions of one feed
// are distinct and below pos_base + 2^44.
#define JF_MAX_MATCHES 4096
__device__ __forceinline__ uint64_t jf_pos(uint64_t base, int64_t li, uint32_t j, bool tail) {
    return base + ((tail ? (1ull << 43) : 0ull) | ((uint64_t)li << 12) | (j & 0xFFFu));
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

// Folds of block partials (see gb_lds_store_partial).  Dense tables (record
// s of every block holds the same key): one block per record folds record s
// of every partial in LDS and merges it into the HBM table once.  Hashed
// tables: block j folds partials [j * per, (j + 1) * per) into an LDS table
// of S slots (a key that finds no slot goes to the HBM table directly) and
// merges that, so a key's HBM record takes one merge per fold block.
__global__ void __launch_bounds__(GB_THREADS) k_gb_part_fold_dense(GbLayout L, const uint64_t *part, int64_t nparts,
                                                                   int S, uint64_t *hot, uint64_t *cold, int64_t cap,
                                                                   uint64_t *ctl) {
    __shared__ uint64_t rec[2 * GB_MAXW];
    const int s = blockIdx.x, RW = L.RH + L.RC;
    const uint64_t *first = part + (size_t)s * RW;
    if ((int)threadIdx.x < RW) {
        const int w = threadIdx.x;
        rec[w] = w == 0 ? first[0]
                 : w < L.RH ? (((L.hot_init_ones >> w) & 1) ? ~0ull : 0ull)
                            : (((L.cold_init_ones >> (w - L.RH)) & 1) ? ~0ull : 0ull);
    }
    __syncthreads();
    for (int64_t b = threadIdx.x; b < nparts; b += GB_THREADS) {
        const uint64_t *p = part + ((size_t)b * (S + 2) + s) * RW;
        if (p[1] == 0) continue;
        gb_merge(L, rec, rec + L.RH, p, p + L.RH);
    }
    __syncthreads();
    if (threadIdx.x != 0 || rec[1] == 0) return;
    const int64_t gs = s < S ? gb_global_slot(L, hot, cap, VG_KEY_REGULAR, rec[0], 0, ctl) : cap + (s - S);
    if (gs >= 0) gb_merge(L, hot + (size_t)gs * L.RH, cold + (size_t)gs * L.RC, rec, rec + L.RH);
}

// DENSE: the partials are dense tables (record s keyed kmin + s), folded
// record s into record s of a dense LDS table
template <bool DENSE>
__global__ void __launch_bounds__(GB_THREADS) k_gb_part_fold(GbLayout L, const uint64_t *part, int64_t nparts, int S,
                                                             int per, uint64_t *hot, uint64_t *cold, int64_t cap,
                                                             uint64_t *ctl, int64_t kmin) {
    extern __shared__ uint64_t dyn_fold[];
    const int RW = L.RH + L.RC;
    uint64_t *lhot = dyn_fold;
    uint64_t *lcold = lhot + (size_t)(S + 2) * L.RH;
    const uint64_t empty = L.key_is_double ? VG_EMPTY_F64 : VG_EMPTY_I64;
    gb_lds_init(L, lhot, lcold, S, empty, DENSE, kmin);
    const int64_t b0 = (int64_t)blockIdx.x * per;
    const int64_t nb = nparts - b0 < per ? nparts - b0 : per;
    const int64_t entries = nb * (S + 2);
    for (int64_t q = threadIdx.x; q < entries; q += GB_THREADS) {
        const int64_t b = b0 + q / (S + 2);
        const int s = (int)(q % (S + 2));
        const uint64_t *r = part + ((size_t)b * (S + 2) + s) * RW;
        if (r[1] == 0) continue;
        const int64_t ls = s < S && !DENSE ? gb_lds_slot(lhot, S, L.RH, VG_KEY_REGULAR, r[0], empty) : s;
        if (ls >= 0) gb_merge(L, lhot + (size_t)ls * L.RH, lcold + (size_t)ls * L.RC, r, r + L.RH);
        else gb_merge_global(L, r, r + L.RH, hot, cold, cap, ctl);
    }
    gb_lds_merge(L, lhot, lcold, S, empty, hot, cold, cap, ctl);
}

// The star shape of most join -> aggregate queries, with as little as
// possible kept per row: a unique-key join table whose 32-byte entries hold
// the payload, the group key an int64 / bool payload column over a short
// range (a dense LDS table: no hashing, no probing), one input (a probe
// column, a payload column or their product) under n / count / sum / mean,
// no filter.  Left joins send an unmatched row to the NA group with an NA
// payload operand (src/join.c:415-432), at the batch-tail position.
struct StarDenseArgs {
    vgpu_col pkey;          // probe key
    vgpu_col pval;          // probe operand (input modes 0 and 2)
    const uint64_t *ent;    // join entries: key, segment word, payload 0, payload 1
    uint64_t jmask;         // join slots - 1 (the EMPTY-valued key's entry at jmask + 1)
    const uint64_t *bloom;  // the table's filter of build keys (nullptr: none)
    uint64_t bloom_mask;
    int64_t n_logical, pos_base;
    int64_t kmin;           // group key of dense record 0
    uint64_t *hot, *cold, *ctl;
    uint64_t *part;         // per-block copies of the dense table (nullptr: merge into HBM directly)
    int64_t cap;
    int32_t S;              // dense records; the NA group at S, the EMPTY-valued key at S + 1
    int32_t RH, RC;
    int32_t w_fpos, w_sum, c_na;   // record words (-1: absent)
    int32_t kslot, wslot;          // payload words of the group key and of the payload operand
    int32_t jkey_double;           // the join key is double
    int32_t ktype;                 // group key type (int64 / bool)
    int32_t ta, tb;                // operand types (probe, payload)
    int32_t sum_f64;               // the input is double
};

__device__ __forceinline__ void star_dense_update(const StarDenseArgs &A, uint64_t *h, uint64_t *c, uint64_t pos,
                                                  InVal v) {
    gb_atomic_add(&h[1], 1);
    if (A.w_fpos >= 0) gb_atomic_min(&h[A.w_fpos], pos);
    if (!v.valid) {
        gb_atomic_add(&c[A.c_na], 1);
    } else if (A.w_sum >= 0) {
        if (A.sum_f64) gb_atomic_addf(&h[A.w_sum], vg_bits_to_f64(v.bits));
        else gb_atomic_add(&h[A.w_sum], v.bits);
    }
}

// One step's probes and updates: every row's filter word and first entry
// load, then the chain walks (rare at load <= 1/2) and the record updates.
// Every row's entry load is issued -- a row whose key is absent or filtered
// out reads slot 0, a line already in cache -- so no load sits in a branch
// whose result the compiler must wait for before issuing the next.
template <int ITEMS, int IN_MODE, bool LEFT>
__device__ __forceinline__ void star_dense_step(const GbLayout &L, const StarDenseArgs &A, uint64_t *lhot,
                                                uint64_t *lcold, int64_t base, int tid, const bool (&live)[ITEMS],
                                                const int (&cls)[ITEMS], const uint64_t (&kn)[ITEMS],
                                                const bool (&pv_ok)[ITEMS], const uint64_t (&pv)[ITEMS]) {
    const uint64_t empty_j = A.jkey_double ? VG_EMPTY_F64 : VG_EMPTY_I64;
    bool found[ITEMS];
    uint64_t slot[ITEMS], bw[ITEMS], bb[ITEMS];
    uint64_t e0[ITEMS], e1[ITEMS], e2[ITEMS], e3[ITEMS];
#pragma unroll
    for (int i = 0; i < ITEMS; i++) {
        found[i] = cls[i] == VG_KEY_REGULAR || cls[i] == VG_KEY_SPECIAL;
        const uint64_t kh = vg_key_hash(kn[i]);
        slot[i] = cls[i] == VG_KEY_SPECIAL ? A.jmask + 1 : (kh & A.jmask);
        // the filter word now; a regular key whose bits are missing is
        // absent from the table and reads no entry line of its own
        bw[i] = ~0ull;
        bb[i] = 0;
        if (A.bloom) {
            const uint64_t w = A.bloom[jt_bloom_word(kh, A.bloom_mask)];
            if (cls[i] == VG_KEY_REGULAR) {
                bw[i] = w;
                bb[i] = jt_bloom_bits(kh);
            }
        }
    }
#pragma unroll
    for (int i = 0; i < ITEMS; i++) {
        found[i] = found[i] && (bw[i] & bb[i]) == bb[i];
        const uint64_t *e = A.ent + (found[i] ? slot[i] : 0) * 4;
        e0[i] = e[0];
        e1[i] = e[1];
        e2[i] = e[2];
        e3[i] = e[3];
    }
#pragma unroll
    for (int i = 0; i < ITEMS; i++) {
        if (!live[i]) continue;
        bool matched = false;
        if (found[i]) {
            if (slot[i] <= A.jmask) {
                uint64_t probes = 0;
                while (e0[i] != kn[i] && e0[i] != empty_j && probes <= A.jmask) {
                    slot[i] = (slot[i] + 1) & A.jmask;
                    const uint64_t *e = A.ent + slot[i] * 4;
                    e0[i] = e[0];
                    e1[i] = e[1];
                    e2[i] = e[2];
                    e3[i] = e[3];
                    probes++;
                }
                matched = e0[i] == kn[i] && (e1[i] & 0xFFFFFFu) != 0;
            } else {
                matched = (e1[i] & 0xFFFFFFu) != 0;
            }
        }
        if (!LEFT && !matched) continue;
        const int64_t li = base + (int64_t)i * GB_THREADS + tid;
        // group record
        const bool kvalid = matched && ((e1[i] >> (24 + A.kslot)) & 1);
        uint64_t gk;
        const int gcls = vg_key_class_bits(A.ktype, kvalid, A.kslot ? e3[i] : e2[i], gk);
        const uint64_t pos = jf_pos((uint64_t)A.pos_base, li, 0, !matched);
        // input
        InVal v;
        if (IN_MODE == 0) {
            v.valid = pv_ok[i];
            v.bits = pv[i];
        } else {
            InVal b;
            b.valid = matched && ((e1[i] >> (24 + A.wslot)) & 1);
            b.bits = A.wslot ? e3[i] : e2[i];
            if (IN_MODE == 1) {
                v = b;
            } else {
                InVal a;
                a.valid = pv_ok[i];
                a.bits = pv[i];
                v = jf_mul(a, A.ta, b, A.tb);
            }
        }
        const int64_t ls = gb_lds_dense_slot(A.S, gcls, gk, A.kmin);
        // one expansion per address space, so the LDS one compiles to
        // ds_* atomics rather than flat ones
        if (ls >= 0) {
            star_dense_update(A, lhot + (size_t)ls * A.RH, lcold + (size_t)ls * A.RC, pos, v);
        } else {
            const int64_t gs = gb_global_slot(L, A.hot, A.cap, gcls, gk, pos, A.ctl);
            if (gs >= 0) star_dense_update(A, A.hot + (size_t)gs * A.RH, A.cold + (size_t)gs * A.RC, pos, v);
        }
    }
}

// IN_MODE: 0 = probe column, 1 = payload column, 2 = their product.
// PLAIN: the probe key (and the probe operand) are 8-byte columns without
// validity: rows past the end read the last row and are masked out, and
// the next step's keys and operands are loaded while this step probes the
// table (a step's probe chain -- key, hash, filter word, entry -- is
// dependent; the streamed loads of the next step are not).
template <int ITEMS, int IN_MODE, bool LEFT, bool PLAIN>
__global__ void __launch_bounds__(GB_THREADS) k_gb_star_dense(GbLayout L, StarDenseArgs A) {
    extern __shared__ uint64_t dyn64[];
    const int tid = threadIdx.x;
    uint64_t *lhot = dyn64;
    uint64_t *lcold = lhot + (size_t)(A.S + 2) * A.RH;
    gb_lds_init(L, lhot, lcold, A.S, VG_EMPTY_I64, true, A.kmin);
    const int64_t step = (int64_t)gridDim.x * GB_THREADS * ITEMS;
    const int64_t base0 = (int64_t)blockIdx.x * GB_THREADS * ITEMS;
    if (PLAIN) {
        const uint64_t *kd = static_cast<const uint64_t *>(A.pkey.data);
        const uint64_t *vd = static_cast<const uint64_t *>(A.pval.data);
        uint64_t ck[ITEMS], cv[ITEMS], nk[ITEMS], nv[ITEMS];
        auto load_step = [&](int64_t b, uint64_t (&k)[ITEMS], uint64_t (&v)[ITEMS]) {
#pragma unroll
            for (int i = 0; i < ITEMS; i++) {
                const int64_t li = b + (int64_t)i * GB_THREADS + tid;
                const int64_t lc = li < A.n_logical ? li : A.n_logical - 1;
                k[i] = __builtin_nontemporal_load(kd + lc);
                v[i] = IN_MODE != 1 ? __builtin_nontemporal_load(vd + lc) : 0;
            }
        };
        if (base0 < A.n_logical) load_step(base0, ck, cv);
        for (int64_t base = base0; base < A.n_logical; base += step) {
            if (base + step < A.n_logical) load_step(base + step, nk, nv);
            bool live[ITEMS], pv_ok[ITEMS];
            int cls[ITEMS];
            uint64_t kn[ITEMS];
#pragma unroll
            for (int i = 0; i < ITEMS; i++) {
                live[i] = base + (int64_t)i * GB_THREADS + tid < A.n_logical;
                cls[i] = vg_key_class_bits(A.pkey.type, live[i], ck[i], kn[i]);
                pv_ok[i] = live[i];
            }
            star_dense_step<ITEMS, IN_MODE, LEFT>(L, A, lhot, lcold, base, tid, live, cls, kn, pv_ok, cv);
#pragma unroll
            for (int i = 0; i < ITEMS; i++) {
                ck[i] = nk[i];
                cv[i] = nv[i];
            }
        }
    } else {
        for (int64_t base = base0; base < A.n_logical; base += step) {
            bool live[ITEMS], pv_ok[ITEMS];
            int cls[ITEMS];
            uint64_t kn[ITEMS], pv[ITEMS];
#pragma unroll
            for (int i = 0; i < ITEMS; i++) {
                const int64_t li = base + (int64_t)i * GB_THREADS + tid;
                live[i] = li < A.n_logical;
                const bool kv = live[i] && vg_valid(A.pkey.validity, li);
                cls[i] = vg_key_class_bits(A.pkey.type, kv, kv ? vg_load_bits_stream(A.pkey, li) : 0, kn[i]);
                pv_ok[i] = false;
                pv[i] = 0;
                if (IN_MODE != 1) {
                    pv_ok[i] = live[i] && vg_valid(A.pval.validity, li);
                    pv[i] = pv_ok[i] ? vg_load_bits_stream(A.pval, li) : 0;
                }
            }
            star_dense_step<ITEMS, IN_MODE, LEFT>(L, A, lhot, lcold, base, tid, live, cls, kn, pv_ok, pv);
        }
    }
    gb_lds_finish(L, lhot, lcold, A.S, VG_EMPTY_I64, A.hot, A.cold, A.cap, A.ctl, A.part);
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------


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

static int gb_read_ctl(vgpu_groupby *g, uint64_t *ctl_host) {
    VG_TRY(hipMemcpyAsync(ctl_host, g->ctl.ptr, GB_CTL_WORDS * 8, hipMemcpyDeviceToHost, vg_stream()));
    VG_TRY(hipStreamSynchronize(vg_stream()));
    return 0;
}

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
    __shared__ unsigned long long nn, lo, hi;
    for (int i = threadIdx.x; i < GB_SK_REGS; i += GB_THREADS) r[i] = 0;
    if (threadIdx.x == 0) {
        nn = 0;
        lo = ~0ull;
        hi = 0;
    }
    __syncthreads();
    // the regular keys' range (order-preserving map), for dense LDS tables
    unsigned long long my_lo = ~0ull, my_hi = 0;
    const bool ranged = key.type == VGPU_INT64 || key.type == VGPU_BOOL;
    unsigned long long my_nan = 0;
    const int64_t stride = (int64_t)gridDim.x * GB_THREADS;
    for (int64_t li = (int64_t)blockIdx.x * GB_THREADS + threadIdx.x; li < n_logical; li += stride) {
        int64_t row = sel ? (int64_t)sel[li] : li;
        uint64_t kn;
        int cls = vg_key_class(key, row, kn);
        if (cls == VG_KEY_NAN) { my_nan++; continue; }
        if (cls != VG_KEY_REGULAR) continue;
        if (ranged) {
            const unsigned long long o = kn ^ 0x8000000000000000ull;
            my_lo = o < my_lo ? o : my_lo;
            my_hi = o > my_hi ? o : my_hi;
        }
        uint64_t h = vg_mix64(kn);
        uint32_t idx = (uint32_t)(h >> (64 - GB_SK_BITS));
        uint64_t rest = (h << GB_SK_BITS) | (1ull << (GB_SK_BITS - 1));
        uint32_t rank = (uint32_t)__builtin_clzll(rest) + 1;
        // repeated keys find their register already set: no atomic
        if (r[idx] < rank) atomicMax(&r[idx], rank);
    }
    if (my_nan) atomicAdd(&nn, my_nan);
    if (my_lo != ~0ull) atomicMin(&lo, my_lo);
    if (my_hi) atomicMax(&hi, my_hi);
    __syncthreads();
    for (int i = threadIdx.x; i < GB_SK_REGS; i += GB_THREADS)
        if (r[i]) atomicMax(&regs[i], r[i]);
    if (threadIdx.x == 0) {
        if (nn) atomicAdd(nan_rows, nn);
        if (lo != ~0ull) atomicMin(&nan_rows[1], lo);
        if (hi) atomicMax(&nan_rows[2], hi);
    }
}

// nan_rows[0] = NaN-key rows, [1] / [2] = min / max of the regular keys
__global__ void k_gb_sketch_clear(uint32_t *regs, unsigned long long *nan_rows) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < GB_SK_REGS) regs[i] = 0;
    if (i == 0) {
        nan_rows[0] = 0;
        nan_rows[1] = ~0ull;
        nan_rows[2] = 0;
    }
}

// estimated groups of a feed: distinct regular keys + one per NaN row + the
// two reserved groups (NA key, EMPTY-valued key)
static int gb_sketch(vgpu_groupby *g, const vgpu_col *key, const int32_t *sel, int64_t n_logical, int64_t *est) {
    size_t bytes = GB_SK_REGS * 4 + 24;
    if (g->sketch.reserve(bytes)) return -1;
    uint32_t *regs = g->sketch.as<uint32_t>();
    auto *nan_rows = reinterpret_cast<unsigned long long *>(regs + GB_SK_REGS);
    k_gb_sketch_clear<<<GB_SK_REGS / 256, 256, 0, vg_stream()>>>(regs, nan_rows);
    VG_LAUNCH_CHECK();
    k_gb_sketch<<<vg_blocks(n_logical, 65536, 2048), GB_THREADS, 0, vg_stream()>>>(*key, sel, n_logical, regs,
                                                                                     nan_rows);
    VG_LAUNCH_CHECK();
    std::vector<uint32_t> h(GB_SK_REGS + 6);
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
    uint64_t nan, lo, hi;
    memcpy(&nan, &h[GB_SK_REGS], 8);
    memcpy(&lo, &h[GB_SK_REGS + 2], 8);
    memcpy(&hi, &h[GB_SK_REGS + 4], 8);
    *est = (int64_t)std::ceil(e) + (int64_t)nan + 2;
    g->sk_ranged = lo <= hi;
    g->sk_lo = g->sk_ranged ? (int64_t)(lo ^ 0x8000000000000000ull) : 0;
    g->sk_hi = g->sk_ranged ? (int64_t)(hi ^ 0x8000000000000000ull) : 0;
    return 0;
}

// hashed slots for `groups` groups at the tuned load of the HBM table
// (knob gb.glb.load_pct, 25..90 percent; default 50)
static int64_t gb_slots_for(int64_t groups) {
    const int64_t pct = vg_tune_for(VG_TUNE_GB_GLB_LOAD_PCT, groups);
    return pow2_at_least((groups * 100 + pct - 1) / pct);
}

// grows the table so that `incoming` more groups keep the tuned load
static int gb_ensure_capacity(vgpu_groupby *g, int64_t incoming) {
    uint64_t ctl[GB_CTL_WORDS];
    if (g->cap == 0) {
        int64_t want = gb_slots_for(std::max<int64_t>(incoming, 32));
        if (g->hint > 0) want = std::max(want, gb_slots_for(g->hint));
        return gb_alloc_table(g, want);
    }
    // a table sized for the caller's group count is trusted until finish,
    // which reports an overflow; no per-feed synchronization
    if (g->hint > 0 && g->cap >= gb_slots_for(g->hint)) return 0;
    if (gb_read_ctl(g, ctl)) return -1;
    if (ctl[1]) {
        vg_set_error("groupby: hash table overflow");
        return -1;
    }
    // new groups of this feed are bounded by its rows (or its sketch)
    int64_t occupied = (int64_t)ctl[0];
    if (g->hint <= 0) g->est_groups = std::max(g->est_groups, occupied);
    int64_t need = gb_slots_for(occupied + incoming);
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

static vgpu_groupby *groupby_create_body(int key_type, int n_inputs, const int32_t *input_types, int n_aggs,
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
    if (gb_layout(g->L, g->order, key_type, n_inputs, input_types, n_aggs, aggs) || g->ctl.reserve(GB_CTL_WORDS * 8)) {
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
    g->off_vals.resize(n_aggs);
    g->off_valid.resize(n_aggs);
    return g;
}

extern "C" vgpu_groupby *vgpu_groupby_create(int key_type, int n_inputs, const int32_t *input_types, int n_aggs,
                                             const vgpu_agg *aggs, int order, int64_t expected_groups) {
    return vg_guard<vgpu_groupby *>(nullptr, "vgpu_groupby_create", [&] { return groupby_create_body(key_type, n_inputs,
        input_types, n_aggs, aggs, order, expected_groups); });
}

// row copies with 16-byte accesses when rows have an even number of words
// (row starts are then 16-byte aligned)
__device__ __forceinline__ void gbp_copy_row(uint64_t *d, const uint64_t *s, int W) {
    if ((W & 1) == 0) {
        for (int k = 0; k < W; k += 2)
            *reinterpret_cast<ulonglong2 *>(d + k) = *reinterpret_cast<const ulonglong2 *>(s + k);
    } else {
        for (int k = 0; k < W; k++) d[k] = s[k];
    }
}

__device__ __forceinline__ uint64_t gbp_hash(int cls, uint64_t kn) {
    if (cls == VG_KEY_NA) return 0x9e3779b97f4a7c15ull;       // every NA row in one partition
    if (cls == VG_KEY_SPECIAL) return 0x7f4a7c159e3779b9ull;  // every EMPTY-valued key row in one partition
    return vg_key_hash(kn);
}

// the word whose top B = B1 + B2 bits are a row's partition: its key's
// hash, or (range partitioning) its key's place in the range scaled onto
// partitions 0 .. P - 2; NA and EMPTY-valued keys alone in the last one
__device__ __forceinline__ uint64_t gbp_pkey(const GbpArgs &G, int cls, uint64_t kn) {
    if (!G.range) return gbp_hash(cls, kn);
    if (cls != VG_KEY_REGULAR) return ~0ull;
    const int B = G.B1 + G.B2;
    // clamped, so that a key outside the sketched range still lands in a
    // partition that keeps the order (first or last)
    const uint64_t o = kn ^ 0x8000000000000000ull;
    uint64_t f = (o > G.omin ? o - G.omin : 0) >> G.rshift;
    f = f < G.fmax ? f : G.fmax;
    // f * (P - 1) / (fmax + 1), below P - 1 and monotone; f * fmul < (P - 1) * 2^32 <= 2^52
    const uint64_t b = (f * G.fmul) >> 32;
    return b << (64 - B);
}

// largest i in [0, n) with a[i] <= x (a ascending, a[0] <= x)
__device__ __forceinline__ int64_t gbp_search(const uint64_t *a, int64_t n, uint64_t x) {
    int64_t lo = 0, hi = n - 1;
    while (lo < hi) {
        int64_t mid = (lo + hi + 1) >> 1;
        if (a[mid] <= x) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// key of logical row li of the feed; false when the predicate drops it
__device__ __forceinline__ bool gbp_source_row(const FeedArgs &A, const PredInfo &pi, const vgpu_prog &prog,
                                               const ColSet &pcs, uint64_t *regs, uint8_t *regv, int64_t li,
                                               int64_t &row, int &cls, uint64_t &kn, uint64_t &raw) {
    row = A.sel ? (int64_t)A.sel[li] : li;
    if (!vg_pred_row(pi, prog, pcs, row, regs, regv, GB_THREADS, threadIdx.x)) return false;
    bool v = vg_valid(A.key.validity, row);
    raw = v ? vg_load_bits(A.key, row) : 0;
    cls = vg_key_class_bits(A.key.type, v, raw, kn);
    if (cls == VG_KEY_NAN) {
        kn = 0x7FF8000000000000ull | ((uint64_t)(A.pos_base + li) & 0x0007FFFFFFFFFFFFull);
        cls = VG_KEY_REGULAR;
    }
    return true;
}

// Plain first pass: no predicate, no selection vector, no validity bitmap,
// an int64 key and 8-byte inputs.  Every row is regular or EMPTY-valued, so
// the loads of four rows per thread go out together (row indexes clamped,
// results of rows past the end unused) instead of one at a time behind the
// validity and type branches.
#define GBP_PLAIN_U 4

__device__ __forceinline__ int gbp_plain_class(uint64_t k) {
    return k == VG_EMPTY_I64 ? VG_KEY_SPECIAL : VG_KEY_REGULAR;
}

template <bool PLAIN>
__global__ void __launch_bounds__(GB_THREADS)
k_gbp_hist1(FeedArgs A, GbpArgs G, PredInfo pi, vgpu_prog prog, ColSet pcs) {
    extern __shared__ uint8_t dyn[];
    uint64_t *regs = reinterpret_cast<uint64_t *>(dyn);
    uint8_t *regv = dyn + (size_t)prog.n_insn * GB_THREADS * 8;
    uint32_t *hist = reinterpret_cast<uint32_t *>(dyn + A.vm_lds);
    const int NB = 1 << G.B1;
    if (blockIdx.x == 0 && threadIdx.x == 0) G.chk[2] = 0;   // the first pass's NaN / zero flag
    for (int b = threadIdx.x; b < NB; b += GB_THREADS) hist[b] = 0;
    __syncthreads();
    const int64_t r0 = (int64_t)blockIdx.x * G.T1;
    const int64_t r1 = r0 + G.T1 < A.n_logical ? r0 + G.T1 : A.n_logical;
    if (PLAIN) {
        const uint64_t *kp = static_cast<const uint64_t *>(A.key.data);
        for (int64_t li = r0 + threadIdx.x; li < r1; li += (int64_t)GB_THREADS * GBP_PLAIN_U) {
            uint64_t k[GBP_PLAIN_U];
#pragma unroll
            for (int u = 0; u < GBP_PLAIN_U; u++) {
                const int64_t rr = li + (int64_t)u * GB_THREADS;
                k[u] = kp[rr < r1 ? rr : r1 - 1];
            }
#pragma unroll
            for (int u = 0; u < GBP_PLAIN_U; u++)
                if (li + (int64_t)u * GB_THREADS < r1)
                    atomicAdd(&hist[gbp_pkey(G, gbp_plain_class(k[u]), k[u]) >> (64 - G.B1)], 1u);
        }
    } else {
        for (int64_t li = r0 + threadIdx.x; li < r1; li += GB_THREADS) {
            int64_t row;
            int cls;
            uint64_t kn, raw;
            if (!gbp_source_row(A, pi, prog, pcs, regs, regv, li, row, cls, kn, raw)) continue;
            atomicAdd(&hist[gbp_pkey(G, cls, kn) >> (64 - G.B1)], 1u);
        }
    }
    __syncthreads();
    for (int b = threadIdx.x; b < NB; b += GB_THREADS) G.H1[(int64_t)b * G.G1 + blockIdx.x] = hist[b];
}

// a staged row: key, inputs, meta (G.W words; 16-byte stores for four)
__device__ __forceinline__ void gbp_store_row(const GbpArgs &G, uint64_t *d, const uint64_t *w, uint64_t meta) {
    if (G.W == 4) {
        reinterpret_cast<ulonglong2 *>(d)[0] = make_ulonglong2(w[0], w[1]);
        reinterpret_cast<ulonglong2 *>(d)[1] = make_ulonglong2(w[2], meta);
    } else {
#pragma unroll
        for (int k = 0; k < VGPU_MAX_INPUTS + 1; k++)
            if (k < G.W - 1) d[k] = w[k];
        d[G.W - 1] = meta;
    }
}

template <bool PLAIN>
__global__ void __launch_bounds__(GB_THREADS)
k_gbp_scatter1(GbLayout L, FeedArgs A, GbpArgs G, InSet ins, PredInfo pi, vgpu_prog prog, ColSet pcs) {
    extern __shared__ uint8_t dyn[];
    uint64_t *regs = reinterpret_cast<uint64_t *>(dyn);
    uint8_t *regv = dyn + (size_t)prog.n_insn * GB_THREADS * 8;
    unsigned long long *cur = reinterpret_cast<unsigned long long *>(dyn + A.vm_lds);
    const int NB = 1 << G.B1;
    for (int b = threadIdx.x; b < NB; b += GB_THREADS) cur[b] = G.O1[(int64_t)b * G.G1 + blockIdx.x];
    __syncthreads();
    const int64_t r0 = (int64_t)blockIdx.x * G.T1;
    const int64_t r1 = r0 + G.T1 < A.n_logical ? r0 + G.T1 : A.n_logical;
    if (PLAIN) {
        const uint64_t *kp = static_cast<const uint64_t *>(A.key.data);
        uint64_t valid_bits = 0;
        for (int i = 0; i < L.n_inputs; i++) valid_bits |= 1ull << (34 + i);
        for (int64_t li = r0 + threadIdx.x; li < r1; li += (int64_t)GB_THREADS * GBP_PLAIN_U) {
            uint64_t w[GBP_PLAIN_U][VGPU_MAX_INPUTS + 1];
#pragma unroll
            for (int u = 0; u < GBP_PLAIN_U; u++) {
                int64_t rr = li + (int64_t)u * GB_THREADS;
                rr = rr < r1 ? rr : r1 - 1;
                w[u][0] = kp[rr];
#pragma unroll
                for (int i = 0; i < VGPU_MAX_INPUTS; i++)
                    w[u][1 + i] = i < L.n_inputs ? static_cast<const uint64_t *>(ins.c[i].data)[rr] : 0;
            }
#pragma unroll
            for (int u = 0; u < GBP_PLAIN_U; u++) {
                const int64_t rr = li + (int64_t)u * GB_THREADS;
                if (rr >= r1) continue;
                const int cls = gbp_plain_class(w[u][0]);
                const uint64_t dst = atomicAdd(&cur[gbp_pkey(G, cls, w[u][0]) >> (64 - G.B1)], 1ull);
                gbp_store_row(G, G.S1 + dst * G.W, w[u], (uint64_t)(uint32_t)rr | ((uint64_t)cls << 32) | valid_bits);
            }
        }
        return;
    }
    const int lane = vg_lane();
    for (int64_t li = r0 + threadIdx.x; li < r1; li += GB_THREADS) {
        int64_t row = 0;
        int cls = VG_KEY_NA;
        uint64_t kn = 0, raw = 0;
        bool ok = gbp_source_row(A, pi, prog, pcs, regs, regv, li, row, cls, kn, raw);
        if (L.key_is_double) {
            // first -0 / +0 key positions decide the sign printed for the 0 key;
            // lanes hold ascending rows, so the lowest flagged lane is the first
            bool z = ok && cls == VG_KEY_REGULAR && kn == 0;
            uint64_t mneg = vg_ballot(z && (raw >> 63)), mpos = vg_ballot(z && !(raw >> 63));
            if (mneg && lane == __ffsll((unsigned long long)mneg) - 1)
                gb_atomic_min(&A.ctl[2], (uint64_t)(A.pos_base + li));
            if (mpos && lane == __ffsll((unsigned long long)mpos) - 1)
                gb_atomic_min(&A.ctl[3], (uint64_t)(A.pos_base + li));
        }
        if (!ok) continue;
        uint64_t dst = atomicAdd(&cur[gbp_pkey(G, cls, kn) >> (64 - G.B1)], 1ull);
        uint64_t w[VGPU_MAX_INPUTS + 2];
        uint64_t meta = (uint64_t)(uint32_t)li | ((uint64_t)cls << 32);
        w[0] = kn;
#pragma unroll
        for (int i = 0; i < VGPU_MAX_INPUTS; i++) {
            if (i < L.n_inputs) {
                bool v = vg_valid(ins.c[i].validity, row);
                w[1 + i] = v ? vg_load_bits(ins.c[i], row) : 0;
                meta |= (uint64_t)v << (34 + i);
            }
        }
        gbp_store_row(G, G.S1 + dst * G.W, w, meta);
    }
}

// blocks of the second pass per first-pass bucket, scanned (one block)
__global__ void __launch_bounds__(1024) k_gbp_plan2(GbpArgs G) {
    __shared__ uint64_t part[1024];
    const int NB = 1 << G.B1;
    const int per = (NB + 1023) / 1024;
    uint64_t local = 0;
    for (int j = 0; j < per; j++) {
        int b = threadIdx.x * per + j;
        if (b < NB) {
            uint64_t size = G.O1[(int64_t)(b + 1) * G.G1] - G.O1[(int64_t)b * G.G1];
            local += (size + GBP_T2 - 1) / GBP_T2;
        }
    }
    part[threadIdx.x] = local;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t run = 0;
        for (int t = 0; t < 1024; t++) {
            uint64_t v = part[t];
            part[t] = run;
            run += v;
        }
        G.blkbase2[NB] = run;
    }
    __syncthreads();
    uint64_t run = part[threadIdx.x];
    for (int j = 0; j < per; j++) {
        int b = threadIdx.x * per + j;
        if (b < NB) {
            G.blkbase2[b] = run;
            uint64_t size = G.O1[(int64_t)(b + 1) * G.G1] - G.O1[(int64_t)b * G.G1];
            run += (size + GBP_T2 - 1) / GBP_T2;
        }
    }
}

// second-pass block -> (bucket, chunk, rows); false for blocks past the end
__device__ __forceinline__ bool gbp_block2(const GbpArgs &G, int64_t blk, int64_t &b, int64_t &chunk, int64_t &nblk,
                                           int64_t &r0, int64_t &r1) {
    const int64_t NB = 1 << G.B1;
    if ((uint64_t)blk >= G.blkbase2[NB]) return false;
    b = gbp_search(G.blkbase2, NB + 1, (uint64_t)blk);
    chunk = blk - (int64_t)G.blkbase2[b];
    nblk = (int64_t)(G.blkbase2[b + 1] - G.blkbase2[b]);
    int64_t bs = (int64_t)G.O1[b * G.G1], be = (int64_t)G.O1[(b + 1) * G.G1];
    r0 = bs + chunk * GBP_T2;
    r1 = r0 + GBP_T2 < be ? r0 + GBP_T2 : be;
    return true;
}

__device__ __forceinline__ int64_t gbp_h2_index(const GbpArgs &G, int64_t b, int64_t sub, int64_t chunk, int64_t nblk) {
    return (int64_t)G.blkbase2[b] * ((int64_t)1 << G.B2) + sub * nblk + chunk;
}

__device__ __forceinline__ uint32_t gbp_sub(const GbpArgs &G, const uint64_t *w) {
    int cls = G.nometa ? gbp_plain_class(w[0]) : (int)((w[G.W - 1] >> 32) & 3);
    return (uint32_t)((gbp_pkey(G, cls, w[0]) >> (64 - G.B1 - G.B2)) & ((1u << G.B2) - 1));
}

__global__ void __launch_bounds__(GB_THREADS) k_gbp_hist2(GbpArgs G) {
    __shared__ uint32_t hist[1 << GBP_MAX_BITS];
    int64_t b, chunk, nblk, r0, r1;
    if (!gbp_block2(G, blockIdx.x, b, chunk, nblk, r0, r1)) return;
    const int NS = 1 << G.B2;
    for (int i = threadIdx.x; i < NS; i += GB_THREADS) hist[i] = 0;
    __syncthreads();
    for (int64_t r = r0 + threadIdx.x; r < r1; r += GB_THREADS) atomicAdd(&hist[gbp_sub(G, G.S1 + r * G.W)], 1u);
    __syncthreads();
    for (int i = threadIdx.x; i < NS; i += GB_THREADS) G.H2[gbp_h2_index(G, b, i, chunk, nblk)] = hist[i];
}

__global__ void __launch_bounds__(GB_THREADS) k_gbp_scatter2(GbpArgs G) {
    __shared__ unsigned long long cur[1 << GBP_MAX_BITS];
    int64_t b, chunk, nblk, r0, r1;
    if (!gbp_block2(G, blockIdx.x, b, chunk, nblk, r0, r1)) return;
    const int NS = 1 << G.B2;
    for (int i = threadIdx.x; i < NS; i += GB_THREADS) cur[i] = G.O2[gbp_h2_index(G, b, i, chunk, nblk)];
    __syncthreads();
    if (G.W == 4) {
        // four rows per thread loaded before any is placed
        for (int64_t r = r0 + threadIdx.x; r < r1; r += (int64_t)GB_THREADS * 4) {
            ulonglong2 a[4], c[4];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const int64_t rr = r + (int64_t)u * GB_THREADS;
                a[u] = c[u] = make_ulonglong2(0, 0);
                if (rr < r1) {
                    const ulonglong2 *src = reinterpret_cast<const ulonglong2 *>(G.S1 + rr * 4);
                    a[u] = src[0];
                    c[u] = src[1];
                }
            }
#pragma unroll
            for (int u = 0; u < 4; u++) {
                if (r + (int64_t)u * GB_THREADS >= r1) continue;
                const uint64_t w[4] = {a[u].x, a[u].y, c[u].x, c[u].y};
                const uint64_t dst = atomicAdd(&cur[gbp_sub(G, w)], 1ull);
                ulonglong2 *d = reinterpret_cast<ulonglong2 *>(G.S2 + dst * 4);
                d[0] = a[u];
                d[1] = c[u];
            }
        }
        return;
    }
    for (int64_t r = r0 + threadIdx.x; r < r1; r += GB_THREADS) {
        const uint64_t *w = G.S1 + r * G.W;
        uint64_t dst = atomicAdd(&cur[gbp_sub(G, w)], 1ull);
        gbp_copy_row(G.S2 + dst * G.W, w, G.W);
    }
}

// ---- LDS-staged scatter ---------------------------------------------------
// A block moves its rows in tiles of SU * GB_THREADS: every row of a tile is
// counted by destination bucket (an LDS counter gives its rank), the tile's
// counts are scanned, the rows are placed in LDS in bucket order, and the
// block then writes the tile out in that order with consecutive lanes on
// consecutive 16-byte chunks (8-byte words for rows of an odd width), so
// that one bucket's rows of the tile land on consecutive addresses and a
// store instruction covers a few runs instead of one scattered row per
// lane.  Rows keep their meta word (feed row index), so the order inside a
// bucket is free.  Four-word rows are also read from staged rows chunk by
// chunk (a lane pair holds one row) in the second pass.
struct GbpStage {
    uint64_t *rows;                 // tile x W words, row-major
    unsigned long long *cur;        // next global row of each bucket
    uint32_t *cnt, *tst;            // the tile's rows per bucket, their scan
    uint16_t *bkt;                  // bucket of each placed row
    uint32_t *scan;                 // block prefix scratch
};

__host__ __device__ __forceinline__ size_t gbp_stage_bytes(int W, int NBk, int tile) {
    return (size_t)tile * W * 8 + (size_t)NBk * 16 + (size_t)tile * 2 + 4 * (GB_THREADS / VG_WAVE);
}

__device__ __forceinline__ GbpStage gbp_stage_carve(uint64_t *base, int W, int NBk, int tile) {
    GbpStage S;
    S.rows = base;
    S.cur = reinterpret_cast<unsigned long long *>(base + (size_t)tile * W);
    S.cnt = reinterpret_cast<uint32_t *>(S.cur + NBk);
    S.tst = S.cnt + NBk;
    S.bkt = reinterpret_cast<uint16_t *>(S.tst + NBk);
    S.scan = reinterpret_cast<uint32_t *>(S.bkt + tile);
    return S;
}

// exclusive scan of the tile's bucket counts into S.tst
__device__ __forceinline__ void gbp_stage_scan(GbpStage &S, int NBk) {
    const int per = (NBk + GB_THREADS - 1) / GB_THREADS;
    uint32_t loc = 0;
    for (int q = 0; q < per; q++) {
        const int s = (int)threadIdx.x * per + q;
        if (s < NBk) loc += S.cnt[s];
    }
    uint32_t total;
    uint32_t run = gb_block_prefix(loc, S.scan, total);
    for (int q = 0; q < per; q++) {
        const int s = (int)threadIdx.x * per + q;
        if (s < NBk) {
            S.tst[s] = run;
            run += S.cnt[s];
        }
    }
    __syncthreads();
}

// A tile in two halves, so that a kernel can issue the next tile's loads
// between them.  gbp_stage_place ranks rows w[u] (u < SU: row
// u * GB_THREADS + threadIdx.x of the tile, live below n) by bucket bk[u]
// and places them in LDS in bucket order (w is free afterwards);
// gbp_stage_write writes the placed rows to out (rows of W words) and
// advances the cursors.  S.cnt is zero on entry and left zero.
template <int WM, int SU>
__device__ __forceinline__ void gbp_stage_place(GbpStage &S, int W, int NBk, int n, const uint64_t (&w)[SU][WM],
                                                const uint32_t (&bk)[SU]) {
    uint32_t rank[SU];
#pragma unroll
    for (int u = 0; u < SU; u++)
        if (u * GB_THREADS + (int)threadIdx.x < n) rank[u] = atomicAdd(&S.cnt[bk[u]], 1u);
    __syncthreads();
    gbp_stage_scan(S, NBk);
#pragma unroll
    for (int u = 0; u < SU; u++) {
        if (u * GB_THREADS + (int)threadIdx.x >= n) continue;
        const uint32_t p = S.tst[bk[u]] + rank[u];
        S.bkt[p] = (uint16_t)bk[u];
        uint64_t *d = S.rows + (size_t)p * W;
        if constexpr (WM == 4) {
            reinterpret_cast<ulonglong2 *>(d)[0] = make_ulonglong2(w[u][0], w[u][1]);
            reinterpret_cast<ulonglong2 *>(d)[1] = make_ulonglong2(w[u][2], w[u][3]);
        } else {
#pragma unroll
            for (int k = 0; k < WM; k++)
                if (k < W) d[k] = w[u][k];
        }
    }
    __syncthreads();
}

// the same for four-word rows held as chunks: chunk c = u * GB_THREADS +
// threadIdx.x of the tile (u < 2 SU) is half c & 1 of row c >> 1, and the
// two lanes of a row hold its bucket
template <int SU>
__device__ __forceinline__ void gbp_stage_place4c(GbpStage &S, int NBk, int n, const ulonglong2 (&ch)[2 * SU],
                                                  const uint32_t (&bk)[2 * SU]) {
    uint32_t rank[2 * SU];
    const bool half = threadIdx.x & 1;
#pragma unroll
    for (int u = 0; u < 2 * SU; u++) {
        const int row = (u * GB_THREADS + (int)threadIdx.x) >> 1;
        uint32_t r = 0;
        if (row < n && !half) r = atomicAdd(&S.cnt[bk[u]], 1u);
        const uint32_t o = __shfl_xor(r, 1, VG_WAVE);
        rank[u] = half ? o : r;
    }
    __syncthreads();
    gbp_stage_scan(S, NBk);
#pragma unroll
    for (int u = 0; u < 2 * SU; u++) {
        const int row = (u * GB_THREADS + (int)threadIdx.x) >> 1;
        if (row >= n) continue;
        const uint32_t p = S.tst[bk[u]] + rank[u];
        if (!half) S.bkt[p] = (uint16_t)bk[u];
        reinterpret_cast<ulonglong2 *>(S.rows + (size_t)p * 4)[half] = ch[u];
    }
    __syncthreads();
}

template <int WM>
__device__ __forceinline__ void gbp_stage_write(GbpStage &S, int W, int NBk, int n, uint64_t *out) {
    if constexpr (WM == 4) {
        for (int c = threadIdx.x; c < 2 * n; c += GB_THREADS) {
            const int p = c >> 1, s = S.bkt[p];
            const uint64_t dst = S.cur[s] + (uint64_t)(p - (int)S.tst[s]);
            reinterpret_cast<ulonglong2 *>(out + dst * 4)[c & 1] =
                reinterpret_cast<const ulonglong2 *>(S.rows + (size_t)p * 4)[c & 1];
        }
    } else if (W % 2 == 0) {
        const int half = W >> 1;
        for (int c = threadIdx.x; c < n * half; c += GB_THREADS) {
            const int p = c / half, k = c - p * half, s = S.bkt[p];
            const uint64_t dst = S.cur[s] + (uint64_t)(p - (int)S.tst[s]);
            reinterpret_cast<ulonglong2 *>(out + dst * W)[k] =
                reinterpret_cast<const ulonglong2 *>(S.rows + (size_t)p * W)[k];
        }
    } else {
        for (int c = threadIdx.x; c < n * W; c += GB_THREADS) {
            const int p = c / W, k = c - p * W, s = S.bkt[p];
            const uint64_t dst = S.cur[s] + (uint64_t)(p - (int)S.tst[s]);
            out[dst * W + k] = S.rows[(size_t)p * W + k];
        }
    }
    __syncthreads();
    for (int s = threadIdx.x; s < NBk; s += GB_THREADS) {
        S.cur[s] += S.cnt[s];
        S.cnt[s] = 0;
    }
    __syncthreads();
}

// rows of the plain first pass's tile at t0 (n rows) with their buckets
template <int WM, int SU>
__device__ __forceinline__ void gbp_load1(const GbLayout &L, const FeedArgs &A, const GbpArgs &G, const InSet &ins,
                                          int64_t t0, int n, uint64_t valid_bits, uint64_t (&w)[SU][WM],
                                          uint32_t (&bk)[SU], bool &nan_or_zero) {
    const uint64_t *kp = static_cast<const uint64_t *>(A.key.data);
    const int W = G.W;
#pragma unroll
    for (int u = 0; u < SU; u++) {
        const int j = u * GB_THREADS + (int)threadIdx.x;
        const int64_t rr = t0 + (j < n ? j : n - 1);
        w[u][0] = kp[rr];
#pragma unroll
        for (int q = 1; q < WM; q++) {
            const int i = q - 1;
            w[u][q] = i < L.n_inputs ? static_cast<const uint64_t *>(ins.c[i].data)[rr] : 0;
        }
    }
#pragma unroll
    for (int u = 0; u < SU; u++) {
        const int j = u * GB_THREADS + (int)threadIdx.x;
        const int64_t rr = t0 + (j < n ? j : n - 1);
        const uint64_t k = w[u][0];
        const int cls = gbp_plain_class(k);
        if (G.nometa) {
            // a NaN or a zero in a double min / max input needs positions
#pragma unroll
            for (int q = 1; q < WM; q++) {
                const int i = q - 1;
                if (i < L.n_inputs && L.c_nan[i] >= 0) {
                    const uint64_t a = w[u][q] & 0x7FFFFFFFFFFFFFFFull;
                    nan_or_zero = nan_or_zero || a > 0x7FF0000000000000ull || a == 0;
                }
            }
        } else {
            const uint64_t meta = (uint64_t)(uint32_t)rr | ((uint64_t)cls << 32) | valid_bits;
#pragma unroll
            for (int q = 1; q < WM; q++)
                if (q == W - 1) w[u][q] = meta;
        }
        bk[u] = (uint32_t)(gbp_pkey(G, cls, k) >> (64 - G.B1));
    }
}

// the staged form of the plain first pass (see k_gbp_scatter1<true>)
template <int WM, int SU>
__global__ void __launch_bounds__(GB_THREADS) k_gbp_scatter1s(GbLayout L, FeedArgs A, GbpArgs G, InSet ins) {
    extern __shared__ uint64_t dyn_st[];
    constexpr int TILE = SU * GB_THREADS;
    const int NB = 1 << G.B1, W = G.W;
    GbpStage S = gbp_stage_carve(dyn_st, W, NB, TILE);
    for (int b = threadIdx.x; b < NB; b += GB_THREADS) {
        S.cur[b] = G.O1[(int64_t)b * G.G1 + blockIdx.x];
        S.cnt[b] = 0;
    }
    __syncthreads();
    const int64_t r0 = (int64_t)blockIdx.x * G.T1;
    const int64_t r1 = r0 + G.T1 < A.n_logical ? r0 + G.T1 : A.n_logical;
    uint64_t valid_bits = 0;
    for (int i = 0; i < L.n_inputs; i++) valid_bits |= 1ull << (34 + i);
    uint64_t w[SU][WM];
    uint32_t bk[SU];
    bool nan_or_zero = false;
    auto rows_at = [&](int64_t t) { return r1 - t < TILE ? (int)(r1 - t) : TILE; };
    if (r0 < r1) gbp_load1<WM, SU>(L, A, G, ins, r0, rows_at(r0), valid_bits, w, bk, nan_or_zero);
    for (int64_t t0 = r0; t0 < r1; t0 += TILE) {
        const int n = rows_at(t0);
        gbp_stage_place<WM, SU>(S, W, NB, n, w, bk);
        if (t0 + TILE < r1)
            gbp_load1<WM, SU>(L, A, G, ins, t0 + TILE, rows_at(t0 + TILE), valid_bits, w, bk, nan_or_zero);
        gbp_stage_write<WM>(S, W, NB, n, G.S1);
    }
    if (nan_or_zero) atomicOr((unsigned long long *)&G.chk[2], 1ull);
}

// staged rows of the second pass's tile at t0 (n rows) with their sub-buckets
template <int WM, int SU>
__device__ __forceinline__ void gbp_load2(const GbpArgs &G, int64_t t0, int n, uint64_t (&w)[SU][WM],
                                          uint32_t (&bk)[SU]) {
    const int W = G.W;
#pragma unroll
    for (int u = 0; u < SU; u++) {
        const int j = u * GB_THREADS + (int)threadIdx.x;
        const int64_t rr = t0 + (j < n ? j : n - 1);
        const uint64_t *src = G.S1 + rr * W;
#pragma unroll
        for (int q = 0; q < WM; q++) w[u][q] = q < W ? src[q] : 0;
    }
#pragma unroll
    for (int u = 0; u < SU; u++) {
        uint64_t meta = 0;
#pragma unroll
        for (int q = 0; q < WM; q++)
            if (q == W - 1) meta = w[u][q];
        const int cls = G.nometa ? gbp_plain_class(w[u][0]) : (int)((meta >> 32) & 3);
        bk[u] = (uint32_t)((gbp_pkey(G, cls, w[u][0]) >> (64 - G.B1 - G.B2)) & ((1u << G.B2) - 1));
    }
}

// four-word staged rows of the second pass's tile at t0 as chunks
template <int SU>
__device__ __forceinline__ void gbp_load2c(const GbpArgs &G, int64_t t0, int n, ulonglong2 (&ch)[2 * SU],
                                           uint32_t (&bk)[2 * SU]) {
    const bool half = threadIdx.x & 1;
#pragma unroll
    for (int u = 0; u < 2 * SU; u++) {
        const int row = (u * GB_THREADS + (int)threadIdx.x) >> 1;
        const int64_t rr = t0 + (row < n ? row : n - 1);
        ch[u] = reinterpret_cast<const ulonglong2 *>(G.S1 + rr * 4)[half];
    }
#pragma unroll
    for (int u = 0; u < 2 * SU; u++) {
        // half 0 holds (key, input 0), half 1 (input 1, meta)
        const uint64_t mine = half ? ch[u].y : ch[u].x;
        const uint32_t lo = __shfl_xor((uint32_t)mine, 1, VG_WAVE);
        const uint32_t hi = __shfl_xor((uint32_t)(mine >> 32), 1, VG_WAVE);
        const uint64_t other = ((uint64_t)hi << 32) | lo;
        const uint64_t key = half ? other : mine, meta = half ? mine : other;
        const int cls = (int)((meta >> 32) & 3);
        bk[u] = (uint32_t)((gbp_pkey(G, cls, key) >> (64 - G.B1 - G.B2)) & ((1u << G.B2) - 1));
    }
}

// the staged form of the second pass (see k_gbp_scatter2)
template <int WM, int SU>
__global__ void __launch_bounds__(GB_THREADS) k_gbp_scatter2s(GbpArgs G) {
    extern __shared__ uint64_t dyn_st[];
    constexpr int TILE = SU * GB_THREADS;
    int64_t b, chunk, nblk, r0, r1;
    if (!gbp_block2(G, blockIdx.x, b, chunk, nblk, r0, r1)) return;
    const int NS = 1 << G.B2, W = G.W;
    GbpStage S = gbp_stage_carve(dyn_st, W, NS, TILE);
    for (int i = threadIdx.x; i < NS; i += GB_THREADS) {
        S.cur[i] = G.O2[gbp_h2_index(G, b, i, chunk, nblk)];
        S.cnt[i] = 0;
    }
    __syncthreads();
    auto rows_at = [&](int64_t t) { return r1 - t < TILE ? (int)(r1 - t) : TILE; };
    if constexpr (WM == 4) {
        ulonglong2 ch[2 * SU];
        uint32_t bk[2 * SU];
        if (r0 < r1) gbp_load2c<SU>(G, r0, rows_at(r0), ch, bk);
        for (int64_t t0 = r0; t0 < r1; t0 += TILE) {
            const int n = rows_at(t0);
            gbp_stage_place4c<SU>(S, NS, n, ch, bk);
            if (t0 + TILE < r1) gbp_load2c<SU>(G, t0 + TILE, rows_at(t0 + TILE), ch, bk);
            gbp_stage_write<4>(S, 4, NS, n, G.S2);
        }
    } else {
        uint64_t w[SU][WM];
        uint32_t bk[SU];
        if (r0 < r1) gbp_load2<WM, SU>(G, r0, rows_at(r0), w, bk);
        for (int64_t t0 = r0; t0 < r1; t0 += TILE) {
            const int n = rows_at(t0);
            gbp_stage_place<WM, SU>(S, W, NS, n, w, bk);
            if (t0 + TILE < r1) gbp_load2<WM, SU>(G, t0 + TILE, rows_at(t0 + TILE), w, bk);
            gbp_stage_write<WM>(S, W, NS, n, G.S2);
        }
    }
}

// row start of every partition (P + 1 entries) and its item count
__global__ void k_gbp_pstart(GbpArgs G) {
    int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p > G.P) return;
    const int64_t NB = 1 << G.B1;
    uint64_t start;
    if (p == G.P) {
        start = G.O1[NB * G.G1];
    } else {
        int64_t b = p >> G.B2, sub = p & (((int64_t)1 << G.B2) - 1);
        int64_t nblk = (int64_t)(G.blkbase2[b + 1] - G.blkbase2[b]);
        start = nblk ? G.O2[gbp_h2_index(G, b, sub, 0, nblk)] : G.O1[b * G.G1];
    }
    G.pstart[p] = start;
    if (p == 0) G.chk[1] = 0;
}

__global__ void k_gbp_nitems(GbpArgs G) {
    int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= G.P) return;
    uint64_t cnt = G.pstart[p + 1] - G.pstart[p];
    const uint32_t items = (uint32_t)((cnt + GBP_CH - 1) / GBP_CH);
    G.nitems[p] = items;
    if (G.range && items > 1 && p < G.P - 1)   // not the NA / EMPTY-valued key partition
        atomicMax((unsigned long long *)&G.chk[1], (unsigned long long)items);
}

// rows of the largest first-pass bucket but the last (range partitioning;
// the last holds the NA and EMPTY-valued key rows)
__global__ void __launch_bounds__(1024) k_gbp_bmax(GbpArgs G) {
    __shared__ unsigned long long m;
    if (threadIdx.x == 0) m = 0;
    __syncthreads();
    const int64_t NB = (int64_t)1 << G.B1;
    unsigned long long mine = 0;
    for (int64_t b = threadIdx.x; b < NB - 1; b += 1024) {
        const unsigned long long size = G.O1[(b + 1) * G.G1] - G.O1[b * G.G1];
        mine = size > mine ? size : mine;
    }
    atomicMax(&m, mine);
    __syncthreads();
    if (threadIdx.x == 0) G.chk[0] = m;
}

// initial (empty) NA and EMPTY-key records, multi-item list counter
__global__ void k_gbp_init_reserved(GbLayout L, GbpArgs G) {
    int w = threadIdx.x;
    const int64_t n = G.n_rows;
    for (int j = 0; j < 2; j++) {
        uint64_t *r = G.rec + (n + j) * G.RW;
        if (w < L.RH) r[w] = w == 0 ? 0 : (((L.hot_init_ones >> w) & 1) ? ~0ull : 0ull);
        else if (w < L.RH + L.RC) r[w] = ((L.cold_init_ones >> (w - L.RH)) & 1) ? ~0ull : 0ull;
    }
    if (w == 0) {
        G.ctl[10] = 0;
        G.ctl[12] = 0;   // k_gbp_agg_loop's item ticket
    }
}

// LDS scratch of the binned key sort of an item's records
#define GBP_SORT_BINS 1024
struct GbpSortLds {
    uint32_t bin[GBP_SORT_BINS + 1];    // counts, then their exclusive scan
    uint16_t tmp[512];                  // record indexes grouped by bin
    uint64_t red[2 * GB_THREADS / VG_WAVE];
};

// Order of n (<= 512) distinct int64 keys kc: slots src[i] into dst at the
// rank of kc[i].  Keys are binned by a monotone function of their place in
// [min, max] (GBP_SORT_BINS bins), the bins counted and scanned, and a key
// ranked inside its bin by comparing it with the bin's other keys only: for
// keys spread over their range (the records of one key-range partition)
// a bin holds one or two, so the sort costs a few LDS operations per key
// instead of n comparisons; any spread gives the exact order.
__device__ __forceinline__ void gbp_sort_binned(GbpSortLds &B, const int64_t *kc, int n, const uint16_t *src,
                                                uint16_t *dst) {
    const uint64_t SIGN = 0x8000000000000000ull;
    uint64_t mn = ~0ull, mx = 0;
    for (int i = threadIdx.x; i < n; i += GB_THREADS) {
        const uint64_t o = (uint64_t)kc[i] ^ SIGN;
        mn = o < mn ? o : mn;
        mx = o > mx ? o : mx;
    }
    for (int d = VG_WAVE / 2; d >= 1; d >>= 1) {
        const uint64_t a = __shfl_xor(mn, d, VG_WAVE), b = __shfl_xor(mx, d, VG_WAVE);
        mn = a < mn ? a : mn;
        mx = b > mx ? b : mx;
    }
    const int wid = threadIdx.x >> 6;
    if (vg_lane() == 0) {
        B.red[2 * wid] = mn;
        B.red[2 * wid + 1] = mx;
    }
    for (int b = threadIdx.x; b <= GBP_SORT_BINS; b += GB_THREADS) B.bin[b] = 0;
    __syncthreads();
    mn = B.red[0];
    mx = B.red[1];
    for (int w = 1; w < GB_THREADS / VG_WAVE; w++) {
        mn = B.red[2 * w] < mn ? B.red[2 * w] : mn;
        mx = B.red[2 * w + 1] > mx ? B.red[2 * w + 1] : mx;
    }
    const uint64_t span = mx - mn;
    const int bits = span ? 64 - __builtin_clzll(span) : 0;
    const int sh = bits > 10 ? bits - 10 : 0;   // (o - mn) >> sh < 1024 = GBP_SORT_BINS
    constexpr int PER = (512 + GB_THREADS - 1) / GB_THREADS;
    uint32_t bn[PER], rk[PER];
#pragma unroll
    for (int j = 0; j < PER; j++) {
        const int i = (int)threadIdx.x + j * GB_THREADS;
        if (i < n) {
            bn[j] = (uint32_t)((((uint64_t)kc[i] ^ SIGN) - mn) >> sh);
            rk[j] = atomicAdd(&B.bin[bn[j]], 1u);
        }
    }
    __syncthreads();
    // exclusive scan of the bin counts (bin[GBP_SORT_BINS] = n afterwards)
    constexpr int BPER = GBP_SORT_BINS / GB_THREADS;
    uint32_t c[BPER], loc = 0;
#pragma unroll
    for (int q = 0; q < BPER; q++) {
        c[q] = B.bin[threadIdx.x * BPER + q];
        loc += c[q];
    }
    uint32_t total;
    uint32_t run = gb_block_prefix(loc, reinterpret_cast<uint32_t *>(B.red), total);
#pragma unroll
    for (int q = 0; q < BPER; q++) {
        B.bin[threadIdx.x * BPER + q] = run;
        run += c[q];
    }
    if (threadIdx.x == 0) B.bin[GBP_SORT_BINS] = total;
    __syncthreads();
#pragma unroll
    for (int j = 0; j < PER; j++) {
        const int i = (int)threadIdx.x + j * GB_THREADS;
        if (i < n) B.tmp[B.bin[bn[j]] + rk[j]] = (uint16_t)i;
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < PER; j++) {
        const int i = (int)threadIdx.x + j * GB_THREADS;
        if (i >= n) continue;
        const int64_t key = kc[i];
        const uint32_t b0 = B.bin[bn[j]], b1 = B.bin[bn[j] + 1];
        uint32_t r = b0;
        for (uint32_t q = b0; q < b1; q++) r += kc[B.tmp[q]] < key;
        dst[r] = src[i];
    }
    __syncthreads();
}

// copies the occupied records of an LDS table (slots [0, S) regular, S at
// most 512) to dst, compacted, with coalesced stores; sorted: in ascending
// order of their int64 keys, each record placed at its rank (the keys of a
// table are distinct), counted against a compact copy of the keys in kc
// whose reads are the same address across a wave.  Returns the count in
// every thread.
__device__ __forceinline__ uint32_t gbp_write_records(const GbLayout &L, uint64_t *lhot, uint64_t *lcold,
                                                      int S, uint64_t empty, uint64_t *dst, int RW, uint16_t *slist,
                                                      uint32_t *scan_lds, bool sorted, int64_t *kc,
                                                      bool reset = false, GbpSortLds *bs = nullptr) {
    const int per = (S + GB_THREADS - 1) / GB_THREADS;
    uint32_t mine = 0;
    for (int j = 0; j < per; j++) {
        int s = threadIdx.x * per + j;
        if (s < S && lhot[(size_t)s * L.RH] != empty) mine++;
    }
    uint32_t total;
    uint32_t off = gb_block_prefix(mine, scan_lds, total);
    for (int j = 0; j < per; j++) {
        int s = threadIdx.x * per + j;
        if (s < S && lhot[(size_t)s * L.RH] != empty) slist[off++] = (uint16_t)s;
    }
    __syncthreads();
    const uint16_t *order = slist;
    if (sorted && total > 1) {
        const int n = (int)total;
        for (int i = threadIdx.x; i < n; i += GB_THREADS) kc[i] = (int64_t)lhot[(size_t)slist[i] * L.RH];
        __syncthreads();
        if (bs) {
            gbp_sort_binned(*bs, kc, n, slist, slist + 512);
        } else {
            for (int i = threadIdx.x; i < n; i += GB_THREADS) {
                const int64_t key = kc[i];
                int rank = 0;
#pragma unroll 8
                for (int j = 0; j < n; j++) rank += kc[j] < key;
                slist[512 + rank] = slist[i];
            }
            __syncthreads();
        }
        order = slist + 512;
    }
    // word q of the output is word q % RW of record q / RW (the quotient by
    // a multiply: exact for q < 2^32 / RW)
    const uint32_t mag = (uint32_t)((0xFFFFFFFFull + (uint64_t)RW) / (uint64_t)RW);
    const int words = (int)total * RW;
    for (int q = threadIdx.x; q < words; q += GB_THREADS) {
        const int k = (int)__umulhi((uint32_t)q, mag), w = q - k * RW;
        const int s = order[k];
        uint64_t *src = w < L.RH ? lhot + (size_t)s * L.RH + w : lcold + (size_t)s * L.RC + (w - L.RH);
        dst[q] = *src;
        // reset: the word back to its initial value (one thread per word)
        if (reset)
            *src = w == 0 ? empty
                   : w < L.RH ? (((L.hot_init_ones >> w) & 1) ? ~0ull : 0ull)
                              : (((L.cold_init_ones >> (w - L.RH)) & 1) ? ~0ull : 0ull);
    }
    return total;
}

// a record of an LDS table back to its initial (empty) value
__device__ __forceinline__ void gbp_reset_record(const GbLayout &L, uint64_t *h, uint64_t *c, uint64_t empty) {
    h[0] = empty;
    for (int w = 1; w < L.RH; w++) h[w] = ((L.hot_init_ones >> w) & 1) ? ~0ull : 0ull;
    for (int w = 0; w < L.RC; w++) c[w] = ((L.cold_init_ones >> w) & 1) ? ~0ull : 0ull;
}

__device__ __forceinline__ void gbp_copy_record(const GbLayout &L, const uint64_t *h, const uint64_t *c, uint64_t *dst) {
    for (int w = 0; w < L.RH; w++) dst[w] = h[w];
    for (int w = 0; w < L.RC; w++) dst[L.RH + w] = c[w];
}

// LDS slot of a staged or partial record's key, or -1 (table full): a
// partition holds a bounded number of groups, so probing may scan the
// whole table
__device__ __forceinline__ int64_t gbp_lds_slot(uint64_t *lhot, int S, int RH, int cls, uint64_t kn, uint64_t empty) {
    if (cls == VG_KEY_NA) return S;
    if (cls == VG_KEY_SPECIAL) return S + 1;
    return gb_slot(lhot, S, RH, kn, empty, S, nullptr);
}

// a staged row (key, inputs, meta word) into one record
__device__ __forceinline__ void gbp_apply_staged(const GbLayout &L, uint64_t *h, uint64_t *c, const uint64_t *w,
                                                 uint64_t meta, uint64_t pos) {
    gb_update_row(L, h, pos);
    for (int i = 0; i < L.n_inputs; i++) {
        InVal v;
        v.valid = (meta >> (34 + i)) & 1;
        v.bits = w[1 + i];
        gb_update_input(L, h, c, i, pos, v);
    }
}

// one staged row (G.W words: key, inputs, meta) into the item's LDS table,
// or the HBM table when the LDS table has no room for its key
__device__ __forceinline__ void gbp_agg_row(const GbLayout &L, const GbpArgs &G, uint64_t *lhot, uint64_t *lcold,
                                            int S, uint64_t empty, const uint64_t *w, bool &spilled) {
    const uint64_t kn = w[0], meta = w[G.W - 1];
    const int cls = (int)((meta >> 32) & 3);
    const uint64_t pos = (uint64_t)G.pos_base + (uint32_t)meta;
    const int64_t ls = gbp_lds_slot(lhot, S, L.RH, cls, kn, empty);
    // one expansion per address space: ds_* atomics for the LDS record
    if (ls >= 0) {
        gbp_apply_staged(L, lhot + (size_t)ls * L.RH, lcold + (size_t)ls * L.RC, w, meta, pos);
    } else {
        const int64_t gs = gb_global_slot(L, G.hot, G.cap, cls, kn, pos, G.ctl);
        spilled = true;
        if (gs >= 0) gbp_apply_staged(L, G.hot + (size_t)gs * L.RH, G.cold + (size_t)gs * L.RC, w, meta, pos);
    }
}

// gbp_apply_staged / gbp_agg_row for rows of a width known at compile time
// (W = inputs + 2): the row's words are indexed by constants only, so they
// stay in registers
template <int W, bool META>
__device__ __forceinline__ void gbp_apply_w(const GbLayout &L, uint64_t *h, uint64_t *c, const uint64_t (&w)[W],
                                            uint64_t meta, uint64_t pos) {
    gb_update_row(L, h, pos);
#pragma unroll
    for (int i = 0; i < (META ? W - 2 : W - 1); i++) {
        InVal v;
        v.valid = (meta >> (34 + i)) & 1;
        v.bits = w[1 + i];
        gb_update_input(L, h, c, i, pos, v);
    }
}

// META false: rows of key and inputs only (GbpArgs::nometa): the key class
// from the key, every input valid, position 0
template <int W, bool META>
__device__ __forceinline__ void gbp_agg_row_w(const GbLayout &L, const GbpArgs &G, uint64_t *lhot, uint64_t *lcold,
                                              int S, uint64_t empty, const uint64_t (&w)[W], bool &spilled) {
    constexpr int NI = META ? W - 2 : W - 1;
    const uint64_t kn = w[0];
    const uint64_t meta = META ? w[W - 1]
                               : ((uint64_t)gbp_plain_class(kn) << 32) | ((((uint64_t)1 << NI) - 1) << 34);
    const int cls = (int)((meta >> 32) & 3);
    const uint64_t pos = (uint64_t)G.pos_base + (uint32_t)meta;
    const int64_t ls = gbp_lds_slot(lhot, S, L.RH, cls, kn, empty);
    // one expansion per address space: ds_* atomics for the LDS record
    if (ls >= 0) {
        gbp_apply_w<W, META>(L, lhot + (size_t)ls * L.RH, lcold + (size_t)ls * L.RC, w, meta, pos);
    } else {
        const int64_t gs = gb_global_slot(L, G.hot, G.cap, cls, kn, pos, G.ctl);
        spilled = true;
        if (gs >= 0) gbp_apply_w<W, META>(L, G.hot + (size_t)gs * L.RH, G.cold + (size_t)gs * L.RC, w, meta, pos);
    }
}

__global__ void __launch_bounds__(GB_THREADS) k_gbp_agg(GbLayout L, GbpArgs G) {
    extern __shared__ uint64_t dynw[];   // 8-byte aligned after the static arrays
    __shared__ uint16_t slist[1024];   // [0, 512) occupied slots, [512, 1024) in key order
    __shared__ int64_t kc[512];
    __shared__ uint32_t scan_lds[GB_THREADS / VG_WAVE];
    const int64_t item = blockIdx.x;
    const int64_t NI = (int64_t)G.itembase[G.P];
    if (item >= NI) {
        if (threadIdx.x == 0 && item < G.NImax) G.icount[item] = 0;
        return;
    }
    const int S = G.S;
    uint64_t *lhot = dynw;
    uint64_t *lcold = lhot + (size_t)(S + 2) * L.RH;
    const uint64_t empty = L.key_is_double ? VG_EMPTY_F64 : VG_EMPTY_I64;
    gb_lds_init(L, lhot, lcold, S, empty);
    const int64_t p = gbp_search(G.itembase, G.P + 1, (uint64_t)item);
    const int64_t i0 = (int64_t)G.itembase[p], nit = (int64_t)G.itembase[p + 1] - i0, chunk = item - i0;
    const int64_t pe = (int64_t)G.pstart[p + 1];
    const int64_t r0 = (int64_t)G.pstart[p] + chunk * GBP_CH;
    const int64_t r1 = r0 + GBP_CH < pe ? r0 + GBP_CH : pe;
    bool spilled = false;
    if (G.W == 4) {
        // the two-input row: four rows per thread loaded before any is applied
        for (int64_t r = r0 + threadIdx.x; r < r1; r += (int64_t)GB_THREADS * 4) {
            ulonglong2 a[4], b[4];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const int64_t rr = r + (int64_t)u * GB_THREADS;
                a[u] = b[u] = make_ulonglong2(0, 0);
                if (rr < r1) {
                    const ulonglong2 *src = reinterpret_cast<const ulonglong2 *>(G.S2 + rr * 4);
                    a[u] = src[0];
                    b[u] = src[1];
                }
            }
#pragma unroll
            for (int u = 0; u < 4; u++) {
                if (r + (int64_t)u * GB_THREADS >= r1) continue;
                const uint64_t w[4] = {a[u].x, a[u].y, b[u].x, b[u].y};
                gbp_agg_row(L, G, lhot, lcold, S, empty, w, spilled);
            }
        }
    } else {
        for (int64_t r = r0 + threadIdx.x; r < r1; r += GB_THREADS)
            gbp_agg_row(L, G, lhot, lcold, S, empty, G.S2 + r * G.W, spilled);
    }
    if (spilled) G.ctl[9] = 1;
    __syncthreads();
    uint32_t total =
        gbp_write_records(L, lhot, lcold, S, empty, G.rec + r0 * G.RW, G.RW, slist, scan_lds, G.range != 0, kc);
    if (threadIdx.x == 0) {
        G.icount[item] = total;
        G.irec[item] = (uint64_t)r0;
        uint32_t fl = 0;
        int64_t at = r0 + total;
        for (int j = 0; j < 2; j++) {
            const uint64_t *h = lhot + (size_t)(S + j) * L.RH;
            if (h[1] == 0) continue;
            const uint64_t *c = lcold + (size_t)(S + j) * L.RC;
            if (nit == 1) {
                gbp_copy_record(L, h, c, G.rec + (G.n_rows + j) * G.RW);
            } else {
                gbp_copy_record(L, h, c, G.rec + at * G.RW);
                at++;
                fl |= 1u << j;
            }
        }
        G.rflags[item] = fl;
        if (nit > 1 && chunk == 0) G.mlist[atomicAdd((unsigned long long *)&G.ctl[10], 1ull)] = (uint32_t)p;
    }
}

// after an item's regular records are written (thread 0): its counts,
// its NA / EMPTY-valued key records (final when the partition has one
// item, partials behind the regular records otherwise), the LDS words of
// those two records back to their initial values, and the partition on
// the merge list when this is its first item of several
__device__ __forceinline__ void gbp_item_tail(const GbLayout &L, const GbpArgs &G, uint64_t *lhot, uint64_t *lcold,
                                              int S, uint64_t empty, int64_t item, int64_t p, int64_t nit,
                                              int64_t chunk, int64_t r0, uint32_t total) {
    G.icount[item] = total;
    G.irec[item] = (uint64_t)r0;
    uint32_t fl = 0;
    int64_t at = r0 + total;
    for (int j = 0; j < 2; j++) {
        uint64_t *h = lhot + (size_t)(S + j) * L.RH;
        if (h[1] == 0) continue;
        uint64_t *c = lcold + (size_t)(S + j) * L.RC;
        if (nit == 1) {
            gbp_copy_record(L, h, c, G.rec + (G.n_rows + j) * G.RW);
        } else {
            gbp_copy_record(L, h, c, G.rec + at * G.RW);
            at++;
            fl |= 1u << j;
        }
        gbp_reset_record(L, h, c, empty);
    }
    G.rflags[item] = fl;
    if (nit > 1 && chunk == 0) G.mlist[atomicAdd((unsigned long long *)&G.ctl[10], 1ull)] = (uint32_t)p;
}

// The aggregation with resident blocks: a block takes items in turn (an
// atomic ticket), reads its partition from ipart instead of searching for
// it, and sets only the LDS words it copied out back to their initial
// values, so the table is cleared once per block rather than once per item.
// Same records, flags and lists as k_gbp_agg.
__global__ void __launch_bounds__(GB_THREADS) k_gbp_agg_loop(GbLayout L, GbpArgs G) {
    extern __shared__ uint64_t dynw[];   // 8-byte aligned after the static arrays
    __shared__ uint16_t slist[1024];   // [0, 512) occupied slots, [512, 1024) in key order
    __shared__ int64_t kc[512];
    __shared__ uint32_t scan_lds[GB_THREADS / VG_WAVE];
    __shared__ int64_t ticket;
    __shared__ GbpSortLds bsort;
    const int S = G.S;
    uint64_t *lhot = dynw;
    uint64_t *lcold = lhot + (size_t)(S + 2) * L.RH;
    const uint64_t empty = L.key_is_double ? VG_EMPTY_F64 : VG_EMPTY_I64;
    gb_lds_init(L, lhot, lcold, S, empty);
    const int64_t NI = (int64_t)G.itembase[G.P];
    bool spilled = false;
    for (;;) {
        if (threadIdx.x == 0) ticket = (int64_t)atomicAdd((unsigned long long *)&G.ctl[12], 1ull);
        __syncthreads();
        const int64_t item = ticket;
        if (item >= NI) break;
        const int64_t p = G.ipart[item];
        const int64_t i0 = (int64_t)G.itembase[p], nit = (int64_t)G.itembase[p + 1] - i0, chunk = item - i0;
        const int64_t pe = (int64_t)G.pstart[p + 1];
        const int64_t r0 = (int64_t)G.pstart[p] + chunk * GBP_CH;
        const int64_t r1 = r0 + GBP_CH < pe ? r0 + GBP_CH : pe;
        if (G.W == 4) {
            for (int64_t r = r0 + threadIdx.x; r < r1; r += (int64_t)GB_THREADS * 4) {
                ulonglong2 a[4], b[4];
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    const int64_t rr = r + (int64_t)u * GB_THREADS;
                    a[u] = b[u] = make_ulonglong2(0, 0);
                    if (rr < r1) {
                        const ulonglong2 *src = reinterpret_cast<const ulonglong2 *>(G.S2 + rr * 4);
                        a[u] = src[0];
                        b[u] = src[1];
                    }
                }
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    if (r + (int64_t)u * GB_THREADS >= r1) continue;
                    const uint64_t w[4] = {a[u].x, a[u].y, b[u].x, b[u].y};
                    gbp_agg_row(L, G, lhot, lcold, S, empty, w, spilled);
                }
            }
        } else {
            for (int64_t r = r0 + threadIdx.x; r < r1; r += GB_THREADS)
                gbp_agg_row(L, G, lhot, lcold, S, empty, G.S2 + r * G.W, spilled);
        }
        __syncthreads();
        uint32_t total = gbp_write_records(L, lhot, lcold, S, empty, G.rec + r0 * G.RW, G.RW, slist, scan_lds,
                                           G.range != 0, kc, true, &bsort);
        if (threadIdx.x == 0) gbp_item_tail(L, G, lhot, lcold, S, empty, item, p, nit, chunk, r0, total);
        __syncthreads();
    }
    if (spilled) G.ctl[9] = 1;
    for (int64_t it = NI + (int64_t)blockIdx.x * GB_THREADS + threadIdx.x; it < G.NImax;
         it += (int64_t)gridDim.x * GB_THREADS)
        G.icount[it] = 0;
}

// U staged rows of W words per thread, rows r, r + GB_THREADS, ...; rows
// at or past r1 read as zeros (never used)
template <int W, int U>
__device__ __forceinline__ void gbp_load_rows(const uint64_t *S2, int64_t r, int64_t r1, uint64_t (&v)[U][W]) {
#pragma unroll
    for (int u = 0; u < U; u++) {
        const int64_t rr = r + (int64_t)u * GB_THREADS;
        if (rr < r1) {
            const uint64_t *src = S2 + rr * W;
            if constexpr ((W & 1) == 0) {
#pragma unroll
                for (int k = 0; k < W; k += 2) {
                    const ulonglong2 q = *reinterpret_cast<const ulonglong2 *>(src + k);
                    v[u][k] = q.x;
                    v[u][k + 1] = q.y;
                }
            } else {
#pragma unroll
                for (int k = 0; k < W; k++) v[u][k] = src[k];
            }
        } else {
#pragma unroll
            for (int k = 0; k < W; k++) v[u][k] = 0;
        }
    }
}

// The aggregation with resident blocks taking items in a fixed stride
// (block b: items b, b + grid, ...), so a block knows its next item while
// it works on the current one: the next item's bounds are loaded while the
// current rows are applied, and its first U rows per thread while the
// current records are sorted and written.  The barriers of that phase
// wait for LDS only, so those loads stay in flight across them.  Same
// records, flags and lists as k_gbp_agg.
template <int W, int U, bool META = true>
__global__ void __launch_bounds__(GB_THREADS) k_gbp_agg_pf(GbLayout L, GbpArgs G, const uint4 *__restrict__ idesc) {
    extern __shared__ uint64_t dynw[];   // 8-byte aligned after the static arrays
    __shared__ uint16_t slist[1024];   // [0, 512) occupied slots, [512, 1024) in key order
    __shared__ int64_t kc[512];
    __shared__ uint32_t scan_lds[GB_THREADS / VG_WAVE];
    __shared__ GbpSortLds bsort;
    const int S = G.S;
    uint64_t *lhot = dynw;
    uint64_t *lcold = lhot + (size_t)(S + 2) * L.RH;
    const uint64_t empty = L.key_is_double ? VG_EMPTY_F64 : VG_EMPTY_I64;
    gb_lds_init(L, lhot, lcold, S, empty);
    const int64_t NI = (int64_t)G.itembase[G.P];
    const int64_t step = (int64_t)gridDim.x;
    bool spilled = false;
    uint64_t v[U][W];
    int64_t item = blockIdx.x;
    uint4 cur = make_uint4(0, 0, 0, 0);
    if (item < NI) {
        cur = idesc[item];
        gbp_load_rows<W, U>(G.S2, (int64_t)cur.x + threadIdx.x, cur.y, v);
    }
    for (; item < NI; item += step) {
        const bool more = item + step < NI;
        // (a scalar load: idesc is read-only here)
        const uint4 nxt = more ? idesc[item + step] : make_uint4(0, 0, 0, 0);
        const int64_t r0 = cur.x, r1 = cur.y;
#pragma unroll
        for (int u = 0; u < U; u++)
            if (r0 + threadIdx.x + (int64_t)u * GB_THREADS < r1)
                gbp_agg_row_w<W, META>(L, G, lhot, lcold, S, empty, v[u], spilled);
        for (int64_t r = r0 + threadIdx.x + (int64_t)U * GB_THREADS; r < r1; r += (int64_t)U * GB_THREADS) {
            uint64_t w[U][W];
            gbp_load_rows<W, U>(G.S2, r, r1, w);
#pragma unroll
            for (int u = 0; u < U; u++)
                if (r + (int64_t)u * GB_THREADS < r1) gbp_agg_row_w<W, META>(L, G, lhot, lcold, S, empty, w[u], spilled);
        }
        if (more) gbp_load_rows<W, U>(G.S2, (int64_t)nxt.x + threadIdx.x, nxt.y, v);
        __syncthreads();
        const uint32_t total = gbp_write_records(L, lhot, lcold, S, empty, G.rec + r0 * G.RW, G.RW, slist,
                                                 scan_lds, G.range != 0, kc, true, &bsort);
        if (threadIdx.x == 0)
            gbp_item_tail(L, G, lhot, lcold, S, empty, item, cur.z, (cur.w & 1) ? 2 : 1, (cur.w & 2) ? 0 : 1, r0,
                          total);
        __syncthreads();
        cur = nxt;
    }
    if (spilled) G.ctl[9] = 1;
    for (int64_t it = NI + (int64_t)blockIdx.x * GB_THREADS + threadIdx.x; it < G.NImax;
         it += (int64_t)gridDim.x * GB_THREADS)
        G.icount[it] = 0;
}

// partition and descriptor of every item (staged rows fit 32 bits: the
// partitioned path takes feeds of fewer than 2^32 - 2 rows)
__global__ void k_gbp_ipart(GbpArgs G) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= G.P) return;
    const uint64_t i0 = G.itembase[p], i1 = G.itembase[p + 1], pe = G.pstart[p + 1];
    for (uint64_t i = i0; i < i1; i++) {
        G.ipart[i] = (uint32_t)p;
        const uint64_t r0 = G.pstart[p] + (i - i0) * GBP_CH, r1 = r0 + GBP_CH < pe ? r0 + GBP_CH : pe;
        G.idesc[i] = make_uint4((uint32_t)r0, (uint32_t)r1, (uint32_t)p, (i1 - i0 > 1 ? 1u : 0u) | (i == i0 ? 2u : 0u));
    }
}

// one block per partition with several items: merge their partial records
__global__ void __launch_bounds__(GB_THREADS) k_gbp_combine(GbLayout L, GbpArgs G) {
    extern __shared__ uint64_t dynw[];   // 8-byte aligned after the static arrays
    __shared__ uint16_t slist[1024];   // [0, 512) occupied slots, [512, 1024) in key order
    __shared__ int64_t kc[512];
    __shared__ uint32_t scan_lds[GB_THREADS / VG_WAVE];
    if ((uint64_t)blockIdx.x >= G.ctl[10]) return;
    const int S = G.S;
    uint64_t *lhot = dynw;
    uint64_t *lcold = lhot + (size_t)(S + 2) * L.RH;
    const uint64_t empty = L.key_is_double ? VG_EMPTY_F64 : VG_EMPTY_I64;
    gb_lds_init(L, lhot, lcold, S, empty);
    const int64_t p = G.mlist[blockIdx.x];
    const int64_t i0 = (int64_t)G.itembase[p], i1 = (int64_t)G.itembase[p + 1];
    bool spilled = false;
    for (int64_t it = i0; it < i1; it++) {
        const int64_t cnt = G.icount[it];
        const uint32_t fl = G.rflags[it];
        const int64_t nres = __builtin_popcount(fl);
        const uint64_t *base = G.rec + (int64_t)G.irec[it] * G.RW;
        for (int64_t k = threadIdx.x; k < cnt + nres; k += GB_THREADS) {
            const uint64_t *r = base + k * G.RW;
            int cls = VG_KEY_REGULAR;
            if (k >= cnt) cls = (k == cnt && (fl & 1)) ? VG_KEY_NA : VG_KEY_SPECIAL;
            int64_t ls = gbp_lds_slot(lhot, S, L.RH, cls, r[0], empty);
            if (ls >= 0) {
                gb_merge(L, lhot + (size_t)ls * L.RH, lcold + (size_t)ls * L.RC, r, r + L.RH);
            } else {
                int64_t gs = gb_global_slot(L, G.hot, G.cap, cls, r[0], 0, G.ctl);
                spilled = true;
                if (gs >= 0) gb_merge(L, G.hot + (size_t)gs * L.RH, G.cold + (size_t)gs * L.RC, r, r + L.RH);
            }
        }
    }
    if (spilled) G.ctl[9] = 1;
    __syncthreads();
    uint32_t total = gbp_write_records(L, lhot, lcold, S, empty, G.rec + (int64_t)G.irec[i0] * G.RW, G.RW, slist,
                                       scan_lds, G.range != 0, kc);
    for (int64_t it = i0 + 1 + threadIdx.x; it < i1; it += GB_THREADS) G.icount[it] = 0;
    if (threadIdx.x == 0) {
        G.icount[i0] = total;
        for (int j = 0; j < 2; j++) {
            const uint64_t *h = lhot + (size_t)(S + j) * L.RH;
            if (h[1]) gbp_copy_record(L, h, lcold + (size_t)(S + j) * L.RC, G.rec + (G.n_rows + j) * G.RW);
        }
    }
}

// finish of a run: (sort key, record) pairs per item at scanned offsets;
// a run of range partitions (records in key order) writes only the records
// from entry 1 on (entry 0 is for an EMPTY-valued key's record)
__global__ void __launch_bounds__(GB_THREADS)
k_gbp_collect(GbLayout L, RecSrc R, int order, const uint32_t *icount, const uint64_t *irec, const uint64_t *ioff,
              uint64_t *skeys, uint32_t *svals, uint64_t *por, uint64_t *pand, int range) {
    __shared__ uint64_t lor[2 * GB_THREADS / VG_WAVE];
    const int64_t item = blockIdx.x;
    const int64_t cnt = icount[item];
    const int64_t base = cnt ? (int64_t)irec[item] : 0;
    const uint64_t off = ioff[item];
    if (range) {
        for (int64_t k = threadIdx.x; k < cnt; k += GB_THREADS) svals[1 + off + k] = (uint32_t)(base + k);
        return;
    }
    uint64_t o = 0, a = ~0ull;
    for (int64_t k = threadIdx.x; k < cnt; k += GB_THREADS) {
        uint64_t sk = 0;
        gb_sort_entry(L, R, order, base + k, sk);
        skeys[off + k] = sk;
        svals[off + k] = (uint32_t)(base + k);
        o |= sk;
        a &= sk;
    }
    gb_block_or_and(o, a, lor);
    if (threadIdx.x == 0) {
        por[item] = o;
        pand[item] = a;
    }
}

// the reserved records of a run join the sort after the regular ones; in a
// run of range partitions (key order: the NA group is placed by the emit)
// an EMPTY-valued key's record takes entry 0, before the regular ones
// (ctl[11] = 1)
__global__ void k_gbp_reserved(GbLayout L, RecSrc R, int order, uint64_t *skeys, uint32_t *svals, uint64_t *ctl,
                               int range) {
    if (threadIdx.x != 0) return;
    ctl[8] = R.hot[R.na_slot * R.hstride + 1];
    ctl[11] = 0;
    uint64_t sk;
    if (range) {
        if (gb_sort_entry(L, R, order, R.sp_slot, sk)) {
            svals[0] = (uint32_t)R.sp_slot;
            ctl[11] = 1;
            ctl[5]++;
        }
        return;
    }
    const int64_t slots[2] = {R.sp_slot, R.na_slot};
    for (int j = 0; j < 2; j++) {
        if (!gb_sort_entry(L, R, order, slots[j], sk)) continue;
        uint64_t at = ctl[5]++;
        skeys[at] = sk;
        svals[at] = (uint32_t)slots[j];
        ctl[6] |= sk;
        ctl[7] &= sk;
    }
}

// a run's records into the hash table (before another feed, or when some
// record did not fit an LDS table)
__global__ void __launch_bounds__(GB_THREADS)
k_gbp_flush(GbLayout L, const uint64_t *rec, int RW, const uint32_t *icount, const uint64_t *irec, int64_t n_rows,
            uint64_t *hot, uint64_t *cold, int64_t cap, uint64_t *ctl) {
    const int64_t item = blockIdx.x;
    const int64_t cnt = icount[item];
    const int64_t base = cnt ? (int64_t)irec[item] : 0;
    for (int64_t k = threadIdx.x; k < cnt; k += GB_THREADS) {
        const uint64_t *r = rec + (base + k) * RW;
        int64_t gs = gb_global_slot(L, hot, cap, VG_KEY_REGULAR, r[0], 0, ctl);
        if (gs >= 0) gb_merge(L, hot + (size_t)gs * L.RH, cold + (size_t)gs * L.RC, r, r + L.RH);
    }
    if (item == 0 && threadIdx.x < 2) {
        const uint64_t *r = rec + (n_rows + threadIdx.x) * RW;
        int64_t gs = cap + threadIdx.x;   // NA -> cap, EMPTY-valued key -> cap + 1
        if (r[1]) gb_merge(L, hot + (size_t)gs * L.RH, cold + (size_t)gs * L.RC, r, r + L.RH);
    }
}

// ---- host side of the partitioned feed

static int gbp_lds_slots(const vgpu_groupby *g) {
    size_t rw = (size_t)(g->L.RH + g->L.RC) * 8;
    int S = (int)vg_tune(VG_TUNE_GB_PART_SLOTS);
    while (S >= 64 && (size_t)(S + 2) * rw > 61440) S >>= 1;
    return S >= 64 ? S : 0;
}

// can this feed take the partitioned path?  (the first feed of a group-by,
// large, with its group count estimated)
static bool gbp_possible(const vgpu_groupby *g, int64_t n_logical, int64_t groups) {
    return !g->any_fed && !g->L.key_none && groups > 0 && n_logical >= GB_SK_MIN_ROWS &&
           n_logical < ((int64_t)1 << 32) - 2 && gbp_lds_slots(g) > 0;
}

static bool gbp_wanted(const vgpu_groupby *g, int64_t n_logical, int64_t groups) {
    return gbp_possible(g, n_logical, groups) && groups >= vg_tune_for(VG_TUNE_GB_PART_MIN_GROUPS, groups);
}

// carves 256-byte aligned pieces out of one buffer
struct Carve {
    uint8_t *base;
    size_t off = 0;
    template <class T> T *take(size_t count) {
        T *p = reinterpret_cast<T *>(base + off);
        off += (count * sizeof(T) + 255) & ~(size_t)255;
        return p;
    }
};

static void gbp_layout(GbpArgs &G, uint8_t *base, size_t *bytes) {
    const int64_t NB = (int64_t)1 << G.B1, NS = (int64_t)1 << G.B2;
    const int64_t ML = std::min<int64_t>(G.P, G.n_rows / GBP_CH + 1);
    Carve c{base};
    G.H1 = c.take<uint32_t>(NB * G.G1);
    G.O1 = c.take<uint64_t>(NB * G.G1 + 1);
    G.blkbase2 = c.take<uint64_t>(NB + 1);
    G.H2 = c.take<uint32_t>(G.NB2max * NS);
    G.O2 = c.take<uint64_t>(G.NB2max * NS + 1);
    G.pstart = c.take<uint64_t>(G.P + 1);
    G.nitems = c.take<uint32_t>(G.P);
    G.itembase = c.take<uint64_t>(G.P + 1);
    G.ipart = c.take<uint32_t>(G.NImax);
    G.idesc = c.take<uint4>(G.NImax);
    G.icount = c.take<uint32_t>(G.NImax);
    G.irec = c.take<uint64_t>(G.NImax);
    G.rflags = c.take<uint32_t>(G.NImax);
    G.mlist = c.take<uint32_t>(ML);
    G.chk = c.take<uint64_t>(3);
    *bytes = c.off;
}

// bits of a range of ordered keys' differences (0 for a range of one key)
static int span_bits(uint64_t span) { return span ? 64 - __builtin_clzll(span) : 0; }

static int gbp_feed(vgpu_groupby *g, FeedArgs A, const InSet &ins, const PredInfo &pi, const vgpu_prog &prog,
                    const ColSet &pcs, int64_t groups, bool sketched) {
    hipStream_t st = vg_stream();
    const GbLayout &L = g->L;
    const int64_t n = A.n_logical;
    GbpArgs G;
    memset(&G, 0, sizeof G);
    G.W = L.n_inputs + 2;
    G.RW = L.RH + L.RC;
    G.S = gbp_lds_slots(g);
    int64_t want = std::max<int64_t>(64, 2 * groups / G.S);
    int B = 6;
    while (((int64_t)1 << B) < want && B < 2 * GBP_MAX_BITS) B++;
    G.B1 = (B + 1) / 2;
    G.B2 = B - G.B1;
    G.P = (int64_t)1 << B;
    G.G1 = (int)std::min<int64_t>(GBP_G1, (n + 4095) / 4096);
    G.T1 = ((n + G.G1 - 1) / G.G1 + GB_THREADS - 1) / GB_THREADS * GB_THREADS;
    G.n_rows = n;
    G.pos_base = A.pos_base;
    G.NB2max = ((int64_t)1 << G.B1) + n / GBP_T2 + 1;
    G.NImax = G.P + n / GBP_CH + 1;
    // key order over int64 keys whose range this feed's sketch saw: range
    // partitions (knob gb.part.range)
    if (sketched && g->sk_ranged && g->order == VGPU_ORDER_KEY && !L.key_is_double &&
        vg_tune_for(VG_TUNE_GB_PART_RANGE, groups) >= 2) {
        G.range = 1;
        G.omin = (uint64_t)g->sk_lo ^ 0x8000000000000000ull;
        // f keeps up to B + 8 bits of a key's place in the range: every
        // partition within about one part in 2^7 of its even share
        const uint64_t span = ((uint64_t)g->sk_hi ^ 0x8000000000000000ull) - G.omin;
        G.rshift = std::max(0, span_bits(span) - (B + 8));
        G.fmax = span >> G.rshift;
        G.fmul = (uint64_t)(((unsigned __int128)(G.P - 1) << 32) / ((unsigned __int128)G.fmax + 1));
    }
    size_t aux = 0;
    gbp_layout(G, nullptr, &aux);
    size_t stage = (size_t)n * G.W * 8;
    if (g->gbp_s1.reserve(stage) || g->gbp_s2.reserve(stage) || g->gbp_aux.reserve(aux) ||
        g->run_rec.reserve((size_t)(n + 2) * G.RW * 8))
        return -1;
    gbp_layout(G, g->gbp_aux.as<uint8_t>(), &aux);
    G.S1 = g->gbp_s1.as<uint64_t>();
    G.S2 = g->gbp_s2.as<uint64_t>();
    G.rec = g->run_rec.as<uint64_t>();
    // records that do not fit an LDS table fall back to the hash table
    if (g->cap == 0 && gb_alloc_table(g, pow2_at_least(2 * std::max<int64_t>(65536, groups / 16)))) return -1;
    G.hot = g->hot.as<uint64_t>();
    G.cold = g->cold.as<uint64_t>();
    G.ctl = g->ctl.as<uint64_t>();
    G.cap = g->cap;
    A.ctl = G.ctl;

    const int64_t NB = (int64_t)1 << G.B1, NS = (int64_t)1 << G.B2;
    size_t lds1 = (size_t)A.vm_lds + NB * 8;
    // Range partitions are only as balanced as the keys are spread over
    // their range.  Two checks, each one small read: the largest first-pass
    // bucket (before any row is moved) and the largest partition in items
    // (after both passes; a partition of many items is merged by one
    // block).  Either beyond 4x its even share plus slack -> hash partitions.
    const uint64_t bucket_limit = 4 * (uint64_t)(n / NB) + 4 * GBP_T2;
    const uint64_t item_limit = 4 * (uint64_t)((n / G.P + GBP_CH - 1) / GBP_CH) + 4;
    // the plain first pass (every row of the feed, no NA, 8-byte columns)
    bool plain = !A.sel && !pi.present && A.key.type == VGPU_INT64 && !A.key.validity;
    for (int i = 0; i < L.n_inputs; i++)
        plain = plain && !ins.c[i].validity && (ins.c[i].type == VGPU_INT64 || ins.c[i].type == VGPU_DOUBLE);
    auto hist1 = plain ? k_gbp_hist1<true> : k_gbp_hist1<false>;
    auto scatter1 = plain ? k_gbp_scatter1<true> : k_gbp_scatter1<false>;
    // staged scatters (knob gb.part.stage): rows of 2 to 4 or up to
    // VGPU_MAX_INPUTS + 2 words
    const int64_t stage_knob = vg_tune_for(VG_TUNE_GB_PART_STAGE, groups);
    // rows without the meta word (knob gb.part.nometa): a plain feed in key
    // order of one or two inputs, staged, aggregated by k_gbp_agg_pf
    G.nometa = plain && stage_knob >= 2 && g->order == VGPU_ORDER_KEY && L.n_inputs >= 1 && L.n_inputs <= 2 &&
               vg_tune_for(VG_TUNE_GB_PART_AGG, groups) >= 3 && vg_tune_for(VG_TUNE_GB_PART_NOMETA, groups) >= 2;
    bool staged = false;
    vg_kt_begin();
    for (;;) {
        G.W = L.n_inputs + (G.nometa ? 1 : 2);
        int64_t stage_mode = stage_knob;
        // 2,048-row tiles of wide rows would not fit a CU's LDS: 1,024-row ones
        if (stage_mode >= 3 && gbp_stage_bytes(G.W, 1 << std::max(G.B1, G.B2), 8 * GB_THREADS) > 96 * 1024)
            stage_mode = 2;
        staged = stage_mode >= 2;
        const int tile = stage_mode >= 3 ? 8 * GB_THREADS : 4 * GB_THREADS;
        auto scatter1s = stage_mode >= 3 ? (G.W == 4 ? k_gbp_scatter1s<4, 8> : G.W <= 3 ? k_gbp_scatter1s<3, 8>
                                                                                       : k_gbp_scatter1s<VGPU_MAX_INPUTS + 2, 8>)
                                         : (G.W == 4 ? k_gbp_scatter1s<4, 4> : G.W <= 3 ? k_gbp_scatter1s<3, 4>
                                                                                       : k_gbp_scatter1s<VGPU_MAX_INPUTS + 2, 4>);
        auto scatter2s = stage_mode >= 3 ? (G.W == 4 ? k_gbp_scatter2s<4, 8> : G.W <= 3 ? k_gbp_scatter2s<3, 8>
                                                                                       : k_gbp_scatter2s<VGPU_MAX_INPUTS + 2, 8>)
                                         : (G.W == 4 ? k_gbp_scatter2s<4, 4> : G.W <= 3 ? k_gbp_scatter2s<3, 4>
                                                                                       : k_gbp_scatter2s<VGPU_MAX_INPUTS + 2, 4>);
        hist1<<<G.G1, GB_THREADS, lds1, st>>>(A, G, pi, prog, pcs);
        VG_LAUNCH_CHECK();
        if (vg_scan_u32_to_u64(G.H1, G.O1, NB * G.G1, G.O1 + NB * G.G1, g->sort_tmp)) return -1;
        if (G.range) {
            uint64_t big = 0;
            k_gbp_bmax<<<1, 1024, 0, st>>>(G);
            VG_LAUNCH_CHECK();
            VG_TRY(hipMemcpyAsync(&big, G.chk, 8, hipMemcpyDeviceToHost, st));
            VG_TRY(hipStreamSynchronize(st));
            if (big > bucket_limit) {
                G.range = 0;
                continue;
            }
        }
        if (plain && staged)
            scatter1s<<<G.G1, GB_THREADS, gbp_stage_bytes(G.W, (int)NB, tile), st>>>(L, A, G, ins);
        else
            scatter1<<<G.G1, GB_THREADS, lds1, st>>>(L, A, G, ins, pi, prog, pcs);
        VG_LAUNCH_CHECK();
        k_gbp_plan2<<<1, 1024, 0, st>>>(G);
        VG_LAUNCH_CHECK();
        k_gbp_hist2<<<(unsigned)G.NB2max, GB_THREADS, 0, st>>>(G);
        VG_LAUNCH_CHECK();
        if (vg_scan_u32_to_u64(G.H2, G.O2, G.NB2max * NS, nullptr, g->sort_tmp)) return -1;
        if (staged)
            scatter2s<<<(unsigned)G.NB2max, GB_THREADS, gbp_stage_bytes(G.W, (int)NS, tile), st>>>(G);
        else
            k_gbp_scatter2<<<(unsigned)G.NB2max, GB_THREADS, 0, st>>>(G);
        VG_LAUNCH_CHECK();
        k_gbp_pstart<<<vg_blocks(G.P + 1, 256), 256, 0, st>>>(G);
        VG_LAUNCH_CHECK();
        k_gbp_nitems<<<vg_blocks(G.P, 256), 256, 0, st>>>(G);
        VG_LAUNCH_CHECK();
        if (G.range || G.nometa) {
            uint64_t chk[2] = {0, 0};
            VG_TRY(hipMemcpyAsync(chk, G.chk + 1, 16, hipMemcpyDeviceToHost, st));
            VG_TRY(hipStreamSynchronize(st));
            if (G.nometa && chk[1]) {   // a NaN or a zero that needs positions: rows with meta words
                G.nometa = 0;
                continue;
            }
            if (G.range && chk[0] > item_limit) {
                G.range = 0;
                continue;
            }
        }
        break;
    }
    if (vg_scan_u32_to_u64(G.nitems, G.itembase, G.P, G.itembase + G.P, g->sort_tmp)) return -1;
    k_gbp_init_reserved<<<1, 128, 0, st>>>(L, G);
    VG_LAUNCH_CHECK();
    size_t lds = (size_t)(G.S + 2) * G.RW * 8;
    if (staged) {
        k_gbp_ipart<<<vg_blocks(G.P, 256), 256, 0, st>>>(G);
        VG_LAUNCH_CHECK();
        // resident blocks: as many as the LDS of a CU holds (table + the
        // kernel's static arrays), at most 8 per CU
        const size_t per_block = lds + 1024 * 2 + 512 * 8 + 64 + sizeof(GbpSortLds);
        const int64_t bpc = std::max<int64_t>(1, std::min<int64_t>(8, (int64_t)(160 * 1024 / per_block)));
        const int64_t grid = std::min<int64_t>(G.NImax, (int64_t)vg_cus() * bpc);
        // knob gb.part.agg: 3 = fixed-stride items with the next item's rows
        // loaded ahead (rows of 3 to 6 words), 2 = items taken by ticket
        const int64_t agg = vg_tune_for(VG_TUNE_GB_PART_AGG, groups);
        auto pf = G.nometa ? (G.W == 2 ? k_gbp_agg_pf<2, 4, false> : G.W == 3 ? k_gbp_agg_pf<3, 4, false> : nullptr)
                : G.W == 3 ? k_gbp_agg_pf<3, 4> : G.W == 4 ? k_gbp_agg_pf<4, 4>
                : G.W == 5 ? k_gbp_agg_pf<5, 2> : G.W == 6 ? k_gbp_agg_pf<6, 2> : nullptr;
        if (agg >= 3 && pf) pf<<<(unsigned)grid, GB_THREADS, lds, st>>>(L, G, G.idesc);
        else k_gbp_agg_loop<<<(unsigned)grid, GB_THREADS, lds, st>>>(L, G);
    } else {
        k_gbp_agg<<<(unsigned)G.NImax, GB_THREADS, lds, st>>>(L, G);
    }
    VG_LAUNCH_CHECK();
    k_gbp_combine<<<(unsigned)std::min<int64_t>(G.P, n / GBP_CH + 1), GB_THREADS, lds, st>>>(L, G);
    VG_LAUNCH_CHECK();
    vg_kt_end();
    g->gbp = G;
    g->run = 1;
    g->any_fed = 1;
    g->path = G.range ? 2 : 1;
    return 0;
}

static RecSrc gbp_run_src(const vgpu_groupby *g) {
    const GbpArgs &G = g->gbp;
    RecSrc R;
    R.hot = G.rec;
    R.cold = G.rec + g->L.RH;
    R.hstride = G.RW;
    R.cstride = G.RW;
    R.na_slot = G.n_rows;
    R.sp_slot = G.n_rows + 1;
    return R;
}

// folds the run into the hash table
static int gbp_flush(vgpu_groupby *g) {
    const GbpArgs &G = g->gbp;
    hipStream_t st = vg_stream();
    if (g->cmp.reserve((size_t)(G.NImax + 1) * 8)) return -1;
    uint64_t *ioff = g->cmp.as<uint64_t>();
    if (vg_scan_u32_to_u64(G.icount, ioff, G.NImax, ioff + G.NImax, g->sort_tmp)) return -1;
    uint64_t groups = 0;
    VG_TRY(hipMemcpyAsync(&groups, ioff + G.NImax, 8, hipMemcpyDeviceToHost, st));
    VG_TRY(hipStreamSynchronize(st));
    if (gb_ensure_capacity(g, (int64_t)groups + 2)) return -1;
    k_gbp_flush<<<(unsigned)G.NImax, GB_THREADS, 0, st>>>(g->L, G.rec, G.RW, G.icount, G.irec, G.n_rows,
                                                          g->hot.as<uint64_t>(), g->cold.as<uint64_t>(), g->cap,
                                                          g->ctl.as<uint64_t>());
    VG_LAUNCH_CHECK();
    g->run = 0;
    return 0;
}

// sort pairs of a run's records (the counterpart of gb_compact)
static int gbp_collect(vgpu_groupby *g, const RecSrc &R) {
    const GbpArgs &G = g->gbp;
    hipStream_t st = vg_stream();
    int64_t NI = G.NImax;
    if (g->sort_keys.reserve((size_t)(G.n_rows + 2) * 8) || g->sort_vals.reserve((size_t)(G.n_rows + 2) * 4) ||
        g->cmp.reserve((size_t)NI * 24 + 256))
        return -1;
    uint64_t *ioff = g->cmp.as<uint64_t>();
    uint64_t *por = ioff + NI, *pand = por + NI;
    uint64_t *ctl = g->ctl.as<uint64_t>();
    if (vg_scan_u32_to_u64(G.icount, ioff, NI, ctl + 5, g->sort_tmp)) return -1;
    k_gbp_collect<<<(unsigned)NI, GB_THREADS, 0, st>>>(g->L, R, g->order, G.icount, G.irec, ioff,
                                                       g->sort_keys.as<uint64_t>(), g->sort_vals.as<uint32_t>(), por,
                                                       pand, G.range);
    VG_LAUNCH_CHECK();
    if (!G.range) {
        k_gb_or_and_final<<<vg_blocks(NI, 4 * GB_THREADS, 256), GB_THREADS, 0, st>>>(por, pand, NI, ctl);
        VG_LAUNCH_CHECK();
    }
    k_gbp_reserved<<<1, 64, 0, st>>>(g->L, R, g->order, g->sort_keys.as<uint64_t>(), g->sort_vals.as<uint32_t>(),
                                     ctl, G.range);
    VG_LAUNCH_CHECK();
    return 0;
}

// LDS slots for privatization: the table must hold the expected groups at
// load <= 1/2 within the LDS budget
// groups the group-by expects (hint, sketch or groups seen; 0 = unknown)
static int64_t gb_groups(const vgpu_groupby *g) {
    return g->L.key_none ? 1 : g->hint > 0 ? g->hint : g->est_groups;
}

static int gb_lds_slots(const vgpu_groupby *g, size_t vm_bytes, bool cold = true) {
    int64_t groups = gb_groups(g);
    if (groups <= 0) return 0;
    size_t rec = (size_t)(g->L.RH + (cold ? g->L.RC : 0)) * 8;
    size_t lim = (size_t)vg_tune_for(VG_TUNE_GB_LDS_BUDGET, groups);
    size_t budget = lim > vm_bytes ? lim - vm_bytes : 0;
    // the NA and EMPTY-valued key groups the count includes have reserved
    // records of their own: only the rest are hashed
    const int64_t hashed = std::max<int64_t>(groups - 2, 1);
    const int64_t max_slots = vg_tune_for(VG_TUNE_GB_LDS_MAX_SLOTS, groups);
    int64_t want = pow2_at_least(vg_tune_for(VG_TUNE_GB_LDS_SLOTS_PER_GROUP, groups) * hashed);
    // a table that does not fit at the tuned load may still fit at a load
    // of up to 0.8: far better than every row updating HBM records
    if (want > max_slots || (size_t)(want + 2) * rec > budget) want = pow2_at_least(hashed + hashed / 4 + 1);
    if (want > max_slots || (size_t)(want + 2) * rec > budget) return 0;
    return (int)want;
}

static unsigned gb_grid(int64_t n_logical, int lds_slots, size_t lds_bytes, int64_t groups) {
    // privatized blocks amortize their merge over many rows; no more of
    // them than fit on the device at once (a second, partial round of
    // blocks would leave most CUs idle)
    if (lds_slots) {
        const int64_t cap = std::min<int64_t>(vg_tune_for(VG_TUNE_GB_LDS_MAX_BLOCKS, groups),
                                              vg_resident_blocks(lds_bytes));
        // a short feed (a row group of a few 100K rows) still spreads over
        // the CUs: at least one block per 2,048 rows, up to one per CU
        const int64_t floor = std::min<int64_t>(vg_cus(), (n_logical + 2047) / 2048);
        return (unsigned)std::max<int64_t>(vg_blocks(n_logical, vg_tune_for(VG_TUNE_GB_LDS_ROWS, groups), cap),
                                           std::min<int64_t>(floor, cap));
    }
    return vg_blocks(n_logical, vg_tune_for(VG_TUNE_GB_GLB_ROWS, groups),
                     vg_tune_for(VG_TUNE_GB_GLB_MAX_BLOCKS, groups));
}

// n / count / sum / mean of at most two inputs, a fast predicate or none
static bool gb_narrow_shape(const GbLayout &L, const PredInfo &pi) {
    if (L.n_inputs > 2 || (pi.present && !pi.fast)) return false;
    for (int i = 0; i < L.n_inputs; i++)
        if (L.w_min[i] >= 0 || L.w_max[i] >= 0 || L.w_fvpos[i] >= 0) return false;
    return true;
}

// block partials of an LDS-table feed of `grid` blocks with S hashed or
// dense slots (knob gb.merge.partials): room for every block's S + 2 full
// records, when the feed has enough blocks for HBM merges to contend, the
// partials stay within 64 MB and a fold block's table fits in 64 KB of LDS;
// *part = nullptr otherwise (the blocks merge into HBM directly)
static int gb_part_setup(vgpu_groupby *g, int64_t grid, int S, uint64_t **part) {
    *part = nullptr;
    const size_t rw = (size_t)(g->L.RH + g->L.RC) * 8;
    const size_t bytes = (size_t)grid * (S + 2) * rw;
    if (S <= 0 || grid < 64 || vg_tune(VG_TUNE_GB_MERGE_PARTIALS) < 2 || bytes > ((size_t)64 << 20) ||
        (size_t)(S + 2) * rw > 65536)
        return 0;
    if (g->part.reserve(bytes)) return -1;
    *part = g->part.as<uint64_t>();
    return 0;
}

#define GB_FOLD_PER 32   // partials per fold block of a hashed table

// folds the partials of `grid` blocks into the HBM table
static int gb_part_fold(vgpu_groupby *g, const uint64_t *part, int64_t grid, int S, bool dense, int64_t kmin) {
    hipStream_t st = vg_stream();
    uint64_t *hot = g->hot.as<uint64_t>(), *cold = g->cold.as<uint64_t>(), *ctl = g->ctl.as<uint64_t>();
    const size_t lds = (size_t)(S + 2) * (g->L.RH + g->L.RC) * 8;
    const unsigned nb = (unsigned)((grid + GB_FOLD_PER - 1) / GB_FOLD_PER);
    // a dense table of many records: one block per record; of few records
    // (too few blocks to fill the device): one block per GB_FOLD_PER partials
    if (dense && S + 2 >= 256) {
        k_gb_part_fold_dense<<<(unsigned)(S + 2), GB_THREADS, 0, st>>>(g->L, part, grid, S, hot, cold, g->cap, ctl);
    } else if (dense) {
        k_gb_part_fold<true><<<nb, GB_THREADS, lds, st>>>(g->L, part, grid, S, GB_FOLD_PER, hot, cold, g->cap, ctl, kmin);
    } else {
        k_gb_part_fold<false><<<nb, GB_THREADS, lds, st>>>(g->L, part, grid, S, GB_FOLD_PER, hot, cold, g->cap, ctl, 0);
    }
    VG_LAUNCH_CHECK();
    return 0;
}

static int gb_feed_narrow(vgpu_groupby *g, const FeedArgs &F, const vgpu_col *inputs, const PredInfo &pi,
                          const vgpu_col *pred_cols, int64_t groups, bool sketched) {
    const GbLayout &L = g->L;
    NarrowArgs A;
    memset(&A, 0, sizeof A);
    A.n_logical = F.n_logical;
    A.pos_base = F.pos_base;
    A.sel = F.sel;
    A.key = F.key;
    for (int i = 0; i < L.n_inputs; i++) {
        A.in[i] = inputs[i];
        A.w_sum[i] = L.w_sum[i];
        A.c_na[i] = L.c_na[i];
        A.in_f64[i] = L.in_type[i] == VGPU_DOUBLE;
    }
    A.pi = pi;
    if (pi.present) A.pcol = pred_cols[pi.col];
    A.hot = F.hot;
    A.cold = F.cold;
    A.ctl = F.ctl;
    A.cap = F.cap;
    A.RH = L.RH;
    A.RC = L.RC;
    A.lds_cold = F.lds_cold;
    A.w_fpos = L.w_fpos;
    A.key_double = L.key_is_double;
    A.key_none = L.key_none;
    const size_t rec = (size_t)(L.RH + (F.lds_cold ? L.RC : 0)) * 8;
    const size_t budget = (size_t)vg_tune_for(VG_TUNE_GB_LDS_BUDGET, std::max<int64_t>(groups, 1));
    // dense when this feed's sketch saw a short range of int64 / bool keys
    bool dense = false;
    if (sketched && g->sk_ranged && !L.key_is_double && !L.key_none) {
        const uint64_t range = (uint64_t)g->sk_hi - (uint64_t)g->sk_lo + 1;
        const int64_t hashed = std::max<int64_t>(gb_groups(g) - 2, 1);
        if (range != 0 && range <= 4096 && (range + 2) * rec <= budget && (int64_t)range <= 4 * hashed + 64) {
            dense = true;
            A.S = (int32_t)range;
            A.kmin = g->sk_lo;
        }
    }
    if (!dense) A.S = gb_lds_slots(g, 0, F.lds_cold != 0);
    const size_t lds = A.S ? (size_t)(A.S + 2) * rec : 0;
    const unsigned grid = gb_grid(F.n_logical, A.S, lds, gb_groups(g));
    hipStream_t st = vg_stream();
    if (gb_part_setup(g, grid, A.S, &A.part)) return -1;
    vg_kt_begin();
    // NI is the exact input count: an instantiation reads inputs 0..NI-1
    if (L.n_inputs == 0) {
        if (dense) k_gb_feed_narrow<0, true><<<grid, GB_THREADS, lds, st>>>(L, A);
        else k_gb_feed_narrow<0, false><<<grid, GB_THREADS, lds, st>>>(L, A);
    } else if (L.n_inputs == 1) {
        if (dense) k_gb_feed_narrow<1, true><<<grid, GB_THREADS, lds, st>>>(L, A);
        else k_gb_feed_narrow<1, false><<<grid, GB_THREADS, lds, st>>>(L, A);
    } else {
        if (dense) k_gb_feed_narrow<2, true><<<grid, GB_THREADS, lds, st>>>(L, A);
        else k_gb_feed_narrow<2, false><<<grid, GB_THREADS, lds, st>>>(L, A);
    }
    VG_LAUNCH_CHECK();
    if (A.part && gb_part_fold(g, A.part, grid, A.S, dense, A.kmin)) return -1;
    vg_kt_end();
    g->any_fed = 1;
    return 0;
}

// ---------------------------------------------------------------------------
// multi-column keys: one 64-bit code per row of the tuple
// ---------------------------------------------------------------------------

#define VG_MAX_KEY_COLS 8

struct KeyCodeArgs {
    vgpu_col k[VG_MAX_KEY_COLS];
    int32_t n_keys;
    int64_t n_rows, pos_base;
    int64_t *codes;
    uint64_t *valid;   // optional: bit r = no component of row r is NA or NaN
};

// Each component's class and normalized bits (NA, -0 as +0) are mixed in
// order; a tuple with a NaN component is a group of its own (NaN equals
// nothing), so its code is drawn from its row position instead.
// Rows are taken 64 to a wave in order (block size and grid stride are
// multiples of 64), so a wave's ballot is one whole word of the bitmap.
__global__ void __launch_bounds__(256) k_key_codes(KeyCodeArgs A) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < A.n_rows; r += stride) {
        uint64_t h = 0x6a09e667f3bcc909ull;
        bool nan = false, na = false;
        for (int j = 0; j < A.n_keys; j++) {
            uint64_t kn;
            const int cls = vg_key_class(A.k[j], r, kn);
            nan |= cls == VG_KEY_NAN;
            na |= cls == VG_KEY_NA;
            const uint64_t comp = cls == VG_KEY_NA ? 0x243f6a8885a308d3ull : kn;
            h = vg_mix64(h ^ vg_mix64(comp + (uint64_t)(cls == VG_KEY_NA) * 0x13198a2e03707344ull + (uint64_t)j));
        }
        if (nan) h = vg_mix64(0xa4093822299f31d0ull ^ (uint64_t)(A.pos_base + r)) ^ 0x082efa98ec4e6c89ull;
        A.codes[r] = (int64_t)h;
        if (A.valid) {
            const uint64_t m = vg_ballot(!nan && !na);
            if (vg_lane() == 0) A.valid[r >> 6] = m;
        }
    }
}

extern "C" int vgpu_key_codes(const vgpu_col *keys, int n_keys, int64_t n_rows, int64_t pos_base, int64_t *codes,
                              uint8_t *valid) {
    if (vg_ensure_init()) return -1;
    if (n_keys < 1 || n_keys > VG_MAX_KEY_COLS) {
        vg_set_error("key codes: %d key columns (1..%d)", n_keys, VG_MAX_KEY_COLS);
        return -1;
    }
    if (n_rows <= 0) return 0;
    KeyCodeArgs A;
    memset(&A, 0, sizeof A);
    for (int j = 0; j < n_keys; j++) {
        if (keys[j].type != VGPU_INT64 && keys[j].type != VGPU_DOUBLE && keys[j].type != VGPU_BOOL) {
            vg_set_error("key codes: key %d has unsupported type %d", j, keys[j].type);
            return -1;
        }
        A.k[j] = keys[j];
    }
    A.n_keys = n_keys;
    A.n_rows = n_rows;
    A.pos_base = pos_base;
    A.codes = codes;
    A.valid = reinterpret_cast<uint64_t *>(valid);
    if (valid && (reinterpret_cast<uintptr_t>(valid) & 7)) {
        vg_set_error("key codes: the validity bitmap must be 8-byte aligned");
        return -1;
    }
    k_key_codes<<<vg_blocks(n_rows, 256, 4096), 256, 0, vg_stream()>>>(A);
    VG_LAUNCH_CHECK();
    return 0;
}

// String keys: two independent 64-bit hashes of each row's bytes -- FNV-1a
// (Fowler, Noll, Vo; the reference hashes strings with it, src/hash.c:35-41)
// and a multiply-xorshift byte hash seeded with the length and finished with
// vg_mix64 -- and the row's position. One thread per row reads its bytes in
// order; neighbouring rows' bytes are neighbours, so a wave's loads share
// cache lines.
__global__ void __launch_bounds__(256) k_string_keys(const int64_t *off, const uint8_t *chars, const uint8_t *valid,
                                                     int64_t n, int64_t pos_base, int64_t *h1, int64_t *h2,
                                                     int64_t *pos) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += stride) {
        uint64_t a = 0, b = 0;
        if (!valid || ((valid[r >> 3] >> (r & 7)) & 1)) {
            const int64_t s = off[r], e = off[r + 1];
            a = 0xcbf29ce484222325ull;
            b = 0x9e3779b97f4a7c15ull ^ (uint64_t)(e - s);
            for (int64_t k = s; k < e; k++) {
                const uint64_t c = chars[k];
                a = (a ^ c) * 0x100000001b3ull;
                b = (b ^ c) * 0xd6e8feb86659fd93ull;
                b ^= b >> 32;
            }
            b = vg_mix64(b);
        }
        h1[r] = (int64_t)a;
        if (h2) h2[r] = (int64_t)b;
        if (pos) pos[r] = pos_base + r;
    }
}

extern "C" int vgpu_string_keys(const int64_t *offsets, const uint8_t *chars, const uint8_t *validity, int64_t n_rows,
                                int64_t pos_base, int64_t *h1, int64_t *h2, int64_t *pos) {
    if (vg_ensure_init()) return -1;
    if (n_rows < 0) {
        vg_set_error("string keys: negative row count");
        return -1;
    }
    if (n_rows == 0) return 0;
    if (!offsets || !h1) {
        vg_set_error("string keys: offsets and h1 are required");
        return -1;
    }
    k_string_keys<<<vg_blocks(n_rows, 256, 4096), 256, 0, vg_stream()>>>(offsets, chars, validity, n_rows, pos_base,
                                                                          h1, h2, pos);
    VG_LAUNCH_CHECK();
    return 0;
}

static int groupby_feed_body(vgpu_groupby *g, const vgpu_col *key, const vgpu_col *inputs, int64_t n_rows,
                             const int32_t *sel, int64_t n_sel, int64_t pos_base, const vgpu_prog *pred,
                             const vgpu_col *pred_cols, int n_pred_cols) {
    if (g->finished) {
        vg_set_error("groupby: feed after finish");
        return -1;
    }
    int64_t n_logical = sel ? n_sel : n_rows;
    if (n_logical <= 0) return 0;
    VgTuneScope tune_scope(g->L.RH + g->L.RC, n_logical);
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
    if (g->run && gbp_flush(g)) return -1;
    g->path = 0;
    // groups created by this feed are bounded by its rows, by the hint, or
    // (a large feed without a hint) by its sketched key count
    int64_t incoming = n_logical;
    int64_t groups = g->hint;
    if (g->hint > 0 && incoming > g->hint) incoming = g->hint;
    bool sketched = false;
    if (g->hint <= 0 && !g->L.key_none && n_logical >= GB_SK_MIN_ROWS) {
        int64_t est;
        if (gb_sketch(g, key, sel, n_logical, &est)) return -1;
        sketched = true;
        g->est_groups = std::max(g->est_groups, est);
        incoming = std::min(incoming, est + est / 4 + 4096);
        groups = est;
    }
    PredInfo pi = vg_pred_info(pred);
    size_t vm = (pi.present && !pi.fast) ? vg_vm_lds_bytes(pred, GB_THREADS) : 0;
    vm = (vm + 15) & ~(size_t)15;
    const bool lds_cold = gb_needs_cold(g->L, inputs);
    // partitioned: many groups, or groups that no LDS table holds (every
    // row would otherwise update a few hundred HBM records in contention)
    const bool partitioned = gbp_wanted(g, n_logical, groups) ||
                             (gbp_possible(g, n_logical, groups) && gb_lds_slots(g, vm, lds_cold) == 0);
    if (!partitioned && gb_ensure_capacity(g, incoming)) return -1;

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
    vgpu_prog prog;
    memset(&prog, 0, sizeof prog);
    if (pred) prog = *pred;
    ColSet pcs;
    memset(&pcs, 0, sizeof pcs);
    for (int i = 0; i < n_pred_cols && pred_cols; i++) pcs.c[i] = pred_cols[i];
    A.vm_lds = (int32_t)vm;
    if (partitioned) return gbp_feed(g, A, ins, pi, prog, pcs, groups, sketched);
    A.lds_cold = lds_cold;
    if (gb_narrow_shape(g->L, pi) && vg_tune(VG_TUNE_GB_NARROW) >= 2)
        return gb_feed_narrow(g, A, inputs, pi, pred_cols, groups, sketched);
    A.lds_slots = gb_lds_slots(g, vm, A.lds_cold);
    const size_t rec = (size_t)(g->L.RH + (A.lds_cold ? g->L.RC : 0)) * 8;
    // a dense table when this feed's sketch saw a short range of int64 /
    // bool keys (knob gb.feed.dense; not below the hashed table's slots)
    bool dense = false;
    if (sketched && g->sk_ranged && !g->L.key_is_double && !g->L.key_none && A.lds_slots &&
        vg_tune_for(VG_TUNE_GB_FEED_DENSE, groups) >= 2) {
        const uint64_t range = (uint64_t)g->sk_hi - (uint64_t)g->sk_lo + 1;
        if (range != 0 && range <= (uint64_t)A.lds_slots && vm + (range + 2) * rec <= (size_t)vg_tune(VG_TUNE_GB_LDS_BUDGET)) {
            dense = true;
            A.lds_slots = (int32_t)range;
            A.kmin = g->sk_lo;
        }
    }
    const size_t table = A.lds_slots ? (size_t)(A.lds_slots + 2) * rec : 0;
    // small tables are replicated while the copies stay within the budget
    A.lds_copies = 1;
    if (A.lds_slots && A.lds_slots <= 256) {
        const size_t budget = (size_t)vg_tune_for(VG_TUNE_GB_LDS_BUDGET, gb_groups(g));
        const int want = (int)vg_tune_for(VG_TUNE_GB_LDS_COPIES, gb_groups(g));
        while (A.lds_copies * 2 <= want && vm + table * A.lds_copies * 2 <= budget) A.lds_copies *= 2;
    }
    size_t lds = vm + table * A.lds_copies;
    const unsigned grid = gb_grid(n_logical, A.lds_slots, lds, gb_groups(g));
    A.part = nullptr;
    if (gb_part_setup(g, grid, A.lds_slots, &A.part)) return -1;
    vg_kt_begin();
    const bool pf = g->L.n_inputs <= 4 && vg_tune(VG_TUNE_GB_FEED_PF) >= 2;
    auto feed = pf ? (dense ? k_gb_feed<4, true> : k_gb_feed<4, false>)
                   : (dense ? k_gb_feed<0, true> : k_gb_feed<0, false>);
    feed<<<grid, GB_THREADS, lds, vg_stream()>>>(g->L, A, ins, pi, prog, pcs);
    VG_LAUNCH_CHECK();
    if (A.part && gb_part_fold(g, A.part, grid, A.lds_slots, dense, A.kmin)) return -1;
    vg_kt_end();
    g->any_fed = 1;
    return 0;
}

extern "C" int vgpu_groupby_feed(vgpu_groupby *g, const vgpu_col *key, const vgpu_col *inputs, int64_t n_rows,
                                 const int32_t *sel, int64_t n_sel, int64_t pos_base, const vgpu_prog *pred,
                                 const vgpu_col *pred_cols, int n_pred_cols) {
    return vg_guard<int>(-1, "vgpu_groupby_feed", [&] { return groupby_feed_body(g, key, inputs, n_rows, sel, n_sel,
        pos_base, pred, pred_cols, n_pred_cols); });
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

// Is this star feed the dense shape k_gb_star_dense handles?  (unique join
// keys, no filter or selection, one input that is a probe column or the
// product of a probe and a payload column, count / sum / mean only, and an
// int64 / bool payload group key whose value range fits an LDS table.)
static bool gb_star_dense_shape(const vgpu_groupby *g, const vgpu_feed_spec *sp, const StarArgs &st,
                                const PredInfo &pi, StarDenseArgs *sd) {
    const GbLayout &L = g->L;
    const vgpu_jtable *t = sp->join;
    if (vg_tune(VG_TUNE_GB_STAR_DENSE) < 2) return false;
    if (pi.present || st.sel || st.key_src != 1 || st.n_inputs != 1 || t->max_dup > 1) return false;
    if (st.ktype != VGPU_INT64 && st.ktype != VGPU_BOOL) return false;
    if (L.w_min[0] >= 0 || L.w_max[0] >= 0 || L.w_fvpos[0] >= 0) return false;
    if (st.pa[0] < 0) return false;   // a payload-only input takes the general kernel
    const int p = st.kslot;
    if (p < 0 || p >= t->n_payload || !t->pay_ranged[p]) return false;
    const uint64_t range = (uint64_t)t->pay_hi[p] - (uint64_t)t->pay_lo[p] + 1;
    const size_t rec = (size_t)(L.RH + L.RC) * 8;
    if (range == 0 || range > 4096 || (range + 2) * rec > 65536) return false;
    memset(sd, 0, sizeof *sd);
    sd->pkey = st.pkey;
    sd->pval = st.pval[st.pa[0]];
    JView jv = t->view();
    sd->ent = jv.ent;
    sd->jmask = (uint64_t)jv.cap - 1;
    if (vg_tune(VG_TUNE_JOIN_BLOOM) >= 2) {
        sd->bloom = jv.bloom;
        sd->bloom_mask = jv.bloom_mask;
    }
    sd->n_logical = st.n_logical;
    sd->pos_base = st.pos_base;
    sd->kmin = t->pay_lo[p];
    sd->hot = st.hot;
    sd->cold = st.cold;
    sd->ctl = st.ctl;
    sd->cap = st.cap;
    sd->S = (int32_t)range;
    sd->RH = L.RH;
    sd->RC = L.RC;
    sd->w_fpos = L.w_fpos;
    sd->w_sum = L.w_sum[0];
    sd->c_na = L.c_na[0];
    sd->kslot = p;
    sd->wslot = st.pb[0] >= 0 ? st.pb[0] : 0;
    sd->jkey_double = jv.key_is_double;
    sd->ktype = st.ktype;
    sd->ta = st.ta[0];
    sd->tb = st.tb[0];
    sd->sum_f64 = L.in_type[0] == VGPU_DOUBLE;
    return true;
}

static int groupby_feed_join_body(vgpu_groupby *g, const vgpu_feed_spec *sp) {
    if (g->finished) {
        vg_set_error("groupby: feed after finish");
        return -1;
    }
    if (g->run && gbp_flush(g)) return -1;
    g->path = 0;
    int64_t n_logical = sp->sel ? sp->n_sel : sp->n_rows;
    if (n_logical <= 0) return 0;
    VgTuneScope tune_scope(g->L.RH + g->L.RC, n_logical);
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
        // a pair's position keeps 12 bits for its match index (jf_pos): more
        // matches per probe row would tie positions, which decide first-seen
        // order and the NaN / zero-sign min / max rules
        if (sp->join->max_dup > JF_MAX_MATCHES) {
            vg_set_error("groupby_feed_join: a build key has %d rows; the fused join -> aggregate orders at most "
                         "%d matches per probe row", (int)sp->join->max_dup, JF_MAX_MATCHES);
            return -1;
        }
        if (n_logical > ((int64_t)1 << 31)) {
            vg_set_error("groupby_feed_join: %lld probe rows in one feed exceed 2^31", (long long)n_logical);
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
    // without a group count, sketch the column the groups come from once: a
    // build column (its distinct values bound the groups, plus the NA group
    // of unmatched rows) or a large probe column
    if (g->hint <= 0 && !g->L.key_none && g->est_groups == 0) {
        int64_t est = 0;
        if (sp->key_side == 1 && sp->join && sp->join->n_build > 0) {
            if (gb_sketch(g, &sp->build_cols[sp->key_col], nullptr, sp->join->n_build, &est)) return -1;
        } else if (sp->key_side == 0 && n_logical >= GB_SK_MIN_ROWS) {
            if (gb_sketch(g, &sp->probe_cols[sp->key_col], sp->sel, n_logical, &est)) return -1;
        }
        if (est > 0) {
            g->est_groups = est + 1;
            incoming = std::min(incoming, est + est / 4 + 4096);
        }
    }
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
    const bool star = gb_star_shape(g, sp, A, pi, &st);
    StarDenseArgs sd;
    if (star && gb_star_dense_shape(g, sp, st, pi, &sd)) {
        const int64_t groups = gb_groups(g);
        const size_t dl = (size_t)(sd.S + 2) * (sd.RH + sd.RC) * 8;
        int items = (int)vg_tune_for(VG_TUNE_GB_STAR_ITEMS, groups);
        items = items <= 2 ? 2 : (items >= 8 ? 8 : 4);
        // blocks per CU (knob gb.star.blocks_per_cu, default 4): more blocks
        // add table inits and merges (measured on 50M probe rows with the
        // general feed: 1,024 blocks 0.97 ms, 2,048 1.22 ms)
        const unsigned grid = vg_blocks(n_logical, vg_tune_for(VG_TUNE_GB_LDS_ROWS, groups),
                                        std::min<int64_t>((int64_t)vg_cus() * vg_tune_for(VG_TUNE_GB_STAR_BPC, groups),
                                                          vg_resident_blocks(dl)));
        const int mode = st.pa[0] >= 0 && st.pb[0] >= 0 ? 2 : 0;
        const bool left = sp->join_kind == VGPU_JOIN_LEFT;
        // both modes read the probe operand: plain when it and the probe key
        // are 8-byte columns without validity
        const bool plain = !sd.pkey.validity && sd.pkey.type != VGPU_BOOL && sd.pval.data && !sd.pval.validity &&
                           sd.pval.type != VGPU_BOOL;
        hipStream_t ks = vg_stream();
        // block partials reduced per record (knob gb.merge.partials)
        if (gb_part_setup(g, grid, sd.S, &sd.part)) return -1;
        vg_kt_begin();
#define STAR_DENSE(I, M, LF)                                                                \
    do {                                                                                    \
        if (plain) k_gb_star_dense<I, M, LF, true><<<grid, GB_THREADS, dl, ks>>>(g->L, sd);  \
        else k_gb_star_dense<I, M, LF, false><<<grid, GB_THREADS, dl, ks>>>(g->L, sd);       \
    } while (0)
        if (mode == 2) {
            if (left) { if (items == 2) STAR_DENSE(2, 2, true); else if (items == 4) STAR_DENSE(4, 2, true); else STAR_DENSE(8, 2, true); }
            else { if (items == 2) STAR_DENSE(2, 2, false); else if (items == 4) STAR_DENSE(4, 2, false); else STAR_DENSE(8, 2, false); }
        } else {
            if (left) { if (items == 2) STAR_DENSE(2, 0, true); else if (items == 4) STAR_DENSE(4, 0, true); else STAR_DENSE(8, 0, true); }
            else { if (items == 2) STAR_DENSE(2, 0, false); else if (items == 4) STAR_DENSE(4, 0, false); else STAR_DENSE(8, 0, false); }
        }
#undef STAR_DENSE
        VG_LAUNCH_CHECK();
        if (sd.part && gb_part_fold(g, sd.part, grid, sd.S, true, sd.kmin)) return -1;
        vg_kt_end();
        g->any_fed = 1;
        return 0;
    }
    if (star) {
        st.lds_slots = A.lds_slots;
        st.vm_lds = 0;
        size_t slds = A.lds_slots ? (size_t)(A.lds_slots + 2) * (g->L.RH + g->L.RC) * 8 : 0;
        const int64_t groups = gb_groups(g);
        int items = (int)vg_tune_for(VG_TUNE_GB_STAR_ITEMS, groups);
        unsigned grid = vg_blocks(n_logical, vg_tune_for(A.lds_slots ? VG_TUNE_GB_LDS_ROWS : VG_TUNE_GB_GLB_ROWS, groups),
                                  vg_tune_for(A.lds_slots ? VG_TUNE_GB_LDS_MAX_BLOCKS : VG_TUNE_GB_GLB_MAX_BLOCKS,
                                              groups));
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
    k_gb_feed_join<<<gb_grid(n_logical, A.lds_slots, lds, gb_groups(g)), GB_THREADS, lds, vg_stream()>>>(g->L, A, cp, jv, pi, prog);
    VG_LAUNCH_CHECK();
    vg_kt_end();
    g->any_fed = 1;
    return 0;
}

extern "C" int vgpu_groupby_feed_join(vgpu_groupby *g, const vgpu_feed_spec *sp) {
    return vg_guard<int>(-1, "vgpu_groupby_feed_join", [&] { return groupby_feed_join_body(g, sp); });
}

static RecSrc gb_table_src(const vgpu_groupby *g) {
    RecSrc R;
    R.hot = g->hot.as<uint64_t>();
    R.cold = g->cold.as<uint64_t>();
    R.hstride = g->L.RH;
    R.cstride = g->L.RC;
    R.na_slot = g->cap;
    R.sp_slot = g->cap + 1;
    return R;
}

// (sort key, slot) pairs of every sorted record of the table into
// sort_keys / sort_vals; count -> ctl[5], OR / AND of the keys -> ctl[6] /
// ctl[7], NA rows -> ctl[8]
static int gb_compact(vgpu_groupby *g, const RecSrc &R, int64_t n_slots) {
    hipStream_t s = vg_stream();
    int64_t nb = (n_slots + GB_CMP_SLOTS - 1) / GB_CMP_SLOTS;
    if (g->cmp.reserve((size_t)nb * 32)) return -1;
    uint64_t *boff = g->cmp.as<uint64_t>();
    uint64_t *por = boff + nb, *pand = por + nb;
    uint32_t *bcount = reinterpret_cast<uint32_t *>(pand + nb);
    uint64_t *ctl = g->ctl.as<uint64_t>();
    k_gb_cmp_count<<<(unsigned)nb, GB_THREADS, 0, s>>>(g->L, R, n_slots, g->order, bcount);
    VG_LAUNCH_CHECK();
    if (vg_scan_u32_to_u64(bcount, boff, nb, ctl + 5, g->sort_tmp)) return -1;
    k_gb_cmp_write<<<(unsigned)nb, GB_THREADS, 0, s>>>(g->L, R, n_slots, g->order, boff, g->sort_keys.as<uint64_t>(),
                                                       g->sort_vals.as<uint32_t>(), por, pand, ctl);
    VG_LAUNCH_CHECK();
    k_gb_or_and_final<<<vg_blocks(nb, 4 * GB_THREADS, 256), GB_THREADS, 0, s>>>(por, pand, nb, ctl);
    VG_LAUNCH_CHECK();
    return 0;
}

static int highest_bit(uint64_t x) { return x ? 63 - __builtin_clzll(x) : -1; }
static int lowest_bit(uint64_t x) { return x ? __builtin_ctzll(x) : 64; }

static int64_t groupby_finish_body(vgpu_groupby *g) {
    if (g->finished) return g->n_groups;
    hipStream_t s = vg_stream();
    if (g->cap == 0 && gb_alloc_table(g, 64)) return -1;
    // the one synchronization of finish (in the common case): group count,
    // sort-key bit range, overflow flag and the NA group's size
    uint64_t ctl[GB_CTL_WORDS];
    RecSrc R;
    bool from_table = true;
    if (g->run) {
        // records of a partitioned feed sort directly, unless some records
        // went to the hash table; then the run is folded into it first
        R = gbp_run_src(g);
        k_gb_ctl_init<<<1, 64, 0, s>>>(g->ctl.as<uint64_t>(), 1);
        VG_LAUNCH_CHECK();
        if (gbp_collect(g, R) || gb_read_ctl(g, ctl)) return -1;
        from_table = ctl[0] != 0 || ctl[9] != 0;
        if (from_table && gbp_flush(g)) return -1;
    }
    if (from_table) {
        int64_t slots = g->cap + 2;
        if (g->sort_keys.reserve((size_t)slots * 8) || g->sort_vals.reserve((size_t)slots * 4)) return -1;
        R = gb_table_src(g);
        k_gb_ctl_init<<<1, 64, 0, s>>>(g->ctl.as<uint64_t>(), 1);
        VG_LAUNCH_CHECK();
        if (gb_compact(g, R, slots) || gb_read_ctl(g, ctl)) return -1;
    }
    if (ctl[1]) {
        vg_set_error("groupby: hash table overflow (more groups than the table was sized for)");
        return -1;
    }
    int64_t n_sorted = (int64_t)ctl[5];
    // NA group of the key-sorted order: appended after the sort
    int na_last = g->order == VGPU_ORDER_KEY && ctl[8] != 0;
    uint64_t diff = ctl[6] ^ ctl[7];
    // a run of range partitions collected its records in key order from
    // entry 1, entry 0 holding an EMPTY-valued key's record if there is one
    const bool in_order = !from_table && g->gbp.range;
    const uint32_t *order_vals = g->sort_vals.as<uint32_t>() + (in_order && !(ctl[11] & 1) ? 1 : 0);
    if (n_sorted > 1 && diff && !in_order) {
        int lo = lowest_bit(diff), hi = highest_bit(diff) + 1;
        if (vg_sort_pairs(g->sort_keys.as<uint64_t>(), g->sort_vals.as<uint32_t>(), n_sorted, lo, hi, g->sort_tmp))
            return -1;
    }
    int64_t n_out = n_sorted + na_last;
    size_t vwords = (size_t)((n_out + 63) / 64 + 1);
    const size_t col_bytes = (size_t)(n_out + 1) * 8, bits_bytes = vwords * 8;
    size_t off = col_bytes;
    g->off_kvalid = off;
    off += bits_bytes;
    for (int a = 0; a < g->aggs.n; a++) {
        g->off_vals[a] = off;
        off += col_bytes;
        g->off_valid[a] = off;
        off += bits_bytes;
    }
    g->out_bytes = off;
    if (g->out.reserve(off)) return -1;
    uint8_t *ob = g->out.as<uint8_t>();
    EmitOut O;
    memset(&O, 0, sizeof O);
    O.keys = ob;
    O.kvalid = reinterpret_cast<uint64_t *>(ob + g->off_kvalid);
    for (int a = 0; a < g->aggs.n; a++) {
        O.vals[a] = reinterpret_cast<double *>(ob + g->off_vals[a]);
        O.valid[a] = reinterpret_cast<uint64_t *>(ob + g->off_valid[a]);
    }
    if (n_out > 0) {
        const int RW = g->L.RH + g->L.RC;
        if (RW <= GB_EMIT_STAGE_WORDS)
            k_gb_emit<true><<<vg_blocks(n_out, GB_THREADS), GB_THREADS, (size_t)GB_THREADS * RW * 8, s>>>(
                g->L, g->aggs, R, order_vals, n_sorted, na_last, n_out, 8, O, g->ctl.as<uint64_t>());
        else
            k_gb_emit<false><<<vg_blocks(n_out, GB_THREADS), GB_THREADS, 0, s>>>(
                g->L, g->aggs, R, order_vals, n_sorted, na_last, n_out, 8, O, g->ctl.as<uint64_t>());
        VG_LAUNCH_CHECK();
    }
    g->n_groups = n_out;
    g->finished = 1;
    g->run = 0;
    return n_out;
}

extern "C" int64_t vgpu_groupby_finish(vgpu_groupby *g) {
    return vg_guard<int64_t>(-1, "vgpu_groupby_finish", [&] { return groupby_finish_body(g); });
}

static int64_t groupby_fetch_body(vgpu_groupby *g, void *keys, uint8_t *key_validity, double *const *agg_values,
                                  uint8_t *const *agg_validity) {
    if (!g->finished) {
        vg_set_error("groupby: fetch before finish");
        return -1;
    }
    hipStream_t s = vg_stream();
    int64_t n = g->n_groups;
    size_t vb = (size_t)(n + 7) / 8;
    const uint8_t *ob = g->out.as<uint8_t>();
    // bool keys are emitted as 8-byte 0/1 words; callers narrow them
    uint8_t *dst_keys = keys && !g->L.key_none ? static_cast<uint8_t *>(keys) : nullptr;
    // a small result: the whole buffer and the control words in two copies
    // to pinned staging, one synchronization, then host copies (each
    // separate copy to pageable memory costs tens of microseconds)
    const size_t staged = g->out_bytes + sizeof(uint64_t) * GB_CTL_WORDS;
    uint8_t *pin = staged <= VG_STAGE_MAX ? static_cast<uint8_t *>(vg_pinned_stage(staged)) : nullptr;
    if (pin) {
        if (n > 0) VG_TRY(hipMemcpyAsync(pin, ob, g->out_bytes, hipMemcpyDeviceToHost, s));
        VG_TRY(hipMemcpyAsync(pin + g->out_bytes, g->ctl.ptr, sizeof(uint64_t) * GB_CTL_WORDS, hipMemcpyDeviceToHost,
                              s));
        VG_TRY(hipStreamSynchronize(s));
        if (n > 0) {
            if (dst_keys) memcpy(dst_keys, pin, (size_t)n * 8);
            if (key_validity) memcpy(key_validity, pin + g->off_kvalid, vb);
            for (int a = 0; a < g->aggs.n; a++) {
                if (agg_values && agg_values[a]) memcpy(agg_values[a], pin + g->off_vals[a], (size_t)n * 8);
                if (agg_validity && agg_validity[a]) memcpy(agg_validity[a], pin + g->off_valid[a], vb);
            }
        }
        uint64_t w4;
        memcpy(&w4, pin + g->out_bytes + 4 * sizeof(uint64_t), sizeof w4);
        g->warn_count = (int64_t)w4;
        return g->warn_count;
    }
    uint64_t ctl[GB_CTL_WORDS];
    if (n > 0) {
        if (dst_keys) VG_TRY(hipMemcpyAsync(dst_keys, ob, (size_t)n * 8, hipMemcpyDeviceToHost, s));
        if (key_validity) VG_TRY(hipMemcpyAsync(key_validity, ob + g->off_kvalid, vb, hipMemcpyDeviceToHost, s));
        for (int a = 0; a < g->aggs.n; a++) {
            if (agg_values && agg_values[a])
                VG_TRY(hipMemcpyAsync(agg_values[a], ob + g->off_vals[a], (size_t)n * 8, hipMemcpyDeviceToHost, s));
            if (agg_validity && agg_validity[a])
                VG_TRY(hipMemcpyAsync(agg_validity[a], ob + g->off_valid[a], vb, hipMemcpyDeviceToHost, s));
        }
    }
    VG_TRY(hipMemcpyAsync(ctl, g->ctl.ptr, sizeof ctl, hipMemcpyDeviceToHost, s));
    VG_TRY(hipStreamSynchronize(s));
    g->warn_count = (int64_t)ctl[4];
    return g->warn_count;
}

extern "C" int64_t vgpu_groupby_fetch(vgpu_groupby *g, void *keys, uint8_t *key_validity, double *const *agg_values,
                                      uint8_t *const *agg_validity) {
    return vg_guard<int64_t>(-1, "vgpu_groupby_fetch", [&] { return groupby_fetch_body(g, keys, key_validity,
        agg_values, agg_validity); });
}

extern "C" int vgpu_groupby_result_device(vgpu_groupby *g, const void **keys, const uint8_t **key_validity, int agg,
                                          const double **values, const uint8_t **validity) {
    if (!g->finished) {
        vg_set_error("groupby: result before finish");
        return -1;
    }
    const uint8_t *ob = g->out.as<uint8_t>();
    if (keys) *keys = ob;
    if (key_validity) *key_validity = ob + g->off_kvalid;
    if (agg >= 0 && agg < g->aggs.n) {
        if (values) *values = reinterpret_cast<const double *>(ob + g->off_vals[agg]);
        if (validity) *validity = ob + g->off_valid[agg];
    }
    return 0;
}

extern "C" int vgpu_groupby_record_words(const vgpu_groupby *g) { return g ? g->L.RH + g->L.RC : -1; }

extern "C" int vgpu_groupby_last_path(const vgpu_groupby *g) { return g ? g->path : -1; }

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
    g->cmp.release();
    g->gbp_s1.release();
    g->gbp_s2.release();
    g->gbp_aux.release();
    g->run_rec.release();
    g->part.release();
    g->out.release();
    delete g;
}
