// vgpu_batch.hip -- many independent filter -> group-by queries of one
// shape (key type, inputs, aggregates, order, predicate program with one
// per-query constant) executed as one batch.
//
// Each query q is what vectra runs for
//     tbl_q |> filter(<pred with constant param_q>) |> group_by(key) |>
//     summarise(<aggs>)
// (src/filter.c:13-84, src/group_agg.c:155-672, src/agg_ops.c:25-269), with
// the same results as one vgpu_groupby per query.
//
// Design.  Planning sizes every query's hash table from a cardinality
// sketch taken in one launch over all queries (one synchronization) and
// lays the tables out back to back in one arena of group records (the
// record model of vgpu_groupby_core.h).  A run is then pure device work,
// capturable as a hipGraph:
//   1. init      every table, every query's control words, the work queue;
//   2. feed      ONE persistent launch: blocks pull (query, row tile) items
//                from an atomic work queue, filter the tile, and aggregate
//                it into an LDS-private table (queries with few groups) or
//                straight into the query's HBM table;
//   3. finish    compaction of every table into (sort key, slot) pairs,
//                a stable radix sort by the order key and then by query, so
//                that every query's groups are contiguous and in its order,
//                and one emit launch for all groups of all queries.
// The sort's bit range comes from the first run (inside plan); later runs
// check on the device that their keys stay in that range and flag the
// batch otherwise (vgpu_batch_counts reports it).
#include "vgpu_groupby_core.h"
#include "vgpu_join_table.h"

#include <algorithm>
#include <cmath>
#include <vector>

#define BT_TILE_MIN 32768          // rows per work item: at least ...
#define BT_TILE_MAX (1 << 21)      // ... and at most; chosen per batch from its total rows
#define BT_SK_BITS 10              // sketch registers per query: 2^10
#define BT_SK_REGS (1 << BT_SK_BITS)
#define BT_LDS_BUDGET 61440        // dynamic LDS of the 256-thread feed blocks
#define BT_LDS_BIG (160 * 1024 - 2048)   // ... of the 1,024-thread blocks of the large-table launch
#define BT_BIG_THREADS 1024
#define BT_BLOCKS_PER_CU 8
#define BT_U 4                     // rows in flight per thread in the feed
#define BT_STILE 16384             // rows per tile of a shared scan
#define BT_CTL_ERR 9               // batch control word: inputs left the plan's key range
#define BT_SPLIT_SLOTS 256         // queries with larger LDS tables run in a later launch of the feed
#define BT_CLASSES 3               // feed launches: small / mid LDS tables (256 threads), large (1,024)
#define BT_MID_BYTES (160 * 1024 / BT_BLOCKS_PER_CU - 512)   // LDS of a mid table: 8 blocks per CU still fit

struct BtQuery {
    vgpu_col key;
    vgpu_col probe_key;  // join batches: the probe side's join key
    InSet ins;
    ColSet pcs;
    PredInfo pi;
    int64_t n_rows;
    int64_t base;       // first arena slot of the query's table
    int64_t cap;        // hashed slots (pow2); NA / EMPTY-valued key records at cap, cap + 1
    int32_t lds_slots;  // 0 = aggregate straight into the HBM table
    int32_t vm;         // predicate evaluated by the VM (program in progs[q])
    int64_t kmin;       // dense tables: the key of record 0
    int32_t dense;      // key k owns record k - kmin of cap records (HBM and LDS tables)
    int32_t pad_;
};

struct BtItem {
    uint32_t q;
    uint32_t part;      // row tile (feed) or slot chunk (compaction)
};

// inner join of every query's probe key against one resident table; the
// group key is an embedded payload column of the table's entries
struct BtJoin {
    JView jv;
    int32_t on;
    int32_t key_payload;   // payload column that is the group key
    int32_t key_type;      // its type
    int32_t pad;
};

struct BtArgs {
    BtJoin jn;
    const BtQuery *qs;
    const vgpu_prog *progs;
    const BtItem *items;
    int64_t n_items;
    unsigned long long *queue;
    uint64_t *hot, *cold;
    uint64_t *ctl;      // GB_CTL_WORDS per query, then the batch's own GB_CTL_WORDS
    int32_t vm_lds;
    int32_t n_queries;
    // shared scans: every query reads the same rows (see k_bt_feed_shared)
    int32_t K, J;
    int64_t n_tiles;
    int32_t lds_cold;   // LDS tables carry cold words (some input has NAs or a double min / max)
    int32_t n_active;   // queries a run feeds (the first n_active)
    int64_t tile;       // rows per work item
};

struct vgpu_batch {
    GbLayout L;
    GbAggs aggs;
    int32_t order;
    int32_t key_type;
    vgpu_prog prog;
    int32_t has_pred = 0;
    int32_t n_pred_cols = 0;
    int32_t param_insn = -1;
    // plan
    int32_t n_queries = 0;
    int64_t n_slots = 0;            // arena slots
    int64_t n_items = 0, n_cblocks = 0;
    int32_t S_max = 0;
    int32_t vm_lds = 0;
    int32_t lds_cold = 1;
    // the feed's items in two classes, one launch each: queries with small
    // or no LDS tables (many blocks per CU) first, then queries with large
    // LDS tables (their LDS size limits the blocks per CU)
    // the feed's launches, one per class of LDS table size: items
    // [cls_first[c], cls_first[c + 1]) of the item list, threads per block,
    // rows per item, largest table, grid
    int64_t cls_first[BT_CLASSES + 1] = {};
    int32_t cls_nt[BT_CLASSES] = {GB_THREADS, GB_THREADS, GB_THREADS};
    int64_t cls_tile[BT_CLASSES] = {};
    int32_t cls_S[BT_CLASSES] = {};
    int64_t cls_grid[BT_CLASSES] = {};
    int64_t shared_resident = 1;    // feed blocks resident at once (LDS and wave limits)
    int64_t tile = BT_TILE_MIN;     // rows per work item
    int32_t planned = 0;
    int64_t n_pairs = 0;            // sorted records of the plan's run
    int32_t lo = 0, hi = 0, qbits = 0;
    uint64_t range_mask = 0;
    DevBuf qs, progs, items, cmap, queue, hot, cold, ctl, qbase, qcount, qoff;
    DevBuf bcount, boff, por, pand, skeys, svals, sort_tmp, scan_tmp;
    DevBuf out_keys, out_kvalid;
    std::vector<DevBuf> out_vals, out_valid;
    std::vector<int64_t> h_count, h_off;
    int32_t counts_valid = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    // join
    BtJoin join;
    int32_t join_key_type = 0;
    vgpu_col join_payload;
    int64_t join_rows = 0;
    vgpu_jtable *jtab = nullptr;      // its direct-addressed map is built on the first bucketed plan
    JDense jd = {nullptr, 0, 0, 0};   // the map a bucketed pass probes instead of the hash table
    // shared scans
    int32_t shared = 0;
    int32_t n_active = 0;            // queries the shared feed runs (rerun may lower it)
    int64_t n_tiles = 0;
    std::vector<BtQuery> h_qs;
    std::vector<vgpu_prog> h_progs;
    // bucketed passes (k_bt_chain_feed)
    int32_t chain = 0;
    int64_t ch_kmin = 0, ch_rpb = 0;
    int32_t ch_E = 0, ch_m = 0, ch_nsum = 0, ch_G = 0, ch_f64 = 0, ch_key_type = 0;
    int32_t ch_sum_in[4] = {0, 0, 0, 0}, ch_sum_pred[4] = {0, 0, 0, 0};
    size_t ch_lds = 0;
    DevBuf ch_aux, ch_slab, ch_tab;
    std::vector<uint8_t> h_chain;   // constants in chain order, then chain positions (uploaded)
};

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------

// per-query cardinality sketch (HyperLogLog registers, NaN-key rows) and
// range of the regular keys as int64 (mm[2q] = max of ~ordered key, i.e.
// the complement of the smallest; mm[2q + 1] = the largest ordered key;
// both start at 0): the feed's items, one block per item at a time
__global__ void __launch_bounds__(GB_THREADS)
k_bt_sketch(BtArgs B, int key_none, uint32_t *regs, unsigned long long *nan_rows, unsigned long long *mm) {
    __shared__ uint32_t r[BT_SK_REGS];
    __shared__ unsigned long long nn, s_lo, s_hi;
    for (int64_t it = blockIdx.x; it < B.n_items; it += gridDim.x) {
        for (int i = threadIdx.x; i < BT_SK_REGS; i += GB_THREADS) r[i] = 0;
        if (threadIdx.x == 0) {
            nn = 0;
            s_lo = 0;
            s_hi = 0;
        }
        __syncthreads();
        unsigned long long lo = 0, hi = 0;
        const BtItem item = B.items[it];
        const BtQuery &Q = B.qs[item.q];
        const int64_t r0 = (int64_t)item.part * B.tile;
        const int64_t r1 = r0 + B.tile < Q.n_rows ? r0 + B.tile : Q.n_rows;
        unsigned long long my_nan = 0;
        if (!key_none)
            for (int64_t row = r0 + threadIdx.x; row < r1; row += GB_THREADS) {
                uint64_t kn;
                int cls = vg_key_class(Q.key, row, kn);
                if (cls == VG_KEY_NAN) { my_nan++; continue; }
                if (cls != VG_KEY_REGULAR) continue;
                const uint64_t o = vg_i64_ordered(kn);
                lo = ~o > lo ? ~o : lo;
                hi = o > hi ? o : hi;
                uint64_t h = vg_mix64(kn);
                uint32_t idx = (uint32_t)(h >> (64 - BT_SK_BITS));
                uint64_t rest = (h << BT_SK_BITS) | (1ull << (BT_SK_BITS - 1));
                uint32_t rank = (uint32_t)__builtin_clzll(rest) + 1;
                if (r[idx] < rank) atomicMax(&r[idx], rank);
            }
        if (my_nan) atomicAdd(&nn, my_nan);
        if (lo) atomicMax(&s_lo, lo);
        if (hi) atomicMax(&s_hi, hi);
        __syncthreads();
        uint32_t *gr = regs + (size_t)item.q * BT_SK_REGS;
        for (int i = threadIdx.x; i < BT_SK_REGS; i += GB_THREADS)
            if (r[i]) atomicMax(&gr[i], r[i]);
        if (threadIdx.x == 0) {
            if (nn) atomicAdd(&nan_rows[item.q], nn);
            if (s_lo) atomicMax(&mm[2 * (size_t)item.q], s_lo);
            if (s_hi) atomicMax(&mm[2 * (size_t)item.q + 1], s_hi);
        }
        __syncthreads();
    }
}

__global__ void k_bt_clear_u32(uint32_t *p, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// control words of every query and of the batch, per-query group counts,
// the work queue
__global__ void k_bt_init_ctl(uint64_t *ctl, int n_queries, uint64_t *qcount, unsigned long long *queue) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t words = (int64_t)(n_queries + 1) * GB_CTL_WORDS;
    if (i < words) {
        int w = (int)(i % GB_CTL_WORDS);
        ctl[i] = (w == 2 || w == 3 || w == 7) ? ~0ull : 0ull;
    }
    if (i < n_queries) qcount[i] = 0;
    if (i == 0) {
        for (int c = 0; c < BT_CLASSES; c++) queue[c] = 0;
    }
}

// Fields of one query a feed block works on: column sets in LDS, scalars
// in registers (block-uniform).
struct BtLocal {
    vgpu_col key, probe_key;
    PredInfo pi;
    int64_t n_rows, cap, tbase, kmin;
    int32_t S, vm, dense;
};

template <int NT>
__device__ __forceinline__ BtLocal bt_load_query(const BtQuery &G, InSet &s_ins, ColSet &s_pcs) {
    for (int w = threadIdx.x; w < (int)(sizeof(InSet) / 4); w += NT)
        reinterpret_cast<uint32_t *>(&s_ins)[w] = reinterpret_cast<const uint32_t *>(&G.ins)[w];
    for (int w = threadIdx.x; w < (int)(sizeof(ColSet) / 4); w += NT)
        reinterpret_cast<uint32_t *>(&s_pcs)[w] = reinterpret_cast<const uint32_t *>(&G.pcs)[w];
    BtLocal Q;
    Q.key = G.key;
    Q.probe_key = G.probe_key;
    Q.pi = G.pi;
    Q.n_rows = G.n_rows;
    Q.cap = G.cap;
    Q.tbase = G.base;
    Q.S = G.lds_slots;
    Q.vm = G.vm;
    Q.kmin = G.kmin;
    Q.dense = G.dense;
    return Q;
}

// record of a key in a query's HBM table: hashed (gb_global_slot), or in a
// dense table the key's own record; a key outside the planned range flags
// the batch (its inputs changed since plan)
__device__ __forceinline__ int64_t bt_table_slot(const GbLayout &L, const BtArgs &B, int dense, int64_t kmin,
                                                 uint64_t *hot, int64_t cap, int cls, uint64_t kn, uint64_t pos,
                                                 uint64_t *ctl) {
    if (!dense || cls != VG_KEY_REGULAR) return gb_global_slot(L, hot, cap, cls, kn, pos, ctl);
    const uint64_t d = kn - (uint64_t)kmin;
    if (d < (uint64_t)cap) return (int64_t)d;
    atomicOr((unsigned long long *)&B.ctl[(size_t)B.n_queries * GB_CTL_WORDS + BT_CTL_ERR], 1ull);
    return -1;
}

// the block's LDS table of a query: dense when its HBM table is (keys
// preset), else hashed
template <int NT>
__device__ __forceinline__ void bt_lds_init(const GbLayout &L, const BtLocal &Q, uint64_t *lhot, uint64_t *lcold,
                                            uint64_t empty) {
    gb_lds_init<NT>(L, lhot, lcold, Q.S, empty, Q.dense != 0, Q.kmin);
}

// the LDS table into the query's HBM table (records no row touched skipped)
template <int NT>
__device__ __forceinline__ void bt_lds_merge(const GbLayout &L, const BtArgs &B, const BtLocal &Q,
                                             const uint64_t *lhot, const uint64_t *lcold, uint64_t empty,
                                             uint64_t *hot, uint64_t *cold, uint64_t *ctl) {
    __syncthreads();
    for (int s = threadIdx.x; s < Q.S + 2; s += NT) {
        const uint64_t *h = lhot + (size_t)s * L.RH;
        int64_t gs;
        if (s < Q.S) {
            if (h[0] == empty || h[1] == 0) continue;
            gs = bt_table_slot(L, B, Q.dense, Q.kmin, hot, Q.cap, VG_KEY_REGULAR, h[0], 0, ctl);
        } else {
            if (h[1] == 0) continue;
            gs = Q.cap + (s - Q.S);
        }
        if (gs >= 0)
            gb_merge(L, hot + (size_t)gs * L.RH, cold + (size_t)gs * L.RC, h, lcold ? lcold + (size_t)s * L.RC : nullptr);
    }
}

// one group update: LDS table when the key has a slot there, else the
// query's HBM table
__device__ __forceinline__ void bt_update(const GbLayout &L, const BtArgs &B, const BtLocal &Q, const InSet &ins,
                                          int64_t row, uint64_t pos, int cls, uint64_t kn, uint64_t *lhot,
                                          uint64_t *lcold, uint64_t *hot, uint64_t *cold, uint64_t *ctl,
                                          uint64_t empty) {
    const int S = Q.S;
    int64_t ls = -1;
    if (S) ls = Q.dense ? gb_lds_dense_slot(S, cls, kn, Q.kmin) : gb_lds_slot(lhot, S, L.RH, cls, kn, empty);
    if (ls >= 0) {
        gb_apply(L, ins, row, pos, lhot + (size_t)ls * L.RH, lcold ? lcold + (size_t)ls * L.RC : nullptr);
    } else {
        int64_t gs = bt_table_slot(L, B, Q.dense, Q.kmin, hot, Q.cap, cls, kn, pos, ctl);
        if (gs >= 0) gb_apply(L, ins, row, pos, hot + (size_t)gs * L.RH, cold + (size_t)gs * L.RC);
    }
}

// rows [r0, r1) of one query: BT_U rows per thread per step, their
// predicate and key (or join probe) loads issued together so each wave
// keeps several loads in flight
template <int NT>
__device__ __forceinline__ void bt_rows(const GbLayout &L, const BtArgs &B, const BtLocal &Q, const InSet &ins,
                                        const ColSet &pcs, const vgpu_prog &prog, uint64_t *regs, uint8_t *regv,
                                        int64_t r0, int64_t r1, uint64_t *lhot, uint64_t *lcold, uint64_t *hot,
                                        uint64_t *cold, uint64_t *ctl) {
    const uint64_t empty = L.key_is_double ? VG_EMPTY_F64 : VG_EMPTY_I64;
    for (int64_t base = r0 + threadIdx.x; base < r1; base += (int64_t)NT * BT_U) {
        bool ok[BT_U];
        int cls[BT_U];
        uint64_t kn[BT_U];
        const uint64_t *ent[BT_U];
#pragma unroll
        for (int u = 0; u < BT_U; u++) {
            int64_t row = base + (int64_t)u * NT;
            ok[u] = row < r1 && vg_pred_row(Q.pi, prog, pcs, row, regs, regv, NT, threadIdx.x);
        }
        if (B.jn.on) {
#pragma unroll
            for (int u = 0; u < BT_U; u++)
                ent[u] = ok[u] ? jt_find(B.jn.jv, Q.probe_key, base + (int64_t)u * NT) : nullptr;
#pragma unroll
            for (int u = 0; u < BT_U; u++) {
                if (!ent[u]) continue;
                int64_t row = base + (int64_t)u * NT;
                const uint64_t segw = ent[u][1];
                const uint32_t cnt = jt_count(B.jn.jv, segw), off = (uint32_t)(segw >> 32);
                const int p = B.jn.key_payload;
                for (uint32_t j = 0; j < cnt; j++) {
                    bool v;
                    uint64_t bits;
                    if (j == 0) {
                        v = (segw >> (24 + p)) & 1;
                        bits = ent[u][2 + p];
                    } else {
                        v = B.jn.jv.cval[p][off + j];
                        bits = B.jn.jv.cpay[p][off + j];
                    }
                    uint64_t k;
                    int c = vg_key_class_bits(B.jn.key_type, v, v ? bits : 0, k);
                    uint64_t pos = ((uint64_t)row << 12) | (j < 0xFFFu ? j : 0xFFFu);
                    if (L.key_is_double && c == VG_KEY_REGULAR && k == 0) gb_note_zero_key(ctl, bits, pos);
                    bt_update(L, B, Q, ins, row, pos, c, k, lhot, lcold, hot, cold, ctl, empty);
                }
            }
            continue;
        }
#pragma unroll
        for (int u = 0; u < BT_U; u++) {
            kn[u] = 0;
            cls[u] = VG_KEY_REGULAR;
            if (ok[u] && !L.key_none) cls[u] = vg_key_class(Q.key, base + (int64_t)u * NT, kn[u]);
        }
#pragma unroll
        for (int u = 0; u < BT_U; u++) {
            if (!ok[u]) continue;
            int64_t row = base + (int64_t)u * NT;
            if (L.key_is_double && cls[u] == VG_KEY_REGULAR && kn[u] == 0)
                gb_note_zero_key(ctl, vg_load_bits(Q.key, row), (uint64_t)row);
            bt_update(L, B, Q, ins, row, (uint64_t)row, cls[u], kn[u], lhot, lcold, hot, cold, ctl, empty);
        }
    }
}

template <int NT>
__global__ void __launch_bounds__(NT) k_bt_feed(GbLayout L, BtArgs B) {
    extern __shared__ uint64_t dynw[];
    __shared__ long long s_item;
    __shared__ InSet s_ins;
    __shared__ ColSet s_pcs;
    uint8_t *dyn = reinterpret_cast<uint8_t *>(dynw);
    uint64_t *regs = dynw;
    const uint64_t empty = L.key_is_double ? VG_EMPTY_F64 : VG_EMPTY_I64;
    uint64_t *lhot = reinterpret_cast<uint64_t *>(dyn + B.vm_lds);
    // The block keeps its LDS table across consecutive items of one query
    // (items come query-major off the queue) and merges it into the query's
    // HBM table only when the query changes: once per block and query
    // rather than once per item.
    int cur = -1;
    BtLocal Q;
    uint64_t *lcold = nullptr, *hot = nullptr, *cold = nullptr, *ctl = nullptr;
    for (;;) {
        if (threadIdx.x == 0) s_item = (long long)atomicAdd(B.queue, 1ull);
        __syncthreads();
        // the item is block-uniform
        const int64_t it = (int64_t)__builtin_amdgcn_readfirstlane((int)s_item);
        __syncthreads();   // every wave has read s_item before thread 0 may overwrite it
        if (it >= B.n_items) break;
        const BtItem item = B.items[it];
        if ((int)item.q >= B.n_active) continue;   // a rerun of fewer queries: block-uniform skip
        if ((int)item.q != cur) {
            if (cur >= 0 && Q.S) bt_lds_merge<NT>(L, B, Q, lhot, lcold, empty, hot, cold, ctl);
            __syncthreads();
            Q = bt_load_query<NT>(B.qs[item.q], s_ins, s_pcs);
            __syncthreads();
            cur = (int)item.q;
            lcold = B.lds_cold ? lhot + (size_t)(Q.S + 2) * L.RH : nullptr;
            hot = B.hot + Q.tbase * L.RH;
            cold = B.cold + Q.tbase * L.RC;
            ctl = B.ctl + (size_t)cur * GB_CTL_WORDS;
            if (Q.S) bt_lds_init<NT>(L, Q, lhot, lcold, empty);
        }
        const BtLocal Qi = Q;   // this item's copy stays in registers through the row loop
        const vgpu_prog &prog = B.progs[Qi.vm ? cur : 0];
        uint8_t *regv = dyn + (size_t)(Qi.vm ? prog.n_insn : 0) * NT * 8;
        const int64_t r0 = (int64_t)item.part * B.tile;
        const int64_t r1 = r0 + B.tile < Qi.n_rows ? r0 + B.tile : Qi.n_rows;
        bt_rows<NT>(L, B, Qi, s_ins, s_pcs, prog, regs, regv, r0, r1, lhot, lcold, hot, cold, ctl);
        __syncthreads();
    }
    if (cur >= 0 && Q.S) bt_lds_merge<NT>(L, B, Q, lhot, lcold, empty, hot, cold, ctl);
}

// Shared scans: every query reads the same rows.  Each block keeps one
// query for the whole launch (its LDS table is initialized and merged
// once) and walks a strided sequence of row tiles; the blocks of all K
// queries on one XCD (blocks are dispatched to XCDs round-robin by index)
// walk the same tiles at the same time, so a tile comes from HBM once per
// XCD and is re-read by the other queries' blocks from that XCD's L2.
#define BT_XCDS 8

__global__ void __launch_bounds__(GB_THREADS) k_bt_feed_shared(GbLayout L, BtArgs B) {
    extern __shared__ uint64_t dynw[];
    __shared__ InSet s_ins;
    __shared__ ColSet s_pcs;
    uint8_t *dyn = reinterpret_cast<uint8_t *>(dynw);
    uint64_t *regs = dynw;
    const uint64_t empty = L.key_is_double ? VG_EMPTY_F64 : VG_EMPTY_I64;
    uint64_t *lhot = reinterpret_cast<uint64_t *>(dyn + B.vm_lds);
    const int xcd = blockIdx.x % BT_XCDS, local = blockIdx.x / BT_XCDS;
    const int q = local % B.K, j = local / B.K;
    if (q >= B.n_queries || j >= B.J) return;
    const BtLocal Q = bt_load_query<GB_THREADS>(B.qs[q], s_ins, s_pcs);
    __syncthreads();
    const int S = Q.S;
    const vgpu_prog &prog = B.progs[Q.vm ? q : 0];
    uint8_t *regv = dyn + (size_t)(Q.vm ? prog.n_insn : 0) * GB_THREADS * 8;
    uint64_t *lcold = B.lds_cold ? lhot + (size_t)(S + 2) * L.RH : nullptr;
    uint64_t *hot = B.hot + Q.tbase * L.RH;
    uint64_t *cold = B.cold + Q.tbase * L.RC;
    uint64_t *ctl = B.ctl + (size_t)q * GB_CTL_WORDS;
    if (S) bt_lds_init<GB_THREADS>(L, Q, lhot, lcold, empty);
    for (int64_t i = 0;; i++) {
        const int64_t t = xcd + (int64_t)BT_XCDS * (j + (int64_t)B.J * i);
        if (t >= B.n_tiles) break;
        const int64_t r0 = t * BT_STILE;
        const int64_t r1 = r0 + BT_STILE < Q.n_rows ? r0 + BT_STILE : Q.n_rows;
        bt_rows<GB_THREADS>(L, B, Q, s_ins, s_pcs, prog, regs, regv, r0, r1, lhot, lcold, hot, cold, ctl);
    }
    if (S) bt_lds_merge<GB_THREADS>(L, B, Q, lhot, lcold, empty, hot, cold, ctl);
}

// sort entry of arena record (query q, table slot s): like gb_sort_entry,
// but the NA group of the key order takes the largest key instead of being
// set aside, so a stable sort puts it after the query's other groups
// (a dense table's record s below cap holds key kmin + s when a row
// reached it: its key word is left empty until the compaction writes it)
__device__ __forceinline__ bool bt_sort_entry(const GbLayout &L, const RecSrc &R, const BtQuery &Q, int order,
                                              int64_t s, uint64_t &sk) {
    if (order == VGPU_ORDER_KEY && s == R.na_slot) {
        if (R.hot[s * R.hstride + 1] == 0) return false;
        sk = ~0ull;
        return true;
    }
    if (Q.dense && s < Q.cap) {
        const uint64_t *h = R.hot + s * R.hstride;
        if (h[1] == 0) return false;
        sk = order == VGPU_ORDER_FIRST_SEEN ? h[L.w_fpos] : vg_i64_ordered((uint64_t)Q.kmin + (uint64_t)s);
        return true;
    }
    return gb_sort_entry(L, R, order, s, sk);
}

__device__ __forceinline__ RecSrc bt_src(const GbLayout &L, const BtArgs &B, const BtQuery &Q) {
    RecSrc R;
    R.hot = B.hot + Q.base * L.RH;
    R.cold = B.cold + Q.base * L.RC;
    R.hstride = L.RH;
    R.cstride = L.RC;
    R.na_slot = Q.cap;
    R.sp_slot = Q.cap + 1;
    return R;
}

// compaction blocks: item = (query, chunk of GB_CMP_SLOTS table slots)
__global__ void __launch_bounds__(GB_THREADS)
k_bt_cmp_count(GbLayout L, BtArgs B, const BtItem *cmap, int order, uint32_t *bcount, uint64_t *qcount) {
    __shared__ uint32_t lds[GB_THREADS / VG_WAVE];
    const BtItem item = cmap[blockIdx.x];
    const BtQuery &Q = B.qs[item.q];
    const RecSrc R = bt_src(L, B, Q);
    const int64_t base = (int64_t)item.part * GB_CMP_SLOTS, n_slots = Q.cap + 2;
    uint32_t c = 0;
    for (int j = 0; j < GB_CMP_PER_THREAD; j++) {
        int64_t s = base + j * GB_THREADS + threadIdx.x;
        uint64_t sk;
        if (s < n_slots && bt_sort_entry(L, R, Q, order, s, sk)) c++;
    }
    uint32_t total;
    gb_block_prefix(c, lds, total);
    if (threadIdx.x == 0) {
        bcount[blockIdx.x] = total;
        if (total) atomicAdd((unsigned long long *)&qcount[item.q], (unsigned long long)total);
    }
}

__global__ void __launch_bounds__(GB_THREADS)
k_bt_cmp_write(GbLayout L, BtArgs B, const BtItem *cmap, int order, const uint64_t *boff, uint64_t *skeys,
               uint32_t *svals, uint64_t *por, uint64_t *pand) {
    __shared__ uint32_t lds[GB_THREADS / VG_WAVE];
    __shared__ uint64_t lor[2 * GB_THREADS / VG_WAVE];
    const BtItem item = cmap[blockIdx.x];
    const BtQuery &Q = B.qs[item.q];
    const RecSrc R = bt_src(L, B, Q);
    const int64_t base = (int64_t)item.part * GB_CMP_SLOTS, n_slots = Q.cap + 2;
    uint64_t sk[GB_CMP_PER_THREAD];
    uint32_t mask = 0;
    uint64_t o = 0, a = ~0ull;
#pragma unroll
    for (int j = 0; j < GB_CMP_PER_THREAD; j++) {
        int64_t s = base + j * GB_THREADS + threadIdx.x;
        sk[j] = 0;
        if (s < n_slots && bt_sort_entry(L, R, Q, order, s, sk[j])) {
            mask |= 1u << j;
            if (Q.dense && s < Q.cap) B.hot[(size_t)(Q.base + s) * L.RH] = (uint64_t)Q.kmin + (uint64_t)s;
            o |= sk[j];
            a &= sk[j];
        }
    }
    uint32_t total;
    uint64_t off = boff[blockIdx.x] + gb_block_prefix(__builtin_popcount(mask), lds, total);
#pragma unroll
    for (int j = 0; j < GB_CMP_PER_THREAD; j++)
        if (mask & (1u << j)) {
            skeys[off] = sk[j];
            svals[off] = (uint32_t)(Q.base + base + j * GB_THREADS + threadIdx.x);
            off++;
        }
    gb_block_or_and(o, a, lor);
    if (threadIdx.x == 0) {
        por[blockIdx.x] = o;
        pand[blockIdx.x] = a;
    }
}

// the run's keys must vary only within the plan's sort range, and the run
// must sort as many records as the plan; otherwise the batch is flagged
__global__ void k_bt_guard(uint64_t *bctl, uint64_t range_mask, const uint64_t *total, uint64_t n_pairs) {
    uint64_t varying = bctl[6] ^ bctl[7];
    if ((varying & ~range_mask) || *total != n_pairs) bctl[BT_CTL_ERR] = 1;
}

// query of every sorted record (the second sort's key)
__global__ void k_bt_qkeys(const uint32_t *svals, int64_t n, const uint64_t *qbase, int n_queries, uint64_t *qk) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint64_t s = svals[i];
    int lo = 0, hi = n_queries - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (qbase[mid] <= s) lo = mid;
        else hi = mid - 1;
    }
    qk[i] = (uint64_t)lo;
}

template <bool STAGED>
__global__ void __launch_bounds__(GB_THREADS)
k_bt_emit(GbLayout L, GbAggs AG, BtArgs B, const uint32_t *svals, const uint64_t *qk, int64_t n_out, EmitOut O) {
    extern __shared__ uint64_t dynw[];
    __shared__ int64_t sidx[GB_THREADS];
    int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool in = g < n_out;
    int q = in ? (int)qk[g] : 0;
    const BtQuery &Q = B.qs[q];
    RecSrc R = bt_src(L, B, Q);
    int64_t s = in ? (int64_t)svals[g] - Q.base : 0;   // slot within the query's table
    const uint64_t *h, *c;
    if (STAGED) {
        RecSrc A;   // the whole arena: staging reads by arena slot
        A.hot = B.hot;
        A.cold = B.cold;
        A.hstride = L.RH;
        A.cstride = L.RC;
        A.na_slot = A.sp_slot = -1;
        sidx[threadIdx.x] = in ? (int64_t)svals[g] : 0;
        __syncthreads();
        const int64_t g0 = (int64_t)blockIdx.x * blockDim.x;
        const int cnt = n_out - g0 < GB_THREADS ? (int)(n_out - g0) : GB_THREADS;
        gb_stage_records(L, A, cnt, [&](int k) { return sidx[k]; }, dynw);
        __syncthreads();
        h = dynw + threadIdx.x * (L.RH + L.RC);
        c = h + L.RH;
    } else {
        h = R.hot + s * R.hstride;
        c = R.cold + s * R.cstride;
    }
    gb_emit_row(L, AG, R, g, in, s, h, c, n_out, 8, O, B.ctl + (size_t)q * GB_CTL_WORDS);
}

// ---------------------------------------------------------------------------
// bucketed passes: queries that differ only in a comparison constant
//
// When every query of a batch reads the same columns and filters them with
// `col op param_q` (op one of <, <=, >, >=), the row sets the queries
// select form a chain under inclusion: ordered by decreasing size (by
// param, with a NaN param at the end its comparison rule puts it: NaN
// compares as "equal", so >= and <= select every row and > and < none),
// a row that passes the query at chain position c passes every query
// before it.  A row is then described by ONE number T, the count of the
// queries it passes (found by binary search over the constants in chain
// order, each comparison made exactly as the per-query filter makes it),
// and the query at position c selects the rows with T >= c.  One scan
// aggregates every selected row once into a dense LDS table keyed by
// (group key, T); the query at position c gets, per group, the sum of the
// group's buckets c..K.  Counts and integer sums are exact; double sums
// are the same values added in another order.
//
// Shape: int64 / bool group key over a dense range (the NA group one more
// entry), count / sum / mean aggregates of inputs without NAs, key order.
// Blocks of 1,024 threads take contiguous row ranges, write their whole
// table to a slab of their own, one launch sums the slabs, one expands
// every (query, group) into the query's table of the arena; the finish is
// the batch's usual one.
// ---------------------------------------------------------------------------

#define BT_CH_THREADS 1024
#define BT_CH_U 8
#define BT_CH_MAX_SUMS 2
#define BT_CH_LDS (160 * 1024 - 2048)   // dynamic LDS a bucketed feed block may take

struct BtChainArgs {
    vgpu_col key;          // the group key column (batches without a join)
    vgpu_col probe_key;    // join batches: the probe key column
    vgpu_col pcol;         // the compared column
    InSet ins;
    BtJoin jn;
    PredInfo pi;           // the comparison (operator, types); constants in ks
    const uint64_t *ks;    // the active queries' constants in chain order
    const int32_t *rank;   // per query: its chain position (1..K)
    uint64_t *slab;        // per block: E*m counts, then E*m*n_sum sums
    uint64_t *tab;         // the slabs summed
    uint64_t *bctl;        // the batch's control words
    int64_t n_rows, rows_per_block;
    int64_t kmin;          // group key of entry 0
    int32_t E;             // entries: the key range, then the NA group
    int32_t m;             // bucket slots per entry (the batch's queries)
    int32_t K;             // active queries
    int32_t n_sum, key_type, G;
    int32_t f64_mask;      // bit s: sum s adds doubles (else int64 / bool bits, wrapping)
    int32_t sum_in[BT_CH_MAX_SUMS];
    int32_t sum_pred[BT_CH_MAX_SUMS];   // sum s reads the compared column itself
    JDense jd;             // join batches: direct-addressed map of the group key payload (map null: hash probe)
};

// the count of the active queries a valid compared value passes
__device__ __forceinline__ int bt_chain_T(const BtChainArgs &A, const uint64_t *s_k, uint64_t bits) {
    int lo = 0, hi = A.K;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (vg_fast_cmp(A.pi, bits, s_k[mid - 1])) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Up to BT_CH_LIN active queries over a plain scan whose comparison and
// column share one type: T counted branch-free against the constants held
// in registers, the operator fixed at compile time.  With cmp = (a>b)-(a<b)
// (a NaN operand gives 0), `>` passes iff v > t, `<` iff v < t, `>=` iff
// !(v < t) and `<=` iff !(v > t): GREATER counts v > t, else v < t, and NEG
// (>= / <=) takes K minus the count.  Slots past K hold a constant that no
// value passes the counted comparison with (+inf / INT64_MAX for v > t,
// -inf / INT64_MIN for v < t), so the count runs over all NL slots (8 for up
// to 8 queries, 16 for up to 16).
#define BT_CH_LIN 16
// TM (T mode): 0 = the search above; 1 + 2*int64 + less: the count
#define BT_TM_SEARCH 0
__host__ __device__ constexpr int bt_tm(bool f64, bool greater) { return 1 + (f64 ? 0 : 2) + (greater ? 0 : 1); }
// TM 5..8 (TM - 4 as above): more than BT_CH_LIN queries on the plain scan:
// a branch-free binary search of the constants in LDS with the comparison
// fixed at compile time (7 steps, up to 127 queries)
#define BT_TM_CSEARCH 4

template <int TM>
__device__ __forceinline__ int bt_chain_T_csearch(const uint64_t *s_k, int K, uint64_t bits, bool neg) {
    constexpr bool F64 = TM == 1 || TM == 2;
    constexpr bool GREATER = TM == 1 || TM == 3;
    // the queries passed form a prefix of the chain: its length, by steps of
    // 64, 32, ..., 1 (a step past K never passes)
    int lo = 0;
#pragma unroll
    for (int step = 64; step >= 1; step >>= 1) {
        const int at = lo + step;
        const int ix = (at <= K ? at : K) - 1;   // the constant the step tests (K >= 1 whenever rows pass)
        const uint64_t t = s_k[ix > 0 ? ix : 0];
        bool base;
        if (F64) {
            const double v = vg_bits_to_f64(bits), c = vg_bits_to_f64(t);
            base = GREATER ? v > c : v < c;
        } else {
            base = GREATER ? (int64_t)bits > (int64_t)t : (int64_t)bits < (int64_t)t;
        }
        if (at <= K && base != neg) lo = at;
    }
    return lo;
}

template <int TM, int NL>
__device__ __forceinline__ int bt_chain_T_lin(const uint64_t (&kr)[BT_CH_LIN], uint64_t bits, int K, bool neg) {
    static_assert(NL == 8 || NL == 16, "8 or 16 compare slots");
    constexpr bool F64 = TM == 1 || TM == 2;
    constexpr bool GREATER = TM == 1 || TM == 3;
    int c = 0;
#pragma unroll
    for (int i = 0; i < NL; i++) {
        bool pass;
        if (F64) {
            const double v = vg_bits_to_f64(bits), t = vg_bits_to_f64(kr[i]);
            pass = GREATER ? v > t : v < t;
        } else {
            const int64_t v = (int64_t)bits, t = (int64_t)kr[i];
            pass = GREATER ? v > t : v < t;
        }
        c += pass ? 1 : 0;
    }
    return neg ? K - c : c;
}

// the constants of a linear count in registers (uniform loads), padded
template <int TM, int NL>
__device__ __forceinline__ void bt_chain_load_lin(const BtChainArgs &A, uint64_t (&kr)[BT_CH_LIN]) {
    constexpr bool F64 = TM == 1 || TM == 2;
    constexpr bool GREATER = TM == 1 || TM == 3;
    const uint64_t pad = F64 ? (GREATER ? 0x7FF0000000000000ull : 0xFFF0000000000000ull)
                             : (GREATER ? 0x7FFFFFFFFFFFFFFFull : 0x8000000000000000ull);
#pragma unroll
    for (int i = 0; i < BT_CH_LIN; i++) kr[i] = i < NL && i < A.K ? A.ks[i] : pad;
}

// dense entry of a group key, the NA group last; -1 outside the plan's range
__device__ __forceinline__ int64_t bt_chain_entry(const BtChainArgs &A, bool valid, uint64_t bits) {
    if (!valid) return A.E - 1;
    const uint64_t d = bits - (uint64_t)A.kmin;
    return d < (uint64_t)(A.E - 1) ? (int64_t)d : -1;
}

__device__ __forceinline__ void bt_chain_add(const BtChainArgs &A, uint32_t *cnt, uint64_t *sums, int64_t e, int T,
                                             const uint64_t (&sv)[BT_CH_MAX_SUMS]) {
    const int64_t idx = e * A.m + (T - 1);
    atomicAdd(&cnt[idx], 1u);
#pragma unroll
    for (int s = 0; s < BT_CH_MAX_SUMS; s++) {
        if (s >= A.n_sum) break;
        uint64_t *p = &sums[idx * A.n_sum + s];
        if ((A.f64_mask >> s) & 1) atomicAdd(reinterpret_cast<double *>(p), vg_bits_to_f64(sv[s]));
        else atomicAdd((unsigned long long *)p, (unsigned long long)sv[s]);
    }
}

// bt_chain_add for one sum (n_sum == 1) of a known type
template <bool F64>
__device__ __forceinline__ void bt_chain_add_one(const BtChainArgs &A, uint32_t *cnt, uint64_t *sums, int64_t e,
                                                 int T, uint64_t v) {
    const int64_t idx = e * A.m + (T - 1);
    atomicAdd(&cnt[idx], 1u);
    if (F64) atomicAdd(reinterpret_cast<double *>(&sums[idx]), vg_bits_to_f64(v));
    else atomicAdd((unsigned long long *)&sums[idx], (unsigned long long)v);
}

// one thread's rows of one step: compared value and group / probe key of
// each, the summed inputs (SP: every sum reads the compared column, so none
// is kept), bit u of ok / kv: row u is in range with a valid compared
// value / a valid key
template <int U, bool SP>
struct BtRows {
    uint64_t pv[U], kb[U];
    uint64_t sv[SP ? 1 : U][BT_CH_MAX_SUMS];
    uint32_t ok, kv;
};

// PLAIN: the compared column and the key are 8-byte columns without
// validity and every sum reads the compared column; rows past r1 read row
// r1 - 1 (the block's range is not empty) and are masked out, so every load
// of the step is issued without a branch and none is waited on here.
template <int U, bool SP, bool PLAIN>
__device__ __forceinline__ void bt_chain_rows(const BtChainArgs &A, const vgpu_col &kc, int64_t base, int64_t r1,
                                              BtRows<U, SP> &R) {
    if (PLAIN) {
        static_assert(!PLAIN || SP, "the plain scan keeps no separate sum inputs");
        const uint64_t *pd = static_cast<const uint64_t *>(A.pcol.data);
        const uint64_t *kd = static_cast<const uint64_t *>(kc.data);
        uint32_t in_mask = 0;
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int64_t row = base + (int64_t)u * BT_CH_THREADS;
            const bool in = row < r1;
            const int64_t rc = in ? row : r1 - 1;
            R.pv[u] = __builtin_nontemporal_load(pd + rc);
            R.kb[u] = __builtin_nontemporal_load(kd + rc);
            in_mask |= (uint32_t)in << u;
        }
        R.ok = in_mask;
        R.kv = in_mask;
        return;
    }
    R.ok = 0;
    R.kv = 0;
#pragma unroll
    for (int u = 0; u < U; u++) {
        const int64_t row = base + (int64_t)u * BT_CH_THREADS;
        const bool in = row < r1;
        const bool ok = in && vg_valid(A.pcol.validity, row);
        const bool kv = in && vg_valid(kc.validity, row);
        R.pv[u] = ok ? vg_load_bits_stream(A.pcol, row) : 0;
        R.kb[u] = kv ? vg_load_bits_stream(kc, row) : 0;
        R.ok |= (uint32_t)ok << u;
        R.kv |= (uint32_t)kv << u;
        if (!SP) {
#pragma unroll
            for (int s = 0; s < BT_CH_MAX_SUMS; s++) {
                uint64_t v = 0;
                if (s < A.n_sum && ok) v = A.sum_pred[s] ? R.pv[u] : vg_load_bits_stream(A.ins.c[A.sum_in[s]], row);
                R.sv[SP ? 0 : u][s] = v;
            }
        }
    }
}

// the bucket of every row of a step that passes at least one query, and
// its aggregation into the block's table
template <int TM, int NL, int U, bool SP>
__device__ __forceinline__ void bt_chain_rows_add(const BtChainArgs &A, const vgpu_col &kc, const uint64_t *s_k,
                                                  const uint64_t (&kr)[BT_CH_LIN], bool neg, uint32_t *cnt,
                                                  uint64_t *sums, const BtRows<U, SP> &R, bool &err) {
    int T[U];
    uint64_t sv[U][BT_CH_MAX_SUMS];
#pragma unroll
    for (int u = 0; u < U; u++) {
        if (TM == BT_TM_SEARCH) T[u] = (R.ok >> u) & 1 ? bt_chain_T(A, s_k, R.pv[u]) : 0;
        else if (TM > BT_TM_CSEARCH)
            T[u] = (R.ok >> u) & 1 ? bt_chain_T_csearch<TM - BT_TM_CSEARCH>(s_k, A.K, R.pv[u], neg) : 0;
        else T[u] = (R.ok >> u) & 1 ? bt_chain_T_lin<TM, NL>(kr, R.pv[u], A.K, neg) : 0;
#pragma unroll
        for (int s = 0; s < BT_CH_MAX_SUMS; s++) sv[u][s] = SP ? R.pv[u] : R.sv[SP ? 0 : u][s];
    }
    if (A.jn.on && A.jd.map) {
        // unique int64 build keys over a short range: one 2-byte load per
        // row; NA probe keys never match (src/join.c:85)
#pragma unroll
        for (int u = 0; u < U; u++) {
            if (!T[u] || !((R.kv >> u) & 1)) continue;
            const uint64_t off = R.kb[u] - (uint64_t)A.jd.kmin;
            if (off >= A.jd.range) continue;
            const uint32_t c = A.jd.map[off];
            if (c == JD_ABSENT) continue;
            const int64_t e = c == JD_NA_PAYLOAD ? A.E - 1 : bt_chain_entry(A, true, (uint64_t)(A.jd.pmin + (int64_t)c));
            if (e < 0) err = true;
            else if (TM == BT_TM_SEARCH) bt_chain_add(A, cnt, sums, e, T[u], sv[u]);
            else bt_chain_add_one<(TM > BT_TM_CSEARCH ? TM - BT_TM_CSEARCH : TM) == 1 || (TM > BT_TM_CSEARCH ? TM - BT_TM_CSEARCH : TM) == 2>(A, cnt, sums, e, T[u], sv[u][0]);
        }
    } else if (A.jn.on) {
        const uint64_t *ent[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            ent[u] = nullptr;
            if (T[u]) {
                uint64_t kn;
                const int cls = vg_key_class_bits(kc.type, (R.kv >> u) & 1, R.kb[u], kn);
                ent[u] = jt_find_key(A.jn.jv, cls, kn);
            }
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            if (!ent[u]) continue;
            // table entries are device memory: global loads (the pointer
            // array alone would be read with flat ones)
            const uint64_t segw = vg_gload<uint64_t>(ent[u], 1);
            const uint32_t n = jt_count(A.jn.jv, segw), off = (uint32_t)(segw >> 32);
            const int p = A.jn.key_payload;
            for (uint32_t j = 0; j < n; j++) {
                bool v;
                uint64_t bits;
                if (j == 0) {
                    v = (segw >> (24 + p)) & 1;
                    bits = vg_gload<uint64_t>(ent[u], 2 + p);
                } else {
                    v = A.jn.jv.cval[p][off + j];
                    bits = A.jn.jv.cpay[p][off + j];
                }
                const int64_t e = bt_chain_entry(A, v, bits);
                if (e < 0) err = true;
                else if (TM == BT_TM_SEARCH) bt_chain_add(A, cnt, sums, e, T[u], sv[u]);
                else bt_chain_add_one<(TM > BT_TM_CSEARCH ? TM - BT_TM_CSEARCH : TM) == 1 || (TM > BT_TM_CSEARCH ? TM - BT_TM_CSEARCH : TM) == 2>(A, cnt, sums, e, T[u], sv[u][0]);
            }
        }
    } else {
#pragma unroll
        for (int u = 0; u < U; u++) {
            if (!T[u]) continue;
            const int64_t e = bt_chain_entry(A, (R.kv >> u) & 1, R.kb[u]);
            if (e < 0) err = true;
            else if (TM == BT_TM_SEARCH) bt_chain_add(A, cnt, sums, e, T[u], sv[u]);
            else bt_chain_add_one<(TM > BT_TM_CSEARCH ? TM - BT_TM_CSEARCH : TM) == 1 || (TM > BT_TM_CSEARCH ? TM - BT_TM_CSEARCH : TM) == 2>(A, cnt, sums, e, T[u], sv[u][0]);
        }
    }
}

// SP: every sum reads the compared column (8 rows per thread and step),
// else 4 rows per thread and step.  The next step's columns are loaded
// before the current step is aggregated, so a wave keeps its loads in
// flight while it works on LDS: with one 1,024-thread block per CU the
// scan is otherwise bound by the loads the 16 waves have outstanding.
// TM != 0 (a linear count, PLAIN only): at most NL active queries,
// one sum; neg: the operator is >= or <=.
template <bool SP, bool PLAIN, int TM, int NL = 8>
__global__ void __launch_bounds__(BT_CH_THREADS) k_bt_chain_feed(BtChainArgs A, int neg) {
    static_assert(TM == BT_TM_SEARCH || PLAIN, "the linear count runs on the plain scan");
    constexpr int U = SP ? BT_CH_U : BT_CH_U / 2;
    uint64_t kr[BT_CH_LIN];
    if (TM != BT_TM_SEARCH && TM <= BT_TM_CSEARCH) bt_chain_load_lin<TM, NL>(A, kr);
    extern __shared__ uint64_t dynw[];
    uint64_t *s_k = dynw;                                   // m constants
    uint64_t *sums = dynw + A.m;                            // E*m*n_sum
    const int64_t Em = (int64_t)A.E * A.m;
    uint32_t *cnt = reinterpret_cast<uint32_t *>(sums + Em * A.n_sum);   // E*m
    for (int i = threadIdx.x; i < A.K; i += BT_CH_THREADS) s_k[i] = A.ks[i];
    for (int64_t i = threadIdx.x; i < Em * A.n_sum; i += BT_CH_THREADS) sums[i] = 0;
    for (int64_t i = threadIdx.x; i < Em; i += BT_CH_THREADS) cnt[i] = 0;
    __syncthreads();
    const int64_t r0 = (int64_t)blockIdx.x * A.rows_per_block;
    const int64_t r1 = r0 + A.rows_per_block < A.n_rows ? r0 + A.rows_per_block : A.n_rows;
    const vgpu_col kc = A.jn.on ? A.probe_key : A.key;   // a copy: a reference to either kernel argument would be read with flat loads
    const int64_t step = (int64_t)BT_CH_THREADS * U;
    bool err = false;
    BtRows<U, SP> cur, nxt;
    if (r0 < r1) {
        bt_chain_rows<U, SP, PLAIN>(A, kc, r0 + threadIdx.x, r1, cur);
        for (int64_t base = r0 + threadIdx.x; base < r1; base += step) {
            bt_chain_rows<U, SP, PLAIN>(A, kc, base + step, r1, nxt);
            bt_chain_rows_add<TM, NL>(A, kc, s_k, kr, neg != 0, cnt, sums, cur, err);
            cur = nxt;
        }
    }
    if (err) atomicOr((unsigned long long *)&A.bctl[BT_CTL_ERR], 1ull);
    __syncthreads();
    uint64_t *slab = A.slab + (size_t)blockIdx.x * (size_t)Em * (1 + A.n_sum);
    for (int64_t i = threadIdx.x; i < Em; i += BT_CH_THREADS) slab[i] = cnt[i];
    for (int64_t i = threadIdx.x; i < Em * A.n_sum; i += BT_CH_THREADS) slab[Em + i] = sums[i];
}

// tab[w] = the sum over the blocks' slabs of word w (in block order)
__global__ void k_bt_chain_reduce(BtChainArgs A) {
    const int64_t Em = (int64_t)A.E * A.m, W = Em * (1 + A.n_sum);
    const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= W) return;
    const bool f64 = w >= Em && ((A.f64_mask >> (int)((w - Em) % A.n_sum)) & 1);
    if (f64) {
        double acc = 0.0;
        for (int g = 0; g < A.G; g++) acc += vg_bits_to_f64(A.slab[(size_t)g * W + w]);
        A.tab[w] = vg_f64_to_bits(acc);
    } else {
        uint64_t acc = 0;
        for (int g = 0; g < A.G; g++) acc += A.slab[(size_t)g * W + w];
        A.tab[w] = acc;
    }
}

// every (active query, entry): the merge of the entry's buckets from the
// query's chain position on, into the query's table when it has rows
__global__ void k_bt_chain_expand(GbLayout L, BtArgs B, BtChainArgs A) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)A.K * A.E) return;
    const int q = (int)(i / A.E), e = (int)(i % A.E);
    const int r = A.rank[q];
    const int64_t Em = (int64_t)A.E * A.m;
    uint64_t n = 0;
    uint64_t su[BT_CH_MAX_SUMS];
    double sf[BT_CH_MAX_SUMS];
#pragma unroll
    for (int s = 0; s < BT_CH_MAX_SUMS; s++) {
        su[s] = 0;
        sf[s] = 0.0;
    }
    for (int b = r - 1; b < A.K; b++) {
        const int64_t idx = (int64_t)e * A.m + b;
        n += A.tab[idx];
#pragma unroll
        for (int s = 0; s < BT_CH_MAX_SUMS; s++) {
            if (s >= A.n_sum) break;
            const uint64_t w = A.tab[Em + idx * A.n_sum + s];
            if ((A.f64_mask >> s) & 1) sf[s] += vg_bits_to_f64(w);
            else su[s] += w;
        }
    }
    if (n == 0) return;
    const BtQuery &Q = B.qs[q];
    uint64_t *hot = B.hot + (size_t)Q.base * L.RH;
    uint64_t *ctl = B.ctl + (size_t)q * GB_CTL_WORDS;
    uint64_t kn = 0;
    const int cls = e == A.E - 1 ? VG_KEY_NA
                                 : vg_key_class_bits(A.key_type, true, (uint64_t)A.kmin + (uint64_t)e, kn);
    const int64_t gs = bt_table_slot(L, B, Q.dense, Q.kmin, hot, Q.cap, cls, kn, 0, ctl);
    if (gs < 0) return;
    uint64_t *h = hot + (size_t)gs * L.RH;
    gb_atomic_add(&h[1], n);
#pragma unroll
    for (int s = 0; s < BT_CH_MAX_SUMS; s++) {
        if (s >= A.n_sum) break;
        uint64_t *w = &h[L.w_sum[A.sum_in[s]]];
        if ((A.f64_mask >> s) & 1) gb_atomic_addf(w, sf[s]);
        else gb_atomic_add(w, su[s]);
    }
}


// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------

static vgpu_batch *batch_create_body(int key_type, int n_inputs, const int32_t *input_types, int n_aggs,
                                     const vgpu_agg *aggs, int order, const vgpu_prog *pred, int n_pred_cols,
                                     int param_insn) {
    if (vg_ensure_init()) return nullptr;
    if (key_type != VGPU_INT64 && key_type != VGPU_DOUBLE && key_type != VGPU_BOOL && key_type != VGPU_KEY_NONE) {
        vg_set_error("batch: unsupported key type %d", key_type);
        return nullptr;
    }
    if (n_inputs < 0 || n_inputs > VGPU_MAX_INPUTS || n_aggs < 0 || n_aggs > VGPU_MAX_AGGS) {
        vg_set_error("batch: %d inputs / %d aggregates exceed the limits (%d / %d)", n_inputs, n_aggs,
                     VGPU_MAX_INPUTS, VGPU_MAX_AGGS);
        return nullptr;
    }
    if (order != VGPU_ORDER_FIRST_SEEN && order != VGPU_ORDER_KEY) {
        vg_set_error("batch: unknown order %d", order);
        return nullptr;
    }
    if (pred && (pred->n_insn <= 0 || pred->n_insn > VGPU_MAX_INSN || n_pred_cols < 0 ||
                 n_pred_cols > VGPU_MAX_COLS)) {
        vg_set_error("batch: invalid predicate program");
        return nullptr;
    }
    if (pred) {
        for (int i = 0; i < pred->n_insn; i++)
            if (pred->insn[i].op == VGPU_OP_COL && (pred->insn[i].col < 0 || pred->insn[i].col >= n_pred_cols)) {
                vg_set_error("batch: predicate reads column slot %d of %d", pred->insn[i].col, n_pred_cols);
                return nullptr;
            }
        if (param_insn >= 0 && (param_insn >= pred->n_insn || pred->insn[param_insn].op != VGPU_OP_CONST)) {
            vg_set_error("batch: parameter instruction %d is not a constant of the predicate", param_insn);
            return nullptr;
        }
    }
    vgpu_batch *b = new vgpu_batch();
    b->order = key_type == VGPU_KEY_NONE ? VGPU_ORDER_FIRST_SEEN : order;
    b->key_type = key_type;
    b->aggs.n = n_aggs;
    for (int a = 0; a < n_aggs; a++) b->aggs.a[a] = aggs[a];
    memset(&b->prog, 0, sizeof b->prog);
    if (pred) {
        b->prog = *pred;
        b->has_pred = 1;
        b->n_pred_cols = n_pred_cols;
        b->param_insn = param_insn;
    }
    if (gb_layout(b->L, b->order, key_type, n_inputs, input_types, n_aggs, aggs)) {
        delete b;
        return nullptr;
    }
    b->out_vals.resize(n_aggs);
    b->out_valid.resize(n_aggs);
    memset(&b->join, 0, sizeof b->join);
    memset(&b->join_payload, 0, sizeof b->join_payload);
    return b;
}

extern "C" vgpu_batch *vgpu_batch_create(int key_type, int n_inputs, const int32_t *input_types, int n_aggs,
                                         const vgpu_agg *aggs, int order, const vgpu_prog *pred, int n_pred_cols,
                                         int param_insn) {
    return vg_guard<vgpu_batch *>(nullptr, "vgpu_batch_create", [&] { return batch_create_body(key_type, n_inputs,
        input_types, n_aggs, aggs, order, pred, n_pred_cols, param_insn); });
}

static int batch_join_body(vgpu_batch *b, const vgpu_jtable *t, int key_payload) {
    if (b->planned) {
        vg_set_error("batch: join after plan");
        return -1;
    }
    if (!t || key_payload < 0 || key_payload >= t->n_payload) {
        vg_set_error("batch: the group key must be an embedded payload column of the join table");
        return -1;
    }
    if (b->L.key_none || t->payload[key_payload].type != b->key_type) {
        vg_set_error("batch: payload column %d has type %d, the batch groups by type %d", key_payload,
                     t->payload[key_payload].type, b->key_type);
        return -1;
    }
    b->join.jv = t->view();
    b->join.on = 1;
    b->join.key_payload = key_payload;
    b->join.key_type = b->key_type;
    b->join_key_type = t->key_type;
    b->join_payload = t->payload[key_payload];
    b->join_rows = t->n_build;
    // the table's lazily built direct-addressed map is internal state of it
    b->jtab = const_cast<vgpu_jtable *>(t);
    return 0;
}

extern "C" int vgpu_batch_join(vgpu_batch *b, const vgpu_jtable *t, int key_payload) {
    return vg_guard<int>(-1, "vgpu_batch_join", [&] { return batch_join_body(b, t, key_payload); });
}

static int bt_enqueue(vgpu_batch *b, bool first);

// estimate of one query's groups from its sketch registers
static int64_t bt_estimate(const uint32_t *r, uint64_t nan_rows) {
    double m = BT_SK_REGS, z = 0;
    int zeros = 0;
    for (int i = 0; i < BT_SK_REGS; i++) {
        z += std::ldexp(1.0, -(int)r[i]);
        zeros += r[i] == 0;
    }
    double e = 0.7213 / (1.0 + 1.079 / m) * m * m / z;
    if (e <= 2.5 * m && zeros > 0) e = m * std::log(m / zeros);
    return (int64_t)std::ceil(e) + (int64_t)nan_rows + 2;
}

// ---- host side of the bucketed passes

// Can the batch's queries (all reading Q0's columns) share one bucketed
// aggregation?  Fills the sums' description when they can.
static bool bt_chain_shape(vgpu_batch *b, const BtQuery &Q0) {
    const GbLayout &L = b->L;
    if (vg_tune(VG_TUNE_BT_BUCKETED) < 2 || !b->has_pred || b->param_insn < 0 || !Q0.pi.fast) return false;
    const int op = Q0.pi.cmp_op;
    if (op != VGPU_OP_LT && op != VGPU_OP_LE && op != VGPU_OP_GT && op != VGPU_OP_GE) return false;
    if (L.key_none || L.key_is_double || b->order != VGPU_ORDER_KEY || L.w_fpos >= 0) return false;
    const int kt = b->join.on ? b->join.key_type : b->key_type;
    if (kt != VGPU_INT64 && kt != VGPU_BOOL) return false;
    for (int a = 0; a < b->aggs.n; a++) {
        const int k = b->aggs.a[a].kind;
        if (k != VGPU_AGG_COUNT_STAR && k != VGPU_AGG_COUNT && k != VGPU_AGG_SUM && k != VGPU_AGG_MEAN) return false;
    }
    const vgpu_col &pc = Q0.pcs.c[Q0.pi.col];
    int ns = 0, f64 = 0;
    for (int i = 0; i < L.n_inputs; i++) {
        if (Q0.ins.c[i].validity) return false;
        if (L.w_sum[i] < 0) continue;
        if (ns == BT_CH_MAX_SUMS) return false;
        b->ch_sum_in[ns] = i;
        b->ch_sum_pred[ns] = Q0.ins.c[i].data == pc.data && Q0.ins.c[i].type == pc.type;
        if (L.in_type[i] == VGPU_DOUBLE) f64 |= 1 << ns;
        ns++;
    }
    b->ch_nsum = ns;
    b->ch_f64 = f64;
    b->ch_key_type = kt;
    return true;
}

// key range of the group key (one synchronization), table and launch
// sizes; leaves b->chain at 0 when the dense table does not fit in LDS
static int bt_chain_plan(vgpu_batch *b, const BtQuery &Q0) {
    hipStream_t st = vg_stream();
    const vgpu_col kc = b->join.on ? b->join_payload : Q0.key;
    const int64_t kn = b->join.on ? b->join_rows : Q0.n_rows;
    DevBuf mm;
    if (mm.reserve(16)) return -1;
    unsigned long long h[2] = {~0ull, 0ull};
    VG_TRY(hipMemcpyAsync(mm.ptr, h, 16, hipMemcpyHostToDevice, st));
    if (kn > 0) {
        k_key_minmax<<<vg_blocks(kn, 256, 2048), 256, 0, st>>>(kc, kn, mm.as<unsigned long long>());
        VG_LAUNCH_CHECK();
    }
    VG_TRY(hipMemcpyAsync(h, mm.ptr, 16, hipMemcpyDeviceToHost, st));
    VG_TRY(hipStreamSynchronize(st));
    int64_t range = 0;   // valid key values lo..hi
    if (h[0] <= h[1]) {
        const unsigned long long d = h[1] - h[0];
        if (d >= (1ull << 24)) return 0;
        range = (int64_t)d + 1;
    }
    const int64_t E = range + 1, m = b->n_queries, Em = E * m;
    const size_t lds = (size_t)m * 8 + (size_t)Em * b->ch_nsum * 8 + (size_t)Em * 4;
    if (lds > BT_CH_LDS) return 0;
    const int64_t step = (int64_t)BT_CH_THREADS * BT_CH_U;
    const int64_t bpc = lds <= 78 * 1024 ? 2 : 1;   // 2 blocks of 1,024 threads fill a CU's 32 waves
    int64_t G = std::min<int64_t>((int64_t)vg_cus() * bpc, std::max<int64_t>(1, (Q0.n_rows + step - 1) / step));
    const int64_t rpb = (Q0.n_rows + G - 1) / G;
    if (rpb > (int64_t)0xFFFFFFFF) return 0;   // LDS counts are 32-bit
    const size_t W = (size_t)Em * (1 + b->ch_nsum);
    if (b->ch_slab.reserve((size_t)G * W * 8) || b->ch_tab.reserve(W * 8) ||
        b->ch_aux.reserve((size_t)m * 12 + 16))
        return -1;
    if (lds > 65536) {
        const void *fns[] = {reinterpret_cast<const void *>(k_bt_chain_feed<true, true, 1>),
                             reinterpret_cast<const void *>(k_bt_chain_feed<true, true, 2>),
                             reinterpret_cast<const void *>(k_bt_chain_feed<true, true, 3>),
                             reinterpret_cast<const void *>(k_bt_chain_feed<true, true, 4>),
                             reinterpret_cast<const void *>(k_bt_chain_feed<true, true, 1, 16>),
                             reinterpret_cast<const void *>(k_bt_chain_feed<true, true, 2, 16>),
                             reinterpret_cast<const void *>(k_bt_chain_feed<true, true, 3, 16>),
                             reinterpret_cast<const void *>(k_bt_chain_feed<true, true, 4, 16>),
                             reinterpret_cast<const void *>(k_bt_chain_feed<true, true, 5>),
                             reinterpret_cast<const void *>(k_bt_chain_feed<true, true, 6>),
                             reinterpret_cast<const void *>(k_bt_chain_feed<true, true, 7>),
                             reinterpret_cast<const void *>(k_bt_chain_feed<true, true, 8>),
                             reinterpret_cast<const void *>(k_bt_chain_feed<true, true, BT_TM_SEARCH>),
                             reinterpret_cast<const void *>(k_bt_chain_feed<true, false, BT_TM_SEARCH>),
                             reinterpret_cast<const void *>(k_bt_chain_feed<false, false, BT_TM_SEARCH>)};
        for (const void *fn : fns)
            VG_TRY(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    }
    b->ch_kmin = h[0] <= h[1] ? (int64_t)(h[0] ^ 0x8000000000000000ull) : 0;
    b->ch_E = (int32_t)E;
    b->ch_m = (int32_t)m;
    b->ch_G = (int32_t)G;
    b->ch_rpb = rpb;
    b->ch_lds = lds;
    b->chain = 1;
    b->jd = {nullptr, 0, 0, 0};
    if (b->join.on && b->jtab && Q0.probe_key.type == VGPU_INT64 &&
        vg_jtable_dense(b->jtab, b->join.key_payload, &b->jd))
        return -1;
    return 0;
}

// The active queries' chain: ordered by decreasing selected set.  A larger
// constant selects fewer rows for > and >=, more for < and <=; a NaN
// constant selects every row for >= and <= (first) and none for > and <
// (last).  Uploads the constants in chain order and every query's position.
static int bt_chain_upload(vgpu_batch *b) {
    const int K = b->n_active, m = b->ch_m;
    const PredInfo &P = b->h_qs[0].pi;
    const bool dbl = P.cmp_type == VGPU_DOUBLE;
    const bool asc = P.cmp_op == VGPU_OP_GT || P.cmp_op == VGPU_OP_GE;
    const bool nan_all = P.cmp_op == VGPU_OP_GE || P.cmp_op == VGPU_OP_LE;
    auto f64 = [](uint64_t k) {
        double d;
        memcpy(&d, &k, 8);
        return d;
    };
    auto cls = [&](uint64_t k) { return (dbl && std::isnan(f64(k))) ? (nan_all ? 0 : 2) : 1; };
    std::vector<int> ord((size_t)K);
    for (int q = 0; q < K; q++) ord[q] = q;
    std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) {
        const uint64_t kx = b->h_qs[x].pi.k, ky = b->h_qs[y].pi.k;
        const int cx = cls(kx), cy = cls(ky);
        if (cx != cy) return cx < cy;
        if (cx != 1) return false;
        if (dbl) return asc ? f64(kx) < f64(ky) : f64(kx) > f64(ky);
        return asc ? (int64_t)kx < (int64_t)ky : (int64_t)kx > (int64_t)ky;
    });
    b->h_chain.assign((size_t)m * 12, 0);
    uint64_t *ks = reinterpret_cast<uint64_t *>(b->h_chain.data());
    int32_t *rank = reinterpret_cast<int32_t *>(b->h_chain.data() + (size_t)m * 8);
    for (int pos = 0; pos < K; pos++) {
        ks[pos] = b->h_qs[ord[pos]].pi.k;
        rank[ord[pos]] = pos + 1;
    }
    VG_TRY(hipMemcpyAsync(b->ch_aux.ptr, b->h_chain.data(), b->h_chain.size(), hipMemcpyHostToDevice, vg_stream()));
    return 0;
}

static int batch_plan_body(vgpu_batch *b, const vgpu_query *queries, int n_queries) {
    if (b->planned) {
        vg_set_error("batch: already planned");
        return -1;
    }
    if (n_queries <= 0) {
        vg_set_error("batch: no queries");
        return -1;
    }
    const GbLayout &L = b->L;
    VgTuneScope tune_scope(L.RH + L.RC, 0);
    hipStream_t st = vg_stream();
    std::vector<BtQuery> qs((size_t)n_queries);
    std::vector<vgpu_prog> progs((size_t)n_queries);
    std::vector<BtItem> items;
    int any_vm = 0;
    // rows per work item: enough items for every resident block several
    // times over, as few as that allows (each item initializes and merges
    // an LDS table)
    {
        // sized for the largest query, which a rerun may run alone: about
        // 16 items per CU keep the rounds of resident blocks balanced
        int64_t largest = 0;
        for (int q = 0; q < n_queries; q++) largest = std::max<int64_t>(largest, queries[q].n_rows);
        int64_t want = largest / ((int64_t)vg_cus() * 16);
        int64_t t = BT_TILE_MIN;
        while (t < want && t < BT_TILE_MAX) t <<= 1;
        b->tile = t;
    }
    for (int q = 0; q < n_queries; q++) {
        const vgpu_query &in = queries[q];
        BtQuery &Q = qs[q];
        memset(&Q, 0, sizeof Q);
        if (in.n_rows < 0 || in.n_rows > (int64_t)0xFFFFFFFF * BT_TILE_MIN) {
            vg_set_error("batch: query %d has %lld rows", q, (long long)in.n_rows);
            return -1;
        }
        if (b->join.on) {
            if (in.probe_key.type != b->join_key_type) {
                vg_set_error("batch: query %d probe key type %d differs from the join table's %d", q,
                             in.probe_key.type, b->join_key_type);
                return -1;
            }
            Q.probe_key = in.probe_key;
        } else if (!L.key_none) {
            if (in.key.type != b->key_type) {
                vg_set_error("batch: query %d key type %d differs from declared %d", q, in.key.type, b->key_type);
                return -1;
            }
            Q.key = in.key;
        }
        for (int i = 0; i < L.n_inputs; i++) {
            if (!in.inputs || in.inputs[i].type != L.in_type[i]) {
                vg_set_error("batch: query %d input %d does not have the declared type", q, i);
                return -1;
            }
            Q.ins.c[i] = in.inputs[i];
        }
        for (int i = 0; i < b->n_pred_cols; i++) {
            if (!in.pred_cols) {
                vg_set_error("batch: query %d has no predicate columns", q);
                return -1;
            }
            Q.pcs.c[i] = in.pred_cols[i];
        }
        Q.n_rows = in.n_rows;
        memset(&progs[q], 0, sizeof progs[q]);
        if (b->has_pred) {
            progs[q] = b->prog;
            if (b->param_insn >= 0) memcpy(&progs[q].insn[b->param_insn].imm, &in.param, 8);
            Q.pi = vg_pred_info(&progs[q]);
            Q.vm = !Q.pi.fast;
            any_vm |= Q.vm;
        }
        for (int64_t t = 0; t * b->tile < in.n_rows; t++) items.push_back(BtItem{(uint32_t)q, (uint32_t)t});
    }
    b->n_queries = n_queries;
    b->n_items = (int64_t)items.size();
    // a shared scan when every query reads the same columns
    auto same_col = [](const vgpu_col &a, const vgpu_col &c) { return a.data == c.data && a.validity == c.validity; };
    bool same = true;
    for (int q = 1; q < n_queries && same; q++) {
        const BtQuery &A = qs[0], &C2 = qs[q];
        same = A.n_rows == C2.n_rows && same_col(A.key, C2.key) && same_col(A.probe_key, C2.probe_key);
        for (int i = 0; i < L.n_inputs && same; i++) same = same_col(A.ins.c[i], C2.ins.c[i]);
        for (int i = 0; i < b->n_pred_cols && same; i++) same = same_col(A.pcs.c[i], C2.pcs.c[i]);
    }
    const bool shared = same && n_queries > 1;
    const bool dense_ok = !L.key_none && !L.key_is_double && vg_tune(VG_TUNE_BT_DENSE) > 1;
    // measured on MI355X with 1B-row tables: the per-row work, not HBM,
    // bounds these queries and query blocks drift apart, so shared scans
    // ran slower than one query after another; they stay behind a knob
    b->shared = shared && vg_tune(VG_TUNE_BT_SHARED) > 1;
    b->n_active = n_queries;
    b->n_tiles = shared ? (qs[0].n_rows + BT_STILE - 1) / BT_STILE : 0;
    size_t vm = any_vm ? (vg_vm_lds_bytes(&b->prog, GB_THREADS) + 15) & ~(size_t)15 : 0;
    b->vm_lds = (int32_t)vm;
    if (b->qs.reserve(qs.size() * sizeof(BtQuery)) || b->progs.reserve(progs.size() * sizeof(vgpu_prog)) ||
        b->items.reserve(std::max<size_t>(items.size(), 1) * sizeof(BtItem)) ||
        b->ctl.reserve((size_t)(n_queries + 1) * GB_CTL_WORDS * 8) || b->queue.reserve(BT_CLASSES * 8) ||
        b->qcount.reserve((size_t)n_queries * 8) || b->qoff.reserve((size_t)(n_queries + 1) * 8) ||
        b->qbase.reserve((size_t)(n_queries + 1) * 8))
        return -1;
    VG_TRY(hipMemcpyAsync(b->progs.ptr, progs.data(), progs.size() * sizeof(vgpu_prog), hipMemcpyHostToDevice, st));
    if (!items.empty())
        VG_TRY(hipMemcpyAsync(b->items.ptr, items.data(), items.size() * sizeof(BtItem), hipMemcpyHostToDevice, st));

    // 1. sketch every query's keys (one launch), size the tables; a join
    // batch sketches the group-key payload column of the build side once, a
    // shared scan the one key column once
    std::vector<int64_t> est((size_t)n_queries, 1);
    std::vector<uint64_t> klo((size_t)n_queries, ~0ull), khi((size_t)n_queries, 0);   // ordered key range
    std::vector<BtQuery> sq;      // the columns the sketch reads
    std::vector<BtItem> sitems;
    if (b->join.on) {
        BtQuery P;
        memset(&P, 0, sizeof P);
        P.key = b->join_payload;
        P.n_rows = b->join_rows;
        sq.push_back(P);
    } else if (shared) {
        sq.push_back(qs[0]);
    } else {
        sq = qs;
    }
    for (int q = 0; q < (int)sq.size(); q++)
        for (int64_t t = 0; t * b->tile < sq[q].n_rows; t++) sitems.push_back(BtItem{(uint32_t)q, (uint32_t)t});
    if (!L.key_none && !sitems.empty()) {
        const int nq = (int)sq.size();
        const int64_t ni = (int64_t)sitems.size();
        DevBuf sk, dq, di;
        size_t regs = (size_t)nq * BT_SK_REGS;
        if (sk.reserve(regs * 4 + (size_t)nq * 24) || dq.reserve(sq.size() * sizeof(BtQuery)) ||
            di.reserve(sitems.size() * sizeof(BtItem)))
            return -1;
        uint32_t *d_regs = sk.as<uint32_t>();
        auto *d_nan = reinterpret_cast<unsigned long long *>(d_regs + regs);
        unsigned long long *d_mm = d_nan + nq;
        k_bt_clear_u32<<<vg_blocks((int64_t)regs + 6 * nq, 256), 256, 0, st>>>(d_regs, (int64_t)regs + 6 * nq);
        VG_LAUNCH_CHECK();
        VG_TRY(hipMemcpyAsync(dq.ptr, sq.data(), sq.size() * sizeof(BtQuery), hipMemcpyHostToDevice, st));
        VG_TRY(hipMemcpyAsync(di.ptr, sitems.data(), sitems.size() * sizeof(BtItem), hipMemcpyHostToDevice, st));
        BtArgs A;
        memset(&A, 0, sizeof A);
        A.qs = dq.as<BtQuery>();
        A.items = di.as<BtItem>();
        A.n_items = ni;
        A.tile = b->tile;
        k_bt_sketch<<<vg_blocks(ni, 1, 4096), GB_THREADS, 0, st>>>(A, L.key_none, d_regs, d_nan, d_mm);
        VG_LAUNCH_CHECK();
        std::vector<uint32_t> h(regs);
        std::vector<uint64_t> hn((size_t)nq), hm((size_t)nq * 2);
        VG_TRY(hipMemcpyAsync(h.data(), d_regs, regs * 4, hipMemcpyDeviceToHost, st));
        VG_TRY(hipMemcpyAsync(hn.data(), d_nan, (size_t)nq * 8, hipMemcpyDeviceToHost, st));
        VG_TRY(hipMemcpyAsync(hm.data(), d_mm, (size_t)nq * 16, hipMemcpyDeviceToHost, st));
        VG_TRY(hipStreamSynchronize(st));
        sk.release();
        dq.release();
        di.release();
        for (int q = 0; q < n_queries; q++) {
            const int sq_i = nq == n_queries ? q : 0;
            est[q] = bt_estimate(&h[(size_t)sq_i * BT_SK_REGS], hn[sq_i]);
            klo[q] = ~hm[2 * (size_t)sq_i];
            khi[q] = hm[2 * (size_t)sq_i + 1];
        }
    }
    // LDS tables drop the cold words when no input can touch them
    bool cold = false;
    for (int i = 0; i < L.n_inputs; i++) cold |= L.c_nan[i] >= 0;
    for (int q = 0; q < n_queries && !cold; q++)
        for (int i = 0; i < L.n_inputs; i++) cold |= qs[q].ins.c[i].validity != nullptr;
    b->lds_cold = cold;
    const size_t rw = (size_t)(L.RH + (cold ? L.RC : 0)) * 8;
    std::vector<uint64_t> qbase((size_t)n_queries + 1);
    std::vector<BtItem> cmap;
    int64_t slots = 0;
    int S_max = 0;
    // Tables above BT_SPLIT_SLOTS records run in launches of their own: those
    // that still let 8 blocks of 256 threads share a CU in one, larger ones
    // in another whose blocks are 1,024 threads sharing one table of up to
    // BT_LDS_BIG bytes, so a table of a few thousand groups still has 16
    // waves per CU (not with a VM predicate, whose LDS registers are sized
    // for 256 threads: its large tables stay in 256-thread blocks within
    // BT_LDS_BUDGET). A 1,024-thread block takes four tiles per item: as
    // many rows per thread as the others.
    const bool big_ok = !shared && !vm;
    const size_t budget = big_ok ? BT_LDS_BIG : BT_LDS_BUDGET;
    const int64_t tile_big = b->tile * (BT_BIG_THREADS / GB_THREADS);
    auto big_table = [&](int64_t S) {
        return big_ok && S > BT_SPLIT_SLOTS && (size_t)(S + 2) * rw + vm > BT_MID_BYTES;
    };
    for (int q = 0; q < n_queries; q++) {
        BtQuery &Q = qs[q];
        int64_t bound = b->join.on ? std::max<int64_t>(Q.n_rows, 1) * std::max<int64_t>(b->join_rows, 1)
                                   : std::max<int64_t>(Q.n_rows, 1);
        int64_t g = std::min<int64_t>(bound, est[q] + est[q] / 4 + 16);
        Q.cap = pow2_at_least(2 * g);
        // a dense table when the keys' range is no longer than the hashed
        // table would be (every key the sketch saw is in the range)
        Q.dense = 0;
        Q.kmin = 0;
        if (dense_ok && klo[q] <= khi[q] && khi[q] - klo[q] < (uint64_t)Q.cap) {
            Q.dense = 1;
            Q.kmin = (int64_t)(klo[q] ^ 0x8000000000000000ull);
            Q.cap = (int64_t)(khi[q] - klo[q]) + 1;
        }
        Q.base = slots;
        qbase[q] = (uint64_t)slots;
        slots += Q.cap + 2;
        // LDS privatization when the query's groups fit and its tiles are long enough to amortize
        // (the estimate counts the NA and EMPTY-valued key groups, which
        // have reserved records)
        const int64_t hashed = std::max<int64_t>(est[q] - 2, 1);
        int64_t S = pow2_at_least(vg_tune_for(VG_TUNE_GB_LDS_SLOTS_PER_GROUP, est[q]) * hashed);
        // (a load of up to 0.8 when the tuned load does not fit)
        if ((size_t)(S + 2) * rw + vm > budget) S = pow2_at_least(hashed + hashed / 4 + 1);
        if (L.key_none) S = 64;
        if (Q.dense) {
            // the LDS table is dense too: one record per key of the range, or
            // the records of the range's first keys when at least half of the
            // range fits (rows of the other keys update the HBM table)
            S = Q.cap;
            const int64_t fit = (int64_t)((budget - std::min(budget, vm)) / rw) - 2;
            if (S > fit && 2 * fit >= S) S = fit;
        }
        Q.lds_slots = 0;
        int64_t tile = shared ? (int64_t)BT_STILE * 64 : b->tile;   // rows one LDS table sees
        if (big_table(S)) tile = tile_big;
        if ((size_t)(S + 2) * rw + vm <= (S > BT_SPLIT_SLOTS ? budget : BT_LDS_BUDGET) &&
            S * 8 <= std::min<int64_t>(Q.n_rows, tile)) {
            Q.lds_slots = (int32_t)S;
            S_max = std::max<int>(S_max, (int)S);
        }
        for (int64_t c = 0; c * GB_CMP_SLOTS < Q.cap + 2; c++) cmap.push_back(BtItem{(uint32_t)q, (uint32_t)c});
    }
    if (slots >= ((int64_t)1 << 32)) {
        vg_set_error("batch: %lld table slots exceed the 2^32 the finish addresses", (long long)slots);
        return -1;
    }
    qbase[n_queries] = (uint64_t)slots;
    b->n_slots = slots;
    b->S_max = S_max;
    b->n_cblocks = (int64_t)cmap.size();
    int dev = 0, cus = 0;
    VG_TRY(hipGetDevice(&dev));
    VG_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    {
        // items of the queries with small or no LDS tables, then the rest
        // (query-major within each class, as the blocks' sticky tables want)
        std::vector<BtItem> ordered;
        ordered.reserve(items.size());
        auto cls_of = [&](const BtQuery &Q) {
            if (Q.lds_slots <= BT_SPLIT_SLOTS) return 0;
            return big_table(Q.lds_slots) ? 2 : 1;
        };
        for (int c = 0; c < BT_CLASSES; c++) {
            b->cls_first[c] = (int64_t)ordered.size();
            b->cls_nt[c] = c == 2 ? BT_BIG_THREADS : GB_THREADS;
            b->cls_tile[c] = c == 2 ? tile_big : b->tile;
            b->cls_S[c] = 0;
            for (int q = 0; q < n_queries; q++) {
                if (cls_of(qs[q]) != c) continue;
                b->cls_S[c] = std::max<int>(b->cls_S[c], qs[q].lds_slots);
                for (int64_t t = 0; t * b->cls_tile[c] < qs[q].n_rows; t++)
                    ordered.push_back(BtItem{(uint32_t)q, (uint32_t)t});
            }
        }
        b->cls_first[BT_CLASSES] = (int64_t)ordered.size();
        b->n_items = (int64_t)ordered.size();
        if (!ordered.empty())
            VG_TRY(hipMemcpyAsync(b->items.ptr, ordered.data(), ordered.size() * sizeof(BtItem), hipMemcpyHostToDevice,
                                  st));
        for (int c = 0; c < BT_CLASSES; c++) {
            const size_t lds = (size_t)b->vm_lds + (b->cls_S[c] ? (size_t)(b->cls_S[c] + 2) * rw : 0);
            const int64_t per_cu = std::max<int64_t>(
                1, std::min<int64_t>(BT_BLOCKS_PER_CU * GB_THREADS / b->cls_nt[c], (int64_t)(160 * 1024 / (lds + 512))));
            b->cls_grid[c] = std::max<int64_t>(
                1, std::min<int64_t>(b->cls_first[c + 1] - b->cls_first[c], (int64_t)cus * per_cu));
            if (b->cls_nt[c] == BT_BIG_THREADS && lds > 65536)
                VG_TRY(hipFuncSetAttribute(reinterpret_cast<const void *>(k_bt_feed<BT_BIG_THREADS>),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        }
    }
    {
        // 160 KB of LDS and 32 waves (8 blocks of 256 threads) per CU
        size_t lds = (size_t)b->vm_lds + (S_max ? (size_t)(S_max + 2) * rw : 0) + 512;
        int64_t per_cu = std::min<int64_t>(8, (int64_t)(160 * 1024 / std::max<size_t>(lds, 1)));
        b->shared_resident = std::max<int64_t>(BT_XCDS, (int64_t)cus * std::max<int64_t>(per_cu, 1));
    }
    int qbits = 0;
    while (((int64_t)1 << qbits) < n_queries) qbits++;
    b->qbits = qbits;
    if (b->hot.reserve((size_t)slots * L.RH * 8) || b->cold.reserve((size_t)slots * L.RC * 8) ||
        b->cmap.reserve(cmap.size() * sizeof(BtItem)) || b->bcount.reserve((size_t)b->n_cblocks * 4) ||
        b->boff.reserve((size_t)b->n_cblocks * 8 + 8) || b->por.reserve((size_t)b->n_cblocks * 8) ||
        b->pand.reserve((size_t)b->n_cblocks * 8) || b->skeys.reserve((size_t)slots * 8) ||
        b->svals.reserve((size_t)slots * 4))
        return -1;
    VG_TRY(hipMemcpyAsync(b->qs.ptr, qs.data(), qs.size() * sizeof(BtQuery), hipMemcpyHostToDevice, st));
    VG_TRY(hipMemcpyAsync(b->cmap.ptr, cmap.data(), cmap.size() * sizeof(BtItem), hipMemcpyHostToDevice, st));
    b->h_qs = qs;
    b->h_progs = progs;
    VG_TRY(hipMemcpyAsync(b->qbase.ptr, qbase.data(), qbase.size() * 8, hipMemcpyHostToDevice, st));

    // 2. queries that differ only in a comparison constant share one
    // bucketed aggregation
    if (same && bt_chain_shape(b, qs[0]) && bt_chain_plan(b, qs[0])) return -1;
    if (b->chain && bt_chain_upload(b)) return -1;

    // 3. the first run fixes the sort range and the record count
    if (bt_enqueue(b, true)) return -1;
    b->planned = 1;
    return 0;
}

extern "C" int vgpu_batch_plan(vgpu_batch *b, const vgpu_query *queries, int n_queries) {
    return vg_guard<int>(-1, "vgpu_batch_plan", [&] { return batch_plan_body(b, queries, n_queries); });
}

// one run's device work; `first` (inside plan) synchronizes once to learn
// the sort range and the number of records
static int bt_enqueue(vgpu_batch *b, bool first) {
    const GbLayout &L = b->L;
    hipStream_t st = vg_stream();
    BtArgs A;
    memset(&A, 0, sizeof A);
    A.qs = b->qs.as<BtQuery>();
    A.progs = b->progs.as<vgpu_prog>();
    A.items = b->items.as<BtItem>();
    A.n_items = b->n_items;
    A.queue = b->queue.as<unsigned long long>();
    A.hot = b->hot.as<uint64_t>();
    A.cold = b->cold.as<uint64_t>();
    A.ctl = b->ctl.as<uint64_t>();
    A.vm_lds = b->vm_lds;
    A.n_queries = b->n_queries;
    A.jn = b->join;
    A.K = b->n_active;
    // shared scans: every query's blocks must be resident together so that
    // they read each tile from the XCD's L2 at about the same time
    A.J = std::max<int>(1, (int)(b->shared_resident / ((int64_t)BT_XCDS * std::max(b->n_active, 1))));
    A.n_tiles = b->n_tiles;
    A.lds_cold = b->lds_cold;
    A.n_active = b->n_active;
    A.tile = b->tile;
    uint64_t *bctl = A.ctl + (size_t)b->n_queries * GB_CTL_WORDS;
    uint64_t *qcount = b->qcount.as<uint64_t>();
    const uint64_t empty = L.key_is_double ? VG_EMPTY_F64 : VG_EMPTY_I64;

    // 1. init
    k_gb_init<<<vg_blocks(b->n_slots, 256, 8192), 256, 0, st>>>(A.hot, A.cold, b->n_slots, L.RH, L.RC,
                                                                 L.hot_init_ones, L.cold_init_ones, empty);
    VG_LAUNCH_CHECK();
    int64_t ctl_words = (int64_t)(b->n_queries + 1) * GB_CTL_WORDS;
    k_bt_init_ctl<<<vg_blocks(std::max<int64_t>(ctl_words, b->n_queries), 256), 256, 0, st>>>(
        A.ctl, b->n_queries, qcount, A.queue);
    VG_LAUNCH_CHECK();
    // 2. one persistent feed launch, or the bucketed pass
    if (b->chain && b->n_items > 0) {
        const BtQuery &Q0 = b->h_qs[0];
        BtChainArgs C;
        memset(&C, 0, sizeof C);
        C.key = Q0.key;
        C.probe_key = Q0.probe_key;
        C.pcol = Q0.pcs.c[Q0.pi.col];
        C.ins = Q0.ins;
        C.jn = b->join;
        C.pi = Q0.pi;
        C.ks = b->ch_aux.as<uint64_t>();
        C.rank = reinterpret_cast<const int32_t *>(b->ch_aux.as<uint8_t>() + (size_t)b->ch_m * 8);
        C.slab = b->ch_slab.as<uint64_t>();
        C.tab = b->ch_tab.as<uint64_t>();
        C.bctl = bctl;
        C.n_rows = Q0.n_rows;
        C.rows_per_block = b->ch_rpb;
        C.kmin = b->ch_kmin;
        C.E = b->ch_E;
        C.m = b->ch_m;
        C.K = b->n_active;
        C.n_sum = b->ch_nsum;
        C.key_type = b->ch_key_type;
        C.G = b->ch_G;
        C.f64_mask = b->ch_f64;
        C.jd = b->jd;
        for (int s2 = 0; s2 < BT_CH_MAX_SUMS; s2++) {
            C.sum_in[s2] = b->ch_sum_in[s2];
            C.sum_pred[s2] = b->ch_sum_pred[s2];
        }
        const int64_t W = (int64_t)C.E * C.m * (1 + C.n_sum);
        vg_kt_begin();
        bool sp = true;
        for (int s2 = 0; s2 < C.n_sum; s2++) sp = sp && C.sum_pred[s2];
        const vgpu_col &kc = C.jn.on ? C.probe_key : C.key;
        const bool plain = sp && !C.pcol.validity && C.pcol.type != VGPU_BOOL && !kc.validity &&
                           kc.type != VGPU_BOOL;
        const int op = C.pi.cmp_op;
        const bool same_type = (C.pi.cmp_type == VGPU_DOUBLE) == (C.pcol.type == VGPU_DOUBLE);
        const int tm0 = bt_tm(C.pi.cmp_type == VGPU_DOUBLE, op == VGPU_OP_GT || op == VGPU_OP_LE);
        const int tm = !(plain && same_type && C.n_sum == 1) ? BT_TM_SEARCH
                       : C.K <= BT_CH_LIN                    ? tm0
                       : C.K <= 127                          ? tm0 + BT_TM_CSEARCH
                                                             : BT_TM_SEARCH;
        const int neg = op == VGPU_OP_GE || op == VGPU_OP_LE;
        const dim3 g((unsigned)C.G), t(BT_CH_THREADS);
        const size_t lds = b->ch_lds;
        // the count over 8 compare slots for up to 8 queries, 16 for up to 16
        switch (plain ? (tm != BT_TM_SEARCH && tm <= BT_TM_CSEARCH && C.K > 8 ? tm + 16 : tm) : -1) {
        case 1: k_bt_chain_feed<true, true, 1><<<g, t, lds, st>>>(C, neg); break;
        case 2: k_bt_chain_feed<true, true, 2><<<g, t, lds, st>>>(C, neg); break;
        case 3: k_bt_chain_feed<true, true, 3><<<g, t, lds, st>>>(C, neg); break;
        case 4: k_bt_chain_feed<true, true, 4><<<g, t, lds, st>>>(C, neg); break;
        case 17: k_bt_chain_feed<true, true, 1, 16><<<g, t, lds, st>>>(C, neg); break;
        case 18: k_bt_chain_feed<true, true, 2, 16><<<g, t, lds, st>>>(C, neg); break;
        case 19: k_bt_chain_feed<true, true, 3, 16><<<g, t, lds, st>>>(C, neg); break;
        case 20: k_bt_chain_feed<true, true, 4, 16><<<g, t, lds, st>>>(C, neg); break;
        case 5: k_bt_chain_feed<true, true, 5><<<g, t, lds, st>>>(C, neg); break;
        case 6: k_bt_chain_feed<true, true, 6><<<g, t, lds, st>>>(C, neg); break;
        case 7: k_bt_chain_feed<true, true, 7><<<g, t, lds, st>>>(C, neg); break;
        case 8: k_bt_chain_feed<true, true, 8><<<g, t, lds, st>>>(C, neg); break;
        case BT_TM_SEARCH: k_bt_chain_feed<true, true, BT_TM_SEARCH><<<g, t, lds, st>>>(C, neg); break;
        default:
            if (sp) k_bt_chain_feed<true, false, BT_TM_SEARCH><<<g, t, lds, st>>>(C, neg);
            else k_bt_chain_feed<false, false, BT_TM_SEARCH><<<g, t, lds, st>>>(C, neg);
        }
        VG_LAUNCH_CHECK();
        vg_kt_end();   // the scan kernel alone; the slab reduce and expand below are small
        k_bt_chain_reduce<<<vg_blocks(W, 256), 256, 0, st>>>(C);
        VG_LAUNCH_CHECK();
        k_bt_chain_expand<<<vg_blocks((int64_t)C.K * C.E, 256), 256, 0, st>>>(L, A, C);
        VG_LAUNCH_CHECK();
    } else if (b->n_items > 0) {
        const size_t rwb = (size_t)(L.RH + (b->lds_cold ? L.RC : 0)) * 8;
        vg_kt_begin();
        if (b->shared) {
            const size_t lds = (size_t)b->vm_lds + (b->S_max ? (size_t)(b->S_max + 2) * rwb : 0);
            k_bt_feed_shared<<<(unsigned)((int64_t)BT_XCDS * A.K * A.J), GB_THREADS, lds, st>>>(L, A);
            VG_LAUNCH_CHECK();
        } else {
            // small-table and HBM-table queries, then mid-size and large LDS tables
            for (int c = 0; c < BT_CLASSES; c++) {
                const int64_t n = b->cls_first[c + 1] - b->cls_first[c];
                if (n <= 0) continue;
                BtArgs Ac = A;
                Ac.items = A.items + b->cls_first[c];
                Ac.n_items = n;
                Ac.tile = b->cls_tile[c];
                Ac.queue = A.queue + c;
                const size_t lds = (size_t)b->vm_lds + (b->cls_S[c] ? (size_t)(b->cls_S[c] + 2) * rwb : 0);
                if (b->cls_nt[c] == BT_BIG_THREADS)
                    k_bt_feed<BT_BIG_THREADS><<<(unsigned)b->cls_grid[c], BT_BIG_THREADS, lds, st>>>(L, Ac);
                else
                    k_bt_feed<GB_THREADS><<<(unsigned)b->cls_grid[c], GB_THREADS, lds, st>>>(L, Ac);
                VG_LAUNCH_CHECK();
            }
        }
        vg_kt_end();
    }
    // 3. finish: compaction into (key, arena slot) pairs
    uint32_t *bcount = b->bcount.as<uint32_t>();
    uint64_t *boff = b->boff.as<uint64_t>();
    uint64_t *total = boff + b->n_cblocks;
    k_bt_cmp_count<<<(unsigned)b->n_cblocks, GB_THREADS, 0, st>>>(L, A, b->cmap.as<BtItem>(), b->order, bcount,
                                                                  qcount);
    VG_LAUNCH_CHECK();
    if (vg_scan_u32_to_u64(bcount, boff, b->n_cblocks, total, b->scan_tmp)) return -1;
    k_bt_cmp_write<<<(unsigned)b->n_cblocks, GB_THREADS, 0, st>>>(L, A, b->cmap.as<BtItem>(), b->order, boff,
                                                                  b->skeys.as<uint64_t>(), b->svals.as<uint32_t>(),
                                                                  b->por.as<uint64_t>(), b->pand.as<uint64_t>());
    VG_LAUNCH_CHECK();
    k_gb_or_and_final<<<vg_blocks(b->n_cblocks, 4 * GB_THREADS, 256), GB_THREADS, 0, st>>>(
        b->por.as<uint64_t>(), b->pand.as<uint64_t>(), b->n_cblocks, bctl);
    VG_LAUNCH_CHECK();
    if (first) {
        uint64_t h[GB_CTL_WORDS], n_pairs = 0;
        VG_TRY(hipMemcpyAsync(h, bctl, sizeof h, hipMemcpyDeviceToHost, st));
        VG_TRY(hipMemcpyAsync(&n_pairs, total, 8, hipMemcpyDeviceToHost, st));
        VG_TRY(hipStreamSynchronize(st));
        uint64_t diff = h[6] ^ h[7];
        b->n_pairs = (int64_t)n_pairs;
        b->lo = diff ? __builtin_ctzll(diff) : 0;
        b->hi = diff ? 64 - __builtin_clzll(diff) : 0;
        b->range_mask = 0;
        for (int i = b->lo; i < b->hi; i++) b->range_mask |= 1ull << i;
        size_t vwords = (size_t)((b->n_pairs + 63) / 64 + 1);
        if (b->out_keys.reserve((size_t)(b->n_pairs + 1) * 8) || b->out_kvalid.reserve(vwords * 8)) return -1;
        for (int a = 0; a < b->aggs.n; a++)
            if (b->out_vals[a].reserve((size_t)(b->n_pairs + 1) * 8) || b->out_valid[a].reserve(vwords * 8))
                return -1;
    }
    k_bt_guard<<<1, 1, 0, st>>>(bctl, b->range_mask, total, (uint64_t)b->n_pairs);
    VG_LAUNCH_CHECK();
    int64_t n = b->n_pairs;
    uint64_t *skeys = b->skeys.as<uint64_t>();
    uint32_t *svals = b->svals.as<uint32_t>();
    // stable sort by the order key, then by query
    if (n > 1 && b->hi > b->lo && vg_sort_pairs(skeys, svals, n, b->lo, b->hi, b->sort_tmp)) return -1;
    if (n > 0) {
        k_bt_qkeys<<<vg_blocks(n, 256), 256, 0, st>>>(svals, n, b->qbase.as<uint64_t>(), b->n_queries, skeys);
        VG_LAUNCH_CHECK();
    }
    if (n > 1 && b->qbits > 0 && vg_sort_pairs(skeys, svals, n, 0, b->qbits, b->sort_tmp)) return -1;
    // per-query output offsets
    if (vg_scan_u64(qcount, b->qoff.as<uint64_t>(), b->n_queries, b->qoff.as<uint64_t>() + b->n_queries,
                    b->scan_tmp))
        return -1;
    // emit every group of every query
    if (n > 0) {
        EmitOut O;
        memset(&O, 0, sizeof O);
        O.keys = b->out_keys.ptr;
        O.kvalid = b->out_kvalid.as<uint64_t>();
        for (int a = 0; a < b->aggs.n; a++) {
            O.vals[a] = b->out_vals[a].as<double>();
            O.valid[a] = b->out_valid[a].as<uint64_t>();
        }
        const int RW = L.RH + L.RC;
        if (RW <= GB_EMIT_STAGE_WORDS)
            k_bt_emit<true><<<vg_blocks(n, GB_THREADS), GB_THREADS, (size_t)GB_THREADS * RW * 8, st>>>(
                L, b->aggs, A, svals, skeys, n, O);
        else
            k_bt_emit<false><<<vg_blocks(n, GB_THREADS), GB_THREADS, 0, st>>>(L, b->aggs, A, svals, skeys, n, O);
        VG_LAUNCH_CHECK();
    }
    b->counts_valid = 0;
    return 0;
}

static int batch_run_body(vgpu_batch *b) {
    if (!b->planned) {
        vg_set_error("batch: run before plan");
        return -1;
    }
    if (b->exec) {
        VG_TRY(hipGraphLaunch(b->exec, vg_stream()));
        b->counts_valid = 0;
        return 0;
    }
    return bt_enqueue(b, false);
}

extern "C" int vgpu_batch_run(vgpu_batch *b) {
    return vg_guard<int>(-1, "vgpu_batch_run", [&] { return batch_run_body(b); });
}

static int batch_rerun_body(vgpu_batch *b, const uint64_t *params, int n_active) {
    if (!b->planned) {
        vg_set_error("batch: rerun before plan");
        return -1;
    }
    if (n_active < 1 || n_active > b->n_queries) {
        vg_set_error("batch: %d active queries of %d", n_active, b->n_queries);
        return -1;
    }
    // a rerun re-derives the sort range and the output sizes: replays of
    // an earlier capture no longer apply
    if (b->exec) {
        (void)hipGraphExecDestroy(b->exec);
        (void)hipGraphDestroy(b->graph);
        b->exec = nullptr;
        b->graph = nullptr;
    }
    if (params && b->has_pred && b->param_insn >= 0)
        for (int q = 0; q < n_active; q++) {
            memcpy(&b->h_progs[q].insn[b->param_insn].imm, &params[q], 8);
            b->h_qs[q].pi = vg_pred_info(&b->h_progs[q]);
        }
    hipStream_t st = vg_stream();
    VG_TRY(hipMemcpyAsync(b->qs.ptr, b->h_qs.data(), b->h_qs.size() * sizeof(BtQuery), hipMemcpyHostToDevice, st));
    VG_TRY(hipMemcpyAsync(b->progs.ptr, b->h_progs.data(), b->h_progs.size() * sizeof(vgpu_prog),
                          hipMemcpyHostToDevice, st));
    b->n_active = n_active;
    if (b->chain && bt_chain_upload(b)) return -1;
    return bt_enqueue(b, true);
}

extern "C" int vgpu_batch_rerun(vgpu_batch *b, const uint64_t *params, int n_active) {
    return vg_guard<int>(-1, "vgpu_batch_rerun", [&] { return batch_rerun_body(b, params, n_active); });
}

static int batch_capture_body(vgpu_batch *b) {
    if (!b->planned) {
        vg_set_error("batch: capture before plan");
        return -1;
    }
    if (b->exec) return 0;
    hipStream_t st = vg_stream();
    VG_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    int rc = bt_enqueue(b, false);
    hipGraph_t graph = nullptr;
    hipError_t e = hipStreamEndCapture(st, &graph);
    if (rc) {
        if (graph) (void)hipGraphDestroy(graph);
        return -1;
    }
    if (e != hipSuccess) {
        vg_set_error("batch: capture failed: %s", hipGetErrorString(e));
        return -1;
    }
    e = hipGraphInstantiate(&b->exec, graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        (void)hipGraphDestroy(graph);
        b->exec = nullptr;
        vg_set_error("batch: graph instantiation failed: %s", hipGetErrorString(e));
        return -1;
    }
    b->graph = graph;
    return 0;
}

extern "C" int vgpu_batch_capture(vgpu_batch *b) {
    return vg_guard<int>(-1, "vgpu_batch_capture", [&] { return batch_capture_body(b); });
}

static int bt_read_counts(vgpu_batch *b) {
    if (b->counts_valid) return 0;
    hipStream_t st = vg_stream();
    int Q = b->n_queries;
    std::vector<uint64_t> cnt((size_t)Q), off((size_t)Q + 1);
    uint64_t bctl[GB_CTL_WORDS];
    VG_TRY(hipMemcpyAsync(cnt.data(), b->qcount.ptr, (size_t)Q * 8, hipMemcpyDeviceToHost, st));
    VG_TRY(hipMemcpyAsync(off.data(), b->qoff.ptr, (size_t)(Q + 1) * 8, hipMemcpyDeviceToHost, st));
    VG_TRY(hipMemcpyAsync(bctl, b->ctl.as<uint64_t>() + (size_t)Q * GB_CTL_WORDS, sizeof bctl,
                          hipMemcpyDeviceToHost, st));
    VG_TRY(hipStreamSynchronize(st));
    if (bctl[BT_CTL_ERR]) {
        vg_set_error("batch: the inputs changed since plan (keys outside the planned sort range or a "
                     "different number of groups); plan a new batch");
        return -1;
    }
    b->h_count.assign(cnt.begin(), cnt.end());
    b->h_off.assign(off.begin(), off.end() - 1);
    b->counts_valid = 1;
    return 0;
}

static int batch_counts_body(vgpu_batch *b, int64_t *n_groups, int64_t *offsets) {
    if (!b->planned) {
        vg_set_error("batch: counts before plan");
        return -1;
    }
    if (bt_read_counts(b)) return -1;
    for (int q = 0; q < b->n_queries; q++) {
        if (n_groups) n_groups[q] = b->h_count[q];
        if (offsets) offsets[q] = b->h_off[q];
    }
    return 0;
}

extern "C" int vgpu_batch_counts(vgpu_batch *b, int64_t *n_groups, int64_t *offsets) {
    return vg_guard<int>(-1, "vgpu_batch_counts", [&] { return batch_counts_body(b, n_groups, offsets); });
}

// copies bits [off, off + n) of a device bitmap into dst (bit 0 = first)
static int bt_copy_bits(uint8_t *dst, const uint8_t *dev_bits, int64_t off, int64_t n) {
    if (n <= 0) return 0;
    int64_t b0 = off / 8, b1 = (off + n + 7) / 8;
    std::vector<uint8_t> tmp((size_t)(b1 - b0 + 1), 0);
    VG_TRY(hipMemcpyAsync(tmp.data(), dev_bits + b0, (size_t)(b1 - b0), hipMemcpyDeviceToHost, vg_stream()));
    VG_TRY(hipStreamSynchronize(vg_stream()));
    int sh = (int)(off & 7);
    memset(dst, 0, (size_t)((n + 7) / 8));
    for (int64_t i = 0; i < n; i++) {
        int64_t src = i + sh;
        if ((tmp[(size_t)(src >> 3)] >> (src & 7)) & 1) dst[i >> 3] |= (uint8_t)(1u << (i & 7));
    }
    return 0;
}

static int64_t batch_fetch_body(vgpu_batch *b, int q, void *keys, uint8_t *key_validity, double *const *agg_values,
                                uint8_t *const *agg_validity) {
    if (!b->planned) {
        vg_set_error("batch: fetch before plan");
        return -1;
    }
    if (q < 0 || q >= b->n_queries) {
        vg_set_error("batch: query %d of %d", q, b->n_queries);
        return -1;
    }
    if (bt_read_counts(b)) return -1;
    hipStream_t st = vg_stream();
    int64_t n = b->h_count[q], off = b->h_off[q];
    if (n > 0) {
        if (keys && !b->L.key_none)
            VG_TRY(hipMemcpyAsync(keys, b->out_keys.as<uint64_t>() + off, (size_t)n * 8, hipMemcpyDeviceToHost, st));
        for (int a = 0; a < b->aggs.n; a++)
            if (agg_values && agg_values[a])
                VG_TRY(hipMemcpyAsync(agg_values[a], b->out_vals[a].as<double>() + off, (size_t)n * 8,
                                      hipMemcpyDeviceToHost, st));
        VG_TRY(hipStreamSynchronize(st));
        if (key_validity && bt_copy_bits(key_validity, b->out_kvalid.as<uint8_t>(), off, n)) return -1;
        for (int a = 0; a < b->aggs.n; a++)
            if (agg_validity && agg_validity[a] && bt_copy_bits(agg_validity[a], b->out_valid[a].as<uint8_t>(), off, n))
                return -1;
    }
    uint64_t qctl[GB_CTL_WORDS];
    VG_TRY(hipMemcpyAsync(qctl, b->ctl.as<uint64_t>() + (size_t)q * GB_CTL_WORDS, sizeof qctl, hipMemcpyDeviceToHost,
                          st));
    VG_TRY(hipStreamSynchronize(st));
    return (int64_t)qctl[4];
}

extern "C" int64_t vgpu_batch_fetch(vgpu_batch *b, int q, void *keys, uint8_t *key_validity, double *const *agg_values,
                                    uint8_t *const *agg_validity) {
    return vg_guard<int64_t>(-1, "vgpu_batch_fetch", [&] { return batch_fetch_body(b, q, keys, key_validity, agg_values,
        agg_validity); });
}

extern "C" int64_t vgpu_batch_device_bytes(const vgpu_batch *b) {
    const DevBuf *bufs[] = {&b->qs, &b->progs, &b->items, &b->cmap, &b->queue, &b->hot, &b->cold, &b->ctl,
                            &b->qbase, &b->qcount, &b->qoff, &b->bcount, &b->boff, &b->por, &b->pand, &b->skeys,
                            &b->svals, &b->sort_tmp, &b->scan_tmp, &b->out_keys, &b->out_kvalid,
                            &b->ch_aux, &b->ch_slab, &b->ch_tab};
    int64_t total = 0;
    for (const DevBuf *p : bufs) total += (int64_t)p->cap;
    for (auto &x : b->out_vals) total += (int64_t)x.cap;
    for (auto &x : b->out_valid) total += (int64_t)x.cap;
    return total;
}

extern "C" int vgpu_batch_bucketed(const vgpu_batch *b) { return b && b->planned && b->chain ? 1 : 0; }

extern "C" void vgpu_batch_destroy(vgpu_batch *b) {
    if (!b) return;
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    // graph replays may still run: the buffers go back to the pool only
    // after the stream drains
    (void)hipStreamSynchronize(vg_stream());
    DevBuf *bufs[] = {&b->qs, &b->progs, &b->items, &b->cmap, &b->queue, &b->hot, &b->cold, &b->ctl, &b->qbase,
                      &b->qcount, &b->qoff, &b->bcount, &b->boff, &b->por, &b->pand, &b->skeys, &b->svals,
                      &b->sort_tmp, &b->scan_tmp, &b->out_keys, &b->out_kvalid};
    for (DevBuf *p : bufs) p->release();
    for (auto &x : b->out_vals) x.release();
    for (auto &x : b->out_valid) x.release();
    delete b;
}
