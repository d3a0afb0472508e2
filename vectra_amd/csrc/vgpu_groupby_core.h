// vgpu_groupby_core.h -- the group-by record model shared by the single
// group-by (vgpu_groupby.hip) and the batched queries (vgpu_batch.hip):
// record layout, atomic record updates, table slots, LDS-private tables,
// and the finish kernels (compaction into sort pairs, emit).
//
// A group is a record of 64-bit words: a "hot" part (key, n, first
// position, per input sum / min / max) and a "cold" part (per input NA
// count and the NaN / signed-zero positions only NA, NaN and zero values
// touch).  Reference semantics: src/hash.c:12-78 (key equality),
// src/agg_ops.c:93-255 (accumulators and their finish).
#pragma once

#include "vgpu_expr.h"
#include "vgpu_keys.h"

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

// merge a whole record (LDS partial -> HBM); sc == nullptr for a partial
// without cold words, kept only when no input can touch them (no input has
// NAs and none has a double min / max)
__device__ __forceinline__ void gb_merge(const GbLayout &L, uint64_t *dh, uint64_t *dc,
                                         const uint64_t *sh, const uint64_t *sc) {
    gb_atomic_add(&dh[1], sh[1]);
    if (L.w_fpos >= 0) gb_atomic_min(&dh[L.w_fpos], sh[L.w_fpos]);
    for (int i = 0; i < L.n_inputs; i++) {
        if (sc && sc[L.c_na[i]]) gb_atomic_add(&dc[L.c_na[i]], sc[L.c_na[i]]);
        if (L.w_fvpos[i] >= 0) gb_atomic_min(&dh[L.w_fvpos[i]], sh[L.w_fvpos[i]]);
        if (L.w_sum[i] >= 0) {
            if (L.in_type[i] == VGPU_DOUBLE) gb_atomic_addf(&dh[L.w_sum[i]], vg_bits_to_f64(sh[L.w_sum[i]]));
            else gb_atomic_add(&dh[L.w_sum[i]], sh[L.w_sum[i]]);
        }
        if (L.w_min[i] >= 0) gb_atomic_min(&dh[L.w_min[i]], sh[L.w_min[i]]);
        if (L.w_max[i] >= 0) gb_atomic_max(&dh[L.w_max[i]], sh[L.w_max[i]]);
        if (sc && L.c_nan[i] >= 0) {
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

// The arrays are written as flat word streams (consecutive lanes store
// consecutive words), each word's value set by its index within a record.
__device__ __forceinline__ void gb_init_words(uint64_t *a, int64_t n_words, int R, uint64_t ones, bool key_word0,
                                              uint64_t key) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int w = (int)(i % R);
    const int step = (int)(stride % R);
    for (; i < n_words; i += stride) {
        a[i] = (key_word0 && w == 0) ? key : (((ones >> w) & 1) ? ~0ull : 0ull);
        w += step;
        if (w >= R) w -= R;
    }
}

static __global__ void k_gb_init(uint64_t *hot, uint64_t *cold, int64_t n_slots, int RH, int RC,
                          uint64_t hot_ones, uint64_t cold_ones, uint64_t empty) {
    gb_init_words(hot, n_slots * RH, RH, hot_ones, true, empty);
    gb_init_words(cold, n_slots * RC, RC, cold_ones, false, 0);
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
    int32_t lds_cold;      // LDS records carry cold words (an input has NAs or a double min / max)
    int32_t lds_copies;    // copies of the LDS table, lane l updating copy l % lds_copies (1, 2, 4, 8)
    uint64_t *part;        // block partials of the LDS table (nullptr: merge into HBM directly)
    int64_t kmin;          // a dense LDS table's record 0 key (DENSE feeds)
};

// cold words can be touched only by NA inputs and by double min / max
static inline bool gb_needs_cold(const GbLayout &L, const vgpu_col *inputs) {
    for (int i = 0; i < L.n_inputs; i++)
        if (L.c_nan[i] >= 0 || (inputs && inputs[i].validity)) return true;
    return false;
}

// The LDS-private table holds S hashed records plus two reserved ones (NA
// key, EMPTY-valued int64 key), like the HBM table's slots cap / cap+1, so
// that no row class funnels every block into one HBM record.
//
// A dense table (dense = true) serves int64 / bool keys of a short range:
// key k owns record k - kmin, its key word preset, so a row finds its
// record without hashing or probing; keys outside the range take the HBM
// table.  The merge skips the records no row touched (n == 0).
template <int NT = GB_THREADS>
__device__ __forceinline__ void gb_lds_init(const GbLayout &L, uint64_t *lhot, uint64_t *lcold, int S,
                                            uint64_t empty, bool dense = false, int64_t kmin = 0) {
    for (int s = threadIdx.x; s < S + 2; s += NT) {
        uint64_t *h = lhot + (size_t)s * L.RH;
        h[0] = (dense && s < S) ? (uint64_t)kmin + (uint64_t)s : empty;
        for (int w = 1; w < L.RH; w++) h[w] = ((L.hot_init_ones >> w) & 1) ? ~0ull : 0ull;
        if (!lcold) continue;   // a table without cold words (see gb_lds_merge)
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

// a regular key's LDS record into the HBM table
__device__ __forceinline__ void gb_merge_global(const GbLayout &L, const uint64_t *h, const uint64_t *c, uint64_t *hot,
                                                uint64_t *cold, int64_t cap, uint64_t *ctl) {
    const int64_t gs = gb_global_slot(L, hot, cap, VG_KEY_REGULAR, h[0], 0, ctl);
    if (gs >= 0) gb_merge(L, hot + (size_t)gs * L.RH, cold + (size_t)gs * L.RC, h, c);
}

// record of a key in a dense table; -1: outside the range (or a NaN key)
__device__ __forceinline__ int64_t gb_lds_dense_slot(int S, int cls, uint64_t kn, int64_t kmin) {
    if (cls == VG_KEY_NA) return S;
    if (cls == VG_KEY_SPECIAL) return S + 1;
    if (cls == VG_KEY_NAN) return -1;
    const uint64_t d = kn - (uint64_t)kmin;
    return d < (uint64_t)S ? (int64_t)d : -1;
}

__device__ __forceinline__ void gb_lds_merge(const GbLayout &L, const uint64_t *lhot, const uint64_t *lcold, int S,
                                             uint64_t empty, uint64_t *hot, uint64_t *cold, int64_t cap,
                                             uint64_t *ctl) {
    __syncthreads();
    for (int s = threadIdx.x; s < S + 2; s += GB_THREADS) {
        const uint64_t *h = lhot + (size_t)s * L.RH;
        int64_t gs;
        if (s < S) {
            if (h[0] == empty || h[1] == 0) continue;
            gs = gb_global_slot(L, hot, cap, VG_KEY_REGULAR, h[0], 0, ctl);
        } else {
            if (h[1] == 0) continue;
            gs = cap + (s - S);
        }
        if (gs >= 0)
            gb_merge(L, hot + (size_t)gs * L.RH, cold + (size_t)gs * L.RC, h, lcold ? lcold + (size_t)s * L.RC : nullptr);
    }
}

// LDS tables merge through per-block partials instead of HBM atomics when
// the host gives them room: every block stores its S + 2 records (RH + RC
// words each; a table without cold words stores their initial values) at
// part + blockIdx.x * (S + 2) * (RH + RC), and k_gb_part_fold folds them,
// so that a record's HBM words take a few merges instead of one
// device-scope atomic per block, which the memory side serializes.
__device__ __forceinline__ void gb_lds_store_partial(const GbLayout &L, const uint64_t *lhot, const uint64_t *lcold,
                                                     int S, uint64_t *part) {
    const int RW = L.RH + L.RC;
    const int words = (S + 2) * RW;
    uint64_t *dst = part + (size_t)blockIdx.x * words;
    for (int q = threadIdx.x; q < words; q += GB_THREADS) {
        const int s = q / RW, w = q - s * RW;
        dst[q] = w < L.RH ? lhot[(size_t)s * L.RH + w]
                 : lcold  ? lcold[(size_t)s * L.RC + (w - L.RH)]
                          : (((L.cold_init_ones >> (w - L.RH)) & 1) ? ~0ull : 0ull);
    }
}

// the end of an LDS-table feed block
__device__ __forceinline__ void gb_lds_finish(const GbLayout &L, const uint64_t *lhot, const uint64_t *lcold, int S,
                                              uint64_t empty, uint64_t *hot, uint64_t *cold, int64_t cap,
                                              uint64_t *ctl, uint64_t *part) {
    if (part) {
        __syncthreads();
        gb_lds_store_partial(L, lhot, lcold, S, part);
    } else {
        gb_lds_merge(L, lhot, lcold, S, empty, hot, cold, cap, ctl);
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

#define GB_CTL_WORDS 16

// control block: [0] claimed slots, [1] overflow flag, [2]/[3] first
// position of a -0 / +0 double key, [4] min/max warnings, [5] compaction
// counter, [6]/[7] OR / AND of the sort keys, [8] rows of the NA group
static __global__ void k_gb_ctl_init(uint64_t *ctl, int finish_only) {
    int i = threadIdx.x;
    if (i >= GB_CTL_WORDS) return;
    if (finish_only) {
        if (i == 5 || i == 6 || i == 8) ctl[i] = 0;
        if (i == 7) ctl[i] = ~0ull;
        return;
    }
    ctl[i] = (i == 2 || i == 3 || i == 7) ? ~0ull : 0ull;
}

// ---------------------------------------------------------------------------
// finish: compact, order, emit
// ---------------------------------------------------------------------------

// Where finish reads the group records from: the hash table (hot and cold
// arrays, reserved NA / EMPTY-key records at slots cap and cap+1) or the
// records of a partitioned feed (hot and cold words of a record side by
// side, reserved records at indices n and n+1).
struct RecSrc {
    const uint64_t *hot;
    const uint64_t *cold;
    int64_t hstride, cstride;
    int64_t na_slot, sp_slot;
};

// Sort key of record s in the finish order; false when the record takes no
// part in the sort (an empty slot, an unused reserved record, or the NA
// group of the key order, which is appended after the sort).
__device__ __forceinline__ bool gb_sort_entry(const GbLayout &L, const RecSrc &R, int order, int64_t s,
                                              uint64_t &sk) {
    const uint64_t *h = R.hot + s * R.hstride;
    if (s == R.na_slot || s == R.sp_slot) {
        if (h[1] == 0) return false;
    } else if (h[0] == (L.key_is_double ? VG_EMPTY_F64 : VG_EMPTY_I64)) {
        return false;
    }
    if (s == R.na_slot && order == VGPU_ORDER_KEY) return false;
    if (order == VGPU_ORDER_FIRST_SEEN) sk = h[L.w_fpos];
    else if (s == R.sp_slot) sk = 0;                          // INT64_MIN: smallest
    else if (L.key_is_double) sk = vg_f64_ordered(h[0]);      // NaN groups sort above +Inf by position
    else sk = vg_i64_ordered(h[0]);
    return true;
}

// exclusive prefix of v over the block; *total = the block's sum
__device__ __forceinline__ uint32_t gb_block_prefix(uint32_t v, uint32_t *lds, uint32_t &total) {
    const int lane = vg_lane(), wid = threadIdx.x >> 6;
    uint32_t x = v;
    for (int d = 1; d < VG_WAVE; d <<= 1) {
        uint32_t y = __shfl_up(x, d, VG_WAVE);
        if (lane >= d) x += y;
    }
    if (lane == VG_WAVE - 1) lds[wid] = x;
    __syncthreads();
    uint32_t base = 0, tot = 0;
    for (int w = 0; w < GB_THREADS / VG_WAVE; w++) {
        uint32_t t = lds[w];
        if (w < wid) base += t;
        tot += t;
    }
    __syncthreads();
    total = tot;
    return base + x - v;
}

// OR and AND of the block's values, valid in thread 0
__device__ __forceinline__ void gb_block_or_and(uint64_t &o, uint64_t &a, uint64_t *lds) {
    for (int d = VG_WAVE / 2; d >= 1; d >>= 1) {
        o |= __shfl_xor(o, d, VG_WAVE);
        a &= __shfl_xor(a, d, VG_WAVE);
    }
    const int wid = threadIdx.x >> 6;
    if (vg_lane() == 0) {
        lds[2 * wid] = o;
        lds[2 * wid + 1] = a;
    }
    __syncthreads();
    if (threadIdx.x == 0)
        for (int w = 1; w < GB_THREADS / VG_WAVE; w++) {
            o |= lds[2 * w];
            a &= lds[2 * w + 1];
        }
    __syncthreads();
}

// Table compaction in two passes over blocks of GB_CMP_SLOTS slots: count
// the sorted records per block, scan, then write (sort key, slot) pairs at
// the scanned offsets with per-block OR / AND partials of the sort keys
// (no contended atomics: one word per block).
#define GB_CMP_PER_THREAD 16
#define GB_CMP_SLOTS (GB_THREADS * GB_CMP_PER_THREAD)

static __global__ void __launch_bounds__(GB_THREADS)
k_gb_cmp_count(GbLayout L, RecSrc R, int64_t n_slots, int order, uint32_t *bcount) {
    __shared__ uint32_t lds[GB_THREADS / VG_WAVE];
    const int64_t base = (int64_t)blockIdx.x * GB_CMP_SLOTS;
    uint32_t c = 0;
#pragma unroll
    for (int j = 0; j < GB_CMP_PER_THREAD; j++) {
        int64_t s = base + j * GB_THREADS + threadIdx.x;
        uint64_t sk;
        if (s < n_slots && gb_sort_entry(L, R, order, s, sk)) c++;
    }
    uint32_t total;
    gb_block_prefix(c, lds, total);
    if (threadIdx.x == 0) bcount[blockIdx.x] = total;
}

static __global__ void __launch_bounds__(GB_THREADS)
k_gb_cmp_write(GbLayout L, RecSrc R, int64_t n_slots, int order, const uint64_t *boff, uint64_t *skeys,
               uint32_t *svals, uint64_t *por, uint64_t *pand, uint64_t *ctl) {
    __shared__ uint32_t lds[GB_THREADS / VG_WAVE];
    __shared__ uint64_t lor[2 * GB_THREADS / VG_WAVE];
    const int64_t base = (int64_t)blockIdx.x * GB_CMP_SLOTS;
    uint64_t sk[GB_CMP_PER_THREAD];
    uint32_t mask = 0;
    uint64_t o = 0, a = ~0ull;
#pragma unroll
    for (int j = 0; j < GB_CMP_PER_THREAD; j++) {
        int64_t s = base + j * GB_THREADS + threadIdx.x;
        sk[j] = 0;
        if (s < n_slots && gb_sort_entry(L, R, order, s, sk[j])) {
            mask |= 1u << j;
            o |= sk[j];
            a &= sk[j];
        }
        if (s == R.na_slot && s < n_slots) ctl[8] = R.hot[s * R.hstride + 1];   // rows of the NA group
    }
    uint32_t total;
    uint64_t off = boff[blockIdx.x] + gb_block_prefix(__builtin_popcount(mask), lds, total);
#pragma unroll
    for (int j = 0; j < GB_CMP_PER_THREAD; j++)
        if (mask & (1u << j)) {
            skeys[off] = sk[j];
            svals[off] = (uint32_t)(base + j * GB_THREADS + threadIdx.x);
            off++;
        }
    gb_block_or_and(o, a, lor);
    if (threadIdx.x == 0) {
        por[blockIdx.x] = o;
        pand[blockIdx.x] = a;
    }
}

// OR / AND of the per-block partials into ctl[6] / ctl[7] (which start at
// 0 / all-ones): a block-strided reduction, one atomic pair per block
static __global__ void __launch_bounds__(GB_THREADS)
k_gb_or_and_final(const uint64_t *por, const uint64_t *pand, int64_t nb, uint64_t *ctl) {
    __shared__ uint64_t lor[2 * GB_THREADS / VG_WAVE];
    uint64_t o = 0, a = ~0ull;
    for (int64_t i = (int64_t)blockIdx.x * GB_THREADS + threadIdx.x; i < nb; i += (int64_t)gridDim.x * GB_THREADS) {
        o |= por[i];
        a &= pand[i];
    }
    gb_block_or_and(o, a, lor);
    if (threadIdx.x == 0) {
        if (o) atomicOr((unsigned long long *)&ctl[6], (unsigned long long)o);
        if (~a) atomicAnd((unsigned long long *)&ctl[7], (unsigned long long)a);
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
        const int64_t v = (int64_t)(o ^ 0x8000000000000000ull);
        out = (ag.reserved & VGPU_AGG_RAW) ? vg_bits_to_f64((uint64_t)v) : (double)v;
    }
}

// Records whose words fit GB_EMIT_STAGE_WORDS are staged through LDS: the
// block copies its records word by word (consecutive lanes, consecutive
// words of one record) before each thread finishes its group, instead of
// every lane gathering the words of its own record.
#define GB_EMIT_STAGE_WORDS 24

// one output row g of the finished result: key, key validity bit, every
// aggregate and its validity bit (bitmap words written by lane 0 of each
// wave from ballots, so every lane of the wave must call it); h / c are
// the group's record words, s its slot in R, ctl its query's control words
__device__ __forceinline__ void gb_emit_row(const GbLayout &L, const GbAggs &AG, const RecSrc &R, int64_t g, bool in,
                                            int64_t s, const uint64_t *h, const uint64_t *c, int64_t n_out,
                                            int key_width, const EmitOut &O, uint64_t *ctl) {
    bool kvalid = in && s != R.na_slot;
    if (in && !L.key_none) {
        uint64_t kb = 0;
        if (s == R.sp_slot) kb = VG_EMPTY_I64;
        else if (s != R.na_slot) {
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

// copies the records of the block's output rows into LDS word by word
// (consecutive lanes, consecutive words of one record); rec_of(k) gives the
// record slot of the block's k-th row
template <class SlotOf>
__device__ __forceinline__ void gb_stage_records(const GbLayout &L, const RecSrc &R, int cnt, SlotOf rec_of,
                                                 uint64_t *dst) {
    const int RW = L.RH + L.RC;
    // q / RW by a multiply (exact for q < 2^32 / RW)
    const uint32_t mag = (uint32_t)((0xFFFFFFFFull + (uint64_t)RW) / (uint64_t)RW);
    for (int q = threadIdx.x; q < cnt * RW; q += GB_THREADS) {
        int k = (int)__umulhi((uint32_t)q, mag), w = q - k * RW;
        int64_t sk = rec_of(k);
        dst[q] = w < L.RH ? R.hot[sk * R.hstride + w] : R.cold[sk * R.cstride + (w - L.RH)];
    }
}

template <bool STAGED>
__global__ void __launch_bounds__(GB_THREADS)
k_gb_emit(GbLayout L, GbAggs AG, RecSrc R, const uint32_t *order_slots, int64_t n_sorted, int na_last, int64_t n_out,
          int key_width, EmitOut O, uint64_t *ctl) {
    extern __shared__ uint64_t dynw[];
    __shared__ int64_t sidx[GB_THREADS];
    int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool in = g < n_out;
    int64_t s = 0;
    if (in) s = (na_last && g == n_sorted) ? R.na_slot : (int64_t)order_slots[g];
    const uint64_t *h, *c;
    if (STAGED) {
        sidx[threadIdx.x] = s;
        __syncthreads();
        const int64_t g0 = (int64_t)blockIdx.x * blockDim.x;
        const int cnt = n_out - g0 < GB_THREADS ? (int)(n_out - g0) : GB_THREADS;
        gb_stage_records(L, R, cnt, [&](int k) { return sidx[k]; }, dynw);
        __syncthreads();
        h = dynw + threadIdx.x * (L.RH + L.RC);
        c = h + L.RH;
    } else {
        h = R.hot + s * R.hstride;
        c = R.cold + s * R.cstride;
    }
    gb_emit_row(L, AG, R, g, in, s, h, c, n_out, key_width, O, ctl);
}

// ---------------------------------------------------------------------------
// host helpers
// ---------------------------------------------------------------------------

static inline int64_t pow2_at_least(int64_t x) {
    int64_t p = 64;
    while (p < x) p <<= 1;
    return p;
}

static inline int gb_layout(GbLayout &L, int order, int key_type, int n_inputs, const int32_t *types, int n_aggs,
                            const vgpu_agg *aggs) {
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
    if (order == VGPU_ORDER_FIRST_SEEN || any_dbl_mm) { L.w_fpos = h; hot_ones |= 1ull << h; h++; }
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
