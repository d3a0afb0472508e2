// vgpu_join.hip -- hash-join build and probe, join output gather, and the
// column cast the join uses to bring keys to a common type.
//
// Reference semantics (gcol33/vectra 0.2.2):
//   build over the whole right side ............. src/join.c:156-220
//   chain order: later build rows first ......... src/join.c:45-61, 115-150
//   probe in logical order; left/full append the
//   batch's unmatched probe rows after all matches src/join.c:262-480
//   full join: unmatched build rows at the end .. src/join.c:487-551
//   NA keys never match ......................... src/join.c:85
//
// Build: keys are claimed in an open-addressing table with 64-bit CAS while
// a per-slot counter counts rows; a scan of the counters gives each key a
// contiguous segment of rows[].  With unique keys (the common dimension
// table) every segment has one row and is filled directly; with duplicates
// a stable radix sort on (slot, ~row) lays each segment out in descending
// build row order.
// Probe: count pass (match count + segment offset per probe row), device
// scans, write pass -- an ordered scatter, so the output order is the
// reference's without any sort.
#include "vgpu_join_table.h"

#include <algorithm>

#define JT_THREADS 256
#define JT_NONE 0xFFFFFFFFu

__global__ void k_jt_init(uint64_t *ent, int64_t n, int E, uint64_t empty, uint32_t *cnt) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) {
        uint64_t *e = ent + i * E;
        e[0] = empty;
        for (int w = 1; w < E; w++) e[w] = 0;
        cnt[i] = 0;
    }
}

__global__ void k_jt_zero(uint64_t *p, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

__global__ void k_jt_insert(vgpu_col key, int64_t n, uint64_t *ent, int E, int64_t cap, uint32_t *cnt,
                            uint32_t *slot_of, uint32_t *err, uint64_t *bloom, uint64_t bloom_mask) {
    int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    uint64_t kn;
    int cls = vg_key_class(key, r, kn);
    if (cls == VG_KEY_NA || cls == VG_KEY_NAN) {
        slot_of[r] = JT_NONE;
        return;
    }
    int64_t s = cap;
    if (cls == VG_KEY_REGULAR) {
        const uint64_t empty = key.type == VGPU_DOUBLE ? VG_EMPTY_F64 : VG_EMPTY_I64;
        uint64_t mask = (uint64_t)cap - 1;
        const uint64_t kh = vg_key_hash(kn);
        if (bloom) atomicOr((unsigned long long *)&bloom[jt_bloom_word(kh, bloom_mask)], jt_bloom_bits(kh));
        uint64_t h = kh & mask;
        s = -1;
        for (int64_t probes = 0; probes < cap; probes++) {
            uint64_t *kp = &ent[h * E];
            uint64_t k = vg_atomic_load(kp);
            if (k == kn) { s = (int64_t)h; break; }
            if (k == empty) {
                uint64_t old = atomicCAS((unsigned long long *)kp, (unsigned long long)empty,
                                         (unsigned long long)kn);
                if (old == empty || old == kn) { s = (int64_t)h; break; }
            }
            h = (h + 1) & mask;
        }
        if (s < 0) {
            atomicOr(err, 1u);
            slot_of[r] = JT_NONE;
            return;
        }
    }
    atomicAdd(&cnt[s], 1u);
    slot_of[r] = (uint32_t)s;
}

__global__ void k_jt_seg(const uint32_t *cnt, const uint64_t *off, int64_t n, uint64_t *ent, int E,
                         uint32_t *max_cnt) {
    int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    uint32_t c = cnt[s];
    ent[s * E + 1] = (off[s] << 32) | c;
    if (c > 1) atomicMax(max_cnt, c);
}

__global__ void k_jt_fill_unique(const uint32_t *slot_of, int64_t n, const uint64_t *ent, int E, uint32_t *rows) {
    int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    uint32_t s = slot_of[r];
    if (s == JT_NONE) return;
    rows[ent[(int64_t)s * E + 1] >> 32] = (uint32_t)r;
}

__global__ void k_jt_sortkeys(const uint32_t *slot_of, int64_t n, int64_t cap, uint64_t *skeys, uint32_t *svals) {
    int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    uint64_t s = slot_of[r] == JT_NONE ? (uint64_t)cap + 1 : (uint64_t)slot_of[r];
    skeys[r] = (s << 32) | (uint64_t)(0xFFFFFFFFu - (uint32_t)r);
    svals[r] = (uint32_t)r;
}

struct PaySet {
    vgpu_col c[JT_MAX_PAYLOAD];
};

// embeds the payload of each key's first (highest) build row in its entry
__global__ void k_jt_embed(uint64_t *ent, int E, int64_t n_slots, const uint32_t *rows, PaySet ps, int n_payload) {
    int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n_slots) return;
    uint64_t *e = ent + s * E;
    uint64_t sg = e[1];
    if ((uint32_t)sg == 0) return;
    uint32_t r0 = rows[sg >> 32];
    uint64_t vbits = 0;
    for (int p = 0; p < n_payload; p++) {
        bool ok = vg_valid(ps.c[p].validity, r0);
        e[2 + p] = ok ? vg_load_bits(ps.c[p], r0) : 0;
        if (ok) vbits |= 1ull << (24 + p);
    }
    e[1] = sg | vbits;
}

// payload in rows[] order, for keys with several build rows
__global__ void k_jt_cluster(const uint32_t *rows, int64_t n_valid, vgpu_col c, uint64_t *pay, uint8_t *val) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_valid) return;
    uint32_t r = rows[i];
    bool ok = vg_valid(c.validity, r);
    pay[i] = ok ? vg_load_bits(c, r) : 0;
    val[i] = ok;
}

static vgpu_jtable *join_build_impl(const vgpu_col *key, int64_t n_build, const vgpu_col *payload, int n_payload) {
    if (vg_ensure_init()) return nullptr;
    if (key->type != VGPU_INT64 && key->type != VGPU_DOUBLE && key->type != VGPU_BOOL) {
        vg_set_error("join_build: unsupported key type %d", key->type);
        return nullptr;
    }
    if (n_build < 0 || n_build >= (int64_t)JT_NONE) {
        vg_set_error("join_build: %lld build rows exceed the 32-bit row index", (long long)n_build);
        return nullptr;
    }
    if (n_payload < 0 || n_payload > JT_MAX_PAYLOAD) {
        vg_set_error("join_build: %d payload columns exceed %d", n_payload, JT_MAX_PAYLOAD);
        return nullptr;
    }
    vgpu_jtable *t = new vgpu_jtable();
    t->key_type = key->type;
    t->n_build = n_build;
    t->n_payload = n_payload;
    for (int p = 0; p < n_payload; p++) t->payload[p] = payload[p];
    t->E = n_payload == 0 ? 2 : (n_payload <= 2 ? 4 : 6);
    if (key->type != VGPU_BOOL) t->bkey = static_cast<const uint64_t *>(key->data);
    // slots: a power of two keeping the load at most join.load_pct percent
    // (tuned per build size and entry width)
    VgTuneScope tune_scope(t->E, n_build);
    const int64_t load = std::min<int64_t>(95, std::max<int64_t>(10, vg_tune_for(VG_TUNE_JOIN_LOAD_PCT, n_build)));
    int64_t cap = 64;
    while (cap * load < 100 * n_build) cap <<= 1;
    t->cap = cap;
    const int E = t->E;
    hipStream_t s = vg_stream();
    int64_t slots = cap + 1;
    DevBuf cnt, slot_of, off, misc;
    auto fail = [&]() -> vgpu_jtable * {
        cnt.release(); slot_of.release(); off.release(); misc.release();
        t->ent.release(); t->rows.release(); t->sort_tmp.release(); t->scan_tmp.release();
        for (int p = 0; p < JT_MAX_PAYLOAD; p++) { t->cpay[p].release(); t->cval[p].release(); }
        delete t;
        return nullptr;
    };
    if (t->ent.reserve((size_t)slots * E * 8) || cnt.reserve((size_t)slots * 4) ||
        slot_of.reserve((size_t)std::max<int64_t>(n_build, 1) * 4) || off.reserve((size_t)slots * 8) ||
        misc.reserve(64))
        return fail();
    uint64_t empty = t->key_type == VGPU_DOUBLE ? VG_EMPTY_F64 : VG_EMPTY_I64;
    if (hipMemsetAsync(misc.ptr, 0, 64, s) != hipSuccess) { vg_set_error("join_build: memset failed"); return fail(); }
    k_jt_init<<<vg_blocks(slots, 256, 8192), 256, 0, s>>>(t->ent.as<uint64_t>(), slots, E, empty, cnt.as<uint32_t>());
    // the filter: about 16 bits per build key, a power of two of words
    // one 64-bit word per join.bloom_keys build keys (tuned; 4: about 16 bits a key)
    const uint64_t kpw = (uint64_t)std::max<int64_t>(1, vg_tune(VG_TUNE_JOIN_BLOOM_KPW));
    t->bloom_words = 64;
    while (t->bloom_words < (uint64_t)n_build / kpw && t->bloom_words < (1ull << 22)) t->bloom_words <<= 1;
    if (t->bloom.reserve(t->bloom_words * 8)) return fail();
    k_jt_zero<<<vg_blocks((int64_t)t->bloom_words, 256), 256, 0, s>>>(t->bloom.as<uint64_t>(),
                                                                     (int64_t)t->bloom_words);
    if (n_build > 0) {
        vgpu_col k = *key;
        k_jt_insert<<<vg_blocks(n_build, JT_THREADS), JT_THREADS, 0, s>>>(k, n_build, t->ent.as<uint64_t>(), E, cap,
                                                                          cnt.as<uint32_t>(), slot_of.as<uint32_t>(),
                                                                          misc.as<uint32_t>(), t->bloom.as<uint64_t>(),
                                                                          t->bloom_words - 1);
    }
    uint64_t *total = misc.as<uint64_t>() + 2;
    if (vg_scan_u32_to_u64(cnt.as<uint32_t>(), off.as<uint64_t>(), slots, total, t->scan_tmp)) return fail();
    k_jt_seg<<<vg_blocks(slots, 256), 256, 0, s>>>(cnt.as<uint32_t>(), off.as<uint64_t>(), slots,
                                                   t->ent.as<uint64_t>(), E, misc.as<uint32_t>() + 1);
    // value ranges of the int64 / bool payload columns, read back with the
    // build's counts (one synchronization)
    DevBuf pmm;
    unsigned long long hr[2 * JT_MAX_PAYLOAD];
    if (n_payload > 0) {
        for (int p = 0; p < n_payload; p++) {
            hr[2 * p] = ~0ull;
            hr[2 * p + 1] = 0;
        }
        if (pmm.reserve(sizeof hr) ||
            hipMemcpyAsync(pmm.ptr, hr, (size_t)n_payload * 16, hipMemcpyHostToDevice, s) != hipSuccess)
            return fail();
        for (int p = 0; p < n_payload; p++)
            if (n_build > 0 && (payload[p].type == VGPU_INT64 || payload[p].type == VGPU_BOOL))
                k_key_minmax<<<vg_blocks(n_build, 256, 1024), 256, 0, s>>>(payload[p], n_build,
                                                                            pmm.as<unsigned long long>() + 2 * p);
        if (hipMemcpyAsync(hr, pmm.ptr, (size_t)n_payload * 16, hipMemcpyDeviceToHost, s) != hipSuccess)
            return fail();
    }
    uint64_t hm[8];
    if (hipMemcpyAsync(hm, misc.ptr, 64, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
        vg_set_error("join_build: %s", hipGetErrorString(hipGetLastError()));
        return fail();
    }
    uint32_t err = (uint32_t)hm[0], maxc = (uint32_t)(hm[0] >> 32);
    if (err) {
        vg_set_error("join_build: hash table overflow");
        return fail();
    }
    if (n_payload > 0 && maxc >= (1u << 24)) {
        vg_set_error("join_build: a key repeats %u times; payload embedding needs fewer than 2^24", maxc);
        return fail();
    }
    for (int p = 0; p < n_payload; p++) {
        const bool ranged = (payload[p].type == VGPU_INT64 || payload[p].type == VGPU_BOOL) && hr[2 * p] <= hr[2 * p + 1];
        t->pay_ranged[p] = ranged;
        t->pay_lo[p] = ranged ? (int64_t)(hr[2 * p] ^ 0x8000000000000000ull) : 0;
        t->pay_hi[p] = ranged ? (int64_t)(hr[2 * p + 1] ^ 0x8000000000000000ull) : 0;
    }
    pmm.release();
    t->n_valid = (int64_t)hm[2];
    t->max_dup = (int32_t)std::max<uint32_t>(maxc, t->n_valid > 0 ? 1u : 0u);
    if (t->rows.reserve((size_t)std::max<int64_t>(n_build, 1) * 4)) return fail();
    if (n_build > 0) {
        if (maxc <= 1) {
            k_jt_fill_unique<<<vg_blocks(n_build, 256), 256, 0, s>>>(slot_of.as<uint32_t>(), n_build,
                                                                      t->ent.as<uint64_t>(), E, t->rows.as<uint32_t>());
        } else {
            DevBuf skeys;
            if (skeys.reserve((size_t)n_build * 8)) return fail();
            k_jt_sortkeys<<<vg_blocks(n_build, 256), 256, 0, s>>>(slot_of.as<uint32_t>(), n_build, cap,
                                                                   skeys.as<uint64_t>(), t->rows.as<uint32_t>());
            // slots (cap + 1 for rows kept out of the table) sit above the 32
            // bits of row complement: sort only the bits they occupy
            int hi = 33;
            while (hi < 64 && ((int64_t)1 << (hi - 32)) <= cap + 1) hi++;
            int rc = vg_sort_pairs(skeys.as<uint64_t>(), t->rows.as<uint32_t>(), n_build, 0, hi, t->sort_tmp);
            skeys.release();   // stream-ordered pool: safe before the sort has run
            if (rc) return fail();
        }
        if (n_payload > 0) {
            PaySet ps;
            memset(&ps, 0, sizeof ps);
            for (int p = 0; p < n_payload; p++) ps.c[p] = payload[p];
            k_jt_embed<<<vg_blocks(slots, 256), 256, 0, s>>>(t->ent.as<uint64_t>(), E, slots, t->rows.as<uint32_t>(),
                                                             ps, n_payload);
            if (maxc > 1) {
                for (int p = 0; p < n_payload; p++) {
                    if (t->cpay[p].reserve((size_t)std::max<int64_t>(t->n_valid, 1) * 8) ||
                        t->cval[p].reserve((size_t)std::max<int64_t>(t->n_valid, 1)))
                        return fail();
                    if (t->n_valid > 0)
                        k_jt_cluster<<<vg_blocks(t->n_valid, 256), 256, 0, s>>>(
                            t->rows.as<uint32_t>(), t->n_valid, payload[p], t->cpay[p].as<uint64_t>(),
                            t->cval[p].as<uint8_t>());
                }
            }
        }
    }
    if (hipGetLastError() != hipSuccess) {
        vg_set_error("join_build: kernel launch failed");
        return fail();
    }
    // scratch returns to the stream-ordered pool; later work on the stream
    // sees the finished table without a host synchronization here
    cnt.release();
    slot_of.release();
    off.release();
    misc.release();
    t->sort_tmp.release();
    return t;
}

extern "C" vgpu_jtable *vgpu_join_build(const vgpu_col *key, int64_t n_build) {
    return vg_guard<vgpu_jtable *>(nullptr, "vgpu_join_build",
                                   [&] { return join_build_impl(key, n_build, nullptr, 0); });
}

extern "C" vgpu_jtable *vgpu_join_build_payload(const vgpu_col *key, int64_t n_build, const vgpu_col *payload,
                                                int n_payload) {
    return vg_guard<vgpu_jtable *>(nullptr, "vgpu_join_build_payload",
                                   [&] { return join_build_impl(key, n_build, payload, n_payload); });
}

extern "C" int64_t vgpu_join_build_rows(const vgpu_jtable *t) { return t ? t->n_build : -1; }
extern "C" int64_t vgpu_join_max_dup(const vgpu_jtable *t) { return t ? (int64_t)t->max_dup : -1; }

extern "C" int64_t vgpu_join_device_bytes(const vgpu_jtable *t) {
    if (!t) return -1;
    int64_t b = (int64_t)(t->ent.cap + t->rows.cap + t->bloom.cap);
    for (int p = 0; p < JT_MAX_PAYLOAD; p++) b += (int64_t)(t->cpay[p].cap + t->cval[p].cap);
    return b;
}

// ---------------------------------------------------------------------------
// probe
// ---------------------------------------------------------------------------

// A left / full probe whose key had to be coerced to the build key's type
// continues the reference's chain walk after the first match comparing the
// ORIGINAL probe key with the build key array's bytes (src/join.c:398-411
// pass the uncoerced probe columns to jht_chain_next, whose join_keys_equal
// switches on the probe column's type): an int64 probe against the coerced
// double build key read as int64, a bool probe's byte against byte `br` of
// the build key array.  Every later row of the segment holds the same key
// value, so this comparison alone decides whether it is still emitted.
__device__ __forceinline__ bool jp_chain_keep(const vgpu_col &orig, int64_t pr, const uint64_t *bkey, uint32_t br) {
    if (!bkey) return false;
    if (orig.type == VGPU_BOOL)
        return static_cast<const uint8_t *>(orig.data)[pr] == reinterpret_cast<const uint8_t *>(bkey)[br];
    return static_cast<const uint64_t *>(orig.data)[pr] == bkey[br];
}

__global__ void __launch_bounds__(JT_THREADS)
k_jp_count(JView v, vgpu_col key, int64_t n_logical, const int32_t *sel, int kind, int coerced, vgpu_col orig,
           uint32_t *cnt_main, uint32_t *un, uint32_t *segoff, uint32_t *segcnt) {
    int64_t li = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (li >= n_logical) return;
    int64_t pr = sel ? (int64_t)sel[li] : li;
    uint32_t off;
    uint32_t c = jt_lookup(v, key, pr, off);
    uint32_t m;
    switch (kind) {
    case VGPU_JOIN_SEMI: m = c > 0; break;
    case VGPU_JOIN_ANTI: m = c == 0; break;
    default:
        m = c;
        if (coerced && c > 1) {
            m = 1;
            for (uint32_t j = 1; j < c; j++) m += jp_chain_keep(orig, pr, v.bkey, v.rows[off + j]);
        }
        break;
    }
    cnt_main[li] = m;
    segoff[li] = off;
    if (segcnt) segcnt[li] = c;
    if (un) un[li] = c == 0;
}

__global__ void __launch_bounds__(JT_THREADS)
k_jp_write(JView v, int64_t n_logical, const int32_t *sel, int kind, int coerced, vgpu_col orig,
           const uint32_t *cnt_main, const uint64_t *off_main, const uint32_t *un, const uint64_t *off_un,
           const uint64_t *totals, const uint32_t *segoff, const uint32_t *segcnt, int32_t *out_probe,
           int64_t *out_build, uint8_t *bmatched) {
    int64_t li = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (li >= n_logical) return;
    int32_t pr = sel ? sel[li] : (int32_t)li;
    uint32_t m = cnt_main[li];
    uint64_t o = off_main[li];
    if (kind == VGPU_JOIN_SEMI || kind == VGPU_JOIN_ANTI) {
        if (m) { out_probe[o] = pr; out_build[o] = -1; }
        return;
    }
    uint32_t so = segoff[li];
    uint32_t c = segcnt[li];
    for (uint32_t j = 0, k = 0; j < c && k < m; j++) {
        uint32_t br = v.rows[so + j];
        // rows past the first are emitted exactly when the count pass
        // counted them (k_jp_count, same rule)
        if (coerced && j > 0 && !jp_chain_keep(orig, pr, v.bkey, br)) continue;
        out_probe[o + k] = pr;
        out_build[o + k] = (int64_t)br;
        k++;
        if (bmatched) bmatched[br] = 1;
    }
    if (un && un[li]) {
        uint64_t p = totals[0] + off_un[li];
        out_probe[p] = pr;
        out_build[p] = -1;
    }
}

static int join_probe_body(vgpu_jtable *t, int kind, const vgpu_col *probe_key, int64_t n_rows, const int32_t *sel,
                           int64_t n_sel, const vgpu_col *uncoerced_key, int64_t *n_out) {
    if (kind < VGPU_JOIN_INNER || kind > VGPU_JOIN_ANTI) {
        vg_set_error("join_probe: unknown join kind %d", kind);
        return -1;
    }
    if (probe_key->type != t->key_type) {
        vg_set_error("join_probe: probe key type %d differs from build key type %d", probe_key->type, t->key_type);
        return -1;
    }
    // the chain rule of a coerced probe applies to left / full joins only
    // (inner joins walk the chain on the coerced key, src/join.c:388-395)
    int coerced = uncoerced_key && (kind == VGPU_JOIN_LEFT || kind == VGPU_JOIN_FULL);
    vgpu_col orig;
    memset(&orig, 0, sizeof orig);
    if (coerced) {
        orig = *uncoerced_key;
        if (orig.type != VGPU_INT64 && orig.type != VGPU_BOOL) {
            vg_set_error("join_probe: an uncoerced probe key is int64 or bool (got type %d)", orig.type);
            return -1;
        }
        if (!t->bkey) {
            vg_set_error("join_probe: a coerced probe needs an int64 or double build key");
            return -1;
        }
    }
    hipStream_t s = vg_stream();
    int64_t n = sel ? n_sel : n_rows;
    *n_out = 0;
    t->last_n_out = 0;
    bool tail = kind == VGPU_JOIN_LEFT || kind == VGPU_JOIN_FULL;
    if (kind == VGPU_JOIN_FULL && !t->bmatched.ptr) {
        if (t->bmatched.reserve((size_t)std::max<int64_t>(t->n_build, 1))) return -1;
        VG_TRY(hipMemsetAsync(t->bmatched.ptr, 0, (size_t)std::max<int64_t>(t->n_build, 1), s));
    }
    if (n <= 0) return 0;
    if (t->cnt.reserve((size_t)n * 4) || t->un.reserve((size_t)n * 4 * 3) || t->off_main.reserve((size_t)n * 8) ||
        t->off_un.reserve((size_t)n * 8) || t->totals.reserve(16))
        return -1;
    uint32_t *segoff = t->un.as<uint32_t>() + n;       // second third of `un`
    uint32_t *segcnt = t->un.as<uint32_t>() + 2 * n;   // last third
    JView v = t->view();
    vgpu_col k = *probe_key;
    unsigned grid = vg_blocks(n, JT_THREADS, 1 << 30);
    k_jp_count<<<grid, JT_THREADS, 0, s>>>(v, k, n, sel, kind, coerced, orig, t->cnt.as<uint32_t>(),
                                           tail ? t->un.as<uint32_t>() : nullptr, segoff, segcnt);
    VG_LAUNCH_CHECK();
    uint64_t *tot = t->totals.as<uint64_t>();
    if (vg_scan_u32_to_u64(t->cnt.as<uint32_t>(), t->off_main.as<uint64_t>(), n, tot, t->scan_tmp)) return -1;
    if (tail) {
        if (vg_scan_u32_to_u64(t->un.as<uint32_t>(), t->off_un.as<uint64_t>(), n, tot + 1, t->scan_tmp)) return -1;
    } else {
        VG_TRY(hipMemsetAsync(tot + 1, 0, 8, s));
    }
    uint64_t h[2];
    VG_TRY(hipMemcpyAsync(h, tot, 16, hipMemcpyDeviceToHost, s));
    VG_TRY(hipStreamSynchronize(s));
    int64_t total = (int64_t)(h[0] + h[1]);
    if (t->out_probe.reserve((size_t)std::max<int64_t>(total, 1) * 4) ||
        t->out_build.reserve((size_t)std::max<int64_t>(total, 1) * 8))
        return -1;
    k_jp_write<<<grid, JT_THREADS, 0, s>>>(v, n, sel, kind, coerced, orig, t->cnt.as<uint32_t>(),
                                           t->off_main.as<uint64_t>(),
                                           tail ? t->un.as<uint32_t>() : nullptr, t->off_un.as<uint64_t>(), tot, segoff,
                                           segcnt, t->out_probe.as<int32_t>(), t->out_build.as<int64_t>(),
                                           kind == VGPU_JOIN_FULL ? t->bmatched.as<uint8_t>() : nullptr);
    VG_LAUNCH_CHECK();
    VG_TRY(hipStreamSynchronize(s));
    *n_out = total;
    t->last_n_out = total;
    return 0;
}

extern "C" int vgpu_join_probe(vgpu_jtable *t, int kind, const vgpu_col *probe_key, int64_t n_rows, const int32_t *sel,
                               int64_t n_sel, const vgpu_col *uncoerced_key, int64_t *n_out) {
    return vg_guard<int>(-1, "vgpu_join_probe", [&] { return join_probe_body(t, kind, probe_key, n_rows, sel, n_sel,
        uncoerced_key, n_out); });
}

extern "C" int vgpu_join_pairs(vgpu_jtable *t, const int32_t **probe_rows, const int64_t **build_rows) {
    if (probe_rows) *probe_rows = t->out_probe.as<int32_t>();
    if (build_rows) *build_rows = t->out_build.as<int64_t>();
    return 0;
}

__global__ void k_unmatched_flags(const uint8_t *bmatched, int64_t n, uint32_t *flags) {
    int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r < n) flags[r] = bmatched[r] == 0;
}

__global__ void k_unmatched_write(const uint32_t *flags, const uint64_t *off, int64_t n, int32_t *out_probe,
                                  int64_t *out_build) {
    int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r < n && flags[r]) {
        out_probe[off[r]] = -1;
        out_build[off[r]] = r;
    }
}

extern "C" int vgpu_join_unmatched_build(vgpu_jtable *t, int64_t *n_out) {
    hipStream_t s = vg_stream();
    int64_t n = t->n_build;
    *n_out = 0;
    t->last_n_out = 0;
    if (n <= 0) return 0;
    if (!t->bmatched.ptr) {
        if (t->bmatched.reserve((size_t)n)) return -1;
        VG_TRY(hipMemsetAsync(t->bmatched.ptr, 0, (size_t)n, s));
    }
    if (t->cnt.reserve((size_t)n * 4) || t->off_main.reserve((size_t)n * 8) || t->totals.reserve(16)) return -1;
    k_unmatched_flags<<<vg_blocks(n, 256, 1 << 30), 256, 0, s>>>(t->bmatched.as<uint8_t>(), n, t->cnt.as<uint32_t>());
    VG_LAUNCH_CHECK();
    uint64_t *tot = t->totals.as<uint64_t>();
    if (vg_scan_u32_to_u64(t->cnt.as<uint32_t>(), t->off_main.as<uint64_t>(), n, tot, t->scan_tmp)) return -1;
    uint64_t h = 0;
    VG_TRY(hipMemcpyAsync(&h, tot, 8, hipMemcpyDeviceToHost, s));
    VG_TRY(hipStreamSynchronize(s));
    if (t->out_probe.reserve((size_t)std::max<uint64_t>(h, 1) * 4) ||
        t->out_build.reserve((size_t)std::max<uint64_t>(h, 1) * 8))
        return -1;
    k_unmatched_write<<<vg_blocks(n, 256, 1 << 30), 256, 0, s>>>(t->cnt.as<uint32_t>(), t->off_main.as<uint64_t>(), n,
                                                                  t->out_probe.as<int32_t>(), t->out_build.as<int64_t>());
    VG_LAUNCH_CHECK();
    VG_TRY(hipStreamSynchronize(s));
    *n_out = (int64_t)h;
    t->last_n_out = (int64_t)h;
    return 0;
}

// ---------------------------------------------------------------------------
// direct-addressed map of a unique-key table
// ---------------------------------------------------------------------------

// acc[0] = min, acc[1] = max of the regular keys (sign-flipped order);
// acc[2] != 0 when a key repeats or the sentinel slot holds rows
__global__ void k_jd_range(const uint64_t *ent, int64_t cap, int E, unsigned long long *acc) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s <= cap; s += stride) {
        const uint64_t *e = ent + (size_t)s * E;
        const uint32_t n = (uint32_t)(e[1] & 0xFFFFFFu);
        if (s == cap) {
            if (n) atomicOr(&acc[2], 1ull);
            continue;
        }
        if (e[0] == VG_EMPTY_I64) continue;
        if (n != 1) atomicOr(&acc[2], 1ull);
        const unsigned long long o = e[0] ^ 0x8000000000000000ull;
        atomicMin(&acc[0], o);
        atomicMax(&acc[1], o);
    }
}

__global__ void k_jd_fill(const uint64_t *ent, int64_t cap, int E, int p, int64_t kmin, int64_t pmin, uint16_t *map) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < cap; s += stride) {
        const uint64_t *e = ent + (size_t)s * E;
        if (e[0] == VG_EMPTY_I64) continue;
        const bool valid = (e[1] >> (24 + p)) & 1;
        map[e[0] - (uint64_t)kmin] = valid ? (uint16_t)((int64_t)e[2 + p] - pmin) : (uint16_t)JD_NA_PAYLOAD;
    }
}

static int jtable_dense_build(vgpu_jtable *t, int p) {
    if (vg_tune(VG_TUNE_JOIN_DENSE) < 2 || t->key_type != VGPU_INT64 || t->max_dup > 1 || p < 0 ||
        p >= t->n_payload || (t->payload[p].type != VGPU_INT64 && t->payload[p].type != VGPU_BOOL) ||
        !t->pay_ranged[p] || (uint64_t)t->pay_hi[p] - (uint64_t)t->pay_lo[p] >= 0xFFFDull || t->n_valid <= 0)
        return 0;
    hipStream_t s = vg_stream();
    DevBuf acc;
    if (acc.reserve(24)) return -1;
    const unsigned long long init[3] = {~0ull, 0ull, 0ull};
    VG_TRY(hipMemcpyAsync(acc.ptr, init, 24, hipMemcpyHostToDevice, s));
    k_jd_range<<<vg_blocks(t->cap + 1, 256, 8192), 256, 0, s>>>(t->ent.as<uint64_t>(), t->cap, t->E,
                                                               acc.as<unsigned long long>());
    VG_LAUNCH_CHECK();
    unsigned long long h[3];
    VG_TRY(hipMemcpyAsync(h, acc.ptr, 24, hipMemcpyDeviceToHost, s));
    VG_TRY(hipStreamSynchronize(s));
    acc.release();
    if (h[2] || h[0] > h[1]) return 0;
    const uint64_t range = h[1] - h[0] + 1;
    if (range > (uint64_t)1 << 30 || range > (uint64_t)16 * (uint64_t)t->n_valid) return 0;
    if (t->dmap.reserve((size_t)range * 2)) return -1;
    const int64_t kmin = (int64_t)(h[0] ^ 0x8000000000000000ull);
    VG_TRY(hipMemsetAsync(t->dmap.ptr, 0xFF, (size_t)range * 2, s));
    k_jd_fill<<<vg_blocks(t->cap, 256, 8192), 256, 0, s>>>(t->ent.as<uint64_t>(), t->cap, t->E, p, kmin,
                                                          t->pay_lo[p], t->dmap.as<uint16_t>());
    VG_LAUNCH_CHECK();
    VG_TRY(hipStreamSynchronize(s));
    t->dm = {t->dmap.as<uint16_t>(), kmin, range, t->pay_lo[p]};
    return 1;
}

int vg_jtable_dense(vgpu_jtable *t, int p, JDense *out) {
    *out = {nullptr, 0, 0, 0};
    std::lock_guard<std::mutex> lk(t->dm_mu);
    // one map per table, for the first payload column asked for: a map in
    // use by other passes is never replaced
    if (t->dm_state == 0) {
        const int rc = jtable_dense_build(t, p);
        if (rc < 0) return -1;
        t->dm_state = rc ? 1 : -1;
        t->dm_payload = p;
    }
    if (t->dm_state == 1 && t->dm_payload == p) *out = t->dm;
    return 0;
}

extern "C" void vgpu_join_destroy(vgpu_jtable *t) {
    if (!t) return;
    DevBuf *bufs[] = {&t->ent, &t->rows, &t->cnt, &t->un, &t->off_main, &t->off_un, &t->out_probe,
                      &t->out_build, &t->totals, &t->scan_tmp, &t->sort_tmp, &t->bmatched, &t->bloom, &t->dmap};
    for (DevBuf *b : bufs) b->release();
    for (int p = 0; p < JT_MAX_PAYLOAD; p++) { t->cpay[p].release(); t->cval[p].release(); }
    delete t;
}

// ---------------------------------------------------------------------------
// gather and cast
// ---------------------------------------------------------------------------

__global__ void k_gather(vgpu_col src, const void *idx, int idx64, int64_t n, void *dst, uint64_t *dvalid) {
    int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool in = g < n;
    int64_t r = -1;
    if (in) r = idx64 ? static_cast<const int64_t *>(idx)[g] : (int64_t)static_cast<const int32_t *>(idx)[g];
    bool ok = in && r >= 0 && vg_valid(src.validity, r);
    if (in) {
        if (src.type == VGPU_BOOL) static_cast<uint8_t *>(dst)[g] = ok ? static_cast<const uint8_t *>(src.data)[r] : 0;
        else static_cast<uint64_t *>(dst)[g] = ok ? static_cast<const uint64_t *>(src.data)[r] : 0;
    }
    uint64_t w = vg_ballot(ok);
    if (vg_lane() == 0 && in) dvalid[g >> 6] = w;
}

extern "C" int vgpu_gather(const vgpu_col *src, const void *idx, int idx_is_64, int64_t n, void *dst_data,
                           uint8_t *dst_validity) {
    if (vg_ensure_init()) return -1;
    if (n <= 0) return 0;
    if ((reinterpret_cast<uintptr_t>(dst_validity) & 7) != 0) {
        vg_set_error("vgpu_gather: validity buffer must be 8-byte aligned");
        return -1;
    }
    vgpu_col c = *src;
    k_gather<<<vg_blocks(n, 256, 1 << 30), 256, 0, vg_stream()>>>(c, idx, idx_is_64, n, dst_data,
                                                                   reinterpret_cast<uint64_t *>(dst_validity));
    VG_LAUNCH_CHECK();
    VG_TRY(hipStreamSynchronize(vg_stream()));
    return 0;
}

// Pairs of a join on tuple codes whose tuples differ: a pair counts when a
// key column is NA on either side or its values differ (double by ==, so -0
// equals +0; int64 and bool by their bits). Both sides of a column carry the
// pair's common type.
#define VG_MAX_PAIR_KEYS 8
struct PairKeys {
    vgpu_col p[VG_MAX_PAIR_KEYS], b[VG_MAX_PAIR_KEYS];
    int32_t n;
};

__global__ void __launch_bounds__(256) k_pairs_unequal(PairKeys K, const int32_t *prow, const int64_t *brow, int64_t n,
                                                       unsigned long long *bad) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    unsigned long long local = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const int64_t b = brow[i];
        if (b < 0) continue;
        const int64_t p = prow[i];
        bool eq = true;
        for (int j = 0; j < K.n && eq; j++) {
            if (!vg_valid(K.p[j].validity, p) || !vg_valid(K.b[j].validity, b)) {
                eq = false;
                break;
            }
            const uint64_t x = vg_load_bits(K.p[j], p), y = vg_load_bits(K.b[j], b);
            eq = K.p[j].type == VGPU_DOUBLE ? vg_bits_to_f64(x) == vg_bits_to_f64(y) : x == y;
        }
        local += !eq;
    }
    if (local) atomicAdd(bad, local);
}

extern "C" int vgpu_pairs_unequal(const vgpu_col *probe_keys, const vgpu_col *build_keys, int n_keys,
                                  const int32_t *probe_rows, const int64_t *build_rows, int64_t n, int64_t *n_unequal) {
    if (vg_ensure_init()) return -1;
    if (n_keys < 1 || n_keys > VG_MAX_PAIR_KEYS) {
        vg_set_error("pairs_unequal: %d key columns (1..%d)", n_keys, VG_MAX_PAIR_KEYS);
        return -1;
    }
    PairKeys K;
    memset(&K, 0, sizeof K);
    for (int j = 0; j < n_keys; j++) {
        const int t = probe_keys[j].type;
        if (t != build_keys[j].type || (t != VGPU_INT64 && t != VGPU_DOUBLE && t != VGPU_BOOL)) {
            vg_set_error("pairs_unequal: key %d has types %d and %d", j, t, build_keys[j].type);
            return -1;
        }
        K.p[j] = probe_keys[j];
        K.b[j] = build_keys[j];
    }
    K.n = n_keys;
    *n_unequal = 0;
    if (n <= 0) return 0;
    hipStream_t s = vg_stream();
    DevBuf cnt;
    if (cnt.reserve(8)) return -1;
    VG_TRY(hipMemsetAsync(cnt.ptr, 0, 8, s));
    k_pairs_unequal<<<vg_blocks(n, 256, 4096), 256, 0, s>>>(K, probe_rows, build_rows, n,
                                                             cnt.as<unsigned long long>());
    VG_LAUNCH_CHECK();
    unsigned long long h = 0;
    VG_TRY(hipMemcpyAsync(&h, cnt.ptr, 8, hipMemcpyDeviceToHost, s));
    VG_TRY(hipStreamSynchronize(s));
    *n_unequal = (int64_t)h;
    return 0;
}

__global__ void k_cast(vgpu_col src, int64_t n, int dst_type, uint64_t *dst) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint64_t b = vg_load_bits(src, i);   // bool widens to 0/1
    if (dst_type == VGPU_DOUBLE && src.type != VGPU_DOUBLE) b = vg_f64_to_bits((double)(int64_t)b);
    dst[i] = b;
}

extern "C" int vgpu_cast(const vgpu_col *src, int64_t n, int dst_type, void *dst_data) {
    if (vg_ensure_init()) return -1;
    bool ok = (src->type == VGPU_INT64 && dst_type == VGPU_DOUBLE) ||
              (src->type == VGPU_BOOL && (dst_type == VGPU_INT64 || dst_type == VGPU_DOUBLE)) ||
              src->type == dst_type;
    if (!ok || dst_type == VGPU_BOOL) {
        vg_set_error("vgpu_cast: unsupported coercion %d -> %d", src->type, dst_type);
        return -1;
    }
    if (n <= 0) return 0;
    vgpu_col c = *src;
    k_cast<<<vg_blocks(n, 256, 1 << 30), 256, 0, vg_stream()>>>(c, n, dst_type, static_cast<uint64_t *>(dst_data));
    VG_LAUNCH_CHECK();
    VG_TRY(hipStreamSynchronize(vg_stream()));
    return 0;
}
