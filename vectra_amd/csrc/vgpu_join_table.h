// vgpu_join_table.h -- layout of the device join table (shared by the join
// probe kernels and the fused join -> aggregate kernel).
//
// ent[s * E + 0] : key bits of slot s (EMPTY = free); slot `cap` is reserved
//                  for the int64 key whose bits equal the EMPTY sentinel.
// ent[s * E + 1] : (offset << 32) | count of the slot's build rows in rows[];
//                  with embedded payload the count keeps 24 bits and bits
//                  24..31 hold the validity of the embedded values.
// ent[s * E + 2 + p] : payload column p of the slot's FIRST build row, so a
//                  probe of a unique-key table reads key, row count and the
//                  build-side values it aggregates in one 16/32/48-byte load.
// rows[]         : build row ids grouped by key, DESCENDING within a key --
//                  the order in which the reference walks a head-inserted
//                  chain (src/join.c:54-57, 124-131).
// cpay[p][i]     : payload p of rows[i] (only built when some key repeats),
//                  cval[p][i] its validity byte.
#pragma once

#include <mutex>

#include "vgpu_keys.h"

#define JT_MAX_PAYLOAD 4

// Direct-addressed view of a table whose int64 keys are unique and span a
// short range: map[key - kmin] (key - kmin < range) holds the key's
// payload column value minus pmin, JD_NA_PAYLOAD for an NA payload and
// JD_ABSENT where no build key lies.  A probe reads two bytes instead of a
// hash slot's line; for a range of up to 16 keys per slot the map takes at
// most 32 bytes a key (2 a slot).
#define JD_ABSENT 0xFFFFu
#define JD_NA_PAYLOAD 0xFFFEu
struct JDense {
    const uint16_t *map;   // nullptr: not built / not applicable
    int64_t kmin;
    uint64_t range;
    int64_t pmin;
};

struct JView {
    const uint64_t *ent;
    const uint32_t *rows;
    const uint64_t *bkey;    // raw build key bits per build row (8-byte keys)
    const uint64_t *cpay[JT_MAX_PAYLOAD];
    const uint8_t *cval[JT_MAX_PAYLOAD];
    int64_t cap;
    int32_t E;               // entry stride in words
    int32_t n_payload;
    int32_t key_is_double;
    int32_t pad;
    const uint64_t *bloom;   // blocked Bloom filter of the regular build keys (jt_bloom_*)
    uint64_t bloom_mask;     // its words - 1
};

// Blocked Bloom filter over the regular build keys: one 64-bit word per
// key, three bits in it, all from the key's table hash above the bits that
// pick its slot (slots use the low bits; tables stay below 2^24 slots of
// hash for the word index to be independent, beyond that it is only less
// selective).  About 16 filter bits per key: a probe of an absent key
// passes with a probability near 1%, and skips the table's random line
// otherwise.
__host__ __device__ __forceinline__ uint64_t jt_bloom_bits(uint64_t h) {
    return (1ull << ((h >> 46) & 63)) | (1ull << ((h >> 52) & 63)) | (1ull << ((h >> 58) & 63));
}
__host__ __device__ __forceinline__ uint64_t jt_bloom_word(uint64_t h, uint64_t mask) { return (h >> 24) & mask; }

struct vgpu_jtable {
    int32_t key_type = VGPU_INT64;
    int64_t n_build = 0;
    int64_t n_valid = 0;
    int64_t cap = 0;
    int32_t max_dup = 0;
    int32_t E = 2;
    int32_t n_payload = 0;
    vgpu_col payload[JT_MAX_PAYLOAD];
    // value range of every int64 / bool payload column (pay_ranged[p] = 0:
    // a double column, or no valid values)
    int64_t pay_lo[JT_MAX_PAYLOAD] = {0}, pay_hi[JT_MAX_PAYLOAD] = {0};
    int32_t pay_ranged[JT_MAX_PAYLOAD] = {0};
    const uint64_t *bkey = nullptr;   // caller's build key column (kept alive by the caller)
    DevBuf ent, rows, bloom;
    uint64_t bloom_words = 0;
    DevBuf cpay[JT_MAX_PAYLOAD], cval[JT_MAX_PAYLOAD];
    // probe scratch
    DevBuf cnt, un, off_main, off_un, out_probe, out_build, totals, scan_tmp, sort_tmp;
    DevBuf bmatched;       // FULL join: one byte per build row
    int64_t last_n_out = 0;
    // direct-addressed map of one payload column (built on first request)
    DevBuf dmap;
    int32_t dm_state = 0;       // 0 not tried, 1 built, -1 not applicable
    int32_t dm_payload = -1;
    JDense dm = {nullptr, 0, 0, 0};
    std::mutex dm_mu;
    JView view() const {
        JView v;
        memset(&v, 0, sizeof v);
        v.ent = ent.as<uint64_t>();
        v.rows = rows.as<uint32_t>();
        v.bkey = bkey;
        for (int p = 0; p < n_payload; p++) {
            v.cpay[p] = cpay[p].as<uint64_t>();
            v.cval[p] = cval[p].as<uint8_t>();
        }
        v.cap = cap;
        v.E = E;
        v.n_payload = n_payload;
        v.key_is_double = key_type == VGPU_DOUBLE;
        v.bloom = bloom_words ? bloom.as<uint64_t>() : nullptr;
        v.bloom_mask = bloom_words ? bloom_words - 1 : 0;
        return v;
    }
};

// Entry of a classified probe key (vg_key_class), or nullptr: NA and NaN
// keys never match (src/join.c:85, ==).
__device__ __forceinline__ const uint64_t *jt_find_key(const JView &v, int cls, uint64_t kn) {
    if (cls == VG_KEY_NA || cls == VG_KEY_NAN) return nullptr;
    if (cls == VG_KEY_SPECIAL) return v.ent + (size_t)v.cap * v.E;
    const uint64_t empty = v.key_is_double ? VG_EMPTY_F64 : VG_EMPTY_I64;
    uint64_t mask = (uint64_t)v.cap - 1;
    uint64_t s = vg_key_hash(kn) & mask;
    for (int64_t probes = 0; probes < v.cap; probes++) {
        const uint64_t *e = v.ent + s * v.E;
        uint64_t k = e[0];
        if (k == kn) return e;
        if (k == empty) return nullptr;
        s = (s + 1) & mask;
    }
    return nullptr;
}

__device__ __forceinline__ const uint64_t *jt_find(const JView &v, const vgpu_col &kc, int64_t row) {
    uint64_t kn;
    const int cls = vg_key_class(kc, row, kn);
    return jt_find_key(v, cls, kn);
}

__device__ __forceinline__ uint32_t jt_count(const JView &v, uint64_t segw) {
    return v.n_payload ? (uint32_t)(segw & 0xFFFFFFu) : (uint32_t)segw;
}

// match count and first rows[] index of a probe key (0 = no match)
__device__ __forceinline__ uint32_t jt_lookup(const JView &v, const vgpu_col &kc, int64_t row, uint32_t &off) {
    const uint64_t *e = jt_find(v, kc, row);
    off = 0;
    if (!e) return 0;
    uint64_t sg = e[1];
    off = (uint32_t)(sg >> 32);
    return jt_count(v, sg);
}

// The direct-addressed map of payload column p (built on the first request,
// under the table's lock, on the caller's stream, which is synchronized
// before the map is shared; one map per table, for the first payload column
// asked for); out->map == nullptr when the table does not take one: keys not int64, a repeated key, an int64 key equal to the
// empty sentinel, a key range above 16 per key or 2^30, a payload not
// int64 / bool or spanning more than 65,533 values, or knob join.dense_map
// off.  Returns -1 with the error set on a device failure.
int vg_jtable_dense(vgpu_jtable *t, int p, JDense *out);
