// vgpu_join_table.h -- layout of the device join table (shared by the join
// probe kernels and the fused join -> aggregate kernel).
//
// keys[s]  : key bits of slot s (EMPTY = free); slot `cap` is reserved for
//            the int64 key whose bits equal the EMPTY sentinel.
// seg[s]   : (offset << 32) | count of the slot's build rows in rows[].
// rows[]   : build row ids grouped by key, DESCENDING within a key -- the
//            order in which the reference walks a head-inserted chain
//            (src/join.c:54-57, 124-131).
#pragma once

#include "vgpu_keys.h"

struct JView {
    const uint64_t *keys;
    const uint64_t *seg;
    const uint32_t *rows;
    int64_t cap;
    int32_t key_is_double;
    int32_t pad;
};

struct vgpu_jtable {
    int32_t key_type = VGPU_INT64;
    int64_t n_build = 0;
    int64_t n_valid = 0;
    int64_t cap = 0;
    int32_t max_dup = 0;
    DevBuf keys, seg, rows;
    // probe scratch
    DevBuf cnt, un, off_main, off_un, out_probe, out_build, totals, scan_tmp, sort_tmp;
    DevBuf bmatched;       // FULL join: one byte per build row
    int64_t last_n_out = 0;
    JView view() const {
        JView v;
        v.keys = keys.as<uint64_t>();
        v.seg = seg.as<uint64_t>();
        v.rows = rows.as<uint32_t>();
        v.cap = cap;
        v.key_is_double = key_type == VGPU_DOUBLE;
        v.pad = 0;
        return v;
    }
};

// Looks up a probe key; returns the match count (0 = none) and the first
// index into rows[].  NA and NaN keys never match (src/join.c:85, ==).
__device__ __forceinline__ uint32_t jt_lookup(const JView &v, const vgpu_col &kc, int64_t row, uint32_t &off) {
    uint64_t kn;
    int cls = vg_key_class(kc, row, kn);
    off = 0;
    if (cls == VG_KEY_NA || cls == VG_KEY_NAN) return 0;
    uint64_t s;
    if (cls == VG_KEY_SPECIAL) {
        s = (uint64_t)v.cap;
    } else {
        const uint64_t empty = v.key_is_double ? VG_EMPTY_F64 : VG_EMPTY_I64;
        uint64_t mask = (uint64_t)v.cap - 1;
        s = vg_key_hash(kn) & mask;
        for (int64_t probes = 0;; probes++) {
            uint64_t k = v.keys[s];
            if (k == kn) break;
            if (k == empty || probes >= v.cap) return 0;
            s = (s + 1) & mask;
        }
    }
    uint64_t sg = v.seg[s];
    off = (uint32_t)(sg >> 32);
    return (uint32_t)sg;
}
