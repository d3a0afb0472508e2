// vgpu_keys.h -- key normalization shared by the group-by and join tables.
//
// Both tables are open-addressing arrays of 8-byte key bits claimed with a
// 64-bit compare-and-swap against an EMPTY sentinel that no regular key can
// take:
//   int64 / bool keys : EMPTY = INT64_MIN's bits; a real INT64_MIN key goes
//                       to a reserved slot instead of the hashed array;
//   double keys       : EMPTY = all-ones (a NaN); NaN keys never reach the
//                       hashed array (they never compare equal), and -0 is
//                       folded onto +0 because -0 == +0 (src/hash.c:23-24,
//                       src/hash.c:64).
#pragma once

#include "vgpu_internal.h"

#define VG_EMPTY_I64 0x8000000000000000ull
#define VG_EMPTY_F64 0xFFFFFFFFFFFFFFFFull

enum { VG_KEY_REGULAR = 0, VG_KEY_NA = 1, VG_KEY_NAN = 2, VG_KEY_SPECIAL = 3 };

// class and normalized bits of a key value already loaded
__device__ __forceinline__ int vg_key_class_bits(int type, bool valid, uint64_t b, uint64_t &kn) {
    if (!valid) { kn = 0; return VG_KEY_NA; }
    if (type == VGPU_DOUBLE) {
        if ((b & 0x7FFFFFFFFFFFFFFFull) > 0x7FF0000000000000ull) { kn = b; return VG_KEY_NAN; }
        if (b == 0x8000000000000000ull) b = 0;
        kn = b;
        return VG_KEY_REGULAR;
    }
    kn = b;
    return b == VG_EMPTY_I64 ? VG_KEY_SPECIAL : VG_KEY_REGULAR;
}

__device__ __forceinline__ int vg_key_class(const vgpu_col &c, int64_t row, uint64_t &kn) {
    bool v = vg_valid(c.validity, row);
    return vg_key_class_bits(c.type, v, v ? vg_load_bits(c, row) : 0, kn);
}

__device__ __forceinline__ uint64_t vg_key_hash(uint64_t kn) { return vg_mix64(kn ^ 0x9e3779b97f4a7c15ull); }

// min and max over the valid values of an int64 / bool column, as
// order-preserving unsigned maps (bits ^ sign): mm[0] takes the min by
// atomicMin, mm[1] the max by atomicMax (start them at ~0 and 0); a column
// without valid values leaves mm[0] > mm[1]
static __global__ void k_key_minmax(vgpu_col c, int64_t n, unsigned long long *mm) {
    __shared__ unsigned long long lo, hi;
    if (threadIdx.x == 0) {
        lo = ~0ull;
        hi = 0;
    }
    __syncthreads();
    unsigned long long a = ~0ull, z = 0;
    for (int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; row < n;
         row += (int64_t)gridDim.x * blockDim.x) {
        if (!vg_valid(c.validity, row)) continue;
        const unsigned long long o = vg_load_bits(c, row) ^ 0x8000000000000000ull;
        a = o < a ? o : a;
        z = o > z ? o : z;
    }
    if (a != ~0ull) atomicMin(&lo, a);
    if (z) atomicMax(&hi, z);
    __syncthreads();
    if (threadIdx.x == 0) {
        if (lo != ~0ull) atomicMin(&mm[0], lo);
        if (hi) atomicMax(&mm[1], hi);
    }
}
