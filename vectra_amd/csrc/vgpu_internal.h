// vgpu_internal.h -- shared host/device helpers of libvgpu (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#include <new>

#include "vgpu.h"

// ---------------------------------------------------------------------------
// host side: error channel, stream, growable device buffers
// ---------------------------------------------------------------------------

void vg_set_error(const char *fmt, ...) __attribute__((format(printf, 1, 2)));

// C entry points whose work uses the C++ library (containers, new) run it
// through vg_guard: an exception becomes the library error and the entry
// point's failure value instead of unwinding through C callers.
template <typename R, typename F>
static inline R vg_guard(R failed, const char *where, F &&body) {
    try {
        return body();
    } catch (const std::bad_alloc &) {
        vg_set_error("%s: host memory exhausted", where);
    } catch (...) {
        vg_set_error("%s: unexpected C++ exception", where);
    }
    return failed;
}

hipStream_t vg_stream();
// the pool's free blocks of a stream being destroyed move to the library stream
void vg_pool_forget_stream(hipStream_t s);
int vg_ensure_init();
// bracket a main kernel launch for vgpu_kernel_time_ms (no-ops unless enabled)
void vg_kt_begin();
void vg_kt_end();

// launch-shape knobs (see vgpu_tune_set)
enum { VG_TUNE_GB_LDS_MAX_BLOCKS, VG_TUNE_GB_LDS_ROWS, VG_TUNE_GB_GLB_MAX_BLOCKS, VG_TUNE_GB_GLB_ROWS,
       VG_TUNE_GB_LDS_MAX_SLOTS, VG_TUNE_GB_LDS_BUDGET, VG_TUNE_GB_STAR_ITEMS, VG_TUNE_GB_PART_MIN_GROUPS,
       VG_TUNE_BT_SHARED, VG_TUNE_GB_LDS_SLOTS_PER_GROUP, VG_TUNE_BT_BUCKETED, VG_TUNE_GB_STAR_DENSE, VG_TUNE_JOIN_BLOOM, VG_TUNE_GB_LDS_COPIES, VG_TUNE_GB_NARROW, VG_TUNE_BT_DENSE, VG_TUNE_SCAN_FEED, VG_TUNE_JOIN_BLOOM_KPW, VG_TUNE_JOIN_LOAD_PCT, VG_TUNE_GB_FUSE_JOIN, VG_TUNE_SD_HASH_BITS, VG_TUNE_GB_GLB_LOAD_PCT, VG_TUNE_JOIN_DENSE, VG_TUNE_GB_STAR_BPC, VG_TUNE_GB_PART_RANGE, VG_TUNE_GB_PART_STAGE, VG_TUNE_GB_PART_SLOTS, VG_TUNE_GB_FEED_PF, VG_TUNE_GB_MERGE_PARTIALS, VG_TUNE_GB_PART_AGG, VG_TUNE_GB_PART_FILL, VG_TUNE_GB_FEED_DENSE, VG_TUNE_GB_PART_NOMETA, VG_TUNE_COUNT };
int64_t vg_tune(int knob);
// the knob for a group-by expecting `groups` groups (tuning table buckets;
// the bucket's other two axes come from the innermost VgTuneScope)
int64_t vg_tune_for(int knob, int64_t groups);
// the tuning context of the work a thread is launching: record width in
// 8-byte words and rows of the launch (0 = unknown, matches any bucket)
struct VgTuneScope {
    int64_t saved_width, saved_rows;
    VgTuneScope(int64_t width, int64_t rows);
    ~VgTuneScope();
};

#define VG_TRY(call)                                                          \
    do {                                                                      \
        hipError_t e_ = (call);                                               \
        if (e_ != hipSuccess) {                                               \
            vg_set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), \
                         __FILE__, __LINE__);                                 \
            return -1;                                                        \
        }                                                                     \
    } while (0)

#define VG_TRY_NULL(call)                                                     \
    do {                                                                      \
        hipError_t e_ = (call);                                               \
        if (e_ != hipSuccess) {                                               \
            vg_set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), \
                         __FILE__, __LINE__);                                 \
            return nullptr;                                                   \
        }                                                                     \
    } while (0)

#define VG_LAUNCH_CHECK()                                                     \
    do {                                                                      \
        hipError_t e_ = hipGetLastError();                                    \
        if (e_ != hipSuccess) {                                               \
            vg_set_error("kernel launch failed: %s (%s:%d)",                 \
                         hipGetErrorString(e_), __FILE__, __LINE__);          \
            return -1;                                                        \
        }                                                                     \
    } while (0)

// Caching device allocator: blocks are power-of-two sized (>= 4 KiB) and
// returned to per-size free lists instead of hipFree, so a steady stream of
// queries allocates nothing and never takes hipFree's implicit device
// synchronization.  vg_pool_trim() hands cached blocks back to the runtime.
void *vg_pool_alloc(size_t bytes, size_t *granted);
// the library's pinned host staging buffer for small device -> host results
// (at least `bytes`, at most VG_STAGE_MAX; nullptr when larger or when the
// allocation fails); valid until the next call, library stream only
#define VG_STAGE_MAX ((size_t)4 << 20)
void *vg_pinned_stage(size_t bytes);
void vg_pool_free(void *ptr, size_t bytes);
void vg_pool_trim();

// A device buffer that only grows; reused across batches so the steady
// state allocates nothing.
struct DevBuf {
    void  *ptr = nullptr;
    size_t cap = 0;
    int reserve(size_t bytes) {
        if (bytes <= cap) return 0;
        if (ptr) {
            vg_pool_free(ptr, cap);
            ptr = nullptr;
            cap = 0;
        }
        size_t got = 0;
        ptr = vg_pool_alloc(bytes, &got);
        if (!ptr) {
            vg_set_error("device allocation of %zu bytes failed", bytes);
            return -1;
        }
        cap = got;
        return 0;
    }
    template <class T> T *as() const { return static_cast<T *>(ptr); }
    void release() {
        if (ptr) vg_pool_free(ptr, cap);
        ptr = nullptr;
        cap = 0;
    }
};

// device-wide exclusive scan of n uint32 counts into uint64 offsets;
// *total (device, may be null) gets the sum.  tmp is scratch.
int vg_scan_u32_to_u64(const uint32_t *in, uint64_t *out, int64_t n,
                       uint64_t *total, DevBuf &tmp);
// the same over caller-provided scratch of tmp_bytes bytes (device): no
// allocation; fails when vg_scan_tmp_bytes(n) exceeds tmp_bytes
int vg_scan_u32_to_u64_in(const uint32_t *in, uint64_t *out, int64_t n,
                          uint64_t *total, void *tmp, size_t tmp_bytes);
size_t vg_scan_tmp_bytes(int64_t n);
// per-device set-up of the no-allocation filter (vgpu_filter.hip), called by
// vgpu_init with the device current
int vg_filter_noalloc_setup(int device);
// device-wide exclusive scan of n uint64 values (in may equal out)
int vg_scan_u64(const uint64_t *in, uint64_t *out, int64_t n, uint64_t *total,
                DevBuf &tmp);
// stable LSD radix sort of (key, value) pairs on bits [begin_bit, end_bit)
int vg_sort_pairs(uint64_t *keys, uint32_t *vals, int64_t n, int begin_bit,
                  int end_bit, DevBuf &tmp);

int vg_cus();
int64_t vg_resident_blocks(size_t lds_bytes_per_block);

// blocks for n items at per_block each, at most cap (kernels with one thread
// per item rely on the default, which covers any grid the hardware takes)
static inline unsigned vg_blocks(int64_t n, int64_t per_block, int64_t cap = 0x7FFFFFFF) {
    int64_t b = (n + per_block - 1) / per_block;
    if (b < 1) b = 1;
    if (b > cap) b = cap;
    return (unsigned)b;
}

// ---------------------------------------------------------------------------
// device side
// ---------------------------------------------------------------------------

#define VG_WAVE 64

__device__ __forceinline__ int vg_lane() { return __lane_id(); }

// number of set bits of mask strictly below this lane
__device__ __forceinline__ int vg_lane_prefix(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                     __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

__device__ __forceinline__ uint64_t vg_ballot(bool p) { return __ballot(p); }

// Column data and validity bitmaps always live in device (global) memory.
// They are read through a global address-space pointer, so a column whose
// descriptor was itself loaded from memory (a batch's query table) still
// gets global loads: through a plain pointer the compiler cannot rule out
// LDS and emits flat loads, which also count against the LDS wait counter.
template <class T>
__device__ __forceinline__ T vg_gload(const void *p, int64_t i) {
    return ((const __attribute__((address_space(1))) T *)p)[i];
}

__device__ __forceinline__ bool vg_valid(const uint8_t *validity, int64_t row) {
    return validity == nullptr || ((vg_gload<uint8_t>(validity, row >> 3) >> (row & 7)) & 1);
}

// raw 8-byte view of a row of an int64/double column, or the 0/1 byte of a
// bool column widened to 8 bytes
__device__ __forceinline__ uint64_t vg_load_bits(const vgpu_col &c, int64_t row) {
    if (c.type == VGPU_BOOL) return (uint64_t)vg_gload<uint8_t>(c.data, row);
    return vg_gload<uint64_t>(c.data, row);
}

// host overloads of the two (the expression VM is also built for the host,
// where columns are host arrays)
__host__ inline bool vg_valid(const uint8_t *validity, int64_t row) {
    return validity == nullptr || ((validity[row >> 3] >> (row & 7)) & 1);
}

__host__ inline uint64_t vg_load_bits(const vgpu_col &c, int64_t row) {
    if (c.type == VGPU_BOOL) return (uint64_t)static_cast<const uint8_t *>(c.data)[row];
    return static_cast<const uint64_t *>(c.data)[row];
}

// the same for a column streamed once: the load is marked non-temporal so
// its lines do not push the probed table's hot lines out of the caches
__device__ __forceinline__ uint64_t vg_load_bits_stream(const vgpu_col &c, int64_t row) {
    if (c.type == VGPU_BOOL) return (uint64_t)static_cast<const uint8_t *>(c.data)[row];
    return __builtin_nontemporal_load(static_cast<const uint64_t *>(c.data) + row);
}

__host__ __device__ __forceinline__ double vg_bits_to_f64(uint64_t b) { return __builtin_bit_cast(double, b); }
__host__ __device__ __forceinline__ uint64_t vg_f64_to_bits(double d) { return __builtin_bit_cast(uint64_t, d); }

// order-preserving map of a double's bits onto unsigned integers
// (-0 maps just below +0; only used for non-NaN values)
__device__ __forceinline__ uint64_t vg_f64_ordered(uint64_t bits) {
    return (bits >> 63) ? ~bits : (bits | 0x8000000000000000ull);
}
__device__ __host__ __forceinline__ uint64_t vg_f64_unordered(uint64_t o) {
    return (o >> 63) ? (o & 0x7FFFFFFFFFFFFFFFull) : ~o;
}
// signed int64 onto unsigned order
__device__ __host__ __forceinline__ uint64_t vg_i64_ordered(uint64_t bits) {
    return bits ^ 0x8000000000000000ull;
}

// 64-bit finalizer (Stafford variant 13 of the MurmurHash3 mixer)
__device__ __host__ __forceinline__ uint64_t vg_mix64(uint64_t x) {
    x ^= x >> 30;
    x *= 0xbf58476d1ce4e5b9ull;
    x ^= x >> 27;
    x *= 0x94d049bb133111ebull;
    x ^= x >> 31;
    return x;
}

template <class T>
__device__ __forceinline__ T vg_atomic_load(const T *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
