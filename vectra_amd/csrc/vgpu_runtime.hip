// vgpu_runtime.hip -- device setup, error channel, memory entry points,
// device-wide scan and the stable LSD radix sort used for result ordering
// and join builds.
#include "vgpu_internal.h"

#include <algorithm>

#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <mutex>
#include <vector>

// ---------------------------------------------------------------------------
// error channel / device state
// ---------------------------------------------------------------------------

static thread_local char g_err[1024];
static hipStream_t g_stream = nullptr;
static int g_device = -1;
static std::mutex g_init_mu;
static hipEvent_t g_t0 = nullptr, g_t1 = nullptr;

void vg_set_error(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
}

extern "C" const char *vgpu_last_error(void) { return g_err; }

// a thread's library calls run on the stream it selected (vgpu_stream_use),
// else on the library stream
static thread_local hipStream_t t_stream = nullptr;

hipStream_t vg_stream() { return t_stream ? t_stream : g_stream; }

extern "C" void *vgpu_stream_create(void) {
    if (vg_ensure_init()) return nullptr;
    hipStream_t s = nullptr;
    hipError_t e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);   // default priority
    if (e != hipSuccess) {
        vg_set_error("vgpu_stream_create: %s", hipGetErrorString(e));
        return nullptr;
    }
    return s;
}

static int vg_h2d_order(hipStream_t st);

extern "C" int vgpu_stream_use(void *stream) {
    if (vg_ensure_init()) return -1;
    t_stream = static_cast<hipStream_t>(stream);
    // work on the selected stream comes after the staged host -> device
    // copies this process has issued (vgpu_memcpy_h2d does not wait for them)
    return vg_h2d_order(vg_stream());
}

extern "C" int vgpu_stream_destroy(void *stream) {
    if (!stream) return 0;
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (t_stream == s) t_stream = nullptr;
    VG_TRY(hipStreamSynchronize(s));
    vg_pool_forget_stream(s);
    VG_TRY(hipStreamDestroy(s));
    return 0;
}

extern "C" int vgpu_error(const char *msg) {
    vg_set_error("%s", msg ? msg : "");
    return -1;
}

extern "C" int vgpu_mem_info(int64_t *free_bytes, int64_t *total_bytes) {
    if (vg_ensure_init()) return -1;
    size_t f = 0, t = 0;
    VG_TRY(hipMemGetInfo(&f, &t));
    if (free_bytes) *free_bytes = (int64_t)f;
    if (total_bytes) *total_bytes = (int64_t)t;
    return 0;
}

extern "C" int vgpu_tune_load(const char *path);

extern "C" int vgpu_init(int device) {
    std::lock_guard<std::mutex> lk(g_init_mu);
    if (g_device == device && g_stream) return 0;
    int n = 0;
    VG_TRY(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) {
        vg_set_error("vgpu_init: device %d not present (%d devices)", device, n);
        return -1;
    }
    VG_TRY(hipSetDevice(device));
    hipDeviceProp_t prop;
    VG_TRY(hipGetDeviceProperties(&prop, device));
    if (strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
        vg_set_error("vgpu_init: libvgpu is built for gfx950, device %d is %s",
                     device, prop.gcnArchName);
        return -1;
    }
    if (!g_stream) VG_TRY(hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking));
    if (vg_filter_noalloc_setup(device)) return -1;
    // an autotuner table named by the environment applies from the start
    const char *tuning = getenv("VECTRA_GPU_TUNING");
    if (tuning && *tuning && vgpu_tune_load(tuning) < 0) return -1;
    if (!g_t0) VG_TRY(hipEventCreate(&g_t0));
    if (!g_t1) VG_TRY(hipEventCreate(&g_t1));
    g_device = device;
    return 0;
}

int vg_ensure_init() {
    if (g_stream) return 0;
    return vgpu_init(0);
}

// compute units of the current device (256 on MI355X)
int vg_cus() {
    static int cus = 0;
    if (!cus) {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
            cus = 256;
    }
    return cus;
}

// blocks of `lds` bytes and 256 threads that fit on the device at once
// (160 KB of LDS and 8 such blocks per CU)
int64_t vg_resident_blocks(size_t lds) {
    int64_t per_cu = 8;
    if (lds > 0) per_cu = std::min<int64_t>(8, (int64_t)((160 * 1024) / lds));
    return (int64_t)vg_cus() * std::max<int64_t>(per_cu, 1);
}

// ---------------------------------------------------------------------------
// caching allocator
// ---------------------------------------------------------------------------

static std::mutex g_pool_mu;

// Free blocks by size class (2^k bytes), kept per stream: a block released
// while work enqueued on stream S may still read it is handed out again
// only to later work on S, which runs after that work.  (With one stream
// this is the plain caching allocator.)
struct PoolSet {
    hipStream_t stream;
    std::vector<void *> cls[64];
};
static std::vector<PoolSet *> g_pools;

static PoolSet *pool_for(hipStream_t s) {   // g_pool_mu held
    for (PoolSet *p : g_pools)
        if (p->stream == s) return p;
    PoolSet *p = new PoolSet();
    p->stream = s;
    g_pools.push_back(p);
    return p;
}

static int pool_class(size_t bytes, size_t *rounded) {
    int k = 12;
    while (((size_t)1 << k) < bytes) k++;
    *rounded = (size_t)1 << k;
    return k;
}

void *vg_pool_alloc(size_t bytes, size_t *granted) {
    size_t r;
    int k = pool_class(bytes ? bytes : 1, &r);
    {
        std::lock_guard<std::mutex> lk(g_pool_mu);
        PoolSet *ps = pool_for(vg_stream());
        if (!ps->cls[k].empty()) {
            void *p = ps->cls[k].back();
            ps->cls[k].pop_back();
            *granted = r;
            return p;
        }
    }
    void *p = nullptr;
    if (hipMalloc(&p, r) != hipSuccess) {
        // give cached blocks back and retry once
        vg_pool_trim();
        if (hipMalloc(&p, r) != hipSuccess) return nullptr;
    }
    *granted = r;
    return p;
}

void vg_pool_free(void *ptr, size_t bytes) {
    if (!ptr) return;
    size_t r;
    int k = pool_class(bytes ? bytes : 1, &r);
    std::lock_guard<std::mutex> lk(g_pool_mu);
    try {
        pool_for(vg_stream())->cls[k].push_back(ptr);
    } catch (const std::bad_alloc &) {
        // no room to cache it: give the block back (hipFree waits for the
        // device work that may still read it)
        hipError_t e = hipFree(ptr);
        (void)e;
    }
}

// a stream about to be destroyed (and synchronized): its free blocks are
// safe for any stream now and move to the library stream's lists
void vg_pool_forget_stream(hipStream_t s) {
    std::lock_guard<std::mutex> lk(g_pool_mu);
    for (size_t i = 0; i < g_pools.size(); i++) {
        PoolSet *p = g_pools[i];
        if (p->stream != s) continue;
        PoolSet *dst = pool_for(g_stream);
        for (int k = 0; k < 64; k++) dst->cls[k].insert(dst->cls[k].end(), p->cls[k].begin(), p->cls[k].end());
        g_pools.erase(g_pools.begin() + (long)i);
        delete p;
        return;
    }
}

void vg_pool_trim() {
    std::lock_guard<std::mutex> lk(g_pool_mu);
    for (PoolSet *ps : g_pools)
        for (auto &v : ps->cls) {
            for (void *p : v) {
                hipError_t e = hipFree(p);
                (void)e;
            }
            v.clear();
        }
}

extern "C" void *vgpu_malloc(size_t bytes) {
    if (vg_ensure_init()) return nullptr;
    void *p = nullptr;
    VG_TRY_NULL(hipMalloc(&p, bytes ? bytes : 1));
    return p;
}

extern "C" int vgpu_free(void *p) {
    if (!p) return 0;
    VG_TRY(hipFree(p));
    return 0;
}

// Host -> device copies of pageable memory go through a ring of pinned
// slots on the caller's library stream: the host copies a piece of the
// source into a slot (first waiting for the device copy that last read
// that slot) and the stream copies the slot to the device.  The call
// returns once the whole source has been read, without waiting for the
// device: later work on the same stream sees the data, the source may be
// reused at once, and the host's copy of the next piece overlaps the
// device copy of this one.  (A pageable hipMemcpyAsync followed by a
// stream synchronization cost a round trip per call and ran at a fraction
// of the link's rate.)
#define VG_H2D_SLOTS 4
#define VG_H2D_SLOT_BYTES ((size_t)4 << 20)
static std::mutex g_h2d_mu;
static uint8_t *g_h2d_buf[VG_H2D_SLOTS];
static hipEvent_t g_h2d_ev[VG_H2D_SLOTS];
static bool g_h2d_busy[VG_H2D_SLOTS];
static int g_h2d_next = 0;

static int vg_h2d_staged(void *dst, const void *src, size_t bytes) {
    std::lock_guard<std::mutex> lk(g_h2d_mu);
    if (!g_h2d_buf[0]) {
        for (int k = 0; k < VG_H2D_SLOTS; k++) {
            void *p = nullptr;
            VG_TRY(hipHostMalloc(&p, VG_H2D_SLOT_BYTES, hipHostMallocDefault));
            g_h2d_buf[k] = static_cast<uint8_t *>(p);
            VG_TRY(hipEventCreateWithFlags(&g_h2d_ev[k], hipEventDisableTiming));
        }
    }
    hipStream_t st = vg_stream();
    for (size_t off = 0; off < bytes; off += VG_H2D_SLOT_BYTES) {
        const size_t len = std::min(VG_H2D_SLOT_BYTES, bytes - off);
        const int k = g_h2d_next;
        g_h2d_next = (k + 1) % VG_H2D_SLOTS;
        if (g_h2d_busy[k]) VG_TRY(hipEventSynchronize(g_h2d_ev[k]));
        memcpy(g_h2d_buf[k], static_cast<const uint8_t *>(src) + off, len);
        VG_TRY(hipMemcpyAsync(static_cast<uint8_t *>(dst) + off, g_h2d_buf[k], len, hipMemcpyHostToDevice, st));
        VG_TRY(hipEventRecord(g_h2d_ev[k], st));
        g_h2d_busy[k] = true;
    }
    return 0;
}

// st waits (on the device) for every staged copy still in flight
static int vg_h2d_order(hipStream_t st) {
    std::lock_guard<std::mutex> lk(g_h2d_mu);
    for (int k = 0; k < VG_H2D_SLOTS; k++)
        if (g_h2d_busy[k]) VG_TRY(hipStreamWaitEvent(st, g_h2d_ev[k], 0));
    return 0;
}

extern "C" int vgpu_memcpy_h2d(void *dst, const void *src, size_t bytes) {
    if (vg_ensure_init()) return -1;
    if (!bytes) return 0;
    return vg_h2d_staged(dst, src, bytes);
}

extern "C" int vgpu_memcpy_d2h(void *dst, const void *src, size_t bytes) {
    if (vg_ensure_init()) return -1;
    if (!bytes) return 0;
    VG_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, vg_stream()));
    VG_TRY(hipStreamSynchronize(vg_stream()));
    return 0;
}

extern "C" int vgpu_memset(void *dst, int value, size_t bytes) {
    if (vg_ensure_init()) return -1;
    if (!bytes) return 0;
    VG_TRY(hipMemsetAsync(dst, value, bytes, vg_stream()));
    VG_TRY(hipStreamSynchronize(vg_stream()));
    return 0;
}

extern "C" int vgpu_synchronize(void) {
    if (vg_ensure_init()) return -1;
    VG_TRY(hipStreamSynchronize(vg_stream()));
    return 0;
}

static void *g_stage = nullptr;
static size_t g_stage_cap = 0;

void *vg_pinned_stage(size_t bytes) {
    if (bytes > VG_STAGE_MAX) return nullptr;
    if (bytes <= g_stage_cap) return g_stage;
    size_t want = g_stage_cap ? g_stage_cap : 65536;
    while (want < bytes) want *= 2;
    void *p = nullptr;
    if (hipHostMalloc(&p, want, hipHostMallocDefault) != hipSuccess) return nullptr;
    if (g_stage) (void)hipHostFree(g_stage);   // no copy into the old buffer is pending: callers synchronize
    g_stage = p;
    g_stage_cap = want;
    return g_stage;
}

extern "C" void *vgpu_host_alloc(size_t bytes) {
    if (vg_ensure_init()) return nullptr;
    void *p = nullptr;
    VG_TRY_NULL(hipHostMalloc(&p, bytes ? bytes : 1, hipHostMallocDefault));
    return p;
}

extern "C" int vgpu_host_free(void *p) {
    if (!p) return 0;
    VG_TRY(hipHostFree(p));
    return 0;
}

// ---- asynchronous uploads ---------------------------------------------
// A second stream (default priority, like the library stream) carries the
// host -> device copies of staged inputs; per slot, one event marks the
// slot's copies done and one marks the library-stream work that last read
// the slot's device buffers.
#define VG_UPLOAD_SLOTS 4
static hipStream_t g_copy = nullptr;
static hipEvent_t g_up_done[VG_UPLOAD_SLOTS], g_up_read[VG_UPLOAD_SLOTS];
static bool g_up_read_set[VG_UPLOAD_SLOTS];

static int vg_upload_slot(int slot) {
    if (vg_ensure_init()) return -1;
    if (slot < 0 || slot >= VG_UPLOAD_SLOTS) {
        vg_set_error("upload slot %d out of range (0..%d)", slot, VG_UPLOAD_SLOTS - 1);
        return -1;
    }
    if (!g_copy) {
        VG_TRY(hipStreamCreateWithFlags(&g_copy, hipStreamNonBlocking));
        for (int s = 0; s < VG_UPLOAD_SLOTS; s++) {
            VG_TRY(hipEventCreateWithFlags(&g_up_done[s], hipEventDisableTiming));
            VG_TRY(hipEventCreateWithFlags(&g_up_read[s], hipEventDisableTiming));
        }
    }
    return 0;
}

extern "C" int vgpu_upload_async(int slot, void *dst, const void *src, size_t bytes) {
    if (vg_upload_slot(slot)) return -1;
    if (g_up_read_set[slot]) VG_TRY(hipStreamWaitEvent(g_copy, g_up_read[slot], 0));
    if (bytes) VG_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, g_copy));
    VG_TRY(hipEventRecord(g_up_done[slot], g_copy));
    return 0;
}

extern "C" int vgpu_upload_acquire(int slot) {
    if (vg_upload_slot(slot)) return -1;
    VG_TRY(hipStreamWaitEvent(vg_stream(), g_up_done[slot], 0));
    return 0;
}

extern "C" int vgpu_upload_release(int slot) {
    if (vg_upload_slot(slot)) return -1;
    VG_TRY(hipEventRecord(g_up_read[slot], vg_stream()));
    g_up_read_set[slot] = true;
    return 0;
}

extern "C" int vgpu_upload_host_wait(int slot) {
    if (vg_upload_slot(slot)) return -1;
    VG_TRY(hipEventSynchronize(g_up_done[slot]));
    return 0;
}

extern "C" int vgpu_timer_start(void) {
    if (vg_ensure_init()) return -1;
    VG_TRY(hipEventRecord(g_t0, vg_stream()));
    return 0;
}

extern "C" double vgpu_timer_stop(void) {
    if (hipEventRecord(g_t1, vg_stream()) != hipSuccess) return -1.0;
    if (hipEventSynchronize(g_t1) != hipSuccess) return -1.0;
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, g_t0, g_t1) != hipSuccess) return -1.0;
    return (double)ms;
}

// launch-shape knobs: global values, and per cardinality bucket overrides
// (the autotuner's table: the first bucket whose max_groups covers a
// group-by's expected groups and sets a knob decides it)
static const char *g_tune_names[VG_TUNE_COUNT] = {"gb.lds.max_blocks", "gb.lds.rows_per_block", "gb.glb.max_blocks",
                                                  "gb.glb.rows_per_block", "gb.lds.max_slots", "gb.lds.budget",
                                                  "gb.star.items", "gb.part.min_groups", "batch.shared_scan",
                                                  "gb.lds.slots_per_group", "batch.bucketed",
                                                  "gb.star.dense", "join.bloom",
                                                  "gb.lds.copies", "gb.narrow", "batch.dense", "scan.device_feed", "join.bloom_keys", "join.load_pct", "gb.fuse_join",
                                                  "strdict.hash_bits", "gb.glb.load_pct",
                                                  "join.dense_map", "gb.star.blocks_per_cu", "gb.part.range",
                                                  "gb.part.stage", "gb.part.slots", "gb.feed.prefetch",
                                                  "gb.merge.partials", "gb.part.agg", "gb.part.fill_pct", "gb.feed.dense", "gb.part.nometa"};
#define VG_TUNE_DEFAULTS {1024, 8192, 16384, 1024, 4096, 65536, 4, 65536, 1, 2, 2, 2, 1, 2, 2, 2, 2, 4, 50, 2, 64, 50, 2, 4, 2, 3, 512, 2, 2, 3, 50, 2, 2}
static int64_t g_tune[VG_TUNE_COUNT] = VG_TUNE_DEFAULTS;

// A bucket covers groups <= max_groups, record widths <= max_width (8-byte
// words) and launches of rows <= max_rows; lookups take the first bucket
// that covers the work and sets the knob, in ascending (max_width,
// max_rows, max_groups) order: the narrowest width class first, then the
// smallest row class, then the smallest cardinality class.
#define TUNE_ANY INT64_MAX
struct TuneBucket {
    int64_t max_groups, max_width, max_rows;
    int64_t val[VG_TUNE_COUNT];
    uint32_t set;   // bit i: knob i overridden
};
static std::vector<TuneBucket> g_buckets;
static std::mutex g_tune_mu;

static bool bucket_before(const TuneBucket &a, const TuneBucket &b) {
    if (a.max_width != b.max_width) return a.max_width < b.max_width;
    if (a.max_rows != b.max_rows) return a.max_rows < b.max_rows;
    return a.max_groups < b.max_groups;
}

static thread_local int64_t t_tune_width = 0, t_tune_rows = 0;

VgTuneScope::VgTuneScope(int64_t width, int64_t rows) : saved_width(t_tune_width), saved_rows(t_tune_rows) {
    t_tune_width = width;
    t_tune_rows = rows;
}
VgTuneScope::~VgTuneScope() {
    t_tune_width = saved_width;
    t_tune_rows = saved_rows;
}

int64_t vg_tune(int knob) { return g_tune[knob]; }

static int64_t tune_at(int knob, int64_t groups, int64_t width, int64_t rows) {
    if (groups > 0)
        for (const TuneBucket &b : g_buckets)
            if (groups <= b.max_groups && width <= b.max_width && rows <= b.max_rows && (b.set & (1u << knob)))
                return b.val[knob];
    return g_tune[knob];
}

int64_t vg_tune_for(int knob, int64_t groups) { return tune_at(knob, groups, t_tune_width, t_tune_rows); }

static int tune_index(const char *name) {
    for (int i = 0; i < VG_TUNE_COUNT; i++)
        if (strcmp(name, g_tune_names[i]) == 0) return i;
    return -1;
}

static int tune_check(int i, const char *name, int64_t value) {
    if (i < 0) {
        vg_set_error("vgpu_tune_set: unknown knob %s", name);
        return -1;
    }
    if (value <= 0) {
        vg_set_error("vgpu_tune_set: %s must be positive", name);
        return -1;
    }
    if (i == VG_TUNE_GB_LDS_BUDGET && value > 65536) {
        vg_set_error("vgpu_tune_set: %s is at most 65536 bytes", name);
        return -1;
    }
    if (i == VG_TUNE_GB_PART_SLOTS && (value < 64 || value > 512 || (value & (value - 1)))) {
        vg_set_error("vgpu_tune_set: %s is a power of two from 64 to 512", name);
        return -1;
    }
    if (i == VG_TUNE_GB_LDS_SLOTS_PER_GROUP && (value < 2 || value > 8)) {
        vg_set_error("vgpu_tune_set: %s is between 2 and 8", name);
        return -1;
    }
    if ((i == VG_TUNE_GB_GLB_LOAD_PCT && (value < 25 || value > 90)) || (i == VG_TUNE_GB_PART_FILL && value > 90) ||
        (i == VG_TUNE_JOIN_LOAD_PCT && (value < 10 || value > 95))) {
        vg_set_error("vgpu_tune_set: %s is out of range", name);
        return -1;
    }
    return 0;
}

extern "C" int vgpu_tune_set(const char *name, int64_t value) {
    int i = tune_index(name);
    if (tune_check(i, name, value)) return -1;
    g_tune[i] = value;
    return 0;
}

extern "C" int vgpu_tune_set_at(int64_t max_groups, int64_t max_width, int64_t max_rows, const char *name,
                                int64_t value) {
    int i = tune_index(name);
    if (tune_check(i, name, value)) return -1;
    if (max_groups <= 0 || max_width <= 0 || max_rows <= 0) {
        vg_set_error("vgpu_tune_set_at: bucket bounds must be positive");
        return -1;
    }
    std::lock_guard<std::mutex> lk(g_tune_mu);
    TuneBucket key;
    memset(&key, 0, sizeof key);
    key.max_groups = max_groups;
    key.max_width = max_width;
    key.max_rows = max_rows;
    auto it = g_buckets.begin();
    while (it != g_buckets.end() && bucket_before(*it, key)) ++it;
    if (it == g_buckets.end() || bucket_before(key, *it)) {
        try {
            it = g_buckets.insert(it, key);
        } catch (const std::bad_alloc &) {
            vg_set_error("vgpu_tune_set_at: host memory exhausted");
            return -1;
        }
    }
    it->val[i] = value;
    it->set |= 1u << i;
    return 0;
}

extern "C" int vgpu_tune_set_for(int64_t max_groups, const char *name, int64_t value) {
    return vgpu_tune_set_at(max_groups, TUNE_ANY, TUNE_ANY, name, value);
}

extern "C" int64_t vgpu_tune_get_at(int64_t groups, int64_t width, int64_t rows, const char *name) {
    int i = tune_index(name);
    if (i < 0) return -1;
    std::lock_guard<std::mutex> lk(g_tune_mu);
    return tune_at(i, groups, width, rows);
}

extern "C" int64_t vgpu_tune_get_for(int64_t groups, const char *name) {
    return vgpu_tune_get_at(groups, 0, 0, name);
}

extern "C" void vgpu_tune_reset(void) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_buckets.clear();
    const int64_t defaults[VG_TUNE_COUNT] = VG_TUNE_DEFAULTS;
    memcpy(g_tune, defaults, sizeof g_tune);
}

// "max_groups name value [w<=W] [rows<=R]" per line (max_groups 0 = the
// global value; a bucket without w<= / rows<= covers any width / rows);
// '#' starts a comment
extern "C" int vgpu_tune_load(const char *path) {
    FILE *f = fopen(path, "r");
    if (!f) {
        vg_set_error("vgpu_tune_load: cannot open %s", path);
        return -1;
    }
    char line[256];
    int n = 0, rc = 0;
    while (rc == 0 && fgets(line, sizeof line, f)) {
        char *h = strchr(line, '#');
        if (h) *h = 0;
        long long mg, v;
        char name[64];
        int used = 0;
        int k = sscanf(line, "%lld %63s %lld%n", &mg, name, &v, &used);
        if (k <= 0) continue;
        if (k != 3) {
            vg_set_error("vgpu_tune_load: %s: malformed line", path);
            rc = -1;
            break;
        }
        long long mw = TUNE_ANY, mr = TUNE_ANY;
        const char *rest = line + used;
        char tok[64];
        int adv = 0;
        while (rc == 0 && sscanf(rest, "%63s%n", tok, &adv) == 1) {
            rest += adv;
            if (sscanf(tok, "w<=%lld", &mw) == 1 || sscanf(tok, "rows<=%lld", &mr) == 1) continue;
            vg_set_error("vgpu_tune_load: %s: unknown bucket bound %s", path, tok);
            rc = -1;
        }
        if (rc) break;
        rc = mg > 0 ? vgpu_tune_set_at(mg, mw, mr, name, v) : vgpu_tune_set(name, v);
        n++;
    }
    fclose(f);
    return rc ? -1 : n;
}

extern "C" int64_t vgpu_tune_get(const char *name) {
    for (int i = 0; i < VG_TUNE_COUNT; i++)
        if (strcmp(name, g_tune_names[i]) == 0) return g_tune[i];
    return -1;
}

// per-launch kernel timing (off by default: it synchronizes each launch)
static int g_kt_on = 0;
static hipEvent_t g_k0 = nullptr, g_k1 = nullptr;
static double g_kt_ms = 0.0;
static int64_t g_kt_n = 0;

extern "C" int vgpu_kernel_timing(int enable) {
    if (vg_ensure_init()) return -1;
    if (!g_k0) VG_TRY(hipEventCreate(&g_k0));
    if (!g_k1) VG_TRY(hipEventCreate(&g_k1));
    g_kt_on = enable;
    return 0;
}

extern "C" double vgpu_kernel_time_ms(int64_t *launches, int reset) {
    double ms = g_kt_ms;
    if (launches) *launches = g_kt_n;
    if (reset) { g_kt_ms = 0.0; g_kt_n = 0; }
    return ms;
}

void vg_kt_begin() {
    if (g_kt_on) (void)hipEventRecord(g_k0, vg_stream());
}

void vg_kt_end() {
    if (!g_kt_on) return;
    if (hipEventRecord(g_k1, vg_stream()) != hipSuccess || hipEventSynchronize(g_k1) != hipSuccess) return;
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, g_k0, g_k1) == hipSuccess) {
        g_kt_ms += ms;
        g_kt_n++;
    }
}

// ---------------------------------------------------------------------------
// block-level scan helpers (256 threads = 4 waves)
// ---------------------------------------------------------------------------

#define SCAN_THREADS 256
#define SCAN_ITEMS 16
#define SCAN_TILE (SCAN_THREADS * SCAN_ITEMS)

// exclusive scan of one value per thread across the block; *total gets
// the block sum.  Uses 5 words of LDS.
__device__ __forceinline__ uint64_t block_exclusive_scan(uint64_t v, uint64_t *lds, uint64_t &total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint64_t x = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        uint64_t y = __shfl_up(x, d, 64);
        if (lane >= d) x += y;
    }
    if (lane == 63) lds[wave] = x;
    __syncthreads();
    uint64_t before = 0;
    for (int w = 0; w < wave; w++) before += lds[w];
    total = lds[0] + lds[1] + lds[2] + lds[3];
    __syncthreads();
    return before + x - v;
}

template <class TIn>
__global__ void __launch_bounds__(SCAN_THREADS) k_scan_reduce(const TIn *in, int64_t n, uint64_t *partials) {
    __shared__ uint64_t lds[4];
    int64_t base = (int64_t)blockIdx.x * SCAN_TILE;
    uint64_t s = 0;
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; i++) {
        int64_t idx = base + (int64_t)i * SCAN_THREADS + threadIdx.x;
        if (idx < n) s += (uint64_t)in[idx];
    }
    uint64_t total;
    block_exclusive_scan(s, lds, total);
    if (threadIdx.x == 0) partials[blockIdx.x] = total;
}

// each thread owns SCAN_ITEMS consecutive elements (blocked layout)
template <class TIn>
__global__ void __launch_bounds__(SCAN_THREADS) k_scan_apply(const TIn *in, uint64_t *out, int64_t n,
                                                             const uint64_t *block_off, uint64_t *total_out) {
    __shared__ uint64_t lds[4];
    int64_t base = (int64_t)blockIdx.x * SCAN_TILE + (int64_t)threadIdx.x * SCAN_ITEMS;
    uint64_t vals[SCAN_ITEMS];
    uint64_t s = 0;
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; i++) {
        int64_t idx = base + i;
        vals[i] = idx < n ? (uint64_t)in[idx] : 0;
        s += vals[i];
    }
    uint64_t total;
    uint64_t run = block_exclusive_scan(s, lds, total) + (block_off ? block_off[blockIdx.x] : 0);
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; i++) {
        int64_t idx = base + i;
        if (idx < n) out[idx] = run;
        run += vals[i];
    }
    if (total_out && blockIdx.x == gridDim.x - 1 && threadIdx.x == SCAN_THREADS - 1)
        *total_out = run;
}

// stream work of the scan and the sort is kernels only (no memset or copy
// nodes), so both can be captured into a hipGraph
__global__ void k_set_u64(uint64_t *p, uint64_t v) { *p = v; }

__global__ void k_copy_pairs(const uint64_t *ks, const uint32_t *vs, uint64_t *kd, uint32_t *vd, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        kd[i] = ks[i];
        vd[i] = vs[i];
    }
}

template <class TIn>
static int scan_impl(const TIn *in, uint64_t *out, int64_t n, uint64_t *total, DevBuf &tmp, size_t tmp_off) {
    hipStream_t s = vg_stream();
    if (n <= 0) {
        if (total) {
            k_set_u64<<<1, 1, 0, s>>>(total, 0);
            VG_LAUNCH_CHECK();
        }
        return 0;
    }
    int64_t nb = (n + SCAN_TILE - 1) / SCAN_TILE;
    if (nb == 1) {
        k_scan_apply<TIn><<<1, SCAN_THREADS, 0, s>>>(in, out, n, nullptr, total);
        VG_LAUNCH_CHECK();
        return 0;
    }
    // partials live in tmp at tmp_off; recursion uses the space after them
    size_t need = tmp_off + (size_t)nb * 8;
    if (tmp.cap < need) {
        vg_set_error("scan: scratch too small (%zu < %zu)", tmp.cap, need);
        return -1;
    }
    uint64_t *partials = reinterpret_cast<uint64_t *>(static_cast<char *>(tmp.ptr) + tmp_off);
    k_scan_reduce<TIn><<<(unsigned)nb, SCAN_THREADS, 0, s>>>(in, n, partials);
    VG_LAUNCH_CHECK();
    if (scan_impl<uint64_t>(partials, partials, nb, nullptr, tmp, need)) return -1;
    k_scan_apply<TIn><<<(unsigned)nb, SCAN_THREADS, 0, s>>>(in, out, n, partials, total);
    VG_LAUNCH_CHECK();
    return 0;
}

static size_t scan_tmp_bytes(int64_t n) {
    size_t b = 4096;
    while (n > SCAN_TILE) {
        n = (n + SCAN_TILE - 1) / SCAN_TILE;
        b += (size_t)n * 8 + 64;
    }
    return b * 2 + 4096;
}

size_t vg_scan_tmp_bytes(int64_t n) { return scan_tmp_bytes(n); }

int vg_scan_u32_to_u64_in(const uint32_t *in, uint64_t *out, int64_t n, uint64_t *total, void *tmp,
                          size_t tmp_bytes) {
    if (scan_tmp_bytes(n) > tmp_bytes) {
        vg_set_error("scan: scratch of %zu bytes too small for %lld values", tmp_bytes, (long long)n);
        return -1;
    }
    DevBuf view;   // borrowed: never reserved or released
    view.ptr = tmp;
    view.cap = tmp_bytes;
    return scan_impl<uint32_t>(in, out, n, total, view, 0);
}

int vg_scan_u32_to_u64(const uint32_t *in, uint64_t *out, int64_t n, uint64_t *total, DevBuf &tmp) {
    if (tmp.reserve(scan_tmp_bytes(n))) return -1;
    return scan_impl<uint32_t>(in, out, n, total, tmp, 0);
}

int vg_scan_u64(const uint64_t *in, uint64_t *out, int64_t n, uint64_t *total, DevBuf &tmp) {
    if (tmp.reserve(scan_tmp_bytes(n))) return -1;
    return scan_impl<uint64_t>(in, out, n, total, tmp, 0);
}

// ---------------------------------------------------------------------------
// stable LSD radix sort, 8-bit digits
// ---------------------------------------------------------------------------

#define RS_THREADS 256
#define RS_ROUNDS 16
#define RS_TILE (RS_THREADS * RS_ROUNDS)

__global__ void __launch_bounds__(RS_THREADS) k_rs_hist(const uint64_t *keys, int64_t n, int shift,
                                                        uint32_t *hist, int64_t n_tiles) {
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    int64_t base = (int64_t)blockIdx.x * RS_TILE;
    for (int r = 0; r < RS_ROUNDS; r++) {
        int64_t idx = base + (int64_t)r * RS_THREADS + threadIdx.x;
        if (idx < n) atomicAdd(&h[(keys[idx] >> shift) & 0xFF], 1u);
    }
    __syncthreads();
    hist[(int64_t)threadIdx.x * n_tiles + blockIdx.x] = h[threadIdx.x];
}

// One pass of the stable LSD sort: the block's tile of RS_TILE pairs is
// ranked by digit (rounds in index order, waves in order, lanes by ballot
// match), placed in LDS in digit order, and written out so that one
// digit's pairs of the tile land on consecutive global positions: runs of
// about RS_TILE / 256 pairs instead of one scattered 8-byte write per pair.
__global__ void __launch_bounds__(RS_THREADS) k_rs_scatter(const uint64_t *keys_in, const uint32_t *vals_in,
                                                           uint64_t *keys_out, uint32_t *vals_out, int64_t n,
                                                           int shift, const uint64_t *offsets, int64_t n_tiles) {
    __shared__ uint64_t sk[RS_TILE];
    __shared__ uint32_t sv[RS_TILE];
    __shared__ uint32_t start[256];   // the digit's first position in the tile
    __shared__ uint32_t fill[256];    // the digit's pairs placed so far
    __shared__ uint32_t wcnt[4][256];
    __shared__ uint64_t gpos[256];    // the digit's first global position of this tile
    const int t = threadIdx.x, wave = t >> 6;
    const int64_t base = (int64_t)blockIdx.x * RS_TILE;
    const int cnt = n - base < RS_TILE ? (int)(n - base) : RS_TILE;
    uint64_t kr[RS_ROUNDS];
    uint32_t vr[RS_ROUNDS];
    fill[t] = 0;
    for (int w = 0; w < 4; w++) wcnt[w][t] = 0;
    gpos[t] = offsets[(int64_t)t * n_tiles + blockIdx.x];
    __syncthreads();
#pragma unroll
    for (int r = 0; r < RS_ROUNDS; r++) {
        const int i = r * RS_THREADS + t;
        const bool in = i < cnt;
        kr[r] = in ? keys_in[base + i] : 0;
        vr[r] = in ? vals_in[base + i] : 0;
        if (in) atomicAdd(&fill[(uint32_t)(kr[r] >> shift) & 0xFF], 1u);
    }
    __syncthreads();
    // exclusive scan of the tile's digit counts
    const uint32_t c = fill[t];
    start[t] = c;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
        const uint32_t x = t >= off ? start[t - off] : 0;
        __syncthreads();
        start[t] += x;
        __syncthreads();
    }
    start[t] -= c;
    fill[t] = 0;
    __syncthreads();
#pragma unroll
    for (int r = 0; r < RS_ROUNDS; r++) {
        const int i = r * RS_THREADS + t;
        const bool in = i < cnt;
        const uint32_t d = (uint32_t)(kr[r] >> shift) & 0xFF;
        uint64_t m = vg_ballot(in);
#pragma unroll
        for (int b = 0; b < 8; b++) {
            const uint64_t bb = vg_ballot((d >> b) & 1);
            m &= ((d >> b) & 1) ? bb : ~bb;
        }
        const int rank = vg_lane_prefix(m);
        if (in && rank == 0) wcnt[wave][d] = (uint32_t)__popcll(m);
        __syncthreads();
        if (in) {
            uint32_t pos = start[d] + fill[d] + (uint32_t)rank;
            for (int w = 0; w < wave; w++) pos += wcnt[w][d];
            sk[pos] = kr[r];
            sv[pos] = vr[r];
        }
        __syncthreads();
        fill[t] += wcnt[0][t] + wcnt[1][t] + wcnt[2][t] + wcnt[3][t];
        for (int w = 0; w < 4; w++) wcnt[w][t] = 0;
        __syncthreads();
    }
    for (int i = t; i < cnt; i += RS_THREADS) {
        const uint64_t k = sk[i];
        const uint32_t d = (uint32_t)(k >> shift) & 0xFF;
        const uint64_t p = gpos[d] + (uint64_t)(i - (int)start[d]);
        keys_out[p] = k;
        vals_out[p] = sv[i];
    }
}

int vg_sort_pairs(uint64_t *keys, uint32_t *vals, int64_t n, int begin_bit, int end_bit, DevBuf &tmp) {
    if (n <= 1 || begin_bit >= end_bit) return 0;
    hipStream_t s = vg_stream();
    int64_t n_tiles = (n + RS_TILE - 1) / RS_TILE;
    int64_t n_hist = n_tiles * 256;
    size_t off_keys = 0;
    size_t off_vals = off_keys + (size_t)n * 8;
    size_t off_hist = (off_vals + (size_t)n * 4 + 255) & ~(size_t)255;
    size_t off_offs = (off_hist + (size_t)n_hist * 4 + 255) & ~(size_t)255;
    size_t off_scan = (off_offs + (size_t)n_hist * 8 + 255) & ~(size_t)255;
    size_t scan_bytes = scan_tmp_bytes(n_hist);
    if (tmp.reserve(off_scan + scan_bytes)) return -1;
    char *base = static_cast<char *>(tmp.ptr);
    uint64_t *k_alt = reinterpret_cast<uint64_t *>(base + off_keys);
    uint32_t *v_alt = reinterpret_cast<uint32_t *>(base + off_vals);
    uint32_t *hist = reinterpret_cast<uint32_t *>(base + off_hist);
    uint64_t *offs = reinterpret_cast<uint64_t *>(base + off_offs);
    DevBuf scan_view;
    scan_view.ptr = base + off_scan;
    scan_view.cap = scan_bytes;
    uint64_t *kin = keys, *kout = k_alt;
    uint32_t *vin = vals, *vout = v_alt;
    int passes = 0;
    for (int shift = begin_bit; shift < end_bit; shift += 8) {
        k_rs_hist<<<(unsigned)n_tiles, RS_THREADS, 0, s>>>(kin, n, shift, hist, n_tiles);
        VG_LAUNCH_CHECK();
        if (scan_impl<uint32_t>(hist, offs, n_hist, nullptr, scan_view, 0)) return -1;
        k_rs_scatter<<<(unsigned)n_tiles, RS_THREADS, 0, s>>>(kin, vin, kout, vout, n, shift, offs, n_tiles);
        VG_LAUNCH_CHECK();
        uint64_t *tk = kin; kin = kout; kout = tk;
        uint32_t *tv = vin; vin = vout; vout = tv;
        passes++;
    }
    if (passes & 1) {
        k_copy_pairs<<<vg_blocks(n, 256), 256, 0, s>>>(kin, vin, keys, vals, n);
        VG_LAUNCH_CHECK();
    }
    scan_view.ptr = nullptr;   // borrowed view: nothing to free
    scan_view.cap = 0;
    return 0;
}

extern "C" int vgpu_sort_pairs_u64(uint64_t *keys, uint32_t *values, int64_t n, int begin_bit, int end_bit) {
    if (vg_ensure_init()) return -1;
    DevBuf tmp;
    int rc = vg_sort_pairs(keys, values, n, begin_bit, end_bit, tmp);
    if (!rc && hipStreamSynchronize(vg_stream()) != hipSuccess) {
        vg_set_error("sort: stream sync failed");
        rc = -1;
    }
    tmp.release();
    return rc;
}
