// vgpu_strdict.hip -- an exact device dictionary of strings.
//
// The reference compares string keys byte for byte wherever it hashes them
// (src/hash.c:58-66 for group-by keys, src/join.c:96-108 for join keys).
// Here every distinct string gets one int64 code: the arena row index of a
// row that holds it.  Two rows get the same code iff their bytes are equal,
// because the table compares bytes; the hash only picks the first slot and
// a 24-bit tag that skips most comparisons.  Grouping or joining on codes is
// therefore exact, whatever the strings' hashes do.
//
// Layout in HBM:
//   arena    every appended row's bytes (achr) and offsets (aoff, rows + 1),
//            in append order: a later batch compares against earlier rows'
//            bytes in place;
//   slots    open addressing, linear probing, one 64-bit word per slot:
//            (arena row << 24) | (hash >> 40), all ones when empty.  A slot
//            is claimed by one compare-and-swap from empty and never changes
//            after that, so a thread that loses the claim reads the winner's
//            word from the swap and compares bytes at once: no thread waits
//            for another (no spin inside a wave).
// The table keeps at most half its slots full (it doubles, re-inserting the
// claimed words, before an append could pass that), so every probe sequence
// meets an empty slot or the string.
#include "vgpu_internal.h"

#include <algorithm>
#include <cstring>

#define SD_EMPTY 0xFFFFFFFFFFFFFFFFull
#define SD_TAG_MASK 0xFFFFFFull
#define SD_MAX_ROWS (int64_t(1) << 40)

struct SdView {
    uint64_t *slots;
    uint64_t mask;                  // slots - 1
    const int64_t *aoff;            // arena offsets
    const uint8_t *achr;            // arena bytes
    unsigned long long *n_new;      // strings claimed by this launch
    int32_t hash_bits;
};

__device__ __forceinline__ uint64_t sd_hash(const uint8_t *p, int64_t len, int bits) {
    uint64_t a = 0xcbf29ce484222325ull;   // FNV-1a
    for (int64_t k = 0; k < len; k++) a = (a ^ p[k]) * 0x100000001b3ull;
    uint64_t h = vg_mix64(a ^ (uint64_t)len);
    if (bits < 64) h &= (uint64_t(1) << bits) - 1;
    return h;
}

__device__ __forceinline__ bool sd_equal(const uint8_t *a, const uint8_t *b, int64_t len) {
    for (int64_t k = 0; k < len; k++)
        if (a[k] != b[k]) return false;
    return true;
}

// the code of the string p[0..len) with hash h: an arena row holding the
// same bytes, found or (claim) inserted as row `mine`; -1 when absent and
// not claimed
template <bool CLAIM>
__device__ __forceinline__ int64_t sd_find(const SdView &V, const uint8_t *p, int64_t len, uint64_t h, int64_t mine) {
    const uint64_t tag = h >> 40;
    uint64_t s = h & V.mask;
    for (uint64_t probe = 0; probe <= V.mask; probe++) {
        uint64_t w = __atomic_load_n(&V.slots[s], __ATOMIC_RELAXED);
        if (w == SD_EMPTY) {
            if (!CLAIM) return -1;
            const uint64_t word = ((uint64_t)mine << 24) | tag;
            const uint64_t old = atomicCAS((unsigned long long *)&V.slots[s], SD_EMPTY, word);
            if (old == SD_EMPTY) {
                atomicAdd(V.n_new, 1ull);
                return mine;
            }
            w = old;
        }
        if ((w & SD_TAG_MASK) == tag) {
            const int64_t j = (int64_t)(w >> 24);
            const int64_t sj = V.aoff[j];
            if (V.aoff[j + 1] - sj == len && sd_equal(p, V.achr + sj, len)) return j;
        }
        s = (s + 1) & V.mask;
    }
    return -1;   // not reached: the table always has an empty slot
}

// arena rows [row0, row0 + n): offsets shifted by delta (the batch's bytes
// now start at the arena's end)
__global__ void k_sd_rebase(int64_t *aoff, int64_t n1, int64_t delta) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n1) aoff[i] += delta;
}

__global__ void __launch_bounds__(256) k_sd_insert(SdView V, const uint8_t *valid, int64_t row0, int64_t n,
                                                   int64_t *codes, int64_t *pos, int64_t pos_base) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += stride) {
        int64_t code = 0;
        if (!valid || ((valid[r >> 3] >> (r & 7)) & 1)) {
            const int64_t i = row0 + r, s = V.aoff[i], len = V.aoff[i + 1] - s;
            const uint8_t *p = V.achr + s;
            code = sd_find<true>(V, p, len, sd_hash(p, len, V.hash_bits), i);
        }
        codes[r] = code;
        if (pos) pos[r] = pos_base + r;
    }
}

__global__ void __launch_bounds__(256) k_sd_lookup(SdView V, const int64_t *off, const uint8_t *chr,
                                                   const uint8_t *valid, int64_t n, int64_t *codes) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += stride) {
        int64_t code = 0;
        if (!valid || ((valid[r >> 3] >> (r & 7)) & 1)) {
            const int64_t s = off[r], len = off[r + 1] - s;
            const uint8_t *p = chr + (s - off[0]);
            code = sd_find<false>(V, p, len, sd_hash(p, len, V.hash_bits), 0);
        }
        codes[r] = code;
    }
}

// the claimed words of an old table into the new one (distinct strings: no
// byte comparisons, the words keep their rows and tags)
__global__ void __launch_bounds__(256) k_sd_rehash(const uint64_t *old, int64_t old_cap, SdView V) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < old_cap; i += stride) {
        const uint64_t w = old[i];
        if (w == SD_EMPTY) continue;
        const int64_t j = (int64_t)(w >> 24), sj = V.aoff[j], len = V.aoff[j + 1] - sj;
        uint64_t s = sd_hash(V.achr + sj, len, V.hash_bits) & V.mask;
        for (uint64_t probe = 0; probe <= V.mask; probe++) {
            if (atomicCAS((unsigned long long *)&V.slots[s], SD_EMPTY, w) == SD_EMPTY) break;
            s = (s + 1) & V.mask;
        }
    }
}

struct vgpu_strdict {
    DevBuf slots, aoff, achr, ctl, look_off, look_chr;
    int64_t cap = 0;        // slots
    int64_t rows = 0;       // arena rows
    int64_t chars = 0;      // arena bytes
    int64_t distinct = 0;
    int32_t hash_bits = 64;
};

// grows b to hold `bytes`, keeping its first `keep` bytes
static int sd_grow(DevBuf &b, size_t bytes, size_t keep) {
    if (bytes <= b.cap) return 0;
    size_t want = std::max<size_t>(bytes, b.cap * 2);
    DevBuf nb;
    if (nb.reserve(want)) return -1;
    if (keep) VG_TRY(hipMemcpyAsync(nb.ptr, b.ptr, keep, hipMemcpyDeviceToDevice, vg_stream()));
    if (b.ptr) {
        // the pool hands the old block out again only to later work on
        // this stream, which runs after the copy
        b.release();
    }
    b = nb;
    return 0;
}

static SdView sd_view(vgpu_strdict *d) {
    SdView V;
    V.slots = d->slots.as<uint64_t>();
    V.mask = (uint64_t)d->cap - 1;
    V.aoff = d->aoff.as<int64_t>();
    V.achr = d->achr.as<uint8_t>();
    V.n_new = d->ctl.as<unsigned long long>();
    V.hash_bits = d->hash_bits;
    return V;
}

// room for `more` new strings at half load
static int sd_reserve_slots(vgpu_strdict *d, int64_t more) {
    const int64_t need = 2 * (d->distinct + more) + 2;
    if (d->cap >= need) return 0;
    int64_t cap = 1024;
    while (cap < 2 * need) cap *= 2;
    hipStream_t st = vg_stream();
    DevBuf ns;
    if (ns.reserve((size_t)cap * 8)) return -1;
    VG_TRY(hipMemsetAsync(ns.ptr, 0xFF, (size_t)cap * 8, st));
    DevBuf old = d->slots;
    const int64_t old_cap = d->cap;
    d->slots = ns;
    d->cap = cap;
    if (old_cap > 0) {
        SdView V = sd_view(d);
        k_sd_rehash<<<vg_blocks(old_cap, 256, 8192), 256, 0, st>>>(old.as<uint64_t>(), old_cap, V);
        VG_LAUNCH_CHECK();
    }
    old.release();
    return 0;
}

static vgpu_strdict *strdict_create_body() {
    if (vg_ensure_init()) return nullptr;
    vgpu_strdict *d = new vgpu_strdict();
    d->hash_bits = (int32_t)std::min<int64_t>(64, std::max<int64_t>(1, vg_tune(VG_TUNE_SD_HASH_BITS)));
    if (d->ctl.reserve(8) || d->aoff.reserve(8)) {
        delete d;
        return nullptr;
    }
    if (hipMemsetAsync(d->aoff.ptr, 0, 8, vg_stream()) != hipSuccess) {
        vg_set_error("strdict: %s", hipGetErrorString(hipGetLastError()));
        delete d;
        return nullptr;
    }
    return d;
}

extern "C" vgpu_strdict *vgpu_strdict_create(void) {
    return vg_guard<vgpu_strdict *>(nullptr, "vgpu_strdict_create", [&] { return strdict_create_body(); });
}

extern "C" void vgpu_strdict_free(vgpu_strdict *d) {
    if (!d) return;
    (void)hipStreamSynchronize(vg_stream());
    d->slots.release();
    d->aoff.release();
    d->achr.release();
    d->ctl.release();
    d->look_off.release();
    d->look_chr.release();
    delete d;
}

static int strdict_check(const vgpu_strdict *d, const int64_t *offsets, const uint8_t *chars, int64_t n,
                         const int64_t *codes, const char *what) {
    if (!d || n < 0 || (n > 0 && (!offsets || !codes))) {
        vg_set_error("%s: invalid arguments", what);
        return -1;
    }
    if (n > 0 && offsets[n] < offsets[0]) {
        vg_set_error("%s: offsets decrease", what);
        return -1;
    }
    if (n > 0 && offsets[n] > offsets[0] && !chars) {
        vg_set_error("%s: no bytes", what);
        return -1;
    }
    return 0;
}

static int strdict_append_body(vgpu_strdict *d, const int64_t *offsets, const uint8_t *chars,
                               const uint8_t *validity, int64_t n, int64_t pos_base, int64_t *codes, int64_t *pos) {
    if (strdict_check(d, offsets, chars, n, codes, "strdict append")) return -1;
    if (n == 0) return 0;
    if (d->rows + n >= SD_MAX_ROWS) {
        vg_set_error("strdict append: more than 2^40 rows");
        return -1;
    }
    hipStream_t st = vg_stream();
    const int64_t bytes = offsets[n] - offsets[0];
    // the batch's offsets and bytes at the arena's end
    if (sd_grow(d->aoff, (size_t)(d->rows + n + 1) * 8, (size_t)(d->rows + 1) * 8) ||
        sd_grow(d->achr, (size_t)std::max<int64_t>(d->chars + bytes, 1), (size_t)d->chars))
        return -1;
    int64_t *aoff = d->aoff.as<int64_t>() + d->rows;
    VG_TRY(hipMemcpyAsync(aoff, offsets, (size_t)(n + 1) * 8, hipMemcpyHostToDevice, st));
    if (bytes > 0)
        VG_TRY(hipMemcpyAsync(d->achr.as<uint8_t>() + d->chars, chars + offsets[0], (size_t)bytes,
                              hipMemcpyHostToDevice, st));
    k_sd_rebase<<<vg_blocks(n + 1, 256), 256, 0, st>>>(aoff, n + 1, d->chars - offsets[0]);
    VG_LAUNCH_CHECK();
    if (sd_reserve_slots(d, n)) return -1;
    VG_TRY(hipMemsetAsync(d->ctl.ptr, 0, 8, st));
    SdView V = sd_view(d);
    k_sd_insert<<<vg_blocks(n, 256, 8192), 256, 0, st>>>(V, validity, d->rows, n, codes, pos, pos_base);
    VG_LAUNCH_CHECK();
    uint64_t claimed = 0;
    VG_TRY(hipMemcpyAsync(&claimed, d->ctl.ptr, 8, hipMemcpyDeviceToHost, st));
    VG_TRY(hipStreamSynchronize(st));
    d->distinct += (int64_t)claimed;
    d->rows += n;
    d->chars += bytes;
    return 0;
}

extern "C" int vgpu_strdict_append(vgpu_strdict *d, const int64_t *offsets, const uint8_t *chars,
                                   const uint8_t *validity, int64_t n_rows, int64_t pos_base, int64_t *codes,
                                   int64_t *pos) {
    return vg_guard<int>(-1, "vgpu_strdict_append",
                         [&] { return strdict_append_body(d, offsets, chars, validity, n_rows, pos_base, codes, pos); });
}

static int strdict_lookup_body(vgpu_strdict *d, const int64_t *offsets, const uint8_t *chars,
                               const uint8_t *validity, int64_t n, int64_t *codes) {
    if (strdict_check(d, offsets, chars, n, codes, "strdict lookup")) return -1;
    if (n == 0) return 0;
    hipStream_t st = vg_stream();
    if (d->cap == 0 && sd_reserve_slots(d, 0)) return -1;   // nothing appended: an empty table
    const int64_t bytes = offsets[n] - offsets[0];
    if (d->look_off.reserve((size_t)(n + 1) * 8) || d->look_chr.reserve((size_t)std::max<int64_t>(bytes, 1)))
        return -1;
    VG_TRY(hipMemcpyAsync(d->look_off.ptr, offsets, (size_t)(n + 1) * 8, hipMemcpyHostToDevice, st));
    if (bytes > 0)
        VG_TRY(hipMemcpyAsync(d->look_chr.ptr, chars + offsets[0], (size_t)bytes, hipMemcpyHostToDevice, st));
    SdView V = sd_view(d);
    k_sd_lookup<<<vg_blocks(n, 256, 8192), 256, 0, st>>>(V, d->look_off.as<int64_t>(), d->look_chr.as<uint8_t>(),
                                                         validity, n, codes);
    VG_LAUNCH_CHECK();
    // the lookup buffers are reused by the next call
    VG_TRY(hipStreamSynchronize(st));
    return 0;
}

extern "C" int vgpu_strdict_lookup(vgpu_strdict *d, const int64_t *offsets, const uint8_t *chars,
                                   const uint8_t *validity, int64_t n_rows, int64_t *codes) {
    return vg_guard<int>(-1, "vgpu_strdict_lookup",
                         [&] { return strdict_lookup_body(d, offsets, chars, validity, n_rows, codes); });
}

extern "C" int64_t vgpu_strdict_size(const vgpu_strdict *d, int64_t *rows) {
    if (!d) return -1;
    if (rows) *rows = d->rows;
    return d->distinct;
}

// ---------------------------------------------------------------------------
// string comparison (vec_cmp on strings, src/scalar_ops.c:77-106)
// ---------------------------------------------------------------------------

struct ScSide {
    const int64_t *off;
    const uint8_t *chr;
    const uint8_t *valid;
    int32_t constant, null_const;
};

__device__ __forceinline__ bool sc_get(const ScSide &S, int64_t r, const uint8_t *&p, int64_t &len) {
    if (S.null_const) return false;
    const int64_t i = S.constant ? 0 : r;
    if (!S.constant && S.valid && !((S.valid[r >> 3] >> (r & 7)) & 1)) return false;
    p = S.chr + S.off[i];
    len = S.off[i + 1] - S.off[i];
    return true;
}

// rows are taken 64 to a wave in order (block size and grid stride are
// multiples of 64), so a wave's ballot is one whole word of the bitmap
__global__ void __launch_bounds__(256) k_string_compare(ScSide A, ScSide B, int64_t n, int64_t *out,
                                                        uint64_t *out_valid) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += stride) {
        const uint8_t *pa = nullptr, *pb = nullptr;
        int64_t la = 0, lb = 0;
        const bool ok = sc_get(A, r, pa, la) & sc_get(B, r, pb, lb);
        int64_t c = 0;
        if (ok) {
            const int64_t m = la < lb ? la : lb;
            for (int64_t k = 0; k < m && c == 0; k++) c = (int64_t)(pa[k] > pb[k]) - (int64_t)(pa[k] < pb[k]);
            if (c == 0) c = (int64_t)(la > lb) - (int64_t)(la < lb);
        }
        out[r] = c;
        const uint64_t w = vg_ballot(ok);
        if (vg_lane() == 0) out_valid[r >> 6] = w;
    }
}

static int sc_side(const vgpu_strcol *s, ScSide &S, const char *name) {
    if (!s) {
        vg_set_error("string compare: %s side missing", name);
        return -1;
    }
    S.off = s->offsets;
    S.chr = s->chars;
    S.valid = s->validity;
    S.constant = s->constant;
    S.null_const = s->null_const;
    if (!S.null_const && !S.off) {
        vg_set_error("string compare: %s side has no offsets", name);
        return -1;
    }
    return 0;
}

extern "C" int vgpu_string_compare(const vgpu_strcol *a, const vgpu_strcol *b, int64_t n_rows, int64_t *out,
                                   uint8_t *out_valid) {
    if (vg_ensure_init()) return -1;
    if (n_rows < 0 || (n_rows > 0 && (!out || !out_valid))) {
        vg_set_error("string compare: invalid arguments");
        return -1;
    }
    if (reinterpret_cast<uintptr_t>(out_valid) & 7) {
        vg_set_error("string compare: the validity bitmap must be 8-byte aligned");
        return -1;
    }
    ScSide A, B;
    if (sc_side(a, A, "left") || sc_side(b, B, "right")) return -1;
    if (n_rows == 0) return 0;
    // one wave's rows share a bitmap word: the grid covers whole words
    k_string_compare<<<vg_blocks(n_rows, 256, 8192), 256, 0, vg_stream()>>>(A, B, n_rows, out,
                                                                            reinterpret_cast<uint64_t *>(out_valid));
    VG_LAUNCH_CHECK();
    return 0;
}
