// vgpu_filter.hip -- predicate -> selection vector (src/filter.c:13-39).
//
// Three launches per batch, no inter-block waiting:
//   1. k_filter_mask : evaluate the predicate over 2048-row tiles, one
//      64-bit ballot word per wave-row-group, and a per-tile count;
//   2. device scan of the tile counts;
//   3. k_filter_write: re-read the (n/8-byte) mask and write ascending
//      physical indices with mbcnt ranks.
// The column is read once; the mask round trip adds n/4 bytes.
#include "vgpu_expr.h"

#include <mutex>

#define FT_THREADS 256
#define FT_ITEMS 8
#define FT_TILE (FT_THREADS * FT_ITEMS)
#define FT_WORDS (FT_TILE / 64)

PredInfo vg_pred_info(const vgpu_prog *p) {
    PredInfo pi;
    memset(&pi, 0, sizeof pi);
    if (!p || p->n_insn <= 0) return pi;
    pi.present = 1;
    const vgpu_insn *in = p->insn;
    int n = p->n_insn;
    const vgpu_insn &c = in[n - 1];
    if (c.op < VGPU_OP_EQ || c.op > VGPU_OP_GE) return pi;
    if (c.type != VGPU_INT64 && c.type != VGPU_DOUBLE) return pi;
    // locate the column side (optionally widened by one CAST) and the const
    auto col_of = [&](int r, int &col, int &ctype) -> bool {
        if (in[r].op == VGPU_OP_COL && in[r].type == c.type) {
            col = in[r].col; ctype = in[r].type; return true;
        }
        if (in[r].op == VGPU_OP_CAST && in[r].type == VGPU_DOUBLE && c.type == VGPU_DOUBLE &&
            in[in[r].a].op == VGPU_OP_COL && in[in[r].a].type == VGPU_INT64) {
            col = in[in[r].a].col; ctype = VGPU_INT64; return true;
        }
        return false;
    };
    int col = -1, ctype = 0;
    int op = c.op;
    uint64_t k;
    if (in[c.b].op == VGPU_OP_CONST && in[c.b].type == c.type && col_of(c.a, col, ctype)) {
        k = (uint64_t)in[c.b].imm.i;
    } else if (in[c.a].op == VGPU_OP_CONST && in[c.a].type == c.type && col_of(c.b, col, ctype)) {
        k = (uint64_t)in[c.a].imm.i;
        switch (op) {   // k op col  ==  col op' k
        case VGPU_OP_LT: op = VGPU_OP_GT; break;
        case VGPU_OP_LE: op = VGPU_OP_GE; break;
        case VGPU_OP_GT: op = VGPU_OP_LT; break;
        case VGPU_OP_GE: op = VGPU_OP_LE; break;
        default: break;
        }
    } else {
        return pi;
    }
    // the fast path must read nothing the program does not: every other
    // instruction has to be the ones consumed above
    int used = 2 + (ctype != c.type ? 1 : 0) + 1;
    if (used != n) return pi;
    pi.fast = 1;
    pi.col = col;
    pi.cmp_op = op;
    pi.cmp_type = c.type;
    pi.col_type = ctype;
    pi.k = k;
    return pi;
}

__global__ void __launch_bounds__(FT_THREADS)
k_filter_mask(PredInfo pi, vgpu_prog prog, ColSet cs, int64_t n_logical, const int32_t *sel_in,
              uint64_t *mask, uint32_t *tile_counts) {
    extern __shared__ uint8_t dyn[];
    __shared__ uint32_t wsum[4];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    uint64_t *regs = reinterpret_cast<uint64_t *>(dyn);
    uint8_t *regv = dyn + (size_t)prog.n_insn * FT_THREADS * 8;
    int64_t base = (int64_t)blockIdx.x * FT_TILE;
    uint32_t cnt = 0;
#pragma unroll
    for (int i = 0; i < FT_ITEMS; i++) {
        int64_t li = base + (int64_t)i * FT_THREADS + tid;
        bool pass = false;
        if (li < n_logical) {
            int64_t row = sel_in ? (int64_t)sel_in[li] : li;
            pass = vg_pred_row(pi, prog, cs, row, regs, regv, FT_THREADS, tid);
        }
        uint64_t w = vg_ballot(pass);
        if (lane == 0) {
            int64_t wi = (base + (int64_t)i * FT_THREADS + wave * 64) >> 6;
            if (wi * 64 < n_logical) mask[wi] = w;
            cnt += (uint32_t)__popcll(w);
        }
    }
    if (lane == 0) wsum[wave] = cnt;
    __syncthreads();
    if (tid == 0) tile_counts[blockIdx.x] = wsum[0] + wsum[1] + wsum[2] + wsum[3];
}

__global__ void __launch_bounds__(FT_THREADS)
k_filter_write(int64_t n_logical, const int32_t *sel_in, const uint64_t *mask,
               const uint64_t *tile_off, int32_t *sel_out) {
    __shared__ uint32_t wpre[FT_WORDS];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    int64_t base = (int64_t)blockIdx.x * FT_TILE;
    int64_t wbase = base >> 6;
    int64_t n_words = (n_logical + 63) >> 6;
    if (tid < 64) {
        // word tid of this tile, popcount, exclusive scan across 32 words
        uint32_t c = 0;
        if (tid < FT_WORDS && wbase + tid < n_words) c = (uint32_t)__popcll(mask[wbase + tid]);
        uint32_t x = c;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            uint32_t y = __shfl_up(x, d, 64);
            if (lane >= d) x += y;
        }
        if (tid < FT_WORDS) wpre[tid] = x - c;
    }
    __syncthreads();
    uint64_t off0 = tile_off[blockIdx.x];
#pragma unroll
    for (int i = 0; i < FT_ITEMS; i++) {
        int64_t li = base + (int64_t)i * FT_THREADS + tid;
        int wi = i * 4 + wave;
        if (wbase + wi >= n_words) continue;
        uint64_t w = mask[wbase + wi];
        if ((w >> lane) & 1) {
            uint64_t pos = off0 + wpre[wi] + (uint64_t)vg_lane_prefix(w);
            sel_out[pos] = sel_in ? sel_in[li] : (int32_t)li;
        }
    }
}

extern "C" int vgpu_eval_host(const vgpu_prog *prog, const vgpu_col *cols, int n_cols, int64_t n_rows,
                              const int32_t *sel_in, int64_t n_sel_in, int32_t *sel_out, int64_t *n_out) {
    if (!prog || prog->n_insn <= 0 || prog->n_insn > VGPU_MAX_INSN || n_cols < 0 || n_cols > VGPU_MAX_COLS) {
        vg_set_error("vgpu_eval_host: invalid program");
        return -1;
    }
    ColSet cs;
    memset(&cs, 0, sizeof cs);
    for (int i = 0; i < n_cols; i++) cs.c[i] = cols[i];
    PredInfo pi = vg_pred_info(prog);
    uint64_t regs[VGPU_MAX_INSN];
    uint8_t regv[VGPU_MAX_INSN];
    int64_t n = sel_in ? n_sel_in : n_rows, k = 0;
    for (int64_t li = 0; li < n; li++) {
        int64_t row = sel_in ? sel_in[li] : li;
        if (vg_pred_row(pi, *prog, cs, row, regs, regv, 1, 0)) sel_out[k++] = (int32_t)row;
    }
    *n_out = k;
    return 0;
}

static std::mutex g_filter_mu;
static DevBuf g_mask, g_counts, g_offs, g_scan, g_total;

// argument checks shared by both filter entry points
static int filter_check(const char *who, const vgpu_prog *prog, int n_cols) {
    if (!prog || prog->n_insn <= 0 || prog->n_insn > VGPU_MAX_INSN) {
        vg_set_error("%s: invalid program", who);
        return -1;
    }
    if (n_cols < 0 || n_cols > VGPU_MAX_COLS) {
        vg_set_error("%s: %d columns exceeds %d", who, n_cols, VGPU_MAX_COLS);
        return -1;
    }
    for (int i = 0; i < prog->n_insn; i++) {
        const vgpu_insn &in = prog->insn[i];
        if (in.op == VGPU_OP_COL && (in.col < 0 || in.col >= n_cols)) {
            vg_set_error("%s: instruction %d reads column slot %d of %d", who, i, in.col, n_cols);
            return -1;
        }
        if (in.op >= VGPU_OP_CAST && (in.a >= i || (in.op >= VGPU_OP_ADD && in.op <= VGPU_OP_OR && in.b >= i))) {
            vg_set_error("%s: instruction %d reads a later register", who, i);
            return -1;
        }
    }
    return 0;
}

// mask -> tile counts -> scan -> ordered write, all into the given buffers
// (mask: n_words words, counts: n_tiles, offs: n_tiles, total: 1 word);
// scan_tmp / scan_buf is the scan's scratch (one of the two)
static int filter_enqueue(const vgpu_prog *prog, const vgpu_col *cols, int n_cols, int64_t n_logical,
                          const int32_t *sel_in, int32_t *sel_out, uint64_t *mask, uint32_t *counts,
                          uint64_t *offs, uint64_t *total, void *scan_tmp, size_t scan_cap, DevBuf *scan_buf,
                          bool timed) {
    ColSet cs;
    memset(&cs, 0, sizeof cs);
    for (int i = 0; i < n_cols; i++) cs.c[i] = cols[i];
    PredInfo pi = vg_pred_info(prog);
    int64_t n_tiles = (n_logical + FT_TILE - 1) / FT_TILE;
    hipStream_t s = vg_stream();
    size_t lds = pi.fast ? 0 : vg_vm_lds_bytes(prog, FT_THREADS);
    if (timed) vg_kt_begin();
    k_filter_mask<<<(unsigned)n_tiles, FT_THREADS, lds, s>>>(pi, *prog, cs, n_logical, sel_in, mask, counts);
    VG_LAUNCH_CHECK();
    if (timed) vg_kt_end();
    int rc = scan_buf ? vg_scan_u32_to_u64(counts, offs, n_tiles, total, *scan_buf)
                      : vg_scan_u32_to_u64_in(counts, offs, n_tiles, total, scan_tmp, scan_cap);
    if (rc) return -1;
    k_filter_write<<<(unsigned)n_tiles, FT_THREADS, 0, s>>>(n_logical, sel_in, mask, offs, sel_out);
    VG_LAUNCH_CHECK();
    return 0;
}

extern "C" int vgpu_filter(const vgpu_prog *prog, const vgpu_col *cols, int n_cols, int64_t n_rows,
                           const int32_t *sel_in, int64_t n_sel_in, int32_t *sel_out, int64_t *n_out) {
    if (vg_ensure_init()) return -1;
    if (filter_check("vgpu_filter", prog, n_cols)) return -1;
    std::lock_guard<std::mutex> lk(g_filter_mu);
    int64_t n_logical = sel_in ? n_sel_in : n_rows;
    *n_out = 0;
    if (n_logical <= 0) return 0;
    int64_t n_tiles = (n_logical + FT_TILE - 1) / FT_TILE;
    int64_t n_words = (n_logical + 63) / 64;
    if (g_mask.reserve((size_t)n_words * 8) || g_counts.reserve((size_t)n_tiles * 4) ||
        g_offs.reserve((size_t)n_tiles * 8) || g_total.reserve(8))
        return -1;
    if (filter_enqueue(prog, cols, n_cols, n_logical, sel_in, sel_out, g_mask.as<uint64_t>(),
                       g_counts.as<uint32_t>(), g_offs.as<uint64_t>(), g_total.as<uint64_t>(), nullptr, 0,
                       &g_scan, true))
        return -1;
    hipStream_t s = vg_stream();
    uint64_t total = 0;
    VG_TRY(hipMemcpyAsync(&total, g_total.ptr, 8, hipMemcpyDeviceToHost, s));
    VG_TRY(hipStreamSynchronize(s));
    *n_out = (int64_t)total;
    return 0;
}

// ---------------------------------------------------------------------------
// The filter without allocation.  Its scratch -- the ballot mask, the tile
// counts, their offsets, the total and the scan's partials -- lives in one
// file-scope device array, `scratch`, sized for VGPU_FILTER_NOALLOC_MAX_ROWS
// logical rows.  Each device holds its own copy of the array; its address is
// looked up once per device by vgpu_init (hipGetSymbolAddress) and only that
// address is handed to kernels.  Every scratch byte a call reads is written
// by the same call first (mask words of the logical rows, one count and one
// offset per tile, the total), so nothing has to start at zero.
// ---------------------------------------------------------------------------

#define NA_MAX_ROWS ((int64_t)VGPU_FILTER_NOALLOC_MAX_ROWS)
#define NA_MAX_TILES ((NA_MAX_ROWS + FT_TILE - 1) / FT_TILE)
#define NA_MASK_OFF ((size_t)0)
#define NA_COUNTS_OFF (NA_MASK_OFF + (size_t)(NA_MAX_ROWS / 64) * 8)
#define NA_OFFS_OFF (NA_COUNTS_OFF + (((size_t)NA_MAX_TILES * 4 + 255) & ~(size_t)255))
#define NA_TOTAL_OFF (NA_OFFS_OFF + (size_t)NA_MAX_TILES * 8)
#define NA_SCAN_OFF (NA_TOTAL_OFF + 256)
#define NA_SCAN_BYTES ((size_t)65536)
#define NA_SCRATCH_BYTES (NA_SCAN_OFF + NA_SCAN_BYTES)

static __device__ __align__(256) uint8_t scratch[NA_SCRATCH_BYTES];

#define NA_MAX_DEVICES 64
static void *g_na_scratch[NA_MAX_DEVICES];   // per device: the address of `scratch`
static uint64_t *g_na_total[NA_MAX_DEVICES];  // per device: pinned host word for the count
static int g_na_device = -1;                  // device of the last set-up (the library's device)
static std::mutex g_na_mu;

int vg_filter_noalloc_setup(int device) {
    if (device < 0 || device >= NA_MAX_DEVICES) {
        vg_set_error("vgpu_init: device %d beyond the %d devices the no-allocation filter supports", device,
                     NA_MAX_DEVICES);
        return -1;
    }
    std::lock_guard<std::mutex> lk(g_na_mu);
    if (!g_na_scratch[device]) {
        void *ptr = nullptr;
        hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) {
            vg_set_error("vgpu_init: address of the filter scratch array on device %d: %s", device,
                         hipGetErrorString(e));
            return -1;
        }
        void *host = nullptr;
        e = hipHostMalloc(&host, 64, hipHostMallocDefault);
        if (e != hipSuccess) {
            vg_set_error("vgpu_init: pinned word for the filter count: %s", hipGetErrorString(e));
            return -1;
        }
        g_na_total[device] = static_cast<uint64_t *>(host);
        g_na_scratch[device] = ptr;
    }
    g_na_device = device;
    return 0;
}

extern "C" int vgpu_filter_noalloc(const vgpu_prog *prog, const vgpu_col *cols, int n_cols, int64_t n_rows,
                                   const int32_t *sel_in, int64_t n_sel_in, int32_t *sel_out, int64_t *n_out) {
    int dev = g_na_device;
    if (dev < 0 || !g_na_scratch[dev]) {
        vg_set_error("vgpu_filter_noalloc: call vgpu_init first (set-up allocates, this entry point does not)");
        return -1;
    }
    if (filter_check("vgpu_filter_noalloc", prog, n_cols)) return -1;
    int64_t n_logical = sel_in ? n_sel_in : n_rows;
    if (n_logical > NA_MAX_ROWS) {
        vg_set_error("vgpu_filter_noalloc: %lld logical rows exceed the maximum of %lld (VGPU_FILTER_NOALLOC_MAX_ROWS)",
                     (long long)n_logical, (long long)NA_MAX_ROWS);
        return -1;
    }
    std::lock_guard<std::mutex> lk(g_na_mu);
    *n_out = 0;
    if (n_logical <= 0) return 0;
    uint8_t *base = static_cast<uint8_t *>(g_na_scratch[dev]);
    if (filter_enqueue(prog, cols, n_cols, n_logical, sel_in, sel_out,
                       reinterpret_cast<uint64_t *>(base + NA_MASK_OFF),
                       reinterpret_cast<uint32_t *>(base + NA_COUNTS_OFF),
                       reinterpret_cast<uint64_t *>(base + NA_OFFS_OFF),
                       reinterpret_cast<uint64_t *>(base + NA_TOTAL_OFF), base + NA_SCAN_OFF, NA_SCAN_BYTES,
                       nullptr, false))
        return -1;
    hipStream_t s = vg_stream();
    uint64_t *total = g_na_total[dev];
    VG_TRY(hipMemcpyAsync(total, base + NA_TOTAL_OFF, 8, hipMemcpyDeviceToHost, s));
    VG_TRY(hipStreamSynchronize(s));
    *n_out = (int64_t)*total;
    return 0;
}
