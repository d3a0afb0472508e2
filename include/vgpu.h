/*
 * vgpu.h -- the thin C ABI between vectra's C11 operator nodes and the
 * hand-written gfx950 kernels (libvgpu.so).  Plain pointers and sizes only:
 * every buffer argument documented "device" is device memory obtained from
 * vgpu_malloc (or any HIP allocation on the current device); every other
 * pointer is host memory.  All calls are synchronous on return unless they
 * say otherwise, run on the library's own stream, and report failure as a
 * negative status with the message in vgpu_last_error().
 *
 * Each entry point stands in for a reference loop (gcol33/vectra 0.2.2):
 *   vgpu_filter ............ src/filter.c:13-39 apply_mask_sel over
 *                            src/expr.c:91-146 + src/scalar_ops.c:11-185
 *   vgpu_groupby_* ......... src/group_agg.c:155-281 (hash path) and
 *                            466-585 (sorted path), src/hash.c:80-165,
 *                            src/agg_ops.c:93-255
 *   vgpu_join_build ........ src/join.c:156-220 (jht_create/jht_insert :23-61)
 *   vgpu_join_probe ........ src/join.c:262-480 (jht_probe :115-135,
 *                            jht_chain_next :137-150)
 *   vgpu_gather ............ the per-value builder appends of
 *                            src/join.c:385-448 (src/builder.c:110-138)
 *   vgpu_groupby_feed_join . the fused form of Join -> Project('*') ->
 *                            GroupAgg (src/join.c:262-480,
 *                            src/project.c:51-252, src/group_agg.c)
 */
#ifndef VGPU_H
#define VGPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* column types use vectra's VecType numbering */
#define VGPU_INT64  0
#define VGPU_DOUBLE 1
#define VGPU_BOOL   2

/* A device column: data has 8 bytes per row (int64/double) or 1 (bool);
 * validity is an LSB-first bitmap (bit i of byte i/8 = row i valid) or NULL
 * for all-valid. */
typedef struct {
    int32_t        type;
    int32_t        reserved;
    const void    *data;
    const uint8_t *validity;
} vgpu_col;

/* ---- runtime --------------------------------------------------------- */
int         vgpu_init(int device);
const char *vgpu_last_error(void);
void       *vgpu_malloc(size_t bytes);
int         vgpu_free(void *ptr);
/* vgpu_memcpy_h2d returns once src has been read (it may be reused at
 * once); the copy itself completes in order on the calling thread's library
 * stream, before any later work there (pinned staging, no synchronization).
 * vgpu_memcpy_d2h returns with the data in dst. */
int         vgpu_memcpy_h2d(void *dst_device, const void *src_host, size_t bytes);
int         vgpu_memcpy_d2h(void *dst_host, const void *src_device, size_t bytes);
int         vgpu_memset(void *dst_device, int value, size_t bytes);
int         vgpu_synchronize(void);
/* Streams (no counterpart in the reference, which runs one query at a time):
 * a thread's library calls go to the library's own stream unless the thread
 * selected another with vgpu_stream_use (NULL selects the library stream
 * again).  Streams are created at the default priority.  Device buffers
 * the library caches are reused only by later work on the stream that
 * released them.  vgpu_stream_destroy waits for the stream's work. */
void       *vgpu_stream_create(void);
int         vgpu_stream_use(void *stream);
int         vgpu_stream_destroy(void *stream);
/* free and total device memory in bytes */
int         vgpu_mem_info(int64_t *free_bytes, int64_t *total_bytes);
/* sets this thread's vgpu_last_error() text (for callers that report their
 * own failures through it); returns -1 */
int         vgpu_error(const char *msg);
/* pinned host staging (hipHostMalloc); never handed to vectra consumers */
void       *vgpu_host_alloc(size_t bytes);
int         vgpu_host_free(void *ptr);

/* Asynchronous uploads (no reference counterpart: the reference reads row
 * groups into host memory, src/vtr1.c:306-392).  A copy stream of the
 * library moves staged host buffers to the device while the library stream
 * runs kernels on earlier ones.  Slot s (0..3) orders one set of buffers:
 *   vgpu_upload_async(s, dst, src, n)  enqueues a copy of n bytes from
 *       pinned src; it starts after the library-stream work marked by the
 *       last vgpu_upload_release(s) (so dst may still be read until then);
 *   vgpu_upload_acquire(s)   later library-stream work waits for the copies
 *       enqueued on slot s so far;
 *   vgpu_upload_release(s)   marks the library-stream work enqueued so far
 *       as the last reader of slot s's device buffers;
 *   vgpu_upload_host_wait(s) blocks until slot s's copies are done (their
 *       pinned sources may be overwritten). */
int         vgpu_upload_async(int slot, void *dst_device, const void *src_pinned, size_t bytes);
int         vgpu_upload_acquire(int slot);
int         vgpu_upload_release(int slot);
int         vgpu_upload_host_wait(int slot);

/* ---- expression programs ---------------------------------------------
 * A straight-line program in SSA form: instruction i writes register i and
 * reads registers a and b (indices < i).  Types are the VGPU_* column
 * types.  vectra's coercion rules (src/coerce.c:7-19) are applied by the
 * compiler that emits the program; the device never coerces implicitly. */
enum {
    VGPU_OP_COL = 1,    /* load column `col` (type = column type)          */
    VGPU_OP_CONST,      /* immediate, valid                                 */
    VGPU_OP_NA,         /* NA of `type`                                     */
    VGPU_OP_CAST,       /* a -> type (bool->int64, bool/int64->double)      */
    VGPU_OP_ADD, VGPU_OP_SUB, VGPU_OP_MUL, VGPU_OP_DIV, VGPU_OP_MOD,
    VGPU_OP_EQ, VGPU_OP_NE, VGPU_OP_LT, VGPU_OP_LE, VGPU_OP_GT, VGPU_OP_GE,
    VGPU_OP_AND, VGPU_OP_OR, VGPU_OP_NOT,
    VGPU_OP_ISNA,
    VGPU_OP_NEG
};

typedef struct {
    uint8_t op;
    uint8_t type;     /* result type for COL/CONST/NA/CAST/arith/NEG,
                         operand type for comparisons */
    uint8_t a, b;     /* operand registers */
    int32_t col;      /* column slot for VGPU_OP_COL */
    union { int64_t i; double d; } imm;
} vgpu_insn;

#define VGPU_MAX_INSN 24
#define VGPU_MAX_COLS 8

typedef struct {
    int32_t   n_insn;   /* result is register n_insn - 1 */
    int32_t   reserved;
    vgpu_insn insn[VGPU_MAX_INSN];
} vgpu_prog;

/* ---- filter -> selection vector -------------------------------------
 * Evaluates prog over the logical rows of a batch (physical rows
 * 0..n_rows-1, or sel_in[0..n_sel_in-1] when sel_in is non-NULL) and
 * writes to sel_out (device, capacity >= logical rows) the ascending
 * physical indices whose result is valid and true; *n_out gets the count
 * (host).  cols (host array of n_cols device columns) are the program's
 * column slots. */
int vgpu_filter(const vgpu_prog *prog, const vgpu_col *cols, int n_cols,
                int64_t n_rows, const int32_t *sel_in, int64_t n_sel_in,
                int32_t *sel_out, int64_t *n_out);

/* The same filter without any allocation.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Its
 * scratch (ballot mask, tile counts and offsets, scan partials; about
 * 8.5 MB) is a fixed device array that each GPU holds from vgpu_init until
 * the process ends; vgpu_init must have run on the device (it looks the
 * array up), and the call allocates nothing -- no hipMalloc / hipFree, no
 * pool growth.  At most VGPU_FILTER_NOALLOC_MAX_ROWS logical rows per call
 * (more is an error).  Calls share the array: the library serializes its
 * own callers, but work that overlaps a call on the same GPU from outside
 * the library (another process's stream excepted, which has its own copy)
 * must not use it.  Results equal vgpu_filter's. */
#define VGPU_FILTER_NOALLOC_MAX_ROWS (1 << 26)
int vgpu_filter_noalloc(const vgpu_prog *prog, const vgpu_col *cols, int n_cols,
                        int64_t n_rows, const int32_t *sel_in, int64_t n_sel_in,
                        int32_t *sel_out, int64_t *n_out);

/* Host execution of the same program interpreter and fast path, over HOST
 * columns: a test hook that lets the expression compiler and the device
 * evaluator's semantics be checked without a GPU.  No operator node calls
 * it. */
int vgpu_eval_host(const vgpu_prog *prog, const vgpu_col *cols, int n_cols,
                   int64_t n_rows, const int32_t *sel_in, int64_t n_sel_in,
                   int32_t *sel_out, int64_t *n_out);

/* ---- group-by --------------------------------------------------------- */
enum { VGPU_AGG_COUNT = 0, VGPU_AGG_COUNT_STAR, VGPU_AGG_SUM, VGPU_AGG_MEAN,
       VGPU_AGG_MIN, VGPU_AGG_MAX };   /* vectra's AggKind numbering */

#define VGPU_KEY_NONE (-1)   /* ungrouped summarise: one group if any row */

enum { VGPU_ORDER_FIRST_SEEN = 0,   /* hash path: first occurrence       */
       VGPU_ORDER_KEY = 1 };        /* sorted path: ascending, NA last    */

typedef struct {
    int32_t kind;     /* VGPU_AGG_* */
    int32_t input;    /* input slot, -1 for COUNT_STAR */
    int32_t na_rm;
    int32_t reserved; /* 0, or VGPU_AGG_RAW */
} vgpu_agg;

/* vgpu_agg.reserved flag: MIN / MAX of an int64 input deliver the exact
 * value's 64 bits in the double output slot (reinterpret them as int64)
 * instead of its conversion to double */
#define VGPU_AGG_RAW 1

#define VGPU_MAX_INPUTS 8
#define VGPU_MAX_AGGS   16

typedef struct vgpu_groupby vgpu_groupby;

/* key_type: VGPU_INT64 / VGPU_DOUBLE / VGPU_BOOL / VGPU_KEY_NONE.
 * input_types: n_inputs VGPU_* types of the aggregate input slots.
 * expected_groups: sizing hint (0 = unknown). */
vgpu_groupby *vgpu_groupby_create(int key_type, int n_inputs,
                                  const int32_t *input_types, int n_aggs,
                                  const vgpu_agg *aggs, int order,
                                  int64_t expected_groups);

/* Feeds the logical rows of one batch.  pos_base is the logical position
 * of the batch's first logical row in the whole input (first-occurrence
 * order and min/max tie rules count positions).  pred (optional, may be
 * NULL) filters rows first, reading pred_cols. */
int vgpu_groupby_feed(vgpu_groupby *g, const vgpu_col *key,
                      const vgpu_col *inputs, int64_t n_rows,
                      const int32_t *sel, int64_t n_sel, int64_t pos_base,
                      const vgpu_prog *pred, const vgpu_col *pred_cols,
                      int n_pred_cols);

/* One 64-bit code per physical row of the tuple of n_keys (1..8) key
 * columns (int64 / double / bool, device), into codes (device, n_rows):
 * equal tuples (NA equal to NA, -0 to +0) get equal codes; a tuple with a
 * NaN component gets a code drawn from its position pos_base + row, as it
 * forms a group of its own (src/hash.c:48-78, NaN equals nothing).  Codes
 * of different tuples collide with probability about 2^-64 per pair: a
 * multi-key group-by over codes checks its groups (src/group_agg.c:155-281
 * with n_keys > 1; the node carries min / max / count of every key). */
int vgpu_key_codes(const vgpu_col *keys, int n_keys, int64_t n_rows, int64_t pos_base, int64_t *codes,
                   uint8_t *valid);
/* (valid: optional device bitmap of (n_rows + 63) / 64 * 8 bytes, 8-byte
 * aligned; bit r set when no component of row r is NA or NaN -- the rows a
 * join may match, src/join.c:77-113) */

/* String key columns (device: offsets[n_rows + 1] into chars, validity
 * optional): per row two independent 64-bit hashes of its bytes into h1 and
 * h2 (h2 may be NULL; NA rows get 0) and its position pos_base + row into
 * pos (may be NULL).  h1 (with the column's validity) stands for the string
 * in vgpu_key_codes and in join tables; h2 and the position let a group-by
 * prove that each group holds one string (min == max of h2) and find a row
 * that holds it (min of the position).  Replaces vec_hash_value's string
 * case (src/hash.c:35-41) and the join's probe hashing of string keys
 * (src/join.c:335-341). */
int vgpu_string_keys(const int64_t *offsets, const uint8_t *chars, const uint8_t *validity, int64_t n_rows,
                     int64_t pos_base, int64_t *h1, int64_t *h2, int64_t *pos);

/* ---- string comparison ------------------------------------------------
 * vec_cmp on two string operands (src/scalar_ops.c:77-106): per row
 * cmp = memcmp over the shorter length, then the shorter string first, as
 * -1 / 0 / 1 into out (int64), NA where either side is NA.  A side is a
 * device string column (offsets n+1 entries, the bytes they index,
 * validity bitmap or NULL) or a constant (constant != 0: offsets[0..1]
 * bound the one string for every row; null_const: the NA string).  out
 * has n_rows entries; out_valid (8-byte aligned, (n_rows + 63) / 64 words)
 * gets a bit per row.  A filter compiles `a op b` on strings to
 * `cmp op 0` over this column. */
typedef struct {
    const int64_t *offsets;
    const uint8_t *chars;
    const uint8_t *validity;
    int32_t constant;
    int32_t null_const;
} vgpu_strcol;
int vgpu_string_compare(const vgpu_strcol *a, const vgpu_strcol *b, int64_t n_rows, int64_t *out,
                        uint8_t *out_valid);

/* ---- exact string dictionary -----------------------------------------
 * Replaces the byte comparisons of the reference's string keys
 * (src/hash.c:58-66 in group-by tables, src/join.c:96-108 in join chains):
 * every distinct string gets one int64 code, the arena row index of a row
 * that holds it, and two rows get the same code iff their bytes are equal
 * (the table compares bytes; hashes only choose slots), so grouping or
 * joining on the codes is exact.  The dictionary keeps every appended
 * row's bytes on the device (offsets 8 bytes a row plus the bytes).
 *   vgpu_strdict_append  appends rows [0, n) of a host string column
 *       (offsets[0..n], chars indexed by them, host memory) as arena rows
 *       rows..rows+n-1, inserting the strings; codes (device, n) gets each
 *       row's code, 0 for NA rows (validity: device bitmap or NULL); pos
 *       (device, may be NULL) gets pos_base + row;
 *   vgpu_strdict_lookup  codes of a host string column without inserting:
 *       -1 for a string that was never appended, 0 for NA rows;
 *   vgpu_strdict_size    distinct strings (and arena rows into *rows). */
typedef struct vgpu_strdict vgpu_strdict;
vgpu_strdict *vgpu_strdict_create(void);
void          vgpu_strdict_free(vgpu_strdict *d);
int           vgpu_strdict_append(vgpu_strdict *d, const int64_t *offsets, const uint8_t *chars,
                                  const uint8_t *validity, int64_t n_rows, int64_t pos_base, int64_t *codes,
                                  int64_t *pos);
int           vgpu_strdict_lookup(vgpu_strdict *d, const int64_t *offsets, const uint8_t *chars,
                                  const uint8_t *validity, int64_t n_rows, int64_t *codes);
int64_t       vgpu_strdict_size(const vgpu_strdict *d, int64_t *rows);

/* Finishes on the device and orders the groups; returns the group count. */
int64_t vgpu_groupby_finish(vgpu_groupby *g);

/* Copies the finished result into host buffers of n_groups entries:
 * keys (8 bytes per group, or 1 for bool keys; may be NULL for KEY_NONE),
 * key_validity (bitmap, (n_groups+7)/8 bytes, may be NULL), and per
 * aggregate a double array plus a validity bitmap.  Returns the number of
 * min/max results that had no value (the reference warns once for each),
 * or a negative status. */
int64_t vgpu_groupby_fetch(vgpu_groupby *g, void *keys, uint8_t *key_validity,
                           double *const *agg_values,
                           uint8_t *const *agg_validity);

/* Device-side view of the finished result (valid until destroy): keys,
 * key validity bitmap, and aggregate a's values / validity bitmap. */
int vgpu_groupby_result_device(vgpu_groupby *g, const void **keys,
                               const uint8_t **key_validity, int agg,
                               const double **values,
                               const uint8_t **validity);

void vgpu_groupby_destroy(vgpu_groupby *g);
/* 8-byte words of one group's record (hot + cold): the width axis of the
 * tuning table's buckets */
int  vgpu_groupby_record_words(const vgpu_groupby *g);
/* How the last feed aggregated: 0 into the hash table, 1 through hash
 * partitions, 2 through key-range partitions (records already in key
 * order at finish).  A diagnostic with no counterpart in vectra. */
int  vgpu_groupby_last_path(const vgpu_groupby *g);

/* ---- hash join -------------------------------------------------------- */
enum { VGPU_JOIN_INNER = 0, VGPU_JOIN_LEFT, VGPU_JOIN_FULL, VGPU_JOIN_SEMI,
       VGPU_JOIN_ANTI };   /* vectra's JoinKind numbering */

typedef struct vgpu_jtable vgpu_jtable;

/* Builds the device hash table over n_build rows of key (device column,
 * already coerced to the common key type).  NA and NaN keys are kept out
 * of the table: they never match (src/join.c:85). */
vgpu_jtable *vgpu_join_build(const vgpu_col *key, int64_t n_build);

/* The same, embedding up to 4 build-side payload columns in the table: each
 * key's entry carries the payload of its first (highest) build row, so a
 * fused probe -> aggregate over a unique-key table reads key, match count
 * and payload in one load.  The key and payload columns must outlive the
 * table (the fused feed matches build columns to payload by data pointer). */
vgpu_jtable *vgpu_join_build_payload(const vgpu_col *key, int64_t n_build,
                                     const vgpu_col *payload, int n_payload);

/* Probes the logical rows of a batch.  For each probe row in logical order
 * the output holds its matches in descending build row order; LEFT/FULL
 * append the batch's unmatched probe rows after all matches; SEMI/ANTI
 * emit matching / non-matching probe rows once.  uncoerced_key (device,
 * optional) is the probe key column before its coercion to the build key
 * type (int64 or bool; probe_key holds the coerced values): LEFT/FULL then
 * keep, after each row's first match, only the build rows whose key bytes
 * equal the uncoerced key's -- an int64 probe against the build double's
 * 64 bits, a bool probe against byte `build_row` of the build key array --
 * the reference's chain walk over uncoerced probe columns
 * (src/join.c:398-411).  Results stay on the device: vgpu_join_pairs()
 * returns them. */
int vgpu_join_probe(vgpu_jtable *t, int kind, const vgpu_col *probe_key,
                    int64_t n_rows, const int32_t *sel, int64_t n_sel,
                    const vgpu_col *uncoerced_key, int64_t *n_out);

/* Device pointers to the last probe's output: probe physical rows and
 * build rows (-1 = no build row). Valid until the next probe/destroy. */
int vgpu_join_pairs(vgpu_jtable *t, const int32_t **probe_rows,
                    const int64_t **build_rows);

/* FULL join: after the last probe, lists build rows never matched, in
 * ascending order, into the pair buffers (probe row -1). */
int vgpu_join_unmatched_build(vgpu_jtable *t, int64_t *n_out);

int64_t vgpu_join_build_rows(const vgpu_jtable *t);
/* the largest number of build rows sharing one (non-NA) key */
int64_t vgpu_join_max_dup(const vgpu_jtable *t);
/* device bytes the table holds (entries, segments, filter, payload) */
int64_t vgpu_join_device_bytes(const vgpu_jtable *t);
void    vgpu_join_destroy(vgpu_jtable *t);

/* ---- cast --------------------------------------------------------------
 * dst_data[i] = src[i] converted to dst_type (int64 -> double, bool ->
 * int64 / double), vec_coerce's widenings (src/coerce.c:21-71).  Validity
 * is unchanged, so the caller keeps using src's bitmap. */
int vgpu_cast(const vgpu_col *src, int64_t n, int dst_type, void *dst_data);

/* ---- gather (join output materialization) ----------------------------
 * dst[i] = src[idx[i]] with validity, idx < 0 giving NA.  idx is int32 or
 * int64 (idx_is_64).  dst_data / dst_validity are device buffers of n rows
 * (validity: (n+7)/8 bytes, fully written). */
int vgpu_gather(const vgpu_col *src, const void *idx, int idx_is_64,
                int64_t n, void *dst_data, uint8_t *dst_validity);

/* ---- fused join -> aggregate -----------------------------------------
 * One pass over the probe rows: optional filter, probe of t, and for each
 * output pair (kind INNER or LEFT) group-by on a key column taken from the
 * probe or build side, with aggregate inputs that are a column of either
 * side or the product of two such columns.  Equivalent to
 * Join -> Project(tmp = a * b) -> GroupAgg on the same rows. */
typedef struct {
    int32_t op;        /* 0 = column a, 1 = a * b */
    int32_t side_a;    /* 0 = probe, 1 = build */
    int32_t col_a;
    int32_t side_b;
    int32_t col_b;
    int32_t reserved;
} vgpu_input;

typedef struct {
    int64_t          n_rows;
    const int32_t   *sel;          /* device, optional */
    int64_t          n_sel;
    int64_t          pos_base;
    const vgpu_prog *pred;         /* optional, over probe_cols */
    vgpu_jtable     *join;         /* optional: NULL = no join */
    int32_t          join_kind;    /* VGPU_JOIN_INNER or VGPU_JOIN_LEFT */
    int32_t          key_side;     /* 0 = probe, 1 = build */
    int32_t          key_col;
    int32_t          probe_key_col;
    int32_t          n_probe_cols;
    int32_t          n_build_cols;
    const vgpu_col  *probe_cols;   /* host array of device columns */
    const vgpu_col  *build_cols;
    int32_t          n_inputs;
    int32_t          reserved;
    const vgpu_input *inputs;      /* host array */
} vgpu_feed_spec;

/* Order of the pairs (first-seen groups, NaN / zero-sign min and max): a
 * pair's position is pos_base + (unmatched << 43 | probe logical row << 12
 * | match index), so one feed takes fewer than 2^31 probe rows and a build
 * key of at most 4,096 rows (more is an error), and a caller that feeds
 * several batches advances pos_base by 2^44 per feed. */
int vgpu_groupby_feed_join(vgpu_groupby *g, const vgpu_feed_spec *spec);

/* Of n join pairs (probe_rows[i], build_rows[i]; device; build row < 0 =
 * no match), the number whose key tuples differ -- a column NA on either
 * side or unequal values (double by ==) -- into *n_unequal.  Each column's
 * probe and build sides have one type (int64 / double / bool), the pair's
 * common type.  The check of a join on tuple codes (src/join.c:77-113 is
 * the comparison it restates): a nonzero count means two tuples shared a
 * code. */
int vgpu_pairs_unequal(const vgpu_col *probe_keys, const vgpu_col *build_keys, int n_keys,
                       const int32_t *probe_rows, const int64_t *build_rows, int64_t n, int64_t *n_unequal);

/* ---- device radix sort (exposed for tests) ----------------------------
 * Stable ascending sort of n (key, value) pairs on key bits
 * [begin_bit, end_bit), device buffers, in place. */
int vgpu_sort_pairs_u64(uint64_t *keys, uint32_t *values, int64_t n,
                        int begin_bit, int end_bit);

/* ---- batched queries ----------------------------------------------------
 * Many independent filter -> group-by queries of one shape run as one
 * batch: per query q, the rows of inputs[] (and pred_cols[]) where the
 * shared predicate holds -- with the immediate of instruction param_insn
 * replaced by param's bits -- grouped by key and aggregated as
 * vgpu_groupby would (same aggregates, order, NA / NaN rules).  Each
 * query is the reference's tbl |> filter() |> group_by() |> summarise()
 * (src/filter.c:13-84, src/group_agg.c:591-672).
 *
 *   vgpu_batch_plan     sizes every query's table from a device sketch of
 *                       its keys (one synchronization), lays out the work
 *                       items and runs the batch once; the columns must
 *                       stay in place until the batch is destroyed
 *   vgpu_batch_run      enqueues the batch again on the same columns:
 *                       device work only (one persistent work-queue feed
 *                       launch, one finish for all queries); queries that
 *                       all read the same columns share one scan of them
 *   vgpu_batch_capture  records that device work as a hipGraph; later
 *                       runs replay it
 *   vgpu_batch_counts   groups and output offset of every query
 *                       (synchronizes); fails if a run's keys left the
 *                       sort range found at plan (columns changed)
 *   vgpu_batch_fetch    query q's result, like vgpu_groupby_fetch        */
typedef struct vgpu_batch vgpu_batch;

typedef struct {
    vgpu_col        key;        /* ignored for VGPU_KEY_NONE and joins    */
    vgpu_col        probe_key;  /* join batches: the probe side's key     */
    const vgpu_col *inputs;     /* n_inputs columns                       */
    const vgpu_col *pred_cols;  /* n_pred_cols predicate column slots     */
    int64_t         n_rows;
    uint64_t        param;      /* bits of the per-query constant         */
} vgpu_query;

vgpu_batch *vgpu_batch_create(int key_type, int n_inputs,
                              const int32_t *input_types, int n_aggs,
                              const vgpu_agg *aggs, int order,
                              const vgpu_prog *pred, int n_pred_cols,
                              int param_insn);
/* Joins (before plan): every query's rows inner-join table t on
 * probe_key (src/join.c:262-370, matches in the reference's chain order);
 * the group key is the table's embedded payload column key_payload. */
int     vgpu_batch_join(vgpu_batch *b, const vgpu_jtable *t, int key_payload);
int     vgpu_batch_plan(vgpu_batch *b, const vgpu_query *queries, int n_queries);
int     vgpu_batch_run(vgpu_batch *b);
/* Runs the first n_active queries again with new constants params[q]
 * (n_active < n_queries only when every query reads the same columns: a
 * shared scan); synchronizes once to size the finish.  Drops a capture. */
int     vgpu_batch_rerun(vgpu_batch *b, const uint64_t *params, int n_active);
int     vgpu_batch_capture(vgpu_batch *b);
int     vgpu_batch_counts(vgpu_batch *b, int64_t *n_groups, int64_t *offsets);
int64_t vgpu_batch_fetch(vgpu_batch *b, int q, void *keys, uint8_t *key_validity,
                         double *const *agg_values, uint8_t *const *agg_validity);
/* device bytes the batch holds (tables, sort and output buffers) */
int64_t vgpu_batch_device_bytes(const vgpu_batch *b);
/* 1 when the planned batch runs as ONE bucketed pass: every query reads
 * the same columns through a `col < <= > >= param` filter whose constant
 * alone differs, the group key is an int64 / bool column of a dense range
 * (at most 2^24 values and one LDS table of (key, bucket) records), and
 * the aggregates are counts, sums and means of inputs without NAs in key
 * order (knob "batch.bucketed"); 0 otherwise */
int     vgpu_batch_bucketed(const vgpu_batch *b);
void    vgpu_batch_destroy(vgpu_batch *b);

/* ---- tuning -----------------------------------------------------------
 * Launch-shape knobs the autotuner sets (defaults in parentheses):
 *   "gb.lds.max_blocks" (1024)  "gb.lds.rows_per_block" (8192)
 *   "gb.glb.max_blocks" (16384) "gb.glb.rows_per_block" (1024)
 *   "gb.lds.max_slots" (4096)   "gb.lds.budget" (65536 bytes)
 *   "gb.star.items" (4: rows in flight per thread of the fused star probe)
 *   "gb.part.min_groups" (65536: groups from which a first feed of at
 *                         least 2^20 rows is hash-partitioned)
 *   "batch.shared_scan"  (1: off; 2: batches whose queries all read the
 *                         same columns share one scan of them)
 *   "gb.lds.slots_per_group" (2: LDS table slots per expected group, the
 *                         table's inverse load factor; 2..8)
 *   "batch.bucketed"     (2: on; 1: off) batches whose queries read the
 *                         same columns and differ only in the constant of
 *                         a `col < <= > >= param` filter aggregate every
 *                         row once, keyed by (group, the number of the
 *                         queries it passes)
 *   "gb.star.dense"      (2: on; 1: off) fused join -> aggregate feeds of
 *                         the common star shape (unique build keys, an
 *                         int64 / bool payload group key of a short range)
 *                         aggregate into a dense LDS table
 *   "join.bloom"         (1: off; 2: on) those feeds test the join table's
 *                         filter of build keys before loading an entry (one
 *                         more dependent random load per probe row: it pays
 *                         only when most probe keys are absent)
 *   "gb.lds.copies"      (2: at most this many copies of an LDS table of up
 *                         to 256 slots, lane l updating copy l % copies, as
 *                         the LDS budget allows; 1 = one table)
 *   "gb.narrow"          (2: on; 1: off) group-by feeds of n / count / sum /
 *                         mean over at most two inputs, with a fast filter or
 *                         none, run a lean kernel (dense LDS table when the
 *                         feed's sketch saw a short int64 / bool key range)
 *   "batch.dense"        (2: on; 1: off) batch queries whose sketch saw an
 *                         int64 / bool key range no longer than their hashed
 *                         table get a dense table instead: key k owns record
 *                         k - min, found without hashing or probing
 *   "scan.device_feed"   (2: on; 1: off) GPU group-by nodes over a .vtr
 *                         scan (or a filter over one) read its row groups
 *                         into pinned buffers and copy them on the copy
 *                         stream, the filter evaluated in the feed
 *   "join.bloom_keys"    (4) build keys per 64-bit word of a join table's
 *                         filter (3 bits a key), read when a table is built
 *   "join.load_pct"      (50) highest load of a join table's slots, in
 *                         percent (10..95), read when a table is built
 *                         (bucket: build rows, entry words)
 *   "gb.glb.load_pct"    (50) highest load of a group-by's HBM table, in
 *                         percent (25..90), read when the table is sized
 *   "join.dense_map"     (2: on; 1: off) a bucketed join pass over unique
 *                         int64 build keys of a short range probes a
 *                         direct-addressed map (built once per table)
 *   "gb.star.blocks_per_cu" (4) blocks per CU of the dense star kernel
 *   "gb.part.range"      (2: on; 1: off) a partitioned group-by in key order
 *                         over int64 keys splits rows by key range instead
 *                         of by hash, so its records come out in key order
 *                         and finish does not sort them (hash partitions
 *                         are kept when the ranges are badly balanced)
 *   "gb.part.stage"      (3) the partition passes of a partitioned
 *                         group-by stage each tile of rows in LDS in bucket
 *                         order and write a bucket's rows of the tile as one
 *                         run, and its aggregation runs on resident blocks:
 *                         2 = tiles of 1,024 rows, 3 = of 2,048 rows, 1 =
 *                         row-wise scatters and one block per item
 *   "gb.part.slots"      (512) LDS table slots of a partition's block in a
 *                         partitioned group-by (64..512, a power of two);
 *                         partitions are sized to half of it
 *   "gb.feed.prefetch"   (2: on; 1: off) the general LDS-table feed loads
 *                         a tile's key and input words (up to four inputs)
 *                         before it applies the first row
 *   "gb.merge.partials"  (2: on; 1: off) blocks of a dense-table feed store
 *                         their tables as partials that one block per record
 *                         folds, instead of merging into the HBM table with
 *                         device-scope atomics on the same few records
 *   "gb.fuse_join"       (2) a GPU GroupAggNode over Project(a * b) over a
 *                         GPU JoinNode runs the three as one fused device
 *                         pass (2), or node by node (1)
 *   "strdict.hash_bits"  (64) bits of the string dictionary's hash that are
 *                         kept (1..64), read when a dictionary is created;
 *                         fewer bits make different strings share slots and
 *                         tags, which the byte comparisons resolve (a test
 *                         of exactness under collisions)
 * Returns -1 for an unknown name or a value <= 0 (or out of its range). */
int     vgpu_tune_set(const char *name, int64_t value);
int64_t vgpu_tune_get(const char *name);
/* A knob value for one bucket of the autotuner's table: work expecting at
 * most max_groups groups, whose records are at most max_width 8-byte words
 * (a group-by's hot + cold record, a join table's entry) and whose launch
 * covers at most max_rows rows.  A lookup takes the first bucket that
 * covers the work and sets the knob, in ascending (max_width, max_rows,
 * max_groups) order, else the global value. */
int     vgpu_tune_set_at(int64_t max_groups, int64_t max_width, int64_t max_rows, const char *name, int64_t value);
/* the same bucket covering any width and row count */
int     vgpu_tune_set_for(int64_t max_groups, const char *name, int64_t value);
/* the value work of `groups` groups, `width` words and `rows` rows uses
 * (0 for width / rows: any); -1 for an unknown name */
int64_t vgpu_tune_get_at(int64_t groups, int64_t width, int64_t rows, const char *name);
int64_t vgpu_tune_get_for(int64_t groups, const char *name);
/* Loads a table written by the autotuner (vectra_amd/autotune.py): lines
 * "max_groups name value [w<=W] [rows<=R]", max_groups 0 for global
 * values.  Returns the number of settings or -1. */
int     vgpu_tune_load(const char *path);
/* Drops every bucket and restores the built-in global values. */
void    vgpu_tune_reset(void);

/* ---- timing ---------------------------------------------------------
 * Event timing of the library's own stream: elapsed milliseconds of all
 * work issued between vgpu_timer_start() and vgpu_timer_stop(). */
int    vgpu_timer_start(void);
double vgpu_timer_stop(void);

/* Per-launch timing of the main kernel of vgpu_filter /
 * vgpu_groupby_feed / vgpu_groupby_feed_join / vgpu_batch_run (the scan
 * kernel of a bucketed pass, the feed kernels of the others): when
 * enabled, each of those calls brackets its kernel with events on the
 * stream it runs on and adds the elapsed time here.  Returns the
 * accumulated milliseconds and launch count since the last reset. */
int    vgpu_kernel_timing(int enable);
double vgpu_kernel_time_ms(int64_t *launches, int reset);

#ifdef __cplusplus
}
#endif
#endif
