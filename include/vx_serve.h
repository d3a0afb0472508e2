/*
 * vx_serve.h -- a request loop that serves vectra queries against
 * device-resident columns on one GPU.
 *
 * No counterpart in the reference: vectra runs one query at a time from R
 * (collect(), src/collect.c:188-206).  The two request classes are the
 * pipelines the reference builds for
 *
 *   class 0:  tbl |> filter(value > t) |> group_by(group) |>
 *             summarise(s = sum(value), n = n())
 *   class 1:  tbl |> filter(value > t) |> inner_join(dim, by = c(key = "key")) |>
 *             group_by(attr) |> summarise(s = sum(value), n = n())
 *
 * (src/filter.c:13-84, src/join.c:262-480, src/group_agg.c:591-672), key
 * order of the groups, NA group last.
 *
 * Two modes over the same requests:
 *   plain     requests run one at a time in arrival order, each as its own
 *             group-by (class 1 probes the resident join table);
 *   batched   continuous batching: up to in_flight passes run on the device
 *             at once, each on a stream of its own; whenever a pass slot is
 *             free, the oldest waiting request's class forms the next pass
 *             from every waiting request of that class (up to max_batch, or
 *             fewer for a class whose shared scan keeps a table of groups x
 *             requests in on-chip memory: the largest pass whose table still
 *             fits; and as many as the per-pass state budget admits).  A pass is ONE
 *             shared scan of the fact columns for all its requests; requests
 *             that arrive while passes run join the next pass as soon as a
 *             slot frees, of either class, while the others still run.
 * Admission by device memory: the resident columns and the join table are
 * counted against hbm_budget, and pass slots are opened only while their
 * query state (pass templates planned for max_batch requests) still fits.
 * The dimension's join table is built once and stays resident.
 */
#ifndef VX_SERVE_H
#define VX_SERVE_H

#include <stdint.h>

#include "vgpu.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int32_t cls;          /* 0 or 1 (see above) */
    int32_t pad;
    double  threshold;    /* t */
    double  arrival_s;    /* seconds after the start of the run */
} vx_request;

typedef struct {
    double   start_s;     /* work began on the device (seconds after start) */
    double   end_s;       /* result on the host */
    int32_t  pass;        /* batched: pass that served it; plain: -1 */
    int32_t  batch;       /* requests in that pass (1 when plain) */
    int64_t  n_groups;
    /* with keep_results: malloc'd, n_groups entries; key validity and the
     * sum's validity as LSB-first bitmaps; free with vx_result_free */
    int64_t *keys;
    uint8_t *key_valid;
    double  *sum;
    uint8_t *sum_valid;
    double  *count;
} vx_result;

typedef struct {
    /* device-resident columns (same length n_rows): probe key, value,
     * group (class 0 key) */
    vgpu_col fact_key, fact_value, fact_group;
    int64_t  n_rows;
    /* dimension (n_dim rows): unique join key and the group-by attribute */
    vgpu_col dim_key, dim_attr;
    int64_t  n_dim;
    int32_t  batched;        /* 0 plain, 1 continuous batching */
    int32_t  max_batch;      /* requests per pass (batched) */
    int64_t  state_budget;   /* device bytes of query state one pass may hold (0: no cap) */
    int32_t  keep_results;
    int32_t  in_flight;      /* batched: passes on the device at once (1..8; 0 = 2) */
    int64_t  hbm_budget;     /* device bytes the run may hold: resident columns, join
                                table, pass state (0: the device's memory) */
} vx_serve_cfg;

typedef struct {
    int64_t n_done;
    int32_t n_passes;
    int32_t max_pass;          /* largest pass */
    double  wall_s;            /* first arrival to last result */
    double  setup_s;           /* resident join table, pass templates */
    int64_t state_bytes;       /* device bytes one admitted request holds */
    int64_t resident_bytes;    /* columns + join table + pass templates */
    int32_t max_in_flight;     /* most passes on the device at once */
    int32_t slots;             /* pass slots the budget admitted */
    int64_t hbm_budget;        /* the budget applied */
} vx_serve_stats;

/* Serves reqs[0..n) (arrival_s ascending) in real time; out[i] is
 * request i's result.  Returns 0, or -1 with vgpu_last_error(). */
int  vx_serve_run(const vx_serve_cfg *cfg, const vx_request *reqs, int n, vx_result *out,
                  vx_serve_stats *stats);
void vx_result_free(vx_result *r);

#ifdef __cplusplus
}
#endif

#endif
