/*
 * devscan.h -- device scan feed: the row groups of a .vtr scan node read
 * straight into pinned staging buffers and copied to the device on the
 * library's copy stream, two groups in flight, so that reading group i + 1
 * from the file and copying it overlap the kernels working on group i.
 *
 * It replaces, for GPU consumers, the reference scan's read of each row
 * group into malloc'd columns (src/vtr1.c:306-392, src/scan.c:119-138) and
 * the nodes' synchronous upload of them; row groups are pruned with the
 * scan's pushed-down predicate exactly as the scan prunes them
 * (src/scan.c:10-117).
 */
#ifndef VX_DEVSCAN_H
#define VX_DEVSCAN_H

#include "gpu_common.h"

#ifdef __cplusplus
extern "C" {
#endif

#define VX_DS_COLS VGPU_MAX_COLS

typedef struct {
    ScanNode *scan;                  /* borrowed; its file handle is read here */
    int       n;                     /* columns fed */
    int       fcol[VX_DS_COLS];      /* their file column indices */
    VecType   type[VX_DS_COLS];
    int64_t   max_rows;              /* largest row group */
    uint8_t  *pinned[2];             /* per slot: per column validity, then values */
    size_t    vb_cap;                /* validity bytes reserved per column (8-aligned) */
    void     *d_data[2][VX_DS_COLS];
    uint8_t  *d_valid[2][VX_DS_COLS];
    int       cur;                   /* slot of the group handed out last (-1: none) */
    uint32_t  next_rg;
    int64_t   rows_fed;              /* rows of the groups handed out before */
    int64_t   groups_fed, groups_pruned;
} VxDevScan;

/* columns fcols[0..n) of the scan's file (int64 / double / bool) */
void vx_devscan_open(VxDevScan *d, ScanNode *scan, const int *fcols, int n);
/* The next row group that survives pruning: device columns (in fcols
 * order) valid for library-stream work enqueued before the next call,
 * its row count and the rows handed out before it.  Returns 0 at the end.
 * Work the caller enqueues on the group must be on the library stream. */
int  vx_devscan_next(VxDevScan *d, vgpu_col *cols, int64_t *n_rows, int64_t *pos_base);
void vx_devscan_close(VxDevScan *d);
/* row groups handed to GPU consumers by device scan feeds so far (a test
 * hook: shows that a node took the device path) */
int64_t vx_scan_feed_groups(void);

#ifdef __cplusplus
}
#endif
#endif
