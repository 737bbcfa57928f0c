/*
 * sable.h — C-ABI of libsable_amd.so, the MI355X (gfx950) VBR SpMV/SpMM
 * executor.
 *
 * The reference has no in-process FFI for this path: its boundary is
 * "generate a C program per matrix, compile it with gcc, run it and parse
 * stdout" (src/codegen.py:293-452 / :827-937, bench_launcher.py:80-85,
 * eval.py:23-45).  Each entry point below replaces one step of that cycle;
 * the reference step it stands in for is cited beside it.  All pointers are
 * plain host or device pointers with explicit sizes; no torch types appear.
 *
 * Errors: every function returns SABLE_OK (0) or a nonzero SABLE_ERR_* code
 * and never exits or aborts; sable_last_error() returns the message of the
 * calling thread's last failure.
 *
 * Threading: a plan is used by one host thread at a time and runs on one HIP
 * stream (its own, or the one set with sable_plan_set_stream); different
 * plans may be used from different threads concurrently.
 */
#ifndef SABLE_H
#define SABLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SABLE_OK 0
#define SABLE_ERR_INVALID 1      /* malformed VBR / bad argument */
#define SABLE_ERR_HIP 2          /* HIP runtime error */
#define SABLE_ERR_OOM 3          /* device or host allocation failed */
#define SABLE_ERR_RTC 4          /* hipRTC compile failure (generic kernel used) */
#define SABLE_ERR_NO_DEVICE 5    /* no GPU / sable_init not called */
#define SABLE_ERR_UNSUPPORTED 6  /* option combination not supported */

/* per-block routes (the reference's binary dense / "unrolled" choice,
 * utils/convert_real_to_vbr.py:52-68, plus the MFMA route for SpMM) */
#define SABLE_ROUTE_DENSE 0       /* FP32 dense block (GEMV for SpMV, FMA for SpMM) */
#define SABLE_ROUTE_SPARSE 1      /* nonzeros only, in-block CSR */
#define SABLE_ROUTE_DENSE_MFMA 2  /* SpMM: bf16 split on MFMA (FP32 accumulate) */

#define SABLE_OP_SPMV 0
#define SABLE_OP_SPMM 1

#define SABLE_MODE_GRID 0         /* one workgroup per tile */
#define SABLE_MODE_STATIC 1       /* persistent workgroups, static round robin */
#define SABLE_MODE_QUEUE 2        /* persistent workgroups, atomic work queue */

/* flags for sable_spmv / sable_spmm */
#define SABLE_HOST_PTRS 1         /* x/y are host pointers: copy in and out */
#define SABLE_NO_SYNC 2           /* return without synchronising the stream */

typedef struct SablePlan SablePlan;
typedef struct SableGroup SableGroup;
typedef struct SableInspect SableInspect;
typedef struct SableCompress SableCompress;
typedef struct SableText SableText;

/* The VBR arrays of src/vbr.py:7-14 (validated: see sable_plan_create). */
typedef struct {
    int32_t nbr;              /* block rows = len(rpntr) - 1 */
    int32_t nbc;              /* block cols = len(cpntr) - 1 */
    const int32_t *rpntr;     /* nbr + 1 */
    const int32_t *cpntr;     /* nbc + 1 */
    const int32_t *bindx;     /* nblocks */
    const int32_t *bpntrb;    /* nbr, -1 for an empty block row */
    const int32_t *bpntre;    /* nbr, -1 for an empty block row */
    const int64_t *indx;      /* nblocks + 1 */
    const float *val;         /* indx[nblocks] values, each block column-major */
    int64_t nblocks;
} SableVbr;

typedef struct {
    int32_t op;               /* SABLE_OP_SPMV / SABLE_OP_SPMM */
    int32_t n_rhs;            /* SpMM RHS width N (multiple of 4); 1 for SpMV */
    int32_t wg;               /* workgroup size: 256 (default) */
    int32_t xcap;             /* max staged x columns per piece (16..4095) */
    int64_t tile_bytes;       /* target algorithmic bytes per tile */
    int32_t max_seg_rows;     /* upper bound on tile height (<= wg) */
    int32_t min_seg_rows;     /* lower bound on tile height when splitting */
    int32_t mode;             /* SABLE_MODE_*; -1 = default (grid) */
    int32_t grid;             /* persistent grid size; 0 = 8 per CU */
    int32_t unroll;           /* SpMV dense stream unroll (1,2,4,8); 0 = 4 */
    int32_t use_rtc;          /* 1 = hipRTC per-plan specialisation (default) */
    int32_t host_only;        /* 1 = build the host plan only (no device) */
    int32_t mfma_split;       /* MFMA route: 3 = exact 3-term bf16 split, 6 passes
                                 (FP32-level, default); 2 = hi+lo, 3 passes */
    int32_t min_tiles;        /* tiles wanted at least: while a plan has fewer, its
                                 tile_bytes is halved (down to 16 KiB) so small
                                 matrices still fill the device; 0 = 1024 for
                                 SpMV, 256 for SpMM */
    int32_t mfma_kchunk;      /* SpMM MFMA tiles: at most about this many 32-column
                                 k-steps per work item; a longer row tile is split
                                 along k into chunks whose partial sums a fixup
                                 launch adds in chunk order (deterministic);
                                 0 or negative = no split */
} SablePlanOpts;

typedef struct {
    int32_t nrows, ncols;
    int32_t ntiles, npieces, nxranges;
    int32_t ntiles_mfma;      /* SpMM: MFMA tiles, run first (the first ntiles_mfma tiles) */
    int32_t n_rhs;
    int64_t nblocks, nblocks_dense, nblocks_sparse, nblocks_mfma;
    int64_t nnz;              /* stored values != 0 (flops = 2 * nnz * N) */
    int64_t dense_values;     /* sum over dense-routed blocks of r*c */
    int64_t dense_padded;     /* dense values as laid out (padding included) */
    int64_t mfma_values;      /* sum over MFMA-routed blocks of r*c */
    int64_t mfma_padded;      /* MFMA-ordered FP32 elements, padding included */
    int64_t sparse_nnz;       /* nonzeros stored in sparse pieces */
    int64_t sparse_padded;    /* sparse slots as laid out (SpMV wave slices pad rows) */
    int64_t rowptr_entries;   /* sparse index entries: CSR row pointers (SpMM) or
                                 row orders and wavefront tables (SpMV) */
    int64_t xmap_entries;     /* SpMV x staging map entries (x column per staged slot) */
    int64_t descriptor_bytes; /* tiles + pieces + ranges */
    int64_t bytes_alg;        /* compulsory bytes of the executed format */
    int64_t flops_alg;        /* 2 * nnz * N */
    int64_t device_bytes;     /* device memory held by the plan */
    int32_t rtc_used;         /* 1 if the launch kernel came from hipRTC */
    int32_t mode, grid, wg;
    int32_t ntiles_split;     /* SpMM: MFMA row tiles split along k */
    int64_t scratch_bytes;    /* device scratch of the split tiles' partial sums (not
                                 counted in bytes_alg) */
    int64_t fixup_bytes;      /* partial sums written and read back per launch */
    int64_t nspans;           /* SpMV row spans: tiles of several short block rows,
                                 every block on the sparse path (0 when they would
                                 leave the plan below min_tiles) */
} SablePlanStats;

/* Library. */
int sable_init(int device);                              /* once per process */
const char *sable_last_error(void);
int sable_device_info(int *n_cu, int64_t *hbm_bytes, char *arch, int arch_len);
int sable_set_cache_dir(const char *dir);                /* hipRTC code-object cache */

/* Replaces: gen_single_threaded_spmv / gen_single_threaded_spmm emitting a
 * per-matrix C program (src/codegen.py:293-452, :827-937) plus the gcc
 * compile (bench_launcher.py:85).  Validates the VBR (block rows laid out
 * consecutively, bindx strictly ascending per row, indx steps equal block
 * sizes), routes every block by route[k], lays the matrix out in HBM and
 * selects/compiles the kernel.  Copies all inputs (the caller keeps them). */
int sable_plan_create(const SableVbr *vbr, const uint8_t *route,
                      const SablePlanOpts *opts, SablePlan **out);
int sable_plan_destroy(SablePlan *plan);
int sable_plan_stats(const SablePlan *plan, SablePlanStats *out);
int sable_plan_set_stream(SablePlan *plan, void *hip_stream);
/* Host copies of the plan's arrays (for inspection / tests).  which: 0 tiles,
 * 1 pieces, 2 xranges, 3 dval (FP32 dense and MFMA pieces), 4 sval, 5 sidx,
 * 6 srowptr, 7 xmap (SpMV).
 * Returns the byte size in *bytes; copies when dst != NULL. */
int sable_plan_export(const SablePlan *plan, int which, void *dst, int64_t *bytes);
/* The kernel specialisation chosen for the plan (no reference counterpart):
 * spec6 = {kind, wg, unroll, xcap, nrhs, rtmax}, kind as in
 * sable_rtc_precompile. */
int sable_plan_kernel_spec(const SablePlan *plan, int32_t *spec6);

/* Replaces: one execution of the generated program's body,
 * src/codegen.py:416-437 (memset y; CSR part; every dense block call).
 * y = A x.  x has ncols floats, y nrows floats (device pointers unless
 * SABLE_HOST_PTRS). */
int sable_spmv(SablePlan *plan, const float *x, float *y, int flags);

/* Replaces: the generated SpMM program body, src/codegen.py:895-926 with N
 * as a parameter.  Y = A X, X [ncols x N] and Y [nrows x N] row-major. */
int sable_spmm(SablePlan *plan, const float *X, float *Y, int flags);

/* No counterpart in the reference: the variant of sable_spmm for callers that
 * cannot allocate (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Y = A X with device X/Y (host pointers
 * are refused), bit-identical to sable_spmm on the same plan; it calls no
 * allocator: the partial sums of the plan's k-split MFMA tiles go to the
 * library's file-scope device scratch array (64 MiB, held by every GPU the
 * library runs on until the process ends) instead of the plan's buffer.  A
 * plan needing more scratch (SablePlanStats.scratch_bytes) is refused with
 * SABLE_ERR_UNSUPPORTED.  Calls that overlap on one GPU share the array and
 * race. */
int sable_spmm_noalloc(SablePlan *plan, const float *X, float *Y, int flags);

/* Replaces: the timed loop of src/codegen.py:414-439 and eval.py:23-45.
 * Runs `warmup` untimed launches then `reps` launches, each timed by HIP
 * events attached to the kernel launch itself (its start and end, without
 * the dispatch latency); times_ms[r] receives each launch's duration.  When flush_bytes > 0 a buffer of that size is overwritten
 * before every launch (outside the events) so each launch starts with cold
 * caches.  x/y (or X/Y) are device pointers. */
int sable_time(SablePlan *plan, const float *x, float *y, int warmup, int reps,
               int64_t flush_bytes, float *times_ms);

/* Batching (no single reference counterpart: replaces running many generated
 * programs, bench_launcher.py:85 / eval.py:30, one process each).  The
 * plans' tiles (each plan with its own x/y) run in launches over one
 * LPT-ordered work list per class of members that share every option of the
 * kernel specialisation (op, n_rhs, wg, mfma_split, unroll, use_rtc) -- one
 * launch per phase when all members share them, e.g. SpMV then SpMM for
 * mixed groups -- captured once into a hipGraph and replayed by
 * sable_group_launch.  The group copies only the tile and piece descriptors:
 * its kernels read each member's matrix from that member's own device
 * arrays, so every member plan must stay alive (not destroyed) until the
 * group is destroyed.  mode as in SablePlanOpts: SABLE_MODE_GRID (one
 * workgroup per tile) measured fastest; STATIC and QUEUE launch persistent
 * workgroups.  Members built for a group need no min_tiles of their own
 * (the group fills the GPU). */
int sable_group_create(SablePlan *const *plans, int nplans, const float *const *x,
                       float *const *y, int mode, int grid, SableGroup **out);
int sable_group_launch(SableGroup *g, int sync);
int sable_group_time(SableGroup *g, int warmup, int reps, int64_t flush_bytes,
                     float *times_ms);
int sable_group_destroy(SableGroup *g);
/* The unbatched counterpart of a group (no reference counterpart; the
 * baseline a group is measured against): every plan's own kernel launched
 * back to back on the first plan's stream, each with its own x/y; times_ms[r]
 * is the whole sequence, HIP events around it, a cache flush of flush_bytes
 * before it. */
int sable_time_plans(SablePlan *const *plans, int nplans, const float *const *x, float *const *y,
                     int warmup, int reps, int64_t flush_bytes, float *times_ms);

/* Router network (2-layer MLP, FP32): forward, backward and Adam kernels.
 * Weights are device pointers laid out [W1 (H x F), b1 (H), W2 (K x H), b2 (K)]. */
int sable_mlp_forward(const float *params, int F, int H, int K, const float *X,
                      int n, float *out, float *hidden);
int sable_mlp_backward(const float *params, int F, int H, int K, const float *X,
                       int n, const float *hidden, const float *dout, float *grad);
/* dout = scale * 2 * (out - target) * mask, elementwise over n values */
int sable_mlp_mse_grad(const float *out, const float *target, const float *mask, int64_t n,
                       float scale, float *dout);
int sable_adam_step(float *params, const float *grad, float *m, float *v, int64_t n,
                    float lr, float beta1, float beta2, float eps, int step);

/* Build tooling (no reference counterpart): compile one kernel
 * specialisation into the code-object cache without a device.
 * kind: 0 SpMV, 1 SpMM FP32, 2 SpMM MFMA (split 3), 3 SpMM MFMA (split 2). */
int sable_rtc_precompile(int kind, int wg, int unroll, int xcap, int nrhs, int rtmax,
                         const char *arch);

/* Achievable HBM rates (no reference counterpart): best of `reps` launches
 * of a streaming read and a streaming copy over `bytes` (GB/s; the copy
 * counts read + write bytes), each launch cold (after a 1 GiB overwrite). */
int sable_bandwidth_probe(int64_t bytes, int reps, float *read_gbs, float *copy_gbs);

/* Block partitioner, one axis (src/autopartition.py:34-68 cut_indices2 with
 * similarity2 :73-74, similarity2_numba :76-94 or similarity1 :70-71; the
 * reference densifies every line on the host).  The lines are the columns of
 * a CSC matrix given in host memory: indptr (nlines + 1 entries, indptr[0] =
 * 0, nondecreasing), indices ascending and unique within a line and below
 * line_len, FP64 data.  cut (nlines + 1 bytes) receives 1 where the
 * reference's scan cuts (always at 0 and nlines); sim (nlines - 1 doubles,
 * may be null) the adjacent-pair similarities, each dot product summed in
 * ascending index order without contraction; kernel_ms (may be null) the
 * time of the two kernels.  kind: 0 similarity2, 1 similarity2_numba,
 * 2 similarity1.  Rows are the same call on the transpose. */
int sable_axis_cuts(const int64_t *indptr, const int32_t *indices, const double *data, int64_t nlines,
                    int64_t line_len, int kind, double threshold, uint8_t *cut, double *sim,
                    float *kernel_ms);

/* Block materialisation on the device (utils/convert_real_to_vbr.py:98-199,
 * convert_sparse_to_vbr: the reference slices every block on the host).
 * Input: a CSC matrix in host memory (indptr of ncols + 1, row indices below
 * nrows, duplicates summed, FP64 data) and the partitions rpntr (nbr + 1
 * entries, 0 .. nrows, strictly increasing) and cpntr (nbc + 1).  A block is
 * kept when it holds a stored entry; kept blocks are numbered block-row
 * major and stored densely column-major, as the reference does.  create runs
 * the device pipeline and returns the sizes (nblocks kept blocks, nval
 * values) and the device time after the upload; fetch copies the VBR arrays
 * out (bpntrb / bpntre: -1 for a block row without blocks; block_nnz: the
 * nonzero values per block, the routing count of
 * utils/convert_real_to_vbr.py:41-47); any output may be null. */
int sable_inspect_create(const int64_t *indptr, const int32_t *indices, const double *data, int64_t nrows,
                         int64_t ncols, const int64_t *rpntr, int32_t nbr, const int64_t *cpntr, int32_t nbc,
                         SableInspect **out, int64_t *nblocks, int64_t *nval, float *device_ms);
int sable_inspect_fetch(SableInspect *h, double *val, int64_t *indx, int64_t *bindx, int64_t *bpntrb,
                        int64_t *bpntre, int32_t *block_nnz);
int sable_inspect_destroy(SableInspect *h);

/* Compressed (.vbrc) arrays on the device for given per-block routes
 * (utils/convert_real_to_vbr.py:18-96, convert_vbr_to_compressed, after its
 * routing step): dense-routed blocks are gathered into val2 / indx2, the
 * ordinals of sparse-routed blocks are ublocks, and their nonzero values form
 * a CSR matrix with sorted columns and the reference's padding entry (row
 * nrows - 1, column ncols - 1, value 0, when the last row or column holds no
 * entry).  Inputs are the VBR arrays in host memory (FP64 val) and one
 * route byte per block (SABLE_ROUTE_*).  create returns the sizes
 * (csr_nnz = -1 when there is no CSR part: no sparse-routed block or no
 * nonzero in them, as the reference writes none) and the device time after
 * the upload; fetch copies the arrays out (indptr nrows + 1, indices and
 * csr_val csr_nnz entries); any output may be null. */
int sable_compress_create(const double *val, int64_t nval, const int64_t *indx, const int64_t *bindx,
                          int64_t nblocks, const int64_t *rpntr, int32_t nbr, const int64_t *cpntr, int32_t nbc,
                          const int64_t *bpntrb, const int64_t *bpntre, const uint8_t *routes,
                          SableCompress **out, int64_t *nval2, int64_t *nublocks, int64_t *csr_nnz,
                          float *device_ms);
int sable_compress_fetch(SableCompress *h, double *val2, int64_t *indx2, int64_t *ublocks, int64_t *indptr,
                         int64_t *indices, double *csr_val);
int sable_compress_destroy(SableCompress *h);
/* Compression straight from an inspector handle: convert_sparse_to_vbr then
 * convert_vbr_to_compressed (utils/convert_real_to_vbr.py:98-199, :18-96)
 * without the VBR arrays leaving the device.  routes has one entry per block
 * of h (as sable_compress_create's); the result is read with
 * sable_compress_fetch and released with sable_compress_destroy. */
int sable_inspect_compress(SableInspect *h, const uint8_t *routes, SableCompress **out, int64_t *nval2,
                           int64_t *nublocks, int64_t *csr_nnz, float *device_ms);

/* Text `.vbr` / `.vbrc` lines `name=[a,b,...]` (host only, no device).
 * sable_text_open + sable_text_read_* replace the readers of
 * utils/fileio.py:55-64 (read_vbr) and :66-93 (read_vbrc): the file is
 * memory-mapped, counts[i] receives the element count of line i (0 for
 * "[]"), and each line is parsed by nthreads threads (0: the CPUs this
 * process may run on) with correctly rounded decimal conversion, so the
 * values equal Python's float() / int() of each token bit for bit.
 * sable_text_write replaces the writers of utils/fileio.py:13-35
 * (write_vbr_matrix) and utils/convert_real_to_vbr.py:84-95 (.vbrc): one line
 * per array, doubles formatted as Python's str() formats them, so a file
 * round-trips text-identically.  Errors of these four functions are read
 * with sable_text_last_error. */
#define SABLE_TEXT_F64 0
#define SABLE_TEXT_I64 1
int sable_text_open(const char *path, int nlines, int nthreads, SableText **out, int64_t *counts);
int sable_text_read_f64(SableText *h, int line, double *out);
int sable_text_read_i64(SableText *h, int line, int64_t *out);
void sable_text_close(SableText *h);
int sable_text_write(const char *path, int nlines, const char *const *names, const int32_t *kinds,
                     const void *const *arrays, const int64_t *counts, int append, int nthreads);
const char *sable_text_last_error(void);

/* Device memory helpers for callers without another allocator. */
int sable_malloc(void **ptr, int64_t bytes);
int sable_free(void *ptr);
int sable_memcpy_h2d(void *dst, const void *src, int64_t bytes);
int sable_memcpy_d2h(void *dst, const void *src, int64_t bytes);
int sable_synchronize(void);

#ifdef __cplusplus
}
#endif
#endif /* SABLE_H */
