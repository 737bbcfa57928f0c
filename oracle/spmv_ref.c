/*
 * C restatement of the reference's generated-C executor — TEST / BASELINE ONLY.
 *
 * The reference emits one C program per matrix (src/codegen.py:293-452 for
 * SpMV, :827-937 for SpMM).  Its timed body is a straight-line sequence of
 * kernel calls whose block coordinates are compile-time constants.  This file
 * restates that body over a call list built once per matrix (the same calls,
 * with their coordinates as data), so it can be timed on the host cores as
 * the CPU baseline and used as the FP64 checker of the GPU kernels.
 *
 *  - oracle_build_calls   block enumeration of src/codegen.py:421-427
 *                         (rows with bpntrb == -1 skipped, block columns in
 *                         ascending order found by membership in bindx)
 *  - spmv kernels         shapes of src/codegen.py:263-291 (general
 *                         column-outer/row-inner, height-1 row dot, width-1
 *                         column axpy)
 *  - CSR part             y = A_csr x with beta = 0 before the dense blocks
 *                         (src/codegen.py:419-420); row sums as in
 *                         src/csr-spmv.cpp:250-262 (MKL is not available here)
 *  - SpMM                 src/codegen.py:851-859 loop nest and per-nonzero
 *                         statements for sparse-routed blocks (:919-922) with
 *                         the RHS width n_rhs as a parameter instead of 512
 *  - multi-threaded SpMV  block rows split over threads by greedy LPT on r*c
 *                         work (src/codegen.py:165-186); each thread owns its
 *                         rows, so no two threads write the same y element.
 *                         The value-offset defect of the reference's threaded
 *                         emitter (src/codegen.py:615-624) is not reproduced.
 *
 * Compiled by oracle/Makefile.  Never linked into the product library.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct {
    int32_t kind;     /* 0 general, 1 height-1, 2 width-1 */
    int32_t i0, i1, j0, j1;
    int32_t brow;     /* block row of the call */
    int64_t off;      /* value offset (indx2 of the dense-only numbering) */
} ocall_t;

/* Enumerate dense-routed blocks in ordinal order.  sparse_ord[k] != 0 marks
 * ordinal k as sparse-routed (the reference's ublocks).  dense_off[c] is the
 * value offset of the c-th dense block (indx2).  Returns the number of calls
 * written, or -1 if more than max_calls would be needed. */
int64_t oracle_build_calls(int32_t nbr, int32_t nbc, const int32_t *rpntr,
                           const int32_t *cpntr, const int32_t *bindx,
                           const int32_t *bpntrb, const int32_t *bpntre,
                           const uint8_t *sparse_ord, const int64_t *dense_off,
                           ocall_t *out, int64_t max_calls)
{
    int64_t ordinal = 0, count = 0;
    uint8_t *mark = (uint8_t *)calloc((size_t)nbc + 1, 1);
    if (!mark) return -1;
    for (int32_t a = 0; a < nbr; ++a) {
        if (bpntrb[a] == -1) continue;
        for (int32_t p = bpntrb[a]; p < bpntre[a]; ++p) mark[bindx[p]] = 1;
        for (int32_t b = 0; b < nbc; ++b) {
            if (!mark[b]) continue;
            if (!sparse_ord[ordinal]) {
                if (count >= max_calls) { free(mark); return -1; }
                ocall_t *c = &out[count];
                c->i0 = rpntr[a]; c->i1 = rpntr[a + 1];
                c->j0 = cpntr[b]; c->j1 = cpntr[b + 1];
                c->brow = a;
                c->off = dense_off[count];
                c->kind = (c->i1 - c->i0 == 1) ? 1 : ((c->j1 - c->j0 == 1) ? 2 : 0);
                ++count;
            }
            ++ordinal;
        }
        for (int32_t p = bpntrb[a]; p < bpntre[a]; ++p) mark[bindx[p]] = 0;
    }
    free(mark);
    return count;
}

#define DEFINE_SPMV(T, SUF)                                                     \
static inline void kern_general_##SUF(T *restrict y, const T *restrict x,      \
        const T *restrict v, int i0, int i1, int j0, int j1)                    \
{                                                                               \
    const int h = i1 - i0;                                                      \
    for (int j = j0; j < j1; ++j)                                               \
        for (int i = i0; i < i1; ++i)                                           \
            y[i] += v[(j - j0) * h + (i - i0)] * x[j];                          \
}                                                                               \
static inline void kern_row_##SUF(T *restrict y, const T *restrict x,          \
        const T *restrict v, int i0, int j0, int j1)                            \
{                                                                               \
    for (int j = j0; j < j1; ++j) y[i0] += v[j - j0] * x[j];                    \
}                                                                               \
static inline void kern_col_##SUF(T *restrict y, const T *restrict x,          \
        const T *restrict v, int i0, int i1, int j0)                            \
{                                                                               \
    const T xj = x[j0];                                                         \
    for (int i = i0; i < i1; ++i) y[i] += v[i - i0] * xj;                       \
}                                                                               \
static void run_calls_##SUF(const ocall_t *calls, int64_t c0, int64_t c1,      \
        const T *val2, const T *x, T *y)                                        \
{                                                                               \
    for (int64_t c = c0; c < c1; ++c) {                                         \
        const ocall_t *k = &calls[c];                                           \
        const T *v = val2 + k->off;                                             \
        if (k->kind == 1) kern_row_##SUF(y, x, v, k->i0, k->j0, k->j1);         \
        else if (k->kind == 2) kern_col_##SUF(y, x, v, k->i0, k->i1, k->j0);    \
        else kern_general_##SUF(y, x, v, k->i0, k->i1, k->j0, k->j1);           \
    }                                                                           \
}                                                                               \
static void csr_rows_##SUF(int r0, int r1, const int32_t *indptr,              \
        const int32_t *indices, const T *csr_val, const T *x, T *y)             \
{                                                                               \
    for (int i = r0; i < r1; ++i) {                                             \
        T s = 0;                                                                \
        for (int32_t p = indptr[i]; p < indptr[i + 1]; ++p)                     \
            s += csr_val[p] * x[indices[p]];                                    \
        y[i] = s;                                                               \
    }                                                                           \
}                                                                               \
void oracle_spmv_##SUF(const ocall_t *calls, int64_t ncalls, const T *val2,    \
        int32_t nrows, int32_t has_csr, const int32_t *indptr,                  \
        const int32_t *indices, const T *csr_val, const T *x, T *y)             \
{                                                                               \
    memset(y, 0, sizeof(T) * (size_t)nrows);                                    \
    if (has_csr) csr_rows_##SUF(0, nrows, indptr, indices, csr_val, x, y);      \
    run_calls_##SUF(calls, 0, ncalls, val2, x, y);                              \
}                                                                               \
/* Threaded: chunk t owns block rows listed in rows_of[chunk_ptr[t]..]; the   \
 * calls of block row a are calls[call_ptr[a]..call_ptr[a+1]). */              \
void oracle_spmv_mt_##SUF(const ocall_t *calls, const int64_t *call_ptr,       \
        const int32_t *rpntr, const int32_t *chunk_ptr, const int32_t *rows_of, \
        int32_t nchunks, const T *val2, int32_t nrows, int32_t has_csr,         \
        const int32_t *indptr, const int32_t *indices, const T *csr_val,        \
        const T *x, T *y)                                                       \
{                                                                               \
    (void)nrows;                                                                \
    _Pragma("omp parallel for schedule(static, 1) num_threads(nchunks)")        \
    for (int t = 0; t < nchunks; ++t) {                                         \
        for (int32_t q = chunk_ptr[t]; q < chunk_ptr[t + 1]; ++q) {             \
            const int32_t a = rows_of[q];                                       \
            const int r0 = rpntr[a], r1 = rpntr[a + 1];                         \
            if (has_csr) csr_rows_##SUF(r0, r1, indptr, indices, csr_val, x, y);\
            else memset(y + r0, 0, sizeof(T) * (size_t)(r1 - r0));              \
            run_calls_##SUF(calls, call_ptr[a], call_ptr[a + 1], val2, x, y);   \
        }                                                                       \
    }                                                                           \
}

DEFINE_SPMV(double, f64)
DEFINE_SPMV(float, f32)

static int64_t now_ns(void)
{
    struct timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return (int64_t)t.tv_sec * 1000000000LL + t.tv_nsec;
}

/* Timed loop of src/codegen.py:414-439: bench+1 iterations, the first one
 * dropped; each iteration zeroes y, then runs the CSR part and every dense
 * call.  nchunks <= 1 runs the single-threaded body. */
#define DEFINE_TIMER(T, SUF)                                                    \
int oracle_time_spmv_##SUF(const ocall_t *calls, int64_t ncalls,               \
        const int64_t *call_ptr, const int32_t *rpntr, const int32_t *chunk_ptr,\
        const int32_t *rows_of, int32_t nchunks, const T *val2, int32_t nrows,  \
        int32_t has_csr, const int32_t *indptr, const int32_t *indices,         \
        const T *csr_val, const T *x, T *y, int32_t bench, int64_t *times_ns)   \
{                                                                               \
    for (int it = 0; it < bench + 1; ++it) {                                    \
        memset(y, 0, sizeof(T) * (size_t)nrows);                                \
        int64_t t1 = now_ns();                                                  \
        if (nchunks <= 1)                                                       \
            oracle_spmv_##SUF(calls, ncalls, val2, nrows, has_csr, indptr,      \
                              indices, csr_val, x, y);                          \
        else                                                                    \
            oracle_spmv_mt_##SUF(calls, call_ptr, rpntr, chunk_ptr, rows_of,    \
                                 nchunks, val2, nrows, has_csr, indptr,         \
                                 indices, csr_val, x, y);                       \
        int64_t t2 = now_ns();                                                  \
        if (it) times_ns[it - 1] = t2 - t1;                                     \
    }                                                                           \
    return 0;                                                                   \
}

DEFINE_TIMER(double, f64)
DEFINE_TIMER(float, f32)

/* SpMM, src/codegen.py:827-937 with n_rhs columns.  route[k] != 0 marks
 * ordinal k as sparse (per-nonzero statements).  val/indx are the full VBR
 * arrays (the reference SpMM reads the .vbr, not the .vbrc).  X is
 * [ncols x n_rhs] row-major, Y [nrows x n_rhs] row-major and is zeroed
 * first.  Blocks are processed in ordinal order. */
int oracle_spmm_f32(int32_t nbr, int32_t nbc, const int32_t *rpntr,
                    const int32_t *cpntr, const int32_t *bindx,
                    const int32_t *bpntrb, const int32_t *bpntre,
                    const int64_t *indx, const float *val, const uint8_t *route,
                    const float *X, int32_t n_rhs, float *Y)
{
    const int N = n_rhs;
    int64_t ordinal = 0;
    uint8_t *mark = (uint8_t *)calloc((size_t)nbc + 1, 1);
    if (!mark) return -1;
    memset(Y, 0, sizeof(float) * (size_t)rpntr[nbr] * (size_t)N);
    for (int32_t a = 0; a < nbr; ++a) {
        if (bpntrb[a] == -1) continue;
        for (int32_t p = bpntrb[a]; p < bpntre[a]; ++p) mark[bindx[p]] = 1;
        for (int32_t b = 0; b < nbc; ++b) {
            if (!mark[b]) continue;
            const int i0 = rpntr[a], i1 = rpntr[a + 1];
            const int j0 = cpntr[b], j1 = cpntr[b + 1];
            const int h = i1 - i0;
            const float *v = val + indx[ordinal];
            const int sparse = route[ordinal] != 0;
            for (int i = i0; i < i1; ++i) {
                float *yr = Y + (size_t)i * N;
                for (int j = j0; j < j1; ++j) {
                    const float a_ij = v[(j - j0) * h + (i - i0)];
                    if (sparse && a_ij == 0.0f) continue;
                    const float *xr = X + (size_t)j * N;
                    for (int k = 0; k < N; ++k) yr[k] += a_ij * xr[k];
                }
            }
            ++ordinal;
        }
        for (int32_t p = bpntrb[a]; p < bpntre[a]; ++p) mark[bindx[p]] = 0;
    }
    free(mark);
    return 0;
}

/* Threaded SpMM (src/codegen.py:939-1075 splits block rows over pthreads by
 * split_chunks): chunk t owns the block rows rows_of[chunk_ptr[t] ..
 * chunk_ptr[t+1]), every block row of the matrix in exactly one chunk, so
 * each Y row is written by one thread in the single-threaded order and the
 * result is bit-identical to oracle_spmm_f32.  Block p of row a is ordinal
 * p (bindx is ascending within a row). */
int oracle_spmm_mt_f32(int32_t nbr, int32_t nbc, const int32_t *rpntr,
                       const int32_t *cpntr, const int32_t *bindx,
                       const int32_t *bpntrb, const int32_t *bpntre,
                       const int64_t *indx, const float *val, const uint8_t *route,
                       const float *X, int32_t n_rhs, float *Y,
                       const int32_t *chunk_ptr, const int32_t *rows_of, int32_t nchunks)
{
    const int N = n_rhs;
    (void)nbc;
    _Pragma("omp parallel for schedule(static, 1) num_threads(nchunks)")
    for (int t = 0; t < nchunks; ++t) {
        for (int32_t q = chunk_ptr[t]; q < chunk_ptr[t + 1]; ++q) {
            const int32_t a = rows_of[q];
            const int i0 = rpntr[a], i1 = rpntr[a + 1];
            const int h = i1 - i0;
            memset(Y + (size_t)i0 * N, 0, sizeof(float) * (size_t)h * (size_t)N);
            if (bpntrb[a] == -1) continue;
            for (int32_t p = bpntrb[a]; p < bpntre[a]; ++p) {
                const int j0 = cpntr[bindx[p]], j1 = cpntr[bindx[p] + 1];
                const float *v = val + indx[p];
                const int sparse = route[p] != 0;
                for (int i = i0; i < i1; ++i) {
                    float *yr = Y + (size_t)i * N;
                    for (int j = j0; j < j1; ++j) {
                        const float a_ij = v[(j - j0) * h + (i - i0)];
                        if (sparse && a_ij == 0.0f) continue;
                        const float *xr = X + (size_t)j * N;
                        for (int k = 0; k < N; ++k) yr[k] += a_ij * xr[k];
                    }
                }
            }
        }
    }
    (void)nbr;
    return 0;
}

int oracle_time_spmm_mt_f32(int32_t nbr, int32_t nbc, const int32_t *rpntr,
                            const int32_t *cpntr, const int32_t *bindx,
                            const int32_t *bpntrb, const int32_t *bpntre,
                            const int64_t *indx, const float *val, const uint8_t *route,
                            const float *X, int32_t n_rhs, float *Y,
                            const int32_t *chunk_ptr, const int32_t *rows_of, int32_t nchunks,
                            int32_t reps, int64_t *times_ns)
{
    for (int it = 0; it < reps; ++it) {
        int64_t t1 = now_ns();
        oracle_spmm_mt_f32(nbr, nbc, rpntr, cpntr, bindx, bpntrb, bpntre, indx, val, route, X, n_rhs, Y,
                           chunk_ptr, rows_of, nchunks);
        int64_t t2 = now_ns();
        times_ns[it] = t2 - t1;
    }
    return 0;
}

int oracle_time_spmm_f32(int32_t nbr, int32_t nbc, const int32_t *rpntr,
                         const int32_t *cpntr, const int32_t *bindx,
                         const int32_t *bpntrb, const int32_t *bpntre,
                         const int64_t *indx, const float *val,
                         const uint8_t *route, const float *X, int32_t n_rhs,
                         float *Y, int32_t reps, int64_t *times_ns)
{
    for (int it = 0; it < reps; ++it) {
        int64_t t1 = now_ns();
        int rc = oracle_spmm_f32(nbr, nbc, rpntr, cpntr, bindx, bpntrb, bpntre,
                                 indx, val, route, X, n_rhs, Y);
        int64_t t2 = now_ns();
        if (rc) return rc;
        times_ns[it] = t2 - t1;
    }
    return 0;
}

int oracle_max_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
