// Hand-written gfx950 kernels of the VBR SpMV / SpMM hot path.
//
// This header is compiled twice: ahead of time by hipcc into libsable_amd.so
// (generic instances, the fallback), and at run time by hipRTC with the
// SB_* macros set per plan (the per-matrix specialisations).  It therefore
// includes nothing under hipRTC and uses only compiler builtins.
//
// Work decomposition (see the plan builder): a tile is a row segment of one
// block row; it owns those output rows and writes each exactly once, so no
// atomics touch y/Y and results are deterministic.  A tile's pieces are the
// column groups of its block row (dense FP32, dense FP32 in MFMA order, or
// sparse CSR), each with an x slice of at most SB_XCAP columns staged in LDS.
#pragma once

#include "sable_device.h"

#ifndef SB_WG
#define SB_WG 256
#endif
#ifndef SB_XCAP
#define SB_XCAP 4096
#endif
#ifndef SB_UNROLL
#define SB_UNROLL 4
#endif
#ifndef SB_NRHS
#define SB_NRHS 0          // 0 = runtime RHS width (generic instance)
#endif

typedef float sb_f4 __attribute__((ext_vector_type(4)));
typedef unsigned int sb_u4 __attribute__((ext_vector_type(4)));

#define SB_MODE_GRID 0      // one workgroup per tile
#define SB_MODE_STATIC 1    // persistent, tile = wg + k * grid
#define SB_MODE_QUEUE 2     // persistent, atomic work queue (self-resetting)

struct SbSpmvArgs {
    const SbTile *tiles;
    const SbPiece *pieces;
    const SbXRange *xr;
    const float *dval;
    const float *sval;
    const uint16_t *sidx;
    const int32_t *srowptr;
    const float *const *xvec;     // per matrix slot (batched plans) or null
    float *const *yvec;           // per matrix slot (batched plans) or null
    const float *x0;              // single plan: x
    float *y0;                    // single plan: y
    int32_t *queue;               // [0] next tile, [1] retired workgroups
    uint64_t *tile_clock;         // optional [2*ntiles]: start/end wall clock
    int32_t ntiles;
    int32_t mode;
};

struct SbSpmmArgs {
    const SbTile *tiles;
    const SbPiece *pieces;
    const SbXRange *xr;
    const float *dval;            // FP32 dense pieces (column-major) and MFMA pieces (register order)
    const float *sval;
    const uint16_t *sidx;
    const int32_t *srowptr;
    const float *const *xmat;     // per matrix slot: X [ncols x n_rhs] row-major, or null
    float *const *ymat;           // per matrix slot: Y [nrows x n_rhs] row-major, or null
    const float *x0;              // single plan: X
    float *y0;                    // single plan: Y
    int32_t *queue;
    uint64_t *tile_clock;
    int32_t ntiles;
    int32_t mode;
    int32_t n_rhs;
};

// ---------------------------------------------------------------------------
// Work distribution helpers

__device__ __forceinline__ void sb_clock(uint64_t *clk, int t, int which)
{
    if (clk && threadIdx.x == 0) clk[2 * t + which] = wall_clock64();
}

// Runs body(t) over the tiles according to mode.  The queue counters are
// reset by the last workgroup to retire, so the kernel can be relaunched or
// replayed from a graph without a memset.
template <typename Body>
__device__ __forceinline__ void sb_for_tiles(int ntiles, int mode, int32_t *queue,
                                             int *next_tile, Body body)
{
    if (mode == SB_MODE_GRID) {
        if ((int)blockIdx.x < ntiles) body((int)blockIdx.x);
        return;
    }
    if (mode == SB_MODE_STATIC) {
        for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
            body(t);
            __syncthreads();
        }
        return;
    }
    for (;;) {
        if (threadIdx.x == 0) *next_tile = atomicAdd(&queue[0], 1);
        __syncthreads();
        const int t = *next_tile;
        __syncthreads();
        if (t >= ntiles) break;
        body(t);
    }
    if (threadIdx.x == 0) {
        __threadfence();
        const int done = atomicAdd(&queue[1], 1);
        if (done == (int)gridDim.x - 1) {
            atomicExch(&queue[0], 0);
            atomicExch(&queue[1], 0);
        }
    }
}

// Copies the x slice of a piece into xs[0..ncols) and zeroes xs[ncols..ncols+3]
// (the flat dense stream of a 1- or 2-row tile can touch up to xs[ncols+2],
// always multiplied by the zero padding of the panel).
__device__ __forceinline__ void sb_stage_x(const SbXRange *__restrict__ xr, int r0, int r1,
                                           const float *__restrict__ x, float *xs, int ncols)
{
    int base = 0;
    for (int r = r0; r < r1; ++r) {
        const SbXRange rg = xr[r];
        for (int i = threadIdx.x; i < rg.len; i += blockDim.x) xs[base + i] = x[rg.col0 + i];
        base += rg.len;
    }
    if (threadIdx.x < 4) xs[ncols + threadIdx.x] = 0.f;
}

// ---------------------------------------------------------------------------
// SpMV
//
// Dense piece = column-major hs x ncols FP32 sub-panel, contiguous, 16-byte
// aligned, zero-padded to a multiple of 4 floats.  It is streamed as float4:
// lane t (t < T') reads float4 q = t + k*T'.  Because 4*T' = M*hs, lane t
// always sees the same four row indices (4t+s) % hs and walks columns
// (4t+s)/hs + k*M, so each lane accumulates four fixed rows in registers and
// the wavefront's loads are fully coalesced 16-byte accesses.  The M partial
// sums of every row are combined by a deterministic LDS tree at the end.
//
// Sparse piece = CSR over the tile's rows with uint16 column indices into the
// staged x slice; 2^k lanes per row stride over the row's entries and reduce
// with cross-lane shuffles (k chosen per tile from the mean row length).

template <int WG, int UNROLL, int XCAP>
__device__ __forceinline__ void sb_spmv_tile(const SbSpmvArgs &a, int t, float *xs, float *red,
                                             float *ysp)
{
    const int tid = threadIdx.x;
    sb_clock(a.tile_clock, t, 0);
    const SbTile tile = a.tiles[t];
    const int hs = tile.nrows;
    const int tpar = tile.dense_tpar;
    const int M = tile.dense_mcols;
    const float *__restrict__ x = a.xvec ? a.xvec[tile.mat] : a.x0;
    float *__restrict__ y = a.yvec ? a.yvec[tile.mat] : a.y0;

    const int p0 = 4 * tid;
    const int c0 = p0 / hs, c1 = (p0 + 1) / hs, c2 = (p0 + 2) / hs, c3 = (p0 + 3) / hs;
    float acc0 = 0.f, acc1 = 0.f, acc2 = 0.f, acc3 = 0.f;
    if (tid < hs) ysp[tid] = 0.f;

    for (int pi = tile.piece_begin; pi < tile.piece_end; ++pi) {
        const SbPiece pc = a.pieces[pi];
        __syncthreads();
        sb_stage_x(a.xr, pc.xr_begin, pc.xr_end, x, xs, pc.ncols);
        __syncthreads();
        if (pc.kind == SB_PIECE_DENSE) {
            const int nq = (hs * pc.ncols + 3) >> 2;
            const sb_f4 *__restrict__ P = reinterpret_cast<const sb_f4 *>(a.dval + pc.data_off);
            if (tid < tpar) {
                int q = tid;
                int kc = 0;
                for (; q + (UNROLL - 1) * tpar < nq; q += UNROLL * tpar, kc += UNROLL * M) {
                    sb_f4 v[UNROLL];
#pragma unroll
                    for (int u = 0; u < UNROLL; ++u) v[u] = __builtin_nontemporal_load(&P[q + u * tpar]);
#pragma unroll
                    for (int u = 0; u < UNROLL; ++u) {
                        const int k = kc + u * M;
                        acc0 += v[u].x * xs[c0 + k];
                        acc1 += v[u].y * xs[c1 + k];
                        acc2 += v[u].z * xs[c2 + k];
                        acc3 += v[u].w * xs[c3 + k];
                    }
                }
                for (; q < nq; q += tpar, kc += M) {
                    const sb_f4 v = __builtin_nontemporal_load(&P[q]);
                    acc0 += v.x * xs[c0 + kc];
                    acc1 += v.y * xs[c1 + kc];
                    acc2 += v.z * xs[c2 + kc];
                    acc3 += v.w * xs[c3 + kc];
                }
            }
        } else {
            const int lg = tile.sp_lanes_log2;
            const int L = 1 << lg;
            const int g = tid >> lg;
            const int l = tid & (L - 1);
            const int per_pass = WG >> lg;
            const int32_t *__restrict__ rp = a.srowptr + pc.rowptr_off;
            const float *__restrict__ sv = a.sval + pc.data_off;
            const uint16_t *__restrict__ si = a.sidx + pc.data_off;
            for (int r = g; r < hs; r += per_pass) {
                const int b = rp[r], e = rp[r + 1];
                float s = 0.f;
                for (int p = b + l; p < e; p += L) s += sv[p] * xs[si[p]];
                for (int o = L >> 1; o > 0; o >>= 1) s += __shfl_xor(s, o);
                if (l == 0) ysp[r] += s;
            }
        }
    }

    __syncthreads();
    if (tid < tpar) {
        red[p0] = acc0;
        red[p0 + 1] = acc1;
        red[p0 + 2] = acc2;
        red[p0 + 3] = acc3;
    }
    __syncthreads();
    // red holds M groups of hs partial sums: red[m*hs + r].  Fold halves.
    int m = M;
    while (m > 1) {
        const int half = (m + 1) >> 1;
        const int cnt = (m - half) * hs;
        for (int i = tid; i < cnt; i += WG) red[i] += red[i + half * hs];
        __syncthreads();
        m = half;
    }
    if (tid < hs) y[tile.row0 + tid] = red[tid] + ysp[tid];
    if (a.tile_clock) {
        __syncthreads();
        sb_clock(a.tile_clock, t, 1);
    }
}

template <int WG, int UNROLL, int XCAP>
__device__ __forceinline__ void sb_spmv_body(const SbSpmvArgs &a)
{
    __shared__ float xs[XCAP + 4];
    __shared__ float red[4 * WG];
    __shared__ float ysp[WG];
    __shared__ int next_tile;
    sb_for_tiles(a.ntiles, a.mode, a.queue, &next_tile,
                 [&](int t) { sb_spmv_tile<WG, UNROLL, XCAP>(a, t, xs, red, ysp); });
}

// ---------------------------------------------------------------------------
// SpMM, FP32 on the vector ALUs (the plain path)
//
// Thread (rg, v): row group rg = tid / NV owns rows rg*rt .. rg*rt+rt-1 of the
// tile and RHS columns 4v..4v+3 (NV = n_rhs/4).  Dense piece values are read
// once per (row, column) and broadcast to the NV lanes of a row group; X rows
// come through L2 (each row is shared by the row groups of the tile).  Sparse
// rows are walked entry by entry with the same thread mapping.

// Expands a piece's x ranges into xrow[0..ncols): the X row of every column.
__device__ __forceinline__ void sb_expand_xrow(const SbXRange *__restrict__ xr, int r0, int r1, int *xrow)
{
    int base = 0;
    for (int r = r0; r < r1; ++r) {
        const SbXRange g = xr[r];
        for (int i = threadIdx.x; i < g.len; i += blockDim.x) xrow[base + i] = g.col0 + i;
        base += g.len;
    }
}

struct SbVecMap {
    int rt;       // rows per row group
    int rbeg;     // first tile row of this thread's group
    int v;        // float4 column of this thread
    bool active;
};

__device__ __forceinline__ SbVecMap sb_vec_map(int hs, int N, int wg)
{
    SbVecMap m;
    const int NV = N >> 2;
    const int ngroups = wg / NV;
    m.rt = (hs + ngroups - 1) / ngroups;
    const int rg = threadIdx.x / NV;
    m.v = threadIdx.x - rg * NV;
    m.active = rg < ngroups;
    m.rbeg = rg * m.rt;
    return m;
}

// Accumulates one FP32-dense or sparse piece into acc (RTMAX row registers).
template <int RTMAX>
__device__ __forceinline__ void sb_spmm_vec_piece(const SbSpmmArgs &a, const SbPiece &pc, int hs, int N,
                                                  const SbVecMap &m, const int *xrow,
                                                  const float *__restrict__ X, float4 *acc)
{
    if (pc.kind == SB_PIECE_DENSE) {
        const float *__restrict__ P = a.dval + pc.data_off;
        for (int c = 0; c < pc.ncols; ++c) {
            const float4 xv = *reinterpret_cast<const float4 *>(X + (size_t)xrow[c] * N + 4 * m.v);
            const float *__restrict__ col = P + (size_t)c * hs;
#pragma unroll
            for (int i = 0; i < RTMAX; ++i) {
                const int r = m.rbeg + i;
                if (i < m.rt && r < hs) {
                    const float av = col[r];
                    acc[i].x += av * xv.x;
                    acc[i].y += av * xv.y;
                    acc[i].z += av * xv.z;
                    acc[i].w += av * xv.w;
                }
            }
        }
    } else if (pc.kind == SB_PIECE_SPARSE) {
        const int32_t *__restrict__ rp = a.srowptr + pc.rowptr_off;
        const float *__restrict__ sv = a.sval + pc.data_off;
        const uint16_t *__restrict__ si = a.sidx + pc.data_off;
#pragma unroll
        for (int i = 0; i < RTMAX; ++i) {
            const int r = m.rbeg + i;
            if (i < m.rt && r < hs) {
                for (int p = rp[r]; p < rp[r + 1]; ++p) {
                    const float av = sv[p];
                    const float4 xv = *reinterpret_cast<const float4 *>(X + (size_t)xrow[si[p]] * N + 4 * m.v);
                    acc[i].x += av * xv.x;
                    acc[i].y += av * xv.y;
                    acc[i].z += av * xv.z;
                    acc[i].w += av * xv.w;
                }
            }
        }
    }
}

template <int WG, int RTMAX, int NRHS>
__device__ __forceinline__ void sb_spmm_f32_tile(const SbSpmmArgs &a, int t, int *xrow)
{
    sb_clock(a.tile_clock, t, 0);
    const int N = NRHS ? NRHS : a.n_rhs;
    const SbTile tile = a.tiles[t];
    const int hs = tile.nrows;
    const SbVecMap m = sb_vec_map(hs, N, WG);
    const float *__restrict__ X = a.xmat ? a.xmat[tile.mat] : a.x0;
    float *__restrict__ Y = a.ymat ? a.ymat[tile.mat] : a.y0;
    float4 acc[RTMAX];
#pragma unroll
    for (int i = 0; i < RTMAX; ++i) acc[i] = make_float4(0.f, 0.f, 0.f, 0.f);

    for (int pi = tile.piece_begin; pi < tile.piece_end; ++pi) {
        const SbPiece pc = a.pieces[pi];
        __syncthreads();
        sb_expand_xrow(a.xr, pc.xr_begin, pc.xr_end, xrow);
        __syncthreads();
        if (m.active) sb_spmm_vec_piece<RTMAX>(a, pc, hs, N, m, xrow, X, acc);
    }
    if (m.active) {
#pragma unroll
        for (int i = 0; i < RTMAX; ++i) {
            const int r = m.rbeg + i;
            if (i < m.rt && r < hs)
                *reinterpret_cast<float4 *>(Y + (size_t)(tile.row0 + r) * N + 4 * m.v) = acc[i];
        }
    }
    if (a.tile_clock) {
        __syncthreads();
        sb_clock(a.tile_clock, t, 1);
    }
}

template <int WG, int RTMAX, int NRHS, int XCAP>
__device__ __forceinline__ void sb_spmm_f32_body(const SbSpmmArgs &a)
{
    __shared__ int xrow[XCAP + 4];
    __shared__ int next_tile;
    sb_for_tiles(a.ntiles, a.mode, a.queue, &next_tile,
                 [&](int t) { sb_spmm_f32_tile<WG, RTMAX, NRHS>(a, t, xrow); });
}
