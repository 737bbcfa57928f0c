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
typedef unsigned int sb_u2 __attribute__((ext_vector_type(2)));

// Global-address-space views of the buffer pointers.  The pointers come
// from descriptor tables (SbMatBufs), so the compiler cannot tell they point
// to global memory and would emit FLAT loads and stores; those count against
// the LDS counter as well, so every wait for an LDS access would also wait
// for every global load in flight, defeating the software pipelining.
#define SB_G __attribute__((address_space(1)))
template <typename T>
__device__ __forceinline__ SB_G T *sb_gp(T *p)
{
    return (SB_G T *)p;
}
template <typename T>
__device__ __forceinline__ const SB_G T *sb_gp(const T *p)
{
    return (const SB_G T *)p;
}

#define SB_MODE_GRID 0      // one workgroup per tile
#define SB_MODE_STATIC 1    // persistent, tile = wg + k * grid
#define SB_MODE_QUEUE 2     // persistent, atomic work queue (self-resetting)

// The device arrays of one matrix of a group: each member keeps its own
// arrays (the group copies none of them) and a tile finds them through its
// matrix slot.  Piece and tile offsets are relative to these arrays.
struct SbMatBufs {
    const float *dval;
    const float *sval;
    const uint16_t *sidx;
    const int32_t *srowptr;
    const int32_t *xmap;          // SpMV only
    const float *x;               // x (SpMV) or X [ncols x n_rhs] (SpMM)
    float *y;                     // y (SpMV) or Y [nrows x n_rhs] (SpMM)
    float *part;                  // SpMM: partial sums of the split MFMA tiles
};

struct SbSpmvArgs {
    const SbTile *tiles;
    const SbPiece *pieces;
    const SbXRange *xr;
    const float *dval;
    const float *sval;
    const uint16_t *sidx;
    const int32_t *srowptr;
    const int32_t *xmap;          // x column per staged slot (-1: zero pad)
    const SbMatBufs *mats;        // per matrix slot (groups), or null: the fields above
    const float *x0;              // single plan: x
    float *y0;                    // single plan: y
    float *part;                  // unused (SpMM only)
    int32_t *queue;               // [0] next tile, [1] retired workgroups
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
    const SbMatBufs *mats;        // per matrix slot (groups), or null: the fields above
    const float *x0;              // single plan: X
    float *y0;                    // single plan: Y
    float *part;                  // single plan: partial sums of the split MFMA tiles
    int32_t *queue;
    int32_t ntiles;
    int32_t mode;
    int32_t n_rhs;
};

// The arrays a tile reads: its matrix slot's in a group, else the plan's.
template <typename Args>
__device__ __forceinline__ SbMatBufs sb_mat_bufs(const Args &a, int mat, const int32_t *xmap)
{
    if (a.mats) return a.mats[mat];
    SbMatBufs b;
    b.dval = a.dval;
    b.sval = a.sval;
    b.sidx = a.sidx;
    b.srowptr = a.srowptr;
    b.xmap = xmap;
    b.x = a.x0;
    b.y = a.y0;
    b.part = a.part;
    return b;
}

// ---------------------------------------------------------------------------
// Work distribution helpers

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

// ---------------------------------------------------------------------------
// SpMV
//
// A tile's pieces are processed in staging groups (formed by the host):
// consecutive pieces whose x slices (each followed by 4 zero pads) fit in the
// LDS buffer are staged in one pass through the group's slot map (x column
// per slot), then processed.  Sparse pieces come first in a group.
//
// Roles: lanes tid < T' stream the dense pieces.  In a tile with both kinds
// the last sp_waves wavefronts take the sparse pieces (their loads are
// latency-bound and run alongside the dense stream); otherwise every
// wavefront takes both in turn.
//
// Dense piece = column-major hs x ncols FP32 sub-panel, contiguous, 16-byte
// aligned, zero-padded to a multiple of 4 floats and followed by one zero
// float4, streamed as float4: lane t
// (t < T') reads float4 q = t + k*T'.  Because 4*T' = M*hs, lane t always
// sees the same four row indices (4t+s) % hs and walks columns (4t+s)/hs +
// k*M, so each lane accumulates four fixed rows in registers and every
// wavefront load is a fully coalesced 16-byte access.  Loads are software
// pipelined (two register batches of UNROLL float4).
//
// Sparse piece = wave slices (laid out by the host plan): the tile's rows,
// longest first, are cut into slices of C = 64/G rows, slice s going to
// sparse wavefront s % nsw, G lanes per row; at step j lane q*G+g holds
// entries 4(jG+g) .. 4(jG+g)+3 of its row (uint16 columns into the staged x
// slice).  A wavefront streams a slice's steps as one coalesced float4 + 4 x
// uint16 load per lane, accumulates in a register, folds the G lanes of a
// row with a fixed xor butterfly and adds the row sum to the row's LDS slot
// spy[r].  The same lane owns a row in every sparse piece of the tile, so
// spy needs no barrier between pieces.
//
// Row r's result is (dense partial sums of the M column groups, in group
// order) + (sparse sums, piece by piece): deterministic, no atomics.

// The waves per SIMD the SpMV kernel is compiled for (its VGPR budget) given
// the dense unroll: 80 VGPRs (6 waves) at unroll 2, 96 (5 waves) at unroll 4.
// SB_SPMV_MINWAVES > 0 sets it for every unroll.
#ifndef SB_SPMV_MINWAVES
#define SB_SPMV_MINWAVES 0
#endif
#define SB_SPMV_WAVES(U) (SB_SPMV_MINWAVES > 0 ? SB_SPMV_MINWAVES : (U) <= 2 ? 6 : (U) <= 4 ? 5 : 1)

struct SbStreamLane {
    int c0, c1, c2, c3;    // column offsets of this lane's four elements
};

// Loads are unconditional and nothing is applied to their values until the
// batch is consumed, so the compiler can wait for one batch while the next
// is still in flight: an index past the end reads the zero float4 the plan
// builder appends after every dense piece (P[nq]).
template <int UNROLL>
__device__ __forceinline__ void sb_dense_load(const SB_G sb_f4 *__restrict__ P, int nq, int q0, int tpar,
                                              sb_f4 *v)
{
#pragma unroll
    for (int u = 0; u < UNROLL; ++u) {
        const int q = q0 + u * tpar;
        v[u] = __builtin_nontemporal_load(&P[q < nq ? q : nq]);
    }
}

// The x index is clamped to the slice's first zero pad (xs[ncols]): past the
// end of the stream the loaded value is zero and so is the x it meets.
template <int UNROLL>
__device__ __forceinline__ void sb_dense_fma(const sb_f4 *v, const float *xs, const SbStreamLane &L, int kc,
                                             int M, int ncols, float &a0, float &a1, float &a2, float &a3)
{
#pragma unroll
    for (int u = 0; u < UNROLL; ++u) {
        const int k = kc + u * M;
        a0 += v[u].x * xs[min(L.c0 + k, ncols)];
        a1 += v[u].y * xs[min(L.c1 + k, ncols)];
        a2 += v[u].z * xs[min(L.c2 + k, ncols)];
        a3 += v[u].w * xs[min(L.c3 + k, ncols)];
    }
}

// Streams one dense piece whose x slice starts at xs.
template <int UNROLL>
__device__ __forceinline__ void sb_dense_piece(const SB_G sb_f4 *__restrict__ P, int nq, int tpar, int M, int ncols,
                                               const float *xs, const SbStreamLane &L, float &a0, float &a1,
                                               float &a2, float &a3)
{
    const int tid = threadIdx.x;
    if (tid >= tpar || tid >= nq) return;
    const int nmine = (nq - 1 - tid) / tpar + 1;             // float4 loads of this lane
    const int nbatch = (nmine + UNROLL - 1) / UNROLL;
    // The register copy cur = nxt waits for the batch it copies, so one
    // batch is in flight per round trip.  A copy-free form (two buffers that
    // trade roles) keeps two in flight but needs about 113 VGPRs (4 waves
    // per SIMD, two 512-thread workgroups per CU) and measured no faster on
    // the 100k matrix: the kernel runs at about 85 % of a cold streaming
    // read of the same bytes either way.
    sb_f4 cur[UNROLL], nxt[UNROLL];
    sb_dense_load<UNROLL>(P, nq, tid, tpar, cur);
    for (int b = 0; b < nbatch; ++b) {
        // always prefetch (the batch after the last one reads the zero float4)
        sb_dense_load<UNROLL>(P, nq, tid + (b + 1) * UNROLL * tpar, tpar, nxt);
        sb_dense_fma<UNROLL>(cur, xs, L, b * UNROLL * M, M, ncols, a0, a1, a2, a3);
#pragma unroll
        for (int u = 0; u < UNROLL; ++u) cur[u] = nxt[u];
    }
}

// Stages a group's x slices into xs through its slot map (x column per slot,
// -1 for the zero pads): two dependent loads per slot, 4 slots in flight per
// thread, no search over descriptors.
template <int WG>
__device__ __forceinline__ void sb_stage_map(const SB_G int32_t *__restrict__ map, int total,
                                             const SB_G float *__restrict__ x, float *xs)
{
    for (int i0 = 0; i0 < total; i0 += 4 * WG) {
        int m[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int i = i0 + u * WG + (int)threadIdx.x;
            m[u] = i < total ? __builtin_nontemporal_load(&map[i]) : -1;
        }
        float v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) v[u] = m[u] >= 0 ? x[m[u]] : 0.f;
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int i = i0 + u * WG + (int)threadIdx.x;
            if (i < total) xs[i] = v[u];
        }
    }
}

// One slice of a sparse piece, streamed by one wavefront: its step range
// from the table (wave-uniform), loads unconditional (the step index is
// clamped to the last step) and software pipelined one batch ahead; steps
// past the slice's count are masked to value 0 at the zero pad slot xs[pad].
// The G lanes of a row are folded with a fixed xor butterfly and the
// row-owner lane adds the sum to spy[row].
template <int SU>
__device__ __forceinline__ void sb_sparse_slice(const SB_G float *__restrict__ sval,
                                                const SB_G uint16_t *__restrict__ sidx,
                                                const SB_G int32_t *__restrict__ tab, const SB_G int32_t *__restrict__ map,
                                                int hs, int glog, int slice, int pad, const float *xs, float *spy)
{
    const int lane = threadIdx.x & 63;
    const int off = tab[2 * slice];
    const int nst = tab[2 * slice + 1];
    if (nst <= 0) return;
    const SB_G sb_f4 *__restrict__ V = reinterpret_cast<const SB_G sb_f4 *>(sval) + (int64_t)off * 64 + lane;
    const SB_G sb_u2 *__restrict__ I = reinterpret_cast<const SB_G sb_u2 *>(sidx) + (int64_t)off * 64 + lane;
    const unsigned padpair = (unsigned)pad | ((unsigned)pad << 16);
    sb_f4 cv[SU];
    sb_u2 ci[SU];
#pragma unroll
    for (int u = 0; u < SU; ++u) {
        const int j = min(u, nst - 1);
        cv[u] = __builtin_nontemporal_load(V + (int64_t)j * 64);
        ci[u] = __builtin_nontemporal_load(I + (int64_t)j * 64);
    }
    float acc = 0.f;
    for (int j0 = 0; j0 < nst; j0 += SU) {
        sb_f4 nv[SU];
        sb_u2 ni[SU];
#pragma unroll
        for (int u = 0; u < SU; ++u) {
            const int j = min(j0 + SU + u, nst - 1);
            nv[u] = __builtin_nontemporal_load(V + (int64_t)j * 64);
            ni[u] = __builtin_nontemporal_load(I + (int64_t)j * 64);
        }
#pragma unroll
        for (int u = 0; u < SU; ++u) {
            const bool ok = j0 + u < nst;
            const unsigned lo = ok ? ci[u].x : padpair;
            const unsigned hi = ok ? ci[u].y : padpair;
            acc += (ok ? cv[u].x : 0.f) * xs[lo & 0xffffu];
            acc += (ok ? cv[u].y : 0.f) * xs[lo >> 16];
            acc += (ok ? cv[u].z : 0.f) * xs[hi & 0xffffu];
            acc += (ok ? cv[u].w : 0.f) * xs[hi >> 16];
        }
#pragma unroll
        for (int u = 0; u < SU; ++u) {
            cv[u] = nv[u];
            ci[u] = ni[u];
        }
    }
    for (int m = 1; m < (1 << glog); m <<= 1) acc += __shfl_xor(acc, m, 64);
    const int i = slice * (64 >> glog) + (lane >> glog);
    if ((lane & ((1 << glog) - 1)) == 0 && i < hs) spy[map[i]] += acc;
}

template <int WG, int UNROLL, int XCAP>
__device__ __forceinline__ void sb_spmv_tile(const SbSpmvArgs &a, int t, float *xs, float *red, float *spy)
{
    const int tid = threadIdx.x;
    const SbTile tile = a.tiles[t];
    const int hs = tile.nrows;
    const int tpar = tile.dense_tpar;
    const int M = tile.dense_mcols;
    const SbMatBufs mb = sb_mat_bufs(a, tile.mat, a.xmap);
    const SB_G float *__restrict__ x = sb_gp(mb.x);
    SB_G float *__restrict__ y = sb_gp(mb.y);

    SbStreamLane L;
    {
        const int p0 = 4 * tid;
        L.c0 = p0 / hs;
        L.c1 = (p0 + 1) / hs;
        L.c2 = (p0 + 2) / hs;
        L.c3 = (p0 + 3) / hs;
    }
    float a0 = 0.f, a1 = 0.f, a2 = 0.f, a3 = 0.f;
    if (tid < hs) spy[tid] = 0.f;   // before the first staging barrier; sparse row owners add after it
    // roles: lanes tid < tpar stream the dense pieces; the last nsw
    // wavefronts (all of them when sp_waves == 0) take the sparse slices
    constexpr int NW = WG / 64;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int nsw = tile.sp_waves > 0 ? tile.sp_waves : NW;
    const int sw = wave - (NW - nsw);                     // sparse wavefront index, < 0: none
    const int nsl = (hs + (64 >> tile.sp_glog) - 1) / (64 >> tile.sp_glog);   // sparse slices
    int pb = tile.piece_begin;
    bool head = true;
    while (pb < tile.piece_end) {
        // staging group [pb, pe), formed by the host (stage_len <= XCAP + 4
        // slots); the first group's map offset and length come with the
        // tile, so its map loads need not wait for the piece descriptor
        const SbPiece first = a.pieces[pb];
        const int smap = head ? tile.stage_map : first.xmap_off;
        const int slen = head ? tile.stage_len : first.stage_len;
        head = false;
        if (slen > XCAP + 4) break;            // never built (the host and kernel choice check)
        __syncthreads();                       // previous group done with xs
        sb_stage_map<WG>(sb_gp(mb.xmap) + smap, slen, x, xs);
        __syncthreads();
        const int pe = first.grp_end;
        if (pe <= pb) break;                   // never built
        int base = 0;
        for (int p = pb; p < pe; ++p) {
            const SbPiece pc = a.pieces[p];
            if (pc.kind == SB_PIECE_DENSE) {
                sb_dense_piece<UNROLL>(reinterpret_cast<const SB_G sb_f4 *>(sb_gp(mb.dval) + pc.data_off),
                                       (hs * pc.ncols + 3) >> 2, tpar, M, pc.ncols, xs + base, L, a0, a1, a2, a3);
            } else if (sw >= 0) {
                // this wavefront's slices sw, sw + nsw, ...
                for (int sl = sw; sl < nsl; sl += nsw) {
                    sb_sparse_slice<SB_SPUNROLL>(sb_gp(mb.sval) + pc.data_off, sb_gp(mb.sidx) + pc.data_off,
                                                 sb_gp(mb.srowptr) + pc.rowptr_off, sb_gp(mb.srowptr) + tile.sp_map, hs,
                                                 tile.sp_glog, sl, pc.ncols, xs + base,
                                                 spy);
                }
            }
            base += pc.ncols + 4;
        }
        pb = pe;
    }

    __syncthreads();
    if (tid < tpar) {
        red[4 * tid] = a0;
        red[4 * tid + 1] = a1;
        red[4 * tid + 2] = a2;
        red[4 * tid + 3] = a3;
    }
    __syncthreads();
    // red holds M groups of hs partial sums: red[m*hs + r].  Fold halves while
    // there are many groups (short tiles), then each row owner sums in order.
    int m = M;
    while (m > 32) {
        const int half = (m + 1) >> 1;
        const int cnt = (m - half) * hs;
        for (int i = tid; i < cnt; i += WG) red[i] += red[i + half * hs];
        __syncthreads();
        m = half;
    }
    if (tid < hs) {
        float s = 0.f;
        for (int g = 0; g < m; ++g) s += red[g * hs + tid];
        y[tile.row0 + tid] = s + spy[tid];
    }
}

template <int WG, int UNROLL, int XCAP>
__device__ __forceinline__ void sb_spmv_body(const SbSpmvArgs &a)
{
    __shared__ float xs[XCAP + 4];
    __shared__ float red[4 * WG];
    __shared__ float spy[WG];
    __shared__ int next_tile;
    sb_for_tiles(a.ntiles, a.mode, a.queue, &next_tile,
                 [&](int t) { sb_spmv_tile<WG, UNROLL, XCAP>(a, t, xs, red, spy); });
}

// ---------------------------------------------------------------------------
// SpMM, FP32 on the vector ALUs (the plain path)
//
// Narrow shape (N not SB_F32_WIDE): thread (rg, v): row group rg = tid / NV
// owns rows rg*rt .. rg*rt+rt-1 of the tile and RHS columns 4v..4v+3 (NV =
// n_rhs/4).  Dense piece values are read once per (row, column) and
// broadcast to the NV lanes of a row group.  The wide shape (below) holds 8
// columns per thread and reads A as float4 row quads.  Both stage a piece's X
// rows in LDS once per tile and walk sparse rows entry by entry with the same
// thread mapping.

// Expands a piece's x ranges into xrow[0..ncols): the X row of every column.
__device__ __forceinline__ void sb_expand_xrow(const SbXRange *__restrict__ xr, int r0, int r1, int *xrow)
{
    int base = 0;
    for (int r = r0; r < r1; ++r) {
        const SB_G SbXRange *g = sb_gp(xr) + r;
        const int col0 = g->col0, len = g->len;
        for (int i = threadIdx.x; i < len; i += blockDim.x) xrow[base + i] = col0 + i;
        base += len;
    }
}

// Stages the X rows of a piece (ncols x N floats, row xrow[c] of X for local
// column c) into xs4[c * N/4 + v], four float4 loads in flight per thread,
// coalesced along N.
template <int WG>
__device__ __forceinline__ void sb_stage_xrows(const SB_G float *__restrict__ X, int N, const int *xrow, int ncols,
                                               float4 *xs4)
{
    const int NV = N >> 2;
    const int tot = ncols * NV;
    for (int i0 = 0; i0 < tot; i0 += 4 * WG) {
        sb_f4 v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int i = i0 + u * WG + (int)threadIdx.x;
            const int c = i < tot ? i / NV : 0;
            const int q = i < tot ? i - c * NV : 0;
            v[u] = *reinterpret_cast<const SB_G sb_f4 *>(X + (size_t)xrow[c] * N + 4 * q);
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int i = i0 + u * WG + (int)threadIdx.x;
            if (i < tot) reinterpret_cast<sb_f4 *>(xs4)[i] = v[u];
        }
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

#ifndef SB_SPMM_SPCHUNK
#define SB_SPMM_SPCHUNK 512
#endif

__device__ __forceinline__ void sb_fma4(float4 &acc, float a, const float4 &x)
{
    acc.x += a * x.x;
    acc.y += a * x.y;
    acc.z += a * x.z;
    acc.w += a * x.w;
}

// Accumulates one FP32-dense or sparse piece into acc (RTMAX row registers).
// The piece's X rows are in LDS (xs4, staged once per tile), so every X read
// is an LDS read.  Dense: columns are processed 4 at a time with every load
// of the group issued before its FMAs.  Sparse: the piece's entries are
// staged through LDS (value and local column) in chunks of SB_SPMM_SPCHUNK,
// then each thread walks its rows' entries.  spv/spx are LDS scratch of
// SB_SPMM_SPCHUNK entries.
template <int WG, int RTMAX>
__device__ __forceinline__ void sb_spmm_vec_piece(const SbMatBufs &a, const SbPiece &pc, int hs, int N,
                                                  const SbVecMap &m, const float4 *xs4, float4 *acc, float *spv,
                                                  int *spx)
{
    const int NV = N >> 2;
    if (pc.kind == SB_PIECE_DENSE) {
        if (!m.active) return;
        const SB_G float *__restrict__ P = sb_gp(a.dval) + pc.data_off;
        int c = 0;
        for (; c + 4 <= pc.ncols; c += 4) {
            float4 xv[4];
            float av[4][RTMAX];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                xv[u] = xs4[(c + u) * NV + m.v];
                const SB_G float *__restrict__ col = P + (size_t)(c + u) * pc.pitch;
#pragma unroll
                for (int i = 0; i < RTMAX; ++i) {
                    const int r = m.rbeg + i;
                    av[u][i] = (i < m.rt && r < hs) ? col[r] : 0.f;
                }
            }
#pragma unroll
            for (int u = 0; u < 4; ++u)
#pragma unroll
                for (int i = 0; i < RTMAX; ++i) sb_fma4(acc[i], av[u][i], xv[u]);
        }
        for (; c < pc.ncols; ++c) {
            const float4 xv = xs4[c * NV + m.v];
            const SB_G float *__restrict__ col = P + (size_t)c * pc.pitch;
#pragma unroll
            for (int i = 0; i < RTMAX; ++i) {
                const int r = m.rbeg + i;
                if (i < m.rt && r < hs) sb_fma4(acc[i], col[r], xv);
            }
        }
        return;
    }
    if (pc.kind != SB_PIECE_SPARSE) return;
    const SB_G int32_t *__restrict__ rp = sb_gp(a.srowptr) + pc.rowptr_off;
    const SB_G float *__restrict__ sv = sb_gp(a.sval) + pc.data_off;
    const SB_G uint16_t *__restrict__ si = sb_gp(a.sidx) + pc.data_off;
    const int nnzp = rp[hs];
    int rlo[RTMAX], rhi[RTMAX];
#pragma unroll
    for (int i = 0; i < RTMAX; ++i) {
        const int r = m.rbeg + i;
        const bool ok = m.active && i < m.rt && r < hs;
        rlo[i] = ok ? rp[r] : 0;
        rhi[i] = ok ? rp[r + 1] : 0;
    }
    for (int e0 = 0; e0 < nnzp; e0 += SB_SPMM_SPCHUNK) {
        const int n = min(SB_SPMM_SPCHUNK, nnzp - e0);
        __syncthreads();
        for (int e = threadIdx.x; e < n; e += WG) {
            spv[e] = sv[e0 + e];
            spx[e] = si[e0 + e];
        }
        __syncthreads();
        if (!m.active) continue;
#pragma unroll
        for (int i = 0; i < RTMAX; ++i) {
            const int lo = max(rlo[i], e0) - e0, hi = min(rhi[i], e0 + n) - e0;
            int e = lo;
            for (; e + 4 <= hi; e += 4) {
                float4 xv[4];
#pragma unroll
                for (int u = 0; u < 4; ++u)
                    xv[u] = xs4[spx[e + u] * NV + m.v];
#pragma unroll
                for (int u = 0; u < 4; ++u) sb_fma4(acc[i], spv[e + u], xv[u]);
            }
            for (; e < hi; ++e)
                sb_fma4(acc[i], spv[e], xs4[spx[e] * NV + m.v]);
        }
    }
}

template <int WG, int RTMAX, int NRHS>
__device__ __forceinline__ void sb_spmm_f32_tile(const SbSpmmArgs &a, int t, int *xrow, float4 *xs4, float *spv,
                                                 int *spx)
{
    const int N = NRHS ? NRHS : a.n_rhs;
    const SbTile tile = a.tiles[t];
    const int hs = tile.nrows;
    const SbVecMap m = sb_vec_map(hs, N, WG);
    const SbMatBufs mb = sb_mat_bufs(a, tile.mat, nullptr);
    const SB_G float *__restrict__ X = sb_gp(mb.x);
    SB_G float *__restrict__ Y = sb_gp(mb.y);
    float4 acc[RTMAX];
#pragma unroll
    for (int i = 0; i < RTMAX; ++i) acc[i] = make_float4(0.f, 0.f, 0.f, 0.f);

    for (int pi = tile.piece_begin; pi < tile.piece_end; ++pi) {
        const SbPiece pc = a.pieces[pi];
        __syncthreads();
        sb_expand_xrow(a.xr, pc.xr_begin, pc.xr_end, xrow);
        __syncthreads();
        sb_stage_xrows<WG>(X, N, xrow, pc.ncols, xs4);
        __syncthreads();
        sb_spmm_vec_piece<WG, RTMAX>(mb, pc, hs, N, m, xs4, acc, spv, spx);
    }
    if (m.active) {
#pragma unroll
        for (int i = 0; i < RTMAX; ++i) {
            const int r = m.rbeg + i;
            if (i < m.rt && r < hs) {
                SB_G sb_f4 *dst = reinterpret_cast<SB_G sb_f4 *>(Y + (size_t)(tile.row0 + r) * N + 4 * m.v);
                sb_f4 o = {acc[i].x, acc[i].y, acc[i].z, acc[i].w};
                if (tile.y_accum) {      // the rows' MFMA tile wrote Y in the previous launch
                    const sb_f4 w = *dst;
                    o = sb_f4{w.x + o.x, w.y + o.y, w.z + o.z, w.w + o.w};
                }
                *dst = o;
            }
        }
    }
}

// Wide shape (SB_F32_WIDE(N)): thread (rg, v) of NV8 = N/8 lanes per row
// group holds rows [rbeg, rbeg + rt) (rt a multiple of 4 in tiles with dense
// pieces) and the column quads 4v and N/2 + 4v, so the lanes of a row group
// read consecutive float4s of the staged X row in each half.  A dense
// column's row quads are float4 loads from the padded panel (pitch: the
// rows rounded up to 4), requested SB_F32_DEPTH column pairs ahead through a
// ring of register sets: per column, one vector-memory load and two LDS
// loads for 32 FMAs at rt = 4.  Only tiles with dense pieces read row quads
// (nq >= 1 for every active thread there).
struct SbWideMap {
    int rt;       // rows per row group (multiple of 4, uniform over the tile)
    int rbeg;     // first tile row of this thread's group
    int v;        // lane in the row group
    int nq;       // row quads of this thread inside the panel
    bool active;
};

// quads: rt rounded up to whole row quads (tiles with dense pieces, whose A
// columns are read as float4 row quads); else rt = ceil(hs / groups)
__device__ __forceinline__ SbWideMap sb_wide_map(int hs, int N, int wg, bool quads)
{
    SbWideMap m;
    const int NV8 = N >> 3;
    const int ngroups = wg / NV8;
    m.rt = (hs + ngroups - 1) / ngroups;
    if (quads) m.rt = (m.rt + 3) & ~3;
    const int rg = threadIdx.x / NV8;
    m.v = threadIdx.x - rg * NV8;
    m.rbeg = rg * m.rt;
    m.active = rg < ngroups && m.rbeg < hs;
    const int pitch = (hs + 3) & ~3;
    m.nq = m.active ? min(m.rt, pitch - m.rbeg) >> 2 : 0;
    return m;
}

// Wide FP32 SpMM tuning: column pairs of A in flight per thread (register
// sets of the ring), sparse entries per unrolled step, and the waves per
// SIMD the launch bound asks of the compiler (4: at most 128 VGPRs, two
// 512-thread workgroups per CU, which the X staging LDS also allows)
#ifndef SB_F32_DEPTH
#define SB_F32_DEPTH 4
#endif
#ifndef SB_F32_SPUNROLL
#define SB_F32_SPUNROLL 4
#endif
#ifndef SB_F32_SPCHUNK
#define SB_F32_SPCHUNK 1536
#endif
#ifndef SB_F32_MINWAVES
#define SB_F32_MINWAVES 4
#endif
#define SB_F32_WAVES(N) (SB_F32_WIDE(N) ? SB_F32_MINWAVES : 1)

// r[j] += v * x[j] over the thread's 8 columns
__device__ __forceinline__ void sb_wide_fma1(float (&r)[8], float v, const float4 &x0, const float4 &x1)
{
    r[0] += v * x0.x;
    r[1] += v * x0.y;
    r[2] += v * x0.z;
    r[3] += v * x0.w;
    r[4] += v * x1.x;
    r[5] += v * x1.y;
    r[6] += v * x1.z;
    r[7] += v * x1.w;
}

// acc[4q + k][j] += a[q][k] * x[j]
template <int RQ>
__device__ __forceinline__ void sb_wide_fma(float (&acc)[4 * RQ][8], const sb_f4 (&a)[RQ], const float4 &x0,
                                            const float4 &x1)
{
#pragma unroll
    for (int q = 0; q < RQ; ++q) {
#pragma unroll
        for (int k = 0; k < 4; ++k) sb_wide_fma1(acc[4 * q + k], a[q][k], x0, x1);
    }
}

// The thread's row quads of panel column c.  Quads past the panel (only in
// the last row group) repeat its last quad: their rows are never stored.
template <int RQ>
__device__ __forceinline__ void sb_wide_load(const SB_G float *__restrict__ P, int pitch, int c, int nq,
                                             sb_f4 (&a)[RQ])
{
    const SB_G sb_f4 *__restrict__ col = reinterpret_cast<const SB_G sb_f4 *>(P + (size_t)c * pitch);
#pragma unroll
    for (int q = 0; q < RQ; ++q) a[q] = col[min(q, nq - 1)];
}

// One dense piece: acc += panel x staged X rows.
template <int RQ>
__device__ __forceinline__ void sb_spmm_wide_dense(const SbMatBufs &a, const SbPiece &pc, int N, const SbWideMap &m,
                                                   const float4 *xs4, float (&acc)[4 * RQ][8])
{
    const int NV = N >> 2, H = N >> 3;
    {
        if (!m.active) return;
        // column pairs through a ring of SB_F32_DEPTH register sets: the
        // set just consumed is refilled with the pair SB_F32_DEPTH ahead
        // (clamped to the last column, so always in bounds)
        constexpr int D = SB_F32_DEPTH;
        const int n = pc.ncols, pitch = pc.pitch, np = n >> 1;
        const SB_G float *__restrict__ P = sb_gp(a.dval) + pc.data_off + m.rbeg;
        sb_f4 buf[D][2][RQ];
#pragma unroll
        for (int d = 0; d < D; ++d) {
            sb_wide_load<RQ>(P, pitch, min(2 * d, n - 1), m.nq, buf[d][0]);
            sb_wide_load<RQ>(P, pitch, min(2 * d + 1, n - 1), m.nq, buf[d][1]);
        }
        int p = 0;
        for (; p + D <= np; p += D) {
#pragma unroll
            for (int d = 0; d < D; ++d) {
                const int c = 2 * (p + d);
                const float4 x00 = xs4[c * NV + m.v], x01 = xs4[c * NV + H + m.v];
                const float4 x10 = xs4[(c + 1) * NV + m.v], x11 = xs4[(c + 1) * NV + H + m.v];
                sb_wide_fma<RQ>(acc, buf[d][0], x00, x01);
                sb_wide_fma<RQ>(acc, buf[d][1], x10, x11);
                sb_wide_load<RQ>(P, pitch, min(c + 2 * D, n - 1), m.nq, buf[d][0]);
                sb_wide_load<RQ>(P, pitch, min(c + 2 * D + 1, n - 1), m.nq, buf[d][1]);
                // keep the refill ahead of the next set's FMAs (the
                // scheduler otherwise sinks it to the end of the loop)
                __builtin_amdgcn_sched_barrier(0);
            }
        }
        // the last (np - p) < D pairs are in buf[0 ..]
#pragma unroll
        for (int d = 0; d < D - 1; ++d) {
            if (p + d < np) {
                const int c = 2 * (p + d);
                sb_wide_fma<RQ>(acc, buf[d][0], xs4[c * NV + m.v], xs4[c * NV + H + m.v]);
                sb_wide_fma<RQ>(acc, buf[d][1], xs4[(c + 1) * NV + m.v], xs4[(c + 1) * NV + H + m.v]);
            }
        }
        if (n & 1) {    // odd column count: the last column on its own
            sb_f4 l[RQ];
            sb_wide_load<RQ>(P, pitch, n - 1, m.nq, l);
            sb_wide_fma<RQ>(acc, l, xs4[(n - 1) * NV + m.v], xs4[(n - 1) * NV + H + m.v]);
        }
        return;
    }
}

// One sparse piece (CSR rows of the tile): its entries are staged in LDS in
// chunks of SPCH (value, local column), then each thread walks its rows'
// entries, U entries' loads in flight at a time.
template <int WG, int RQ, int SPCH, int U>
__device__ __forceinline__ void sb_spmm_wide_sparse(const SbMatBufs &a, const SbPiece &pc, int hs, int N,
                                                    const SbWideMap &m, const float4 *xs4,
                                                    float (&acc)[4 * RQ][8], float *spv, int *spx)
{
    const int NV = N >> 2, H = N >> 3;
    const SB_G int32_t *__restrict__ rp = sb_gp(a.srowptr) + pc.rowptr_off;
    const SB_G float *__restrict__ sv = sb_gp(a.sval) + pc.data_off;
    const SB_G uint16_t *__restrict__ si = sb_gp(a.sidx) + pc.data_off;
    const int nnzp = rp[hs];
    int rlo[4 * RQ], rhi[4 * RQ];
#pragma unroll
    for (int i = 0; i < 4 * RQ; ++i) {
        const int r = m.rbeg + i;
        const bool ok = m.active && i < m.rt && r < hs;
        rlo[i] = ok ? rp[r] : 0;
        rhi[i] = ok ? rp[r + 1] : 0;
    }
    for (int e0 = 0; e0 < nnzp; e0 += SPCH) {
        const int cnt = min(SPCH, nnzp - e0);
        __syncthreads();
        for (int e = threadIdx.x; e < cnt; e += WG) {
            spv[e] = sv[e0 + e];
            spx[e] = si[e0 + e];
        }
        __syncthreads();
        if (!m.active) continue;
#pragma unroll
        for (int i = 0; i < 4 * RQ; ++i) {
            const int lo = max(rlo[i], e0) - e0, hi = min(rhi[i], e0 + cnt) - e0;
            // SB_F32_SPUNROLL entries' loads in flight, then their FMAs in
            // entry order
            int e = lo;
            for (; e + U <= hi; e += U) {
                float v[U];
                float4 x0[U], x1[U];
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    v[u] = spv[e + u];
                    const int ci = spx[e + u];
                    x0[u] = xs4[ci * NV + m.v];
                    x1[u] = xs4[ci * NV + H + m.v];
                }
#pragma unroll
                for (int u = 0; u < U; ++u) sb_wide_fma1(acc[i], v[u], x0[u], x1[u]);
            }
            for (; e < hi; ++e) {
                const int ci = spx[e];
                sb_wide_fma1(acc[i], spv[e], xs4[ci * NV + m.v], xs4[ci * NV + H + m.v]);
            }
        }
    }
}

// Stages the X rows of piece pc into xs4 (after every thread is done with
// the previous piece's).
template <int WG>
__device__ __forceinline__ void sb_wide_stage(const SbSpmmArgs &a, const SbPiece &pc, const SB_G float *X, int N,
                                              int *xrow, float4 *xs4)
{
    __syncthreads();
    sb_expand_xrow(a.xr, pc.xr_begin, pc.xr_end, xrow);
    __syncthreads();
    sb_stage_xrows<WG>(X, N, xrow, pc.ncols, xs4);
    __syncthreads();
}

// One tile with up to 4 * RQ rows per thread group: its dense pieces first,
// then its sparse pieces (two loops, so the registers the dense loop keeps
// are free in the sparse one).  DENSE false: the tile has no dense pieces.
template <int WG, int RQ, int NRHS, bool DENSE>
__device__ __forceinline__ void sb_spmm_wide_tile_q(const SbSpmmArgs &a, const SbTile &tile, const SbWideMap &m,
                                                    int *xrow, float4 *xs4, float *spv, int *spx)
{
    const int N = NRHS ? NRHS : a.n_rhs;
    const int hs = tile.nrows;
    const SbMatBufs mb = sb_mat_bufs(a, tile.mat, nullptr);
    const SB_G float *__restrict__ X = sb_gp(mb.x);
    SB_G float *__restrict__ Y = sb_gp(mb.y);
    float acc[4 * RQ][8];
#pragma unroll
    for (int i = 0; i < 4 * RQ; ++i)
#pragma unroll
        for (int j = 0; j < 8; ++j) acc[i][j] = 0.f;

    if constexpr (DENSE) {
        for (int pi = tile.piece_begin; pi < tile.piece_end; ++pi) {
            const SbPiece pc = a.pieces[pi];
            if (pc.kind != SB_PIECE_DENSE) continue;
            sb_wide_stage<WG>(a, pc, X, N, xrow, xs4);
            sb_spmm_wide_dense<RQ>(mb, pc, N, m, xs4, acc);
        }
    }
    for (int pi = tile.piece_begin; pi < tile.piece_end; ++pi) {
        const SbPiece pc = a.pieces[pi];
        if (pc.kind != SB_PIECE_SPARSE) continue;
        sb_wide_stage<WG>(a, pc, X, N, xrow, xs4);
        sb_spmm_wide_sparse<WG, RQ, SB_F32_SPCHUNK, SB_F32_SPUNROLL>(mb, pc, hs, N, m, xs4, acc, spv, spx);
    }
    if (!m.active) return;
#pragma unroll
    for (int i = 0; i < 4 * RQ; ++i) {
        const int r = m.rbeg + i;
        if (i < m.rt && r < hs) {
            SB_G float *row = Y + (size_t)(tile.row0 + r) * N + 4 * m.v;
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                SB_G sb_f4 *dst = reinterpret_cast<SB_G sb_f4 *>(row + h * (N >> 1));
                sb_f4 o = {acc[i][4 * h], acc[i][4 * h + 1], acc[i][4 * h + 2], acc[i][4 * h + 3]};
                if (tile.y_accum) {      // the rows' MFMA tile wrote Y in the previous launch
                    const sb_f4 w = *dst;
                    o = sb_f4{w.x + o.x, w.y + o.y, w.z + o.z, w.w + o.w};
                }
                *dst = o;
            }
        }
    }
}

// rt (uniform over the tile, at most 4 * RQ) picks the instance: 4 or 8
// rows per thread group, with or without the dense loop.
template <int WG, int RQ, int NRHS>
__device__ __forceinline__ void sb_spmm_wide_tile(const SbSpmmArgs &a, int t, int *xrow, float4 *xs4, float *spv,
                                                  int *spx)
{
    static_assert(RQ == 1 || RQ == 2, "the wide shape holds 4 or 8 rows per thread group");
    const int N = NRHS ? NRHS : a.n_rhs;
    const SbTile tile = a.tiles[t];
    bool dense = false;
    for (int pi = tile.piece_begin; pi < tile.piece_end; ++pi) dense = dense || a.pieces[pi].kind == SB_PIECE_DENSE;
    const SbWideMap m = sb_wide_map(tile.nrows, N, WG, dense);
    if (dense) {
        if (RQ == 2 && m.rt > 4) sb_spmm_wide_tile_q<WG, RQ, NRHS, true>(a, tile, m, xrow, xs4, spv, spx);
        else sb_spmm_wide_tile_q<WG, 1, NRHS, true>(a, tile, m, xrow, xs4, spv, spx);
    } else {
        if (RQ == 2 && m.rt > 4) sb_spmm_wide_tile_q<WG, RQ, NRHS, false>(a, tile, m, xrow, xs4, spv, spx);
        else sb_spmm_wide_tile_q<WG, 1, NRHS, false>(a, tile, m, xrow, xs4, spv, spx);
    }
}

template <int WG, int RTMAX, int NRHS, int XCAP>
__device__ __forceinline__ void sb_spmm_f32_body(const SbSpmmArgs &a)
{
    // staged X rows of the current piece: pieces have at most
    // SB_SPMM_XS_FLOATS / N columns (the plan builder cuts them so)
    constexpr int XS4 = NRHS ? (XCAP < SB_SPMM_XS_FLOATS / NRHS ? XCAP : SB_SPMM_XS_FLOATS / NRHS) * (NRHS / 4)
                             : SB_SPMM_XS_FLOATS / 4;
    __shared__ int xrow[XCAP + 4];
    __shared__ float4 xs4[XS4];
    // sparse entry chunks: SB_F32_SPCHUNK entries in the wide shape,
    // SB_SPMM_SPCHUNK in the narrow one (the generic instance runs either)
    constexpr int SPC = NRHS == 0 ? (SB_F32_SPCHUNK > SB_SPMM_SPCHUNK ? SB_F32_SPCHUNK : SB_SPMM_SPCHUNK)
                        : SB_F32_WIDE(NRHS) ? SB_F32_SPCHUNK : SB_SPMM_SPCHUNK;
    __shared__ float spv[SPC];
    __shared__ int spx[SPC];
    __shared__ int next_tile;
    // wide: tiles hold at most SB_F32_WIDE_ROWS rows per thread group (the
    // plan builder cuts them so); RTMAX <= 4 leaves out the 8-row instance
    constexpr int RQ = RTMAX > 4 ? 2 : 1;
    if constexpr (NRHS != 0 && SB_F32_WIDE(NRHS)) {
        sb_for_tiles(a.ntiles, a.mode, a.queue, &next_tile,
                     [&](int t) { sb_spmm_wide_tile<WG, RQ, NRHS>(a, t, xrow, xs4, spv, spx); });
    } else if constexpr (NRHS != 0) {
        sb_for_tiles(a.ntiles, a.mode, a.queue, &next_tile,
                     [&](int t) { sb_spmm_f32_tile<WG, RTMAX, NRHS>(a, t, xrow, xs4, spv, spx); });
    } else if (SB_F32_WIDE(a.n_rhs)) {     // generic instance: the shape follows N
        sb_for_tiles(a.ntiles, a.mode, a.queue, &next_tile,
                     [&](int t) { sb_spmm_wide_tile<WG, SB_F32_WIDE_ROWS / 4, 0>(a, t, xrow, xs4, spv, spx); });
    } else {
        sb_for_tiles(a.ntiles, a.mode, a.queue, &next_tile,
                     [&](int t) { sb_spmm_f32_tile<WG, RTMAX, NRHS>(a, t, xrow, xs4, spv, spx); });
    }
}
