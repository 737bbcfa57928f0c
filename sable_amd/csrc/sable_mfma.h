// SpMM with dense blocks on MFMA: bf16 splits, FP32 accumulation.
//
// Every FP32 value a is split in registers into bf16 terms by round to
// nearest even: a1 = bf16(a), a2 = bf16(a - a1), a3 = bf16(a - a1 - a2).
// Each residual is exact in FP32, and a - a1 - a2 has at most 8 significant
// bits, so a1 + a2 + a3 == a exactly (for normal numbers).  X is split the
// same way while it is staged.  The product is formed from
// v_mfma_f32_16x16x32_bf16 passes over the term pairs whose order is above
// 2^-24 (every bf16 x bf16 product is exact in FP32):
//   SPLIT = 3 (6 passes): a1x1 + a1x2 + a2x1 + a1x3 + a2x2 + a3x1
//                         -> dropped terms ~2^-24 relative: FP32-level
//   SPLIT = 2 (3 passes): a1x1 + a1x2 + a2x1 (the hi+lo scheme)
//                         -> worst case ~3 * 2^-16 relative per product
//
// A (the matrix) is stored by the plan builder as FP32 in MFMA register
// order: for 16-row tile m and 32-column k-step s, lane l holds A[16m + l%16]
// [32s + 8(l/16) + j], j = 0..7, as 8 consecutive floats (two 16-byte
// loads).  Storing FP32 keeps the A stream at 4 bytes per value for either
// split.  X is staged per k-step: the 32 X rows of the piece are read
// (coalesced along N), split, and written to LDS as [n][k] term arrays with a
// 40-element pitch so the B-operand reads (lane l: n = l%16, k = 8(l/16) ..
// +7) are conflict-free 16-byte LDS reads.
//
// A tile may also hold FP32-dense and sparse pieces (other blocks of the
// same block row); those run on the vector ALUs with the FP32 kernel's
// thread mapping, and both partial results meet in an LDS Y tile.
#pragma once

#include "sable_kernels.h"

typedef __attribute__((ext_vector_type(8))) __bf16 sb_bf16x8;
typedef __attribute__((ext_vector_type(4))) float sb_f32x4;

#define SB_KP 40            // LDS pitch of a staged X column (bf16 elements)
#define SB_MFMA_MAXROWS 64  // tile height limit of the MFMA kernel

#ifndef SB_SPLIT
#define SB_SPLIT 3
#endif

__device__ __forceinline__ uint32_t sb_bf16_bits(float f)
{
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x007fffffu)) return (u >> 16) | 0x40u;
    u += 0x7fffu + ((u >> 16) & 1u);
    return u >> 16;
}

__device__ __forceinline__ float sb_bf16_float(uint32_t h) { return __uint_as_float(h << 16); }

__device__ __forceinline__ sb_bf16x8 sb_as_bf16x8(sb_u4 v) { return __builtin_bit_cast(sb_bf16x8, v); }

// Splits 8 floats into SPLIT bf16x8 terms.
template <int SPLIT>
__device__ __forceinline__ void sb_split8(const float *f, sb_bf16x8 *t)
{
    uint32_t w[SPLIT][4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        float ra = f[2 * j], rb = f[2 * j + 1];
#pragma unroll
        for (int s = 0; s < SPLIT; ++s) {
            const uint32_t ha = sb_bf16_bits(ra), hb = sb_bf16_bits(rb);
            w[s][j] = ha | (hb << 16);
            ra -= sb_bf16_float(ha);
            rb -= sb_bf16_float(hb);
        }
    }
#pragma unroll
    for (int s = 0; s < SPLIT; ++s) t[s] = sb_as_bf16x8(sb_u4{w[s][0], w[s][1], w[s][2], w[s][3]});
}

// X staging of an MFMA piece, split in two phases so the global loads of
// k-step s+1 are in flight while k-step s computes: sb_fetch_b loads the
// thread's X values of a k-step into registers, sb_store_b splits them and
// writes B[s][n * SB_KP + k] (k < 32, s < SPLIT) to LDS.
template <int WG, int UPT>
__device__ __forceinline__ void sb_fetch_b(const float *__restrict__ X, int N, const int *xrow, int ncols, int ks,
                                           float (*f)[8])
{
    const int units = N * 4;
#pragma unroll
    for (int uu = 0; uu < UPT; ++uu) {
        const int u = threadIdx.x + uu * WG;
        const int kg = u / N;
        const int n = u - kg * N;
        const int k0 = ks * 32 + kg * 8;
#pragma unroll
        for (int j = 0; j < 8; ++j)
            f[uu][j] = (u < units && k0 + j < ncols) ? X[(size_t)xrow[k0 + j] * N + n] : 0.f;
    }
}

template <int WG, int UPT, int SPLIT>
__device__ __forceinline__ void sb_store_b(int N, float (*f)[8], uint16_t *B, int term_stride)
{
    const int units = N * 4;
#pragma unroll
    for (int uu = 0; uu < UPT; ++uu) {
        const int u = threadIdx.x + uu * WG;
        if (u >= units) continue;
        const int kg = u / N;
        const int n = u - kg * N;
        sb_bf16x8 t[SPLIT];
        sb_split8<SPLIT>(f[uu], t);
#pragma unroll
        for (int s = 0; s < SPLIT; ++s)
            *reinterpret_cast<sb_bf16x8 *>(B + s * term_stride + n * SB_KP + kg * 8) = t[s];
    }
}

template <int SPLIT>
__device__ __forceinline__ sb_f32x4 sb_mfma_terms(const sb_bf16x8 *a, const sb_bf16x8 *b, sb_f32x4 acc)
{
    // largest-order terms last so the small ones accumulate first
    if (SPLIT == 3) {
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[2], b[0], acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[1], b[1], acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0], b[2], acc, 0, 0, 0);
    }
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[1], b[0], acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0], b[1], acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0], b[0], acc, 0, 0, 0);
    return acc;
}

template <int WG, int NRHS, int SPLIT>
struct SbMfmaCfg {
    static constexpr int NMAX = NRHS ? NRHS : 256;
    static constexpr int NTMAX = NMAX / 16;
    // vector-path rows per thread for a 64-row tile: ceil(64 * (N/4) / WG)
    static constexpr int RT0 = (SB_MFMA_MAXROWS * (NMAX / 4) + WG - 1) / WG;
    static constexpr int RTMAX = RT0 < 1 ? 1 : RT0;
    static constexpr int TERM_STRIDE = NMAX * SB_KP;            // bf16 elements per staged term
    static constexpr int UPT = (NMAX * 4 + WG - 1) / WG;         // X staging units per thread
    static constexpr int BSTAGE_BYTES = SPLIT * TERM_STRIDE * 2;
    static constexpr int YTILE_BYTES = SB_MFMA_MAXROWS * NMAX * 4;
    static constexpr int SMEM0 = BSTAGE_BYTES > YTILE_BYTES ? BSTAGE_BYTES : YTILE_BYTES;
    static constexpr int SMEM_BYTES = SMEM0 > 8 * SB_SPMM_SPCHUNK ? SMEM0 : 8 * SB_SPMM_SPCHUNK;
};

template <int WG, int NRHS, int SPLIT>
__device__ __forceinline__ void sb_spmm_mfma_tile(const SbSpmmArgs &a, int t, int *xrow, unsigned char *smem)
{
    using Cfg = SbMfmaCfg<WG, NRHS, SPLIT>;
    sb_clock(a.tile_clock, t, 0);
    const int N = NRHS ? NRHS : a.n_rhs;
    const int nt = N >> 4;
    const SbTile tile = a.tiles[t];
    const int hs = tile.nrows;
    const int mt = (hs + 15) >> 4;
    const int wave = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    constexpr int NWAVES = WG / 64;
    int ng = 1;
    while (ng * 2 <= nt && mt * ng * 2 <= NWAVES) ng <<= 1;
    const int mi = wave / ng;
    const int gi = wave - mi * ng;
    const bool mactive = mi < mt;
    const int ntw = nt / ng;
    const int j0 = gi * ntw;
    const float *__restrict__ X = a.xmat ? a.xmat[tile.mat] : a.x0;
    float *__restrict__ Y = a.ymat ? a.ymat[tile.mat] : a.y0;
    uint16_t *B = reinterpret_cast<uint16_t *>(smem);
    float *ytile = reinterpret_cast<float *>(smem);

    sb_f32x4 acc[Cfg::NTMAX];
#pragma unroll
    for (int j = 0; j < Cfg::NTMAX; ++j) acc[j] = sb_f32x4{0.f, 0.f, 0.f, 0.f};
    const SbVecMap vm = sb_vec_map(hs, N, WG);
    float4 vacc[Cfg::RTMAX];
#pragma unroll
    for (int i = 0; i < Cfg::RTMAX; ++i) vacc[i] = make_float4(0.f, 0.f, 0.f, 0.f);
    bool has_mfma = false, has_vec = false;

    for (int pi = tile.piece_begin; pi < tile.piece_end; ++pi) {
        const SbPiece pc = a.pieces[pi];
        __syncthreads();
        sb_expand_xrow(a.xr, pc.xr_begin, pc.xr_end, xrow);
        __syncthreads();
        if (pc.kind == SB_PIECE_MFMA) {
            has_mfma = true;
            const sb_f4 *__restrict__ A = reinterpret_cast<const sb_f4 *>(a.dval + pc.data_off);
            float xf[Cfg::UPT][8];
            sb_f4 a0 = {0.f, 0.f, 0.f, 0.f}, a1 = {0.f, 0.f, 0.f, 0.f};
            sb_f4 n0 = a0, n1 = a1;
            sb_fetch_b<WG, Cfg::UPT>(X, N, xrow, pc.ncols, 0, xf);
            if (mactive) {
                const int off = (mi * pc.ksteps * 64 + lane) * 2;
                n0 = __builtin_nontemporal_load(&A[off]);
                n1 = __builtin_nontemporal_load(&A[off + 1]);
            }
            for (int ks = 0; ks < pc.ksteps; ++ks) {
                __syncthreads();                       // B free: previous k-step consumed
                sb_store_b<WG, Cfg::UPT, SPLIT>(N, xf, B, Cfg::TERM_STRIDE);
                a0 = n0;
                a1 = n1;
                __syncthreads();
                if (ks + 1 < pc.ksteps) {              // next k-step's loads overlap this one's MFMAs
                    sb_fetch_b<WG, Cfg::UPT>(X, N, xrow, pc.ncols, ks + 1, xf);
                    if (mactive) {
                        const int off = ((mi * pc.ksteps + ks + 1) * 64 + lane) * 2;
                        n0 = __builtin_nontemporal_load(&A[off]);
                        n1 = __builtin_nontemporal_load(&A[off + 1]);
                    }
                }
                if (mactive) {
                    const float af[8] = {a0.x, a0.y, a0.z, a0.w, a1.x, a1.y, a1.z, a1.w};
                    sb_bf16x8 at[SPLIT];
                    sb_split8<SPLIT>(af, at);
                    const int nb = lane & 15, kb = (lane >> 4) * 8;
#pragma unroll
                    for (int jj = 0; jj < Cfg::NTMAX; ++jj) {
                        if (jj < ntw) {
                            const int n = (j0 + jj) * 16 + nb;
                            sb_bf16x8 bt[SPLIT];
#pragma unroll
                            for (int s = 0; s < SPLIT; ++s)
                                bt[s] = *reinterpret_cast<const sb_bf16x8 *>(B + s * Cfg::TERM_STRIDE + n * SB_KP + kb);
                            acc[jj] = sb_mfma_terms<SPLIT>(at, bt, acc[jj]);
                        }
                    }
                }
            }
        } else {
            has_vec = true;
            // LDS scratch for the sparse entries: the B staging area is free here
            sb_spmm_vec_piece<WG, Cfg::RTMAX>(a, pc, hs, N, vm, xrow, X, vacc, reinterpret_cast<float *>(smem),
                                              reinterpret_cast<int *>(smem) + SB_SPMM_SPCHUNK);
        }
    }

    // D fragment of n-tile j: lane holds rows 16*mi + 4*(lane/16) + i, column 16*j + lane%16
    const int rbase = mi * 16 + 4 * (lane >> 4);
    if (has_mfma && !has_vec) {
        if (mactive) {
#pragma unroll
            for (int jj = 0; jj < Cfg::NTMAX; ++jj) {
                if (jj < ntw) {
                    const int col = (j0 + jj) * 16 + (lane & 15);
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        const int r = rbase + i;
                        if (r < hs) Y[(size_t)(tile.row0 + r) * N + col] = acc[jj][i];
                    }
                }
            }
        }
    } else {
        if (has_mfma) {
            __syncthreads();
            if (mactive) {
#pragma unroll
                for (int jj = 0; jj < Cfg::NTMAX; ++jj) {
                    if (jj < ntw) {
                        const int col = (j0 + jj) * 16 + (lane & 15);
#pragma unroll
                        for (int i = 0; i < 4; ++i) {
                            const int r = rbase + i;
                            if (r < hs) ytile[r * N + col] = acc[jj][i];
                        }
                    }
                }
            }
            __syncthreads();
        }
        if (vm.active) {
#pragma unroll
            for (int i = 0; i < Cfg::RTMAX; ++i) {
                const int r = vm.rbeg + i;
                if (i < vm.rt && r < hs) {
                    float4 o = vacc[i];
                    if (has_mfma) {
                        const float4 m4 = *reinterpret_cast<const float4 *>(ytile + r * N + 4 * vm.v);
                        o.x += m4.x;
                        o.y += m4.y;
                        o.z += m4.z;
                        o.w += m4.w;
                    }
                    *reinterpret_cast<float4 *>(Y + (size_t)(tile.row0 + r) * N + 4 * vm.v) = o;
                }
            }
        }
    }
    if (a.tile_clock) {
        __syncthreads();
        sb_clock(a.tile_clock, t, 1);
    }
}

template <int WG, int NRHS, int XCAP, int SPLIT>
__device__ __forceinline__ void sb_spmm_mfma_body(const SbSpmmArgs &a)
{
    using Cfg = SbMfmaCfg<WG, NRHS, SPLIT>;
    __shared__ int xrow[XCAP + 4];
    __shared__ __attribute__((aligned(16))) unsigned char smem[Cfg::SMEM_BYTES];
    __shared__ int next_tile;
    sb_for_tiles(a.ntiles, a.mode, a.queue, &next_tile,
                 [&](int t) { sb_spmm_mfma_tile<WG, NRHS, SPLIT>(a, t, xrow, smem); });
}
