// SpMM with dense blocks on MFMA: bf16 hi+lo split, FP32 accumulation.
//
// Every value a is split as a = hi + lo with hi = bf16(a), lo = bf16(a - hi)
// (round to nearest even), likewise every X element, and the product is
// formed from three v_mfma_f32_16x16x32_bf16 passes:
//     A X ~= Ahi Xhi + Ahi Xlo + Alo Xhi        (Alo Xlo dropped)
// Each bf16 product is exact in FP32; the representation error of hi+lo is
// at most 2^-18 relative, so the result has FP32-level accuracy.
//
// A (the matrix) is pre-split by the plan builder into MFMA register order:
// for 16-row tile m and 32-column k-step s, lane l holds A[16m + l%16]
// [32s + 8(l/16) + j], j = 0..7, as 8 consecutive bf16 (one 16-byte load).
// X is split while it is staged: for each k-step the 32 X rows of the piece
// are read (coalesced along N), converted, and written to LDS as [n][k] with
// a 40-element pitch so that the B-operand reads (lane l: n = l%16,
// k = 8(l/16) .. +7) are conflict-free 16-byte LDS reads.
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

__device__ __forceinline__ uint32_t sb_bf16_bits(float f)
{
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x007fffffu)) return (u >> 16) | 0x40u;
    u += 0x7fffu + ((u >> 16) & 1u);
    return u >> 16;
}

__device__ __forceinline__ float sb_bf16_float(uint32_t h) { return __uint_as_float(h << 16); }

__device__ __forceinline__ sb_bf16x8 sb_as_bf16x8(sb_u4 v) { return __builtin_bit_cast(sb_bf16x8, v); }

// Stages k-step ks of a bf16 piece: Bhi/Blo[n * SB_KP + k] for k < 32.
template <int WG>
__device__ __forceinline__ void sb_stage_b(const float *__restrict__ X, int N, const int *xrow, int ncols,
                                           int ks, uint16_t *Bhi, uint16_t *Blo)
{
    const int units = N * 4;
    for (int u = threadIdx.x; u < units; u += WG) {
        const int kg = u / N;
        const int n = u - kg * N;
        const int k0 = ks * 32 + kg * 8;
        uint32_t hw[4], lw[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int ka = k0 + 2 * j, kb = ka + 1;
            const float fa = ka < ncols ? X[(size_t)xrow[ka] * N + n] : 0.f;
            const float fb = kb < ncols ? X[(size_t)xrow[kb] * N + n] : 0.f;
            const uint32_t ha = sb_bf16_bits(fa), hb = sb_bf16_bits(fb);
            const uint32_t la = sb_bf16_bits(fa - sb_bf16_float(ha));
            const uint32_t lb = sb_bf16_bits(fb - sb_bf16_float(hb));
            hw[j] = ha | (hb << 16);
            lw[j] = la | (lb << 16);
        }
        *reinterpret_cast<sb_u4 *>(Bhi + n * SB_KP + kg * 8) = sb_u4{hw[0], hw[1], hw[2], hw[3]};
        *reinterpret_cast<sb_u4 *>(Blo + n * SB_KP + kg * 8) = sb_u4{lw[0], lw[1], lw[2], lw[3]};
    }
}

template <int WG, int NRHS>
struct SbMfmaCfg {
    static constexpr int NMAX = NRHS ? NRHS : 256;
    static constexpr int NTMAX = NMAX / 16;
    // vector-path rows per thread for a 64-row tile: ceil(64 * (N/4) / WG)
    static constexpr int RT0 = (SB_MFMA_MAXROWS * (NMAX / 4) + WG - 1) / WG;
    static constexpr int RTMAX = RT0 < 1 ? 1 : RT0;
    // LDS: max(B staging hi+lo, Y tile)
    static constexpr int BSTAGE_BYTES = 2 * NMAX * SB_KP * 2;
    static constexpr int YTILE_BYTES = SB_MFMA_MAXROWS * NMAX * 4;
    static constexpr int SMEM_BYTES = BSTAGE_BYTES > YTILE_BYTES ? BSTAGE_BYTES : YTILE_BYTES;
};

template <int WG, int NRHS>
__device__ __forceinline__ void sb_spmm_mfma_tile(const SbSpmmArgs &a, int t, int *xrow, unsigned char *smem)
{
    using Cfg = SbMfmaCfg<WG, NRHS>;
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
    uint16_t *Bhi = reinterpret_cast<uint16_t *>(smem);
    uint16_t *Blo = Bhi + Cfg::NMAX * SB_KP;
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
        if (pc.kind == SB_PIECE_DENSE_BF16) {
            has_mfma = true;
            const sb_u4 *__restrict__ Ahi = reinterpret_cast<const sb_u4 *>(a.dhi + pc.data_off);
            const sb_u4 *__restrict__ Alo = reinterpret_cast<const sb_u4 *>(a.dlo + pc.data_off);
            for (int ks = 0; ks < pc.ksteps; ++ks) {
                sb_u4 ah = {0u, 0u, 0u, 0u}, al = {0u, 0u, 0u, 0u};
                if (mactive) {
                    const int off = (mi * pc.ksteps + ks) * 64 + lane;
                    ah = __builtin_nontemporal_load(&Ahi[off]);
                    al = __builtin_nontemporal_load(&Alo[off]);
                }
                sb_stage_b<WG>(X, N, xrow, pc.ncols, ks, Bhi, Blo);
                __syncthreads();
                if (mactive) {
                    const sb_bf16x8 ahv = sb_as_bf16x8(ah), alv = sb_as_bf16x8(al);
                    const int nb = lane & 15, kb = (lane >> 4) * 8;
#pragma unroll
                    for (int jj = 0; jj < Cfg::NTMAX; ++jj) {
                        if (jj < ntw) {
                            const int n = (j0 + jj) * 16 + nb;
                            const sb_bf16x8 bh = sb_as_bf16x8(*reinterpret_cast<const sb_u4 *>(Bhi + n * SB_KP + kb));
                            const sb_bf16x8 bl = sb_as_bf16x8(*reinterpret_cast<const sb_u4 *>(Blo + n * SB_KP + kb));
                            acc[jj] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ahv, bh, acc[jj], 0, 0, 0);
                            acc[jj] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ahv, bl, acc[jj], 0, 0, 0);
                            acc[jj] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(alv, bh, acc[jj], 0, 0, 0);
                        }
                    }
                }
                __syncthreads();
            }
        } else {
            has_vec = true;
            if (vm.active) sb_spmm_vec_piece<Cfg::RTMAX>(a, pc, hs, N, vm, xrow, X, vacc);
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

template <int WG, int NRHS, int XCAP>
__device__ __forceinline__ void sb_spmm_mfma_body(const SbSpmmArgs &a)
{
    using Cfg = SbMfmaCfg<WG, NRHS>;
    __shared__ int xrow[XCAP + 4];
    __shared__ __attribute__((aligned(16))) unsigned char smem[Cfg::SMEM_BYTES];
    __shared__ int next_tile;
    sb_for_tiles(a.ntiles, a.mode, a.queue, &next_tile,
                 [&](int t) { sb_spmm_mfma_tile<WG, NRHS>(a, t, xrow, smem); });
}
