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
// [32s + 8(l/16) + j], j = 0..7.  The fragment's 512 floats are j = 0..3 of
// lanes 0..63, then j = 4..7 of lanes 0..63, so the lane's two 16-byte loads
// each belong to one contiguous 1 KiB wavefront read.  Storing FP32 keeps the
// A stream at 4 bytes per value for either split.  X is staged per k-step: the 32 X rows of the piece are read
// (coalesced along N), split, and written to LDS as [n][k] term arrays with a
// 40-element pitch so the B-operand reads (lane l: n = l%16, k = 8(l/16) ..
// +7) are conflict-free 16-byte LDS reads.
//
// MFMA tiles hold only MFMA pieces and write their Y rows; the block row's
// FP32-dense and sparse pieces run afterwards, in a second launch of the
// FP32 kernel whose tiles add to those rows.
#pragma once

#include "sable_kernels.h"

typedef __attribute__((ext_vector_type(8))) __bf16 sb_bf16x8;
typedef __attribute__((ext_vector_type(4))) float sb_f32x4;

#define SB_KP 40            // LDS pitch of a staged X column (bf16 elements)

#ifndef SB_SPLIT
#define SB_SPLIT 3
#endif

// k-steps of A loads in flight ahead of the MFMAs (A streams from HBM)
#ifndef SB_MFMA_DEPTH
#define SB_MFMA_DEPTH 2     // 1 to 4
#endif
// k-steps of X loads in flight ahead of their staging (X rows are shared by
// many tiles and mostly come from L2): divides SB_MFMA_DEPTH
#ifndef SB_MFMA_XDEPTH
#define SB_MFMA_XDEPTH 2
#endif
// the compiler's scheduling strategy for a k-step's basic block
// (__builtin_amdgcn_iglp_opt), -1: none.  Strategy 3 ran the 50k SpMM at
// N = 128 6 % faster than none (116-117 vs 124 us, bit-identical) and the
// 256-matrix batch at N = 16 the same; with four staged buffers, 114-115
// against 120-121 us without it
#ifndef SB_MFMA_IGLP
#define SB_MFMA_IGLP 3
#endif
// 1: the issue order of a k-step is set explicitly (MFMAs interleaved with
// the splits' VALU work); 0: the compiler's schedule
#ifndef SB_MFMA_INTERLEAVE
#define SB_MFMA_INTERLEAVE 0
#endif
// 16-row units to wavefronts: 0 round robin (unit u to wavefront u % NW),
// 1 packed (UNITS consecutive units per wavefront: fewer wavefronts read
// the staged X fragments, each for more units)
#ifndef SB_MFMA_PACK
#define SB_MFMA_PACK 0
#endif
// 1: four staged k-step buffers and one barrier per two k-steps (where
// they fit; SB_MFMA_DEPTH even)
#ifndef SB_MFMA_PAIRS
#define SB_MFMA_PAIRS 1
#endif
// 1: one 16-row unit per wavefront and pass (fewer registers; a taller
// tile takes more passes over its pieces)
#ifndef SB_MFMA_UNITS1
#define SB_MFMA_UNITS1 0
#endif
// 1: a scheduling fence after each column pair's MFMAs
#ifndef SB_MFMA_FENCE
#define SB_MFMA_FENCE 0
#endif
// waves per SIMD the MFMA kernels are compiled for: two (256 registers per
// lane, accumulators included), so two 256-thread workgroups share a CU
#ifndef SB_MFMA_WAVES
#define SB_MFMA_WAVES 2
#endif

__device__ __forceinline__ sb_bf16x8 sb_as_bf16x8(sb_u4 v) { return __builtin_bit_cast(sb_bf16x8, v); }

// Splits 8 floats into SPLIT bf16x8 terms: t[0] = bf16(f) by the hardware
// conversion (round to nearest even, one v_cvt_pk_bf16_f32 per pair), then
// each term of the exact FP32 residual.  The bf16 pair p converts back to
// FP32 without rounding: element 0 is p << 16, element 1 is p & 0xffff0000.
typedef __attribute__((ext_vector_type(2))) __bf16 sb_bf16x2;

template <int SPLIT>
__device__ __forceinline__ void sb_split8(const float *f, sb_bf16x8 *t)
{
    float r[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) r[j] = f[j];
#pragma unroll
    for (int s = 0; s < SPLIT; ++s) {
        sb_u4 w;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const sb_bf16x2 h = {(__bf16)r[2 * q], (__bf16)r[2 * q + 1]};
            uint32_t p = __builtin_bit_cast(uint32_t, h);
            asm("" : "+v"(p));   // empty: keeps the compiler from re-deriving p << 16 by a second conversion
            w[q] = p;
            if (s + 1 < SPLIT) {
                r[2 * q] -= __builtin_bit_cast(float, p << 16);
                r[2 * q + 1] -= __builtin_bit_cast(float, p & 0xffff0000u);
            }
        }
        t[s] = sb_as_bf16x8(w);
    }
}

// X staging of an MFMA piece, split in two phases so the global loads of a
// later k-step are in flight while the current one computes: sb_fetch_b
// loads the thread's X values of a k-step into registers, sb_store_b zeroes
// the rows past the piece, splits the values and writes B[s][n * SB_KP + k]
// (k < 32, s < SPLIT) to LDS.  Unit u of the k-step is (k group kg = u / N
// of 8 rows, column n = u % N); its 8 X rows come from xrow in two 16-byte
// LDS reads, and every load is unconditional (xrow holds row 0 for the
// columns past the piece, up to the next multiple of 32), so there is no
// divergent control flow, no bounds test in the address and no use of a
// loaded value before sb_store_b.  X is read through a global-address
// pointer with 32-bit offsets.
typedef const __attribute__((address_space(1))) float *sb_gfloat;

template <int WG, int UPT, int NRHS>
__device__ __forceinline__ void sb_fetch_b(const float *X, int N, const int *xrow, int ncols, int ks,
                                           float (*f)[8])
{
    (void)ncols;   // xrow holds row 0 for the columns past the piece (sb_expand_xrow_mfma)
    const sb_gfloat gX = (sb_gfloat)X;
    constexpr bool EXACT = NRHS != 0 && UPT * WG == NRHS * 4;   // every unit in range
    const int units = N * 4;
#pragma unroll
    for (int uu = 0; uu < UPT; ++uu) {
        const int u = threadIdx.x + uu * WG;
        const int uc = EXACT || u < units ? u : 0;
        const int kg = uc / N;
        const int n = uc - kg * N;
        const int k0 = ks * 32 + kg * 8;
        const int4 ra = *reinterpret_cast<const int4 *>(xrow + k0);
        const int4 rb = *reinterpret_cast<const int4 *>(xrow + k0 + 4);
        const int rows[8] = {ra.x, ra.y, ra.z, ra.w, rb.x, rb.y, rb.z, rb.w};
#pragma unroll
        for (int j = 0; j < 8; ++j) f[uu][j] = gX[(uint32_t)rows[j] * (uint32_t)N + (uint32_t)n];
    }
}

// The X row of every staged column of an MFMA piece, then row 0 for the
// columns up to the next multiple of 32 (the last k-step's fetches read it;
// sb_store_b zeroes those values).
__device__ __forceinline__ void sb_expand_xrow_mfma(const SbXRange *__restrict__ xr, int r0, int r1, int ncols,
                                                    int *xrow)
{
    sb_expand_xrow(xr, r0, r1, xrow);
    const int pad = ((ncols + 31) & ~31) - ncols;
    for (int i = threadIdx.x; i < pad; i += blockDim.x) xrow[ncols + i] = 0;
}

template <int WG, int UPT, int SPLIT, int NRHS, bool MASK = true>
__device__ __forceinline__ void sb_store_b(int N, int ncols, int ks, float (*f)[8], uint16_t *B, int term_stride)
{
    constexpr bool EXACT = NRHS != 0 && UPT * WG == NRHS * 4;
    const int units = N * 4;
#pragma unroll
    for (int uu = 0; uu < UPT; ++uu) {
        const int u = threadIdx.x + uu * WG;
        if (!EXACT && u >= units) continue;
        const int kg = u / N;
        const int n = u - kg * N;
        const int k0 = ks * 32 + kg * 8;
        float v[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) v[j] = !MASK || k0 + j < ncols ? f[uu][j] : 0.f;
        sb_bf16x8 t[SPLIT];
        sb_split8<SPLIT>(v, t);
#pragma unroll
        for (int s = 0; s < SPLIT; ++s)
            *reinterpret_cast<sb_bf16x8 *>(B + s * term_stride + n * SB_KP + kg * 8) = t[s];
    }
}

// The product passes over the term pairs (a_i, x_j) with i + j < SPLIT,
// smallest terms first: SPLIT = 3 -> a3x1 a2x2 a1x3 a2x1 a1x2 a1x1,
// SPLIT = 2 -> a2x1 a1x2 a1x1.
template <int V>
struct SbInt {
    static constexpr int value = V;
};

template <int SPLIT>
struct SbPasses {
    static constexpr int N = SPLIT == 3 ? 6 : 3;
    static constexpr int a(int p) { return SPLIT == 3 ? (p == 0 ? 2 : p == 1 ? 1 : p == 2 ? 0 : p == 3 ? 1 : 0)
                                                      : (p == 0 ? 1 : 0); }
    static constexpr int b(int p) { return SPLIT == 3 ? (p == 0 ? 0 : p == 1 ? 1 : p == 2 ? 2 : p == 3 ? 0 : p == 4 ? 1 : 0)
                                                      : (p == 1 ? 1 : 0); }
};

template <int WG, int NRHS, int SPLIT>
struct SbMfmaCfg {
    static constexpr int NMAX = NRHS ? NRHS : 256;
    static constexpr int NTMAX = NMAX / 16;
    static constexpr int NW = WG / 64;
    static constexpr int UNITS = SB_MFMA_UNITS1 ? 1 : SB_MFMA_UNITS(NMAX);   // row tiles per wavefront and pass
    static constexpr int TERM_STRIDE = NMAX * SB_KP;             // bf16 elements per staged term
    static constexpr int UPT = (NMAX * 4 + WG - 1) / WG;         // X staging units per thread
    static constexpr int BUF_STRIDE = SPLIT * TERM_STRIDE;        // one staged k-step (bf16 elements)
    // staged k-step buffers: 4 (one barrier per two k-steps) when
    // SB_MFMA_PAIRS asks for it (with A and X prefetch depths of at least
    // two), a CU holds one workgroup by registers
    // (SB_MFMA_WAVES waves per SIMD), N is at least 64 and the 4 buffers fit
    // its LDS (16.5 KiB allowed for xrow); else 2.  At 512 threads and N = 128 they took the
    // every-block-on-MFMA plan of the 50k SpMM from 120-122 to 116-118 us
    // (118-119 to 114-115 us in a run timing two copies of each);
    // with 256 threads (two workgroups per CU, which 4 buffers would cut to
    // one) they were 30 % slower, and on the 256-matrix batch's N = 16
    // members (128 threads, tiles of one or two k-steps) 4.7 % slower.
    static constexpr int WGS_PER_CU = (4 * SB_MFMA_WAVES) / NW > 0 ? (4 * SB_MFMA_WAVES) / NW : 1;
    static constexpr int NBUF =
        SB_MFMA_PAIRS && SB_MFMA_DEPTH % 2 == 0 && SB_MFMA_XDEPTH >= 2 && WGS_PER_CU == 1 && NMAX >= 64 &&
                4 * BUF_STRIDE * 2 + 16896 <= 163840
            ? 4
            : 2;
    static constexpr int SMEM_BYTES = NBUF * BUF_STRIDE * 2;
};

// Fixup of the MFMA row tiles split along k (a launch after the MFMA
// launch): the workgroups (t, g), g < gridDim.y, take fix tile t, whose
// ksplit chunks stored their partial sums as (16-row unit, column tile)
// 1 KiB blocks in accumulator order; a wavefront takes every
// (gridDim.y * WG / 64)-th block, loads its chunks four at a time and adds
// them in chunk order 0, 1, ..., ksplit-1 (the same order whatever order
// the chunks ran in), then writes the block's rows.
template <int WG>
__device__ __forceinline__ void sb_mfma_fixup_body(const SbTile *fix_tiles, int nfix, const SbMatBufs *mats,
                                                   const float *part0, float *y0, int N)
{
    const int t = blockIdx.x;
    if (t >= nfix) return;
    const SbTile tile = fix_tiles[t];
    const SB_G float *part = sb_gp(mats ? mats[tile.mat].part : part0);
    SB_G float *Y = sb_gp(mats ? mats[tile.mat].y : y0);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const int nt = N >> 4;
    const int nblk = ((tile.nrows + 15) >> 4) * nt;
    const int nwav = gridDim.y * (WG / 64);
    const int ks = tile.ksplit;
    const SB_G sb_f32x4 *__restrict__ P = reinterpret_cast<const SB_G sb_f32x4 *>(part) + (int64_t)tile.part_blk * 64;
    for (int b = blockIdx.y * (WG / 64) + wave; b < nblk; b += nwav) {
        sb_f32x4 s = P[(int64_t)b * 64 + lane];
        for (int c = 1; c < ks; c += 4) {
            sb_f32x4 v[4];
#pragma unroll
            for (int i = 0; i < 4; ++i)
                if (c + i < ks) v[i] = P[((int64_t)(c + i) * nblk + b) * 64 + lane];
#pragma unroll
            for (int i = 0; i < 4; ++i)
                if (c + i < ks) s += v[i];
        }
        const int m = b / nt, j = b - m * nt;
        const int rbase = m * 16 + 4 * (lane >> 4);
        const int col = j * 16 + (lane & 15);
#pragma unroll
        for (int r = 0; r < 4; ++r)
            if (rbase + r < tile.nrows) Y[(size_t)(tile.row0 + rbase + r) * N + col] = s[r];
    }
}

// Workgroups per fix tile of the fixup launch: enough 4-wavefront
// workgroups for one block per wavefront of the tallest fix tile (at most
// 64).
inline int sb_fixup_groups(const SbTile *fix, size_t nfix, int n_rhs)
{
    int most = 0;
    for (size_t i = 0; i < nfix; ++i) {
        const int nblk = ((fix[i].nrows + 15) >> 4) * (n_rhs >> 4);
        most = nblk > most ? nblk : most;
    }
    const int g = (most + 3) / 4;
    return g < 1 ? 1 : g > 64 ? 64 : g;
}

// The k-loop of one MFMA piece for a wavefront holding NU units (16-row
// tiles umi[0..NU) across all N/16 column tiles; NU = 0: the wavefront only
// helps stage X).  Every wavefront runs the same k-steps and barriers
// whatever its NU, so the workgroup's waves may run different instances.
// A k-step is one basic block: the MFMAs of k-step ks (A terms `at`, B
// fragments read from LDS one column pair ahead), then -- independent of
// them, so their issue can interleave -- the reload of the A slot, the
// split of k-step ks+1's A values into the next `at`, the staging of X for
// k-step ks+1 into the free LDS buffer and the X loads of k-step ks+1+D; one
// barrier ends the step.  A is loaded D k-steps ahead, X DX k-steps ahead of
// its staging (SB_MFMA_DEPTH, SB_MFMA_XDEPTH).  Every load is unconditional (k-steps past the piece clamp to
// its last one), so the compiler's waits stay exact.
template <int WG, int NRHS, int SPLIT, int NU>
__device__ __forceinline__ void sb_mfma_kloop(const SbSpmmArgs &a, const SbMatBufs &mb, const SbPiece &pc,
                                              const int *umi, int N, int ntw, const int *xrow, uint16_t *B,
                                              sb_f32x4 (*acc)[SbMfmaCfg<WG, NRHS, SPLIT>::NTMAX])
{
    using Cfg = SbMfmaCfg<WG, NRHS, SPLIT>;
    constexpr int D = SB_MFMA_DEPTH;
    constexpr int DX = SB_MFMA_XDEPTH;
    static_assert(D % DX == 0, "the X depth divides the A depth");
    constexpr int NUA = NU > 0 ? NU : 1;
    constexpr int JB = Cfg::NTMAX >= 2 ? 2 : 1;
    constexpr int NP = (Cfg::NTMAX + JB - 1) / JB;             // column pairs
    const int lane = threadIdx.x & 63;
    const int nb = lane & 15, kb = (lane >> 4) * 8;
    const float *__restrict__ X = mb.x;
    const int K = pc.ksteps;
    // a global-address-space pointer: loads through a generic pointer would
    // be FLAT loads, which count against the LDS counter as well, so every
    // wait for a B fragment would also wait for the A loads in flight
    const SB_G sb_f4 *__restrict__ A = reinterpret_cast<const SB_G sb_f4 *>(sb_gp(mb.dval) + pc.data_off);
    float xf[DX][Cfg::UPT][8];
    sb_f4 ar[D][NUA][2];
    sb_bf16x8 at[NUA][SPLIT];
    auto load_a = [&](sb_f4 (*dst)[2], int k) {
#pragma unroll
        for (int i = 0; i < NU; ++i) {
            const int off = (umi[i] * K + k) * 128 + lane;
            dst[i][0] = __builtin_nontemporal_load(&A[off]);
            dst[i][1] = __builtin_nontemporal_load(&A[off + 64]);
        }
    };
    auto split_a = [&](sb_f4 (*src)[2]) {
#pragma unroll
        for (int i = 0; i < NU; ++i) {
            const float af[8] = {src[i][0].x, src[i][0].y, src[i][0].z, src[i][0].w,
                                 src[i][1].x, src[i][1].y, src[i][1].z, src[i][1].w};
            sb_split8<SPLIT>(af, at[i]);
        }
    };
    // B fragments of column pair jp (a second column past ntw repeats the first)
    auto read_pair = [&](const uint16_t *Bc, int jp, sb_bf16x8 (*bt)[SPLIT]) {
#pragma unroll
        for (int c = 0; c < JB; ++c) {
            const int j = jp * JB + c;
            const int n = ((NRHS || j < ntw) && j < Cfg::NTMAX ? j : jp * JB) * 16 + nb;
#pragma unroll
            for (int s = 0; s < SPLIT; ++s)
                bt[c][s] = *reinterpret_cast<const sb_bf16x8 *>(Bc + s * Cfg::TERM_STRIDE + n * SB_KP + kb);
        }
    };
#pragma unroll
    for (int sl = 0; sl < D; ++sl) load_a(ar[sl], sl < K ? sl : K - 1);
    // staged k-steps ahead of the one multiplied: 1 with two buffers, 2 with
    // four (k-step s in buffer s % NBUF; the barrier after every
    // (NBUF / 2)-th step orders each buffer's writes after its last reads)
    constexpr int SA = Cfg::NBUF / 2;
    static_assert(SA <= DX, "the X slots cover the staged k-steps");
#pragma unroll
    for (int j = 0; j < SA; ++j) {
        sb_fetch_b<WG, Cfg::UPT, NRHS>(X, N, xrow, pc.ncols, j < K ? j : K - 1, xf[j % DX]);
        sb_store_b<WG, Cfg::UPT, SPLIT, NRHS>(N, pc.ncols, j, xf[j % DX], B + j * Cfg::BUF_STRIDE, Cfg::TERM_STRIDE);
    }
#pragma unroll
    for (int j = SA; j < SA + DX; ++j)
        sb_fetch_b<WG, Cfg::UPT, NRHS>(X, N, xrow, pc.ncols, j < K ? j : K - 1, xf[j % DX]);
    split_a(ar[0]);
    load_a(ar[0], D < K ? D : K - 1);
    __syncthreads();
    // k-step ks: `at` holds its A terms; A slot sl (= ks % D) holds the raw
    // A of k-step ks + D, slot (sl + 1) % D that of k-step ks + 1
    // MASK: the step may stage the last k-step, whose columns past the piece
    // are zeroed (only the last steps of the piece, peeled off the loop)
    auto step = [&](auto slc, auto maskc, const int ks) {
        constexpr int sl = decltype(slc)::value;
        constexpr bool MASK = decltype(maskc)::value != 0;
        const uint16_t *Bc = B + (ks & (Cfg::NBUF - 1)) * Cfg::BUF_STRIDE;
        if constexpr (NU > 0) {
            using Pt = SbPasses<SPLIT>;
            sb_bf16x8 bt[2][JB][SPLIT];
            read_pair(Bc, 0, bt[0]);
#pragma unroll
            for (int jp = 0; jp < NP; ++jp) {
                if (jp + 1 < NP) read_pair(Bc, jp + 1, bt[(jp + 1) & 1]);
                if (NRHS || jp * JB < ntw) {
                    // per pass, the units and the pair's columns: consecutive
                    // MFMAs go to different accumulators
#pragma unroll
                    for (int p = 0; p < Pt::N; ++p)
#pragma unroll
                        for (int i = 0; i < NU; ++i)
#pragma unroll
                            for (int c = 0; c < JB; ++c)
                                if (jp * JB + c < Cfg::NTMAX)
                                    acc[i][jp * JB + c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                                        at[i][Pt::a(p)], bt[jp & 1][c][Pt::b(p)], acc[i][jp * JB + c], 0, 0, 0);
                }
#if SB_MFMA_FENCE
                // keeps the next pairs' B reads from being hoisted above
                // these MFMAs (register budget of the higher occupancies)
                __builtin_amdgcn_sched_barrier(0);
#endif
            }
            // the next k-step's A terms, then the slot refilled D ahead
            constexpr int nsl = (sl + 1) % D;
            split_a(ar[nsl]);
            load_a(ar[nsl], ks + 1 + D < K ? ks + 1 + D : K - 1);
#if SB_MFMA_IGLP >= 0
            __builtin_amdgcn_iglp_opt(SB_MFMA_IGLP);
#endif
        }
        // stage k-step ks + SA into its buffer (past the last k-step this
        // writes zeros nobody reads), then load the X values of k-step
        // ks + SA + DX into the slot it frees
        constexpr int xs = (sl + SA) % DX;
        sb_store_b<WG, Cfg::UPT, SPLIT, NRHS, MASK>(N, pc.ncols, ks + SA, xf[xs],
                                                    B + ((ks + SA) & (Cfg::NBUF - 1)) * Cfg::BUF_STRIDE,
                                                    Cfg::TERM_STRIDE);
        sb_fetch_b<WG, Cfg::UPT, NRHS>(X, N, xrow, pc.ncols, ks + SA + DX < K ? ks + SA + DX : K - 1, xf[xs]);
#if SB_MFMA_INTERLEAVE
        // issue order of the step: the first column pair's B reads, then
        // every MFMA followed by two VALU (the splits), a B read every
        // fourth MFMA
        if constexpr (NU > 0 && NRHS != 0) {
            constexpr int NM = NU * Cfg::NTMAX * SbPasses<SPLIT>::N;
            __builtin_amdgcn_sched_group_barrier(0x100, JB * SPLIT, 0);
#pragma unroll
            for (int i = 0; i < NM; ++i) {
                __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
                __builtin_amdgcn_sched_group_barrier(0x002, 2, 0);
                if (i % 4 == 3) __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);
            }
        }
#endif
        if (SA == 1 || (ks & 1)) __syncthreads();
    };
    // main loop: the steps that stage k-steps below K - 1 (every column in
    // the piece); then the last (at most D + SA) steps with the mask
    int ks = 0;
    for (; ks + D + SA < K; ks += D) {
        step(SbInt<0>{}, SbInt<0>{}, ks);
        if constexpr (D > 1) step(SbInt<(D > 1 ? 1 : 0)>{}, SbInt<0>{}, ks + 1);
        if constexpr (D > 2) step(SbInt<(D > 2 ? 2 : 0)>{}, SbInt<0>{}, ks + 2);
        if constexpr (D > 3) step(SbInt<(D > 3 ? 3 : 0)>{}, SbInt<0>{}, ks + 3);
    }
    if (ks < K) step(SbInt<0>{}, SbInt<1>{}, ks);
    if constexpr (D > 1) if (ks + 1 < K) step(SbInt<(D > 1 ? 1 : 0)>{}, SbInt<1>{}, ks + 1);
    if constexpr (D > 2) if (ks + 2 < K) step(SbInt<(D > 2 ? 2 : 0)>{}, SbInt<1>{}, ks + 2);
    if constexpr (D > 3) if (ks + 3 < K) step(SbInt<(D > 3 ? 3 : 0)>{}, SbInt<1>{}, ks + 3);
    if (ks + D < K) step(SbInt<0>{}, SbInt<1>{}, ks + D);
    if constexpr (SA > 1) if (ks + D + 1 < K) step(SbInt<(D > 1 ? 1 : 0)>{}, SbInt<1>{}, ks + D + 1);
}

// One pass of a wavefront holding NU units umi[0..NU) over every piece of
// the work item, with its accumulators, then its D fragments stored: column
// tile j of unit u, lane l holds rows 16u + 4(l/16) + r, column 16j + l%16;
// a chunk of a split tile stores the fragment as its partial-sum block
// (unit, column tile) in that same lane order instead.
template <int WG, int NRHS, int SPLIT, int NU>
__device__ __forceinline__ void sb_mfma_pass(const SbSpmmArgs &a, const SbMatBufs &mb, const SbTile &tile,
                                             const int *umi, int N, int nt, int mt, int *xrow, uint16_t *B)
{
    using Cfg = SbMfmaCfg<WG, NRHS, SPLIT>;
    constexpr int NUA = NU > 0 ? NU : 1;
    const int lane = threadIdx.x & 63;
    sb_f32x4 acc[NUA][Cfg::NTMAX];
#pragma unroll
    for (int u = 0; u < NUA; ++u)
#pragma unroll
        for (int j = 0; j < Cfg::NTMAX; ++j) acc[u][j] = sb_f32x4{0.f, 0.f, 0.f, 0.f};
    for (int pi = tile.piece_begin; pi < tile.piece_end; ++pi) {
        const SbPiece pc = a.pieces[pi];
        if (pc.kind != SB_PIECE_MFMA) continue;          // never built into an MFMA tile
        __syncthreads();                               // previous piece done with xrow and B
        sb_expand_xrow_mfma(a.xr, pc.xr_begin, pc.xr_end, pc.ncols, xrow);
        __syncthreads();
        sb_mfma_kloop<WG, NRHS, SPLIT, NU>(a, mb, pc, umi, N, nt, xrow, B, acc);
    }
    const bool split = tile.ksplit > 1;
#pragma unroll
    for (int i = 0; i < NU; ++i) {
        const int rbase = umi[i] * 16 + 4 * (lane >> 4);
#pragma unroll
        for (int jj = 0; jj < Cfg::NTMAX; ++jj) {
            if (NRHS || jj < nt) {
                if (split) {
                    const int64_t blk = (int64_t)tile.part_blk + (int64_t)(tile.kpart * mt + umi[i]) * nt + jj;
                    reinterpret_cast<SB_G sb_f32x4 *>(sb_gp(mb.part))[blk * 64 + lane] = acc[i][jj];
                } else {
                    const int col = jj * 16 + (lane & 15);
#pragma unroll
                    for (int r = 0; r < 4; ++r)
                        if (rbase + r < tile.nrows)
                            sb_gp(mb.y)[(size_t)(tile.row0 + rbase + r) * N + col] = acc[i][jj][r];
                }
            }
        }
    }
}

// One MFMA work item (only MFMA pieces): a row tile, or one k chunk of a
// split row tile.  Its 16-row units go round robin to the wavefronts, UNITS
// per wavefront and pass (one pass for tiles the host cut to
// SB_MFMA_MAX_ROWS; more only if a generic kernel meets a taller tile); each
// unit covers all N/16 column tiles.  Per 32-column k-step the workgroup
// stages the X rows once (split into bf16 terms in one of two LDS buffers)
// and every unit multiplies its A fragment against them (sb_mfma_pass in
// the instance for the wavefront's unit count; every instance runs the same
// barriers).
template <int WG, int NRHS, int SPLIT>
__device__ __forceinline__ void sb_spmm_mfma_tile(const SbSpmmArgs &a, int t, int *xrow, unsigned char *smem)
{
    using Cfg = SbMfmaCfg<WG, NRHS, SPLIT>;
    constexpr int NW = Cfg::NW;
    constexpr int UN = Cfg::UNITS;
    const int N = NRHS ? NRHS : a.n_rhs;
    const int nt = N >> 4;
    const SbTile tile = a.tiles[t];
    const int mt = (tile.nrows + 15) >> 4;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int mine = mt > wave ? (mt - wave + NW - 1) / NW : 0;   // units of this wavefront (round robin)
    const SbMatBufs mb = sb_mat_bufs(a, tile.mat, nullptr);
    uint16_t *B = reinterpret_cast<uint16_t *>(smem);
    // every wavefront runs the same number of passes (the staging barriers)
    int all_passes = (((mt + NW - 1) / NW) + UN - 1) / UN;
    if (all_passes < 1) all_passes = 1;
    for (int pass = 0; pass < all_passes; ++pass) {
        // this pass's units: index i -> unit wave + NW * (pass * UN + i)
        // (round robin) or NW * UN * pass + UN * wave + i (packed); the
        // active ones are a prefix of nact
        int umi[UN];
        int nact = 0;
#pragma unroll
        for (int i = 0; i < UN; ++i) {
            const int k = pass * UN + i;
            const int u = SB_MFMA_PACK ? (pass * NW + wave) * UN + i : wave + NW * k;
            const bool act = SB_MFMA_PACK ? u < mt : k < mine;
            umi[i] = act ? u : 0;
            nact += act;
        }
        if (nact >= UN)
            sb_mfma_pass<WG, NRHS, SPLIT, UN>(a, mb, tile, umi, N, nt, mt, xrow, B);
        else if (UN > 1 && nact == 1)
            sb_mfma_pass<WG, NRHS, SPLIT, 1>(a, mb, tile, umi, N, nt, mt, xrow, B);
        else
            sb_mfma_pass<WG, NRHS, SPLIT, 0>(a, mb, tile, umi, N, nt, mt, xrow, B);
    }
}

template <int WG, int NRHS, int XCAP, int SPLIT>
__device__ __forceinline__ void sb_spmm_mfma_body(const SbSpmmArgs &a)
{
    using Cfg = SbMfmaCfg<WG, NRHS, SPLIT>;
    __shared__ __attribute__((aligned(16))) int xrow[XCAP + 32];   // k-steps read up to 31 past ncols
    __shared__ __attribute__((aligned(16))) unsigned char smem[Cfg::SMEM_BYTES];
    __shared__ int next_tile;
    sb_for_tiles(a.ntiles, a.mode, a.queue, &next_tile,
                 [&](int t) { sb_spmm_mfma_tile<WG, NRHS, SPLIT>(a, t, xrow, smem); });
}
