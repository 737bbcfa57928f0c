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

// k-steps of A and X loads in flight ahead of the MFMAs
#ifndef SB_MFMA_DEPTH
#define SB_MFMA_DEPTH 2     // 1 to 4
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
// LDS reads, and every load is unconditional (rows past the piece read
// column n of X row 0), so there is no divergent control flow and no use of
// a loaded value before sb_store_b.  X is read through a global-address
// pointer with 32-bit offsets.
typedef const __attribute__((address_space(1))) float *sb_gfloat;

template <int WG, int UPT, int NRHS>
__device__ __forceinline__ void sb_fetch_b(const float *X, int N, const int *xrow, int ncols, int ks,
                                           float (*f)[8])
{
    const sb_gfloat gX = (sb_gfloat)X;
    constexpr bool EXACT = NRHS != 0 && UPT * WG == NRHS * 4;   // every unit in range
    const int units = N * 4;
#pragma unroll
    for (int uu = 0; uu < UPT; ++uu) {
        const int u = threadIdx.x + uu * WG;
        const bool uok = EXACT || u < units;
        const int uc = uok ? u : 0;
        const int kg = uc / N;
        const int n = uc - kg * N;
        const int k0 = ks * 32 + kg * 8;
        const int4 ra = *reinterpret_cast<const int4 *>(xrow + k0);
        const int4 rb = *reinterpret_cast<const int4 *>(xrow + k0 + 4);
        const int rows[8] = {ra.x, ra.y, ra.z, ra.w, rb.x, rb.y, rb.z, rb.w};
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const bool ok = uok && k0 + j < ncols;
            const uint32_t off = (uint32_t)(ok ? rows[j] : 0) * (uint32_t)N + (uint32_t)n;
            f[uu][j] = gX[off];
        }
    }
}

template <int WG, int UPT, int SPLIT, int NRHS>
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
        for (int j = 0; j < 8; ++j) v[j] = k0 + j < ncols ? f[uu][j] : 0.f;
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
    static constexpr int UNITS = SB_MFMA_UNITS(NMAX);            // row tiles per wavefront and pass
    static constexpr int TERM_STRIDE = NMAX * SB_KP;             // bf16 elements per staged term
    static constexpr int UPT = (NMAX * 4 + WG - 1) / WG;         // X staging units per thread
    static constexpr int BUF_STRIDE = SPLIT * TERM_STRIDE;        // one staged k-step (bf16 elements)
    static constexpr int SMEM_BYTES = 2 * BUF_STRIDE * 2;          // double-buffered
};

// The end of a chunk of a split MFMA tile (the chunks of one row tile are
// separate work items over consecutive k ranges).  Every chunk has stored
// its partial sums; it makes them visible device-wide (release fence), then
// counts itself in the tile's counter.  The chunk that arrives last (any
// chunk: the order of arrival does not matter) makes the others' stores
// visible to itself (acquire fence), adds the partials in chunk order
// 0, 1, ..., ksplit-1 -- so Y is the same whichever chunk finishes last --,
// writes the rows and resets the counter for the next launch.  No wave waits
// on another workgroup.
template <int NW>
__device__ __forceinline__ void sb_mfma_fixup(const SbTile &tile, const SbMatBufs &mb, float *Y, int N, int nt,
                                              int mt, int *fixflag)
{
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0) *fixflag = atomicAdd(&mb.fix[tile.fix], 1) == tile.ksplit - 1;
    __syncthreads();
    const bool last = *fixflag != 0;
    __syncthreads();                               // fixflag is read before the next tile reuses it
    if (!last) return;
    __threadfence();
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const int nblk = mt * nt;
    const sb_f32x4 *__restrict__ part = reinterpret_cast<const sb_f32x4 *>(mb.part) + (int64_t)tile.part_blk * 64;
    for (int b = wave; b < nblk; b += NW) {
        sb_f32x4 s = part[(int64_t)b * 64 + lane];
        for (int c = 1; c < tile.ksplit; ++c) s += part[((int64_t)c * nblk + b) * 64 + lane];
        const int m = b / nt, j = b - m * nt;
        const int rbase = m * 16 + 4 * (lane >> 4);
        const int col = j * 16 + (lane & 15);
#pragma unroll
        for (int r = 0; r < 4; ++r)
            if (rbase + r < tile.nrows) Y[(size_t)(tile.row0 + rbase + r) * N + col] = s[r];
    }
    if (threadIdx.x == 0) atomicExch(&mb.fix[tile.fix], 0);
}

// One MFMA tile (only MFMA pieces; it writes its Y rows).  Work units are
// (16-row tile mi, column group gi) pairs: ng column groups of nt/ng
// 16-column tiles, ng the smallest power of two dividing nt that gives every
// wavefront a unit when the tile is short (or the largest such).  Wavefront w takes units w, w + NW, ...,
// UNITS of them per pass (one pass for tiles the host cut to
// SB_MFMA_MAX_ROWS; more only if a generic kernel meets a taller tile).  Per
// 32-column k-step the workgroup stages the X rows once (split into bf16
// terms in one of two LDS buffers) and every unit multiplies its A fragment
// against them.  The X values of k-step s+1 are split into the other buffer
// after k-step s's MFMAs are issued (the two overlap), their global loads
// were issued a k-step earlier, and one barrier per k-step orders the
// buffers.
template <int WG, int NRHS, int SPLIT>
__device__ __forceinline__ void sb_spmm_mfma_tile(const SbSpmmArgs &a, int t, int *xrow, unsigned char *smem,
                                                  int *fixflag)
{
    using Cfg = SbMfmaCfg<WG, NRHS, SPLIT>;
    constexpr int NW = Cfg::NW;
    constexpr int UN = Cfg::UNITS;
    const int N = NRHS ? NRHS : a.n_rhs;
    const int nt = N >> 4;
    const SbTile tile = a.tiles[t];
    const int hs = tile.nrows;
    const int mt = (hs + 15) >> 4;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    // ng divides nt (every column tile is in exactly one group) and NW
    // (both powers of two, ng <= NW), so the units of one wavefront
    // (w, w + NW, ...) share the column group w % ng
    int ng = 1;
    while (ng * 2 <= nt && nt % (ng * 2) == 0 && mt * ng < NW) ng <<= 1;
    const int gj0 = (wave % ng) * (nt / ng);
    const int ntw = nt / ng;
    const int units = mt * ng;
    const int mine = units > wave ? (units - wave + NW - 1) / NW : 0;   // units of this wavefront
    const int passes = (mine + UN - 1) / UN;
    const SbMatBufs mb = sb_mat_bufs(a, tile.mat, nullptr);
    const float *__restrict__ X = mb.x;
    float *__restrict__ Y = mb.y;
    uint16_t *B = reinterpret_cast<uint16_t *>(smem);
    const int nb = lane & 15, kb = (lane >> 4) * 8;
    const bool split = tile.ksplit > 1;

    // every wavefront runs the same number of passes (the staging barriers)
    int all_passes = (((units + NW - 1) / NW) + UN - 1) / UN;
    if (all_passes < 1) all_passes = 1;
    for (int pass = 0; pass < all_passes; ++pass) {
        sb_f32x4 acc[UN][Cfg::NTMAX];
#pragma unroll
        for (int u = 0; u < UN; ++u)
#pragma unroll
            for (int j = 0; j < Cfg::NTMAX; ++j) acc[u][j] = sb_f32x4{0.f, 0.f, 0.f, 0.f};
        // this pass's units: index i -> unit wave + NW * (pass * UN + i)
        int umi[UN];
        bool uact[UN];
#pragma unroll
        for (int i = 0; i < UN; ++i) {
            const int k = pass * UN + i;
            const int u = wave + NW * k;
            uact[i] = k < mine && pass < passes;
            umi[i] = uact[i] ? u / ng : 0;
        }
        for (int pi = tile.piece_begin; pi < tile.piece_end; ++pi) {
            const SbPiece pc = a.pieces[pi];
            if (pc.kind != SB_PIECE_MFMA) continue;          // never built into an MFMA tile
            __syncthreads();                               // previous piece done with xrow and B
            sb_expand_xrow(a.xr, pc.xr_begin, pc.xr_end, xrow);
            __syncthreads();
            const int K = pc.ksteps;
            const sb_f4 *__restrict__ A = reinterpret_cast<const sb_f4 *>(mb.dval + pc.data_off);
            // pipeline, DEPTH k-steps deep: A of k-step s sits in register
            // slot s % DEPTH and X of k-step s in slot s % DEPTH, each loaded
            // DEPTH k-steps before its use (A is streamed from HBM, so one
            // k-step of compute cannot cover its latency).  B of k-step s is
            // in LDS buffer s & 1; X of k-step s + 1 is split into the free
            // buffer after k-step s's MFMAs are issued.  Every load is
            // unconditional (indices past the piece clamp to its last k-step)
            // so the waits on them stay exact; one barrier per k-step.
            constexpr int D = SB_MFMA_DEPTH;
            float xf[D][Cfg::UPT][8];
            sb_f4 ar[D][UN][2];
#pragma unroll
            for (int sl = 0; sl < D; ++sl) {
                const int kl = sl < K ? sl : K - 1;
#pragma unroll
                for (int i = 0; i < UN; ++i) {
                    const int off = (umi[i] * K + kl) * 128 + lane;
                    ar[sl][i][0] = __builtin_nontemporal_load(&A[off]);
                    ar[sl][i][1] = __builtin_nontemporal_load(&A[off + 64]);
                }
            }
            sb_fetch_b<WG, Cfg::UPT, NRHS>(X, N, xrow, pc.ncols, 0, xf[0]);
            sb_store_b<WG, Cfg::UPT, SPLIT, NRHS>(N, pc.ncols, 0, xf[0], B, Cfg::TERM_STRIDE);
#pragma unroll
            for (int j = 1; j <= D; ++j)
                sb_fetch_b<WG, Cfg::UPT, NRHS>(X, N, xrow, pc.ncols, j < K ? j : K - 1, xf[j % D]);
            __syncthreads();
            // one k-step with A in register slot sl and the X values of k-step
            // ks + 1 in slot (sl + 1) % D
            auto step = [&](auto slc, const int ks) {
                constexpr int sl = decltype(slc)::value;
                const uint16_t *Bc = B + (ks & 1) * Cfg::BUF_STRIDE;
                // A terms of every unit (the slot is then reloaded with
                // k-step ks + D), then each B fragment read once from LDS
                // and applied to all units (units of a wavefront share
                // their column group)
                sb_bf16x8 at[UN][SPLIT];
#pragma unroll
                for (int i = 0; i < UN; ++i) {
                    const float af[8] = {ar[sl][i][0].x, ar[sl][i][0].y, ar[sl][i][0].z, ar[sl][i][0].w,
                                         ar[sl][i][1].x, ar[sl][i][1].y, ar[sl][i][1].z, ar[sl][i][1].w};
                    sb_split8<SPLIT>(af, at[i]);
                }
                {
                    const int kn = ks + D < K ? ks + D : K - 1;
#pragma unroll
                    for (int i = 0; i < UN; ++i) {
                        const int off = (umi[i] * K + kn) * 128 + lane;
                        ar[sl][i][0] = __builtin_nontemporal_load(&A[off]);
                        ar[sl][i][1] = __builtin_nontemporal_load(&A[off + 64]);
                    }
                }
                // two column tiles per step: per unit, its 2 x passes MFMAs
                // alternate between the two accumulators, so no MFMA waits on
                // the one before it.  A second column past ntw reads column
                // j0's B and is never stored.  A unit's branch is uniform
                // over the wavefront (inactive units issue nothing).
                constexpr int JB = Cfg::NTMAX >= 2 ? 2 : 1;
#pragma unroll
                for (int j0 = 0; j0 < Cfg::NTMAX; j0 += JB) {
                    if (j0 < ntw) {
                        sb_bf16x8 bt[JB][SPLIT];
#pragma unroll
                        for (int c = 0; c < JB; ++c) {
                            const int n = (gj0 + (j0 + c < ntw ? j0 + c : j0)) * 16 + nb;
#pragma unroll
                            for (int s = 0; s < SPLIT; ++s)
                                bt[c][s] = *reinterpret_cast<const sb_bf16x8 *>(Bc + s * Cfg::TERM_STRIDE +
                                                                             n * SB_KP + kb);
                        }
                        using Pt = SbPasses<SPLIT>;
#pragma unroll
                        for (int i = 0; i < UN; ++i) {
                            if (!uact[i]) continue;
#pragma unroll
                            for (int p = 0; p < Pt::N; ++p)
#pragma unroll
                                for (int c = 0; c < JB; ++c)
                                    if (j0 + c < Cfg::NTMAX)   // odd NTMAX: the last pair has one column
                                        acc[i][j0 + c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                                            at[i][Pt::a(p)], bt[c][Pt::b(p)], acc[i][j0 + c], 0, 0, 0);
                        }
                    }
                }
                // stage k-step ks + 1 into the free buffer (after the last
                // k-step this writes zeros nobody reads), then load the X
                // values of k-step ks + 1 + D into the slot it frees
                constexpr int xs = (sl + 1) % D;
                sb_store_b<WG, Cfg::UPT, SPLIT, NRHS>(N, pc.ncols, ks + 1, xf[xs],
                                                      B + ((ks + 1) & 1) * Cfg::BUF_STRIDE, Cfg::TERM_STRIDE);
                sb_fetch_b<WG, Cfg::UPT, NRHS>(X, N, xrow, pc.ncols, ks + 1 + D < K ? ks + 1 + D : K - 1, xf[xs]);
                __syncthreads();
            };
            // groups of D k-steps in a loop with one back-edge and no other
            // branch around a load, then the remaining k-steps
            int ks = 0;
            for (; ks + D <= K; ks += D) {
                step(SbInt<0>{}, ks);
                if constexpr (D > 1) step(SbInt<(D > 1 ? 1 : 0)>{}, ks + 1);
                if constexpr (D > 2) step(SbInt<(D > 2 ? 2 : 0)>{}, ks + 2);
                if constexpr (D > 3) step(SbInt<(D > 3 ? 3 : 0)>{}, ks + 3);
            }
            if constexpr (D > 1) if (ks < K) step(SbInt<0>{}, ks);
            if constexpr (D > 2) if (ks + 1 < K) step(SbInt<(D > 2 ? 1 : 0)>{}, ks + 1);
            if constexpr (D > 3) if (ks + 2 < K) step(SbInt<(D > 3 ? 2 : 0)>{}, ks + 2);
        }
        // D fragment of column tile j: lane holds rows 16*mi + 4*(lane/16) + r,
        // column 16*j + lane%16.  A chunk of a split tile stores it as its
        // partial-sum block (unit, column tile) in that same order instead.
#pragma unroll
        for (int i = 0; i < UN; ++i) {
            if (!uact[i]) continue;
            const int j0 = gj0;
            const int rbase = umi[i] * 16 + 4 * (lane >> 4);
#pragma unroll
            for (int jj = 0; jj < Cfg::NTMAX; ++jj) {
                if (jj < ntw) {
                    if (split) {
                        const int64_t blk = (int64_t)tile.part_blk + (int64_t)(tile.kpart * mt + umi[i]) * nt + j0 + jj;
                        reinterpret_cast<sb_f32x4 *>(mb.part)[blk * 64 + lane] = acc[i][jj];
                    } else {
                        const int col = (j0 + jj) * 16 + nb;
#pragma unroll
                        for (int r = 0; r < 4; ++r)
                            if (rbase + r < hs) Y[(size_t)(tile.row0 + rbase + r) * N + col] = acc[i][jj][r];
                    }
                }
            }
        }
    }
    if (split) sb_mfma_fixup<NW>(tile, mb, Y, N, nt, mt, fixflag);
}

template <int WG, int NRHS, int XCAP, int SPLIT>
__device__ __forceinline__ void sb_spmm_mfma_body(const SbSpmmArgs &a)
{
    using Cfg = SbMfmaCfg<WG, NRHS, SPLIT>;
    __shared__ __attribute__((aligned(16))) int xrow[XCAP + 32];   // k-steps read up to 31 past ncols
    __shared__ __attribute__((aligned(16))) unsigned char smem[Cfg::SMEM_BYTES];
    __shared__ int next_tile;
    __shared__ int fixflag;
    sb_for_tiles(a.ntiles, a.mode, a.queue, &next_tile,
                 [&](int t) { sb_spmm_mfma_tile<WG, NRHS, SPLIT>(a, t, xrow, smem, &fixflag); });
}
