// Adjacent-line similarity and cut flags for the block partitioner
// (reference: src/autopartition.py:34-68 cut_indices2, :70-94 similarity1 /
// similarity2 / similarity2_numba).
//
// One axis at a time: the lines are the columns of a CSC matrix (indptr,
// indices ascending and unique within a line, FP64 values); rows are done on
// the transpose.  For every adjacent pair (i, i+1) one wavefront forms
//   d0 = sum a[r] b[r],  d1 = sum a[r+1] b[r],  d2 = sum a[r] b[r+1]
// in ascending r, each product and sum rounded on its own (contraction is
// switched off: HIP otherwise fuses a * b + c into one FMA),
// so the result is bit-identical to the host restatement that sums the same
// products in the same order.  The similarity is then
//   similarity2:       max(d0, d1, d2) / max(ca, cb)      (0/0 = NaN)
//   similarity2_numba: 0 if ca == 0 or cb == 0, else as similarity2
//   similarity1:       ((d0 + d1) + d2) / (3 max(ca, cb))
// with ca, cb the counts of nonzero values, and only d0 for one-long lines.
//
// The reference's cut scan (skip a run of lines with no stored entry and cut
// after it, else cut between i and i+1 when the similarity is below the
// threshold) has a closed form per position p = i + 1, so it runs in
// parallel: cut[0] = cut[n] = 1 and, for 0 <= i < n - 1,
//   cut[i+1] = stored(i) ? sim(i) < threshold : stored(i+1)
// (NaN never cuts; inside an empty run no cut; at its end a cut).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define SB_SIM_S2 0
#define SB_SIM_NUMBA 1
#define SB_SIM_S1 2

// sum of a[k] b[l] over the entry pairs with ix[k] - s == ix[l], ascending
__device__ __forceinline__ double sb_shift_dot(const int32_t *__restrict__ ix, const double *__restrict__ v,
                                               int64_t a0, int64_t a1, int64_t b0, int64_t b1, int s)
{
#pragma clang fp contract(off)   // HIP contracts a * b + c into one FMA by default; the
                                 // pragma covers only operations written in this scope
    double acc = 0.0;
    int64_t ka = a0, kb = b0;
    while (ka < a1 && kb < b1) {
        const int64_t x = (int64_t)ix[ka] - s, y = ix[kb];
        if (x < y) {
            ++ka;
        } else if (y < x) {
            ++kb;
        } else {
            const double prod = v[ka] * v[kb];
            acc = acc + prod;
            ++ka;
            ++kb;
        }
    }
    return acc;
}

// One wavefront per adjacent pair (four pairs per 256-thread workgroup).
// When line b has at most SB_SIM_STAGE entries its row indices are staged in
// LDS first (one coalesced pass, together with its nonzero count), so the
// searches below read LDS; longer lines are searched in global memory.
// Per chunk of 64 entries of line a, lane l takes entry k = chunk + l (a
// coalesced read, requested one chunk ahead), binary-searches row ra - 1 in
// line b and checks the next three entries, forming its products for d0
// (rb = ra), d1 (rb = ra - 1) and d2 (rb = ra + 1).  Lanes 0-2 then add the chunk's products of d0 / d1 / d2
// from LDS in entry order, so every sum is the sequential ascending-row sum
// of the scalar merge.  Nonzero counts are wavefront sums of 0/1 flags.
#define SB_SIM_STAGE 2048
__global__ void __launch_bounds__(256) sb_adjacent_sim(const int64_t *__restrict__ ip,
                                                       const int32_t *__restrict__ ix,
                                                       const double *__restrict__ v, int64_t nlines,
                                                       int64_t line_len, int kind, double *__restrict__ sim)
{
#pragma clang fp contract(off)
    __shared__ double prod[4][3][64];
    __shared__ int32_t brow[4][SB_SIM_STAGE];
    const int wave = (int)(threadIdx.x >> 6), lane = (int)(threadIdx.x & 63);
    const int64_t i = (int64_t)blockIdx.x * 4 + wave;
    if (i + 1 >= nlines) return;                      // the whole wavefront
    const int64_t a0 = ip[i], a1 = ip[i + 1], b0 = a1, b1 = ip[i + 2];
    const bool staged = b1 - b0 <= SB_SIM_STAGE;      // wave-uniform
    int32_t *bs = brow[wave];
    int ca = 0, cb = 0;
    for (int64_t k = a0 + lane; k < a1; k += 64) ca += v[k] != 0.0;
    for (int64_t k = b0 + lane; k < b1; k += 64) {
        cb += v[k] != 0.0;
        if (staged) bs[k - b0] = ix[k];
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        ca += __shfl_xor(ca, o);
        cb += __shfl_xor(cb, o);
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const int nb = (int)(b1 - b0);
    double acc = 0.0;                                 // lane 0: d0, lane 1: d1, lane 2: d2
    // this chunk's entry of line a, loaded one chunk ahead (indices past the
    // line read its last entry; such lanes form no product)
    int32_t ra = 0;
    double va = 0.0;
    if (a1 > a0) {
        const int64_t k0 = a0 + lane < a1 ? a0 + lane : a1 - 1;
        ra = ix[k0];
        va = v[k0];
    }
    for (int64_t c0 = a0; c0 < a1; c0 += 64) {
        const int64_t k = c0 + lane;
        const int64_t kn = k + 64 < a1 ? k + 64 : a1 - 1;
        const int32_t ra_next = ix[kn];
        const double va_next = v[kn];
        double p0 = 0.0, p1 = 0.0, p2 = 0.0;
        if (k < a1) {
            if (staged) {
                int lo = 0, hi = nb;
                while (lo < hi) {                      // first entry of b with row >= ra - 1
                    const int mid = (lo + hi) >> 1;
                    if (bs[mid] < ra - 1) lo = mid + 1;
                    else hi = mid;
                }
                int32_t rb[3];
                double vb[3];
#pragma unroll
                for (int q = 0; q < 3; ++q) {          // independent loads, issued together
                    const bool in = lo + q < nb;
                    rb[q] = in ? bs[lo + q] : INT32_MAX;
                    vb[q] = in ? v[b0 + lo + q] : 0.0;
                }
                for (int q = 0; q < 3; ++q) {
                    if (rb[q] > ra + 1) break;
                    const double pr = va * vb[q];
                    if (rb[q] == ra) p0 = pr;
                    else if (rb[q] == ra - 1) p1 = pr;
                    else p2 = pr;
                }
            } else {
                int64_t lo = b0, hi = b1;
                while (lo < hi) {                      // first entry of b with row >= ra - 1
                    const int64_t mid = (lo + hi) >> 1;
                    if (ix[mid] < ra - 1) lo = mid + 1;
                    else hi = mid;
                }
                for (int q = 0; q < 3 && lo + q < b1; ++q) {
                    const int32_t rb = ix[lo + q];
                    if (rb > ra + 1) break;
                    const double pr = va * v[lo + q];
                    if (rb == ra) p0 = pr;
                    else if (rb == ra - 1) p1 = pr;
                    else p2 = pr;
                }
            }
        }
        prod[wave][0][lane] = p0;
        prod[wave][1][lane] = p1;
        prod[wave][2][lane] = p2;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        if (lane < 3) {
            const int n = a1 - c0 < 64 ? (int)(a1 - c0) : 64;
            for (int j = 0; j < n; ++j) acc = acc + prod[wave][lane][j];
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        ra = ra_next;
        va = va_next;
    }
    const double d0 = __shfl(acc, 0), d1 = __shfl(acc, 1), d2 = __shfl(acc, 2);
    if (lane != 0) return;
    const double den = (double)(ca > cb ? ca : cb);
    double out;
    if (line_len > 1) {
        if (kind == SB_SIM_S1) {
            out = ((d0 + d1) + d2) / (3.0 * den);
        } else {
            const double m12 = d1 > d2 ? d1 : d2;
            out = (d0 > m12 ? d0 : m12) / den;
        }
    } else {
        out = kind == SB_SIM_S1 ? d0 / (3.0 * den) : d0 / den;
    }
    if (kind == SB_SIM_NUMBA && (ca == 0 || cb == 0)) out = 0.0;
    sim[i] = out;
}

__global__ void __launch_bounds__(256) sb_cut_flags(const int64_t *__restrict__ ip,
                                                    const double *__restrict__ sim, int64_t nlines,
                                                    double threshold, uint8_t *__restrict__ cut)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p > nlines) return;
    uint8_t c = 1;
    if (p > 0 && p < nlines) {
        const int64_t i = p - 1;
        const bool si = ip[i + 1] > ip[i], sn = ip[i + 2] > ip[i + 1];
        c = si ? (uint8_t)(sim[i] < threshold) : (uint8_t)sn;
    }
    cut[p] = c;
}
