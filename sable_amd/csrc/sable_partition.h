// Adjacent-line similarity and cut flags for the block partitioner
// (reference: src/autopartition.py:34-68 cut_indices2, :70-94 similarity1 /
// similarity2 / similarity2_numba).
//
// One axis at a time: the lines are the columns of a CSC matrix (indptr,
// indices ascending and unique within a line, FP64 values); rows are done on
// the transpose.  For every adjacent pair (i, i+1) three lanes each walk the
// two sorted index lists once and form
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

// Four lanes per adjacent pair: lanes 0-2 walk the merges for d0, d1, d2,
// lane 3 counts the nonzero values; lane 0 gathers them (4-lane shuffles)
// and writes the similarity.  Launch 4 * (nlines - 1) threads.
__global__ void __launch_bounds__(256) sb_adjacent_sim(const int64_t *__restrict__ ip,
                                                       const int32_t *__restrict__ ix,
                                                       const double *__restrict__ v, int64_t nlines,
                                                       int64_t line_len, int kind, double *__restrict__ sim)
{
#pragma clang fp contract(off)
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t i = t >> 2;
    const int role = (int)(t & 3);
    const bool live = i + 1 < nlines;
    double d = 0.0, nz = 0.0;
    if (live) {
        const int64_t a0 = ip[i], a1 = ip[i + 1], b1 = ip[i + 2];
        if (role == 3) {
            int64_t ca = 0, cb = 0;
            for (int64_t k = a0; k < a1; ++k) ca += v[k] != 0.0;
            for (int64_t k = a1; k < b1; ++k) cb += v[k] != 0.0;
            d = (double)ca;
            nz = (double)cb;
        } else if (role == 0 || line_len > 1) {
            d = sb_shift_dot(ix, v, a0, a1, a1, b1, role == 0 ? 0 : role == 1 ? 1 : -1);
        }
    }
    // the 4 lanes of a pair are consecutive lanes of one wavefront
    const double d0 = __shfl(d, 0, 4), d1 = __shfl(d, 1, 4), d2 = __shfl(d, 2, 4);
    const double ca = __shfl(d, 3, 4), cb = __shfl(nz, 3, 4);
    if (!live || role != 0) return;
    const double den = ca > cb ? ca : cb;
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
    if (kind == SB_SIM_NUMBA && (ca == 0.0 || cb == 0.0)) out = 0.0;
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
