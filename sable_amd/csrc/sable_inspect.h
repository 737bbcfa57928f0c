// Block materialisation on the device (reference:
// utils/convert_real_to_vbr.py:98-199, convert_sparse_to_vbr): from a CSC
// matrix and the row / column partitions to the VBR arrays.
//
// A block (a, b) is kept when it holds at least one stored entry; kept
// blocks are numbered in block-row-major order and stored densely in
// column-major order.  On the device:
//   1. row -> block row and column -> block column maps (binary searches of
//      rpntr / cpntr), and the column of every entry;
//   2. a bit grid of the kept blocks: row a of the grid is rw = ceil(nbc/32)
//      32-bit words, set with atomicOr by the first entry of every run of a
//      column's entries in one block row;
//   3. per block row, the count of kept blocks before each word (exclusive,
//      within the row) and the row's total; bpntrb / bpntre follow on the
//      host from the row totals (nbr values);
//   4. the ordinal of a kept block = bpntrb[a] + word prefix + popcount of
//      the lower bits of its word; per kept block its bindx and its size
//      h * w, whose exclusive scan is indx;
//   5. every entry is written to val[indx[ord] + (j - c0) h + (i - r0)]
//      (val zeroed first), and the nonzero values are counted per block
//      (one atomic add per run).
// Integer work and value copies only: the result equals the host
// restatement bit for bit.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

// line -> partition index: the last p with part[p] <= line
__global__ void __launch_bounds__(256) sb_part_map(const int64_t *__restrict__ part, int32_t nparts, int64_t n,
                                                   int32_t *__restrict__ map)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int32_t lo = 0, hi = nparts;                      // part[lo] <= i < part[hi]
    while (hi - lo > 1) {
        const int32_t mid = (lo + hi) >> 1;
        if (part[mid] <= i) lo = mid;
        else hi = mid;
    }
    map[i] = lo;
}

// column of every entry (one wavefront per column)
__global__ void __launch_bounds__(256) sb_entry_cols(const int64_t *__restrict__ ip, int64_t ncols,
                                                     int32_t *__restrict__ col)
{
    const int64_t j = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (j >= ncols) return;
    for (int64_t k = ip[j] + (threadIdx.x & 63); k < ip[j + 1]; k += 64) col[k] = (int32_t)j;
}

__global__ void __launch_bounds__(256) sb_mark_blocks(const int32_t *__restrict__ ix,
                                                      const int32_t *__restrict__ col, int64_t nnz,
                                                      const int32_t *__restrict__ rmap,
                                                      const int32_t *__restrict__ cmap, int64_t rw,
                                                      uint32_t *__restrict__ bits)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nnz) return;
    const int32_t j = col[k];
    const int64_t a = rmap[ix[k]];
    // a column's entries are in row order: only the first entry of each run
    // in one block row marks the block (one atomic per run, not per entry)
    if (k > 0 && col[k - 1] == j && rmap[ix[k - 1]] == a) return;
    const int64_t b = cmap[j];
    atomicOr(&bits[a * rw + (b >> 5)], 1u << (b & 31));
}

// per block row: exclusive count of kept blocks before each word of the row,
// and the row total
__global__ void __launch_bounds__(256) sb_row_prefix(const uint32_t *__restrict__ bits, int32_t nbr, int64_t rw,
                                                     int32_t *__restrict__ wpre, int32_t *__restrict__ rowcnt)
{
    const int32_t a = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
    if (a >= nbr) return;
    int32_t run = 0;
    for (int64_t w = 0; w < rw; ++w) {
        wpre[(int64_t)a * rw + w] = run;
        run += __popc(bits[(int64_t)a * rw + w]);
    }
    rowcnt[a] = run;
}

__device__ __forceinline__ int64_t sb_block_ordinal(const uint32_t *bits, const int32_t *wpre,
                                                    const int64_t *rowbeg, int64_t rw, int64_t a, int64_t b)
{
    const int64_t w = a * rw + (b >> 5);
    const uint32_t below = (b & 31) ? (bits[w] & ((1u << (b & 31)) - 1u)) : 0u;
    return rowbeg[a] + wpre[w] + __popc(below);
}

// one thread per grid word: bindx and size of every kept block in it
__global__ void __launch_bounds__(256) sb_kept_blocks(const uint32_t *__restrict__ bits,
                                                      const int32_t *__restrict__ wpre,
                                                      const int64_t *__restrict__ rowbeg, int32_t nbr, int64_t rw,
                                                      const int64_t *__restrict__ rpntr,
                                                      const int64_t *__restrict__ cpntr,
                                                      int64_t *__restrict__ bindx, int64_t *__restrict__ size)
{
    const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= (int64_t)nbr * rw) return;
    const int64_t a = w / rw;
    uint32_t m = bits[w];
    int64_t ord = rowbeg[a] + wpre[w];
    const int64_t h = rpntr[a + 1] - rpntr[a];
    while (m) {
        const int bit = __ffs(m) - 1;
        m &= m - 1u;
        const int64_t b = (w - a * rw) * 32 + bit;
        bindx[ord] = b;
        size[ord] = h * (cpntr[b + 1] - cpntr[b]);
        ++ord;
    }
}

// exclusive scan of int64 values in three launches: per-block scans with
// block totals, a scan of the totals in one block, then the block offsets
#define SB_SCAN_BLOCK 1024

__global__ void __launch_bounds__(SB_SCAN_BLOCK) sb_scan_blocks(const int64_t *__restrict__ in, int64_t n,
                                                                int64_t *__restrict__ out, int64_t *__restrict__ tot)
{
    __shared__ int64_t s[SB_SCAN_BLOCK];
    const int64_t g = (int64_t)blockIdx.x * SB_SCAN_BLOCK + threadIdx.x;
    const int64_t x = g < n ? in[g] : 0;
    s[threadIdx.x] = x;
    __syncthreads();
    for (int o = 1; o < SB_SCAN_BLOCK; o <<= 1) {     // Hillis-Steele, inclusive
        const int64_t y = threadIdx.x >= (unsigned)o ? s[threadIdx.x - o] : 0;
        __syncthreads();
        s[threadIdx.x] += y;
        __syncthreads();
    }
    if (g < n) out[g] = s[threadIdx.x] - x;
    if (threadIdx.x == SB_SCAN_BLOCK - 1) tot[blockIdx.x] = s[threadIdx.x];
}

__global__ void __launch_bounds__(SB_SCAN_BLOCK) sb_scan_totals(int64_t *__restrict__ tot, int64_t nb)
{
    // one block: serial carry over chunks of SB_SCAN_BLOCK totals
    __shared__ int64_t s[SB_SCAN_BLOCK];
    __shared__ int64_t carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int64_t c0 = 0; c0 < nb; c0 += SB_SCAN_BLOCK) {
        const int64_t g = c0 + threadIdx.x;
        const int64_t x = g < nb ? tot[g] : 0;
        s[threadIdx.x] = x;
        __syncthreads();
        for (int o = 1; o < SB_SCAN_BLOCK; o <<= 1) {
            const int64_t y = threadIdx.x >= (unsigned)o ? s[threadIdx.x - o] : 0;
            __syncthreads();
            s[threadIdx.x] += y;
            __syncthreads();
        }
        if (g < nb) tot[g] = carry + s[threadIdx.x] - x;
        __syncthreads();
        if (threadIdx.x == SB_SCAN_BLOCK - 1) carry += s[threadIdx.x];
        __syncthreads();
    }
}

__global__ void __launch_bounds__(SB_SCAN_BLOCK) sb_scan_add(int64_t *__restrict__ out, int64_t n,
                                                             const int64_t *__restrict__ tot)
{
    const int64_t g = (int64_t)blockIdx.x * SB_SCAN_BLOCK + threadIdx.x;
    if (g < n) out[g] += tot[blockIdx.x];
}

__global__ void __launch_bounds__(256) sb_scatter_values(const int32_t *__restrict__ ix,
                                                         const int32_t *__restrict__ col,
                                                         const double *__restrict__ v, int64_t nnz,
                                                         const int32_t *__restrict__ rmap,
                                                         const int32_t *__restrict__ cmap,
                                                         const uint32_t *__restrict__ bits,
                                                         const int32_t *__restrict__ wpre,
                                                         const int64_t *__restrict__ rowbeg, int64_t rw,
                                                         const int64_t *__restrict__ rpntr,
                                                         const int64_t *__restrict__ cpntr,
                                                         const int64_t *__restrict__ indx,
                                                         double *__restrict__ val, int32_t *__restrict__ bnnz)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = (int)(threadIdx.x & 63);
    const bool live = k < nnz;                        // live lanes are a prefix of the wavefront
    int64_t ord = -1;
    int32_t a = -1, j = -1, cnt = 0;
    if (live) {
        const int32_t i = ix[k];
        j = col[k];
        a = rmap[i];
        const int32_t b = cmap[j];
        ord = sb_block_ordinal(bits, wpre, rowbeg, rw, a, b);
        const int64_t h = rpntr[a + 1] - rpntr[a];
        const double x = v[k];
        val[indx[ord] + (j - cpntr[b]) * h + (i - rpntr[a])] = x;
        cnt = x != 0.0;
    }
    // Nonzero count per run (consecutive entries of one column in one block
    // row, i.e. of one block): runs are contiguous, so a segmented suffix sum
    // over the wavefront's lanes gives each run's count at its first lane in
    // this wavefront, which adds it once.  A run that crosses a wavefront
    // boundary is added in parts.
    const int32_t jp = __shfl_up(j, 1), ap = __shfl_up(a, 1);
    const bool head = lane == 0 || !live || jp != j || ap != a;
    const unsigned long long starts = __ballot(head);
    const unsigned long long upto = lane == 63 ? starts : starts & ((2ull << lane) - 1ull);
    const int seg = 63 - __builtin_clzll(upto);        // this lane's run start (lane 0 is always one)
    int32_t sum = cnt;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int32_t s2 = __shfl_down(sum, o);
        const int g2 = __shfl_down(seg, o);
        if (lane + o < 64 && g2 == seg) sum += s2;
    }
    if (live && lane == seg && sum) atomicAdd(&bnnz[ord], sum);
}

// ---- compressed (.vbrc) arrays on the device (reference:
// utils/convert_real_to_vbr.py:18-96, convert_vbr_to_compressed, for given
// routes).  Dense-routed blocks are gathered into val2 (one wavefront per
// block); the nonzero values of sparse-routed blocks become a CSR matrix.
// One thread per row walks its block row's sparse blocks in bindx order and,
// inside a block, its columns in order, so every row comes out sorted by
// column, as scipy's COO -> CSR conversion leaves it.  The reference's
// padding entry (row nrows - 1, column ncols - 1, value 0, when the last row
// or the last column holds no entry) is appended by the host.

__global__ void __launch_bounds__(256) sb_gather_dense(const double *__restrict__ val,
                                                       const int64_t *__restrict__ indx,
                                                       const int64_t *__restrict__ dblk,
                                                       const int64_t *__restrict__ indx2, int64_t ndense,
                                                       double *__restrict__ val2)
{
    const int64_t d = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (d >= ndense) return;
    const int64_t k = dblk[d], src = indx[k], n = indx[k + 1] - src, dst = indx2[d];
    // four loads in flight per lane, then their stores
    for (int64_t e = threadIdx.x & 63; e < n; e += 256) {
        double x[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) x[u] = e + 64 * u < n ? val[src + e + 64 * u] : 0.0;
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (e + 64 * u < n) val2[dst + e + 64 * u] = x[u];
    }
}

// per row: the nonzero values of its sparse-routed blocks (count pass when
// cols == nullptr, fill pass otherwise); lastcol is set when one of them is
// in the last column
__global__ void __launch_bounds__(256) sb_csr_rows(const double *__restrict__ val,
                                                   const int64_t *__restrict__ indx,
                                                   const int64_t *__restrict__ bindx,
                                                   const int64_t *__restrict__ rpntr,
                                                   const int64_t *__restrict__ cpntr,
                                                   const int64_t *__restrict__ bpntrb,
                                                   const int64_t *__restrict__ bpntre,
                                                   const uint8_t *__restrict__ sparse, const int32_t *__restrict__ rmap,
                                                   int64_t nrows, int64_t ncols, int64_t *__restrict__ cnt,
                                                   const int64_t *__restrict__ rowptr, int64_t *__restrict__ cols,
                                                   double *__restrict__ vals, int32_t *__restrict__ lastcol)
{
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrows) return;
    const int64_t a = rmap[r];
    int64_t n = 0, pos = cols ? rowptr[r] : 0;
    if (bpntrb[a] >= 0) {
        const int64_t h = rpntr[a + 1] - rpntr[a], lr = r - rpntr[a];
        for (int64_t k = bpntrb[a]; k < bpntre[a]; ++k) {
            if (!sparse[k]) continue;
            const int64_t b = bindx[k], c0 = cpntr[b], w = cpntr[b + 1] - c0;
            const double *__restrict__ blk = val + indx[k] + lr;
            // eight columns per step, their loads issued together (a
            // column past the block reads its last column again)
            for (int64_t c = 0; c < w; c += 8) {
                double xv[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) xv[u] = blk[(c + u < w ? c + u : w - 1) * h];
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    if (c + u >= w || xv[u] == 0.0) continue;
                    if (cols) {
                        cols[pos] = c0 + c + u;
                        vals[pos] = xv[u];
                        ++pos;
                    } else {
                        ++n;
                        if (c0 + c + u == ncols - 1) *lastcol = 1;
                    }
                }
            }
        }
    }
    if (!cols) cnt[r] = n;
}
