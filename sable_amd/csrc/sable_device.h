// Device-visible plan descriptors shared by the host plan builder and the
// kernels.  Plain C layout (no HIP types) so the builder compiles with g++.
#pragma once
#ifdef __HIPCC_RTC__
// hipRTC: fixed-width types without system headers (LP64, as on the host)
typedef signed char int8_t;
typedef int int32_t;
typedef long int64_t;
typedef unsigned char uint8_t;
typedef unsigned short uint16_t;
typedef unsigned int uint32_t;
typedef unsigned long uint64_t;
#else
#include <stdint.h>
#endif

#define SB_PIECE_DENSE 0        // column-major hs x ncols FP32 sub-panel
#define SB_PIECE_SPARSE 1       // CSR over hs rows, uint16 local columns
#define SB_PIECE_MFMA 2         // FP32 sub-panel in MFMA A-operand order (SpMM)

// One unit of work: a row segment [row0, row0+nrows) of one block row of one
// matrix.  The tile owns those y rows (y is written once, by this tile).
struct SbTile {
    int32_t row0;
    int32_t nrows;
    int32_t piece_begin;     // pieces[piece_begin, piece_end)
    int32_t piece_end;
    int32_t mat;             // matrix slot (batched plans); 0 otherwise
    int32_t dense_tpar;      // T': active lanes of the flat dense stream
    int32_t dense_mcols;     // M : columns advanced per stream step (4T' = M*nrows)
    int32_t sp_glog;         // SpMV sparse: log2 of the lanes per row (G)
    int32_t sp_map;          // SpMV sparse: srowptr offset of the row order
                             // (nrows local rows, longest sparse row first); -1 if none
    int32_t sp_waves;        // SpMV: wavefronts reserved for the sparse pieces (the
                             // last ones; the rest stream the dense pieces), or 0
                             // when every wavefront does both in turn
    int32_t y_accum;         // SpMM: 1 = add to Y rows an MFMA tile wrote before
    int32_t stage_map;       // SpMV: xmap_off and stage_len of the first staging group
    int32_t stage_len;       // (copies, so its map loads follow the tile read directly)
    // SpMM MFMA row tiles split along k: this work item is chunk kpart of
    // ksplit (1 = not split).  Every chunk writes its partial sums to scratch
    // blocks part_blk + kpart * (16-row units * N/16) (1 KiB blocks, in
    // accumulator order); a fixup launch after the MFMA launch adds the
    // chunks of each split row tile in chunk order and writes its rows.
    int32_t ksplit;
    int32_t kpart;
    int32_t part_blk;
};

// SpMV sparse slices: steps per pipelined batch; and the batches the SpMV
// stream loops (dense and sparse) complete per round trip (their register
// copies of the next batch wait for it).  The plan builder balances the
// sparse and dense wavefronts' dependent chains with them.
#define SB_SPUNROLL 2
#define SB_SPMV_INFLIGHT 1

// SpMV row spans: runs of short block rows (at most SB_SPMV_SHORT_ROWS rows
// and SB_SPMV_SHORT_VALUES stored values each) are planned as one tile of at
// most SB_SPMV_SPAN_ROWS rows (and the workgroup size) and
// SB_SPMV_SPAN_VALUES stored values, every block on the sparse path.
#define SB_SPMV_SHORT_ROWS 32
#define SB_SPMV_SHORT_VALUES 8192
#define SB_SPMV_SPAN_ROWS 256
#define SB_SPMV_SPAN_VALUES 65536

// LDS for the staged X rows of a FP32 SpMM piece (64 KiB): the plan builder
// gives such a piece at most SB_SPMM_XS_FLOATS / N columns.
#define SB_SPMM_XS_FLOATS 16384

// FP32 SpMM thread shape.  Wide (N a multiple of 8, at least 64): a thread
// holds SB_F32_WIDE_ROWS rows (4 or 8) x 8 columns (two float4 column quads,
// one in each half of N), so a tile has at most (WG / (N/8)) *
// SB_F32_WIDE_ROWS rows.  Otherwise: up to 8 rows x one float4 column quad,
// at most (WG / (N/4)) * 8 rows and 64.
#define SB_F32_WIDE(N) ((N) >= 64 && (N) % 8 == 0)
#define SB_F32_COLS(N) (SB_F32_WIDE(N) ? 8 : 4)
#ifndef SB_F32_WIDE_ROWS
#define SB_F32_WIDE_ROWS 4
#endif

// SpMM MFMA tiles: a wavefront holds SB_MFMA_UNITS(N) 16-row tiles across
// all N/16 column tiles (their FP32 accumulators plus the double-buffered A
// fragments), which bounds a tile at SB_MFMA_MAX_ROWS(N, WG) rows (and 256).
#define SB_MFMA_UNITS(N) ((N) >= 256 ? 1 : 2)
#define SB_MFMA_MAX_ROWS(N, WG) \
    (16 * ((WG) / 64) * SB_MFMA_UNITS(N) < 256 ? 16 * ((WG) / 64) * SB_MFMA_UNITS(N) : 256)

// A panel piece of a tile: the tile's rows restricted to a group of blocks
// whose columns (ncols <= the x staging capacity) are staged in LDS.
struct SbPiece {
    int64_t data_off;        // dense: float offset (multiple of 4) into dval
                             // sparse: entry offset into sval/sidx
                             // mfma: float offset (multiple of 8) into dval
    int32_t kind;            // SB_PIECE_*
    int32_t ncols;           // staged columns (x slice length)
    int32_t xr_begin;        // staging ranges xr[xr_begin, xr_end)
    int32_t xr_end;
    int32_t rowptr_off;      // sparse, SpMM: CSR srowptr[rowptr_off .. +nrows] (relative
                             // to data_off); SpMV: per-wavefront {step offset,
                             // steps} pairs at srowptr[rowptr_off ..]
    int32_t ksteps;          // mfma: number of 32-wide k steps (ncols padded)
    // SpMV staging: pieces are staged in groups of consecutive pieces whose
    // slices (ncols + 4 zero pads each) fit the x buffer together.
    int32_t grp_end;         // first piece of a group: one past its last piece; else 0
    int32_t stage_len;       // first piece of a group: staged slots of the group
    int32_t xmap_off;        // xmap[xmap_off .. + ncols + 4): x column of each staged
                             // slot, -1 for the zero pads; consecutive pieces of a
                             // tile are consecutive in xmap
    int32_t pitch;           // dense: floats from one panel column to the next (SpMV:
                             // the tile's rows; SpMM: the rows rounded up to 4, the
                             // extra rows zero, so a column's row quads are float4s)
};

// A contiguous run of x (or X rows) copied into the staging buffer.
struct SbXRange {
    int32_t col0;
    int32_t len;
};
