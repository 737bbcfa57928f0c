// Host-side plan builder: VBR + per-block routes -> HBM layout + tile list.
// Plain C++17 (no HIP), so it is built and tested on machines without a GPU.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "sable_device.h"
#include "../../include/sable.h"

namespace sable {

struct HostPlan {
    int32_t op = SABLE_OP_SPMV;
    int32_t n_rhs = 1;
    int32_t nrows = 0, ncols = 0;
    int32_t wg = 256;
    std::vector<SbTile> tiles;
    std::vector<SbPiece> pieces;
    std::vector<SbXRange> xr;
    std::vector<float> dval;      // FP32 dense pieces (both layouts)
    std::vector<float> sval;
    std::vector<uint16_t> sidx;
    std::vector<int32_t> srowptr;
    std::vector<int32_t> xmap;    // SpMV: x column per staged slot (-1 = zero pad)
    std::vector<double> tile_cost;   // estimated bytes per tile (ordering)
    std::vector<int32_t> tile_phase; // 0: SpMM MFMA tile (first launch), 1: the rest
    int64_t part_blocks = 0;         // 1 KiB partial-sum blocks of the split MFMA tiles
    std::vector<SbTile> fix_tiles;   // one per split MFMA row tile: its rows, ksplit, part_blk
    SablePlanStats stats{};
};

// SpMV x staging capacity the host groups pieces for: a group's staged slots
// (ncols + 4 per piece) never exceed SB_STAGE_CAP + 4.
constexpr int SB_STAGE_CAP = 4096;

// Fills defaults for zero / negative fields.
SablePlanOpts normalize_opts(const SablePlanOpts *in);

// Returns SABLE_OK or an error code with a message in err.
int validate_vbr(const SableVbr &v, std::string &err);

// Builds the host plan.  route may be null (all dense).
int build_host_plan(const SableVbr &v, const uint8_t *route, const SablePlanOpts &o,
                    HostPlan &out, std::string &err);


}  // namespace sable
