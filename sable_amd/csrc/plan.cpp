// Host plan builder.  See plan.h.
//
// Layout decisions (all per block row, after routing):
//  * dense-routed blocks of a block row are concatenated in ordinal order
//    (each sub-panel is followed by one zero float4 that the kernel's
//    unconditional loads read past the end):
//    in the VBR they already form one column-major h x W panel (same height,
//    consecutive storage), so a row segment of that panel becomes one
//    contiguous column-major sub-panel per column group;
//  * sparse-routed blocks keep only their nonzeros (value != 0.0, the
//    reference's routing count, utils/convert_real_to_vbr.py:41-47) as a CSR
//    over the segment's rows with uint16 column indices into the group's x
//    slice (entries of a row in ascending global column order);
//  * MFMA-routed blocks (SpMM) are stored as FP32 in the exact register order
//    of v_mfma_f32_16x16x32_bf16's A operand (16-row tile m, 32-column k-step
//    s, lane l: A[16m + l%16][32s + 8(l/16) + j], j = 0..7).  A fragment
//    (m, s) is 512 floats: j = 0..3 of lanes 0..63, then j = 4..7 of lanes
//    0..63, so each of a lane's two 16-byte loads is part of one contiguous
//    1 KiB wavefront read; the lane splits its 8 k-values in registers.
//  * column groups never exceed the LDS staging capacity (xcap columns), and
//    blocks wider than that are split by columns.
#include "plan.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>

namespace sable {

SablePlanOpts normalize_opts(const SablePlanOpts *in)
{
    SablePlanOpts o{};
    if (in) o = *in;
    if (o.op != SABLE_OP_SPMM) o.op = SABLE_OP_SPMV;
    if (o.op == SABLE_OP_SPMV) o.n_rhs = 1;
    if (o.wg <= 0) o.wg = 256;
    if (o.xcap <= 0) o.xcap = 4095;
    if (o.tile_bytes <= 0) o.tile_bytes = (o.op == SABLE_OP_SPMV) ? 128 * 1024 : 512 * 1024;
    if (o.max_seg_rows <= 0) o.max_seg_rows = o.wg;
    if (o.min_seg_rows <= 0) o.min_seg_rows = (o.op == SABLE_OP_SPMV) ? 8 : 16;
    if (o.mode < 0 || o.mode > 2) o.mode = SABLE_MODE_GRID;
    if (o.unroll <= 0) o.unroll = 4;
    if (o.mfma_split != 2) o.mfma_split = 3;
    // SpMV tiles are light (stream-bound); SpMM tiles carry MFMA work and
    // gain from tall tiles (each staged X k-step serves every row)
    if (o.min_tiles <= 0) o.min_tiles = (o.op == SABLE_OP_SPMV) ? 1024 : 256;
    // k-split of MFMA row tiles only on request: on the 50k SpMM at N = 128
    // 32-k-step chunks measured 1-3.5 % slower than unsplit tiles (the fixup
    // launch and the partial-sum traffic outweigh the shorter chains)
    if (o.mfma_kchunk == 0) o.mfma_kchunk = -1;
    return o;
}

int validate_vbr(const SableVbr &v, std::string &err)
{
    if (v.nbr < 1 || v.nbc < 1) { err = "need at least one block row and block column"; return SABLE_ERR_INVALID; }
    if (!v.rpntr || !v.cpntr || !v.bpntrb || !v.bpntre || !v.indx) { err = "null VBR array"; return SABLE_ERR_INVALID; }
    if (v.nblocks > 0 && (!v.bindx || !v.val)) { err = "null bindx/val"; return SABLE_ERR_INVALID; }
    if (v.rpntr[0] != 0 || v.cpntr[0] != 0) { err = "rpntr/cpntr must start at 0"; return SABLE_ERR_INVALID; }
    for (int a = 0; a < v.nbr; ++a)
        if (v.rpntr[a + 1] <= v.rpntr[a]) { err = "rpntr must be strictly increasing"; return SABLE_ERR_INVALID; }
    for (int b = 0; b < v.nbc; ++b)
        if (v.cpntr[b + 1] <= v.cpntr[b]) { err = "cpntr must be strictly increasing"; return SABLE_ERR_INVALID; }
    int64_t pos = 0;
    for (int a = 0; a < v.nbr; ++a) {
        const int32_t b0 = v.bpntrb[a], b1 = v.bpntre[a];
        if (b0 == -1 || b1 == -1) {
            if (b0 != b1) { err = "bpntrb/bpntre must mark empty rows with -1 together"; return SABLE_ERR_INVALID; }
            continue;
        }
        if (b0 != pos || b1 < b0) { err = "block rows must be stored consecutively in row order"; return SABLE_ERR_INVALID; }
        for (int32_t p = b0; p < b1; ++p) {
            if (p >= v.nblocks) { err = "bpntre exceeds the number of blocks"; return SABLE_ERR_INVALID; }
            const int32_t b = v.bindx[p];
            if (b < 0 || b >= v.nbc) { err = "bindx out of range"; return SABLE_ERR_INVALID; }
            if (p > b0 && b <= v.bindx[p - 1]) { err = "bindx must be strictly ascending within a block row"; return SABLE_ERR_INVALID; }
            const int64_t sz = (int64_t)(v.rpntr[a + 1] - v.rpntr[a]) * (v.cpntr[b + 1] - v.cpntr[b]);
            if (v.indx[p + 1] - v.indx[p] != sz) { err = "indx steps must equal block height x width"; return SABLE_ERR_INVALID; }
        }
        pos = b1;
    }
    if (pos != v.nblocks) { err = "block count of the rows differs from nblocks"; return SABLE_ERR_INVALID; }
    if (v.indx[0] != 0) { err = "indx must start at 0"; return SABLE_ERR_INVALID; }
    return SABLE_OK;
}

namespace {

struct ColSeg {       // columns [cb, cb+n) of block k
    int64_t k;
    int32_t cb, n;
    int32_t gcol;     // global column of cb
};

// Splits the blocks of one route into column groups of at most cap columns.
std::vector<std::vector<ColSeg>> group_columns(const std::vector<int64_t> &blocks,
                                               const SableVbr &v, int cap)
{
    std::vector<std::vector<ColSeg>> groups;
    std::vector<ColSeg> cur;
    int width = 0;
    for (int64_t k : blocks) {
        const int32_t b = v.bindx[k];
        const int32_t w = v.cpntr[b + 1] - v.cpntr[b];
        int32_t cb = 0;
        while (cb < w) {
            if (width == cap) { groups.push_back(cur); cur.clear(); width = 0; }
            const int32_t n = std::min(w - cb, cap - width);
            cur.push_back({k, cb, n, v.cpntr[b] + cb});
            width += n;
            cb += n;
        }
    }
    if (!cur.empty()) groups.push_back(cur);
    return groups;
}

int group_width(const std::vector<ColSeg> &g)
{
    int w = 0;
    for (auto &s : g) w += s.n;
    return w;
}

void push_xranges(HostPlan &P, const std::vector<ColSeg> &g, SbPiece &pc)
{
    pc.xr_begin = (int32_t)P.xr.size();
    for (auto &s : g) {
        if ((int32_t)P.xr.size() > pc.xr_begin) {
            SbXRange &last = P.xr.back();
            if (last.col0 + last.len == s.gcol) { last.len += s.n; continue; }
        }
        P.xr.push_back({s.gcol, s.n});
    }
    pc.xr_end = (int32_t)P.xr.size();
}

void dense_stream_params(int hs, int wg, int32_t &tpar, int32_t &mcols)
{
    // 4*T' must be a multiple of lcm(hs, 4) and T' <= wg.
    const int l = std::lcm(hs, 4);
    const int groups = (4 * wg) / l;
    tpar = groups * l / 4;
    mcols = 4 * tpar / hs;
}

}  // namespace

// SpMV staging of one tile (pieces already ordered sparse first): consecutive
// pieces are grouped while their slices
// (ncols + 4 zero pads each) fit SB_STAGE_CAP + 4 slots, and each piece's
// slice is described slot by slot in xmap (x column, -1 for a pad), pieces
// of the tile back to back, so a group's map is one contiguous run.
void stage_groups(HostPlan &P, const SbTile &t)
{
    int32_t p = t.piece_begin;
    while (p < t.piece_end) {
        int32_t q = p, used = 0;
        while (q < t.piece_end && used + P.pieces[q].ncols + 4 <= SB_STAGE_CAP + 4) used += P.pieces[q++].ncols + 4;
        P.pieces[p].grp_end = q;
        P.pieces[p].stage_len = used;
        p = q;
    }
    for (int32_t k = t.piece_begin; k < t.piece_end; ++k) {
        SbPiece &pc = P.pieces[k];
        pc.xmap_off = (int32_t)P.xmap.size();
        for (int32_t r = pc.xr_begin; r < pc.xr_end; ++r)
            for (int32_t c = 0; c < P.xr[r].len; ++c) P.xmap.push_back(P.xr[r].col0 + c);
        P.xmap.insert(P.xmap.end(), 4, -1);
    }
}

// Nonzeros of rows [s0, s0+hs) of a row span restricted to one column group,
// row by row in column order (local column = position in the group's
// concatenated columns).  The span is one block row, or a run of short block
// rows planned as one (SpMV): geo gives each block's first row in the span
// and its height.  keep[k]: the block's zeros are entries too (a dense-routed
// block of such a run, so its products are the dense path's).  Returns false
// when there are none.
struct SparseRows {
    int32_t group = 0;
    std::vector<int32_t> rp;      // hs + 1 row pointers
    std::vector<float> val;
    std::vector<uint16_t> col;
};

struct SpanGeo {
    const SableVbr *v;
    const int32_t *brow;          // block row of every block
    int32_t r0;                   // first row of the span
    const std::vector<char> *keep;
    int32_t off(int64_t k) const { return v->rpntr[brow[k]] - r0; }
    int32_t height(int64_t k) const { return v->rpntr[brow[k] + 1] - v->rpntr[brow[k]]; }
    bool keeps(int64_t k) const { return keep && (*keep)[k]; }
};

bool collect_sparse_rows(const SableVbr &v, const std::vector<ColSeg> &g, const SpanGeo &geo, int32_t s0, int32_t hs,
                         SparseRows &out)
{
    // tile rows [s0, s0+hs) that block k covers: [lo, hi), block row br = r + s0 - off
    auto span = [&](int64_t k, int32_t &lo, int32_t &hi, int32_t &shift) {
        const int32_t off = geo.off(k), bh = geo.height(k);
        lo = std::max(0, off - s0);
        hi = std::min(hs, off + bh - s0);
        shift = s0 - off;
    };
    std::vector<int32_t> cnt(hs, 0);
    for (auto &cs : g) {
        const float *bv = v.val + v.indx[cs.k];
        const int32_t bh = geo.height(cs.k);
        const bool keep = geo.keeps(cs.k);
        int32_t lo, hi, shift;
        span(cs.k, lo, hi, shift);
        for (int32_t c = cs.cb; c < cs.cb + cs.n; ++c) {
            const float *col = bv + (int64_t)c * bh + shift;
            for (int32_t r = lo; r < hi; ++r) cnt[r] += (keep || col[r] != 0.0f);
        }
    }
    out.rp.assign(hs + 1, 0);
    for (int32_t r = 0; r < hs; ++r) out.rp[r + 1] = out.rp[r] + cnt[r];
    const int32_t tot = out.rp[hs];
    if (tot == 0) return false;
    out.val.resize(tot);
    out.col.resize(tot);
    std::vector<int32_t> cur(out.rp.begin(), out.rp.end() - 1);
    int32_t lc = 0;
    for (auto &cs : g) {
        const float *bv = v.val + v.indx[cs.k];
        const int32_t bh = geo.height(cs.k);
        const bool keep = geo.keeps(cs.k);
        int32_t lo, hi, shift;
        span(cs.k, lo, hi, shift);
        for (int32_t c = cs.cb; c < cs.cb + cs.n; ++c, ++lc) {
            const float *col = bv + (int64_t)c * bh + shift;
            for (int32_t r = lo; r < hi; ++r) {
                if (keep || col[r] != 0.0f) {
                    out.val[cur[r]] = col[r];
                    out.col[cur[r]] = (uint16_t)lc;
                    ++cur[r];
                }
            }
        }
    }
    return true;
}

// SpMV sparse layout ("wave slices").  The tile's rows are ordered by total
// sparse length (longest first, ties by row) and cut into slices of C = 64/G
// consecutive rows; each row gets G lanes of a wavefront.  Slice s belongs
// to sparse wavefront s % nsw.  A slice step is 256 consecutive entries (64
// lanes x 4): lane q*G + g takes 4-entry chunk c = j*G + g of its row at step
// j, so every step is one coalesced 1 KiB value load and 512 B index load.
// A row is padded to its slice's step count with value 0 at the zero pad
// slot of the staged x slice (local column ncols).  Per piece the table
// holds {step offset, steps} for each slice.  G is chosen per tile to
// minimise padded slots plus half the slots the busiest sparse wavefront
// keeps the tile waiting for (a slice start counts two steps).  Returns the
// cost.
// The tile's rows by total sparse length, longest first (ties by row).
std::vector<int32_t> sparse_row_order(const std::vector<SparseRows> &sp, int32_t hs)
{
    std::vector<int64_t> tot(hs, 0);
    for (auto &sr : sp)
        for (int32_t r = 0; r < hs; ++r) tot[r] += sr.rp[r + 1] - sr.rp[r];
    std::vector<int32_t> order(hs);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int32_t x, int32_t y) { return tot[x] > tot[y]; });
    return order;
}

// Steps of every (piece, slice) for lanes-per-row 2^glog and nsw sparse
// wavefronts: returns {padded slots, steps of the busiest wavefront, a slice
// start counting start_steps}.
std::pair<int64_t, int64_t> slice_widths(const std::vector<SparseRows> &sp, const std::vector<int32_t> &order,
                                         int32_t hs, int glog, int nsw, std::vector<std::vector<int32_t>> *out,
                                         int start_steps = 2)
{
    const int G = 1 << glog, C = 64 >> glog;
    const int nsl = (hs + C - 1) / C;
    int64_t padded = 0, crit = 0;
    std::vector<int64_t> wsteps(nsw, 0);
    for (size_t p = 0; p < sp.size(); ++p) {
        std::vector<int32_t> w(nsl, 0);
        for (int s = 0; s < nsl; ++s) {
            int32_t m = 0;
            for (int q = 0; q < C && s * C + q < hs; ++q) {
                const int32_t r = order[s * C + q];
                const int32_t len = sp[p].rp[r + 1] - sp[p].rp[r];
                m = std::max(m, (len + 4 * G - 1) / (4 * G));
            }
            w[s] = m;
            wsteps[s % nsw] += m + start_steps;
            padded += 256LL * m;
        }
        if (out) out->push_back(w);
    }
    for (int k = 0; k < nsw; ++k) crit = std::max(crit, wsteps[k]);
    return std::make_pair(padded, crit);
}

// G minimising padded slots plus half the slots the busiest sparse wavefront
// keeps the tile waiting for; *crit = that wavefront's steps.
int best_slice_glog(const std::vector<SparseRows> &sp, const std::vector<int32_t> &order, int32_t hs, int nsw,
                    int64_t *crit)
{
    int best = 0;
    double best_cost = 0;
    int64_t best_crit = 0;
    for (int glog = 0; glog <= 6; ++glog) {
        const auto pc = slice_widths(sp, order, hs, glog, nsw, nullptr);
        const double c = (double)pc.first + 0.5 * 256.0 * nsw * (double)pc.second;
        if (glog == 0 || c < best_cost) { best = glog; best_cost = c; best_crit = pc.second; }
    }
    if (crit) *crit = best_crit;
    return best;
}

// glog < 0: G by best_slice_glog.
double emit_wave_slices(HostPlan &P, SablePlanStats &S, SbTile &t, const std::vector<SparseRows> &sp,
                        const std::vector<std::vector<ColSeg>> &sgroups, int32_t hs, int nsw, int glog)
{
    const std::vector<int32_t> order = sparse_row_order(sp, hs);
    const int best = glog >= 0 ? glog : best_slice_glog(sp, order, hs, nsw, nullptr);
    std::vector<std::vector<int32_t>> W;
    slice_widths(sp, order, hs, best, nsw, &W);
    const int G = 1 << best, C = 64 >> best;
    const int nsl = (hs + C - 1) / C;
    t.sp_glog = best;
    t.sp_map = (int32_t)P.srowptr.size();
    P.srowptr.insert(P.srowptr.end(), order.begin(), order.end());
    S.rowptr_entries += hs;
    double cost = 4.0 * hs;
    for (size_t p = 0; p < sp.size(); ++p) {
        const SparseRows &sr = sp[p];
        SbPiece pc{};
        pc.kind = SB_PIECE_SPARSE;
        pc.ncols = group_width(sgroups[sr.group]);
        pc.data_off = (int64_t)P.sval.size();
        pc.rowptr_off = (int32_t)P.srowptr.size();
        int64_t steps = 0;
        for (int s = 0; s < nsl; ++s) {
            P.srowptr.push_back((int32_t)steps);
            P.srowptr.push_back(W[p][s]);
            steps += W[p][s];
        }
        S.rowptr_entries += 2 * nsl;
        P.sval.resize(P.sval.size() + 256 * steps, 0.0f);
        P.sidx.resize(P.sidx.size() + 256 * steps, (uint16_t)pc.ncols);
        float *dv = P.sval.data() + pc.data_off;
        uint16_t *di = P.sidx.data() + pc.data_off;
        int64_t soff = 0;
        for (int s = 0; s < nsl; ++s) {
            for (int q = 0; q < C && s * C + q < hs; ++q) {
                const int32_t r = order[s * C + q];
                for (int32_t e = sr.rp[r]; e < sr.rp[r + 1]; ++e) {
                    const int32_t i = e - sr.rp[r];          // position in the row
                    const int32_t c = i / 4, k = i % 4;      // chunk, slot
                    const int64_t pos = (soff + c / G) * 256 + (int64_t)(q * G + c % G) * 4 + k;
                    dv[pos] = sr.val[e];
                    di[pos] = sr.col[e];
                }
            }
            soff += W[p][s];
        }
        S.sparse_nnz += (int64_t)sr.val.size();
        push_xranges(P, sgroups[sr.group], pc);
        P.pieces.push_back(pc);
        cost += 6.0 * 256.0 * (double)steps + 8.0 * nsl;
    }
    return cost;
}

namespace {
int build_host_plan_once(const SableVbr &v, const uint8_t *route, const SablePlanOpts &o, bool spans, HostPlan &P,
                         std::string &err);
}

int build_host_plan(const SableVbr &v, const uint8_t *route, const SablePlanOpts &o,
                    HostPlan &P, std::string &err)
{
    // Small matrices: halve the tile target until there are min_tiles tiles
    // (or tiles are 16 KiB), so the launch still spreads over the device.
    // Row spans trade tiles for less fixed cost per row; a plan they would
    // leave below min_tiles is built without them (a small matrix launched
    // on its own needs the parallelism; plans built for a group take
    // min_tiles 1 and keep them).
    for (int spans = 1; spans >= 0; --spans) {
        SablePlanOpts oo = o;
        for (;;) {
            const int rc = build_host_plan_once(v, route, oo, spans != 0, P, err);
            if (rc) return rc;
            if ((int64_t)P.tiles.size() >= oo.min_tiles || oo.tile_bytes <= 16 * 1024) break;
            oo.tile_bytes = std::max<int64_t>(16 * 1024, oo.tile_bytes / 2);
        }
        if ((int64_t)P.tiles.size() >= o.min_tiles || !P.stats.nspans) return SABLE_OK;
    }
    return SABLE_OK;
}

namespace {
int build_host_plan_once(const SableVbr &v, const uint8_t *route, const SablePlanOpts &o, bool spans, HostPlan &P,
                         std::string &err)
{
    int rc = validate_vbr(v, err);
    if (rc) return rc;
    if (o.op == SABLE_OP_SPMM && o.wg != 64 && o.wg != 128 && o.wg != 256 && o.wg != 512) {
        err = "SpMM needs wg 64, 128, 256 or 512";
        return SABLE_ERR_INVALID;
    }
    if (o.wg < 64 || o.wg > 1024 || o.wg % 64) { err = "wg must be a multiple of 64 in [64, 1024]"; return SABLE_ERR_INVALID; }
    if (o.xcap < 16 || o.xcap > 4095) { err = "xcap must be in [16, 4095]"; return SABLE_ERR_INVALID; }
    const bool spmm = o.op == SABLE_OP_SPMM;
    if (spmm && (o.n_rhs < 4 || o.n_rhs % 4 != 0 || o.n_rhs / 4 > o.wg)) {
        err = "SpMM needs n_rhs a positive multiple of 4 with n_rhs/4 <= wg";
        return SABLE_ERR_UNSUPPORTED;
    }
    P = HostPlan();
    P.op = o.op;
    P.n_rhs = spmm ? o.n_rhs : 1;
    P.nrows = v.rpntr[v.nbr];
    P.ncols = v.cpntr[v.nbc];
    P.wg = o.wg;
    SablePlanStats &S = P.stats;
    S.nrows = P.nrows;
    S.ncols = P.ncols;
    S.n_rhs = P.n_rhs;
    S.nblocks = v.nblocks;

    // rows per tile bound: SpMV needs hs <= wg (flat stream lanes).  The FP32
    // SpMM kernel's wide shape takes 4 rows per thread group, (wg / (N/8)) * 4
    // rows (8 rows per group measured slower: twice the registers and half
    // the groups busy); its narrow shape at most 8, (wg / (N/4)) * 8 rows and
    // 64.  MFMA tiles are cut separately: multiples of 16 rows, at most
    // mfma_max_rows.
    int max_rows = std::min(o.max_seg_rows, o.wg);
    if (spmm)
        max_rows = SB_F32_WIDE(o.n_rhs) ? std::min(max_rows, (o.wg / (o.n_rhs / 8)) * SB_F32_WIDE_ROWS)
                                        : std::min({max_rows, (o.wg / (o.n_rhs / 4)) * 8, 64});
    const int mfma_rows = spmm && o.n_rhs % 16 == 0
                              ? std::max(16, std::min(SB_MFMA_MAX_ROWS(o.n_rhs, o.wg), (o.max_seg_rows / 16) * 16))
                              : 16;
    max_rows = std::max(max_rows, 1);
    const int min_rows = std::max(1, std::min(o.min_seg_rows, max_rows));

    // count nonzeros (plan-invariant flops basis)
    for (int64_t i = 0; i < v.indx[v.nblocks]; ++i) S.nnz += (v.val[i] != 0.0f);

    // SpMV: a run of consecutive short block rows (at most SB_SPMV_SHORT_ROWS
    // rows and SB_SPMV_SHORT_VALUES stored values each) is planned as one row
    // span of at most SB_SPMV_SPAN_ROWS rows (and the tile height) and
    // SB_SPMV_SPAN_VALUES stored values: one tile in which every block takes
    // the sparse path (a dense-routed block's zeros included, so its
    // products are the dense path's).  A block row of a few rows would
    // otherwise be a tile of its own, whose workgroup's fixed chain (x
    // staging, barriers, fold) costs more than its few hundred values.  (In
    // SpMM the same spans on the FP32 path measured slower than one MFMA
    // tile per block row, padding included.)
    std::vector<int32_t> brow(std::max<int64_t>(v.nblocks, 1), 0);   // block row of every block
    for (int32_t a = 0; a < v.nbr; ++a)
        if (v.bpntrb[a] != -1)
            for (int64_t k = v.bpntrb[a]; k < v.bpntre[a]; ++k) brow[k] = a;
    std::vector<int32_t> span_end(v.nbr);
    {
        auto rows = [&](int32_t a) { return v.rpntr[a + 1] - v.rpntr[a]; };
        auto stored = [&](int32_t a) { return v.bpntrb[a] == -1 ? 0 : v.indx[v.bpntre[a]] - v.indx[v.bpntrb[a]]; };
        const int32_t span_rows = std::min(max_rows, SB_SPMV_SPAN_ROWS);
        for (int32_t a = 0; a < v.nbr;) {
            int32_t e = a + 1;
            auto is_short = [&](int32_t b) { return rows(b) <= SB_SPMV_SHORT_ROWS && stored(b) <= SB_SPMV_SHORT_VALUES; };
            if (spans && !spmm && is_short(a)) {
                int64_t nr = rows(a), nv = stored(a);
                while (e < v.nbr && is_short(e) && nr + rows(e) <= span_rows && nv + stored(e) <= SB_SPMV_SPAN_VALUES) {
                    nr += rows(e);
                    nv += stored(e);
                    ++e;
                }
            }
            span_end[a] = e;
            a = e;
        }
    }
    std::vector<char> keep;         // dense-routed blocks of a span of several block rows
    for (int32_t a = 0, a_end; a < v.nbr; a = a_end) {
        a_end = span_end[a];
        const bool span = a_end > a + 1;
        S.nspans += span;
        const int32_t r0 = v.rpntr[a], h = v.rpntr[a_end] - r0;
        std::vector<int64_t> dense, sparse, mfma;
        for (int32_t b = a; b < a_end; ++b) {
            if (v.bpntrb[b] == -1) continue;
            for (int64_t k = v.bpntrb[b]; k < v.bpntre[b]; ++k) {
                const uint8_t rt = route ? route[k] : (uint8_t)SABLE_ROUTE_DENSE;
                if (rt == SABLE_ROUTE_SPARSE) sparse.push_back(k);
                else if (span && (rt == SABLE_ROUTE_DENSE || rt == SABLE_ROUTE_DENSE_MFMA)) {
                    // a span's dense-routed blocks are entries with their zeros
                    if (keep.empty()) keep.assign(v.nblocks, 0);
                    keep[k] = 1;
                    sparse.push_back(k);
                    S.nblocks_dense += 1;
                }
                else if (rt == SABLE_ROUTE_DENSE_MFMA && spmm) {
                    if (o.n_rhs % 16 != 0 || o.n_rhs > 256) {
                        err = "the MFMA route needs n_rhs a multiple of 16 and at most 256";
                        return SABLE_ERR_UNSUPPORTED;
                    }
                    mfma.push_back(k);
                }
                else if (rt == SABLE_ROUTE_DENSE || rt == SABLE_ROUTE_DENSE_MFMA) dense.push_back(k);
                else { err = "unknown route value"; return SABLE_ERR_INVALID; }
            }
        }
        const SpanGeo geo{&v, brow.data(), r0, keep.empty() ? nullptr : &keep};
        S.nblocks_dense += (int64_t)dense.size();
        S.nblocks_sparse += (int64_t)sparse.size() - (span ? (int64_t)std::count_if(sparse.begin(), sparse.end(),
                                                                                    [&](int64_t k) { return geo.keeps(k); })
                                                              : 0);
        S.nblocks_mfma += (int64_t)mfma.size();
        // FP32 SpMM pieces stage their X rows in LDS (SB_SPMM_XS_FLOATS)
        const int vcap = spmm ? std::max(1, std::min(o.xcap, SB_SPMM_XS_FLOATS / o.n_rhs)) : o.xcap;
        auto dgroups = group_columns(dense, v, vcap);
        auto sgroups = group_columns(sparse, v, vcap);
        // MFMA columns: with k chunks, groups of at most kper 32-column
        // k-steps, kper = ceil(K / ceil(K / kchunk)) for the block row's K
        // k-steps, so its chunks are even; each group is then one work item
        int mcap = o.xcap;
        const bool ksplit = spmm && !mfma.empty() && o.mfma_kchunk > 0;
        if (ksplit) {
            int64_t W = 0;
            for (int64_t k : mfma) W += v.cpntr[v.bindx[k] + 1] - v.cpntr[v.bindx[k]];
            const int64_t K = (W + 31) / 32;
            const int64_t nc = std::max<int64_t>(1, (K + o.mfma_kchunk - 1) / o.mfma_kchunk);
            const int64_t kper = (K + nc - 1) / nc;
            mcap = (int)std::min<int64_t>((o.xcap / 32) * 32, 32 * kper);
        }
        auto mgroups = group_columns(mfma, v, mcap);

        // sparse entries of the row span (for costing)
        int64_t sp_total = 0;
        for (int64_t k : sparse) {
            const float *bv = v.val + v.indx[k];
            const int64_t n = (int64_t)(v.cpntr[v.bindx[k] + 1] - v.cpntr[v.bindx[k]]) * geo.height(k);
            if (geo.keeps(k)) sp_total += n;
            else for (int64_t i = 0; i < n; ++i) sp_total += (bv[i] != 0.0f);
        }
        int64_t wd = 0, wm = 0;
        for (auto &g : dgroups) wd += group_width(g);
        for (auto &g : mgroups) wm += group_width(g);
        const double work = 4.0 * h * (double)(wd + wm) + 6.0 * sp_total + 4.0 * (h + 1) * sgroups.size();

        // Emits the tile of rows [s0, s0+hs) of this block row with the
        // pieces of the selected routes (vector routes: dense FP32 and
        // sparse; MFMA).  y_accum: the tile adds to Y rows an earlier MFMA
        // tile of the same rows wrote.
        const std::vector<std::vector<ColSeg>> none;
        auto emit = [&](int32_t s0, int32_t hs, bool use_vec, bool use_mfma, int32_t y_accum, int mg_only = -1) {
            SbTile t{};
            t.row0 = r0 + s0;
            t.nrows = hs;
            t.mat = 0;
            t.y_accum = y_accum;
            t.ksplit = 1;
            t.kpart = 0;
            t.part_blk = 0;
            // sparse rows of the segment, per sparse column group
            std::vector<SparseRows> sp;
            for (size_t gi = 0; use_vec && gi < sgroups.size(); ++gi) {
                SparseRows sr;
                if (collect_sparse_rows(v, sgroups[gi], geo, s0, hs, sr)) {
                    sr.group = (int32_t)gi;
                    sp.push_back(std::move(sr));
                }
            }
            // SpMV with dense and sparse pieces: the last sp_waves wavefronts
            // take the sparse slices while the others stream the dense
            // pieces.  Both sides are chains of dependent loads, so the split
            // k and the sparse lanes per row G are chosen together for the
            // shortest longest chain, in round trips: a dense lane's float4
            // loads / (unroll * batches in flight) against the busiest sparse
            // wavefront's steps / (SB_SPUNROLL * batches in flight) plus two
            // per slice (its table entry, then its first batch); among
            // choices within 5 % of it, the one with the fewest padded
            // sparse slots.  The dense stream needs at least hs lanes.
            const int nw = o.wg / 64;
            t.sp_waves = 0;
            int sp_glog = -1;
            if (!spmm && !sp.empty() && use_vec && !dgroups.empty() && nw >= 2) {
                int64_t dq = 0;
                for (auto &g : dgroups) dq += ((int64_t)hs * group_width(g) + 3) / 4;
                const std::vector<int32_t> order = sparse_row_order(sp, hs);
                struct Choice { int k, glog; double chain; int64_t padded; };
                std::vector<Choice> cs;
                for (int k = 1; k < nw && 64 * (nw - k) >= hs; ++k) {
                    int32_t tp = 0, mc = 0;
                    dense_stream_params(std::min<int32_t>(hs, 64 * (nw - k)), 64 * (nw - k), tp, mc);
                    const double dch = std::ceil((double)dq / std::max(tp, 1)) /
                                       (std::max(o.unroll, 1) * SB_SPMV_INFLIGHT);
                    constexpr int per_rt = SB_SPUNROLL * SB_SPMV_INFLIGHT;   // sparse steps per round trip
                    for (int g = 0; g <= 6; ++g) {
                        const auto pw = slice_widths(sp, order, hs, g, k, nullptr, 2 * per_rt);
                        cs.push_back({k, g, std::max(dch, (double)pw.second / per_rt), pw.first});
                    }
                }
                if (!cs.empty()) {
                    double best = cs[0].chain;
                    for (auto &c : cs) best = std::min(best, c.chain);
                    const Choice *pick = nullptr;
                    for (auto &c : cs)
                        if (c.chain <= 1.05 * best && (!pick || c.padded < pick->padded)) pick = &c;
                    t.sp_waves = pick->k;
                    sp_glog = pick->glog;
                }
            }
            const int dlanes = 64 * (nw - t.sp_waves);
            dense_stream_params(std::min<int32_t>(hs, dlanes), dlanes, t.dense_tpar, t.dense_mcols);
            t.piece_begin = (int32_t)P.pieces.size();
            double cost = 64.0;
            // dense FP32 pieces
            for (auto &g : (use_vec ? dgroups : none)) {
                SbPiece pc{};
                pc.kind = SB_PIECE_DENSE;
                pc.ncols = group_width(g);
                // SpMM: columns padded to a multiple of 4 rows (float4 row quads)
                pc.pitch = spmm ? (hs + 3) / 4 * 4 : hs;
                while (P.dval.size() % 4) P.dval.push_back(0.f);
                pc.data_off = (int64_t)P.dval.size();
                for (auto &cs : g) {
                    const float *bv = v.val + v.indx[cs.k];
                    for (int32_t c = cs.cb; c < cs.cb + cs.n; ++c) {
                        const float *col = bv + (int64_t)c * h + s0;
                        P.dval.insert(P.dval.end(), col, col + hs);
                        P.dval.insert(P.dval.end(), pc.pitch - hs, 0.f);
                    }
                }
                while (P.dval.size() % 4) P.dval.push_back(0.f);
                P.dval.insert(P.dval.end(), 4, 0.f);   // the zero float4 read past the end
                S.dense_values += (int64_t)hs * pc.ncols;
                push_xranges(P, g, pc);
                P.pieces.push_back(pc);
                cost += 4.0 * hs * pc.ncols;
            }
            // MFMA pieces (SpMM): FP32 in A-operand register order
            for (size_t gi = 0; use_mfma && gi < mgroups.size(); ++gi) {
                if (mg_only >= 0 && (int)gi != mg_only) continue;
                const auto &g = mgroups[gi];
                SbPiece pc{};
                pc.kind = SB_PIECE_MFMA;
                pc.ncols = group_width(g);
                pc.ksteps = (pc.ncols + 31) / 32;
                const int32_t mt = (hs + 15) / 16;
                while (P.dval.size() % 8) P.dval.push_back(0.f);
                pc.data_off = (int64_t)P.dval.size();
                const int64_t n = (int64_t)mt * pc.ksteps * 64 * 8;
                P.dval.resize(P.dval.size() + n, 0.f);
                float *dst = P.dval.data() + pc.data_off;
                int32_t c = 0;
                for (auto &cs : g) {
                    const float *bv = v.val + v.indx[cs.k];
                    for (int32_t cc = cs.cb; cc < cs.cb + cs.n; ++cc, ++c) {
                        const float *col = bv + (int64_t)cc * h + s0;
                        const int32_t ks = c / 32, kk = c % 32;
                        for (int32_t r = 0; r < hs; ++r) {
                            const int32_t lane = (r % 16) + 16 * (kk / 8);
                            dst[((int64_t)(r / 16) * pc.ksteps + ks) * 512 + ((kk % 8) / 4) * 256 + lane * 4 +
                                (kk % 4)] = col[r];
                        }
                    }
                }
                S.mfma_values += (int64_t)hs * pc.ncols;
                S.mfma_padded += n;
                push_xranges(P, g, pc);
                P.pieces.push_back(pc);
                cost += 4.0 * hs * pc.ncols;
            }
            t.sp_glog = 0;
            t.sp_map = -1;
            if (!sp.empty()) {
                if (spmm) {
                    for (auto &sr : sp) {
                        SbPiece pc{};
                        pc.kind = SB_PIECE_SPARSE;
                        pc.ncols = group_width(sgroups[sr.group]);
                        pc.data_off = (int64_t)P.sval.size();
                        pc.rowptr_off = (int32_t)P.srowptr.size();
                        P.srowptr.insert(P.srowptr.end(), sr.rp.begin(), sr.rp.end());
                        P.sval.insert(P.sval.end(), sr.val.begin(), sr.val.end());
                        P.sidx.insert(P.sidx.end(), sr.col.begin(), sr.col.end());
                        const int64_t tot = (int64_t)sr.val.size();
                        S.sparse_nnz += tot;
                        S.rowptr_entries += hs + 1;
                        push_xranges(P, sgroups[sr.group], pc);
                        P.pieces.push_back(pc);
                        cost += 6.0 * tot + 4.0 * (hs + 1);
                    }
                } else {
                    cost += emit_wave_slices(P, S, t, sp, sgroups, hs, t.sp_waves ? t.sp_waves : nw, sp_glog);
                }
            }
            t.piece_end = (int32_t)P.pieces.size();
            if (!spmm) {
                // SpMV: sparse pieces first in every staging group
                std::stable_partition(P.pieces.begin() + t.piece_begin, P.pieces.begin() + t.piece_end,
                                      [](const SbPiece &pc) { return pc.kind == SB_PIECE_SPARSE; });
                stage_groups(P, t);
                // the first group's map, copied into the tile
                if (t.piece_end > t.piece_begin) {
                    t.stage_map = P.pieces[t.piece_begin].xmap_off;
                    t.stage_len = P.pieces[t.piece_begin].stage_len;
                }
            }
            if (spmm) cost += 2.0 * hs * P.n_rhs;   // Y write share
            P.tiles.push_back(t);
            P.tile_cost.push_back(cost);
            P.tile_phase.push_back(use_mfma && !use_vec ? 0 : 1);
        };
        // segment starts: nseg segments, optionally at multiples of 16 rows
        auto cut = [&](int64_t nseg, bool by16) {
            std::vector<int32_t> st;
            if (by16) {
                const int64_t m16 = (h + 15) / 16;
                nseg = std::max<int64_t>(1, std::min<int64_t>(nseg, m16));
                for (int64_t s = 0; s < nseg; ++s) st.push_back((int32_t)(16 * ((m16 * s) / nseg)));
            } else {
                nseg = std::max<int64_t>(1, std::min<int64_t>(nseg, h));
                for (int64_t s = 0; s < nseg; ++s) st.push_back((int32_t)((h * s) / nseg));
            }
            st.push_back(h);
            return st;
        };
        auto vec_segments = [&](double work_v) {
            int64_t nseg = (h + max_rows - 1) / max_rows;
            const int64_t want = (int64_t)std::ceil(work_v / (double)o.tile_bytes);
            return cut(std::max<int64_t>(nseg, std::min<int64_t>(want, (h + min_rows - 1) / min_rows)), false);
        };
        if (spmm && !mgroups.empty()) {
            // MFMA tiles (first launch): tall, multiples of 16 rows, only the
            // MFMA pieces, writing Y; then tiles of the other pieces (second
            // launch) adding to Y
            int64_t nseg = (h + mfma_rows - 1) / mfma_rows;
            const int64_t want = (int64_t)std::ceil(4.0 * h * (double)wm / (double)o.tile_bytes);
            const auto st = cut(std::max(nseg, want), true);
            const int32_t nc = (int32_t)mgroups.size();
            for (size_t s = 0; s + 1 < st.size(); ++s) {
                if (st[s + 1] <= st[s]) continue;
                const int32_t hs = st[s + 1] - st[s];
                if (!ksplit || nc == 1) {
                    emit(st[s], hs, false, true, 0);
                    continue;
                }
                // one work item per k chunk; scratch for nc partials of
                // (16-row units x N/16 column tiles) 1 KiB blocks; the row
                // tile itself goes to the fixup list
                const int64_t blocks = (int64_t)((hs + 15) / 16) * (P.n_rhs / 16);
                if (P.part_blocks + nc * blocks > INT32_MAX) {
                    err = "partial-sum scratch exceeds 2^31 KiB";
                    return SABLE_ERR_UNSUPPORTED;
                }
                for (int32_t c = 0; c < nc; ++c) {
                    emit(st[s], hs, false, true, 0, c);
                    SbTile &tt = P.tiles.back();
                    tt.ksplit = nc;
                    tt.kpart = c;
                    tt.part_blk = (int32_t)P.part_blocks;
                }
                SbTile ft{};
                ft.row0 = r0 + st[s];
                ft.nrows = hs;
                ft.ksplit = nc;
                ft.part_blk = (int32_t)P.part_blocks;
                P.fix_tiles.push_back(ft);
                P.part_blocks += nc * blocks;
                S.ntiles_split += 1;
                S.fixup_bytes += (2 * nc) * blocks * 1024;
            }
            if (!dgroups.empty() || !sgroups.empty()) {
                const auto sv = vec_segments(4.0 * h * (double)wd + 6.0 * sp_total + 4.0 * (h + 1) * sgroups.size());
                for (size_t s = 0; s + 1 < sv.size(); ++s)
                    if (sv[s + 1] > sv[s]) emit(sv[s], sv[s + 1] - sv[s], true, false, 1);
            }
        } else {
            const auto sv = vec_segments(work);
            for (size_t s = 0; s + 1 < sv.size(); ++s)
                if (sv[s + 1] > sv[s]) emit(sv[s], sv[s + 1] - sv[s], true, false, 0);
        }
    }
    // largest tiles first (LPT order for the persistent queue)
    std::vector<int32_t> order(P.tiles.size());
    std::iota(order.begin(), order.end(), 0);
    // MFMA tiles (first launch) before the others; largest first in each
    std::stable_sort(order.begin(), order.end(), [&](int32_t x, int32_t y) {
        if (P.tile_phase[x] != P.tile_phase[y]) return P.tile_phase[x] < P.tile_phase[y];
        return P.tile_cost[x] > P.tile_cost[y];
    });
    std::vector<SbTile> tiles(P.tiles.size());
    std::vector<double> cost(P.tiles.size());
    for (size_t i = 0; i < order.size(); ++i) {
        tiles[i] = P.tiles[order[i]];
        cost[i] = P.tile_cost[order[i]];
    }
    P.tiles.swap(tiles);
    P.tile_cost.swap(cost);
    S.ntiles_mfma = 0;
    for (int32_t ph : P.tile_phase) S.ntiles_mfma += (ph == 0);
    std::sort(P.tile_phase.begin(), P.tile_phase.end());

    S.ntiles = (int32_t)P.tiles.size();
    S.npieces = (int32_t)P.pieces.size();
    S.nxranges = (int32_t)P.xr.size();
    S.dense_padded = (int64_t)P.dval.size() - S.mfma_padded;
    S.sparse_padded = (int64_t)P.sval.size();
    if (P.xmap.size() > (size_t)INT32_MAX) { err = "x staging map exceeds 2^31 entries"; return SABLE_ERR_UNSUPPORTED; }
    S.xmap_entries = (int64_t)P.xmap.size();
    // what the kernels read: SpMV stages x through xmap, SpMM through the ranges
    S.descriptor_bytes = (int64_t)(P.tiles.size() * sizeof(SbTile) + P.pieces.size() * sizeof(SbPiece)) +
                         (spmm ? (int64_t)(P.xr.size() * sizeof(SbXRange)) : 4 * S.xmap_entries);
    S.bytes_alg = 4 * S.dense_values + 4 * S.mfma_values + 6 * S.sparse_nnz + 4 * S.rowptr_entries +
                  S.descriptor_bytes + 4LL * P.ncols * P.n_rhs + 4LL * P.nrows * P.n_rhs;
    S.flops_alg = 2 * S.nnz * (int64_t)P.n_rhs;
    S.scratch_bytes = P.part_blocks * 1024;
    S.wg = o.wg;
    S.mode = o.mode;
    return SABLE_OK;
}
}  // namespace

}  // namespace sable
