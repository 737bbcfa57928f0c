"""The native plan builder on the CPU (host-only plans): the HBM layout it
produces must reconstruct the matrix exactly, cover every row once, respect
the kernels' launch assumptions, and reject malformed VBR input."""
import os
import re

import numpy as np
import pytest

from sable_amd import synth
from sable_amd.convert import vbr_to_dense, threshold_routes, ROUTE_DENSE, ROUTE_SPARSE

_lib = pytest.importorskip("sable_amd._lib")
from sable_amd.runtime import Plan  # noqa: E402
from sable_amd._lib import SableError  # noqa: E402

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def bf16_rne(x):
    u = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return u.astype(np.uint16)


def bf16_to_f32(h):
    return (h.astype(np.uint32) << 16).view(np.float32)


def slice_rows(t, p, srow, sval, sidx):
    """Decodes one SpMV wave-sliced sparse piece into {local row: (cols, vals)}
    and checks the layout: every row of the tile appears once in the row
    order, slices of C = 64/G rows are back to back, each row's entries are a
    prefix of its lane chunks, and all other slots are padding (value 0 at the
    zero pad column ncols)."""
    hs, glog = int(t["nrows"]), int(t["sp_glog"])
    G, C = 1 << glog, 64 >> glog
    nsl = (hs + C - 1) // C
    order = srow[t["sp_map"]:t["sp_map"] + hs]
    assert sorted(order.tolist()) == list(range(hs))
    tab = srow[p["rowptr_off"]:p["rowptr_off"] + 2 * nsl].reshape(nsl, 2)
    off = int(p["data_off"])
    assert off % 256 == 0
    assert np.array_equal(tab[:, 0], np.concatenate([[0], np.cumsum(tab[:-1, 1])]))
    ncols = int(p["ncols"])
    rows = {}
    for w in range(nsl):
        o, W = int(tab[w, 0]), int(tab[w, 1])
        blk_v = sval[off + o * 256: off + (o + W) * 256].reshape(W, 64, 4)
        blk_i = sidx[off + o * 256: off + (o + W) * 256].reshape(W, 64, 4)
        for q in range(C):
            lanes = slice(q * G, q * G + G)
            # chunk c = j*G + g -> entries in row order
            v = blk_v[:, lanes, :].reshape(-1)
            c = blk_i[:, lanes, :].reshape(-1)
            real = c < ncols
            n = int(real.sum())
            assert np.all(real[:n]) and not np.any(real[n:]), "entries must be a prefix"
            assert np.all(v[n:] == 0) and np.all(c[n:] == ncols)
            if w * C + q < hs:
                rows[int(order[w * C + q])] = (c[:n].astype(np.int64), v[:n])
            else:
                assert n == 0
    return rows


def reconstruct(plan):
    tiles = plan.export("tiles")
    pieces = plan.export("pieces")
    xr = plan.export("xranges")
    dval = plan.export("dval")
    sval = plan.export("sval")
    sidx = plan.export("sidx")
    srow = plan.export("srowptr")
    M = np.zeros((plan.nrows, plan.ncols), dtype=np.float64)
    Mb = np.zeros_like(M)   # bf16 hi+lo parts
    cover = np.zeros(plan.nrows, dtype=np.int64)      # tiles that write each Y row
    mfma_rows = np.zeros(plan.nrows, dtype=bool)
    nm = plan.stats["ntiles_mfma"]
    for ti, t in enumerate(tiles):
        r0, hs = int(t["row0"]), int(t["nrows"])
        kinds = {int(pieces[k]["kind"]) for k in range(t["piece_begin"], t["piece_end"])}
        if ti < nm:
            # SpMM MFMA tiles come first, hold only MFMA pieces, 16-row multiples
            assert kinds <= {2} and t["y_accum"] == 0 and (r0 - 0) >= 0
            mfma_rows[r0:r0 + hs] = True
        else:
            assert 2 not in kinds
        if t["y_accum"]:
            assert ti >= nm and np.all(mfma_rows[r0:r0 + hs])   # adds to rows an MFMA tile wrote
        else:
            cover[r0:r0 + hs] += 1
        for p in pieces[t["piece_begin"]:t["piece_end"]]:
            cols = np.concatenate([np.arange(c0, c0 + n) for c0, n in xr[p["xr_begin"]:p["xr_end"]]])
            assert len(cols) == p["ncols"]
            off = int(p["data_off"])
            if p["kind"] == 0:
                assert off % 4 == 0
                pitch = int(p["pitch"])
                # SpMM panels pad each column to a multiple of 4 rows with zeros
                assert pitch == ((hs + 3) // 4 * 4 if plan.op == "spmm" else hs)
                pan = dval[off:off + pitch * len(cols)].reshape(len(cols), pitch)
                assert np.all(pan[:, hs:] == 0)
                M[r0:r0 + hs][:, cols] += pan[:, :hs].T
            elif p["kind"] == 1 and plan.op == "spmv":
                for r, (lc, v) in slice_rows(t, p, srow, sval, sidx).items():
                    assert np.all(np.diff(lc) > 0)  # ascending columns in a row
                    M[r0 + r, cols[lc]] += v
            elif p["kind"] == 1:
                rp = srow[p["rowptr_off"]:p["rowptr_off"] + hs + 1]
                for r in range(hs):
                    e = slice(off + rp[r], off + rp[r + 1])
                    lc = sidx[e].astype(np.int64)
                    assert np.all(np.diff(lc) > 0)  # ascending columns in a row
                    M[r0 + r, cols[lc]] += sval[e]
            else:
                assert off % 8 == 0
                ks = int(p["ksteps"])
                assert ks == (len(cols) + 31) // 32
                mt = (hs + 15) // 16
                n = mt * ks * 64 * 8
                seen = np.zeros(n, dtype=bool)
                for r in range(hs):
                    for c in range(len(cols)):
                        lane = r % 16 + 16 * ((c % 32) // 8)
                        o = ((r // 16) * ks + c // 32) * 512 + ((c % 8) // 4) * 256 + lane * 4 + c % 4
                        Mb[r0 + r, cols[c]] += dval[off + o]
                        seen[o] = True
                # every slot not holding a value is zero padding
                assert np.all(dval[off:off + n][~seen] == 0)
    return M, Mb, cover


def _small_mixed(seed, n=900, parts=12):
    vbr, x, cls = synth.mixed_vbr(n, parts, 20, 140, 0.3, 0.5, (0.8, 1.0), (0.02, 0.3), seed)
    return vbr, x


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_host_plan_reconstructs_matrix(seed):
    vbr, x = _small_mixed(seed)
    routes = threshold_routes(vbr.val, vbr.indx, 15)
    plan = Plan.from_vbr(vbr, routes, host_only=True, tile_bytes=4096)
    M, _, cover = reconstruct(plan)
    D = vbr_to_dense(vbr.val, vbr.indx, vbr.bindx, vbr.rpntr, vbr.cpntr, vbr.bpntrb, vbr.bpntre)
    assert np.array_equal(M, D.astype(np.float32).astype(np.float64))
    assert np.all(cover == 1)
    st = plan.stats
    assert st["nnz"] == int(np.count_nonzero(vbr.val))
    assert st["sparse_nnz"] == int(np.count_nonzero(vbr.val[np.repeat(routes == ROUTE_SPARSE, np.diff(vbr.indx))]))
    assert st["nblocks_dense"] + st["nblocks_sparse"] == len(vbr.bindx)


def test_host_plan_launch_invariants():
    vbr, x = _small_mixed(5, n=1500, parts=30)
    plan = Plan.from_vbr(vbr, threshold_routes(vbr.val, vbr.indx, 15), host_only=True, tile_bytes=2048)
    for t in plan.export("tiles"):
        hs, tp, M = int(t["nrows"]), int(t["dense_tpar"]), int(t["dense_mcols"])
        wg = plan.stats["wg"]
        assert 1 <= hs <= wg
        assert 4 * tp == M * hs and 0 < tp <= wg
        if t["sp_map"] >= 0:
            assert 0 <= t["sp_glog"] <= 6
        nw = plan.stats["wg"] // 64
        assert 0 <= t["sp_waves"] < nw
        if t["sp_waves"] > 0:
            # split roles: the dense lanes fit the non-sparse wavefronts and cover the rows
            assert tp <= 64 * (nw - t["sp_waves"]) and 64 * (nw - t["sp_waves"]) >= hs
    for p in plan.export("pieces"):
        assert p["ncols"] <= 4096
    check_staging(plan)


def check_staging(plan):
    """SpMV staging: every tile's pieces are covered by consecutive groups of
    at most 4096 + 4 slots, and the slot map of each piece is its x ranges
    followed by four -1 pads, pieces of a tile back to back."""
    tiles, pieces = plan.export("tiles"), plan.export("pieces")
    xr, xmap = plan.export("xranges"), plan.export("xmap")
    assert plan.stats["xmap_entries"] == len(xmap)
    for t in tiles:
        p = int(t["piece_begin"])
        # the first group's map, copied into the tile
        if t["piece_end"] > t["piece_begin"]:
            assert (t["stage_map"], t["stage_len"]) == (pieces[p]["xmap_off"], pieces[p]["stage_len"])
        while p < t["piece_end"]:
            q = int(pieces[p]["grp_end"])
            assert p < q <= t["piece_end"]
            slots = sum(int(pieces[k]["ncols"]) + 4 for k in range(p, q))
            assert pieces[p]["stage_len"] == slots <= 4100
            kinds = [int(pieces[k]["kind"]) for k in range(p, q)]
            assert kinds == sorted(kinds, key=lambda k: k != 1)   # sparse pieces first
            for k in range(p + 1, q):
                assert pieces[k]["grp_end"] == 0
                assert pieces[k]["xmap_off"] == pieces[k - 1]["xmap_off"] + pieces[k - 1]["ncols"] + 4
            p = q
        for k in range(t["piece_begin"], t["piece_end"]):
            pc = pieces[k]
            cols = np.concatenate([np.arange(c0, c0 + n) for c0, n in xr[pc["xr_begin"]:pc["xr_end"]]])
            m = xmap[pc["xmap_off"]:pc["xmap_off"] + pc["ncols"] + 4]
            assert np.array_equal(m[:-4], cols) and np.all(m[-4:] == -1)

def test_host_plan_splits_wide_and_tall_blocks():
    """Blocks wider than the staging capacity are split by columns, block
    rows taller than the workgroup by rows; the layout still reconstructs."""
    rng = np.random.default_rng(3)
    h, w = 300, 70
    val = rng.uniform(-1, 1, h * w).astype(np.float32)
    val[rng.random(h * w) < 0.5] = 0
    rp, cp = [0, h], [0, w]
    args = (val, [0, h * w], [0], rp, cp, [0], [1])
    for route in (ROUTE_DENSE, ROUTE_SPARSE):
        plan = Plan(*args, routes=[route], host_only=True, xcap=16)
        M, _, cover = reconstruct(plan)
        check_staging(plan)
        assert np.array_equal(M, val.reshape(w, h).T.astype(np.float64))
        assert np.all(cover == 1)
        assert all(p["ncols"] <= 16 for p in plan.export("pieces"))


def test_host_plan_empty_rows_and_unit_blocks():
    vbr, x = synth.vbr_matrix_gen_seeded(60, 50, 20, 20, 25, seed=7)   # sizes may be 1
    assert np.any(np.asarray(vbr.bpntrb) == -1)
    plan = Plan.from_vbr(vbr, None, host_only=True)
    M, _, cover = reconstruct(plan)
    D = vbr_to_dense(vbr.val, vbr.indx, vbr.bindx, vbr.rpntr, vbr.cpntr, vbr.bpntrb, vbr.bpntre)
    assert np.array_equal(M, D) and np.all(cover == 1)


def test_short_block_row_spans():
    """SpMV: consecutive block rows of at most 32 rows and 8192 stored values form
    one tile of at most 256 rows (or the workgroup size) and 65536 stored
    values, in which every block takes the sparse path; a dense-routed block
    keeps its zeros as entries.  Spans are left out when they would leave the
    plan below min_tiles.  The plan reconstructs the matrix exactly, writes
    every row once, and counts the entries it stores."""
    vbr, x, _ = synth.mixed_vbr(1500, 300, 1, 10, 0.03, 0.5, (0.8, 1.0), (0.05, 0.3), seed=40)
    vbr.val[: len(vbr.val) // 3][::7] = 0.0          # zeros inside dense-routed blocks too
    nbr = len(vbr.rpntr) - 1
    thr = threshold_routes(vbr.val, vbr.indx, 15).astype(np.uint8)
    D = vbr_to_dense(vbr.val, vbr.indx, vbr.bindx, vbr.rpntr, vbr.cpntr, vbr.bpntrb, vbr.bpntre)
    # with the default min_tiles (1024) spans would leave this small plan with
    # too few tiles to spread over the device: it is built without them
    plan = Plan.from_vbr(vbr, thr, host_only=True)
    assert plan.stats["nspans"] == 0 and plan.stats["ntiles"] >= nbr - int(np.sum(np.asarray(vbr.bpntrb) == -1))
    M, _, cover = reconstruct(plan)
    assert np.array_equal(M, D) and np.all(cover == 1)
    for wg in (64, 256):
        plan = Plan.from_vbr(vbr, thr, wg=wg, min_tiles=1, host_only=True)
        assert plan.stats["nspans"] > 0
        M, _, cover = reconstruct(plan)
        assert np.array_equal(M, D) and np.all(cover == 1)
        tiles = plan.export("tiles")
        assert len(tiles) < nbr // 4
        assert all(t["nrows"] <= min(256, wg) for t in tiles)
        st = plan.stats
        assert st["dense_values"] == 0
        sizes = np.diff(np.asarray(vbr.indx))
        want = sum(int(sizes[k]) if thr[k] == ROUTE_DENSE else
                   int(np.count_nonzero(vbr.val[vbr.indx[k]:vbr.indx[k + 1]])) for k in range(len(thr)))
        assert st["sparse_nnz"] == want
        assert st["nblocks_dense"] == int(np.sum(thr == ROUTE_DENSE))
    # SpMM plans keep one tile per block row segment
    plan = Plan.from_vbr(vbr, thr, op="spmm", n_rhs=16, min_tiles=1, host_only=True)
    assert plan.stats["nspans"] == 0
    assert len(plan.export("tiles")) >= nbr - int(np.sum(np.asarray(vbr.bpntrb) == -1))


def test_host_plan_mfma_layout():
    vbr, x = _small_mixed(11, n=400, parts=6)
    routes = np.where(threshold_routes(vbr.val, vbr.indx, 50) == ROUTE_DENSE, 2, 1).astype(np.uint8)
    plan = Plan.from_vbr(vbr, routes, op="spmm", n_rhs=16, host_only=True)
    M, Mb, cover = reconstruct(plan)
    D = vbr_to_dense(vbr.val, vbr.indx, vbr.bindx, vbr.rpntr, vbr.cpntr, vbr.bpntrb, vbr.bpntre)
    assert np.all(cover == 1)
    assert np.array_equal(M + Mb, D.astype(np.float32).astype(np.float64))
    assert np.count_nonzero(Mb) > 0
    tiles = plan.export("tiles")
    nm = plan.stats["ntiles_mfma"]
    assert nm > 0
    # MFMA tiles: at most 16 * (wg/64) * 2 rows (two 16-row units per wavefront), 16-row multiples
    # except at a block row's end; the vector-route tiles keep the FP32 kernel's 64-row bound
    wg = plan.stats["wg"]
    assert all(t["nrows"] <= min(256, 16 * (wg // 64) * 2) for t in tiles[:nm])
    assert all(t["nrows"] <= 64 for t in tiles[nm:])


def test_bad_vbr_rejected():
    good = dict(val=np.ones(4, np.float32), indx=[0, 4], bindx=[0], rpntr=[0, 2], cpntr=[0, 2],
                bpntrb=[0], bpntre=[1])
    Plan(**good, host_only=True)
    bad = [dict(good, indx=[0, 3]), dict(good, rpntr=[0, 0]), dict(good, bindx=[1]),
           dict(good, bpntrb=[-1]), dict(good, bpntre=[2])]
    for b in bad:
        with pytest.raises((SableError, ValueError)):
            Plan(**b, host_only=True)
    # unsorted bindx within a row
    with pytest.raises(SableError):
        Plan(np.ones(8, np.float32), [0, 4, 8], [1, 0], [0, 2], [0, 2, 4], [0], [2], host_only=True)
    # SpMM width must be a multiple of 4
    with pytest.raises(SableError):
        Plan(**good, op="spmm", n_rhs=6, host_only=True)


def test_library_exports_every_declared_symbol():
    hdr = open(os.path.join(ROOT, "include", "sable.h")).read()
    declared = set(re.findall(r"^(?:int|const char \*)\s*(sable_\w+)\(", hdr, re.M))
    assert declared, "no declarations parsed"
    for name in declared:
        assert hasattr(_lib.lib, name), name
    assert declared <= set(_lib.EXPORTED)
