"""Seeded parity sweep: random small VBR matrices (unit and wide partitions,
non-square blocks, empty block rows, the reference's generator law and the
mixed dense/sparse law) with random routes, workgroup sizes and RHS widths,
SpMV with and without row spans, SpMM on the FP32 and MFMA routes, each
against the FP64 oracle (oracle/c_oracle.py, pinned by tests/test_oracle*.py)."""
import numpy as np
import pytest

from oracle import c_oracle
from sable_amd import synth
from sable_amd.convert import ROUTE_DENSE, ROUTE_SPARSE, ROUTE_DENSE_MFMA

pytestmark = pytest.mark.gpu

TOL = 1e-5
CASES = 24


@pytest.fixture(scope="module")
def rt():
    from sable_amd import runtime
    runtime.init()
    return runtime


def _matrix(seed):
    """(VBR, x) of case seed: the reference generator's law for even seeds,
    the mixed law (with short block rows for every fourth) for odd ones."""
    g = np.random.default_rng([7, seed])
    if seed % 2 == 0:
        m, n = int(g.integers(20, 2500)), int(g.integers(20, 2500))
        rs, cs = int(g.integers(1, min(m, 300) + 1)), int(g.integers(1, min(n, 300) + 1))
        nb = int(g.integers(1, max(2, rs * cs // 3)))
        return synth.vbr_matrix_gen_seeded(m, n, rs, cs, nb, seed=1000 + seed)
    n = int(g.integers(200, 3000))
    lo, hi = (1, 12) if seed % 4 == 1 else (8, 160)
    parts = max(1, min(n // ((lo + hi) // 2), 600))
    vbr, x, _ = synth.mixed_vbr(n, parts, lo, hi, float(g.uniform(0.01, 0.2)), 0.5, (0.8, 1.0), (0.02, 0.3),
                                seed=2000 + seed)
    return vbr, x


def _routes(g, nblocks, mfma):
    choices = [ROUTE_DENSE, ROUTE_SPARSE] + ([ROUTE_DENSE_MFMA] if mfma else [])
    return g.choice(choices, nblocks).astype(np.uint8)


@pytest.mark.parametrize("seed", range(CASES))
def test_random_vbr_sweep(rt, seed):
    vbr, x = _matrix(seed)
    g = np.random.default_rng([8, seed])
    nb = len(vbr.bindx)
    x = np.asarray(x, np.float32)
    y64, ay = c_oracle.spmv_f64_dense_blocks(vbr.val, vbr.indx, vbr.bindx, vbr.rpntr, vbr.cpntr, vbr.bpntrb,
                                             vbr.bpntre, x)
    for min_tiles in (0, 1):
        wg = int(g.choice([64, 256, 512]))
        routes = _routes(g, nb, False)
        print(f"case: seed={seed} spmv n={len(vbr.rpntr) - 1} blocks={nb} wg={wg} min_tiles={min_tiles}", flush=True)
        plan = rt.Plan.from_vbr(vbr, routes, wg=wg, min_tiles=min_tiles)
        y = plan.spmv(x)
        assert c_oracle.scaled_error(y, y64, ay)["componentwise"] <= TOL
    n_rhs = int(g.choice([4, 12, 16, 64, 128]))
    X = g.uniform(-1, 1, (int(vbr.cpntr[-1]), n_rhs)).astype(np.float32)
    Y64, AY = c_oracle.spmm_f64_dense_blocks(vbr.val, vbr.indx, vbr.bindx, vbr.rpntr, vbr.cpntr, vbr.bpntrb,
                                             vbr.bpntre, X)
    wg = int(g.choice([128, 256, 512]))
    routes = _routes(g, nb, n_rhs % 16 == 0)
    print(f"case: seed={seed} spmm n_rhs={n_rhs} wg={wg} mfma={int(np.sum(routes == ROUTE_DENSE_MFMA))}",
          flush=True)
    Y = rt.Plan.from_vbr(vbr, routes, op="spmm", n_rhs=n_rhs, wg=wg).spmm(X)
    assert c_oracle.scaled_error(Y.ravel(), Y64.ravel(), AY.ravel())["componentwise"] <= TOL


GROUP_CASES = 6


@pytest.mark.parametrize("seed", range(GROUP_CASES))
def test_random_group_sweep(rt, seed):
    """Groups of four random members (SpMV or SpMM, random routes, workgroup
    sizes, RHS widths and MFMA k-split chunks, so one group launches several
    kernel classes; the first member is an SpMM at N = 128 in 512-thread
    workgroups, the four-buffer MFMA k-loop): every member's group output is bit-identical to its own
    launch and within TOL of the FP64 oracle."""
    g = np.random.default_rng([9, seed])
    plans, xs, ys, own, ref = [], [], [], [], []
    for i in range(4):
        vbr, x = _matrix(100 + 4 * seed + i)
        nb = len(vbr.bindx)
        if i > 0 and g.random() < 0.4:
            wg = int(g.choice([64, 256, 512]))
            p = rt.Plan.from_vbr(vbr, _routes(g, nb, False), wg=wg, min_tiles=1)
            inp = np.asarray(x, np.float32)
            out = p.spmv(inp)
            y64, ay = c_oracle.spmv_f64_dense_blocks(vbr.val, vbr.indx, vbr.bindx, vbr.rpntr, vbr.cpntr,
                                                     vbr.bpntrb, vbr.bpntre, inp)
            print(f"case: group seed={seed} member {i} spmv wg={wg}", flush=True)
        else:
            n_rhs = 128 if i == 0 else int(g.choice([12, 16, 64, 128]))
            wg = 512 if i == 0 else int(g.choice([128, 256, 512]))
            kchunk = int(g.choice([0, 1, 3]))
            routes = _routes(g, nb, n_rhs % 16 == 0)
            p = rt.Plan.from_vbr(vbr, routes, op="spmm", n_rhs=n_rhs, wg=wg, mfma_kchunk=kchunk, min_tiles=1)
            inp = g.uniform(-1, 1, (int(vbr.cpntr[-1]), n_rhs)).astype(np.float32)
            out = p.spmm(inp)
            y64, ay = c_oracle.spmm_f64_dense_blocks(vbr.val, vbr.indx, vbr.bindx, vbr.rpntr, vbr.cpntr,
                                                     vbr.bpntrb, vbr.bpntre, inp)
            print(f"case: group seed={seed} member {i} spmm n_rhs={n_rhs} wg={wg} kchunk={kchunk} "
                  f"mfma={int(np.sum(routes == ROUTE_DENSE_MFMA))}", flush=True)
        plans.append(p)
        xs.append(rt.DeviceArray.from_numpy(inp))
        ys.append(rt.DeviceArray(out.shape))
        own.append(out)
        ref.append((y64, ay))
    grp = rt.Group(plans, xs, ys)
    grp.launch()
    for y, o, (y64, ay) in zip(ys, own, ref):
        got = y.numpy()
        assert np.array_equal(got, o)
        assert c_oracle.scaled_error(got.ravel(), np.ravel(y64), np.ravel(ay))["componentwise"] <= TOL
    grp.close()
