"""GPU block partitioner (sable_axis_cuts) against the host partitioner and
the oracle (marked gpu): brute force on tiny 0/1 patterns, bit-identical
similarities and cuts on random real-valued matrices for every similarity
kind (empty lines and stored zeros included), one-long lines, an empty
matrix, and the reference's own test matrix tests/Franz8.mtx."""
import gzip
import os

import numpy as np
import pytest
import scipy.io
import scipy.sparse

from oracle import vbr_oracle as O
from sable_amd import autopartition as AP

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def rt():
    from sable_amd import runtime
    runtime.init()
    return runtime


def _random(seed, n, m, density, integer=False):
    rng = np.random.default_rng(seed)
    A = scipy.sparse.random(n, m, density=density, format="csc", random_state=rng,
                            data_rvs=(lambda k: rng.integers(-3, 4, k).astype(float)) if integer
                            else (lambda k: rng.standard_normal(k)))
    A = A.tolil()
    A[:, rng.random(m) < 0.1] = 0            # empty columns and runs of them
    A[rng.random(n) < 0.1, :] = 0            # empty rows
    A = A.tocsc()
    A.eliminate_zeros()
    # a few stored zeros (they count as stored lines, not as nonzero values)
    if A.nnz:
        z = rng.choice(A.nnz, max(1, A.nnz // 50), replace=False)
        A.data[z] = 0.0
    return A


@pytest.mark.parametrize("seed", range(8))
def test_tiny_patterns_match_oracle(seed):
    rng = np.random.default_rng(seed)
    n, m = int(rng.integers(2, 12)), int(rng.integers(2, 12))
    P = (rng.random((n, m)) < rng.uniform(0.1, 0.7)).astype(float)
    P[:, rng.random(m) < 0.2] = 0
    P[rng.random(n) < 0.2, :] = 0
    A = scipy.sparse.csc_matrix(P)
    print(f"case: tiny pattern seed {seed}", flush=True)
    for thr in (0.1, 0.2, 0.5):
        ref = O.cut_indices2(P.tolist(), (P != 0).tolist(), thr)
        assert list(AP.cut_indices2(A, thr, AP.similarity2, device=True)) == list(ref)


@pytest.mark.parametrize("kind", ["s2", "numba", "s1"])
@pytest.mark.parametrize("integer", [False, True])
def test_random_matrices_bit_identical_to_host(kind, integer):
    sim_fn = {"s2": AP.similarity2, "numba": AP.similarity2_numba, "s1": AP.similarity1}[kind]
    for seed, (n, m, d) in enumerate([(3000, 2000, 0.003), (800, 5000, 0.01), (2000, 300, 0.05)]):
        A = _random(seed, n, m, d, integer)
        print(f"case: random {n}x{m} kind {kind} integer {integer}", flush=True)
        for M in (A, A.T.tocsc()):
            cuts, sim = AP._axis_cuts_gpu(M, 0.2, sim_fn, return_sim=True)
            Mh = M.tocsc()
            Mh.sum_duplicates()
            host_sim = AP._pair_similarities(Mh.astype(np.float64), {"s2": "s2", "numba": "numba", "s1": "s1"}[kind])
            both_nan = np.isnan(sim) & np.isnan(host_sim)
            assert np.array_equal(sim[~both_nan], host_sim[~both_nan])
            assert cuts == AP._axis_cuts(M, 0.2, sim_fn)


def test_edge_shapes():
    rng = np.random.default_rng(4)
    for shape in [(1, 40), (40, 1), (1, 1), (7, 0), (0, 5)]:
        P = (rng.random(shape) < 0.5).astype(float)
        A = scipy.sparse.csc_matrix(P, shape=shape)
        print(f"case: edge shape {shape}", flush=True)
        assert AP.cut_indices2(A, 0.2, AP.similarity2, device=True) == AP.cut_indices2(A, 0.2, AP.similarity2)
    Z = scipy.sparse.csc_matrix((30, 20))
    assert AP.cut_indices2(Z, 0.2, AP.similarity2, device=True) == ([0, 20], [0, 30])


def test_franz8_matches_reference_algorithm(golden):
    with gzip.open(os.path.join(golden, "Franz8.mtx.gz"), "rt") as f:
        A = scipy.sparse.csc_matrix(scipy.io.mmread(f))
    print("case: Franz8", flush=True)
    cols, rows = AP.cut_indices2(A, 0.2, AP.similarity2, device=True)
    rcols, rrows = O.cut_indices2_dense_lines(A, 0.2)
    assert cols == rcols and rows == rrows


def test_dense_to_vbrc_pipeline_on_device(golden, tmp_path):
    """my_convert_dense_to_vbrc (partition, materialisation, compression) with
    device=True returns what the host pipeline returns, on Franz8."""
    src = tmp_path / "Franz8.mtx"
    with gzip.open(os.path.join(golden, "Franz8.mtx.gz"), "rt") as f:
        src.write_text(f.read())
    print("case: Franz8 pipeline", flush=True)
    outs = []
    for dev in (False, True):
        d = tmp_path / ("dev" if dev else "host")
        d.mkdir()
        outs.append(AP.my_convert_dense_to_vbrc((str(src), str(d / "Franz8.vbrc")), 0.2, AP.cut_indices2,
                                                AP.similarity2, allow_empty_rows=True, device=dev))
    for h, g in zip(*outs):
        assert np.array_equal(np.asarray(h), np.asarray(g))
