"""The oracle's gen_sparse_ham_fast_lookup restatement against brute force:
every pair of determinants is classified by popcount(XOR) and stored under the
reference's rules (singles iff |H| > 1e-12, doubles iff H != 0)."""
import os

import numpy as np
import pytest

from helpers import csr_to_dense, hubbard_space, mol_space
from risq_amd.dets import full_space, to_int_pairs
from risq_amd.fcidump import read_fcidump, synthetic_molecular
from sq_brute import ham_matrix

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def brute_pattern(oracle_lib, ints, dets, n_stored=0):
    n = dets.shape[0]
    H = oracle_lib.OracleHam(ints)
    ii, jj = np.triu_indices(n, 1)
    keep = jj >= n_stored
    ii, jj = ii[keep], jj[keep]
    val, deg = H.off_diag(dets[ii], dets[jj])
    store = ((deg == 1) & (np.abs(val) > 1e-12)) | ((deg == 2) & (val != 0.0))
    return ii[store], jj[store], val[store]


def csr_pairs(rp, col, val):
    rows = np.repeat(np.arange(rp.size - 1), np.diff(rp))
    return rows, col.astype(np.int64), val


@pytest.mark.parametrize("space", [(16, 32, 0, 4, 4, 700), (12, 16, 9, 3, 4, 600)])
def test_fast_lookup_pattern_is_brute_force(oracle_lib, space):
    ints, dets = mol_space(*space)
    H = oracle_lib.OracleHam(ints)
    diag = H.ham_diag(dets)
    S = oracle_lib.OracleSparse().build(H, dets, diag)
    rp, col, val, dg = S.export()
    r, c, v = csr_pairs(rp, col, val)
    bi, bj, bv = brute_pattern(oracle_lib, ints, dets)
    order = np.lexsort((bj, bi))
    assert np.array_equal(r, bi[order]) and np.array_equal(c, bj[order])
    assert np.array_equal(v, bv[order])
    assert np.array_equal(dg, diag)
    # rows sorted, no duplicates (sparse.rs:44-56)
    for k in range(rp.size - 1):
        seg = col[rp[k]:rp[k + 1]]
        assert np.all(np.diff(seg) > 0) and np.all(seg > k)


def test_incremental_equals_full(oracle_lib):
    ints, dets = mol_space(16, 32, 0, 4, 4, 700)
    H = oracle_lib.OracleHam(ints)
    diag = H.ham_diag(dets)
    inc = oracle_lib.OracleSparse()
    inc.build(H, dets[:300], diag[:300])
    inc.build(H, dets, diag, n_stored=300)
    full = oracle_lib.OracleSparse().build(H, dets, diag)
    for a, b in zip(inc.export(), full.export()):
        assert np.array_equal(a, b)


def test_hubbard_pattern(oracle_lib):
    ints, dets = hubbard_space(3, 3, 4, 4, 500)
    H = oracle_lib.OracleHam(ints)
    S = oracle_lib.OracleSparse().build(H, dets, H.ham_diag(dets))
    rp, col, val, _ = S.export()
    bi, bj, bv = brute_pattern(oracle_lib, ints, dets)
    assert S.nnz == bi.size and np.all(np.abs(val) == 1.0)


def test_spmv_and_davidson_vs_dense(oracle_lib):
    ints, dets = mol_space(16, 32, 0, 4, 4, 900)
    H = oracle_lib.OracleHam(ints)
    S = oracle_lib.OracleSparse().build(H, dets, H.ham_diag(dets))
    A = csr_to_dense(*S.export())
    x = np.random.default_rng(2).standard_normal(A.shape[0])
    assert np.max(np.abs(S.spmv(x) - A @ x)) < 1e-12
    w = np.linalg.eigvalsh(A)
    rc, e0, c, iters, nsig = S.davidson(tol=1e-6, etol=1e-10, max_iters=300)
    assert rc == 0 and abs(e0 - w[0]) < 1e-8
    assert np.linalg.norm(A @ c - e0 * c) < 1e-5


def test_davidson_reference_control_flow_small(oracle_lib):
    """n <= 11: the search space is the whole matrix (davidson.rs:64-68)."""
    ints, dets = mol_space(16, 32, 0, 4, 4, 9)
    H = oracle_lib.OracleHam(ints)
    S = oracle_lib.OracleSparse().build(H, dets, H.ham_diag(dets))
    w = np.linalg.eigvalsh(csr_to_dense(*S.export()))
    rc, e0, _, _, _ = S.davidson(tol=1e-8, etol=1e-12, max_iters=50)
    assert rc == 0 and abs(e0 - w[0]) < 1e-10


def test_sym_eigen_matches_numpy(oracle_lib):
    rng = np.random.default_rng(6)
    for k in (1, 2, 5, 11):
        a = rng.standard_normal((k, k))
        a = a + a.T
        vals, vecs = oracle_lib.sym_eigen(a)
        assert np.allclose(vals, np.linalg.eigvalsh(a), atol=1e-12)
        assert np.allclose(a @ vecs, vecs * vals, atol=1e-10)


def test_mock_fcidump_full_ci(oracle_lib):
    """The reference's 4-orbital mock FCIDUMP (src/var/fast_ham_tests.rs:39):
    reader placement and a full-CI build against second quantisation."""
    ints = read_fcidump(os.path.join(GOLDEN, "mock_fcidump_4orb.txt"), 4)
    assert ints.eri(0, 0, 0, 0) == 0.5 and ints.eri(0, 1, 0, 1) == 0.1 and ints.eri(2, 3, 3, 2) == 0.05
    assert ints.h1(0, 1) == 0.1 and ints.h1(3, 3) == -0.4 and ints.e_core == 0.0
    dets = full_space(4, 2, 2)
    H = oracle_lib.OracleHam(ints)
    S = oracle_lib.OracleSparse().build(H, dets, H.ham_diag(dets))
    A = csr_to_dense(*S.export())
    h, eri = ints.dense()
    B = ham_matrix(h, eri, ints.e_core, to_int_pairs(dets))
    assert np.max(np.abs(A - B)) < 1e-12
