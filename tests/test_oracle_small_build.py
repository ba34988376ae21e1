"""The oracle's restatement of the reference's n <= 100 build
(gen_sparse_ham_fast with opp_algo = same_algo = 1, src/var/ham_gen.rs:60-292,
374-433, 621-637; add_el_and_spin_flipped, src/var/off_diag.rs:73-102) against
brute force: on a space closed under spin flip, every pair (i < j) with at
least one index >= n_stored that is a single or double excitation is stored
iff H != 0 (off_diag.rs:33; no 1e-12 rule for singles), with the value of
ham_off_diag_no_excite, each row sorted without duplicates."""
import numpy as np
import pytest

from helpers import mol_space, spin_flip_closed, tiny_single_molecular
from risq_amd.dets import full_space


def brute_nonzero(oracle_lib, ints, dets, n_stored=0):
    n = dets.shape[0]
    H = oracle_lib.OracleHam(ints)
    ii, jj = np.triu_indices(n, 1)
    keep = jj >= n_stored
    ii, jj = ii[keep], jj[keep]
    val, deg = H.off_diag(dets[ii], dets[jj])
    store = ((deg == 1) | (deg == 2)) & (val != 0.0)
    order = np.lexsort((jj[store], ii[store]))
    return ii[store][order], jj[store][order], val[store][order]


def rows_of(rp):
    return np.repeat(np.arange(rp.size - 1), np.diff(rp))


def small_build(oracle_lib, ints, dets, n_stored=0, prev=None):
    H = oracle_lib.OracleHam(ints)
    S = prev if prev is not None else oracle_lib.OracleSparse()
    return S.build_small(H, dets, H.ham_diag(dets), n_stored=n_stored)


def test_full_ci_with_tiny_singles_is_brute_force(oracle_lib):
    ints = tiny_single_molecular()
    dets = full_space(8, 3, 3)
    rp, col, val, _ = small_build(oracle_lib, ints, dets).export()
    bi, bj, bv = brute_nonzero(oracle_lib, ints, dets)
    assert np.array_equal(rows_of(rp), bi) and np.array_equal(col, bj)
    assert np.array_equal(val.view(np.uint64), bv.view(np.uint64))
    # the singles of magnitude 1e-13 are there, and the fast-lookup rule drops them
    tiny = np.abs(val) <= 1e-12
    assert tiny.sum() > 0 and np.all(np.abs(val[tiny]) == 1e-13)
    H = oracle_lib.OracleHam(ints)
    fast = oracle_lib.OracleSparse().build(H, dets, H.ham_diag(dets))
    assert fast.nnz == val.size - tiny.sum()


@pytest.mark.parametrize("n_first", [1, 30, 61])
def test_incremental_small_build_is_brute_force(oracle_lib, n_first):
    """Two steps, as the HCI loop takes them: the first n_first determinants,
    then the rest added with n_stored = n_first (only pairs with a new index)."""
    ints, base = mol_space(16, 32, 0, 4, 4, 60)
    dets = spin_flip_closed(base)[:100]
    dets = spin_flip_closed(dets)
    first = spin_flip_closed(dets[:n_first])
    n1 = first.shape[0]
    assert np.array_equal(first, dets[:n1])
    H = oracle_lib.OracleHam(ints)
    diag = H.ham_diag(dets)
    S = oracle_lib.OracleSparse()
    S.build_small(H, dets[:n1], diag[:n1])
    S.build_small(H, dets, diag, n_stored=n1)
    rp, col, val, dg = S.export()
    bi, bj, bv = brute_nonzero(oracle_lib, ints, dets)
    assert np.array_equal(rows_of(rp), bi) and np.array_equal(col, bj)
    assert np.array_equal(val.view(np.uint64), bv.view(np.uint64))
    assert np.array_equal(dg, diag)


def test_small_build_needs_spin_flip_closure(oracle_lib):
    ints, dets = mol_space(16, 32, 0, 4, 4, 40)
    closed = spin_flip_closed(dets)
    # drop one flipped partner of a determinant whose up and down strings differ
    k = next(i for i in range(1, closed.shape[0]) if closed[i, 0] != closed[i, 2])
    partner = np.array([closed[k, 2], closed[k, 3], closed[k, 0], closed[k, 1]], dtype=np.uint64)
    broken = closed[~np.all(closed == partner, axis=1)]
    with pytest.raises(ValueError, match="spin flip"):
        small_build(oracle_lib, ints, broken)
