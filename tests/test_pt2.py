"""The PT2 oracle (oracle/pt2_oracle.py) against closed forms on complete spaces.

With eps below every contribution, dtm_pt (src/pt/mod.rs:81-122) is the
Epstein-Nesbet sum over all determinants outside the space, which a dense
full-CI Hamiltonian gives directly: v = H[:, space] c and
e2 = sum_a v_a^2 / (e0 - H_aa).  With core orbitals the sum runs over the
determinants that keep them doubly occupied.  With a finite eps, the same dense
matrix with every pair (i, a) screened by the reference's rule (doubles:
|H_ai| |c_i| >= eps; singles: |H_ai c_i| >= eps) gives the screened sum."""
import numpy as np
import pytest

from helpers import csr_to_dense
from oracle import pt2_oracle, pyoracle
from risq_amd.dets import energy_truncated_space, full_space
from risq_amd.fcidump import synthetic_molecular


def _system(n_orb, n_el, n_var, seed, core=None):
    ints = synthetic_molecular(n_orb, 2 * n_orb, seed)
    eps = np.array([ints.h1(p, p) for p in range(n_orb)])
    fci = full_space(n_orb, n_el, n_el)
    H = pyoracle.OracleHam(ints)
    S = pyoracle.OracleSparse()
    S.build(H, fci, H.ham_diag(fci))
    h_full = csr_to_dense(*S.export())
    var = energy_truncated_space(eps, n_el, n_el, n_var)
    if core is not None:  # a selected space never empties its core orbitals
        var = var[(var[:, 0] >> np.uint64(core)) & (var[:, 2] >> np.uint64(core)) & np.uint64(1) == 1]
    index = {tuple(w): k for k, w in enumerate(fci)}
    in_space = np.array([index[tuple(w)] for w in var])
    vals, vecs = np.linalg.eigh(h_full[np.ix_(in_space, in_space)])
    return ints, fci, h_full, var, in_space, vals[0], vecs[:, 0]


@pytest.mark.parametrize("n_core", [0, 1])
def test_oracle_matches_dense_epstein_nesbet(n_core):
    core = 0  # h_pp = -1 + 2p/15 + 0: orbital 0 is the lowest
    ints, fci, h_full, var, in_space, e0, c = _system(8, 3, 120, 21, core if n_core else None)
    assert int(np.argmin([ints.h1(p, p) for p in range(ints.n_orb)])) == core
    e2, n_ext = pt2_oracle.dtm_pt(ints, var, c, e0, 1e-300, n_core=n_core)
    allowed = None
    if n_core:
        allowed = ((fci[:, 0] >> np.uint64(core)) & (fci[:, 2] >> np.uint64(core)) & np.uint64(1)) == 1
    ref = pt2_oracle.epstein_nesbet_dense(h_full, in_space, c, e0, allowed)
    assert e2 < 0
    assert abs(e2 - ref) < 1e-11 * abs(ref)
    # every external determinant with a nonzero numerator is counted
    v = h_full[:, in_space] @ c
    out = ~np.isin(np.arange(fci.shape[0]), in_space)
    if allowed is not None:
        out &= allowed
    assert n_ext == int(np.count_nonzero(v[out]))


def test_oracle_screening_matches_dense_rule():
    ints, fci, h_full, var, in_space, e0, c = _system(8, 3, 120, 22)
    eps = 2e-4
    e2, n_ext = pt2_oracle.dtm_pt(ints, var, c, e0, eps)
    # degree of every (a, i) pair: singles screen the product, doubles |H| |c|
    up, dn = fci[:, 0], fci[:, 2]
    deg = np.zeros((fci.shape[0], in_space.size), dtype=int)
    for k, i in enumerate(in_space):
        du = np.array([bin(int(x)).count("1") for x in (up ^ up[i])]) // 2
        dd = np.array([bin(int(x)).count("1") for x in (dn ^ dn[i])]) // 2
        deg[:, k] = du + dd
    hv = h_full[:, in_space]
    prod = hv * c[None, :]
    keep = np.where(deg == 1, np.abs(prod) >= eps, np.abs(hv) * np.abs(c)[None, :] >= eps) & (deg >= 1) & (deg <= 2)
    v = np.sum(np.where(keep, prod, 0.0), axis=1)
    out = ~np.isin(np.arange(fci.shape[0]), in_space)
    contrib = out & keep.any(axis=1)
    ref = float(np.sum(v[contrib] ** 2 / (e0 - np.diag(h_full)[contrib])))
    assert n_ext == int(np.count_nonzero(contrib))
    assert abs(e2 - ref) < 1e-12 * abs(ref)
    # a looser threshold keeps fewer determinants and less energy
    e2b, nb = pt2_oracle.dtm_pt(ints, var, c, e0, 10 * eps)
    assert nb < n_ext and abs(e2b) < abs(e2)
