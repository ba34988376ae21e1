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


@pytest.mark.parametrize("eps,n_core", [(1e-300, 0), (2e-4, 0), (2e-4, 1)])
def test_external_product_oracle_matches_dense(eps, n_core):
    """matmul_external_dtm (approx_matmul_external_dtm_only, src/wf/mod.rs:186-339)
    against the dense matrix: every external determinant the screened rule
    reaches (doubles: |H_ai| >= eps/|c_i|; singles: the exact |H_ai| >= eps/|c_i|),
    with coefficient sum_i H_ai c_i over the pairs that pass; determinants in
    first-discovery order (the first discoverer never decreases); and the
    remainder c_i^2 * sum over cut lists of the squared magnitudes after the
    first entry below eps/|c_i|, from the list magnitudes directly."""
    core = 0
    ints, fci, h_full, var, in_space, e0, c = _system(8, 3, 120, 23, core if n_core else None)
    words, vals, first, rem = pt2_oracle.matmul_external_dtm(ints, var, c, eps, n_core, 3, 3)
    up, dn = fci[:, 0], fci[:, 2]
    order = sorted(range(ints.n_orb), key=lambda p: ints.h1(p, p))
    valence = sum(1 << p for p in order[n_core:])
    deg_ok = np.zeros((fci.shape[0], in_space.size), dtype=bool)
    single = np.zeros_like(deg_ok)
    for k, i in enumerate(in_space):
        du = np.array([bin(int(x)).count("1") for x in (up ^ up[i])]) // 2
        dd = np.array([bin(int(x)).count("1") for x in (dn ^ dn[i])]) // 2
        # excitations only among valence orbitals: the changed orbitals are valence
        chg = np.array([(int(a) ^ int(up[i])) | (int(b) ^ int(dn[i])) for a, b in zip(up, dn)], dtype=object)
        inval = np.array([(x & ~valence) == 0 for x in chg])
        deg_ok[:, k] = (du + dd >= 1) & (du + dd <= 2) & inval
        single[:, k] = (du + dd == 1)
    hv = h_full[:, in_space]
    local = eps / np.abs(c)[None, :]
    keep = deg_ok & (np.abs(hv) >= local)
    v = np.sum(np.where(keep, hv * c[None, :], 0.0), axis=1)
    out = ~np.isin(np.arange(fci.shape[0]), in_space) & keep.any(axis=1)
    index = {tuple(w): k for k, w in enumerate(fci)}
    got = np.array([index[tuple(w)] for w in words])
    assert set(got.tolist()) == set(np.nonzero(out)[0].tolist())
    assert np.allclose(vals, v[got], rtol=1e-12, atol=1e-15)
    srcs = [k for k, _ in first]
    assert all(a <= b for a, b in zip(srcs, srcs[1:]))
    # remainders from the list magnitudes
    _, opp, same, sing, mo, ms, mg = pt2_oracle._excite_lists(ints, n_core, 3, 3)

    def cut(lst, loc):
        a = np.array([e[0] for e in lst])
        below = a[a < loc]
        return float(np.sum(below ** 2) - below.max() ** 2) if below.size else 0.0

    for k, w in enumerate(var):
        u, d, ck = int(w[0]), int(w[2]), float(c[k])
        loc = eps / abs(ck)
        ref = 0.0
        bu = [p for p in range(ints.n_orb) if (valence & u) >> p & 1]
        bd = [p for p in range(ints.n_orb) if (valence & d) >> p & 1]
        if mo >= loc:
            ref += sum(cut(opp[(i, j)], loc) for i in bu for j in bd)
        if ms >= loc:
            for b in (bu, bd):
                ref += sum(cut(same[(b[x], b[y])], loc) for x in range(len(b)) for y in range(x + 1, len(b)))
        if mg >= loc:
            ref += sum(cut(sing[i], loc) for i in bu + bd)
        assert abs(rem[k] - ck * ck * ref) <= 1e-12 * max(1.0, abs(ck * ck * ref)), (k, rem[k], ck * ck * ref)
