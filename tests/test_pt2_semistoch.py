"""The oracle of the reference's default PT2 deterministic stage
(pt2_oracle.semistoch_dtm, restating new_semistoch_enpt2_dtm_diag_singles,
src/semistoch/mod.rs:262-318) against checks that do not share its loop:

* h_psi: the lists are sorted by magnitude, so the entries read before a
  list's cut are exactly its valid external entries with magnitude * |c_i| >=
  eps -- dtm_pt's candidate set (src/excite/iterator.rs:33-73).  h_psi then
  holds the doubles among them and the singles with |H_ai c_i| >= eps, i.e.
  the determinants and numerators of dtm_pt, whose oracle is pinned by dense
  Epstein-Nesbet sums (tests/test_pt2.py); at eps below every term both are
  the dense sum itself;
* rel_probs[i] = max |H_ai c_i| over those entries, from the dense full-CI
  matrix (doubles: their magnitude is |H_ai|; singles: their stored bound);
* the sample records: for each determinant and each of its valence pairs and
  electrons, the cut is the first valid external entry below the threshold,
  found here with vectorised masks over the list, and its sums are the
  magnitudes (and squares) after it."""
import numpy as np
import pytest

from helpers import csr_to_dense
from oracle import pt2_oracle, pyoracle
from risq_amd.dets import energy_truncated_space, full_space
from risq_amd.fcidump import synthetic_molecular


def _system(n_orb, n_el, n_var, seed):
    ints = synthetic_molecular(n_orb, 2 * n_orb, seed)
    eps = np.array([ints.h1(p, p) for p in range(n_orb)])
    fci = full_space(n_orb, n_el, n_el)
    H = pyoracle.OracleHam(ints)
    S = pyoracle.OracleSparse()
    S.build(H, fci, H.ham_diag(fci))
    h_full = csr_to_dense(*S.export())
    var = energy_truncated_space(eps, n_el, n_el, n_var)
    index = {tuple(int(x) for x in w): k for k, w in enumerate(fci)}
    in_space = np.array([index[tuple(int(x) for x in w)] for w in var])
    vals, vecs = np.linalg.eigh(h_full[np.ix_(in_space, in_space)])
    return ints, fci, h_full, var, in_space, index, vals[0], vecs[:, 0]


def test_semistoch_dtm_at_tiny_eps_is_dense_epstein_nesbet():
    ints, fci, h_full, var, in_space, index, e0, c = _system(8, 3, 120, 31)
    words, nums, firsts, rel, records, e2 = pt2_oracle.semistoch_dtm(ints, var, c, e0, 1e-300, 0, 3, 3)
    ref = pt2_oracle.epstein_nesbet_dense(h_full, in_space, c, e0)
    assert abs(e2 - ref) < 1e-11 * abs(ref)
    v = h_full[:, in_space] @ c
    out = ~np.isin(np.arange(fci.shape[0]), in_space)
    assert words.shape[0] == int(np.count_nonzero(v[out]))
    # numerators: H[a, space] c for every external a
    rows = np.array([index[tuple(int(x) for x in w)] for w in words])
    assert np.max(np.abs(nums - v[rows])) < 1e-13
    # rel_probs: max over external a of |H_ai c_i|
    ext = np.nonzero(out)[0]
    want = np.max(np.abs(h_full[np.ix_(ext, in_space)] * c[None, :]), axis=0)
    assert np.allclose(rel, want, rtol=1e-12, atol=0)
    # lists are cut only at zero magnitudes here: nothing remains after them
    assert all(r[5] == 0.0 and r[6] == 0.0 for r in records)


@pytest.mark.parametrize("eps", [3e-5, 3e-4])
def test_semistoch_dtm_h_psi_is_dtm_pt(eps):
    ints, fci, h_full, var, in_space, index, e0, c = _system(8, 3, 150, 32)
    words, nums, firsts, rel, records, e2 = pt2_oracle.semistoch_dtm(ints, var, c, e0, eps, 0, 3, 3)
    o2, on = pt2_oracle.dtm_pt(ints, var, c, e0, eps)
    assert words.shape[0] == on
    assert abs(e2 - o2) < 1e-12 * abs(o2)


def test_semistoch_dtm_rel_probs_and_records_from_the_lists():
    ints, fci, h_full, var, in_space, index, e0, c = _system(8, 3, 150, 33)
    eps = 2e-4
    words, nums, firsts, rel, records, e2 = pt2_oracle.semistoch_dtm(ints, var, c, e0, eps, 0, 3, 3)
    valence, opp, same, sing, *_ = pt2_oracle._excite_lists(ints, 0, 3, 3)
    space = {(int(w[0]), int(w[2])) for w in var}
    H = pyoracle.OracleHam(ints)
    got = {(r[0], r[1], r[2], r[3], r[4]): r[5:] for r in records}
    want_rel = np.zeros(len(var))
    n_rec = 0
    for k, w in enumerate(var):
        up, dn, ck = int(w[0]), int(w[2]), float(c[k])
        groups = []
        ou = [p for p in range(8) if (up >> p) & 1]
        od = [p for p in range(8) if (dn >> p) & 1]
        for i in ou:
            for j in od:
                lst = opp[(i, j)]
                t = [((up ^ 1 << i) | 1 << r, (dn ^ 1 << j) | 1 << s) for _, (r, s), _, _ in lst]
                ok = [not ((up >> r) & 1 or (dn >> s) & 1) for _, (r, s), _, _ in lst]
                groups.append(((k, 0, -1, i, j), lst, t, ok))
        for occ, cfg, alpha in ((ou, up, 1), (od, dn, 0)):
            for x in range(len(occ)):
                for y in range(x + 1, len(occ)):
                    i, j = occ[x], occ[y]
                    lst = same[(i, j)]
                    nw = [(cfg ^ 1 << i ^ 1 << j) | 1 << r | 1 << s for _, (r, s), _, _ in lst]
                    t = [(q, dn) if alpha else (up, q) for q in nw]
                    ok = [not ((cfg >> r) & 1 or (cfg >> s) & 1) for _, (r, s), _, _ in lst]
                    groups.append(((k, 1, alpha, i, j), lst, t, ok))
        for occ, cfg, alpha in ((ou, up, 1), (od, dn, 0)):
            for i in occ:
                lst = sing[i]
                nw = [(cfg ^ 1 << i) | 1 << r for _, r, _, _ in lst]
                t = [(q, dn) if alpha else (up, q) for q in nw]
                ok = [not (cfg >> r) & 1 for _, r, _, _ in lst]
                groups.append(((k, 2, alpha, i, -1), lst, t, ok))
        for key, lst, t, ok in groups:
            a = np.array([e[0] for e in lst])
            ext = np.array(ok) & np.array([tt not in space for tt in t])
            below = ext & (a * abs(ck) < eps)
            read = ext & ~below
            if read.any():
                idx = np.nonzero(read)[0]
                tw = np.array([[tt[0], 0, tt[1], 0] for tt in (t[q] for q in idx)], dtype=np.uint64)
                hv, _ = H.off_diag(np.tile(np.asarray(w, dtype=np.uint64), (idx.size, 1)), tw)
                want_rel[k] = max(want_rel[k], float(np.max(np.abs(hv * ck))))
            if below.any():
                cut = int(np.argmax(below))
                assert not read[cut:].any()  # nothing is read past the cut
                n_rec += 1
                r = got[key]
                assert np.isclose(r[0], a[cut + 1:].sum(), rtol=1e-12, atol=1e-300)
                assert np.isclose(r[1], (a[cut + 1:] ** 2).sum(), rtol=1e-12, atol=1e-300)
                assert np.isclose(r[2], abs(ck) * r[0], rtol=1e-15, atol=0)
                assert np.isclose(r[3], ck * ck * r[1], rtol=1e-15, atol=0)
            else:
                assert key not in got
    assert n_rec == len(records) > 0
    assert np.array_equal(rel, want_rel)
