"""Deterministic second-order energy by enumeration.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the result the library's
hci_pt2_dtm is compared against.  Never used by risq_amd/.

Restates dtm_pt (src/pt/mod.rs:81-122) over the candidates of
dets_excites_and_excited_dets (src/excite/iterator.rs:33-73), but enumerates
every single and double excitation among the valence orbitals instead of
walking the sorted excitation lists:

* an opposite-spin double (p up, q dn -> r up, s dn) is a candidate iff
  |(pr|qs)| * |c_i| >= eps, and a same-spin double iff
  |(pr|qs) - (ps|qr)| * |c_i| >= eps.  The lists hold exactly these
  magnitudes sorted in descending order and are read while the product stays
  >= eps (iterator.rs:44-46), so they yield the same set.  Both magnitudes
  equal |H_ai| (the element is the signed magnitude), so the test is
  |H_ai| * |c_i| >= eps;
* a single is a candidate iff |H_ai * c_i| >= eps (src/pt/mod.rs:97-102):
  its list is cut by an upper bound of |H_ai| first, which keeps every such
  single (the bound dominates every element; tests/test_hci_select.py checks
  it);
* targets inside the space are dropped (iterator.rs:58-60); core orbitals (the
  n_core lowest h_pp) are neither excited nor filled;
* contributions H_ai c_i are summed per target in discovery order (source
  determinant by source determinant) and e2 = sum_a num_a^2 / (e0 - H_aa),
  summed in first-discovery order (src/pt/mod.rs:132-136).  H_aa comes from
  Ham::ham_diag (the reference updates the parent's diagonal instead,
  src/wf/det.rs:249-260; the two agree to rounding).

Elements and diagonals come from the C++ oracle (pyoracle), whose elements are
pinned by brute-force second quantisation.
"""
from __future__ import annotations

import numpy as np

from . import pyoracle


def _occ(x: int, orbs: list[int]) -> list[int]:
    return [p for p in orbs if (x >> p) & 1]


_M64 = (1 << 64) - 1


def _words(up: list[int], dn: list[int]) -> np.ndarray:
    w = np.zeros((len(up), 4), dtype=np.uint64)
    w[:, 0] = [x & _M64 for x in up]
    w[:, 1] = [x >> 64 for x in up]
    w[:, 2] = [x & _M64 for x in dn]
    w[:, 3] = [x >> 64 for x in dn]
    return w


def dtm_pt(ints, dets: np.ndarray, coeffs: np.ndarray, e0: float, eps: float, n_core: int = 0):
    """(e2, number of external determinants) for the space `dets` ((n, 4)
    uint64 words) with coefficients `coeffs`."""
    n = ints.n_orb
    H = pyoracle.OracleHam(ints)
    order = sorted(range(n), key=lambda p: ints.h1(p, p))  # stable: ties keep index order
    valence = sorted(order[n_core:])
    space = {(int(w[0]) | int(w[1]) << 64, int(w[2]) | int(w[3]) << 64) for w in dets}
    h_psi: dict[tuple[int, int], float] = {}
    for w, c in zip(dets, coeffs):
        up, dn, c = int(w[0]) | int(w[1]) << 64, int(w[2]) | int(w[3]) << 64, float(c)
        ou, od = _occ(up, valence), _occ(dn, valence)
        vu = [r for r in valence if not (up >> r) & 1]
        vd = [r for r in valence if not (dn >> r) & 1]
        tu, td, single = [], [], []
        for p in ou:
            for q in od:
                for r in vu:
                    for s in vd:
                        tu.append(up ^ (1 << p) | (1 << r))
                        td.append(dn ^ (1 << q) | (1 << s))
                        single.append(False)
        for cfg, occ, vir, spin in ((up, ou, vu, 0), (dn, od, vd, 1)):
            for a in range(len(occ)):
                for b in range(a + 1, len(occ)):
                    for x in range(len(vir)):
                        for y in range(x + 1, len(vir)):
                            nw = cfg ^ (1 << occ[a]) ^ (1 << occ[b]) | (1 << vir[x]) | (1 << vir[y])
                            tu.append(nw if spin == 0 else up)
                            td.append(dn if spin == 0 else nw)
                            single.append(False)
        for cfg, occ, vir, spin in ((up, ou, vu, 0), (dn, od, vd, 1)):
            for p in occ:
                for r in vir:
                    nw = cfg ^ (1 << p) | (1 << r)
                    tu.append(nw if spin == 0 else up)
                    td.append(dn if spin == 0 else nw)
                    single.append(True)
        if not tu:
            continue
        tgt = _words(tu, td)
        src = np.tile(np.array(w, dtype=np.uint64), (tgt.shape[0], 1))
        h, _ = H.off_diag(src, tgt)
        single = np.array(single)
        hc = h * c
        keep = np.where(single, np.abs(hc) >= eps, np.abs(h) * abs(c) >= eps)
        for k in np.nonzero(keep)[0]:
            key = (tu[k], td[k])
            if key in space:
                continue
            h_psi[key] = h_psi.get(key, 0.0) + float(hc[k])
    if not h_psi:
        return 0.0, 0
    keys = list(h_psi)
    ext = _words([k[0] for k in keys], [k[1] for k in keys])
    diag = H.ham_diag(ext)
    e2 = 0.0
    for k, d in zip(keys, diag):
        e2 += h_psi[k] * h_psi[k] / (e0 - d)
    return e2, len(keys)


def epstein_nesbet_dense(h_full: np.ndarray, in_space: np.ndarray, coeffs: np.ndarray, e0: float,
                         allowed: np.ndarray | None = None) -> float:
    """Closed form on a complete space: v = H[:, space] c, summed over the
    determinants outside the space (and inside `allowed`, a boolean mask),
    v_a^2 / (e0 - H_aa)."""
    v = h_full[:, in_space] @ coeffs
    out = ~np.isin(np.arange(h_full.shape[0]), in_space)
    if allowed is not None:
        out &= allowed
    return float(np.sum(v[out] ** 2 / (e0 - np.diag(h_full)[out])))
