"""Brute-force second quantisation: an independent check of Slater-Condon rules.

The Hamiltonian H = sum h_pq a+_ps a_qs + 1/2 sum (pq|rs) a+_ps a+_rt a_st a_qs
is applied operator by operator to determinants written as one bitmask over
2n spin orbitals (alpha orbitals 0..n-1, then beta orbitals n..2n-1), with the
fermionic sign (-1)^(number of occupied spin orbitals below k) for each a_k /
a+_k.  Nothing here uses the oracle's or the library's formulas.
"""
import numpy as np


def _annihilate(occ: int, k: int):
    if not (occ >> k) & 1:
        return 0, 0
    sign = -1 if bin(occ & ((1 << k) - 1)).count("1") & 1 else 1
    return sign, occ & ~(1 << k)


def _create(occ: int, k: int):
    if (occ >> k) & 1:
        return 0, 0
    sign = -1 if bin(occ & ((1 << k) - 1)).count("1") & 1 else 1
    return sign, occ | (1 << k)


def ham_matrix(h: np.ndarray, eri: np.ndarray, e_core: float, dets: list[tuple[int, int]]) -> np.ndarray:
    """Dense <D_j|H|D_i> over the given (up, dn) determinants."""
    n = h.shape[0]
    index = {}
    occs = []
    for k, (u, d) in enumerate(dets):
        occ = u | (d << n)
        index[occ] = k
        occs.append(occ)
    m = len(dets)
    H = np.zeros((m, m))
    one = [(p, q, h[p, q]) for p in range(n) for q in range(n) if h[p, q] != 0.0]
    two = [(p, q, r, s, eri[p, q, r, s]) for p in range(n) for q in range(n) for r in range(n)
           for s in range(n) if eri[p, q, r, s] != 0.0]
    for i, occ in enumerate(occs):
        H[i, i] += e_core
        for sig in (0, n):
            for p, q, v in one:
                s1, o1 = _annihilate(occ, q + sig)
                if not s1:
                    continue
                s2, o2 = _create(o1, p + sig)
                if s2 and o2 in index:
                    H[index[o2], i] += s1 * s2 * v
        for sig in (0, n):
            for tau in (0, n):
                for p, q, r, s, v in two:
                    s1, o1 = _annihilate(occ, q + sig)
                    if not s1:
                        continue
                    s2, o2 = _annihilate(o1, s + tau)
                    if not s2:
                        continue
                    s3, o3 = _create(o2, r + tau)
                    if not s3:
                        continue
                    s4, o4 = _create(o3, p + sig)
                    if s4 and o4 in index:
                        H[index[o4], i] += 0.5 * s1 * s2 * s3 * s4 * v
    return H
