"""Shared builders of test systems (seeded, deterministic)."""
from functools import lru_cache

import numpy as np

from risq_amd.dets import energy_truncated_space
from risq_amd.fcidump import hubbard_2d, synthetic_molecular


@lru_cache(maxsize=None)
def molecular(n_orb, n_chol, seed):
    return synthetic_molecular(n_orb, n_chol, seed)


def orb_energies(ints):
    return np.array([ints.h1(p, p) for p in range(ints.n_orb)])


@lru_cache(maxsize=None)
def mol_space(n_orb, n_chol, seed, n_up, n_dn, n_det):
    ints = molecular(n_orb, n_chol, seed)
    dets = energy_truncated_space(orb_energies(ints), n_up, n_dn, n_det)
    return ints, dets


@lru_cache(maxsize=None)
def hubbard_space(lx, ly, n_up, n_dn, n_det, seed=2):
    ints = hubbard_2d(lx, ly, 1.0, 4.0, 0.1, seed)
    # site basis: order strings by a fixed pseudo-energy (site index spread)
    eps = orb_energies(ints) + 0.01 * np.arange(ints.n_orb)
    dets = energy_truncated_space(eps, n_up, n_dn, n_det)
    return ints, dets


def csr_to_dense(rp, col, val, diag):
    n = rp.size - 1
    a = np.zeros((n, n))
    for i in range(n):
        a[i, col[rp[i]:rp[i + 1]]] = val[rp[i]:rp[i + 1]]
    return a + a.T + np.diag(diag)


@lru_cache(maxsize=None)
def tiny_single_molecular(n_orb=8, n_chol=8, seed=3, orb=5, partner=2, h_small=1e-13):
    """Molecular integrals in which orbital `orb` couples to the others only
    through h[orb, partner] = h_small: every (pq|rs) holding `orb` an odd number
    of times is zero and h[orb, q] = 0 for q != partner, orb.  A single
    partner -> orb then has |H| = h_small exactly (every other term of ham_sing
    is an exact zero), other singles into `orb` are exact zeros, so the two
    single-storage rules (H != 0 and |H| > 1e-12) give different patterns."""
    from risq_amd.fcidump import Integrals, combine_2, combine_4
    base = synthetic_molecular(n_orb, n_chol, seed)
    one = base.one_body.copy()
    two = base.two_body.copy()
    for q in range(n_orb):
        if q != orb:
            one[combine_2(orb + 1, q + 1)] = h_small if q == partner else 0.0
    for p in range(n_orb):
        for q in range(n_orb):
            for r in range(n_orb):
                for s in range(n_orb):
                    if [p, q, r, s].count(orb) % 2 == 1:
                        two[combine_4(p + 1, q + 1, r + 1, s + 1)] = 0.0
    return Integrals(n_orb, base.e_core, one, two)


def spin_flip_closed(dets):
    """Each determinant followed by its spin-flipped partner when that is new
    (the order push_and_spin_flipped gives, src/wf/mod.rs:99-116, up to which
    of the two comes first)."""
    seen = set()
    out = []
    for w in dets:
        t = tuple(int(x) for x in w)
        for d in (t, (t[2], t[3], t[0], t[1])):
            if d not in seen:
                seen.add(d)
                out.append(d)
    return np.array(out, dtype=np.uint64)
