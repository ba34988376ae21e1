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
