"""Determinant lists in the boundary's word layout, and a deterministic generator
of HCI-like determinant spaces for kernel benchmarks and parity tests.

Word layout: each determinant is four uint64 words ``[up_lo, up_hi, dn_lo, dn_hi]``
(the reference's ``Config{up: u128, dn: u128}``, src/wf/det.rs:16-19; bit p is
spatial orbital p).

``energy_truncated_space`` builds the set of determinants (a, b) whose orbital
excitation energy above the Hartree-Fock string pair is at most a cutoff,
tuned by bisection so the space holds ``n_target`` determinants, ordered by that
energy (HF first), mimicking the order in which HCI adds determinants.  Its
alpha/beta string structure (a few large alpha groups near HF, a long tail of
small ones) is the structure the connection search has to handle.
"""
from __future__ import annotations

import itertools

import numpy as np


def pack(up: np.ndarray, dn: np.ndarray) -> np.ndarray:
    """up/dn as uint64 (n_orb <= 64) or as pairs (lo, hi) -> (n, 4) uint64 words."""
    up = np.asarray(up)
    dn = np.asarray(dn)
    n = up.shape[0]
    out = np.zeros((n, 4), dtype=np.uint64)
    if up.ndim == 1:
        out[:, 0] = up.astype(np.uint64)
        out[:, 2] = dn.astype(np.uint64)
    else:
        out[:, 0:2] = up.astype(np.uint64)
        out[:, 2:4] = dn.astype(np.uint64)
    return out


def from_orbitals(dets: list[tuple[list[int], list[int]]]) -> np.ndarray:
    """[(up_orbitals, dn_orbitals), ...] -> (n, 4) uint64 words."""
    out = np.zeros((len(dets), 4), dtype=np.uint64)
    for i, (u, d) in enumerate(dets):
        for w, orbs in ((0, u), (2, d)):
            lo = hi = 0
            for o in orbs:
                if o < 64:
                    lo |= 1 << o
                else:
                    hi |= 1 << (o - 64)
            out[i, w] = lo
            out[i, w + 1] = hi
    return out


def to_int_pairs(words: np.ndarray) -> list[tuple[int, int]]:
    """(n, 4) words -> [(up, dn)] as Python ints."""
    w = words.astype(object)
    return [(int(r[0]) | (int(r[1]) << 64), int(r[2]) | (int(r[3]) << 64)) for r in w]


def _strings_below(eps: np.ndarray, n_e: int, cut: float) -> tuple[np.ndarray, np.ndarray]:
    """All n_e-electron strings whose orbital-energy sum above the lowest string
    is <= cut.  Returns (bitstrings as Python ints, energies) sorted by energy."""
    order = np.argsort(eps, kind="stable")
    e_sorted = eps[order]
    n = len(eps)
    base = float(np.sum(e_sorted[:n_e]))
    out_s: list[int] = []
    out_e: list[float] = []

    def rec(start: int, left: int, acc: float, bits: int):
        if left == 0:
            out_s.append(bits)
            out_e.append(acc - base)
            return
        for k in range(start, n - left + 1):
            # lowest possible completion from position k
            low = acc + float(np.sum(e_sorted[k:k + left]))
            if low - base > cut + 1e-12:
                break
            rec(k + 1, left - 1, acc + float(e_sorted[k]), bits | (1 << int(order[k])))

    rec(0, n_e, 0.0, 0)
    e = np.array(out_e)
    idx = np.argsort(e, kind="stable")
    return np.array(out_s, dtype=object)[idx], e[idx]


def energy_truncated_space(orb_energies: np.ndarray, n_up: int, n_dn: int,
                           n_target: int) -> np.ndarray:
    """Deterministic HCI-like space of exactly min(n_target, |FCI|) determinants."""
    eps = np.asarray(orb_energies, dtype=np.float64)

    def count(cut: float, ea: np.ndarray, eb: np.ndarray) -> int:
        ea = ea[ea <= cut + 1e-12]
        return int(np.sum(np.searchsorted(eb, cut - ea + 1e-12, side="right")))

    lo, hi = 0.0, 0.5
    while True:
        sa, ea = _strings_below(eps, n_up, hi)
        sb, eb = _strings_below(eps, n_dn, hi)
        if count(hi, ea, eb) >= n_target or hi > 1e4:
            break
        lo, hi = hi, hi * 2.0
    for _ in range(60):
        mid = 0.5 * (lo + hi)
        if count(mid, ea, eb) >= n_target:
            hi = mid
        else:
            lo = mid
    cut = hi
    keep_a = ea <= cut + 1e-12
    sa, ea = sa[keep_a], ea[keep_a]
    keep_b = eb <= cut + 1e-12
    sb, eb = sb[keep_b], eb[keep_b]
    nb = np.searchsorted(eb, cut - ea + 1e-12, side="right")
    ia = np.repeat(np.arange(len(sa)), nb)
    starts = np.cumsum(nb) - nb
    ib = np.arange(int(nb.sum())) - np.repeat(starts, nb)
    etot = ea[ia] + eb[ib]
    order = np.lexsort((ib, ia, etot))[:n_target]
    ia, ib = ia[order], ib[order]
    n_orb = len(eps)
    if n_orb <= 64:
        up = np.array([int(x) for x in sa], dtype=np.uint64)[ia]
        dn = np.array([int(x) for x in sb], dtype=np.uint64)[ib]
        return pack(up, dn)
    m = (1 << 64) - 1
    up = np.array([[int(x) & m, int(x) >> 64] for x in sa], dtype=np.uint64)[ia]
    dn = np.array([[int(x) & m, int(x) >> 64] for x in sb], dtype=np.uint64)[ib]
    return pack(up, dn)


def full_space(n_orb: int, n_up: int, n_dn: int) -> np.ndarray:
    """Every determinant (full CI space), for tiny systems."""
    ups = [sum(1 << o for o in c) for c in itertools.combinations(range(n_orb), n_up)]
    dns = [sum(1 << o for o in c) for c in itertools.combinations(range(n_orb), n_dn)]
    m = (1 << 64) - 1
    out = np.zeros((len(ups) * len(dns), 4), dtype=np.uint64)
    k = 0
    for a in ups:
        for b in dns:
            out[k] = [a & m, a >> 64, b & m, b >> 64]
            k += 1
    return out
