"""Integral tables, seeded synthetic FCIDUMPs and the FCIDUMP text format.

Layout follows the reference exactly: one-body integrals live at
``one_body[combine_2(p, q)]`` and two-electron integrals (pq|rs) at
``two_body[combine_4(p, q, r, s)]`` with 1-based orbital indices
(reference ``src/utils/ints.rs:38-68``; table sizes ``src/ham/read_ints.rs:78-80``).

Synthetic generator (documented PRNG, so any implementation can reproduce it):
  * splitmix64 stream seeded with ``seed``;
  * uniform u = (x >> 11) * 2**-53 in [0, 1);
  * normal via Box-Muller from two uniforms: sqrt(-2 ln(1 - u1)) * cos(2 pi u2)
    (one normal per pair of draws; the sine partner is discarded);
  * draw order: off-diagonal h_pq (p < q, row-major upper triangle), then for
    K = 1..n_chol the upper triangle (p <= q, row-major) of symmetric L^K;
  * (pq|rs) = sum_K L^K_pq L^K_rs accumulated in K order.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

_MASK64 = (1 << 64) - 1


def combine_2(p: int, q: int) -> int:
    """Triangular packing of a 1-based pair (reference src/utils/ints.rs:38-46)."""
    i, j = abs(p), abs(q)
    if i > j:
        return (i * (i - 1)) // 2 + j
    return (j * (j - 1)) // 2 + i


def _combine_2u(i: int, j: int) -> int:
    if i > j:
        return (i * (i - 1)) // 2 + j
    return (j * (j - 1)) // 2 + i


def combine_4(p: int, q: int, r: int, s: int) -> int:
    """Nested triangular packing (reference src/utils/ints.rs:66-68)."""
    return _combine_2u(combine_2(p, q), combine_2(r, s))


def table_sizes(n_orb: int) -> tuple[int, int]:
    """Sizes the reference allocates (src/ham/read_ints.rs:78-80)."""
    n1 = combine_2(n_orb + 1, n_orb + 1)
    n2 = combine_4(n_orb + 1, n_orb + 1, n_orb + 1, n_orb + 1)
    return n1, n2


@dataclass
class Integrals:
    n_orb: int
    e_core: float
    one_body: np.ndarray  # float64, combine_2 layout
    two_body: np.ndarray  # float64, combine_4 layout

    def h1(self, p: int, q: int) -> float:
        """One-body h_pq, 0-based (reference src/ham/mod.rs:65-67)."""
        return float(self.one_body[combine_2(p + 1, q + 1)])

    def eri(self, p: int, q: int, r: int, s: int) -> float:
        """(pq|rs), 0-based chemist's notation."""
        return float(self.two_body[combine_4(p + 1, q + 1, r + 1, s + 1)])

    def dense(self) -> tuple[np.ndarray, np.ndarray]:
        """Dense h[p,q] and eri[p,q,r,s] (0-based), for small test systems."""
        n = self.n_orb
        c2 = np.zeros((n, n), dtype=np.int64)
        for p in range(n):
            for q in range(n):
                c2[p, q] = combine_2(p + 1, q + 1)
        h = self.one_body[c2]
        a = c2.reshape(n * n)
        hi = np.maximum(a[:, None], a[None, :])
        lo = np.minimum(a[:, None], a[None, :])
        c4 = (hi * (hi - 1)) // 2 + lo
        eri = self.two_body[c4].reshape(n, n, n, n)
        return h, eri


class SplitMix64:
    """splitmix64 (Steele, Lea, Flood 2014) with the constants of its published form."""

    def __init__(self, seed: int):
        self.state = seed & _MASK64

    def next_u64(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & _MASK64
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _MASK64
        return z ^ (z >> 31)

    def uniform(self) -> float:
        return (self.next_u64() >> 11) * (1.0 / 9007199254740992.0)

    def normal(self) -> float:
        u1 = self.uniform()
        u2 = self.uniform()
        return math.sqrt(-2.0 * math.log(1.0 - u1)) * math.cos(2.0 * math.pi * u2)

    def normals(self, m: int) -> np.ndarray:
        return np.array([self.normal() for _ in range(m)], dtype=np.float64)


def _pack(n_orb: int, e_core: float, h: np.ndarray, eri_pair: np.ndarray) -> Integrals:
    """eri_pair[P, Q] holds (pq|rs) for pair indices P=(p>=q) Q=(r>=s) in
    combine_2 order (P = combine_2(p+1, q+1) - 1)."""
    n1, n2 = table_sizes(n_orb)
    one = np.zeros(n1, dtype=np.float64)
    two = np.zeros(n2, dtype=np.float64)
    for p in range(n_orb):
        for q in range(p + 1):
            one[combine_2(p + 1, q + 1)] = h[p, q]
    m = eri_pair.shape[0]
    P = np.arange(1, m + 1, dtype=np.int64)
    hi = np.maximum(P[:, None], P[None, :])
    lo = np.minimum(P[:, None], P[None, :])
    idx = (hi * (hi - 1)) // 2 + lo
    two[idx.ravel()] = eri_pair.ravel()
    return Integrals(n_orb, e_core, one, two)


def _pair_index(n_orb: int) -> np.ndarray:
    """pidx[p, q] = combine_2(p+1, q+1) - 1."""
    pidx = np.zeros((n_orb, n_orb), dtype=np.int64)
    for p in range(n_orb):
        for q in range(n_orb):
            pidx[p, q] = combine_2(p + 1, q + 1) - 1
    return pidx


def synthetic_molecular(n_orb: int, n_chol: int, seed: int, eps_span: float = 15.0,
                        h_sigma: float = 0.05, l_sigma: float = 0.08) -> Integrals:
    """Seeded synthetic molecular-like integrals.

    h_pp = -1 + 2 p / eps_span; h_pq ~ N(0, h_sigma^2) symmetric;
    (pq|rs) = sum_K L^K_pq L^K_rs with symmetric L^K ~ N(0, l_sigma^2).
    The result has 8-fold symmetry and a positive semidefinite ERI matrix.
    """
    rng = SplitMix64(seed)
    h = np.zeros((n_orb, n_orb))
    for p in range(n_orb):
        h[p, p] = -1.0 + 2.0 * p / eps_span
    for p in range(n_orb):
        for q in range(p + 1, n_orb):
            v = h_sigma * rng.normal()
            h[p, q] = v
            h[q, p] = v
    pidx = _pair_index(n_orb)
    npair = n_orb * (n_orb + 1) // 2
    up = [(p, q) for p in range(n_orb) for q in range(p, n_orb)]
    pos = np.array([pidx[p, q] for (p, q) in up], dtype=np.int64)
    eri = np.zeros((npair, npair))
    for _ in range(n_chol):
        vals = l_sigma * rng.normals(len(up))
        lk = np.zeros(npair)
        lk[pos] = vals
        eri += np.outer(lk, lk)
    return _pack(n_orb, 0.0, h, eri)


def hubbard_2d(lx: int, ly: int, t: float, u: float, disorder: float, seed: int) -> Integrals:
    """Periodic lx*ly Hubbard lattice as integrals: h_pq = -t on bonds,
    h_pp = eps_p ~ U(-disorder, disorder), (pp|pp) = U, all other ERIs zero."""
    n = lx * ly
    rng = SplitMix64(seed)
    h = np.zeros((n, n))
    for p in range(n):
        h[p, p] = -disorder + 2.0 * disorder * rng.uniform()
    for x in range(lx):
        for y in range(ly):
            p = x * ly + y
            for q in (((x + 1) % lx) * ly + y, x * ly + (y + 1) % ly):
                if q != p:
                    h[p, q] = -t
                    h[q, p] = -t
    npair = n * (n + 1) // 2
    eri = np.zeros((npair, npair))
    pidx = _pair_index(n)
    for p in range(n):
        eri[pidx[p, p], pidx[p, p]] = u
    return _pack(n, 0.0, h, eri)


def write_fcidump(path: str, ints: Integrals, n_elec: int, ms2: int = 0) -> None:
    """FCIDUMP text with %.17e values (round-trips every double exactly).
    No blank lines: the reference reader rejects them (src/ham/read_ints.rs:95-100)."""
    n = ints.n_orb
    h, eri = ints.dense()
    lines = [f" &FCI NORB={n:4d},NELEC={n_elec:2d},MS2={ms2},",
             "  ORBSYM=" + ",".join(["1"] * n) + ",", "  ISYM=1,", " &END"]
    for p in range(n):
        for q in range(p + 1):
            for r in range(n):
                for s in range(r + 1):
                    if (p * (p + 1) // 2 + q) < (r * (r + 1) // 2 + s):
                        continue
                    v = eri[p, q, r, s]
                    if v != 0.0:
                        lines.append(f"{v:.17e} {p + 1:4d} {q + 1:4d} {r + 1:4d} {s + 1:4d}")
    for p in range(n):
        for q in range(p + 1):
            if h[p, q] != 0.0:
                lines.append(f"{h[p, q]:.17e} {p + 1:4d} {q + 1:4d}    0    0")
    lines.append(f"{ints.e_core:.17e}    0    0    0    0")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


class FcidumpError(ValueError):
    def __init__(self, message: str, line: int):
        super().__init__(f"FCIDUMP line {line}: {message}")
        self.line = line


def read_fcidump(path: str, n_orb: int) -> Integrals:
    """Reader with the reference's rules (src/ham/read_ints.rs:73-190): a line
    whose first token does not parse as a float is a header line and skipped;
    a line with a float must carry four integer indices in [0, n_orb];
    (0,0,0,0) is the core energy, (p,q,0,0) one-body, otherwise (pq|rs)."""
    n1, n2 = table_sizes(n_orb)
    one = np.zeros(n1)
    two = np.zeros(n2)
    nuc = 0.0
    with open(path) as f:
        for ln, line in enumerate(f, start=1):
            toks = line.split()
            if not toks:
                raise FcidumpError("Missing integral value", ln)
            try:
                val = float(toks[0])
            except ValueError:
                continue
            idx = []
            for name, k in zip("pqrs", range(1, 5)):
                if k >= len(toks):
                    raise FcidumpError(f"Missing {name} index", ln)
                try:
                    idx.append(int(toks[k]))
                except ValueError:
                    raise FcidumpError(f"Invalid {name} index: '{toks[k]}'", ln) from None
            for name, v in zip("pqrs", idx):
                if v < 0 or v > n_orb:
                    raise FcidumpError(f"Index {name} out of range [0, {n_orb}]: {v}", ln)
            p, q, r, s = idx
            if p == 0 and q == 0 and r == 0 and s == 0:
                nuc = val
            elif r == 0 and s == 0:
                one[combine_2(p, q)] = val
            else:
                two[combine_4(p, q, r, s)] = val
    return Integrals(n_orb, nuc, one, two)
