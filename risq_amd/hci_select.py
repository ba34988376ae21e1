"""Heat-bath CI variational selection: the host loop around the GPU library.

The selection restates the reference's variational stage so that its
determinant lists, and therefore its variational energies, can be reproduced:

* excitation generator (src/excite/init.rs:40-249): for every valence pair the
  opposite-spin double targets sorted by |(pr|qs)|, the same-spin double targets
  sorted by |(pr|qs) - (ps|qr)|, and for every valence orbital the single
  targets sorted by an upper bound of |H_single|;
* epsilon schedule (src/wf/eps.rs:12-171): start just below the largest double
  excitation from Hartree-Fock, divide by 10 per step down to eps_var; the
  schedule is advanced twice per iteration (src/wf/mod.rs:1424 and
  src/var/mod.rs:123);
* selection (src/wf/mod.rs:1418-1631): from every determinant D, every
  excitation whose sorted magnitude reaches eps / |c_D| and that leads outside
  the current space; singles only once eps has reached eps_var and only if the
  exact element passes; each new determinant is added after its spin-flipped
  partner (src/wf/mod.rs:99-116);
* the loop (src/var/mod.rs:61-134): select, stop when nothing is added at
  eps_var, otherwise build H incrementally and run Davidson warm-started from
  the previous coefficients (src/var/davidson.rs:43-113); at most 20 iterations.

Matrix elements, diagonals, the sparse H build and Davidson go through a
solver object; `GpuSolver` runs them through the HIP library.  Determinants
are (up, dn) pairs of Python ints (bit p = spatial orbital p).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from .fcidump import Integrals


def _bits(x: int):
    while x:
        low = x & -x
        yield low.bit_length() - 1
        x ^= low


def _bit_pairs(x: int):
    b = list(_bits(x))
    for a in range(len(b)):
        for c in range(a + 1, len(b)):
            yield b[a], b[c]


class ExciteGenerator:
    """Sorted excitation lists (reference ExciteGenerator, src/excite/init.rs:40-249).
    Core orbitals are the n_core lowest by h_pp (src/ham/read_ints.rs:161-184)."""

    def __init__(self, ints: Integrals, n_core: int, n_up: int, n_dn: int):
        n = ints.n_orb
        h, eri = ints.dense()
        order = sorted(range(n), key=lambda p: h[p, p])  # stable, like the reference's sort_by
        self.core = order[:n_core]
        self.valence_orbs = order[n_core:]
        self.valence = sum(1 << p for p in self.valence_orbs)
        val = self.valence_orbs
        # direct(p, q, r, s) = (pr|qs); direct_plus_exchange = (pr|qs) - (ps|qr)  (src/ham/mod.rs:69-82)
        self.max_opp_doub = 0.0
        self.opp = {}
        for p in val:
            for q in val:
                v = []
                for r in val:
                    if p == r:
                        continue
                    for s in val:
                        if q == s:
                            continue
                        a = abs(float(eri[p, r, q, s]))
                        self.max_opp_doub = max(self.max_opp_doub, a)
                        v.append((a, r, s))
                v.sort(key=lambda e: -e[0])
                self.opp[(p, q)] = v
        self.max_same_doub = 0.0
        self.same = {}
        for p in val:
            for q in val:
                if p >= q:
                    continue
                v = []
                for r in val:
                    if r in (p, q):
                        continue
                    for s in val:
                        if r >= s or s in (p, q):
                            continue
                        a = abs(float(eri[p, r, q, s]) - float(eri[p, s, q, r]))
                        self.max_same_doub = max(self.max_same_doub, a)
                        v.append((a, r, s))
                v.sort(key=lambda e: -e[0])
                self.same[(p, q)] = v
        # singles: |h_pr + sum over the n_up - 1 same-spin and n_dn opposite-spin
        # occupied q| bounded by the smallest and the largest partial sums
        self.sing = {}
        for p in val:
            v = []
            for r in val:
                if p == r:
                    continue
                v_same, v_opp = [], []
                for q in range(n):
                    if q == p or q == r:
                        continue
                    v_same.append(float(eri[p, r, q, q]) - float(eri[p, q, q, r]))
                    v_opp.append(float(eri[p, r, q, q]))
                v_same.sort()
                v_opp.sort()
                hpr = float(h[p, r])
                lo = hpr + sum(v_same[:n_up - 1]) + sum(v_opp[:n_dn])
                hi = hpr + sum(v_same[len(v_same) - (n_up - 1):]) + sum(v_opp[len(v_same) - n_dn:])
                v.append((abs(lo) if abs(lo) > abs(hi) else abs(hi), r))
            v.sort(key=lambda e: -e[0])
            self.sing[p] = v
        self.max_sing = max((self.sing[p][0][0] for p in val if self.sing[p]), default=0.0)


class EpsSchedule:
    """src/wf/eps.rs:12-31: yields the current value, then divides by 10 down to the target."""

    def __init__(self, first: float, target: float):
        self.next_one = first
        self.target = target

    def next(self) -> float:
        cur = self.next_one
        proposed = self.next_one * 0.1
        self.next_one = proposed if proposed > self.target else self.target
        return cur


def init_eps(dets, gen: ExciteGenerator, eps_var: float, z_sym: int = 1) -> EpsSchedule:
    """src/wf/eps.rs:44-171: the largest double-excitation magnitude reachable from the start."""
    max_sym = max_asym = max_asym_flip = eps_var
    for up, dn in dets:
        for i in _bits(gen.valence & up):
            for j in _bits(gen.valence & dn):
                found_sym = found_asym = False
                for a, r, s in gen.opp[(i, j)]:
                    if (up >> r) & 1 or (dn >> s) & 1:
                        continue
                    if i == j and r == s:
                        found_sym = True
                        max_sym = max(max_sym, a)
                    else:
                        found_asym = True
                        max_asym = max(max_asym, a)
                        if i == j or r == s:
                            max_asym_flip = max(max_asym_flip, a)
                    if found_sym and found_asym:
                        break
        for cfg in (up, dn):
            for i, j in _bit_pairs(gen.valence & cfg):
                for a, r, s in gen.same[(i, j)]:
                    if not ((cfg >> r) & 1 or (cfg >> s) & 1):
                        max_asym = max(max_asym, a)
                        break
    max_doub = (max_sym if max_sym > max_asym else max_asym) if z_sym == 1 else max_asym_flip
    return EpsSchedule(max_doub - 1e-9, eps_var)


_M64 = (1 << 64) - 1


def _words(dets) -> np.ndarray:
    """[(up, dn)] -> (n, 4) uint64 words [up_lo, up_hi, dn_lo, dn_hi]."""
    out = np.zeros((len(dets), 4), dtype=np.uint64)
    for k, (u, d) in enumerate(dets):
        out[k] = (u & _M64, u >> 64, d & _M64, d >> 64)
    return out


def _pairs(words: np.ndarray):
    """(n, 4) uint64 words -> [(up, dn)] Python ints."""
    w = words.astype(object)
    return [(int(r[0]) | (int(r[1]) << 64), int(r[2]) | (int(r[3]) << 64)) for r in w]


def new_diag(ints: Integrals, parent: tuple[int, int], diag: float, exc: tuple) -> float:
    """Det::new_diag_opp / new_diag_same / new_diag_sing (src/wf/det.rs:271-420):
    the diagonal of the determinant `exc` leads to, updated from its parent's in
    O(N) with the reference's order of operations.  exc is ("opp", (p, q), (r, s))
    for p up, q dn -> r up, s dn; ("same", (p, q), (r, s), alpha); ("sing", p, r,
    alpha)."""
    def direct(p, q, r, s):  # src/ham/mod.rs:69-73: (pr|qs)
        return ints.eri(p, r, q, s)

    def dpe(p, q, r, s):  # src/ham/mod.rs:75-82: (pr|qs) - (ps|qr)
        return ints.eri(p, r, q, s) - ints.eri(p, s, q, r)

    up, dn = parent
    if exc[0] == "sing":
        _, p, r, alpha = exc
        out = diag + ints.h1(r, r) - ints.h1(p, p)
        same, other = (up, dn) if alpha else (dn, up)
        for i in _bits(same):
            if i == p:
                continue
            out += dpe(i, r, i, r) - dpe(i, p, i, p)
        for i in _bits(other):
            out += direct(i, r, i, r) - direct(i, p, i, p)
        return out
    (p, q), (r, s) = exc[1], exc[2]
    out = diag + ints.h1(r, r) + ints.h1(s, s) - ints.h1(p, p) - ints.h1(q, q)
    if exc[0] == "opp":
        out += direct(r, s, r, s) - direct(p, q, p, q)
        for i in _bits(up):
            if i == p:
                continue
            out += dpe(i, r, i, r) - dpe(i, p, i, p)
            out += direct(i, s, i, s) - direct(i, q, i, q)
        for i in _bits(dn):
            if i == q:
                continue
            out += dpe(i, s, i, s) - dpe(i, q, i, q)
            out += direct(i, r, i, r) - direct(i, p, i, p)
        return out
    alpha = exc[3]
    out += dpe(r, s, r, s) - dpe(p, q, p, q)
    same, other = (up, dn) if alpha else (dn, up)
    for i in _bits(same):
        if i == p or i == q:
            continue
        out += dpe(i, r, i, r) + dpe(i, s, i, s) - dpe(i, p, i, p) - dpe(i, q, i, q)
    for i in _bits(other):
        out += direct(i, r, i, r) + direct(i, s, i, s) - direct(i, p, i, p) - direct(i, q, i, q)
    return out


def find_new_dets(dets, coeffs, index: dict, gen: ExciteGenerator, eps: float, singles: bool, elements,
                  diags=None, ints: Integrals | None = None):
    """src/wf/mod.rs:1467-1631 (iterate_excites without matmul): new determinants in
    the order the reference discovers them.  `elements(words1, words2)` gives the
    exact singles' matrix elements (one batched call).  With the space's `diags`
    (and `ints`) it returns (new determinants, their diagonals from the first
    discoverer, src/wf/mod.rs:1525,1569,1618)."""
    cands = []  # (config, single slot or -1 for a double, parent index, excitation) in discovery order
    single_pairs = []
    for k, (up, dn) in enumerate(dets):
        c = abs(coeffs[k])
        local = eps / c if c > 0 else math.inf
        if gen.max_opp_doub >= local:
            for i in _bits(gen.valence & up):
                for j in _bits(gen.valence & dn):
                    for a, r, s in gen.opp[(i, j)]:
                        if a < local:
                            break
                        if (up >> r) & 1 or (dn >> s) & 1:
                            continue
                        cands.append((((up ^ (1 << i)) | (1 << r), (dn ^ (1 << j)) | (1 << s)), -1, k,
                                      ("opp", (i, j), (r, s))))
        if gen.max_same_doub >= local:
            for cfg, alpha in ((up, True), (dn, False)):
                for i, j in _bit_pairs(gen.valence & cfg):
                    for a, r, s in gen.same[(i, j)]:
                        if a < local:
                            break
                        if (cfg >> r) & 1 or (cfg >> s) & 1:
                            continue
                        new = (cfg ^ (1 << i) ^ (1 << j)) | (1 << r) | (1 << s)
                        cands.append(((new, dn) if alpha else (up, new), -1, k, ("same", (i, j), (r, s), alpha)))
        if singles and gen.max_sing >= local:
            for cfg, alpha in ((up, True), (dn, False)):
                for i in _bits(gen.valence & cfg):
                    for a, r in gen.sing[i]:
                        if a < local:
                            break
                        if (cfg >> r) & 1:
                            continue
                        new = (cfg ^ (1 << i)) | (1 << r)
                        d = (new, dn) if alpha else (up, new)
                        if d in index:
                            continue
                        cands.append((d, len(single_pairs), k, ("sing", i, r, alpha)))
                        single_pairs.append((k, d, local))
    ok = np.ones(len(single_pairs), dtype=bool)
    if single_pairs:
        vals = elements(_words([dets[k] for k, _, _ in single_pairs]), _words([d for _, d, _ in single_pairs]))
        ok = np.abs(vals) > np.array([loc for _, _, loc in single_pairs])
    out, out_diag, seen = [], [], set()
    for d, si, k, exc in cands:
        if d in index or d in seen:
            continue
        if si >= 0 and not ok[si]:
            continue
        seen.add(d)
        out.append(d)
        if diags is not None:
            out_diag.append(new_diag(ints, dets[k], float(diags[k]), exc))
    return (out, out_diag) if diags is not None else out


@dataclass
class VariationalResult:
    energy: float
    dets: list
    coeffs: np.ndarray
    iterations: int
    converged: bool
    history: list = field(default_factory=list)  # (eps, n_dets, energy) per iteration


def variational(solver, ints: Integrals, n_up: int, n_dn: int, eps_var: float, n_core: int = 0,
                z_sym: int = 1, max_iters: int = 20, log=None) -> VariationalResult:
    """The reference's variational stage (src/var/mod.rs:61-134, src/wf/mod.rs:1687-1740)."""
    if z_sym != 1:
        raise NotImplementedError("only z_sym = 1 (the reference's default) is supported")
    hf = ((1 << n_up) - 1, (1 << n_dn) - 1)
    dets = [hf]
    index = {hf: 0}
    coeffs = np.ones(1)
    diag = solver.ham_diag(_words(dets))
    energy = float(diag[0])
    # new determinants carry the diagonal their first discoverer computes
    # incrementally (new_diag_*), as the reference's Det.diag does
    if hasattr(solver, "select"):
        # lists and selection in the library
        sched = EpsSchedule(solver.select_init(n_core, n_up, n_dn, eps_var), eps_var)

        def finder(eps, singles):
            w, dg = solver.select(_words(dets), coeffs, eps, singles, diag)
            return [tuple(d) for d in _pairs(w)], list(dg)
    else:
        gen = ExciteGenerator(ints, n_core, n_up, n_dn)
        sched = init_eps(dets, gen, eps_var, z_sym)

        def finder(eps, singles):
            return find_new_dets(dets, coeffs, index, gen, eps, singles, solver.elements, diags=diag, ints=ints)
    history = []
    converged = False
    it = 0
    while True:
        it += 1
        if it > max_iters:
            break
        eps = sched.next()
        new, new_diag_vals = finder(eps, eps == eps_var)
        n_old = len(dets)
        added = []
        for d, dg in zip(new, new_diag_vals):  # push_and_spin_flipped (src/wf/mod.rs:99-116)
            for e in ((d[1], d[0]), d) if d[0] != d[1] else (d,):
                if e not in index:
                    index[e] = len(dets)
                    dets.append(e)
                    added.append(dg)  # the flipped copy is a clone, diagonal included
        if eps == eps_var and not new:
            converged = True
            break
        coeffs = np.concatenate([coeffs, np.zeros(len(dets) - n_old)])
        diag = np.concatenate([diag, np.array(added, dtype=np.float64)])
        # The reference warm-starts from the previous coefficients from the
        # second iteration on (src/var/mod.rs:104).  When the previous space
        # was the single start determinant (no double reaches eps, e.g. the
        # site-basis Hubbard model), that vector is an exact eigenvector of
        # the old space, its DPR correction is 0/0 and the reference's
        # Davidson::new panics for N = 1; such a solve starts cold here.
        warm = it > 1 and n_old > 1
        e, c, ok = solver.optimize(_words(dets), diag, coeffs if warm else None)
        if ok:
            energy, coeffs = e, c
        history.append((eps, len(dets), energy))
        if log:
            log(f"HCI iteration {it}: eps={eps:.3e} dets={len(dets)} E={energy:.10f}")
        sched.next()
    return VariationalResult(energy, dets, coeffs, it, converged, history)


class GpuSolver:
    """Selection, elements, diagonals, incremental H builds and Davidson on the
    GPU library (reference semantics: FP64 values, Davidson tol 1e-4 / etol
    1e-8, max 50 iterations, restart to the Ritz vector).  With
    library_select=False the selection runs in this module (find_new_dets).
    storage: tiled by default (rebuilt whole at each iteration, keeping the
    rule each pair was stored under), or FULL (entries of the previous
    iteration kept and renumbered)."""

    def __init__(self, ctx, ints: Integrals, library_select: bool = True, storage: int | None = None):
        from .hci import STORE_TILED
        self.ctx = ctx
        ctx.set_integrals(ints)
        self.h = None
        self.storage = STORE_TILED if storage is None else storage
        if library_select:
            self.select_init = lambda n_core, n_up, n_dn, eps_var: ctx.excite_init(n_core, n_up, n_dn, eps_var)
            self.select = lambda words, coeffs, eps, singles, diag: ctx.select(words, coeffs, eps, singles,
                                                                               diag=diag)

    def ham_diag(self, words):
        return self.ctx.ham_diag(words)

    def elements(self, w1, w2):
        return self.ctx.elements(w1, w2)[0]

    def optimize(self, words, diag, init):
        from .hci import sparse_optimize
        n_stored = self.h.n if self.h is not None else 0
        h, res = sparse_optimize(self.ctx, words, diag, prev=self.h, n_stored=n_stored, init=init,
                                 storage=self.storage)
        if self.h is not None:
            self.h.free()
        self.h = h
        return res.energy, res.coeffs, res.converged

    def close(self):
        if self.h is not None:
            self.h.free()
            self.h = None
