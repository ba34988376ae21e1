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


class _Lists(dict):
    """Excitation lists built on first use (a walk touches only the pairs of
    occupied orbitals)."""

    def __init__(self, make):
        super().__init__()
        self._make = make

    def __missing__(self, key):
        v = self._make(key)
        self[key] = v
        return v


def _excite_lists(ints, n_core: int, n_up: int, n_dn: int):
    """The sorted excitation lists of init_excite_generator
    (src/excite/init.rs:81-249): valence = the orbitals after the n_core lowest
    h_pp (stable), targets looped in valence order and stably sorted by
    descending magnitude; each entry carries sum_remaining_h_squared
    (compute_sum_remaining, :288-309: 0 for the last entry, then a running sum
    of the following entries' squares from the end, added one at a time).
    Lists are built when first read; the three maxima over all lists are
    computed up front.  Entries are (magnitude, target, sum_remaining_h_squared,
    sum_remaining_abs_h)."""
    n = ints.n_orb
    h, eri = ints.dense()
    order = sorted(range(n), key=lambda p: ints.h1(p, p))
    val = np.array(order[n_core:], dtype=np.int64)
    vlist = val.tolist()

    def finish(a, targets):
        idx = np.argsort(-a, kind="stable")
        a = a[idx]
        rem = np.zeros(a.size)
        rabs = np.zeros(a.size)
        if a.size > 1:
            rem[:-1] = np.cumsum((a[1:] * a[1:])[::-1])[::-1]  # sequential, from the end
            rabs[:-1] = np.cumsum(a[1:][::-1])[::-1]  # sum_remaining_abs_h
        return [(float(x), targets[i], float(r), float(b)) for x, i, r, b in zip(a, idx, rem, rabs)]

    vr, vs = np.meshgrid(val, val, indexing="ij")  # r outer, s inner

    def make_opp(key):
        p, q = key
        m = (vr != p) & (vs != q)
        return finish(np.abs(eri[p, vr[m], q, vs[m]]), list(zip(vr[m].tolist(), vs[m].tolist())))

    def make_same(key):
        p, q = key
        m = (vr != p) & (vr != q) & (vs != p) & (vs != q) & (vr < vs)  # r < s as orbital indices
        return finish(np.abs(eri[p, vr[m], q, vs[m]] - eri[p, vs[m], q, vr[m]]), list(zip(vr[m].tolist(), vs[m].tolist())))

    def sing_bound(p, r):
        qs = [q for q in range(n) if q not in (p, r)]
        vsame = sorted((eri[p, r, qs, qs] - eri[p, qs, qs, r]).tolist())
        vopp = sorted(eri[p, r, qs, qs].tolist())
        lo = float(h[p, r]) + sum(vsame[:n_up - 1]) + sum(vopp[:n_dn])
        hi = float(h[p, r]) + sum(vsame[len(vsame) - (n_up - 1):]) + sum(vopp[len(vsame) - n_dn:])
        return abs(lo) if abs(lo) > abs(hi) else abs(hi)

    sing_all = {p: [(sing_bound(p, r), r) for r in vlist if r != p] for p in vlist}

    def make_sing(p):
        v = sing_all[p]
        return finish(np.array([x for x, _ in v]), [r for _, r in v])

    # maxima over every list (the reference's max_opp_doub, max_same_doub, max_sing)
    sub = eri[np.ix_(val, val, val, val)]  # [p, r, q, s] over valence positions
    k = np.arange(val.size)
    a_opp = np.abs(sub).copy()
    a_opp[k, k, :, :] = 0.0  # r == p
    a_opp[:, :, k, k] = 0.0  # s == q
    max_opp = float(a_opp.max()) if val.size > 1 else 0.0
    max_same = 0.0
    for p in vlist:
        for q in vlist:
            if p < q:
                lst = make_same((p, q))
                if lst:
                    max_same = max(max_same, lst[0][0])
    max_sing = max((max(x for x, _ in v) for v in sing_all.values() if v), default=0.0)
    return (sum(1 << int(p) for p in val), _Lists(make_opp), _Lists(make_same), _Lists(make_sing), max_opp,
            max_same, max_sing)


def _bits_asc(x: int):
    while x:
        low = x & -x
        yield low.bit_length() - 1
        x ^= low


def matmul_external_dtm(ints, dets: np.ndarray, coeffs: np.ndarray, eps: float, n_core: int, n_up: int,
                        n_dn: int):
    """Wf::approx_matmul_external_dtm_only (src/wf/mod.rs:186-339): for each
    determinant (in order) with local = eps / |c|, every list whose maximum is
    >= local is walked until its first entry below local (that entry's
    sum_remaining_h_squared is added to the determinant's remainder), targets
    already occupied are skipped, singles are kept when |H_ai| >= local, and
    every target outside the space receives H_ai c_i (summed in discovery
    order; targets in first-discovery order, add_det_with_coeff,
    :144-165).  Returns (external words (k, 4), coefficients, the index of each
    target's first discoverer and its excitation, c_i^2 * remainder per
    determinant).  Elements come from the C++ oracle."""
    valence, opp, same, sing, max_opp, max_same, max_sing = _excite_lists(ints, n_core, n_up, n_dn)
    H = pyoracle.OracleHam(ints)
    space = {(int(w[0]) | int(w[1]) << 64, int(w[2]) | int(w[3]) << 64) for w in dets}
    cand = []  # (target, source index, excitation, single)
    rem_out = np.zeros(len(dets))
    for k, w in enumerate(dets):
        up, dn, c = int(w[0]) | int(w[1]) << 64, int(w[2]) | int(w[3]) << 64, float(coeffs[k])
        local = eps / abs(c) if c != 0.0 else np.inf
        rem = 0.0
        if max_opp >= local:
            for i in _bits_asc(valence & up):
                for j in _bits_asc(valence & dn):
                    for a, (r, s), rr, _ in opp[(i, j)]:
                        if a < local:
                            rem += rr
                            break
                        if (up >> r) & 1 or (dn >> s) & 1:
                            continue
                        cand.append(((up ^ (1 << i) | (1 << r), dn ^ (1 << j) | (1 << s)), k, ("opp", (i, j), (r, s)),
                                     False))
        if max_same >= local:
            for cfg, alpha in ((up, True), (dn, False)):
                b = list(_bits_asc(valence & cfg))
                for x in range(len(b)):
                    for y in range(x + 1, len(b)):
                        i, j = b[x], b[y]
                        for a, (r, s), rr, _ in same[(i, j)]:
                            if a < local:
                                rem += rr
                                break
                            if (cfg >> r) & 1 or (cfg >> s) & 1:
                                continue
                            nw = cfg ^ (1 << i) ^ (1 << j) | (1 << r) | (1 << s)
                            cand.append(((nw, dn) if alpha else (up, nw), k, ("same", (i, j), (r, s), alpha), False))
        if max_sing >= local:
            for cfg, alpha in ((up, True), (dn, False)):
                for i in _bits_asc(valence & cfg):
                    for a, r, rr, _ in sing[i]:
                        if a < local:
                            rem += rr
                            break
                        if (cfg >> r) & 1:
                            continue
                        nw = cfg ^ (1 << i) | (1 << r)
                        cand.append(((nw, dn) if alpha else (up, nw), k, ("sing", i, r, alpha), True))
        rem_out[k] = c * c * rem
    if not cand:
        return np.zeros((0, 4), dtype=np.uint64), np.zeros(0), [], rem_out
    src = np.array([dets[k] for _, k, _, _ in cand], dtype=np.uint64)
    tgt = _words([t[0] for t, _, _, _ in cand], [t[1] for t, _, _, _ in cand])
    h, _ = H.off_diag(src, tgt)
    out: dict[tuple[int, int], list] = {}
    for (t, k, exc, single), hv in zip(cand, h):
        c = float(coeffs[k])
        local = eps / abs(c)
        if single and not abs(hv) >= local:
            continue
        if t in space:
            continue
        if t in out:
            out[t][0] += float(hv) * c
        else:
            out[t] = [float(hv) * c, k, exc]
    keys = list(out)
    words = _words([t[0] for t in keys], [t[1] for t in keys])
    return words, np.array([out[t][0] for t in keys]), [(out[t][1], out[t][2]) for t in keys], rem_out


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


def semistoch_dtm(ints, dets: np.ndarray, coeffs: np.ndarray, e0: float, eps: float, n_core: int, n_up: int,
                  n_dn: int):
    """The deterministic stage of new_semistoch_enpt2_dtm_diag_singles
    (src/semistoch/mod.rs:221-334), the reference's default PT2
    (src/config.rs:253, src/pt/mod.rs:52-56).  For each determinant i (in
    order) and each of its valence electron pairs and electrons in the order
    of valence_elecs_and_epairs (opposite-spin pairs, same-spin pairs up then
    down, electrons up then down; src/utils/bits.rs:76-123), the list is read
    through excites_from_det_and_orbs (src/excite/iterator.rs:113-130): targets
    already occupied and targets inside the space are skipped before anything
    else.  The first remaining entry with |stored magnitude| * |c_i| < eps ends
    the list and records (i, orbitals, spin, sum_remaining_abs_h,
    sum_remaining_h_squared, |c_i| sum_remaining_abs_h,
    c_i^2 sum_remaining_h_squared) -- the DetOrbSample the screened sampler is
    built from (src/stoch/mod.rs:155).  Every entry before it gives
    h = H_ai c_i: rel_probs[i] = max |h|, and h joins h_psi unless the entry
    is a single with |h| < eps (the first discoverer of a target creates it).
    Returns (h_psi words in first-discovery order, their sums, first
    discoverers (i, excitation), rel_probs, records, e2 with H_aa from
    ham_diag).  Elements from the C++ oracle."""
    valence, opp, same, sing, max_opp, max_same, max_sing = _excite_lists(ints, n_core, n_up, n_dn)
    H = pyoracle.OracleHam(ints)
    space = {(int(w[0]) | int(w[1]) << 64, int(w[2]) | int(w[3]) << 64) for w in dets}
    n = len(dets)
    rel = np.zeros(n)
    records = []
    psi: dict[tuple[int, int], list] = {}

    def elem(src_w, tu, td):
        v, _ = H.off_diag(np.asarray(src_w, dtype=np.uint64)[None, :], _words([tu], [td]))
        return float(v[0])

    for k, w in enumerate(dets):
        up, dn, c = int(w[0]) | int(w[1]) << 64, int(w[2]) | int(w[3]) << 64, float(coeffs[k])
        ac = abs(c)
        groups = []  # (kind, is_alpha, p, q, list, target -> (up', dn'))
        for i in _bits_asc(valence & up):
            for j in _bits_asc(valence & dn):
                groups.append((0, -1, i, j, opp[(i, j)],
                               lambda t, i=i, j=j: (up ^ (1 << i) | (1 << t[0]), dn ^ (1 << j) | (1 << t[1]))))
        for cfg, alpha in ((up, 1), (dn, 0)):
            b = list(_bits_asc(valence & cfg))
            for x in range(len(b)):
                for y in range(x + 1, len(b)):
                    i, j = b[x], b[y]

                    def mk(t, i=i, j=j, cfg=cfg, alpha=alpha):
                        nw = cfg ^ (1 << i) ^ (1 << j) | (1 << t[0]) | (1 << t[1])
                        return (nw, dn) if alpha else (up, nw)
                    groups.append((1, alpha, i, j, same[(i, j)], mk))
        for cfg, alpha in ((up, 1), (dn, 0)):
            for i in _bits_asc(valence & cfg):
                def mk1(r, i=i, cfg=cfg, alpha=alpha):
                    nw = cfg ^ (1 << i) | (1 << r)
                    return (nw, dn) if alpha else (up, nw)
                groups.append((2, alpha, i, -1, sing[i], mk1))
        for kind, alpha, i, j, lst, make in groups:
            for a, tgt, rsq, rab in lst:
                if kind == 0:
                    if (up >> tgt[0]) & 1 or (dn >> tgt[1]) & 1:
                        continue
                elif kind == 1:
                    cfg = up if alpha else dn
                    if (cfg >> tgt[0]) & 1 or (cfg >> tgt[1]) & 1:
                        continue
                else:
                    cfg = up if alpha else dn
                    if (cfg >> tgt) & 1:
                        continue
                t = make(tgt)
                if t in space:
                    continue
                if a * ac < eps:
                    records.append((k, kind, alpha, i, j, rab, rsq, ac * rab, c * c * rsq))
                    break
                h = elem(w, t[0], t[1]) * c
                if abs(h) > rel[k]:
                    rel[k] = abs(h)
                if not (kind == 2 and abs(h) < eps):
                    if t in psi:
                        psi[t][0] += h
                    else:
                        exc = (kind, alpha, i, j, tgt)
                        psi[t] = [h, k, exc]
    keys = list(psi)
    words = _words([t[0] for t in keys], [t[1] for t in keys]) if keys else np.zeros((0, 4), dtype=np.uint64)
    nums = np.array([psi[t][0] for t in keys])
    e2 = 0.0
    if keys:
        for num, d in zip(nums, H.ham_diag(words)):
            e2 += num * num / (e0 - d)
    return words, nums, [(psi[t][1], psi[t][2]) for t in keys], rel, records, e2


def matmul_external_sampled(ints, dets: np.ndarray, coeffs: np.ndarray, eps: float, n_core: int, n_up: int,
                            n_dn: int, variant: str):
    """The external products that set up a screened sampler, restated from
    src/wf/mod.rs: `variant` is "separate_doubles_and_singles" (:341-533),
    "skip_singles" (:535-711), "semistoch_singles" (:714-1033) or
    "no_singles" (:1035-1193).  For each determinant (in order) with
    local = eps / |c|, its lists in the order of valence_elecs_and_epairs
    (opposite-spin pairs, same-spin pairs up then down, electrons up then
    down) are walked as the named function walks them; candidates are the
    targets outside the space, summed per target in discovery order
    (add_det_with_coeff, :144-165), and every DetOrbSample pushed is kept as
    (i, kind, is_alpha, p, q, sum_abs_h, sum_h_squared, sum_abs_hc,
    sum_hc_squared) in push order.  Returns (external words (k, 4),
    coefficients, first discoverers (i, excitation), records).  Elements from
    the C++ oracle."""
    valence, opp, same, sing, max_opp, max_same, max_sing = _excite_lists(ints, n_core, n_up, n_dn)
    H = pyoracle.OracleHam(ints)
    space = {(int(w[0]) | int(w[1]) << 64, int(w[2]) | int(w[3]) << 64) for w in dets}
    cut_first = variant in ("separate_doubles_and_singles", "no_singles")
    gate = variant == "separate_doubles_and_singles"  # lists entered only when a kind's maximum >= local
    factor = 1e-6  # single_sample_factor (:895)
    records = []
    cand = []  # (target, source index, excitation), in discovery order

    def elem(src_w, t):
        v, _ = H.off_diag(np.asarray(src_w, dtype=np.uint64)[None, :], _words([t[0]], [t[1]]))
        return float(v[0])

    def add(t, k, exc):
        if t not in space:
            cand.append((t, k, exc))

    for k, w in enumerate(dets):
        up, dn, c = int(w[0]) | int(w[1]) << 64, int(w[2]) | int(w[3]) << 64, float(coeffs[k])
        local = eps / abs(c) if c != 0.0 else np.inf

        def record(kind, alpha, i, j, rab, rsq, scale=1.0):
            if scale == 1.0:
                records.append((k, kind, alpha, i, j, rab, rsq, abs(c) * rab, c * c * rsq))
            else:
                records.append((k, kind, alpha, i, j, rab * scale, rsq * scale, abs(c) * rab * scale,
                                c * c * rsq * scale))

        doubles = []  # (kind, alpha, i, j, list, occupied test, target)
        for i in _bits_asc(valence & up):
            for j in _bits_asc(valence & dn):
                doubles.append((0, -1, i, j, opp[(i, j)],
                                lambda t: (up >> t[0]) & 1 or (dn >> t[1]) & 1,
                                lambda t, i=i, j=j: (up ^ (1 << i) | (1 << t[0]), dn ^ (1 << j) | (1 << t[1]))))
        for cfg, alpha in ((up, 1), (dn, 0)):
            b = list(_bits_asc(valence & cfg))
            for x in range(len(b)):
                for y in range(x + 1, len(b)):
                    i, j = b[x], b[y]

                    def mk(t, i=i, j=j, cfg=cfg, alpha=alpha):
                        nw = cfg ^ (1 << i) ^ (1 << j) | (1 << t[0]) | (1 << t[1])
                        return (nw, dn) if alpha else (up, nw)
                    doubles.append((1, alpha, i, j, same[(i, j)],
                                    lambda t, cfg=cfg: (cfg >> t[0]) & 1 or (cfg >> t[1]) & 1, mk))
        for kind, alpha, i, j, lst, occupied, make in doubles:
            if gate and (max_opp if kind == 0 else max_same) < local:
                continue
            for a, tgt, rsq, rab in lst:
                if cut_first and a < local:
                    record(kind, alpha, i, j, rab, rsq)
                    break
                if occupied(tgt):
                    continue
                t = make(tgt)
                if not cut_first and a < local:
                    if t in space:
                        continue
                    record(kind, alpha, i, j, rab, rsq)
                    break
                add(t, k, (kind, alpha, i, j, tgt))
        if variant == "skip_singles":
            continue
        for cfg, alpha in ((up, 1), (dn, 0)):
            for i in _bits_asc(valence & cfg):
                lst = sing[i]

                def make1(r, i=i, cfg=cfg, alpha=alpha):
                    nw = cfg ^ (1 << i) | (1 << r)
                    return (nw, dn) if alpha else (up, nw)
                if variant == "no_singles":
                    _, _, rsq, rab = lst[0]
                    record(2, alpha, i, -1, rab, rsq)
                    continue
                if max_sing < local:
                    continue
                stored = False
                for a, r, rsq, rab in lst:
                    if variant == "separate_doubles_and_singles":
                        if a < local:
                            record(2, alpha, i, -1, rab, rsq)
                            break
                        if (cfg >> r) & 1:
                            continue
                        t = make1(r)
                        add(t, k, (2, alpha, i, -1, r))
                        continue
                    # semistoch_singles
                    if a >= local:
                        if (cfg >> r) & 1:
                            continue
                        t = make1(r)
                        if t in space:
                            continue
                        hs = elem(w, t)
                        if abs(hs) >= local:
                            add(t, k, (2, alpha, i, -1, r))
                        elif not stored:
                            record(2, alpha, i, -1, rab, rsq)
                            stored = True
                    elif stored:
                        break
                    else:
                        if (cfg >> r) & 1:
                            continue
                        t = make1(r)
                        if t in space:
                            continue
                        record(2, alpha, i, -1, rab, rsq, factor)
                        break
    out: dict[tuple[int, int], list] = {}
    if cand:
        src = np.array([dets[k] for _, k, _ in cand], dtype=np.uint64)
        h, _ = H.off_diag(src, _words([t[0] for t, _, _ in cand], [t[1] for t, _, _ in cand]))
        for (t, k, exc), hv in zip(cand, h):
            x = float(hv) * float(coeffs[k])
            if t in out:
                out[t][0] += x
            else:
                out[t] = [x, k, exc]
    keys = list(out)
    words = _words([t[0] for t in keys], [t[1] for t in keys]) if keys else np.zeros((0, 4), dtype=np.uint64)
    return words, np.array([out[t][0] for t in keys]), [(out[t][1], out[t][2]) for t in keys], records
