"""The oracle's Slater-Condon restatement against brute-force second quantisation,
and the reference's own known answers for the bit helpers."""
import itertools

import numpy as np
import pytest

from risq_amd.dets import full_space, to_int_pairs
from risq_amd.fcidump import combine_2, combine_4, synthetic_molecular
from sq_brute import ham_matrix


# reference src/utils/bits.rs:198-207 (test_parity inputs; expected = brute-force bit count)
PARITY_KAT = [14, 15, 27, 1919, 4958202, 15 << 64, 1 << 127, (1 << 126) + (1 << 65)]


def test_parity_known_answers(oracle_lib):
    for v in PARITY_KAT:
        expect = 1 - 2 * (bin(v).count("1") & 1)
        assert oracle_lib.parity(v) == expect


def test_permute_is_sign_of_single_excitation(oracle_lib):
    """permute(d1, d2) equals the sign of a+_j a_i |d1> (i removed, j added)."""
    rng = np.random.default_rng(0)
    for _ in range(3000):
        n = 70
        occ = rng.choice(n, 6, replace=False)
        d1 = sum(1 << int(o) for o in occ)
        i = int(rng.choice(occ))
        j = int(rng.choice([o for o in range(n) if not (d1 >> o) & 1]))
        d2 = (d1 & ~(1 << i)) | (1 << j)
        s1 = -1 if bin(d1 & ((1 << i) - 1)).count("1") & 1 else 1
        o = d1 & ~(1 << i)
        s2 = -1 if bin(o & ((1 << j) - 1)).count("1") & 1 else 1
        assert oracle_lib.permute(d1, d2) == s1 * s2
        assert oracle_lib.permute(d2, d1) == s1 * s2


def test_permute_2_is_sign_of_double_excitation(oracle_lib):
    """permute_2 with ind = [p<q removed, r<s added] equals the sign of
    a+_r a+_s a_q a_p |d1> (reference src/utils/ints.rs:92-97)."""
    rng = np.random.default_rng(1)
    for _ in range(3000):
        n = 90
        occ = rng.choice(n, 8, replace=False)
        d1 = sum(1 << int(o) for o in occ)
        p, q = sorted(int(x) for x in rng.choice(occ, 2, replace=False))
        free = [o for o in range(n) if not (d1 >> o) & 1]
        r, s = sorted(int(x) for x in rng.choice(free, 2, replace=False))
        d2 = (d1 & ~(1 << p) & ~(1 << q)) | (1 << r) | (1 << s)
        sign, o = 1, d1
        for k, create in ((p, False), (q, False), (s, True), (r, True)):
            sign *= -1 if bin(o & ((1 << k) - 1)).count("1") & 1 else 1
            o = (o | (1 << k)) if create else (o & ~(1 << k))
        assert o == d2
        assert oracle_lib.permute_2(d1, d2, [p, q, r, s]) == sign


def test_combine_closed_form_and_symmetry(oracle_lib):
    for p in range(1, 9):
        for q in range(1, 9):
            assert oracle_lib.lib().or_combine_2(p, q) == combine_2(p, q) == \
                (max(p, q) * (max(p, q) - 1)) // 2 + min(p, q)
    seen = {}
    for p, q, r, s in itertools.product(range(1, 6), repeat=4):
        c = combine_4(p, q, r, s)
        assert oracle_lib.lib().or_combine_4(p, q, r, s) == c
        for perm in [(q, p, r, s), (p, q, s, r), (r, s, p, q), (s, r, q, p)]:
            assert combine_4(*perm) == c
        canon = tuple(sorted([tuple(sorted((p, q))), tuple(sorted((r, s)))]))
        assert seen.setdefault(c, canon) == canon  # injective on canonical quadruples


@pytest.mark.parametrize("n_orb,n_up,n_dn,seed", [(5, 2, 2, 3), (5, 3, 1, 4), (6, 2, 3, 5)])
def test_elements_match_second_quantisation(oracle_lib, n_orb, n_up, n_dn, seed):
    ints = synthetic_molecular(n_orb, 6, seed)
    ints.e_core = 0.37
    h, eri = ints.dense()
    words = full_space(n_orb, n_up, n_dn)
    ref = ham_matrix(h, eri, ints.e_core, to_int_pairs(words))
    H = oracle_lib.OracleHam(ints)
    diag = H.ham_diag(words)
    assert np.allclose(diag, np.diag(ref), rtol=0, atol=1e-12)
    m = words.shape[0]
    a, b = np.meshgrid(np.arange(m), np.arange(m), indexing="ij")
    a, b = a.ravel(), b.ravel()
    off = a != b
    val, deg = H.off_diag(words[a[off]], words[b[off]])
    expect = ref[b[off], a[off]]
    assert np.max(np.abs(val - expect)) < 1e-12
    # zero exactly when more than a double, and signs of nonzero elements agree
    assert np.all(val[deg == 3] == 0.0)
    nz = np.abs(expect) > 1e-10
    assert np.array_equal(np.sign(val[nz]), np.sign(expect[nz]))


def test_double_connection_truth_table(oracle_lib):
    """Reference src/var/fast_ham_tests.rs:165-178: alpha-string pairs and
    whether they are double-connected; here read through the excitation degree
    of determinants that differ only in the alpha string."""
    from risq_amd.dets import from_orbitals
    cases = [([0, 1], [2, 3], True), ([0, 1], [0, 2], False), ([0, 1], [0, 1], False),
             ([0, 1, 2], [0, 3, 4], True), ([0, 1, 2], [0, 1, 3], False), ([0, 1, 2], [3, 4, 5], False),
             ([0, 1, 2, 3], [0, 1, 4, 5], True)]
    ints = synthetic_molecular(6, 4, 0)
    H = oracle_lib.OracleHam(ints)
    for a, b, expect in cases:
        d1 = from_orbitals([(a, [0])])
        d2 = from_orbitals([(b, [0])])
        _, deg = H.off_diag(d1, d2)
        assert (deg[0] == 2) == expect, (a, b)
