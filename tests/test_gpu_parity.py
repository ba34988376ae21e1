"""GPU parity: the HIP library (through its C ABI) against the CPU oracle.

Bar: integer/pattern/sign work bit-exact; FP64 matrix elements bit-exact (the
element kernels keep the reference's summation order); sigma products within
1e-12 relative (atomic FP64 reordering); eigenvalues within 1 uHa.
"""
import numpy as np
import pytest

from helpers import csr_to_dense, hubbard_space, mol_space
from risq_amd.dets import from_orbitals, full_space
from risq_amd.hci import PREC_F32, SIGMA_F32, SINGLES_NONZERO, STORE_FULL, STORE_TILED, STORE_UPPER, HciError

pytestmark = pytest.mark.gpu

MOL0 = (16, 32, 0, 4, 4, 3000)     # config-0 recipe, energy-truncated space


def _oracle_build(oracle_lib, ints, dets, diag, n_stored=0, prev=None):
    H = oracle_lib.OracleHam(ints)
    S = prev if prev is not None else oracle_lib.OracleSparse()
    S.build(H, dets, diag, n_stored=n_stored)
    return S


def _random_pairs(dets, rng, m):
    n = dets.shape[0]
    a = rng.integers(0, n, m)
    b = rng.integers(0, n, m)
    return dets[a], dets[b]


def test_diag_bit_exact(gpu_ctx, oracle_lib):
    ints, dets = mol_space(*MOL0)
    gpu_ctx.set_integrals(ints)
    g = gpu_ctx.ham_diag(dets)
    o = oracle_lib.OracleHam(ints).ham_diag(dets)
    assert np.array_equal(g.view(np.uint64), o.view(np.uint64))


def test_elements_bit_exact(gpu_ctx, oracle_lib):
    ints, dets = mol_space(*MOL0)
    gpu_ctx.set_integrals(ints)
    rng = np.random.default_rng(7)
    d1, d2 = _random_pairs(dets, rng, 200000)
    # make sure every degree is present: neighbours in the energy ordering are close
    d1 = np.concatenate([d1, dets[:-1]])
    d2 = np.concatenate([d2, dets[1:]])
    gv, gd = gpu_ctx.elements(d1, d2)
    ov, od = oracle_lib.OracleHam(ints).off_diag(d1, d2)
    assert np.array_equal(gd, od)
    assert set(np.unique(gd)) >= {0, 1, 2, 3}
    assert np.array_equal(gv.view(np.uint64), ov.view(np.uint64))
    # symmetry H_ij == H_ji bitwise on connected pairs (up to the sign of zero)
    gv2, _ = gpu_ctx.elements(d2, d1)
    nz = gv != 0
    assert np.array_equal(gv[nz].view(np.uint64), gv2[nz].view(np.uint64))


@pytest.mark.parametrize("storage", [STORE_FULL, STORE_UPPER, STORE_TILED])
@pytest.mark.parametrize("space", [MOL0, (16, 32, 3, 3, 5, 2500)])
def test_build_matches_oracle(gpu_ctx, oracle_lib, space, storage):
    ints, dets = mol_space(*space)
    gpu_ctx.set_integrals(ints)
    diag = oracle_lib.OracleHam(ints).ham_diag(dets)
    h = gpu_ctx.build_h(dets, diag, storage=storage)
    rp, col, val, dg = h.download()
    S = _oracle_build(oracle_lib, ints, dets, diag)
    orp, ocol, oval, odg = S.export()
    assert h.nnz == S.nnz
    assert np.array_equal(rp, orp)
    assert np.array_equal(col, ocol)
    assert np.array_equal(val.view(np.uint64), oval.view(np.uint64))
    assert np.array_equal(dg, odg)


def test_build_u128_strings(gpu_ctx, oracle_lib):
    """n_orb > 64: two-word strings through the same kernels."""
    ints, dets = mol_space(70, 8, 5, 3, 3, 1500)
    # force use of high orbitals: relabel orbitals p -> 69 - p
    words = []
    for up_lo, up_hi, dn_lo, dn_hi in dets.astype(object):
        u = int(up_lo) | (int(up_hi) << 64)
        d = int(dn_lo) | (int(dn_hi) << 64)
        u2 = sum(1 << (69 - p) for p in range(70) if (u >> p) & 1)
        d2 = sum(1 << (69 - p) for p in range(70) if (d >> p) & 1)
        m = (1 << 64) - 1
        words.append([u2 & m, u2 >> 64, d2 & m, d2 >> 64])
    dets = np.array(words, dtype=np.uint64)
    gpu_ctx.set_integrals(ints)
    diag = oracle_lib.OracleHam(ints).ham_diag(dets)
    assert np.array_equal(gpu_ctx.ham_diag(dets).view(np.uint64), diag.view(np.uint64))
    h = gpu_ctx.build_h(dets, diag)
    rp, col, val, _ = h.download()
    orp, ocol, oval, _ = _oracle_build(oracle_lib, ints, dets, diag).export()
    assert np.array_equal(rp, orp) and np.array_equal(col, ocol)
    assert np.array_equal(val.view(np.uint64), oval.view(np.uint64))


@pytest.mark.parametrize("storage", [STORE_FULL, STORE_UPPER, STORE_TILED])
def test_build_hubbard_doubles_vanish(gpu_ctx, oracle_lib, storage):
    """Hubbard: every double is exactly zero and dropped; singles are +-t hops.
    Hops between non-adjacent sites are exact zeros that the storage rule drops
    (the tiled layout does not place them at all: their element vanishes for
    every occupation)."""
    ints, dets = hubbard_space(3, 3, 4, 4, 2000)
    gpu_ctx.set_integrals(ints)
    diag = oracle_lib.OracleHam(ints).ham_diag(dets)
    h = gpu_ctx.build_h(dets, diag, storage=storage)
    rp, col, val, _ = h.download()
    S = _oracle_build(oracle_lib, ints, dets, diag)
    orp, ocol, oval, _ = S.export()
    assert h.nnz == S.nnz
    assert np.array_equal(rp, orp) and np.array_equal(col, ocol)
    assert np.array_equal(val.view(np.uint64), oval.view(np.uint64))
    assert np.all(np.isin(np.abs(val), [1.0]))
    x = np.random.default_rng(3).standard_normal(len(dets))
    y = h.matrix_vector_prod(x)
    yo = S.spmv(x)
    assert np.max(np.abs(y - yo)) <= 1e-12 * np.max(np.abs(yo))


@pytest.mark.parametrize("storage", [STORE_FULL, STORE_UPPER])
def test_incremental_build(gpu_ctx, oracle_lib, storage):
    ints, dets = mol_space(*MOL0)
    gpu_ctx.set_integrals(ints)
    diag = oracle_lib.OracleHam(ints).ham_diag(dets)
    n1 = 1200
    h1 = gpu_ctx.build_h(dets[:n1], diag[:n1], storage=storage)
    h2 = gpu_ctx.build_h(dets, diag, prev=h1, n_stored=n1, storage=storage)
    full = gpu_ctx.build_h(dets, diag, storage=storage)
    a, b = h2.download(), full.download()
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
    S = oracle_lib.OracleSparse()
    H = oracle_lib.OracleHam(ints)
    S.build(H, dets[:n1], diag[:n1])
    S.build(H, dets, diag, n_stored=n1)
    o = S.export()
    for x, y in zip(a, o):
        assert np.array_equal(x, y)


def test_build_rejects_bad_input(gpu_ctx, oracle_lib):
    ints, dets = mol_space(*MOL0)
    gpu_ctx.set_integrals(ints)
    diag = np.zeros(dets.shape[0])
    with pytest.raises(HciError, match="duplicate"):
        gpu_ctx.build_h(np.concatenate([dets[:10], dets[:1]]), np.zeros(11))
    bad2 = dets[:10].copy()
    bad2[4, 0] = np.uint64(0b1)  # one up electron instead of four
    with pytest.raises(HciError, match="electron counts"):
        gpu_ctx.build_h(bad2, diag[:10])
    with pytest.raises(HciError):
        gpu_ctx.build_h(dets[:10], diag[:10], n_stored=3)


@pytest.mark.parametrize("storage", [STORE_FULL, STORE_UPPER, STORE_TILED])
def test_small_n_build_matches_oracle(gpu_ctx, oracle_lib, storage):
    """The reference's n <= 100 build (gen_sparse_ham_fast, algo 1/1,
    src/var/ham_gen.rs:60-292) stores singles iff H != 0 (off_diag.rs:33).  On a
    spin-flip-closed space whose integrals make some singles exactly 1e-13 and
    others exact zeros, the library's HCI_SINGLES_NONZERO build equals the
    oracle's restatement of that build bit for bit (and keeps the tiny singles
    the |H| > 1e-12 rule drops)."""
    from helpers import tiny_single_molecular
    ints = tiny_single_molecular()
    dets = full_space(8, 3, 3)
    gpu_ctx.set_integrals(ints)
    H = oracle_lib.OracleHam(ints)
    diag = H.ham_diag(dets)
    o = oracle_lib.OracleSparse().build_small(H, dets, diag).export()
    h = gpu_ctx.build_h(dets, diag, single_rule=SINGLES_NONZERO, storage=storage)
    g = h.download()
    assert np.array_equal(g[0], o[0]) and np.array_equal(g[1], o[1])
    assert np.array_equal(g[2].view(np.uint64), o[2].view(np.uint64))
    assert np.any(np.abs(g[2]) == 1e-13)
    x = np.random.default_rng(12).standard_normal(dets.shape[0])
    ref = oracle_lib.spmv_csr(*o, x)
    y = h.matrix_vector_prod(x)
    assert np.max(np.abs(y - ref)) <= 1e-12 * np.max(np.abs(ref))


@pytest.mark.parametrize("storage", [STORE_FULL, STORE_UPPER])
def test_small_n_then_fast_incremental(gpu_ctx, oracle_lib, storage):
    """The HCI loop's rule switch: rows built while the space held <= 100
    determinants keep their H != 0 singles; pairs added later, with a new index
    >= n_stored, follow the |H| > 1e-12 rule (src/var/davidson.rs:55-61).  The
    library's two incremental builds equal the oracle's n <= 100 build followed
    by its fast-lookup build."""
    from helpers import spin_flip_closed, tiny_single_molecular
    ints = tiny_single_molecular()
    dets = spin_flip_closed(full_space(8, 3, 3))
    n1 = 99  # a spin-flip-closed prefix
    assert np.array_equal(spin_flip_closed(dets[:n1]), dets[:n1])
    gpu_ctx.set_integrals(ints)
    H = oracle_lib.OracleHam(ints)
    diag = H.ham_diag(dets)
    S = oracle_lib.OracleSparse()
    S.build_small(H, dets[:n1], diag[:n1])
    S.build(H, dets, diag, n_stored=n1)
    o = S.export()
    h1 = gpu_ctx.build_h(dets[:n1], diag[:n1], single_rule=SINGLES_NONZERO, storage=storage)
    h2 = gpu_ctx.build_h(dets, diag, prev=h1, n_stored=n1, storage=storage)
    g = h2.download()
    assert np.array_equal(g[0], o[0]) and np.array_equal(g[1], o[1])
    assert np.array_equal(g[2].view(np.uint64), o[2].view(np.uint64))
    rows = np.repeat(np.arange(dets.shape[0]), np.diff(g[0]))
    tiny = np.abs(g[2]) <= 1e-12
    assert tiny.any() and np.all(rows[tiny] < n1) and np.all(g[1][tiny] < n1)


@pytest.mark.parametrize("storage", [STORE_FULL, STORE_UPPER, STORE_TILED])
def test_sigma_matches_oracle(gpu_ctx, oracle_lib, storage):
    ints, dets = mol_space(*MOL0)
    gpu_ctx.set_integrals(ints)
    diag = oracle_lib.OracleHam(ints).ham_diag(dets)
    h = gpu_ctx.build_h(dets, diag, storage=storage)
    S = _oracle_build(oracle_lib, ints, dets, diag)
    rng = np.random.default_rng(3)
    X = rng.standard_normal((dets.shape[0], 3))
    Y = h.matrix_vector_prod(X)
    for c in range(3):
        ref = S.spmv(X[:, c])
        assert np.max(np.abs(Y[:, c] - ref)) <= 1e-12 * np.max(np.abs(ref))


def test_sigma_linearity_and_symmetry(gpu_ctx, oracle_lib):
    """Size-independent checks: H(ax+by) = aHx + bHy and x.Hy = y.Hx."""
    ints, dets = mol_space(16, 32, 0, 4, 4, 20000)
    gpu_ctx.set_integrals(ints)
    diag = gpu_ctx.ham_diag(dets)
    h = gpu_ctx.build_h(dets, diag)
    rng = np.random.default_rng(4)
    x, y = rng.standard_normal((2, dets.shape[0]))
    hx, hy = h.matrix_vector_prod(x), h.matrix_vector_prod(y)
    hz = h.matrix_vector_prod(2.0 * x - 3.0 * y)
    assert np.max(np.abs(hz - (2 * hx - 3 * hy))) < 1e-10 * np.max(np.abs(hz))
    assert abs(x @ hy - y @ hx) < 1e-10 * abs(x @ hy)


def test_davidson_energy(gpu_ctx, oracle_lib):
    ints, dets = mol_space(16, 32, 0, 4, 4, 1500)
    gpu_ctx.set_integrals(ints)
    diag = oracle_lib.OracleHam(ints).ham_diag(dets)
    h = gpu_ctx.build_h(dets, diag)
    ref = np.linalg.eigvalsh(csr_to_dense(*h.download()))[0]
    res = gpu_ctx.davidson(h, tol=1e-5, etol=1e-9, max_iters=300)
    assert res.converged
    assert abs(res.energy - ref) < 1e-6
    hx = h.matrix_vector_prod(res.coeffs)
    assert np.linalg.norm(hx - res.energy * res.coeffs) < 1e-4
    # the oracle's Davidson (reference control flow) reaches the same root
    rc, e0, _, _, _ = _oracle_build(oracle_lib, ints, dets, diag).davidson(tol=1e-5, etol=1e-9, max_iters=300)
    assert rc == 0 and abs(e0 - ref) < 1e-6


def test_davidson_warm_start(gpu_ctx, oracle_lib):
    ints, dets = mol_space(16, 32, 0, 4, 4, 1500)
    gpu_ctx.set_integrals(ints)
    diag = oracle_lib.OracleHam(ints).ham_diag(dets)
    h = gpu_ctx.build_h(dets, diag)
    cold = gpu_ctx.davidson(h, tol=1e-5, etol=1e-9, max_iters=300)
    warm = gpu_ctx.davidson(h, init=cold.coeffs, tol=1e-5, etol=1e-9, max_iters=300)
    assert warm.converged and warm.stats["iters"] <= 3
    assert abs(warm.energy - cold.energy) < 1e-8


def test_davidson_not_converged_keeps_outputs(gpu_ctx, oracle_lib):
    ints, dets = mol_space(16, 32, 0, 4, 4, 1500)
    gpu_ctx.set_integrals(ints)
    h = gpu_ctx.build_h(dets, gpu_ctx.ham_diag(dets))
    res = gpu_ctx.davidson(h, max_iters=2)
    assert not res.converged and res.energy == 0.0 and not res.coeffs.any()


def test_fp32_storage_and_mixed_davidson(gpu_ctx, oracle_lib):
    ints, dets = mol_space(16, 32, 0, 4, 4, 1500)
    gpu_ctx.set_integrals(ints)
    diag = oracle_lib.OracleHam(ints).ham_diag(dets)
    h64 = gpu_ctx.build_h(dets, diag)
    h32 = gpu_ctx.build_h(dets, diag, value_prec=PREC_F32)
    a, b = h64.download(), h32.download()
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    # hi+lo keeps ~48 mantissa bits
    assert np.max(np.abs(a[2] - b[2]) / np.abs(a[2])) < 1e-13
    x = np.random.default_rng(1).standard_normal(dets.shape[0])
    y64 = h64.matrix_vector_prod(x)
    y32 = h32.matrix_vector_prod(x, mode=SIGMA_F32)
    assert np.max(np.abs(y64 - y32)) < 1e-5 * np.max(np.abs(y64))
    ref = gpu_ctx.davidson(h64, tol=1e-5, etol=1e-9, max_iters=300)
    mix = gpu_ctx.davidson(h32, tol=1e-5, etol=1e-9, max_iters=300, mixed=True, refine_tol=1e-3)
    assert mix.converged and mix.stats["n_sigma_f32"] > 0 and mix.stats["refine_iter"] >= 0
    assert abs(mix.energy - ref.energy) < 1e-6


def test_mixed_floor_hands_over_to_refinement(gpu_ctx, oracle_lib):
    """With etol = 0 no phase can meet the energy test, so a reduced-precision
    phase ends at a stall or at the 1e-9 correction floors
    (davidson.rs:210-218, modified_gram_schmidt.rs:29).  Those floors must hand
    over to full-precision sigmas (only they may declare convergence): the
    result is the eigenvalue of the stored hi + lo matrix, not of its FP32 or
    FP16 roundings."""
    from risq_amd.hci import PREC_F16
    ints, dets = mol_space(16, 32, 0, 4, 4, 1500)
    gpu_ctx.set_integrals(ints)
    diag = oracle_lib.OracleHam(ints).ham_diag(dets)
    h32 = gpu_ctx.build_h(dets, diag, value_prec=PREC_F32)
    ref = np.linalg.eigvalsh(csr_to_dense(*h32.download()))[0]
    r = gpu_ctx.davidson(h32, tol=1e-5, etol=0.0, max_iters=800, mixed=1)
    assert r.converged and r.stats["n_sigma_full"] > 0 and r.stats["refine_iter"] >= 0, r.stats
    assert abs(r.energy - ref) < 1e-10, (r.energy, ref)
    h16 = gpu_ctx.build_h(dets, diag, value_prec=PREC_F16, storage=STORE_TILED)
    r = gpu_ctx.davidson(h16, tol=1e-5, etol=0.0, max_iters=800, mixed=2, restart_keep=2)
    assert r.converged and r.stats["n_sigma_full"] > 0 and r.stats["f16_iter"] >= 0, r.stats
    assert abs(r.energy - ref) < 1e-10, (r.energy, ref)


@pytest.mark.parametrize("storage", [STORE_FULL, STORE_UPPER, STORE_TILED])
def test_set_diagonal(gpu_ctx, oracle_lib, storage):
    """MatrixOperations::set_diagonal: the new caller-order diagonal reads back
    unchanged, and the product and Davidson use it (against the oracle's matrix
    with the same diagonal)."""
    ints, dets = mol_space(*MOL0)
    gpu_ctx.set_integrals(ints)
    diag = oracle_lib.OracleHam(ints).ham_diag(dets)
    h = gpu_ctx.build_h(dets, diag, storage=storage)
    shifted = diag + np.random.default_rng(13).uniform(0.0, 0.5, diag.size)
    h.set_diagonal(shifted)
    assert np.array_equal(h.diagonal(), shifted)
    S = _oracle_build(oracle_lib, ints, dets, shifted)
    x = np.random.default_rng(15).standard_normal(dets.shape[0])
    ref = S.spmv(x)
    assert np.max(np.abs(h.matrix_vector_prod(x) - ref)) <= 1e-12 * np.max(np.abs(ref))
    e_ref = np.linalg.eigvalsh(csr_to_dense(*S.export()))[0]
    r = gpu_ctx.davidson(h, tol=1e-6, etol=1e-10, max_iters=400)
    assert r.converged and abs(r.energy - e_ref) < 1e-6
    h.free()


def test_full_ci_tiny(gpu_ctx, oracle_lib):
    """Full CI of a 6-orbital system: every determinant, dense eigh reference."""
    from risq_amd.fcidump import synthetic_molecular
    ints = synthetic_molecular(6, 6, 11)
    dets = full_space(6, 3, 3)
    gpu_ctx.set_integrals(ints)
    diag = gpu_ctx.ham_diag(dets)
    h = gpu_ctx.build_h(dets, diag, single_rule=SINGLES_NONZERO)
    ref = np.linalg.eigvalsh(csr_to_dense(*h.download()))[0]
    res = gpu_ctx.davidson(h, tol=1e-6, etol=1e-10, max_iters=300)
    assert res.converged and abs(res.energy - ref) < 1e-8


def test_two_dets(gpu_ctx):
    """n <= 2: the reference declares convergence at once (davidson.rs:179)."""
    from risq_amd.fcidump import synthetic_molecular
    ints = synthetic_molecular(4, 4, 1)
    dets = from_orbitals([([0, 1], [0, 1]), ([0, 2], [0, 1])])
    gpu_ctx.set_integrals(ints)
    h = gpu_ctx.build_h(dets, gpu_ctx.ham_diag(dets), single_rule=SINGLES_NONZERO)
    res = gpu_ctx.davidson(h)
    assert res.converged


@pytest.mark.parametrize("storage,prec", [(STORE_FULL, 64), (STORE_UPPER, 64), (STORE_UPPER, 32), (STORE_TILED, 64),
                                          (STORE_TILED, 32), (STORE_TILED, 16)])
def test_unconnected_determinants(gpu_ctx, storage, prec):
    """Spaces with no stored pair: one determinant, and three that are
    triple excitations of one another.  Nothing is stored, every product mode
    gives diag * x exactly, and Davidson returns the lowest diagonal."""
    from risq_amd.fcidump import synthetic_molecular
    from risq_amd.hci import SIGMA_F16, SIGMA_FULL
    ints = synthetic_molecular(8, 8, 3)
    gpu_ctx.set_integrals(ints)
    modes = [SIGMA_FULL] + ([SIGMA_F32] if prec <= 32 else []) + ([SIGMA_F16] if prec == 16 else [])
    for orbs in ([([0, 1, 2], [0, 1, 2])],
                 [([0, 1, 2], [0, 1, 2]), ([3, 4, 5], [0, 1, 2]), ([0, 1, 2], [3, 4, 5])]):
        dets = from_orbitals(orbs)
        diag = gpu_ctx.ham_diag(dets)
        h = gpu_ctx.build_h(dets, diag, value_prec=prec, storage=storage)
        assert h.nnz == 0
        x = np.random.default_rng(5).standard_normal(dets.shape[0])
        for m in modes:
            assert np.array_equal(h.matrix_vector_prod(x, mode=m), diag * x), m
        res = gpu_ctx.davidson(h, tol=1e-8, etol=1e-12, max_iters=50)
        assert res.converged and res.energy == pytest.approx(float(diag.min()), abs=1e-12)
        h.free()


def test_full_storage_is_deterministic(gpu_ctx, oracle_lib):
    """FULL storage: builds and sigma products are bitwise reproducible."""
    ints, dets = mol_space(16, 32, 0, 4, 4, 20000)
    gpu_ctx.set_integrals(ints)
    diag = gpu_ctx.ham_diag(dets)
    a = gpu_ctx.build_h(dets, diag)
    b = gpu_ctx.build_h(dets, diag)
    x = np.random.default_rng(5).standard_normal(dets.shape[0])
    ya, ya2, yb = a.matrix_vector_prod(x), a.matrix_vector_prod(x), b.matrix_vector_prod(x)
    assert np.array_equal(ya.view(np.uint64), ya2.view(np.uint64))
    assert np.array_equal(ya.view(np.uint64), yb.view(np.uint64))


def test_row_accessor_matches_download(gpu_ctx, oracle_lib):
    ints, dets = mol_space(*MOL0)
    gpu_ctx.set_integrals(ints)
    h = gpu_ctx.build_h(dets, gpu_ctx.ham_diag(dets))
    rp, col, val, _ = h.download()
    for i in [0, 1, 7, 100, dets.shape[0] - 2, dets.shape[0] - 1]:
        c, v = h.row(i)
        assert np.array_equal(c, col[rp[i]:rp[i + 1]])
        assert np.array_equal(v, val[rp[i]:rp[i + 1]])


def test_thick_restart_same_energy(gpu_ctx, oracle_lib):
    ints, dets = mol_space(16, 32, 0, 4, 4, 1500)
    gpu_ctx.set_integrals(ints)
    h = gpu_ctx.build_h(dets, gpu_ctx.ham_diag(dets))
    ref = np.linalg.eigvalsh(csr_to_dense(*h.download()))[0]
    a = gpu_ctx.davidson(h, tol=1e-5, etol=1e-9, max_iters=400)
    b = gpu_ctx.davidson(h, tol=1e-5, etol=1e-9, max_iters=400, restart_keep=2)
    assert a.converged and b.converged
    assert abs(a.energy - ref) < 1e-6 and abs(b.energy - ref) < 1e-6
    assert b.stats["iters"] <= a.stats["iters"]


@pytest.mark.parametrize("max_sub,keep", [(4, 1), (8, 3), (11, 2), (16, 4), (17, 4), (32, 8)])
def test_subspace_sizes_same_energy(gpu_ctx, oracle_lib, max_sub, keep):
    """The subspace's vector kernels come in three widths (up to 8, 16 and 32
    columns, chosen from max_sub): every width, at its edges, converges to the
    dense lowest eigenvalue, and its vector is an eigenvector."""
    ints, dets = mol_space(16, 32, 0, 4, 4, 1500)
    gpu_ctx.set_integrals(ints)
    h = gpu_ctx.build_h(dets, gpu_ctx.ham_diag(dets))
    ref = np.linalg.eigvalsh(csr_to_dense(*h.download()))[0]
    r = gpu_ctx.davidson(h, tol=1e-5, etol=1e-9, max_iters=600, max_sub=max_sub, restart_keep=keep)
    assert r.converged
    assert abs(r.energy - ref) < 1e-6
    hx = h.matrix_vector_prod(r.coeffs)
    assert np.linalg.norm(hx - r.energy * r.coeffs) < 1e-4


@pytest.mark.parametrize("prec", [64, 32])
def test_tiled_storage_large_groups(gpu_ctx, oracle_lib, prec):
    """Alpha groups larger than a tile (256 rows) and column groups wider than a
    column window (4096) split into several sub-blocks; sigma and download agree
    with full storage."""
    ints, dets = mol_space(16, 32, 0, 4, 4, 40000)
    gpu_ctx.set_integrals(ints)
    diag = gpu_ctx.ham_diag(dets)
    f = gpu_ctx.build_h(dets, diag, value_prec=prec)
    t = gpu_ctx.build_h(dets, diag, value_prec=prec, storage=STORE_TILED)
    assert t.nnz == f.nnz
    assert t.build_stats["n_alpha"] == f.build_stats["n_alpha"]
    x = np.random.default_rng(8).standard_normal(dets.shape[0])
    for mode in ([0, 1] if prec == 32 else [0]):
        a, b = f.matrix_vector_prod(x, mode=mode), t.matrix_vector_prod(x, mode=mode)
        assert np.max(np.abs(a - b)) < 1e-12 * np.max(np.abs(a))
    for u, v in zip(f.download(), t.download()):
        assert np.array_equal(u, v)
    ra = gpu_ctx.davidson(f, tol=1e-5, etol=1e-9, max_iters=400)
    rb = gpu_ctx.davidson(t, tol=1e-5, etol=1e-9, max_iters=400)
    assert abs(ra.energy - rb.energy) < 1e-8


@pytest.mark.parametrize("prec", [64, 32])
def test_tiled_incremental_build(gpu_ctx, oracle_lib, prec):
    """The HCI loop's chain of builds on tiled storage: an n <= 100 build
    (singles iff H != 0), then two incremental builds (|H| > 1e-12 for pairs
    with a new index).  Downloads equal the FULL-storage chain's and the
    oracle's (its n <= 100 build, then its fast-lookup builds with n_stored);
    the tiny singles of the first build survive, none is added later; the
    product matches the oracle's."""
    from helpers import spin_flip_closed, tiny_single_molecular
    ints = tiny_single_molecular()
    dets = spin_flip_closed(full_space(8, 3, 3))
    n1, n2, n3 = 99, 1500, dets.shape[0]
    gpu_ctx.set_integrals(ints)
    H = oracle_lib.OracleHam(ints)
    diag = H.ham_diag(dets)
    S = oracle_lib.OracleSparse()
    S.build_small(H, dets[:n1], diag[:n1])
    S.build(H, dets[:n2], diag[:n2], n_stored=n1)
    S.build(H, dets, diag, n_stored=n2)
    o = S.export()
    chain = {}
    for storage in (STORE_FULL, STORE_TILED):
        h1 = gpu_ctx.build_h(dets[:n1], diag[:n1], single_rule=SINGLES_NONZERO, storage=storage, value_prec=prec)
        h2 = gpu_ctx.build_h(dets[:n2], diag[:n2], prev=h1, n_stored=n1, storage=storage, value_prec=prec)
        h1.free()
        h3 = gpu_ctx.build_h(dets[:n3], diag[:n3], prev=h2, n_stored=n2, storage=storage, value_prec=prec)
        h2.free()
        chain[storage] = h3
    t = chain[STORE_TILED].download()
    for u, v in zip(chain[STORE_FULL].download(), t):
        assert np.array_equal(u, v)
    assert np.array_equal(t[0], o[0]) and np.array_equal(t[1], o[1])
    if prec == 64:
        assert np.array_equal(t[2].view(np.uint64), o[2].view(np.uint64))
    else:
        assert np.max(np.abs(t[2] - o[2]) / np.abs(o[2])) < 1e-13
    rows = np.repeat(np.arange(n3), np.diff(t[0]))
    tiny = np.abs(t[2]) <= 1e-12
    assert tiny.any() and np.all(rows[tiny] < n1) and np.all(t[1][tiny] < n1)
    x = np.random.default_rng(14).standard_normal(n3)
    ref = oracle_lib.spmv_csr(*o, x)
    y = chain[STORE_TILED].matrix_vector_prod(x)
    assert np.max(np.abs(y - ref)) <= 1e-12 * np.max(np.abs(ref))
    with pytest.raises(HciError, match="another storage kind"):
        gpu_ctx.build_h(dets, diag, prev=chain[STORE_FULL], n_stored=n3, storage=STORE_TILED)
    for h in chain.values():
        h.free()


@pytest.mark.timeout(300)
@pytest.mark.parametrize("prec", [64, 32])
def test_tiled_mirror_pairs(gpu_ctx, oracle_lib, prec):
    """Dense windows between two tiles' whole row ranges are stored once for
    both triangles (the n_orb-40 workload's first 1e5 determinants have many):
    the download equals FULL storage's, products agree with FULL storage's in
    every sigma mode (FP64 and FP32 hi + lo within 1e-12; FP32 hi within its
    rounding), the row norms of the learned correction's features come out
    the same (the learned-correction Davidson reaches the FULL-storage E0).
    Builds with FP16 copies store both halves: download and hi + lo product
    equal FULL storage's, the FP32 and FP16 products within their
    roundings."""
    from risq_amd.hci import PREC_F16, PRECOND_MLP, SIGMA_F16, SIGMA_F32
    from risq_amd import precond_train
    ints, dets = mol_space(40, 80, 1, 6, 6, 100000)
    gpu_ctx.set_integrals(ints)
    precond_train.load(gpu_ctx)
    diag = gpu_ctx.ham_diag(dets)
    full = gpu_ctx.build_h(dets, diag, value_prec=prec)
    t = gpu_ctx.build_h(dets, diag, value_prec=prec, storage=STORE_TILED)
    print(f"mirror pairs {t.build_stats['mirror_pairs']} nnz_u {t.nnz}")
    assert t.build_stats["mirror_pairs"] > 0
    assert t.nnz == full.nnz
    for u, v in zip(full.download(), t.download()):
        assert np.array_equal(u, v)
    x = np.random.default_rng(16).standard_normal(dets.shape[0])
    a = full.matrix_vector_prod(x)
    off = np.linalg.norm(a - diag * x)
    assert np.max(np.abs(t.matrix_vector_prod(x) - a)) < 1e-12 * np.max(np.abs(a))
    if prec == 32:
        assert np.linalg.norm(t.matrix_vector_prod(x, mode=SIGMA_F32) - a) / off < 1e-6
    ref = gpu_ctx.davidson(full, tol=1e-5, etol=1e-9, max_iters=600)
    r = gpu_ctx.davidson(t, tol=1e-5, etol=1e-9, max_iters=600, mixed=1 if prec == 32 else 0, restart_keep=4,
                         max_sub=16, precond=PRECOND_MLP)
    assert ref.converged and r.converged and abs(r.energy - ref.energy) < 1e-6
    t.free()
    f16 = gpu_ctx.build_h(dets, diag, value_prec=PREC_F16, storage=STORE_TILED)
    assert f16.build_stats["mirror_pairs"] == 0 and f16.nnz == full.nnz
    if prec == 32:
        for u, v in zip(full.download(), f16.download()):
            assert np.array_equal(u, v)
        assert np.max(np.abs(f16.matrix_vector_prod(x) - a)) < 1e-12 * np.max(np.abs(a))
        assert np.linalg.norm(f16.matrix_vector_prod(x, mode=SIGMA_F32) - a) / off < 1e-6
        assert np.linalg.norm(f16.matrix_vector_prod(x, mode=SIGMA_F16) - a) / off < 2e-3
    f16.free()
    full.free()


def test_recycled_entry_arrays(gpu_ctx, oracle_lib):
    """A recycled tiled matrix's arrays serve the next tiled builds (FP32 after
    FP32, FP64 after FP32, a smaller space after a larger one): downloads equal
    and sigma products agree with those of builds on fresh memory."""
    ints, dets = mol_space(16, 32, 0, 4, 4, 20000)
    gpu_ctx.set_integrals(ints)
    diag = gpu_ctx.ham_diag(dets)
    x = np.random.default_rng(9).standard_normal(dets.shape[0])
    fresh = {}
    for prec, n in ((32, 20000), (64, 20000), (32, 12000)):
        h = gpu_ctx.build_h(dets[:n], diag[:n], value_prec=prec, storage=STORE_TILED)
        fresh[(prec, n)] = (h.download(), h.matrix_vector_prod(x[:n]))
        h.free()
    h = gpu_ctx.build_h(dets, diag, value_prec=32, storage=STORE_TILED)
    for prec, n in ((32, 20000), (64, 20000), (32, 12000)):
        h.free(recycle=True)
        h = gpu_ctx.build_h(dets[:n], diag[:n], value_prec=prec, storage=STORE_TILED)
        d, y = fresh[(prec, n)]
        for u, v in zip(d, h.download()):
            assert np.array_equal(u, v)
        # products accumulate in LDS atomics, whose order varies from run to run
        assert np.max(np.abs(y - h.matrix_vector_prod(x[:n]))) < 1e-12 * np.max(np.abs(y))
    h.free(recycle=True)
    # a full-storage build releases what a recycled matrix left behind
    f = gpu_ctx.build_h(dets, diag)
    assert f.nnz == fresh[(32, 20000)][0][1].size
    f.free()


@pytest.mark.parametrize("prec", [64, 32])
def test_tiled_hubbard_large_groups(gpu_ctx, oracle_lib, prec):
    """Every double vanishes: the tiled build finds beta singles inside alpha
    groups and alpha singles inside beta groups, by key lookup in groups wider
    than 48 and in groups split over several tiles (> 256 rows) and column
    blocks (> 4096); pattern, values and products equal full storage's."""
    ints, dets = hubbard_space(5, 4, 6, 6, 300000)
    _, ca = np.unique(dets[:, 0], return_counts=True)
    assert ca.max() > 4096 and np.sum((ca > 1) & (ca <= 48)) > 0
    gpu_ctx.set_integrals(ints)
    diag = gpu_ctx.ham_diag(dets)
    f = gpu_ctx.build_h(dets, diag, value_prec=prec)
    t = gpu_ctx.build_h(dets, diag, value_prec=prec, storage=STORE_TILED)
    assert t.nnz == f.nnz
    for u, v in zip(f.download(), t.download()):
        assert np.array_equal(u, v)
    x = np.random.default_rng(10).standard_normal(dets.shape[0])
    for mode in ([0, 1] if prec == 32 else [0]):
        a, b = f.matrix_vector_prod(x, mode=mode), t.matrix_vector_prod(x, mode=mode)
        assert np.max(np.abs(a - b)) < 1e-12 * np.max(np.abs(a))


@pytest.mark.parametrize("lattice", [(4, 4, 5, 5, 20000), (9, 8, 3, 3, 6000)])
def test_tiled_skips_singles_that_always_vanish(gpu_ctx, oracle_lib, lattice):
    """A single excitation whose element is zero for every occupation (on the
    lattice: a hop that is not a bond) takes no slot in tiled storage.  The
    same lattice with every non-bond h_pq set to 1e-14 (never above the
    1e-12 rule, but no longer identically zero) stores the same pairs, and
    its tiled matrix holds those hops as explicit zeros: more slots, the same
    download and products.  One-word (16 sites) and two-word (72 sites)
    strings."""
    from risq_amd.fcidump import Integrals, combine_2
    ints, dets = hubbard_space(*lattice)
    one = ints.one_body.copy()
    for p in range(ints.n_orb):
        for q in range(p):
            if one[combine_2(p + 1, q + 1)] == 0.0:
                one[combine_2(p + 1, q + 1)] = 1e-14
    tiny = Integrals(ints.n_orb, ints.e_core, one, ints.two_body)
    x = np.random.default_rng(11).standard_normal(dets.shape[0])
    got = []
    for I in (ints, tiny):
        gpu_ctx.set_integrals(I)
        diag = gpu_ctx.ham_diag(dets)
        t = gpu_ctx.build_h(dets, diag, storage=STORE_TILED)
        S = oracle_lib.OracleSparse().build(oracle_lib.OracleHam(I), dets, diag)
        for u, v in zip(t.download(), S.export()):
            assert np.array_equal(u, v)
        got.append((t.nnz, t.build_stats["stored_entries"], t.matrix_vector_prod(x)))
        t.free()
    (n0, s0, y0), (n1, s1, y1) = got
    # the non-bond hops among these determinants are many: each electron
    # has 4 bonds out of 15 possible hops
    assert n0 == n1 and s1 - s0 > 0.1 * 2 * n0, (n0, n1, s0, s1)
    assert np.max(np.abs(y0 - y1)) < 1e-12 * np.max(np.abs(y0))


@pytest.mark.parametrize("system", ["mol", "u128", "hubbard"])
def test_fp16_copy_and_three_phase_davidson(gpu_ctx, oracle_lib, system):
    """HCI_PREC_F16 tiled storage: the FP32 hi + lo values download unchanged,
    FP16-value products stay within FP16 rounding of the FP64 product, and the
    FP16 -> FP32 -> refinement Davidson converges to the FP64 eigenvalue.
    Molecular (one-word strings), 70 orbitals (two-word strings) and a 4x4
    Hubbard lattice whose alpha groups are wider than the keyed threshold
    (doubles vanish: singles found by keys inside alpha and beta groups)."""
    from risq_amd.hci import PREC_F16, SIGMA_F16
    if system == "mol":
        ints, dets = mol_space(*MOL0)
    elif system == "u128":
        ints, dets = mol_space(70, 8, 5, 3, 3, 1500)
    else:
        ints, dets = hubbard_space(4, 4, 5, 5, 20000)
    gpu_ctx.set_integrals(ints)
    diag = gpu_ctx.ham_diag(dets)
    f32 = gpu_ctx.build_h(dets, diag, value_prec=32, storage=STORE_TILED)
    f16 = gpu_ctx.build_h(dets, diag, value_prec=PREC_F16, storage=STORE_TILED)
    for u, v in zip(f32.download(), f16.download()):
        assert np.array_equal(u, v)
    x = np.random.default_rng(11).standard_normal(dets.shape[0])
    ref = f32.matrix_vector_prod(x)
    y16 = f16.matrix_vector_prod(x, mode=SIGMA_F16)
    rel = np.linalg.norm(y16 - ref) / np.linalg.norm(ref - diag * x)
    assert 0 < rel < 2e-3, rel
    if dets.shape[0] <= 5000:
        H = csr_to_dense(*_oracle_build(oracle_lib, ints, dets, diag).export())
        e_ref = np.linalg.eigvalsh(H)[0]
    else:  # too large for a dense solve: the FP64 FULL-storage Davidson
        full = gpu_ctx.build_h(dets, diag)
        r64 = gpu_ctx.davidson(full, tol=1e-7, etol=1e-11, max_iters=600)
        assert r64.converged
        e_ref = r64.energy
        full.free()
    r = gpu_ctx.davidson(f16, tol=1e-5, etol=1e-9, max_iters=400, mixed=2, restart_keep=2)
    assert r.converged and abs(r.energy - e_ref) < 1e-6, (r.energy, e_ref)
    assert r.stats["n_sigma_f16"] > 0 and r.stats["f16_iter"] >= 0 and r.stats["n_sigma_full"] > 0
    with pytest.raises(HciError, match="needs HCI_PREC_F16"):
        gpu_ctx.davidson(f32, mixed=2)
    with pytest.raises(HciError, match="needs HCI_STORE_TILED"):
        gpu_ctx.build_h(dets, diag, value_prec=PREC_F16)
