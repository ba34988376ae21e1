"""Deterministic PT2 on the GPU (hci_pt2_dtm, through the C ABI) against the
enumeration oracle (oracle/pt2_oracle.py, itself pinned by dense
Epstein-Nesbet sums in tests/test_pt2.py).  Bar: the same number of external
determinants and e2 within 1e-11 relative (each numerator is summed in the
oracle's order, discovery order; the terms are summed in another).""" 
import json
import os

import numpy as np
import pytest

from helpers import mol_space
from oracle import pt2_oracle
from risq_amd import hci_select as hs
from risq_amd.fcidump import read_fcidump

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")
CASES = json.load(open(os.path.join(GOLD, "hci_cases.json")))


def _solve(ctx, ints, dets):
    ctx.set_integrals(ints)
    h = ctx.build_h(dets, ctx.ham_diag(dets))
    r = ctx.davidson(h, tol=1e-8, etol=1e-12, max_iters=500)
    h.free()
    return r.energy, r.coeffs


@pytest.mark.parametrize("eps", [1e-5, 1e-7])
@pytest.mark.parametrize("n_core", [0, 1])
def test_pt2_matches_oracle(gpu_ctx, eps, n_core):
    ints, dets = mol_space(16, 32, 0, 4, 4, 3000)
    if n_core:  # orbital 0 is the core: keep it doubly occupied
        dets = dets[(dets[:, 0] & dets[:, 2] & np.uint64(1)) == 1][:2000]
    else:
        dets = dets[:2000]
    e0, c = _solve(gpu_ctx, ints, dets)
    gpu_ctx.excite_init(n_core, 4, 4, 1e-3)
    e2, n_ext = gpu_ctx.pt2_dtm(dets, c, e0, eps)
    o2, on = pt2_oracle.dtm_pt(ints, dets, c, e0, eps, n_core=n_core)
    print(f"eps={eps:g} n_core={n_core}: e2={e2:.12f} (oracle {o2:.12f}) external={n_ext}")
    assert n_ext == on
    assert abs(e2 - o2) < 1e-11 * abs(o2)


def test_pt2_parts_agree(gpu_ctx):
    """A small candidate budget splits the external determinants into hash
    parts (each walks every source determinant): same determinants, same
    energy up to the order of the final sum."""
    ints, dets = mol_space(16, 32, 0, 4, 4, 3000)
    e0, c = _solve(gpu_ctx, ints, dets)
    gpu_ctx.excite_init(0, 4, 4, 1e-3)
    e2, n = gpu_ctx.pt2_dtm(dets, c, e0, 1e-6)
    for batch in (2000000, 200000, 20000):
        e2b, nb = gpu_ctx.pt2_dtm(dets, c, e0, 1e-6, batch=batch)
        assert nb == n
        assert abs(e2b - e2) < 1e-13 * abs(e2)


def test_pt2_two_word_strings(gpu_ctx):
    """n_orb > 64: two-word strings through the same kernels."""
    ints, dets = mol_space(70, 8, 5, 3, 3, 1500)
    dets = dets[:120]
    e0, c = _solve(gpu_ctx, ints, dets)
    gpu_ctx.excite_init(0, 3, 3, 1e-3)
    e2, n_ext = gpu_ctx.pt2_dtm(dets, c, e0, 1e-5)
    o2, on = pt2_oracle.dtm_pt(ints, dets, c, e0, 1e-5)
    assert n_ext == on and n_ext > 0
    assert abs(e2 - o2) < 1e-11 * abs(o2)


@pytest.mark.timeout(300)
@pytest.mark.parametrize("name", ["be", "li2", "be2"])
def test_pt2_after_hci_on_reference_examples(gpu_ctx, name):
    """The reference examples' variational wavefunctions (HCI loop on the GPU)
    and their deterministic PT2 at eps 1e-6 (the examples' eps_pt_dtm), with
    the examples' frozen core and with every orbital active.  With the core
    frozen these small spaces already hold every external determinant that
    symmetry lets H reach (e2 = 0 from both sides); with it active they do
    not."""
    cs = CASES[name]
    ints = read_fcidump(os.path.join(GOLD, cs["fcidump"]), cs["norb"])
    solver = hs.GpuSolver(gpu_ctx, ints)
    r = hs.variational(solver, ints, cs["nup"], cs["ndn"], cs["eps_var"], n_core=cs["norb_core"])
    solver.close()
    words = hs._words(r.dets)
    for n_core in (cs["norb_core"], 0):
        gpu_ctx.excite_init(n_core, cs["nup"], cs["ndn"], cs["eps_var"])
        e2, n_ext = gpu_ctx.pt2_dtm(words, r.coeffs, r.energy, 1e-6)
        o2, on = pt2_oracle.dtm_pt(ints, words, r.coeffs, r.energy, 1e-6, n_core=n_core)
        print(f"{name} core={n_core}: E_var={r.energy:.10f} e2={e2:.12f} (oracle {o2:.12f}) external={n_ext}")
        assert n_ext == on
        assert abs(e2 - o2) <= 1e-11 * abs(o2)
    assert n_ext > 0 and e2 < 0


def _check_external(gpu_ctx, ints, dets, c, eps, n_core, n_up, n_dn):
    from risq_amd.hci_select import new_diag
    diag = gpu_ctx.ham_diag(dets)
    gpu_ctx.excite_init(n_core, n_up, n_dn, 1e-3)
    w, v, d, rem = gpu_ctx.matmul_external(dets, c, diag, eps)
    ow, ov, first, orem = pt2_oracle.matmul_external_dtm(ints, dets, c, eps, n_core, n_up, n_dn)
    print(f"eps={eps:g} n_core={n_core}: {w.shape[0]} external determinants (oracle {ow.shape[0]})")
    assert w.shape[0] == ow.shape[0] > 0
    assert np.array_equal(w, ow)  # the same determinants in the same first-discovery order
    # each coefficient sums the same products in the same (discovery) order
    assert np.max(np.abs(v - ov) / np.maximum(np.abs(ov), 1e-300)) < 1e-14
    # the first discoverer's new_diag, as the host restatement computes it
    ref = np.array([new_diag(ints, (int(dets[k][0]) | int(dets[k][1]) << 64, int(dets[k][2]) | int(dets[k][3]) << 64),
                             float(diag[k]), exc) for k, exc in first])
    assert np.array_equal(d, ref)
    assert np.max(np.abs(d - gpu_ctx.ham_diag(w))) < 1e-10
    assert np.max(np.abs(rem - orem) / np.maximum(np.abs(orem), 1e-300)) < 1e-13


@pytest.mark.parametrize("eps", [1e-5, 1e-7])
@pytest.mark.parametrize("n_core", [0, 1])
def test_external_product_matches_oracle(gpu_ctx, eps, n_core):
    """hci_matmul_external (approx_matmul_external_dtm_only) against the
    list-walking oracle: determinants and order exact, coefficients and
    remainders to the last bits, diagonals bit-identical to the new_diag
    restatement."""
    ints, dets = mol_space(16, 32, 0, 4, 4, 3000)
    if n_core:
        dets = dets[(dets[:, 0] & dets[:, 2] & np.uint64(1)) == 1][:600]
    else:
        dets = dets[:600]  # the oracle walks ~1e5 candidates per 100 sources in Python
    e0, c = _solve(gpu_ctx, ints, dets)
    _check_external(gpu_ctx, ints, dets, c, eps, n_core, 4, 4)


def test_external_product_two_word_strings(gpu_ctx):
    ints, dets = mol_space(70, 8, 5, 3, 3, 1500)
    dets = dets[:12]  # already ~3e5 external determinants (every double passes at 70 orbitals)
    e0, c = _solve(gpu_ctx, ints, dets)
    _check_external(gpu_ctx, ints, dets, c, 1e-5, 0, 3, 3)


def test_external_product_capacity(gpu_ctx):
    """Too small an output capacity fails with the needed count; the
    remainders are written either way."""
    import ctypes as C
    from risq_amd.hci import HCI_OK, MEM_HOST
    ints, dets = mol_space(16, 32, 0, 4, 4, 3000)
    dets = dets[:300]
    e0, c = _solve(gpu_ctx, ints, dets)
    diag = gpu_ctx.ham_diag(dets)
    gpu_ctx.excite_init(0, 4, 4, 1e-3)
    w, v, d, rem = gpu_ctx.matmul_external(dets, c, diag, 1e-6)
    out = np.empty((4, 4), dtype=np.uint64)
    oc, od, rem2 = np.empty(4), np.empty(4), np.empty(dets.shape[0])
    cnt = C.c_int64()
    dw = np.ascontiguousarray(dets, dtype=np.uint64)
    cc = np.ascontiguousarray(c, dtype=np.float64)
    rc = gpu_ctx.lib.hci_matmul_external(gpu_ctx.handle, dw.ctypes.data, cc.ctypes.data, diag.ctypes.data,
                                         dets.shape[0], 1e-6, out.ctypes.data, oc.ctypes.data, od.ctypes.data, 4,
                                         C.byref(cnt), rem2.ctypes.data, MEM_HOST)
    assert rc != HCI_OK and cnt.value == w.shape[0] > 4
    assert np.array_equal(rem2, rem)


def test_external_product_device_memory(gpu_ctx):
    """Inputs and outputs in device memory (HCI_MEM_DEVICE) give the same
    arrays as the host-memory call."""
    import ctypes as C
    from risq_amd.hci import DeviceBuffer, HCI_OK, MEM_DEVICE
    ints, dets = mol_space(16, 32, 0, 4, 4, 3000)
    dets = np.ascontiguousarray(dets[:400])
    e0, c = _solve(gpu_ctx, ints, dets)
    diag = gpu_ctx.ham_diag(dets)
    gpu_ctx.excite_init(0, 4, 4, 1e-3)
    w, v, d, rem = gpu_ctx.matmul_external(dets, c, diag, 1e-6)
    k = w.shape[0]
    dd, dc, dg = (DeviceBuffer.from_array(gpu_ctx, a) for a in (dets, np.ascontiguousarray(c), diag))
    ow, ov, od, orem = (DeviceBuffer(gpu_ctx, nb) for nb in (k * 32, k * 8, k * 8, dets.shape[0] * 8))
    cnt = C.c_int64()
    rc = gpu_ctx.lib.hci_matmul_external(gpu_ctx.handle, dd.ptr, dc.ptr, dg.ptr, dets.shape[0], 1e-6, ow.ptr, ov.ptr,
                                         od.ptr, k, C.byref(cnt), orem.ptr, MEM_DEVICE)
    assert rc == HCI_OK and cnt.value == k
    assert np.array_equal(ow.to_array(np.uint64, 4 * k).reshape(k, 4), w)
    assert np.array_equal(ov.to_array(np.float64, k), v)
    assert np.array_equal(od.to_array(np.float64, k), d)
    assert np.array_equal(orem.to_array(np.float64, dets.shape[0]), rem)


def _check_semistoch(gpu_ctx, ints, dets, c, e0, eps, n_core, n_up, n_dn):
    """hci_semistoch_dtm (the deterministic stage of the reference's default
    PT2) against pt2_oracle.semistoch_dtm: the same h_psi size, e2 with the
    first discoverers' new_diag within 1e-12 relative, rel_probs and every
    sample record equal."""
    from risq_amd.hci_select import new_diag
    diag = gpu_ctx.ham_diag(dets)
    gpu_ctx.excite_init(n_core, n_up, n_dn, 1e-3)
    e2, n_ext, rel, smp = gpu_ctx.semistoch_dtm(dets, c, diag, e0, eps)
    ow, on, firsts, orel, orec, _ = pt2_oracle.semistoch_dtm(ints, dets, c, e0, eps, n_core, n_up, n_dn)
    print(f"eps={eps:g} n_core={n_core}: h_psi {n_ext} (oracle {ow.shape[0]}), records {smp.size} "
          f"(oracle {len(orec)}), e2={e2:.12f}")
    assert n_ext == ow.shape[0] > 0
    names = {0: "opp", 1: "same", 2: "sing"}
    o2 = 0.0
    for num, (k, (kind, alpha, i, j, tgt)) in zip(on, firsts):
        parent = (int(dets[k][0]) | int(dets[k][1]) << 64, int(dets[k][2]) | int(dets[k][3]) << 64)
        if kind == 0:
            exc = ("opp", (i, j), tgt)
        elif kind == 1:
            exc = ("same", (i, j), tgt, alpha == 1)
        else:
            exc = ("sing", i, tgt, alpha == 1)
        o2 += num * num / (e0 - new_diag(ints, parent, float(diag[k]), exc))
    assert names
    assert abs(e2 - o2) < 1e-12 * abs(o2)
    assert np.array_equal(rel, orel)
    assert smp.size == len(orec) > 0
    orec = np.array(orec, dtype=object)
    for f, col in (("det", 0), ("kind", 1), ("is_alpha", 2), ("p", 3), ("q", 4)):
        assert np.array_equal(smp[f], orec[:, col].astype(np.int64)), f
    for f, col in (("sum_abs_h", 5), ("sum_h_squared", 6), ("sum_abs_hc", 7), ("sum_hc_squared", 8)):
        want = orec[:, col].astype(np.float64)
        assert np.all(np.abs(smp[f] - want) <= 1e-14 * np.abs(want)), f


@pytest.mark.parametrize("eps", [1e-5, 1e-6])
@pytest.mark.parametrize("n_core", [0, 1])
def test_semistoch_dtm_matches_oracle(gpu_ctx, eps, n_core):
    ints, dets = mol_space(16, 32, 0, 4, 4, 3000)
    if n_core:
        dets = dets[(dets[:, 0] & dets[:, 2] & np.uint64(1)) == 1][:300]
    else:
        dets = dets[:300]
    e0, c = _solve(gpu_ctx, ints, dets)
    _check_semistoch(gpu_ctx, ints, dets, c, e0, eps, n_core, 4, 4)


def test_semistoch_dtm_two_word_strings(gpu_ctx):
    ints, dets = mol_space(70, 8, 5, 3, 3, 1500)
    dets = dets[:12]
    e0, c = _solve(gpu_ctx, ints, dets)
    _check_semistoch(gpu_ctx, ints, dets, c, e0, 1e-5, 0, 3, 3)


def test_semistoch_dtm_h_psi_equals_dtm_pt(gpu_ctx):
    """The entries read before the cuts are dtm_pt's candidates, so h_psi is
    dtm_pt's external space: the same count and e2 up to the diagonals
    (first discoverer's new_diag against ham_diag)."""
    ints, dets = mol_space(16, 32, 0, 4, 4, 3000)
    dets = dets[:2000]
    e0, c = _solve(gpu_ctx, ints, dets)
    gpu_ctx.excite_init(0, 4, 4, 1e-3)
    e2, n_ext, rel, smp = gpu_ctx.semistoch_dtm(dets, c, gpu_ctx.ham_diag(dets), e0, 1e-6)
    p2, pn = gpu_ctx.pt2_dtm(dets, c, e0, 1e-6)
    assert n_ext == pn
    assert abs(e2 - p2) < 1e-11 * abs(p2)


def _check_sampled(gpu_ctx, ints, dets, c, eps, n_core, n_up, n_dn, variant):
    """hci_matmul_external_sampled against pt2_oracle.matmul_external_sampled:
    the same external determinants in the same first-discovery order,
    coefficients to the last bits (the same products summed in the same
    order), the first discoverer's new_diag bit-identical to the host
    restatement, and every DetOrbSample record equal (integers exact, sums
    within 1e-14 relative)."""
    from risq_amd.hci_select import new_diag
    diag = gpu_ctx.ham_diag(dets)
    gpu_ctx.excite_init(n_core, n_up, n_dn, 1e-3)
    w, v, d, smp = gpu_ctx.matmul_external_sampled(variant, dets, c, diag, eps)
    ow, ov, firsts, orec = pt2_oracle.matmul_external_sampled(ints, dets, c, eps, n_core, n_up, n_dn, variant)
    print(f"{variant} eps={eps:g} n_core={n_core}: {w.shape[0]} external determinants (oracle {ow.shape[0]}), "
          f"records {smp.size} (oracle {len(orec)})", flush=True)
    assert w.shape[0] == ow.shape[0] > 0
    assert np.array_equal(w, ow)
    assert np.max(np.abs(v - ov) / np.maximum(np.abs(ov), 1e-300)) < 1e-14
    ref = []
    for k, (kind, alpha, i, j, tgt) in firsts:
        parent = (int(dets[k][0]) | int(dets[k][1]) << 64, int(dets[k][2]) | int(dets[k][3]) << 64)
        exc = (("opp", (i, j), tgt) if kind == 0 else ("same", (i, j), tgt, alpha == 1) if kind == 1
               else ("sing", i, tgt, alpha == 1))
        ref.append(new_diag(ints, parent, float(diag[k]), exc))
    assert np.array_equal(d, np.array(ref))
    assert smp.size == len(orec) > 0
    orec = np.array(orec, dtype=object)
    for f, col in (("det", 0), ("kind", 1), ("is_alpha", 2), ("p", 3), ("q", 4)):
        assert np.array_equal(smp[f], orec[:, col].astype(np.int64)), f
    for f, col in (("sum_abs_h", 5), ("sum_h_squared", 6), ("sum_abs_hc", 7), ("sum_hc_squared", 8)):
        want = orec[:, col].astype(np.float64)
        assert np.all(np.abs(smp[f] - want) <= 1e-14 * np.abs(want)), f


SAMPLED = ["separate_doubles_and_singles", "skip_singles", "semistoch_singles", "no_singles"]


@pytest.mark.parametrize("variant", SAMPLED)
@pytest.mark.parametrize("eps,n_core", [(1e-5, 0), (1e-4, 1)])
def test_external_sampled_matches_oracle(gpu_ctx, variant, eps, n_core):
    """The four screened-sampler external products (src/wf/mod.rs:341-1193)
    on a 16-orbital space of 300 determinants."""
    ints, dets = mol_space(16, 32, 0, 4, 4, 3000)
    if n_core:
        dets = dets[(dets[:, 0] & dets[:, 2] & np.uint64(1)) == 1][:300]
    else:
        dets = dets[:300]
    e0, c = _solve(gpu_ctx, ints, dets)
    _check_sampled(gpu_ctx, ints, dets, c, eps, n_core, 4, 4, variant)


@pytest.mark.parametrize("variant", SAMPLED)
def test_external_sampled_two_word_strings(gpu_ctx, variant):
    ints, dets = mol_space(70, 8, 5, 3, 3, 1500)
    dets = dets[:8]
    e0, c = _solve(gpu_ctx, ints, dets)
    _check_sampled(gpu_ctx, ints, dets, c, 1e-4, 0, 3, 3, variant)


def test_external_sampled_capacity_and_device_memory(gpu_ctx):
    """Too small a record capacity fails with the needed count and writes no
    determinants; device-memory inputs and outputs give the host call's
    arrays; an unknown variant is refused."""
    import ctypes as C
    from risq_amd.hci import DET_ORB_SAMPLE, DeviceBuffer, EXT_SEMISTOCH_SINGLES, HCI_OK, MEM_DEVICE, MEM_HOST
    ints, dets = mol_space(16, 32, 0, 4, 4, 3000)
    dets = np.ascontiguousarray(dets[:300])
    e0, c = _solve(gpu_ctx, ints, dets)
    c = np.ascontiguousarray(c)
    diag = gpu_ctx.ham_diag(dets)
    gpu_ctx.excite_init(0, 4, 4, 1e-3)
    w, v, d, smp = gpu_ctx.matmul_external_sampled(EXT_SEMISTOCH_SINGLES, dets, c, diag, 1e-5)
    k, ns = w.shape[0], smp.size
    cnt, nsv = C.c_int64(-1), C.c_int64()
    out = np.empty((k, 4), dtype=np.uint64)
    oc, od = np.empty(k), np.empty(k)
    small = np.empty(4, dtype=DET_ORB_SAMPLE)
    rc = gpu_ctx.lib.hci_matmul_external_sampled(gpu_ctx.handle, EXT_SEMISTOCH_SINGLES, dets.ctypes.data,
                                                 c.ctypes.data, diag.ctypes.data, dets.shape[0], 1e-5,
                                                 out.ctypes.data, oc.ctypes.data, od.ctypes.data, k, C.byref(cnt),
                                                 small.ctypes.data, 4, C.byref(nsv), MEM_HOST)
    assert rc != HCI_OK and nsv.value == ns > 4 and cnt.value == -1
    rc = gpu_ctx.lib.hci_matmul_external_sampled(gpu_ctx.handle, 7, dets.ctypes.data, c.ctypes.data,
                                                 diag.ctypes.data, dets.shape[0], 1e-5, out.ctypes.data,
                                                 oc.ctypes.data, od.ctypes.data, k, C.byref(cnt), small.ctypes.data,
                                                 4, C.byref(nsv), MEM_HOST)
    assert rc != HCI_OK
    dd, dc, dg = (DeviceBuffer.from_array(gpu_ctx, a) for a in (dets, c, diag))
    ow, ov, od2, osm = (DeviceBuffer(gpu_ctx, nb) for nb in (k * 32, k * 8, k * 8, ns * DET_ORB_SAMPLE.itemsize))
    rc = gpu_ctx.lib.hci_matmul_external_sampled(gpu_ctx.handle, EXT_SEMISTOCH_SINGLES, dd.ptr, dc.ptr, dg.ptr,
                                                 dets.shape[0], 1e-5, ow.ptr, ov.ptr, od2.ptr, k, C.byref(cnt),
                                                 osm.ptr, ns, C.byref(nsv), MEM_DEVICE)
    assert rc == HCI_OK and cnt.value == k and nsv.value == ns
    assert np.array_equal(ow.to_array(np.uint64, 4 * k).reshape(k, 4), w)
    assert np.array_equal(ov.to_array(np.float64, k), v)
    assert np.array_equal(od2.to_array(np.float64, k), d)
    raw = osm.to_array(np.uint8, ns * DET_ORB_SAMPLE.itemsize)
    assert np.array_equal(raw.view(DET_ORB_SAMPLE), smp)


@pytest.mark.parametrize("variant", SAMPLED)
def test_external_sampled_nothing_deterministic(gpu_ctx, variant):
    """eps above every magnitude: no external determinant; the records (none
    for separate_doubles_and_singles, one per list with a valid external
    target for skip_singles and semistoch_singles' doubles, one per list for
    no_singles) equal the oracle's."""
    ints, dets = mol_space(16, 32, 0, 4, 4, 3000)
    dets = dets[:200]
    e0, c = _solve(gpu_ctx, ints, dets)
    diag = gpu_ctx.ham_diag(dets)
    gpu_ctx.excite_init(0, 4, 4, 1e-3)
    w, v, d, smp = gpu_ctx.matmul_external_sampled(variant, dets, c, diag, 1e6)
    ow, ov, firsts, orec = pt2_oracle.matmul_external_sampled(ints, dets, c, 1e6, 0, 4, 4, variant)
    assert w.shape[0] == ow.shape[0] == 0
    assert smp.size == len(orec)
    if orec:
        orec = np.array(orec, dtype=object)
        for f, col in (("det", 0), ("kind", 1), ("is_alpha", 2), ("p", 3), ("q", 4)):
            assert np.array_equal(smp[f], orec[:, col].astype(np.int64)), f
        for f, col in (("sum_abs_h", 5), ("sum_h_squared", 6), ("sum_abs_hc", 7), ("sum_hc_squared", 8)):
            want = orec[:, col].astype(np.float64)
            assert np.all(np.abs(smp[f] - want) <= 1e-14 * np.abs(want)), f
    w2, v2, d2, rem = gpu_ctx.matmul_external(dets, c, diag, 1e6)
    assert w2.shape[0] == 0 and np.all(rem >= 0)
