"""Python host binding of the MI355X HCI library (``librisq_hci.so``, C ABI in
``include/hci_gpu.h``).

The names mirror the reference's host interface so callers read the same:
``gen_sparse_ham_fast_lookup`` (src/var/fast_ham.rs:527-575),
``SparseMatUpperTri.matrix_vector_prod`` (src/var/sparse.rs:64-78),
``Davidson`` / ``sparse_optimize`` (src/var/eigenvalues/algorithms/davidson.rs:106-277,
src/var/davidson.rs:43-113).  Everything runs in the HIP library; there is no
CPU fallback: if the library is missing or no GPU is present, calls raise.
"""
from __future__ import annotations

import ctypes as C
import os
import weakref
from dataclasses import dataclass

import numpy as np

from .fcidump import Integrals

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "librisq_hci.so")

HCI_OK = 0
HCI_NOT_CONVERGED = 1
MEM_HOST = 0
MEM_DEVICE = 1
PREC_F64 = 64
PREC_F32 = 32
PREC_F16 = 16
SINGLES_ABS_GT_1E12 = 0
SINGLES_NONZERO = 1
SIGMA_FULL = 0
SIGMA_F32 = 1
SIGMA_F16 = 2
PRECOND_DPR = 0
PRECOND_MLP = 1
STORE_FULL = 0
STORE_UPPER = 1
STORE_TILED = 2
# external-product variants of hci_matmul_external_sampled (the reference's
# Wf::approx_matmul_external_* with a screened sampler)
EXT_SEPARATE = 1
EXT_SKIP_SINGLES = 2
EXT_SEMISTOCH_SINGLES = 3
EXT_NO_SINGLES = 4
EXT_VARIANTS = {"separate_doubles_and_singles": EXT_SEPARATE, "skip_singles": EXT_SKIP_SINGLES,
                "semistoch_singles": EXT_SEMISTOCH_SINGLES, "no_singles": EXT_NO_SINGLES}


class HciError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"hci error {code}: {msg}")
        self.code = code


class BuildOpts(C.Structure):
    _fields_ = [("value_prec", C.c_int), ("single_rule", C.c_int), ("mem", C.c_int), ("storage", C.c_int)]


class BuildStats(C.Structure):
    _fields_ = [("nnz_u", C.c_int64), ("n_alpha", C.c_int64), ("n_beta", C.c_int64),
                ("alpha_pairs", C.c_int64), ("seconds", C.c_double), ("mirror_pairs", C.c_int64),
                ("stored_entries", C.c_int64), ("matrix_bytes", C.c_int64)]


class DavOpts(C.Structure):
    _fields_ = [("tol", C.c_double), ("etol", C.c_double), ("max_iters", C.c_int),
                ("max_sub", C.c_int), ("precond", C.c_int), ("mixed", C.c_int),
                ("refine_tol", C.c_double), ("mem", C.c_int), ("restart_keep", C.c_int),
                ("f16_tol", C.c_double)]


class DavStats(C.Structure):
    _fields_ = [("iters", C.c_int), ("n_sigma_f32", C.c_int), ("n_sigma_full", C.c_int),
                ("converged", C.c_int), ("mlp_applied", C.c_int), ("mlp_fallbacks", C.c_int),
                ("refine_iter", C.c_int), ("residual", C.c_double), ("seconds", C.c_double),
                ("sigma_ms", C.c_double), ("n_sigma_f16", C.c_int), ("f16_iter", C.c_int),
                ("orth_second", C.c_int)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


MLP_NFEAT = 24
MLP_HIDDEN = 128
MLP_NPARAM = 19841


class MlpTrainOpts(C.Structure):
    _fields_ = [("epochs", C.c_int), ("batch", C.c_int), ("lr", C.c_float), ("seed", C.c_uint64),
                ("standardize", C.c_int), ("mem", C.c_int)]


class MlpTrainStats(C.Structure):
    _fields_ = [("first_epoch_loss", C.c_double), ("last_epoch_loss", C.c_double), ("seconds", C.c_double)]


# every symbol include/hci_gpu.h declares: name -> (restype, argtypes)
_P = C.c_void_p
SIGNATURES = {
    "hci_create": (C.c_int, [C.c_int, C.POINTER(_P)]),
    "hci_destroy": (None, [_P]),
    "hci_last_error": (C.c_char_p, [_P]),
    "hci_device_sync": (C.c_int, [_P]),
    "hci_set_integrals": (C.c_int, [_P, C.c_int, C.c_double, _P, C.c_size_t, _P, C.c_size_t]),
    "hci_ham_diag": (C.c_int, [_P, _P, C.c_int64, _P, C.c_int]),
    "hci_elements": (C.c_int, [_P, _P, _P, C.c_int64, _P, _P, C.c_int]),
    "hci_build_h": (C.c_int, [_P, _P, _P, C.c_int64, _P, C.c_int64, C.POINTER(BuildOpts),
                              C.POINTER(_P), C.POINTER(BuildStats)]),
    "hci_csr_info": (C.c_int, [_P, C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(C.c_int)]),
    "hci_csr_download": (C.c_int, [_P, _P, _P, _P, _P]),
    "hci_csr_row": (C.c_int, [_P, C.c_int64, _P, _P, C.c_int64, C.POINTER(C.c_int64)]),
    "hci_csr_free": (None, [_P]),
    "hci_csr_recycle": (C.c_int, [_P]),
    "hci_sigma": (C.c_int, [_P, C.c_int, _P, _P, C.c_int, C.c_int]),
    "hci_sigma_noalloc": (C.c_int, [_P, C.c_int, _P, _P, C.c_int]),
    "hci_csr_set_diagonal": (C.c_int, [_P, _P, C.c_int]),
    "hci_dav_default_opts": (None, [C.POINTER(DavOpts)]),
    "hci_davidson": (C.c_int, [_P, _P, _P, C.POINTER(DavOpts), C.POINTER(C.c_double), _P,
                               C.POINTER(DavStats)]),
    "hci_mlp_init": (C.c_int, [_P, C.c_uint64]),
    "hci_mlp_set_params": (C.c_int, [_P, _P, _P, _P]),
    "hci_mlp_get_params": (C.c_int, [_P, _P, _P, _P]),
    "hci_mlp_forward": (C.c_int, [_P, _P, C.c_int64, _P, C.c_int]),
    "hci_mlp_loss_grad": (C.c_int, [_P, _P, _P, _P, C.c_int64, _P, C.POINTER(C.c_double), C.c_int]),
    "hci_mlp_default_train_opts": (None, [C.POINTER(MlpTrainOpts)]),
    "hci_mlp_train": (C.c_int, [_P, _P, _P, _P, C.c_int64, C.POINTER(MlpTrainOpts), C.POINTER(MlpTrainStats)]),
    "hci_mlp_collect": (C.c_int, [_P, _P, _P, C.POINTER(DavOpts), C.c_int64, _P, _P, _P, C.c_int64,
                                  C.POINTER(C.c_int64)]),
    "hci_dev_alloc": (C.c_int, [_P, C.c_size_t, C.POINTER(_P)]),
    "hci_dev_free": (C.c_int, [_P, _P]),
    "hci_memcpy": (C.c_int, [_P, _P, _P, C.c_size_t, C.c_int, C.c_int]),
    "hci_event_record": (C.c_int, [_P, C.c_int]),
    "hci_event_elapsed_ms": (C.c_int, [_P, C.c_int, C.c_int, C.POINTER(C.c_float)]),
    "hci_excite_init": (C.c_int, [_P, C.c_int, C.c_int, C.c_int, C.c_double, C.POINTER(C.c_double)]),
    "hci_excite_info": (C.c_int, [_P, C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "hci_select": (C.c_int, [_P, _P, _P, _P, C.c_int64, C.c_double, C.c_int, _P, _P, C.c_int64,
                             C.POINTER(C.c_int64), C.c_int]),
    "hci_pt2_dtm": (C.c_int, [_P, _P, _P, C.c_int64, C.c_double, C.c_double, C.c_int64, C.POINTER(C.c_double),
                              C.POINTER(C.c_int64), C.c_int]),
    "hci_matmul_external": (C.c_int, [_P, _P, _P, _P, C.c_int64, C.c_double, _P, _P, _P, C.c_int64,
                                      C.POINTER(C.c_int64), _P, C.c_int]),
    "hci_semistoch_dtm": (C.c_int, [_P, _P, _P, _P, C.c_int64, C.c_double, C.c_double, C.POINTER(C.c_double),
                                    C.POINTER(C.c_int64), _P, _P, C.c_int64, C.POINTER(C.c_int64), C.c_int]),
    "hci_matmul_external_sampled": (C.c_int, [_P, C.c_int, _P, _P, _P, C.c_int64, C.c_double, _P, _P, _P,
                                              C.c_int64, C.POINTER(C.c_int64), _P, C.c_int64,
                                              C.POINTER(C.c_int64), C.c_int]),
}

# hci_det_orb_sample (include/hci_gpu.h), the reference's DetOrbSample
DET_ORB_SAMPLE = np.dtype([("det", np.int64), ("kind", np.int32), ("is_alpha", np.int32), ("p", np.int32),
                           ("q", np.int32), ("sum_abs_h", np.float64), ("sum_h_squared", np.float64),
                           ("sum_abs_hc", np.float64), ("sum_hc_squared", np.float64)])

_lib = None


def load_library(path: str = LIB_PATH) -> C.CDLL:
    """Load the HIP library.  Raises if it is missing: there is no fallback."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is not built; run __graft_entry__.build()")
    lib = C.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


class DeviceBuffer:
    """Library-owned device allocation (inputs kept resident in HBM)."""

    def __init__(self, ctx: "HciContext", nbytes: int):
        self.ctx = ctx
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        ctx._check(ctx.lib.hci_dev_alloc(ctx.handle, self.nbytes, C.byref(p)))
        self.ptr = p.value
        ctx._owned.add(self)

    @classmethod
    def from_array(cls, ctx: "HciContext", a: np.ndarray) -> "DeviceBuffer":
        a = np.ascontiguousarray(a)
        b = cls(ctx, a.nbytes)
        ctx._check(ctx.lib.hci_memcpy(ctx.handle, b.ptr, _ptr(a), a.nbytes, MEM_DEVICE, MEM_HOST))
        return b

    def to_array(self, dtype, count: int) -> np.ndarray:
        out = np.empty(count, dtype=dtype)
        self.ctx._check(self.ctx.lib.hci_memcpy(self.ctx.handle, _ptr(out), self.ptr, out.nbytes,
                                                MEM_HOST, MEM_DEVICE))
        return out

    def free(self):
        if self.ptr and self.ctx.handle:
            self.ctx.lib.hci_dev_free(self.ctx.handle, self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class HciContext:
    """One GPU context (hci_ctx): integral tables, a HIP stream, the last error."""

    def __init__(self, device: int = 0):
        self.lib = load_library()
        h = C.c_void_p()
        rc = self.lib.hci_create(device, C.byref(h))
        if rc != HCI_OK:
            raise HciError(rc, f"hci_create(device={device}) failed (no GPU?)")
        self.handle = h.value
        self.ints: Integrals | None = None
        # objects holding library handles of this context: freed before it
        self._owned: weakref.WeakSet = weakref.WeakSet()

    def close(self):
        if self.handle:
            for obj in list(self._owned):
                obj.free()
            self.lib.hci_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc: int) -> int:
        if rc < 0:
            raise HciError(rc, self.lib.hci_last_error(self.handle).decode())
        return rc

    # -- integrals and elements ------------------------------------------------
    def set_integrals(self, ints: Integrals) -> None:
        one = np.ascontiguousarray(ints.one_body, dtype=np.float64)
        two = np.ascontiguousarray(ints.two_body, dtype=np.float64)
        self._check(self.lib.hci_set_integrals(self.handle, ints.n_orb, ints.e_core, _ptr(one), one.size,
                                               _ptr(two), two.size))
        self.ints = ints

    def ham_diag(self, dets: np.ndarray) -> np.ndarray:
        dets = np.ascontiguousarray(dets, dtype=np.uint64)
        out = np.empty(dets.shape[0])
        self._check(self.lib.hci_ham_diag(self.handle, _ptr(dets), dets.shape[0], _ptr(out), MEM_HOST))
        return out

    def elements(self, det1: np.ndarray, det2: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
        det1 = np.ascontiguousarray(det1, dtype=np.uint64)
        det2 = np.ascontiguousarray(det2, dtype=np.uint64)
        m = det1.shape[0]
        val = np.empty(m)
        deg = np.empty(m, dtype=np.int8)
        self._check(self.lib.hci_elements(self.handle, _ptr(det1), _ptr(det2), m, _ptr(val), _ptr(deg),
                                          MEM_HOST))
        return val, deg

    # -- heat-bath selection --------------------------------------------------------
    def excite_init(self, n_core: int, n_up: int, n_dn: int, eps_var: float) -> float:
        """Builds the sorted excitation lists; returns the schedule's first epsilon."""
        e = C.c_double()
        self._check(self.lib.hci_excite_init(self.handle, n_core, n_up, n_dn, eps_var, C.byref(e)))
        return e.value

    def excite_info(self) -> tuple[float, float, float]:
        a, b, c = C.c_double(), C.c_double(), C.c_double()
        self._check(self.lib.hci_excite_info(self.handle, C.byref(a), C.byref(b), C.byref(c)))
        return a.value, b.value, c.value

    def select(self, dets: np.ndarray, coeffs: np.ndarray, eps: float, singles: bool, diag=None):
        """New determinants (n_new, 4) uint64 in the reference's discovery order;
        with the space's diagonal, also the new determinants' diagonals
        (new_diag_* from their first discoverer): (dets, diags)."""
        dets = np.ascontiguousarray(dets, dtype=np.uint64)
        coeffs = np.ascontiguousarray(coeffs, dtype=np.float64)
        d_in = None if diag is None else np.ascontiguousarray(diag, dtype=np.float64)
        cnt = C.c_int64()
        cap = max(1024, 4 * dets.shape[0])
        while True:
            out = np.empty((cap, 4), dtype=np.uint64)
            d_out = np.empty(cap) if d_in is not None else None
            rc = self.lib.hci_select(self.handle, _ptr(dets), _ptr(coeffs), None if d_in is None else _ptr(d_in),
                                     dets.shape[0], eps, int(singles), _ptr(out),
                                     None if d_out is None else _ptr(d_out), cap, C.byref(cnt), MEM_HOST)
            if rc == HCI_OK:
                if d_in is None:
                    return out[:cnt.value]
                return out[:cnt.value], d_out[:cnt.value]
            if cnt.value > cap:
                cap = cnt.value
                continue
            self._check(rc)

    def pt2_dtm(self, dets: np.ndarray, coeffs: np.ndarray, e0: float, eps: float,
                batch: int = 0) -> tuple[float, int]:
        """Deterministic PT2 energy of the space and the number of external
        determinants that contribute (dtm_pt, src/pt/mod.rs:81-122); needs
        excite_init on the same integrals.  batch: candidate excitations per
        pass over the source determinants (0: the library's default); more
        candidates split the external determinants into hash parts."""
        dets = np.ascontiguousarray(dets, dtype=np.uint64)
        coeffs = np.ascontiguousarray(coeffs, dtype=np.float64)
        e2 = C.c_double()
        n_ext = C.c_int64()
        self._check(self.lib.hci_pt2_dtm(self.handle, _ptr(dets), _ptr(coeffs), dets.shape[0], e0, eps, batch,
                                         C.byref(e2), C.byref(n_ext), MEM_HOST))
        return e2.value, n_ext.value

    def matmul_external(self, dets: np.ndarray, coeffs: np.ndarray, diag: np.ndarray, eps: float):
        """The screened external product of Wf::approx_matmul_external_dtm_only
        (src/wf/mod.rs:186-339): (external determinants (k, 4) uint64 in
        first-discovery order, their coefficients sum_i H_ai c_i, their
        diagonals from the first discoverer, and per source determinant
        c_i^2 times the squared magnitudes left below the cut).  Needs
        excite_init on the same integrals."""
        dets = np.ascontiguousarray(dets, dtype=np.uint64)
        coeffs = np.ascontiguousarray(coeffs, dtype=np.float64)
        diag = np.ascontiguousarray(diag, dtype=np.float64)
        n = dets.shape[0]
        cnt = C.c_int64()
        rem = np.empty(n)
        cap = max(1024, 8 * n)
        while True:
            out = np.empty((cap, 4), dtype=np.uint64)
            oc = np.empty(cap)
            od = np.empty(cap)
            rc = self.lib.hci_matmul_external(self.handle, _ptr(dets), _ptr(coeffs), _ptr(diag), n, eps, _ptr(out),
                                              _ptr(oc), _ptr(od), cap, C.byref(cnt), _ptr(rem), MEM_HOST)
            if rc == HCI_OK:
                k = cnt.value
                return out[:k], oc[:k], od[:k], rem
            if cnt.value > cap:
                cap = cnt.value
                continue
            self._check(rc)

    def matmul_external_sampled(self, variant, dets: np.ndarray, coeffs: np.ndarray, diag: np.ndarray,
                                eps: float):
        """The external products that set up a screened sampler
        (Wf::approx_matmul_external_separate_doubles_and_singles,
        _skip_singles, _semistoch_singles, _no_singles; src/wf/mod.rs:341-1193),
        `variant` an EXT_* code or its name: (external determinants (k, 4)
        uint64 in first-discovery order, their coefficients sum_i H_ai c_i,
        their diagonals from the first discoverer, the DetOrbSample records
        as a DET_ORB_SAMPLE array in push order).  Needs excite_init on the
        same integrals."""
        v = EXT_VARIANTS[variant] if isinstance(variant, str) else int(variant)
        dets = np.ascontiguousarray(dets, dtype=np.uint64)
        coeffs = np.ascontiguousarray(coeffs, dtype=np.float64)
        diag = np.ascontiguousarray(diag, dtype=np.float64)
        n = dets.shape[0]
        cnt = C.c_int64()
        ns = C.c_int64()
        cap = max(1024, 8 * n)
        scap = max(1024, 64 * n)
        while True:
            out = np.empty((cap, 4), dtype=np.uint64)
            oc = np.empty(cap)
            od = np.empty(cap)
            smp = np.empty(scap, dtype=DET_ORB_SAMPLE)
            rc = self.lib.hci_matmul_external_sampled(self.handle, v, _ptr(dets), _ptr(coeffs), _ptr(diag), n, eps,
                                                      _ptr(out), _ptr(oc), _ptr(od), cap, C.byref(cnt), _ptr(smp),
                                                      scap, C.byref(ns), MEM_HOST)
            if rc == HCI_OK:
                k = cnt.value
                return out[:k], oc[:k], od[:k], smp[:ns.value]
            if ns.value > scap:
                scap = ns.value
                continue
            if cnt.value > cap:
                cap = cnt.value
                continue
            self._check(rc)

    def semistoch_dtm(self, dets: np.ndarray, coeffs: np.ndarray, diag: np.ndarray, e0: float, eps: float):
        """The deterministic stage of new_semistoch_enpt2_dtm_diag_singles
        (src/semistoch/mod.rs:262-318): (e2 of h_psi, its size, rel_probs per
        determinant, the DetOrbSample records as a DET_ORB_SAMPLE array).
        Needs excite_init on the same integrals."""
        dets = np.ascontiguousarray(dets, dtype=np.uint64)
        coeffs = np.ascontiguousarray(coeffs, dtype=np.float64)
        diag = np.ascontiguousarray(diag, dtype=np.float64)
        n = dets.shape[0]
        e2 = C.c_double()
        n_ext = C.c_int64()
        ns = C.c_int64()
        rel = np.empty(n)
        cap = max(1024, 64 * n)
        while True:
            smp = np.empty(cap, dtype=DET_ORB_SAMPLE)
            rc = self.lib.hci_semistoch_dtm(self.handle, _ptr(dets), _ptr(coeffs), _ptr(diag), n, e0, eps,
                                            C.byref(e2), C.byref(n_ext), _ptr(rel), _ptr(smp), cap, C.byref(ns),
                                            MEM_HOST)
            if rc == HCI_OK:
                return e2.value, n_ext.value, rel, smp[:ns.value]
            if ns.value > cap:
                cap = ns.value
                continue
            self._check(rc)

    # -- timing ------------------------------------------------------------------
    def record(self, slot: int) -> None:
        self._check(self.lib.hci_event_record(self.handle, slot))

    def elapsed_ms(self, a: int, b: int) -> float:
        ms = C.c_float()
        self._check(self.lib.hci_event_elapsed_ms(self.handle, a, b, C.byref(ms)))
        return float(ms.value)

    def sync(self) -> None:
        self._check(self.lib.hci_device_sync(self.handle))

    # -- sparse H --------------------------------------------------------------------
    def build_h(self, dets, diag, n: int | None = None, prev: "SparseMatUpperTri | None" = None,
                n_stored: int = 0, value_prec: int = PREC_F64,
                single_rule: int = SINGLES_ABS_GT_1E12, storage: int = STORE_FULL) -> "SparseMatUpperTri":
        """hci_build_h.  dets/diag are numpy arrays (host) or DeviceBuffers."""
        if isinstance(dets, DeviceBuffer):
            mem = MEM_DEVICE
            pd, pg = dets.ptr, diag.ptr
            assert n is not None
        else:
            mem = MEM_HOST
            dets = np.ascontiguousarray(dets, dtype=np.uint64)
            diag = np.ascontiguousarray(diag, dtype=np.float64)
            n = dets.shape[0]
            pd, pg = _ptr(dets), _ptr(diag)
        opts = BuildOpts(value_prec, single_rule, mem, storage)
        out = C.c_void_p()
        st = BuildStats()
        self._check(self.lib.hci_build_h(self.handle, pd, pg, n, prev.handle if prev else None,
                                         n_stored, C.byref(opts), C.byref(out), C.byref(st)))
        return SparseMatUpperTri(self, out.value, st)

    # -- learned preconditioner ------------------------------------------------------
    def mlp_init(self, seed: int = 7) -> None:
        self._check(self.lib.hci_mlp_init(self.handle, seed))

    def mlp_set_params(self, params: np.ndarray, mean: np.ndarray, std: np.ndarray) -> None:
        p = np.ascontiguousarray(params, dtype=np.float32)
        mu = np.ascontiguousarray(mean, dtype=np.float32)
        sd = np.ascontiguousarray(std, dtype=np.float32)
        assert p.size == MLP_NPARAM and mu.size == MLP_NFEAT and sd.size == MLP_NFEAT
        self._check(self.lib.hci_mlp_set_params(self.handle, _ptr(p), _ptr(mu), _ptr(sd)))

    def mlp_get_params(self) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
        p = np.empty(MLP_NPARAM, dtype=np.float32)
        mu = np.empty(MLP_NFEAT, dtype=np.float32)
        sd = np.empty(MLP_NFEAT, dtype=np.float32)
        self._check(self.lib.hci_mlp_get_params(self.handle, _ptr(p), _ptr(mu), _ptr(sd)))
        return p, mu, sd

    def mlp_forward(self, X: np.ndarray) -> np.ndarray:
        X = np.ascontiguousarray(X, dtype=np.float32)
        out = np.empty(X.shape[0], dtype=np.float32)
        self._check(self.lib.hci_mlp_forward(self.handle, _ptr(X), X.shape[0], _ptr(out), MEM_HOST))
        return out

    def mlp_loss_grad(self, X, m, y) -> tuple[float, np.ndarray]:
        X = np.ascontiguousarray(X, dtype=np.float32)
        m = np.ascontiguousarray(m, dtype=np.float32)
        y = np.ascontiguousarray(y, dtype=np.float32)
        g = np.empty(MLP_NPARAM, dtype=np.float32)
        loss = C.c_double()
        self._check(self.lib.hci_mlp_loss_grad(self.handle, _ptr(X), _ptr(m), _ptr(y), X.shape[0], _ptr(g),
                                               C.byref(loss), MEM_HOST))
        return loss.value, g

    def mlp_train(self, X, m, y, epochs: int = 20, batch: int = 4096, lr: float = 1e-3, seed: int = 7,
                  standardize: bool = True) -> dict:
        X = np.ascontiguousarray(X, dtype=np.float32)
        m = np.ascontiguousarray(m, dtype=np.float32)
        y = np.ascontiguousarray(y, dtype=np.float32)
        o = MlpTrainOpts()
        self.lib.hci_mlp_default_train_opts(C.byref(o))
        o.epochs, o.batch, o.lr, o.seed, o.standardize, o.mem = epochs, batch, lr, seed, int(standardize), MEM_HOST
        st = MlpTrainStats()
        self._check(self.lib.hci_mlp_train(self.handle, _ptr(X), _ptr(m), _ptr(y), X.shape[0], C.byref(o),
                                           C.byref(st)))
        return {k: getattr(st, k) for k, _ in st._fields_}

    def mlp_collect(self, h: "SparseMatUpperTri", x_exact: np.ndarray, per_iter: int, cap: int,
                    max_iters: int = 300, tol: float = 1e-5, etol: float = 1e-9, restart_keep: int = 1):
        x_exact = np.ascontiguousarray(x_exact, dtype=np.float64)
        opts = DavOpts()
        self.lib.hci_dav_default_opts(C.byref(opts))
        opts.tol, opts.etol, opts.max_iters = tol, etol, max_iters
        opts.restart_keep = restart_keep
        X = np.empty((cap, MLP_NFEAT), dtype=np.float32)
        m = np.empty(cap, dtype=np.float32)
        y = np.empty(cap, dtype=np.float32)
        cnt = C.c_int64()
        self._check(self.lib.hci_mlp_collect(self.handle, h.handle, _ptr(x_exact), C.byref(opts), per_iter,
                                             _ptr(X), _ptr(m), _ptr(y), cap, C.byref(cnt)))
        k = cnt.value
        return X[:k], m[:k], y[:k]

    def davidson(self, h: "SparseMatUpperTri", init: np.ndarray | None = None, tol: float = 1e-4,
                 etol: float = 1e-8, max_iters: int = 50, max_sub: int = 11,
                 precond: int = PRECOND_DPR, mixed: int = 0,
                 refine_tol: float = 0.0, restart_keep: int = 1, f16_tol: float = 0.0) -> "DavidsonResult":
        """mixed: 0 full-precision sigmas; 1 (or True) FP32 then refinement;
        2 FP16 (HCI_PREC_F16 storage) then FP32 then refinement."""
        opts = DavOpts()
        self.lib.hci_dav_default_opts(C.byref(opts))
        opts.tol, opts.etol, opts.max_iters, opts.max_sub = tol, etol, max_iters, max_sub
        opts.precond, opts.mixed, opts.refine_tol, opts.mem = precond, int(mixed), refine_tol, MEM_HOST
        opts.restart_keep = restart_keep
        opts.f16_tol = f16_tol
        coeffs = np.zeros(h.n)
        e0 = C.c_double(0.0)
        st = DavStats()
        pin = None
        if init is not None:
            init = np.ascontiguousarray(init, dtype=np.float64)
            pin = _ptr(init)
        rc = self._check(self.lib.hci_davidson(self.handle, h.handle, pin, C.byref(opts), C.byref(e0),
                                               _ptr(coeffs), C.byref(st)))
        return DavidsonResult(rc == HCI_OK, e0.value, coeffs, st.as_dict())


@dataclass
class DavidsonResult:
    converged: bool
    energy: float
    coeffs: np.ndarray
    stats: dict


class SparseMatUpperTri:
    """Device-resident upper-triangle CSR (reference SparseMatUpperTri,
    src/var/sparse.rs:24-36): diagonal + rows of (col > row, H_ij)."""

    def __init__(self, ctx: HciContext, handle: int, stats: BuildStats):
        self.ctx = ctx
        self.handle = handle
        n = C.c_int64()
        nnz = C.c_int64()
        prec = C.c_int()
        ctx._check(ctx.lib.hci_csr_info(handle, C.byref(n), C.byref(nnz), C.byref(prec)))
        self.n, self.nnz, self.prec = n.value, nnz.value, prec.value
        self.build_stats = {k: getattr(stats, k) for k, _ in stats._fields_}
        ctx._owned.add(self)

    def free(self, recycle: bool = False):
        """Release the matrix; with recycle=True a tiled matrix's entry arrays
        stay with the context for its next tiled build (hci_csr_recycle)."""
        if self.handle and self.ctx.handle:
            if recycle:
                self.ctx._check(self.ctx.lib.hci_csr_recycle(self.handle))
            else:
                self.ctx.lib.hci_csr_free(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def download(self) -> tuple[np.ndarray, np.ndarray, np.ndarray, np.ndarray]:
        rp = np.empty(self.n + 1, dtype=np.int64)
        col = np.empty(max(self.nnz, 1), dtype=np.int32)
        val = np.empty(max(self.nnz, 1))
        diag = np.empty(self.n)
        self.ctx._check(self.ctx.lib.hci_csr_download(self.handle, _ptr(rp), _ptr(col), _ptr(val), _ptr(diag)))
        return rp, col[:self.nnz], val[:self.nnz], diag

    def row(self, i: int) -> tuple[np.ndarray, np.ndarray]:
        """Upper-triangle row i in caller order: (columns j > i ascending, H_ij)."""
        cnt = C.c_int64()
        self.ctx._check(self.ctx.lib.hci_csr_row(self.handle, i, None, None, 0, C.byref(cnt)))
        cols = np.empty(max(cnt.value, 1), dtype=np.int32)
        vals = np.empty(max(cnt.value, 1))
        self.ctx._check(self.ctx.lib.hci_csr_row(self.handle, i, _ptr(cols), _ptr(vals), cnt.value, C.byref(cnt)))
        return cols[:cnt.value], vals[:cnt.value]

    def set_diagonal(self, diag: np.ndarray) -> None:
        """MatrixOperations::set_diagonal (src/var/eigenvalues/matrix_operations.rs:14-27)."""
        diag = np.ascontiguousarray(diag, dtype=np.float64)
        assert diag.shape == (self.n,)
        self.ctx._check(self.ctx.lib.hci_csr_set_diagonal(self.handle, _ptr(diag), MEM_HOST))

    def diagonal(self) -> np.ndarray:
        return self.download()[3]

    def matrix_vector_prod(self, x: np.ndarray, mode: int = SIGMA_FULL) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.float64)
        ncol = 1 if x.ndim == 1 else x.shape[1]
        xc = np.asfortranarray(x.reshape(self.n, ncol))
        y = np.zeros((self.n, ncol), order="F")
        self.ctx._check(self.ctx.lib.hci_sigma(self.handle, ncol, xc.ctypes.data, y.ctypes.data, mode, MEM_HOST))
        return y[:, 0] if x.ndim == 1 else y

    def sigma_device(self, x: DeviceBuffer, y: DeviceBuffer, mode: int = SIGMA_FULL) -> None:
        self.ctx._check(self.ctx.lib.hci_sigma(self.handle, 1, x.ptr, y.ptr, mode, MEM_DEVICE))


def gen_sparse_ham_fast_lookup(ctx: HciContext, dets: np.ndarray, diag: np.ndarray,
                               prev: SparseMatUpperTri | None = None, n_stored: int = 0,
                               value_prec: int = PREC_F64) -> SparseMatUpperTri:
    """The reference's H build (src/var/fast_ham.rs:527-575), on the GPU."""
    return ctx.build_h(dets, diag, prev=prev, n_stored=n_stored, value_prec=value_prec)


def sparse_optimize(ctx: HciContext, dets: np.ndarray, diag: np.ndarray,
                    prev: SparseMatUpperTri | None = None, n_stored: int = 0,
                    init: np.ndarray | None = None, coeff_eps: float = 1e-4,
                    energy_eps: float = 1e-8, storage: int = STORE_TILED,
                    **dav) -> tuple[SparseMatUpperTri, DavidsonResult]:
    """Build H then run Davidson (reference src/var/davidson.rs:43-113).  The
    n <= 100 small-system path of the reference stores singles iff H != 0.
    storage: tiled (the fast product; an incremental build rebuilds it whole)
    or FULL (keeps the previous entries, bitwise reproducible products)."""
    rule = SINGLES_NONZERO if dets.shape[0] <= 100 else SINGLES_ABS_GT_1E12
    h = ctx.build_h(dets, diag, prev=prev, n_stored=n_stored, single_rule=rule, storage=storage)
    res = ctx.davidson(h, init=init, tol=coeff_eps, etol=energy_eps, **dav)
    return h, res
